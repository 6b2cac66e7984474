/*
 * orc.h -- CPU oracle for the edge-assembly + pressure-Poisson hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * (Exawind/nalu-wind) algorithms on the hot path.  It exists so that tests,
 * the smoke check and the CPU-baseline leg of the benchmark can check the
 * HIP implementation against it.  Nothing in the product library links to,
 * loads or calls this code; the product path fails when its HIP library is
 * missing instead of falling back to this file.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * FP64 and follows the reference expression order where it is visible, so
 * that golden values transcribed from the reference unit tests are matched
 * to ~1e-15.
 *
 * Pinning (how each function is checked independently of itself):
 *   geometry  : reference unit-cube values (dual volume 0.125, |A|^2 0.0625),
 *               closure sum_e +-A_e = 0 at interior nodes, sum of dual volumes
 *               equals the box volume, uniform-box closed forms.
 *   edge algs : reference golden LHS/RHS of the continuity, scalar
 *               (Tpetra CSR) and momentum unit tests; orientation invariance;
 *               zero row sums of pure diffusion; symmetry of the continuity LHS.
 *   scatter   : reference CSR golds (row offsets, columns), FixPressureAtNode
 *               and applyDirichletBCs golds; dense-vs-CSR agreement.
 *   peclet    : reference PecletFunction unit-test values.
 *   gmres     : dense LU solve on tiny meshes; explicit residual.
 *   mpnn      : finite-difference gradient check of the analytic backward.
 */
#ifndef NALU_AMD_ORACLE_ORC_H
#define NALU_AMD_ORACLE_ORC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Master-element geometry (Hex8)                                     */
/* ------------------------------------------------------------------ */

/* SCS area vectors of a Hex8 (12 sub-control surfaces, one per edge),
 * restating HexSCS::determinant_scs (src/master_element/Hex8CVFEM.C:565-591),
 * subdivide_hex_8 and quad_area_by_triangulation
 * (include/master_element/Hex8GeometryFunctions.h:258-330, 36-83). */
void orc_hex_scs_areav(const double coords[24], double areav[36]);

/* SCV volumes of a Hex8 (8 sub-volumes), restating HexSCV::determinant_scv
 * (src/master_element/Hex8CVFEM.C:365-389) and hex_volume_grandy
 * (include/master_element/Hex8GeometryFunctions.h:86-152). */
void orc_hex_scv_volume(const double coords[24], double vol[8]);

/* Element-local L/R node pairs of the 12 Hex8 SCS (HexSCS lrscv_,
 * include/master_element/Hex8CVFEM.h:292-293). */
extern const int orc_hex_lrscv[24];

/* Tet4 / Wed6 SCS area vectors (6 / 9 sub-control surfaces, one per edge),
 * restating TetSCS::determinant_scs (src/master_element/Tet4CVFEM.C:
 * 522-618) and WedSCS::determinant_scs (src/master_element/Wed6CVFEM.C:
 * 541-646), and their L/R tables (include/master_element/Tet4CVFEM.h:266,
 * Wed6CVFEM.h:268). */
void orc_tet_scs_areav(const double coords[12], double areav[18]);
void orc_wed_scs_areav(const double coords[18], double areav[27]);
extern const int orc_tet_lrscv[12];
extern const int orc_wed_lrscv[18];

/* Tet4 / Wed6 SCV volumes (4 / 6 sub-control volumes, sub-volume i at
 * node i), restating TetSCV::determinant_scv (src/master_element/
 * Tet4CVFEM.C:243-343) and WedSCV::determinant_scv (src/master_element/
 * Wed6CVFEM.C:267-369): hex_volume_grandy of eight sub-points each. */
void orc_tet_scv_volume(const double coords[12], double vol[4]);
void orc_wed_scv_volume(const double coords[18], double vol[6]);

/* Pyr5 SCS area vectors (12 sub-control surfaces, L/R in orc_pyr_lrscv:
 * the base edges one each, the apex edges two each) and SCV volumes (4
 * bent hexahedra at the base nodes, an octahedron at the apex), restating
 * PyrSCS::determinant_scs and PyrSCV::determinant_scv
 * (src/master_element/Pyr5CVFEM.C:437-571, 771-900) and lrscv_
 * (include/master_element/Pyr5CVFEM.h:300-301). */
void orc_pyr_scs_areav(const double coords[15], double areav[36]);
void orc_pyr_scv_volume(const double coords[15], double vol[5]);
extern const int orc_pyr_lrscv[24];

/* Dual nodal volumes of a mixed mesh (type 0 Hex8, 1 Tet4, 2 Wed6, 3 Pyr5;
 * conn [nelem][8]), src/ngp_algorithms/GeometryInteriorAlg.C:70-112. */
void orc_dual_volume_mixed(int64_t nelem, const int32_t* type,
                           const int32_t* conn, const double* coords,
                           int64_t nnode, double* dualvol);

/* Edge area vectors of a mixed mesh (type 0 Hex8, 1 Tet4, 2 Wed6, 3 Pyr5;
 * conn [nelem][8]); edges sorted (node[0] < node[1]).  Returns -1 if an element
 * edge is missing from the list. */
int orc_edge_area_vec_mixed(int64_t nelem, const int32_t* type,
                            const int32_t* conn, const double* coords,
                            const int32_t* edges, int64_t nedge, double* areav);

/* ------------------------------------------------------------------ */
/* Mesh: edges and edge/node geometry                                  */
/* ------------------------------------------------------------------ */

/* Unique edges of a Hex8 mesh in first-encounter order (element order,
 * then element-local edge ordinal); node[0] is the node with the smaller
 * global id (the polarity STK create_edges gives a LINE_2 edge).  Writes
 * elem_edges[nelem*12] = edge index of each element-local edge.  Returns
 * the edge count, or -1 if edges_cap is too small. */
int64_t orc_create_edges_hex(int64_t nelem, const int32_t* conn,
                             int64_t nnode, int32_t* edges, int64_t edges_cap,
                             int32_t* elem_edges);

/* Edge area vectors: sum over elements of the SCS area vector, sign +1
 * when the element-local left node equals edge node[0]
 * (src/ngp_algorithms/GeometryInteriorAlg.C:163-225). */
void orc_edge_area_vec_hex(int64_t nelem, const int32_t* conn,
                           const double* coords, const int32_t* elem_edges,
                           const int32_t* edges, int64_t nedge, double* areav);

/* Dual nodal volume = sum of SCV volumes
 * (src/ngp_algorithms/GeometryInteriorAlg.C:70-112). */
void orc_dual_volume_hex(int64_t nelem, const int32_t* conn,
                         const double* coords, int64_t nnode, double* dualvol);

/* Edge-based nodal gradient, interior-edge part only
 * (src/ngp_algorithms/NodalGradEdgeAlg.C:58-111).  phi is [N][dim1],
 * grad is [N][dim1*3] and is overwritten. */
void orc_nodal_grad_edge(int64_t nnode, int64_t nedge, const int32_t* edges,
                         const double* areav, const double* dualvol, int dim1,
                         const double* phi, double* grad);

/* Edge mass flow rate 0.5*(rhoL uL + rhoR uR).A
 * (unit_tests/kernels/UnitTestKernelUtils.C:725-769). */
void orc_mdot_edge(int64_t nedge, const int32_t* edges, const double* areav,
                   const double* rho, const double* vel, double* mdot);
/* The balanced buoyancy force option (use_balanced_buoyancy_force_) of
 * ContinuityEdgeSolverAlg (src/edge_kernels/ContinuityEdgeSolverAlg.C:
 * 58-88,147-152,168-173) and MdotEdgeAlg (src/ngp_algorithms/MdotEdgeAlg.C:
 * 81-102,147-154,163-169): gravity vector, node field "buoyancy_source"
 * [n][3] and node field "buoyancy_source_mask" [n]. */
typedef struct {
  double gravity[3];
  const double* source;
  const double* mask;
} orc_buoyancy;

/* MdotEdgeAlg (src/ngp_algorithms/MdotEdgeAlg.C:121-186), interior edges,
 * without the mesh-motion (GCL) term; buoy NULL = balanced buoyancy off. */
void orc_mdot_edge_alg(int64_t nedge, const int32_t* edges,
                       const double* coords, const double* vel,
                       const double* dpdx, const double* rho, const double* p,
                       const double* udiag, const double* areav, double noc,
                       double interp, const orc_buoyancy* buoy, double* mdot);

/* ------------------------------------------------------------------ */
/* Peclet blending and limiter                                         */
/* ------------------------------------------------------------------ */

/* ClassicPecletFunction::execute, src/PecletFunction.C:41-45 */
double orc_peclet_classic(double hf, double pe);
/* TanhFunction::execute, src/PecletFunction.C:68-71 */
double orc_peclet_tanh(double c1, double c2, double x);
/* van_leer, include/edge_kernels/EdgeKernelUtils.h:18-24 */
double orc_van_leer(double dqm, double dqp, double eps);

/* Peclet function selector (EquationSystem::ngp_create_peclet_function,
 * include/EquationSystem.h:401-417): form 0 = classic(hf), 1 = tanh(c1,c2). */
typedef struct {
  int form;
  double hybrid; /* hf of the classic form */
  double c1;     /* tanh transition */
  double c2;     /* tanh width */
} orc_peclet;

double orc_peclet_eval(const orc_peclet* p, double pe);

/* MomentumEdgePecletAlg::execute
 * (src/edge_kernels/MomentumEdgePecletAlg.C:49-102).  Writes per-edge
 * Peclet number and factor. */
void orc_momentum_edge_peclet(int64_t nedge, const int32_t* edges,
                              const double* coords, const double* vel,
                              const double* rho, const double* visc,
                              const orc_peclet* pf, double* pecnum,
                              double* pecfac);

/* ------------------------------------------------------------------ */
/* Edge solver algorithms: per-edge blocks                             */
/* ------------------------------------------------------------------ */

typedef struct {
  double alpha, alpha_upw, upw, relax;
  int use_limiter;
  orc_peclet peclet;
} orc_scalar_opts;

/* One edge of ScalarEdgeSolverAlg::execute
 * (src/edge_kernels/ScalarEdgeSolverAlg.C:56-206).  lhs is [2][2]
 * row-major, rhs is [2]; both are overwritten. */
void orc_scalar_edge(const orc_scalar_opts* o, int64_t e, int32_t nL,
                     int32_t nR, const double* coords, const double* vrtm,
                     const double* q, const double* dqdx, const double* rho,
                     const double* dflux, const double* areav,
                     const double* mdot, double lhs[4], double rhs[2]);

typedef struct {
  double alpha, alpha_upw, upw, relax, include_divU;
  int use_limiter;
} orc_momentum_opts;

/* One edge of MomentumEdgeSolverAlg::execute
 * (src/edge_kernels/MomentumEdgeSolverAlg.C:71-313), no VOF.
 * lhs is [6][6] row-major (row/col = node*3 + component), rhs [6]. */
void orc_momentum_edge(const orc_momentum_opts* o, int64_t e, int32_t nL,
                       int32_t nR, const double* coords, const double* vel,
                       const double* dudx, const double* visc,
                       const double* rho, const double* areav,
                       const double* mdot, const double* pecfac,
                       const double* wallmask, double lhs[36], double rhs[6]);

typedef struct {
  double dt, gamma1;
  double noc;            /* 1 if non-orthogonal correction is used */
  double interp_together;/* mdot interp rho*U together (1) or not (0) */
  double incompressible; /* solve incompressible continuity (1) or not */
  const orc_buoyancy* buoy; /* balanced buoyancy force, NULL = off */
} orc_continuity_opts;

/* One edge of ContinuityEdgeSolverAlg::execute
 * (src/edge_kernels/ContinuityEdgeSolverAlg.C:38-195), with the balanced
 * buoyancy forcing when o->buoy is set; no GCL term. lhs [2][2], rhs [2]. */
void orc_continuity_edge(const orc_continuity_opts* o, int64_t e,
                         int32_t nL, int32_t nR, const double* coords,
                         const double* vel, const double* dpdx,
                         const double* rho, const double* p,
                         const double* udiag, const double* areav,
                         double lhs[4], double rhs[2]);

/* ------------------------------------------------------------------ */
/* Linear system: CSR graph, sum_into scatter, row resets              */
/* ------------------------------------------------------------------ */

/* Point CSR in the Tpetra layout: row r = node*ndof + d; every row of a
 * node holds the columns of each connected node as ndof consecutive
 * entries, nodes in ascending order (self included). */
typedef struct {
  int64_t nnode;
  int ndof;
  int64_t nrows;
  int64_t nnz;
  int64_t* row_ptr; /* nrows+1 */
  int32_t* cols;    /* nnz */
  double* vals;     /* nnz */
  double* rhs;      /* nrows */
  /* hypre backend only (else NULL): [nnode] skipped-row flags of the
   * Dirichlet nodes (HypreLinearSystem::buildDirichletNodeGraph) and the
   * applier's checkSkippedRows_ flag */
  const unsigned char* skip;
  int skip_check;
} orc_csr;

/* HypreLinearSystem::zeroSystem + resetCoeffApplierData
 * (src/HypreLinearSystem.C:1892-1925, 1386-1392): values and rhs to zero,
 * the skipped-row check back on. */
void orc_csr_zero(orc_csr* A);

/* HypreLinearSystem::applyDirichletBCs (src/HypreLinearSystem.C:2408-2456)
 * on skipped rows, whose hypre graph row holds only the diagonal: the
 * diagonal becomes 1, rhs = bc - sol; the row's other entries (zero: the
 * scatter skipped them) are left. */
void orc_csr_apply_dirichlet_hypre(orc_csr* A, int64_t nnodes,
                                   const int32_t* nodes, const double* sol,
                                   const double* bc, int ncomp);

/* Graph of self + edge neighbours, sorted and expanded to ndof
 * (TpetraLinearSystem::buildEdgeToNodeGraph / addConnections /
 * finalizeLinearSystem, src/TpetraLinearSystem.C:420-440, 489-494,
 * 1137-1272).  Allocates the arrays; values and rhs are zero. */
int orc_csr_build_edge_graph(int64_t nnode, int64_t nedge,
                             const int32_t* edges, int ndof, orc_csr* A);
void orc_csr_free(orc_csr* A);

/* sum_into / sum_into_row / sum_into_row_vec_3
 * (src/TpetraLinearSystem.C:1293-1459): sort the entity column ids,
 * walk each sorted row once, add; entries with no graph column are
 * dropped.  lhs is [(n*ndof)][(n*ndof)] row-major, rhs [n*ndof].  With
 * skipped rows set and their check on, the rows of skipped nodes are left
 * alone (HypreLinSysCoeffApplier::sum_into_1DoF,
 * src/HypreLinearSystem.C:2190-2195). */
void orc_csr_sum_into(orc_csr* A, int nent, const int32_t* nodes,
                      const double* lhs, const double* rhs);

/* reset_rows (src/TpetraLinearSystem.C:1461-1520): row -> diag_value on
 * the diagonal, zero elsewhere; rhs -> rhs_residual.  Lifts the
 * skipped-row check (HypreLinSysCoeffApplier::resetRows,
 * src/HypreLinearSystem.C:2326). */
void orc_csr_reset_rows(orc_csr* A, int64_t nnodes, const int32_t* nodes,
                        int begin, int end, double diag_value,
                        double rhs_residual);

/* applyDirichletBCs for owned rows (src/TpetraLinearSystem.C:1686-1757):
 * row -> identity, rhs = bc - sol.  sol/bc are [N][ncomp]. */
void orc_csr_apply_dirichlet(orc_csr* A, int64_t nnodes,
                             const int32_t* nodes, int begin, int end,
                             const double* sol, const double* bc, int ncomp);

/* FixPressureAtNodeAlgorithm::execute (src/FixPressureAtNodeAlgorithm.C:
 * 57-124): reset the node's row(s), then sum lhs = 1, rhs = pref - p. */
void orc_csr_fix_node(orc_csr* A, int32_t node, double pref, double pnode);

/* Dense capture used by the reference unit-test mock coefficient applier
 * (unit_tests/UnitTestLinearSystem.h:42-72): only same-component
 * couplings (d,d) are accumulated.  lhs_dense is [N*ndof][N*ndof]. */
void orc_dense_sum_into(int64_t nnode, int ndof, double* lhs_dense,
                        double* rhs_dense, int nent, const int32_t* nodes,
                        const double* lhs, const double* rhs);

/* ------------------------------------------------------------------ */
/* Assembly drivers: loop the edges in order, apply each block          */
/* ------------------------------------------------------------------ */

/* target: A != NULL -> CSR sum_into; else dense mock (lhs_dense/rhs_dense). */
void orc_assemble_scalar(const orc_scalar_opts* o, int64_t nnode,
                         int64_t nedge, const int32_t* edges,
                         const double* coords, const double* vrtm,
                         const double* q, const double* dqdx,
                         const double* rho, const double* dflux,
                         const double* areav, const double* mdot, orc_csr* A,
                         double* lhs_dense, double* rhs_dense);

void orc_assemble_momentum(const orc_momentum_opts* o, int64_t nnode,
                           int64_t nedge, const int32_t* edges,
                           const double* coords, const double* vel,
                           const double* dudx, const double* visc,
                           const double* rho, const double* areav,
                           const double* mdot, const double* pecfac,
                           const double* wallmask, orc_csr* A,
                           double* lhs_dense, double* rhs_dense);

/* Threads used by the orc_assemble_* loops over a CSR (default 1).  With
 * more than one, edges are processed in parallel and the scatter adds are
 * atomic (the reference's OpenMP execution); the sum order then varies. */
void orc_set_threads(int n);

void orc_assemble_continuity(const orc_continuity_opts* o, int64_t nnode,
                             int64_t nedge, const int32_t* edges,
                             const double* coords, const double* vel,
                             const double* dpdx, const double* rho,
                             const double* p, const double* udiag,
                             const double* areav, orc_csr* A,
                             double* lhs_dense, double* rhs_dense);

/* ------------------------------------------------------------------ */
/* Krylov solve                                                         */
/* ------------------------------------------------------------------ */

/* y = A x (point CSR) */
void orc_spmv(const orc_csr* A, const double* x, double* y);

enum { ORC_PREC_NONE = 0, ORC_PREC_JACOBI = 1, ORC_PREC_CHEBYSHEV = 2 };

typedef struct {
  int kspace;      /* restart length (Belos "Num Blocks") */
  int max_iters;   /* total iterations cap */
  double tol;      /* on ||b - A x||_2 / ||b||_2 with x0 given */
  int precond;     /* ORC_PREC_* (applied on the right) */
  int cheb_degree; /* degree of the Jacobi-scaled Chebyshev polynomial */
  double cheb_ratio;/* lambda_max / lambda_min of the Chebyshev interval */
} orc_gmres_opts;

typedef struct {
  int iters;
  double implicit_rel; /* last Arnoldi residual estimate / ||r0|| */
  double final_abs;    /* explicit ||A x - b||_2 */
  double bnorm;
  int converged;
} orc_gmres_stats;

/* Restarted right-preconditioned GMRES(m) with classical Gram-Schmidt and
 * one full reorthogonalisation (Belos ICGS with x0 given and right
 * preconditioning, settings of src/LinearSolverConfig.C:37-140 and
 * src/LinearSolver.C:209-254).  x holds x0 on entry. */
int orc_gmres(const orc_csr* A, const double* b, double* x,
              const orc_gmres_opts* o, orc_gmres_stats* st);

/* Dense LU with partial pivoting: solves A x = b in place (A n x n
 * row-major is destroyed; b becomes x).  Returns 0 or -1 if singular. */
int orc_dense_solve(int n, double* A, double* b);

/* Largest eigenvalue estimate of D^-1 A by power iteration. */
double orc_jacobi_lambda_max(const orc_csr* A, int iters, uint64_t seed);

/* ------------------------------------------------------------------ */
/* Learned initial guess: message-passing network (orc_mpnn.c)         */
/* ------------------------------------------------------------------ */

/* H*5 + H + L*(2H^2 + H) + 2H + 2 */
int64_t orc_mpnn_nparams(int H, int L);
/* w[nnz]: sqrt(|a_nm a_mn|)/sqrt(|a_nn a_mm|) off the diagonal, 0 on it */
void orc_mpnn_slot_weights(const orc_csr* A, double* w);
/* xb: the base guess, polynomial extrapolation in time of the newest
 * min(nh, 3) of the history hist[0..nh-1] (hist[0] the newest). */
void orc_mpnn_base_guess(int64_t n, const double* const* hist, int nh,
                         double* xb);
/* feat[N][5]; scales = {s_b, s_x, s_r, s_e, s_o} (s_o: output scale);
 * hist[0..nh-1] the guesses of the latest solves, newest first, nh 1..5 */
void orc_mpnn_features(const orc_csr* A, const double* b,
                       const double* const* hist, int nh, double* feat,
                       double* scales);
void orc_mpnn_forward(const double* P, int H, int L, const orc_csr* A,
                      const double* w, const double* feat, double* delta,
                      double* acts);
/* The same with bfloat16 activations, messages and layer weights, as the
 * NW_MPNN_BF16 prediction computes them. */
void orc_mpnn_forward_bf16(const double* P, int H, int L, const orc_csr* A,
                           const double* w, const double* feat, double* delta);
double orc_mpnn_loss_grad(const double* P, int H, int L, const orc_csr* A,
                          const double* w, const double* feat,
                          const double* target, double* grad);
/* Residual loss: mean over rows of (D^-1 A u)^2 with u = delta - t on the
 * rows with couplings (0 on Dirichlet rows); grad receives dLoss/dP. */
double orc_mpnn_loss_grad_residual(const double* P, int H, int L,
                                   const orc_csr* A, const double* w,
                                   const double* feat, const double* target,
                                   double* grad);

#ifdef __cplusplus
}
#endif

#endif
