/*
 * nalu_amd.h -- C ABI of the MI355X edge-assembly + pressure-solve library.
 *
 * One context = one GPU, one HIP stream, one mesh (edge graph), its node and
 * edge fields, one linear system and its solver workspace.  Every call is
 * stream-ordered on the context's stream, returns 0 on success or a negative
 * NW_E_* code, and never throws; nw_last_error() gives the message of the
 * last failure on the calling thread.  Arrays are plain pointers with sizes;
 * "host" arrays are copied, device arrays are the caller's to keep alive.
 *
 * Each entry point names the reference interface it replaces
 * (Exawind/nalu-wind, paths relative to its source root).
 */
#ifndef NALU_AMD_H
#define NALU_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nw_ctx nw_ctx;

/* ---- status codes ------------------------------------------------------ */
enum {
  NW_OK = 0,
  NW_E_ARG = -1,       /* bad argument / size */
  NW_E_STATE = -2,     /* call out of order (no mesh, no graph, ...) */
  NW_E_HIP = -3,       /* HIP runtime error */
  NW_E_NOGRAPHCOL = -4,/* sumInto column absent from the graph */
  NW_E_NONFINITE = -5, /* NaN/Inf found where the call checks */
  NW_E_NOCONV = -6,    /* solver did not reach the tolerance */
  NW_E_ALLOC = -7      /* device allocation failed */
};

/* ---- node and edge fields ------------------------------------------------
 * Upload/download arrays are [n][ncomp] doubles (components of a node
 * adjacent, as in STK field data), n = nnode for node fields and nedge for
 * edge fields.  The device copy has the same layout, so a kernel gathering
 * a neighbour's vector or gradient reads one contiguous record, and lanes on
 * consecutive nodes or edges still read consecutive addresses;
 * nw_field_device_ptr exposes that layout.  Field names follow the
 * reference's registered field names. */
enum {
  NW_F_COORDINATES = 0, /* 3  "coordinates" */
  NW_F_VELOCITY,        /* 3  "velocity" (== velocity_rtm on static mesh) */
  NW_F_DENSITY,         /* 1  "density" (state NP1) */
  NW_F_PRESSURE,        /* 1  "pressure" */
  NW_F_DPDX,            /* 3  "dpdx" */
  NW_F_UDIAG,           /* 1  "momentum_diag" */
  NW_F_VISCOSITY,       /* 1  "viscosity" / "effective_viscosity_u" */
  NW_F_DUDX,            /* 9  "dudx" */
  NW_F_WALL_MASK,       /* 1  "abl_wall_no_slip_wall_func_node_mask" */
  NW_F_SCALAR_Q,        /* 1  transported scalar (e.g. "mixture_fraction") */
  NW_F_DQDX,            /* 3  its gradient */
  NW_F_DIFF_FLUX_COEFF, /* 1  scalar diffusion coefficient */
  NW_F_DUAL_VOLUME,     /* 1  "dual_nodal_volume" */
  NW_F_BC_VALUE,        /* 3  Dirichlet values ("*_bc") */
  NW_E_AREA_VECTOR,     /* 3  "edge_area_vector" (edge field) */
  NW_E_MASS_FLOW_RATE,  /* 1  "mass_flow_rate" (edge field) */
  NW_E_PECLET_FACTOR,   /* 1  "peclet_factor" (edge field) */
  NW_E_PECLET_NUMBER,   /* 1  "peclet_number" (edge field) */
  NW_F_PRESSURE_BC,     /* 1  "pressure_bc" (Dirichlet values of pressure) */
  NW_F_UTMP,            /* 3  "uTmp" (the dp/dx saved by the projection) */
  NW_F_BUOYANCY_SOURCE, /* 3  "buoyancy_source" (balanced buoyancy force) */
  NW_F_BUOYANCY_SOURCE_MASK, /* 1  "buoyancy_source_mask" */
  NW_NUM_FIELDS
};

/* ---- scheme options (SolutionOptions values the edge algs read) -------- */
typedef struct {
  int form;      /* 0 classic (hybrid factor), 1 tanh */
  double hybrid; /* classic: hf (default 0) */
  double c1;     /* tanh transition (default 2) */
  double c2;     /* tanh width (default 4) */
} nw_peclet;

typedef struct {
  double alpha;     /* default 0 */
  double alpha_upw; /* default 1 */
  double upw;       /* default 1 */
  double relax;     /* relaxation factor, default 1 */
  int use_limiter;  /* van Leer limiter on the upwind extrapolation */
  nw_peclet peclet;
  int accumulate;   /* 0: rows overwritten (fused zeroSystem), 1: sumInto */
} nw_scalar_opts;

typedef struct {
  double alpha, alpha_upw, upw, relax;
  double include_divU; /* default 0 */
  int use_limiter;
  int accumulate;
  /* NGPApplyCoeff::extract_diagonal (src/SolverAlgorithm.C:89-105), on
   * when tscaleType_ == TSCALE_UDIAGINV (src/LowMachEquationSystem.C:2736):
   * the (0, 0) entry of each node's summed (node, node) edge block, x-velocity
   * row, is added to NW_F_UDIAG ("momentum_diag"), one add per node */
  int extract_diagonal;
} nw_momentum_opts;

typedef struct {
  double dt, gamma1;
  double noc;             /* non-orthogonal correction on (1) / off (0) */
  double interp_together; /* mdot interpolates rho*U together (default 1) */
  double incompressible;  /* incompressible continuity (default 0) */
  int accumulate;
  /* use_balanced_buoyancy_force_ (ContinuityEdgeSolverAlg.C:58-88,147-152,
   * 168-173): tmdot += tau g.A rho_ip (m_L + m_R)/2 and
   * G_ip -= (m_R s_R/Udiag_R + m_L s_L/Udiag_L)/2, with s the node field
   * NW_F_BUOYANCY_SOURCE and m NW_F_BUOYANCY_SOURCE_MASK; 0 = off */
  int balanced_buoyancy;
  double gravity[3];
} nw_continuity_opts;

/* ---- lifecycle ----------------------------------------------------------- */

/* Context on `device` with its own HIP stream. */
int nw_ctx_create(int device, nw_ctx** ctx);
int nw_ctx_destroy(nw_ctx* ctx);
/* The context's hipStream_t, for callers that time or order work on it. */
int nw_ctx_stream(nw_ctx* ctx, void** stream);
int nw_ctx_synchronize(nw_ctx* ctx);
/* Input check of the edge assemblies (the reference checks nothing; its
 * device asserts, STK_NGP_ThrowRequireMsg, catch only structural errors):
 * with `on`, every nw_assemble_* call scans the assembled values and rhs
 * afterwards and returns NW_E_NONFINITE, naming the first row, when an
 * entry is NaN or Inf (a NaN or Inf in any input field the kernel reads
 * reaches the rows of that node and its neighbours).  The scan reads the
 * system once more and synchronises the stream, so it is off by default. */
int nw_ctx_set_check_finite(nw_ctx* ctx, int on);
const char* nw_last_error(void);

/* ---- mesh ---------------------------------------------------------------
 * Replaces the STK edge/node connectivity the edge algorithms loop over
 * (stk::mesh::create_edges, src/Realm.C:2304-2321; edge node[0] = L,
 * node[1] = R as in include/AssembleEdgeSolverAlgorithm.h:86-96).
 * edge_nodes is host [nedge][2] with 0 <= node < nnode, L != R, no
 * duplicate edges. */
int nw_mesh_set(nw_ctx* ctx, int64_t nnode, int64_t nedge,
                const int32_t* edge_nodes);

/* ---- fields (STK NgpField sync_to_device / sync_to_host) ----------------- */
int nw_field_upload(nw_ctx* ctx, int field, const double* host, int64_t count);
int nw_field_download(nw_ctx* ctx, int field, double* host, int64_t count);
int nw_field_fill(nw_ctx* ctx, int field, double value);
/* Device pointer of the field (allocated on first use), [n][ncomp] like
 * the host arrays. */
int nw_field_device_ptr(nw_ctx* ctx, int field, double** dptr);
/* The coordinates were written through their device pointer (a moving
 * mesh): the row kernels' locality order is rebuilt from them on the next
 * call (an upload of NW_F_COORDINATES does this by itself).  Results never
 * depend on that order, only the speed of the row kernels does.  No
 * counterpart in the reference, whose edge loops keep no locality order. */
int nw_mesh_coordinates_changed(nw_ctx* ctx);
int nw_field_components(int field);

/* ---- linear system (LinearSystem / TpetraLinearSystem) -------------------
 * nw_linsys_build_edge_graph: buildEdgeToNodeGraph + finalizeLinearSystem
 *   (src/TpetraLinearSystem.C:489-494, 420-440, 1137-1272): rows = nodes,
 *   columns = self + edge neighbours, sorted; ndof 1 (scalar/continuity)
 *   or 3 (momentum, 3x3 blocks). */
int nw_linsys_build_edge_graph(nw_ctx* ctx, int ndof);
/* zeroSystem (src/TpetraLinearSystem.C:1274-1291): values, rhs, solution. */
int nw_linsys_zero(nw_ctx* ctx);
/* loadComplete (src/TpetraLinearSystem.C:1788-1809): fillComplete + export
 * of shared rows; on one GPU every row is owned, so this only orders. */
int nw_linsys_load_complete(nw_ctx* ctx);
/* Sizes of the point-CSR view (Tpetra layout: nrows = nnode*ndof). */
int nw_linsys_sizes(nw_ctx* ctx, int64_t* nrows, int64_t* nnz);
/* Download in the Tpetra point-CSR layout: row r = node*ndof + d, the
 * columns of each connected node as ndof consecutive entries, sorted.
 * Any pointer may be NULL. row_ptr has nrows+1 entries. */
int nw_linsys_download(nw_ctx* ctx, int64_t* row_ptr, int32_t* cols,
                       double* vals, double* rhs);
/* Device pointers of the block-CSR arrays: block row_ptr [nnode+1] and
 * cols [nnzb] (int32), values [nnzb][ndof][ndof] and rhs [nnode][ndof]. */
int nw_linsys_device_arrays(nw_ctx* ctx, int32_t** row_ptr, int32_t** cols,
                            double** vals, double** rhs);

/* CoeffApplier::operator() / LinearSystem::sumInto
 * (include/LinearSystem.h:53-70, src/TpetraLinearSystem.C:1381-1459):
 * nblocks blocks of nent nodes each, lhs [nblocks][(nent*ndof)^2] row-major
 * (row/col = node_slot*ndof + dof), rhs [nblocks][nent*ndof]; additive, in
 * block order (deterministic).  A column absent from the graph is an error
 * (NW_E_NOGRAPHCOL) and nothing of that call is applied. Host arrays. */
int nw_linsys_sum_into(nw_ctx* ctx, int nblocks, int nent,
                       const int32_t* nodes, const double* lhs,
                       const double* rhs);
/* The same scatter on device arrays (the device-resident CoeffApplier,
 * include/LinearSystem.h:53-70, src/TpetraLinearSystem.C:1381-1459, as the
 * reference's device algorithms call it per entity): d_nodes
 * [nblocks][nent], d_lhs [nblocks][(nent*ndof)^2], d_rhs [nblocks][nent*ndof]
 * on the device, stream-ordered on the context's stream.  Deterministic:
 * the contributions to each entry are summed in block order and the result
 * equals nw_linsys_sum_into's bitwise.  A column absent from the graph
 * (NW_E_NOGRAPHCOL) or a node id out of range (NW_E_ARG) is found on the
 * device before anything is applied, and nothing of that call is applied
 * (the call synchronises the stream once to learn this).  Its scratch is a
 * context buffer that is allocated only when a call needs more than every
 * call before it; nent <= 64. */
int nw_linsys_sum_into_device(nw_ctx* ctx, int nblocks, int nent,
                              const int32_t* d_nodes, const double* d_lhs,
                              const double* d_rhs);
/* No counterpart in the reference: nw_linsys_sum_into_device for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and result (bitwise); its scratch is a fixed 32 MiB device array per GPU
 * (held until the process ends), so it allocates nothing, and a call that
 * needs more (about 900 thousand scalar contributions, nblocks * nent *
 * ndof * (nent * ndof + 1)) fails with NW_E_ARG.  All calls on one GPU share
 * that array: calls that can overlap on one GPU (contexts on different
 * streams) race and must be ordered by the caller. */
int nw_linsys_sum_into_device_noalloc(nw_ctx* ctx, int nblocks, int nent,
                                      const int32_t* d_nodes,
                                      const double* d_lhs, const double* d_rhs);
/* CoeffApplier::resetRows / LinearSystem::resetRows
 * (src/TpetraLinearSystem.C:1461-1520, 1759-1786). Host node list. */
int nw_linsys_reset_rows(nw_ctx* ctx, int64_t nnodes, const int32_t* nodes,
                         int begin, int end, double diag_value,
                         double rhs_residual);
/* HypreLinearSystem::buildDirichletNodeGraph (src/HypreLinearSystem.C:
 * 720-805) with the skip semantics of its coefficient applier
 * (HypreLinSysCoeffApplier::sum_into_1DoF, src/HypreLinearSystem.C:
 * 2165-2239): the listed nodes become skipped rows.  While the skip check
 * is on -- after nw_linsys_zero, until a nw_linsys_reset_rows call lifts it
 * (the applier's checkSkippedRows_ flag, :2326, :1392) -- the edge
 * assemblies and nw_linsys_sum_into add nothing to a skipped row: it keeps
 * zero entries until nw_linsys_apply_dirichlet writes its diagonal 1 and
 * rhs = bc - solution (HypreLinearSystem::applyDirichletBCs, :2408-2456).
 * Columns of skipped nodes in other rows are assembled as usual.  0 nodes
 * clears the list; a graph rebuild clears it too. Host node list. */
int nw_linsys_set_skipped_rows(nw_ctx* ctx, int64_t nnodes,
                               const int32_t* nodes);
/* applyDirichletBCs (src/TpetraLinearSystem.C:1686-1757): rows -> identity,
 * rhs = bc_field - sol_field for dofs [begin, end). Host node list; the two
 * node fields have the same component count. */
int nw_linsys_apply_dirichlet(nw_ctx* ctx, int64_t nnodes,
                              const int32_t* nodes, int sol_field,
                              int bc_field, int begin, int end);
/* FixPressureAtNodeAlgorithm::execute (src/FixPressureAtNodeAlgorithm.C:
 * 57-124): reset the node's row, then sumInto lhs = 1, rhs = pref - p. */
int nw_fix_pressure_at_node(nw_ctx* ctx, int32_t node, double pref);

/* ---- edge solver algorithms ---------------------------------------------
 * Each replaces one AssembleEdgeSolverAlgorithm subclass execute(): all
 * edges in one launch, deterministic (no atomics), rows written whole. */
/* ScalarEdgeSolverAlg::execute (src/edge_kernels/ScalarEdgeSolverAlg.C:
 * 56-206): reads NW_F_COORDINATES, VELOCITY, SCALAR_Q, DQDX, DENSITY,
 * DIFF_FLUX_COEFF and NW_E_AREA_VECTOR, NW_E_MASS_FLOW_RATE. */
int nw_assemble_scalar(nw_ctx* ctx, const nw_scalar_opts* opts);
/* MomentumEdgeSolverAlg::execute (src/edge_kernels/MomentumEdgeSolverAlg.C:
 * 71-313): reads COORDINATES, VELOCITY, DUDX, VISCOSITY, WALL_MASK and
 * AREA_VECTOR, MASS_FLOW_RATE, PECLET_FACTOR; needs an ndof = 3 graph. */
int nw_assemble_momentum(nw_ctx* ctx, const nw_momentum_opts* opts);
/* ContinuityEdgeSolverAlg::execute (src/edge_kernels/ContinuityEdgeSolverAlg.C:
 * 38-195): reads COORDINATES, VELOCITY, DPDX, DENSITY, PRESSURE, UDIAG and
 * AREA_VECTOR. */
int nw_assemble_continuity(nw_ctx* ctx, const nw_continuity_opts* opts);
/* MomentumEdgePecletAlg::execute (src/edge_kernels/MomentumEdgePecletAlg.C:
 * 49-102): writes NW_E_PECLET_NUMBER and NW_E_PECLET_FACTOR. */
int nw_momentum_edge_peclet(nw_ctx* ctx, const nw_peclet* peclet);
/* MomentumEdgePecletAlg::execute (src/edge_kernels/MomentumEdgePecletAlg.C:
 * 49-102) and MomentumEdgeSolverAlg::execute (src/edge_kernels/
 * MomentumEdgeSolverAlg.C:71-313) run as one pass: the Peclet factor of
 * each edge is computed where the momentum assembly evaluates the edge
 * (same operations as nw_momentum_edge_peclet, which the solver algorithm
 * would otherwise read back), and NW_E_PECLET_NUMBER / NW_E_PECLET_FACTOR
 * are still written.  Also reads NW_F_DENSITY. */
int nw_assemble_momentum_with_peclet(nw_ctx* ctx, const nw_momentum_opts* opts,
                                     const nw_peclet* peclet);

/* ---- edge algorithms around the pressure solve ----------------------------
 * MdotEdgeAlg::execute (src/ngp_algorithms/MdotEdgeAlg.C:43-194), interior
 * edges: the mass flow rate from the new pressure,
 *   mdot = -tau (p_R - p_L) |A|^2/(A.dx) + sum_d [(rho u)_ip + G_ip] A_d
 *          - noc * sum_d k_d G_ip,   tau = (1/Udiag_L + 1/Udiag_R)/2,
 *   G_ip = (dpdx_R/Udiag_R + dpdx_L/Udiag_L)/2, k = A - |A|^2/(A.dx) dx,
 * with (rho u)_ip interpolated as rho*U together (interp_together = 1) or
 * as rho_ip u_ip.  Reads COORDINATES, VELOCITY, DPDX, DENSITY, PRESSURE,
 * UDIAG and AREA_VECTOR; writes NW_E_MASS_FLOW_RATE (one value per edge,
 * one launch over the edges, no graph needed).  With balanced_buoyancy set
 * it also reads NW_F_BUOYANCY_SOURCE and NW_F_BUOYANCY_SOURCE_MASK.  The
 * mesh-motion (GCL) term of the reference is not built (static meshes). */
typedef struct {
  double noc;             /* non-orthogonal correction on (1) / off (0) */
  double interp_together; /* 1: (rho u)_ip = avg(rho u); 0: avg(rho) avg(u) */
  /* the balanced buoyancy force (MdotEdgeAlg.C:81-102,147-154,163-169), as
   * in nw_continuity_opts; 0 = off */
  int balanced_buoyancy;
  double gravity[3];
} nw_mdot_opts;
int nw_compute_mdot(nw_ctx* ctx, const nw_mdot_opts* opts);
/* NodalGradEdgeAlg::execute (src/ngp_algorithms/NodalGradEdgeAlg.C:58-111),
 * interior edges: grad_n = sum over the edges of n of +-q_ip A / V_n
 * (+ for n = L, - for n = R), q_ip = (q_L + q_R)/2, V the dual volume
 * (NW_F_DUAL_VOLUME).  field_in with c components (PRESSURE, VELOCITY,
 * SCALAR_Q), field_out with 3c (DPDX, DUDX, DQDX), component 3 j + i =
 * d q_j / d x_i; overwritten (accumulate = 0) or added to (accumulate = 1:
 * the reference's driver zeroes the field and every interior and boundary
 * algorithm adds its part, so an edge gradient that runs after boundary
 * contributions must accumulate).  One pass over the rows of the context's
 * edge graph (built by nw_linsys_build_edge_graph): each node's sum is
 * formed by one lane in slot order, so the result is bitwise reproducible.
 * Boundary-face contributions (NodalGradBndryElemAlg) are not built. */
int nw_nodal_grad(nw_ctx* ctx, int field_in, int field_out, int accumulate);

/* ---- the projection step around the pressure solve ----------------------
 * EquationSystem::solution_update (src/EquationSystem.C:556-573):
 *   field = delta_frac * delta + field_frac * field
 * over all nodes (or edges) and components of `field`; delta is a field of
 * the same shape, or NW_LINSYS_SOLUTION for the context's linear-system
 * solution (the reference's pTmp / uTmp increment fields), which needs a
 * system of as many dofs as `field` has components.  The pressure update
 * after the continuity solve (src/LowMachEquationSystem.C:800-804) is
 * nw_solution_update(ctx, 1, NW_LINSYS_SOLUTION, 1, NW_F_PRESSURE). */
enum { NW_LINSYS_SOLUTION = -1 };
int nw_solution_update(nw_ctx* ctx, double delta_frac, int delta,
                       double field_frac, int field);
/* LowMachEquationSystem::project_nodal_velocity (src/LowMachEquationSystem.C:
 * 882-939): saves dp/dx into NW_F_UTMP, recomputes NW_F_DPDX from
 * NW_F_PRESSURE by the edge nodal gradient (nw_nodal_grad, which needs the
 * edge graph and NW_F_DUAL_VOLUME), then projects the velocity at every
 * node: u -= (dp/dx_new - dp/dx_old) / (rho * Udiag) with NW_F_DENSITY and
 * NW_F_UDIAG.  (The reference's not-projected parts and directions are
 * boundary options, not built.) */
int nw_project_nodal_velocity(nw_ctx* ctx);

/* ---- geometry (GeometryInteriorAlg, src/ngp_algorithms/
 * GeometryInteriorAlg.C:70-225) ---------------------------------------------
 * nw_mesh_elements: the mesh's elements over the nodes and edges of
 *   nw_mesh_set: type [nelem] (0 Hex8, 1 Tet4, 2 Wed6, 3 Pyr5; NULL = all
 *   Hex8), conn [nelem][8] in exodus node order (unused slots -1).  Builds, on the
 *   host, the lists of element sub-control surfaces per edge and
 *   sub-control volumes per node (element order) that the device sums
 *   through.  NW_E_ARG when a node is out of range, an element edge is not
 *   in the edge list, or nelem > 2^26.  Once per mesh; it also allocates
 *   the per-element scratch of nw_compute_geometry (44 f64 per element).
 * nw_compute_geometry: impl_compute_edge_area_vector and
 *   impl_compute_dual_nodal_volume on the device from NW_F_COORDINATES:
 *   writes NW_E_AREA_VECTOR (sum of the element sub-control-surface area
 *   vectors of each edge, Hex/Tet/Wed/PyrSCS::determinant_scs, oriented
 *   from node[0] to node[1]) and NW_F_DUAL_VOLUME (sum of the sub-control
 *   volumes of each node, Hex/Tet/Wed/PyrSCV::determinant_scv).  Sums run
 *   in element order (the reference adds with atomics), so the result is
 *   reproducible.  Call again after the coordinates move: stream-ordered,
 *   allocates nothing. */
int nw_mesh_elements(nw_ctx* ctx, int64_t nelem, const int32_t* type,
                     const int32_t* conn);
int nw_compute_geometry(nw_ctx* ctx);

/* ---- linear solve (TpetraLinearSystem::solve -> Belos GMRES) ------------ */
enum {
  NW_PREC_NONE = 0,
  NW_PREC_JACOBI = 1,    /* Ifpack2 RELAXATION Jacobi, 1 sweep */
  NW_PREC_CHEBYSHEV = 2, /* Jacobi-scaled Chebyshev polynomial */
  NW_PREC_AMG = 3        /* smoothed-aggregation AMG V-cycle (the role of
                            MueLu with reg_tests/xml/milestone.xml) */
};
enum {
  NW_SOLVE_FP64 = 0,       /* plain FP64 GMRES(m) */
  NW_SOLVE_FP32_REFINE = 1 /* FP32 inner GMRES + FP64 iterative refinement */
};
enum {
  NW_X0_ZERO = 0,     /* x0 = 0, the reference's choice */
  NW_X0_SOLUTION = 1, /* current contents of the solution vector */
  NW_X0_LEARNED = 2   /* message-passing network prediction (nw_mpnn_*) */
};

typedef struct {
  int kspace;         /* restart length ("Num Blocks", default 50) */
  int max_iters;      /* total Krylov iterations cap */
  double tol;         /* ||b - A x||_2 / ||b||_2 target */
  int precond;        /* NW_PREC_* (right preconditioning) */
  int cheb_degree;    /* Chebyshev polynomial degree */
  double cheb_ratio;  /* lambda_max / lambda_min of the Chebyshev interval */
  int mode;           /* NW_SOLVE_* */
  double inner_tol;   /* FP32 inner relative tolerance per refinement step */
  int x0;             /* NW_X0_* */
  int reuse_preconditioner; /* AMG: keep an existing hierarchy (the
                               reference's reuse_preconditioner option) */
} nw_solver_cfg;

typedef struct {
  int iters;            /* Krylov iterations (all refinement steps) */
  int refinements;      /* restart cycles = FP64 refinement steps */
  int reorthogonalizations; /* second Gram-Schmidt passes taken */
  int converged;
  double final_rel;     /* explicit ||b - A x||_2 / ||b||_2 in FP64 */
  double final_abs;     /* explicit ||b - A x||_2 (the reference's report) */
  double bnorm;
  double ms;            /* device time of the solve on the context stream,
                           preconditioner setup excluded */
  double setup_ms;      /* preconditioner setup of this call, wall time (0
                           if reused): for AMG the matrix download, the host
                           build and the upload of the hierarchy, or the
                           FP32 copy made on a first FP32-inner solve */
  int precond_levels;   /* AMG levels (0 for other preconditioners) */
  int guess;            /* NW_X0_* of the guess GMRES started from (a learned
                           guess that did not lower the residual falls back
                           to NW_X0_SOLUTION) */
  double x0_rel;        /* ||b - A x0||_2 / ||b||_2 of that guess */
  double guess_ms;      /* wall time of the learned prediction (in ms) */
} nw_solve_stats;

/* Solve A x = rhs into the solution vector (the pressure increment). */
int nw_linsys_solve(nw_ctx* ctx, const nw_solver_cfg* cfg,
                    nw_solve_stats* stats);
int nw_linsys_solution_upload(nw_ctx* ctx, const double* host);
int nw_linsys_solution_download(nw_ctx* ctx, double* host);
int nw_linsys_solution_device_ptr(nw_ctx* ctx, double** dptr);
/* y = A x on device arrays of nrows doubles (FP64 residual checks). */
int nw_linsys_spmv(nw_ctx* ctx, const double* d_x, double* d_y);

/* Rows and stored entries of level `level` of the context's AMG hierarchy
 * (0 = the fine matrix; built by the first AMG-preconditioned solve), the
 * hierarchy summary MueLu prints at setup (its operator complexity is
 * sum(nnz) / nnz(level 0)).  NW_E_ARG past the last level or without a
 * hierarchy. */
int nw_linsys_amg_level(nw_ctx* ctx, int level, int64_t* rows, int64_t* nnz);

/* ------------------------------------------------------------------ */
/* Learned initial guess for the pressure solve                        */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference, whose pressure solve always starts
 * from x0 = 0 (src/LinearSolver.C:209-254, Belos setProblem with a zero
 * LHS).  A message-passing network on the graph of the context's 1-dof
 * system predicts x = xb + s * delta from the features of (A, b, x0, xp,
 * xpp), where x0 is the previous-step solution, xp and xpp the two before
 * it and xb = 2 x0 - xp the linear extrapolation in time (xb = x0 without
 * history); rows without off-diagonal entries get b / a_nn.  GMRES to the requested tolerance
 * still decides the answer; a prediction whose residual is not below that
 * of x0 is discarded (nw_solve_stats.guess reports which guess was used).
 * A network belongs to one context and must be destroyed before it. */
typedef struct nw_mpnn nw_mpnn;

typedef struct {
  int epochs;       /* passes over the recorded samples */
  int batch;        /* samples per Adam step */
  double lr;        /* Adam step size */
  double beta1, beta2, eps; /* Adam moment decay and denominator guard */
  uint64_t seed;    /* sample order shuffling */
  int holdout;      /* most recent samples kept for validation: after each
                       epoch the validation loss is measured and the
                       parameters with the lowest one (the initial ones
                       included) are kept (0: no validation) */
} nw_mpnn_train_cfg;

/* Parameter count of a network with the given hidden width and layers. */
int64_t nw_mpnn_nparams(int hidden, int layers);
/* hidden in {32, 64}, layers in [1, 8]; seeded He-uniform initialisation;
 * the output head starts at w_out = 0, b_out = (1, -1), so the untrained
 * prediction is the quadratic extrapolation 3 x0 - 3 xp + xpp. */
int nw_mpnn_create(nw_ctx* ctx, int hidden, int layers, uint64_t seed,
                   nw_mpnn** out);
void nw_mpnn_destroy(nw_mpnn* net);
/* Flat FP64 parameters: W_in[H][5], b_in[H], per layer W_self[H][H],
 * W_nbr[H][H], b[H], then w_out[2][H], b_out[2]: the network's two
 * outputs are per-node coefficients of the history differences
 * (x0 - xp)/s and (xp - xpp)/s added to xb. */
int nw_mpnn_set_params(nw_mpnn* net, const double* params, int64_t n);
int nw_mpnn_get_params(const nw_mpnn* net, double* params, int64_t n);
/* Records one training pair from the context's last solve: the features
 * of (A, b, x0, xp, xpp) with x0 the guess that solve started from and
 * xp, xpp the guesses of the two solves before it, and the target
 * (x - xb) / s with x the current solution vector.  The guess history
 * advances only on solves that start from the solution vector with a guess
 * different from the last one (a solve from x0 = 0, or a second method on
 * the same step from the same guess, leaves it unchanged). */
int nw_mpnn_record(nw_mpnn* net);
int nw_mpnn_samples(const nw_mpnn* net);
int nw_mpnn_clear_samples(nw_mpnn* net);
/* Training loss (nw_mpnn_loss_grad, nw_mpnn_train).  With u the error of
 * the prediction in units of the output scale, u = delta - t on the rows
 * with couplings (t = (x - xb)/s the recorded target):
 *   NW_MPNN_LOSS_RESIDUAL (default): mean((D^-1 A u)^2), the Jacobi-scaled
 *     residual of the prediction error -- what GMRES starts from; the
 *     context's current matrix is used (the pressure matrix of a static
 *     mesh does not change between steps);
 *   NW_MPNN_LOSS_SOLUTION: mean((delta - t)^2) over all rows. */
enum {
  NW_MPNN_LOSS_SOLUTION = 0,
  NW_MPNN_LOSS_RESIDUAL = 1
};
int nw_mpnn_set_loss(nw_mpnn* net, int kind);
/* Loss of recorded sample s and its gradient (nparams doubles, may be
 * NULL). */
int nw_mpnn_loss_grad(nw_mpnn* net, int sample, double* loss, double* grad);
/* Adam over the recorded samples; *last_loss = the best validation loss
 * (holdout > 0) or the mean training loss of the last epoch. */
int nw_mpnn_train(nw_mpnn* net, const nw_mpnn_train_cfg* cfg,
                  double* last_loss);
/* Prediction from the current system with x0 = the solution vector and
 * xp, xpp = the guesses the last two solves started from (x0 if none):
 * writes the prediction to x_host (nrows doubles) and the scale s to
 * *scale. */
int nw_mpnn_predict(nw_mpnn* net, double* x_host, double* scale);
/* Arithmetic of the prediction (training always runs in FP32 against the
 * FP64 master parameters).  NW_MPNN_BF16, the default: each hidden layer is
 * one fused kernel that gathers the messages, rounds its inputs [h, m] and
 * the layer weights to bfloat16 and multiplies them on the matrix cores
 * with FP32 accumulation; activations are stored in bfloat16 and the output
 * head runs in FP32 on the last layer's FP32 result.  NW_MPNN_FP32: the
 * FP32 training forward. */
enum {
  NW_MPNN_FP32 = 0,
  NW_MPNN_BF16 = 1
};
int nw_mpnn_set_precision(nw_mpnn* net, int precision);
/* Network used by solves with x0 = NW_X0_LEARNED (NULL detaches). */
int nw_linsys_set_learned_guess(nw_ctx* ctx, nw_mpnn* net);

#ifdef __cplusplus
}
#endif

#endif
