/*
 * nalu_amd_mesh.h -- host-side mesh and geometry producers for the edge
 * path: a structured Hex8 box generator (uniform / stretched / jittered),
 * edge extraction, CVFEM edge area vectors and dual volumes, the edge
 * nodal gradient and the edge mass flow rate.  These produce the inputs the
 * edge solver algorithms consume; they run multithreaded on the host.
 * All arrays are host arrays; return 0 or a negative NW_E_* code.  The
 * last entry generates a synthetic velocity field on the device.
 */
#ifndef NALU_AMD_MESH_H
#define NALU_AMD_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node and element counts of an nx x ny x nz cell box. */
int nw_hex_box_sizes(int nx, int ny, int nz, int64_t* nnode, int64_t* nelem);

/* Box [0,1]^3: nodes numbered x fastest; z spacing stretched geometrically
 * toward z = 0 by `stretch_z` per cell (1 = uniform); interior nodes moved
 * by a seeded uniform offset of +-jitter times the local spacing.
 * coords [nnode][3], conn [nelem][8] in exodus Hex8 node order. */
int nw_hex_box(int nx, int ny, int nz, double stretch_z, double jitter,
               uint64_t seed, double* coords, int32_t* conn);

/* Unique edges of a Hex8 mesh, node[0] = smaller node id (the polarity of
 * stk::mesh::create_edges), sorted by (node[0], node[1]); elem_edges
 * [nelem][12] gives the edge of each Hex8 edge ordinal. edges must hold
 * 2*cap ints; *nedge receives the count. */
int nw_hex_edges(int64_t nelem, const int32_t* conn, int64_t nnode,
                 int32_t* edges, int64_t cap, int64_t* nedge,
                 int32_t* elem_edges);

/* GeometryInteriorAlg::impl_compute_edge_area_vector
 * (src/ngp_algorithms/GeometryInteriorAlg.C:163-225) with HexSCS areas
 * (src/master_element/Hex8CVFEM.C:567-591): areav [nedge][3]. */
int nw_hex_edge_area_vector(int64_t nelem, const int32_t* conn,
                            const double* coords, const int32_t* elem_edges,
                            const int32_t* edges, int64_t nedge,
                            double* areav);

/* GeometryInteriorAlg::impl_compute_dual_nodal_volume
 * (src/ngp_algorithms/GeometryInteriorAlg.C:70-112) with HexSCV volumes
 * (src/master_element/Hex8CVFEM.C:365-389). */
int nw_hex_dual_volume(int64_t nelem, const int32_t* conn,
                       const double* coords, int64_t nnode, double* dualvol);

/* NodalGradEdgeAlg::execute (src/ngp_algorithms/NodalGradEdgeAlg.C:58-111),
 * interior edges: grad [nnode][dim1*3] is overwritten. */
int nw_nodal_grad_edge(int64_t nnode, int64_t nedge, const int32_t* edges,
                       const double* areav, const double* dualvol, int dim1,
                       const double* phi, double* grad);

/* Edge mass flow rate 0.5*(rho_L u_L + rho_R u_R).A (the form of
 * unit_tests/kernels/UnitTestKernelUtils.C:725-769). */
int nw_mdot_edge(int64_t nedge, const int32_t* edges, const double* areav,
                 const double* rho, const double* vel, double* mdot);

/* ---- unstructured mixed meshes -------------------------------------------- */

/* Conforming mixed Hex8 / Tet4 / Wed6 / Pyr5 mesh of [0,1]^3 on an n^3
 * cell lattice: the cells with the lowest values of a seeded smooth random
 * field stay Hex8 (fraction f_hex), the next f_tet are split into 6 Kuhn
 * Tet4, the rest into 2 Wed6 (end faces z); a Hex8 or Wed6 cell with a
 * whole face against a neighbour that cuts it into triangles becomes a
 * transition cell: a node at its centre, a Pyr5 on each whole face and two
 * Tet4 on each cut face (the neighbour's diagonal), so every face of the
 * mesh is shared by matching elements.  Interior lattice nodes are moved
 * by +-jitter/n; nodes are randomly permuted and then renumbered by reverse
 * Cuthill-McKee.  nnode = (n+1)^3 + the number of transition cells.  type
 * [nelem] (0 Hex8, 1 Tet4, 2 Wed6, 3 Pyr5), conn [nelem][8] (exodus node
 * order, the apex of a Pyr5 last, unused slots -1).  With coords == NULL
 * only *nnode and *nelem are returned. */
int nw_mixed_box(int n, double f_hex, double f_tet, double jitter,
                 uint64_t seed, int64_t* nnode, int64_t* nelem,
                 double* coords, int32_t* type, int32_t* conn);

/* Unique edges of a mixed mesh, node[0] < node[1], sorted; with edges ==
 * NULL only *nedge is returned. */
int nw_mixed_edges(int64_t nelem, const int32_t* type, const int32_t* conn,
                   int64_t nnode, int32_t* edges, int64_t cap, int64_t* nedge);

/* GeometryInteriorAlg::impl_compute_edge_area_vector
 * (src/ngp_algorithms/GeometryInteriorAlg.C:163-225) over Hex8, Tet4
 * (src/master_element/Tet4CVFEM.C:522-618), Wed6
 * (src/master_element/Wed6CVFEM.C:541-646) and Pyr5
 * (src/master_element/Pyr5CVFEM.C:771-900) sub-control surfaces. */
int nw_mixed_edge_area_vector(int64_t nelem, const int32_t* type,
                              const int32_t* conn, const double* coords,
                              const int32_t* edges, int64_t nedge,
                              double* areav);

/* ---- synthetic workload input (no reference counterpart) ---------------- */
typedef struct nw_ctx nw_ctx;

/* Sets the context's NW_F_VELOCITY field, on the device, to the Fourier sum
 * u(x) = sum_q b_q sin(k_q . x + phase_q + omega_q t) at the node
 * coordinates (NW_F_COORDINATES must be set): k, b [nmodes][3], phase,
 * omega [nmodes], nmodes <= 32.  Used to step the pressure workloads in
 * time without host traffic. */
int nw_fourier_velocity(nw_ctx* ctx, int nmodes, const double* k,
                        const double* b, const double* phase,
                        const double* omega, double t);

#ifdef __cplusplus
}
#endif

#endif
