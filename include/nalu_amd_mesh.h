/*
 * nalu_amd_mesh.h -- host-side mesh and geometry producers for the edge
 * path: a structured Hex8 box generator (uniform / stretched / jittered),
 * edge extraction, CVFEM edge area vectors and dual volumes, the edge
 * nodal gradient and the edge mass flow rate, a conforming mixed-element
 * box generator, and Exodus II mesh input and output.  These produce the
 * inputs the edge solver algorithms consume; they run multithreaded on the
 * host.  All arrays are host arrays; return 0 or a negative NW_E_* code.
 * The last entry generates a synthetic velocity field on the device.
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

/* ---- on-disk mesh input and output: Exodus II ------------------------------
 * The reference reads its mesh through stk::io::StkMeshIoBroker
 * (Realm::create_mesh, src/Realm.C:2048-2094: add_mesh_database +
 * create_input_mesh; Realm::initialize_prolog, src/Realm.C:450-466:
 * populate_mesh, then create_edges) and writes results through create_output_mesh
 * (src/Realm.C:2100-2215).  These entry points read and write the same
 * Exodus II files when they are stored in the netCDF classic format (CDF-1)
 * or its 64-bit-offset form (CDF-2); netCDF-4 (HDF5) and CDF-5 files are
 * refused with NW_E_ARG and the reason (nccopy -k 64-bit-offset converts
 * them).  Node and element numbers are 0-based here (1-based on disk);
 * elements keep the file's order, block by block.  Host arrays; errors
 * carry a message for nw_last_error(). */
typedef struct nw_exo nw_exo;

enum { NW_EXO_BLOCK = 0, NW_EXO_NODE_SET = 1, NW_EXO_SIDE_SET = 2, NW_EXO_NODAL_VAR = 3 };

typedef struct nw_exo_info {
  int64_t nnode, nelem;
  int32_t nblock, nnodeset, nsideset, nnodvar, ntime;
  int32_t format; /* 1 CDF-1, 2 CDF-2 (64-bit offsets) */
} nw_exo_info;

/* Opens the file and reads its header (the mesh database of
 * add_mesh_database / create_input_mesh). */
int nw_exo_open(const char* path, nw_exo** exo);
void nw_exo_close(nw_exo* exo);
int nw_exo_sizes(const nw_exo* exo, nw_exo_info* info);

/* Entity `index` (0-based, file order) of `kind`: its id, its count
 * (elements of a block, nodes of a node set, sides of a side set, nodes for
 * a nodal variable), the topology of a block (0 Hex8, 1 Tet4, 2 Wed6,
 * 3 Pyr5; -1 for another element type, and for other kinds) and its name
 * ("" when the file gives none; at most name_cap - 1 bytes).  Any pointer
 * may be NULL. */
int nw_exo_entity(const nw_exo* exo, int kind, int32_t index, int32_t* id,
                  int64_t* count, int32_t* topo, char* name, int32_t name_cap);
/* The index of the entity of `kind` named `name` (NW_E_ARG if none). */
int nw_exo_find(const nw_exo* exo, int kind, const char* name, int32_t* index);

/* populate_mesh: coords [nnode][3]; type [nelem] (0 Hex8, 1 Tet4, 2 Wed6,
 * 3 Pyr5); conn [nelem][8] in Exodus node order (unused slots -1); block
 * [nelem] the block index of each element (may be NULL).  A block of any
 * other element type (shells, 2-D, higher order) is an error naming it. */
int nw_exo_read_mesh(const nw_exo* exo, double* coords, int32_t* type,
                     int32_t* conn, int32_t* block);
/* Node set `index`: nodes [count], file order. */
int nw_exo_node_set(const nw_exo* exo, int32_t index, int32_t* nodes);
/* Side set `index`: elem [count] and side ordinal [count] (0-based: side s
 * of a Hex8 is nw_side_nodes(0, s)). */
int nw_exo_side_set(const nw_exo* exo, int32_t index, int32_t* elem,
                    int32_t* side);
/* Nodal variable `index` at time step `step`: values [nnode], *time (may be
 * NULL) the step's time (the restart/initial-condition read of
 * populate_field_data, src/Realm.C:476-484). */
int nw_exo_read_nodal_var(const nw_exo* exo, int32_t index, int32_t step,
                          double* values, double* time);

/* The nodes of side `side` of topology `type` (element-local ordinals),
 * ordered so that the right-hand rule gives the outward normal
 * (sideNodeOrdinals_, include/master_element/Hex8CVFEM.h:395-401,
 * Tet4CVFEM.h:258-263, Wed6CVFEM.h:259-265, Pyr5CVFEM.h:291-297): returns
 * the count (3 or 4), or NW_E_ARG. */
int nw_side_nodes(int32_t type, int32_t side, int32_t* nodes);
/* The nodes of a list of element sides (a side set), sorted and unique:
 * the node list a Dirichlet condition on that surface applies to.  With
 * nodes == NULL only *count is returned. */
int nw_sides_to_nodes(int64_t nelem, const int32_t* type, const int32_t* conn,
                      int64_t nnode, int64_t nside, const int32_t* elem,
                      const int32_t* side, int32_t* nodes, int64_t* count);

/* Sets for nw_exo_write: set i holds entries [offsets[i], offsets[i+1]):
 * nodes (node sets) or elements with their 0-based side ordinals (side
 * sets); ids and names may be NULL (ids 1.., no names). */
typedef struct nw_exo_sets {
  int32_t count;
  const int32_t* ids;
  const char* const* names;
  const int64_t* offsets; /* [count + 1] */
  const int32_t* entries;
  const int32_t* sides; /* side sets only */
} nw_exo_sets;

/* create_output_mesh + one process_output_request: writes a CDF-2 Exodus II
 * file holding the mesh (one element block per topology present, named
 * block_hex8 / block_tet4 / block_wedge6 / block_pyramid5, elements grouped
 * by topology in their given order), the node and side sets (either may be
 * NULL) and `nvar` nodal variables (var_values [nvar][nnode]) at one time
 * step of time `time` (none when nvar == 0).  elem_order [nelem] (may be
 * NULL) receives, for each element of the file, its index in `type`/`conn`. */
int nw_exo_write(const char* path, const char* title, int64_t nnode,
                 const double* coords, int64_t nelem, const int32_t* type,
                 const int32_t* conn, const nw_exo_sets* node_sets,
                 const nw_exo_sets* side_sets, int32_t nvar,
                 const char* const* var_names, const double* var_values,
                 double time, int64_t* elem_order);

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
