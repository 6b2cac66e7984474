// nw_geom.h -- element geometry shared by the host mesh producers and the
// host setup of the device geometry.
#pragma once

#include <cstdint>
#include <vector>

namespace nw {
namespace geom {

// Area vectors of the 12 Hex8 sub-control surfaces, SCS order of
// HexSCS::determinant_scs (src/master_element/Hex8CVFEM.C:567-591).
void hex_scs(const double c[8][3], double a[12][3]);

// Nodes and sub-control surfaces of topology t (0 Hex8, 1 Tet4, 2 Wed6,
// 3 Pyr5).
int topo_nodes(int t);
int topo_scs(int t);

// Incidence lists of a mesh of nelem elements (type [nelem], NULL = all
// Hex8; conn [nelem][8], unused slots -1) over its nedge edges
// (edges [nedge][2]): for edge e, e_inc[e_ptr[e] .. e_ptr[e+1]) holds
// 2 (12 el + s) + flip for every sub-control surface s of element el on
// that edge, in element order, flip = 1 when the surface's L node is the
// edge's node[1]; for node n, n_inc[n_ptr[n] .. n_ptr[n+1]) holds 8 el + a
// for every element el whose local node a it is, in element order.
// Returns NW_E_ARG for bad input, an element edge missing from the edge
// list or more than 2^26 elements.
int build_incidence(int64_t nelem, const int32_t* type, const int32_t* conn,
                    int64_t nnode, int64_t nedge, const int32_t* edges,
                    std::vector<int32_t>& e_ptr, std::vector<int32_t>& e_inc,
                    std::vector<int32_t>& n_ptr, std::vector<int32_t>& n_inc);

}  // namespace geom
}  // namespace nw
