// nw_geom_setup.cpp -- host setup of the device geometry: for every edge the
// list of element sub-control surfaces (SCS) that make up its area vector,
// and for every node the list of element sub-control volumes (SCV) that make
// up its dual volume, both in element order.  Built once per mesh; the
// device then sums the per-element values through these lists in a fixed
// order instead of scattering them with atomics.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "nalu_amd.h"
#include "nw_geom.h"

namespace {

// Element-local L/R node of each SCS, in SCS order: Hex8
// (include/master_element/Hex8CVFEM.h:292-293), Tet4 (Tet4CVFEM.h:266),
// Wed6 (Wed6CVFEM.h:268), Pyr5 (Pyr5CVFEM.h:300-301: two surfaces on each
// apex edge).
const int kLr[4][12][2] = {
  {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {4, 5}, {5, 6}, {6, 7}, {4, 7}, {0, 4},
   {1, 5}, {2, 6}, {3, 7}},
  {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}},
  {{0, 1}, {1, 2}, {0, 2}, {3, 4}, {4, 5}, {3, 5}, {0, 3}, {1, 4}, {2, 5}},
  {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {0, 4}, {0, 4}, {1, 4}, {1, 4}, {2, 4},
   {2, 4}, {3, 4}, {3, 4}}};

}  // namespace

namespace nw {
namespace geom {

int topo_nodes(int t) { return t == 0 ? 8 : (t == 1 ? 4 : (t == 2 ? 6 : 5)); }
int topo_scs(int t) { return t == 0 ? 12 : (t == 1 ? 6 : (t == 2 ? 9 : 12)); }

int build_incidence(int64_t nelem, const int32_t* type, const int32_t* conn,
                    int64_t nnode, int64_t nedge, const int32_t* edges,
                    std::vector<int32_t>& e_ptr, std::vector<int32_t>& e_inc,
                    std::vector<int32_t>& n_ptr, std::vector<int32_t>& n_inc)
{
  if (nelem < 0 || nelem >= (int64_t{1} << 26)) return NW_E_ARG;
  for (int64_t el = 0; el < nelem; ++el) {
    const int t = type ? type[el] : 0;
    if (t < 0 || t > 3) return NW_E_ARG;
    for (int a = 0; a < topo_nodes(t); ++a) {
      const int32_t n = conn[8 * el + a];
      if (n < 0 || n >= nnode) return NW_E_ARG;
    }
  }
  // edges by their smaller node: (larger node, edge id), sorted
  std::vector<int64_t> ap(nnode + 1, 0);
  for (int64_t e = 0; e < nedge; ++e)
    ap[std::min(edges[2 * e], edges[2 * e + 1]) + 1]++;
  for (int64_t n = 0; n < nnode; ++n) ap[n + 1] += ap[n];
  std::vector<std::pair<int32_t, int32_t>> adj(nedge);
  {
    std::vector<int64_t> pos(ap.begin(), ap.end() - 1);
    for (int64_t e = 0; e < nedge; ++e) {
      const int32_t a = edges[2 * e], b = edges[2 * e + 1];
      adj[pos[std::min(a, b)]++] = {std::max(a, b), (int32_t)e};
    }
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t n = 0; n < nnode; ++n)
    std::sort(adj.begin() + ap[n], adj.begin() + ap[n + 1]);
  // SCS s of element el -> code 2 (12 el + s) + (L is not the edge's node[0])
  std::vector<int32_t> code((size_t)nelem * 12, -1), eid((size_t)nelem * 12, -1);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t el = 0; el < nelem; ++el) {
    const int t = type ? type[el] : 0;
    for (int s = 0; s < topo_scs(t); ++s) {
      const int32_t nl = conn[8 * el + kLr[t][s][0]], nr = conn[8 * el + kLr[t][s][1]];
      const int32_t lo = std::min(nl, nr), hi = std::max(nl, nr);
      const auto b = adj.begin() + ap[lo], en = adj.begin() + ap[lo + 1];
      const auto it = std::lower_bound(b, en, std::make_pair(hi, (int32_t)-1));
      if (it == en || it->first != hi) {
        bad = 1;
        continue;
      }
      const int32_t e = it->second;
      eid[12 * el + s] = e;
      code[12 * el + s] = (int32_t)(2 * (12 * el + s) + (edges[2 * e] == nl ? 0 : 1));
    }
  }
  if (bad) return NW_E_ARG;
  e_ptr.assign(nedge + 1, 0);
  for (int64_t k = 0; k < nelem * 12; ++k)
    if (eid[k] >= 0) e_ptr[eid[k] + 1]++;
  for (int64_t e = 0; e < nedge; ++e) e_ptr[e + 1] += e_ptr[e];
  e_inc.resize(e_ptr[nedge]);
  {
    std::vector<int32_t> pos(e_ptr.begin(), e_ptr.end() - 1);
    for (int64_t k = 0; k < nelem * 12; ++k)
      if (eid[k] >= 0) e_inc[pos[eid[k]]++] = code[k];
  }
  // SCV a of element el -> 8 el + a, by node
  n_ptr.assign(nnode + 1, 0);
  for (int64_t el = 0; el < nelem; ++el) {
    const int t = type ? type[el] : 0;
    for (int a = 0; a < topo_nodes(t); ++a) n_ptr[conn[8 * el + a] + 1]++;
  }
  for (int64_t n = 0; n < nnode; ++n) n_ptr[n + 1] += n_ptr[n];
  n_inc.resize(n_ptr[nnode]);
  {
    std::vector<int32_t> pos(n_ptr.begin(), n_ptr.end() - 1);
    for (int64_t el = 0; el < nelem; ++el) {
      const int t = type ? type[el] : 0;
      for (int a = 0; a < topo_nodes(t); ++a)
        n_inc[pos[conn[8 * el + a]]++] = (int32_t)(8 * el + a);
    }
  }
  return NW_OK;
}

}  // namespace geom
}  // namespace nw
