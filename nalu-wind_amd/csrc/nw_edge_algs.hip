// nw_edge_algs.hip -- the edge algorithms on either side of the pressure
// solve, on gfx950:
//
//   k_mdot       MdotEdgeAlg (src/ngp_algorithms/MdotEdgeAlg.C:121-186,
//                with or without the balanced buoyancy force):
//                one lane per edge, edge-ordered (coalesced) area-vector
//                reads and mass-flow-rate writes, the two nodes' fields
//                gathered as node-major records;
//   k_nodal_grad NodalGradEdgeAlg (src/ngp_algorithms/NodalGradEdgeAlg.C:
//                58-111): one lane per row of the edge graph, rows taken in
//                the assembly's locality-ordered 64-row chunks; the lane
//                walks the row's slots (slot code 2e + side names the edge
//                and which end the row is), sums +-q_ip A_e / V and writes
//                the row's 3c components once.  Each node's sum is owned by one
//                lane, so no atomics are needed and the result does not
//                depend on scheduling (the reference scatters both ends of
//                every edge with atomic adds).
// Both are gather-bound: their roofline is HBM bandwidth.
#include <algorithm>
#include <cmath>

#include "nw_internal.h"

namespace nw {

namespace {

struct MdotArgs {
  const double *coords, *vel, *dpdx, *rho, *p, *udiag, *areav;
  int64_t N, E;
  double noc, interp;
  // the balanced buoyancy force: buoyancy_source, its mask, gravity
  const double *src, *mask;
  double grav[3];
};

// BUOY: the balanced buoyancy force (MdotEdgeAlg.C:147-154,163-169)
template <bool BUOY>
__global__ __launch_bounds__(kBlock) void k_mdot(MdotArgs a,
                                                 const int2* __restrict__ edges,
                                                 const int32_t* __restrict__ order,
                                                 double* __restrict__ mdot)
{
  const int64_t i = (int64_t)xcd_block((int)gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= a.E) return;
  const int64_t e = order ? (int64_t)order[i] : i;
  const int2 lr = edges[e];
  const int64_t L = lr.x, R = lr.y;
  const double om_interp = 1.0 - a.interp;
  double av[3];
  for (int d = 0; d < 3; ++d) av[d] = a.areav[3 * e + d];
  const double udL = a.udiag[L], udR = a.udiag[R];
  const double rhoL = a.rho[L], rhoR = a.rho[R];
  const double projTimeScale = 0.5 * (1.0 / udL + 1.0 / udR);
  const double rhoIp = 0.5 * (rhoL + rhoR);
  double dx[3], axdx = 0.0, asq = 0.0;
  for (int d = 0; d < 3; ++d) {
    dx[d] = a.coords[3 * R + d] - a.coords[3 * L + d];
    asq += av[d] * av[d];
    axdx += av[d] * dx[d];
  }
  const double inv_axdx = 1.0 / axdx;
  double tmdot = -projTimeScale * (a.p[R] - a.p[L]) * asq * inv_axdx;
  double mL = 0.0, mR = 0.0;
  if constexpr (BUOY) {
    mL = a.mask[L];
    mR = a.mask[R];
    const double masked_weights = 0.5 * (mL + mR);
    for (int d = 0; d < 3; ++d)
      tmdot += projTimeScale * av[d] * a.grav[d] * rhoIp * masked_weights;
  }
  for (int d = 0; d < 3; ++d) {
    const double kxj = av[d] - asq * inv_axdx * dx[d];
    const double uR = a.vel[3 * R + d], uL = a.vel[3 * L + d];
    const double rhoUjIp = 0.5 * (rhoR * uR + rhoL * uL);
    const double ujIp = 0.5 * (uR + uL);
    double GjIp = 0.5 * (a.dpdx[3 * R + d] / udR + a.dpdx[3 * L + d] / udL);
    if constexpr (BUOY)
      GjIp -= 0.5 * ((mR * a.src[3 * R + d]) / udR + (mL * a.src[3 * L + d]) / udL);
    tmdot += (a.interp * rhoUjIp + om_interp * rhoIp * ujIp + GjIp) * av[d] -
             kxj * GjIp * a.noc;
  }
  mdot[e] = tmdot;
}

// Slots of a chunk staged in LDS: 512 (64 Hex8 rows hold 448) or 1024
// (mixed meshes), chosen from the graph's 99th-percentile chunk; larger
// chunks read codes and columns from HBM directly.
constexpr int kGradSlotsSmall = 512, kGradSlotsLarge = 1024;

// grad[3C n + 3 j + i] = sum_slots sign * (q_ip,j A_e,i) / V_n, summed
// in slot order.  A wave takes one 64-row chunk of the assembly's locality
// schedule (XCD-aware block order): the chunk's slot codes and columns are
// staged in LDS by coalesced loads, and each lane issues the neighbour and
// edge gathers of U slots before it uses any of them.
template <int C, int U, int SCAP>
__global__ __launch_bounds__(kBlock) void k_nodal_grad(
  const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ code,
  const int32_t* __restrict__ cols, const SchedChunk* __restrict__ sched,
  int nchunks, const double* __restrict__ q, const double* __restrict__ areav,
  const double* __restrict__ dualvol, double* __restrict__ grad, int64_t N,
  int64_t E, int accumulate)
{
  constexpr int WPB = kBlock / kWave;
  constexpr int NG = 3 * C;  // gradient components per node
  __shared__ int code_s[WPB][SCAP];
  __shared__ int col_s[WPB][SCAP];
  __shared__ double out_s[WPB][kWave * NG];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = (int)(threadIdx.x / kWave);
  const int chunk_id = xcd_block((int)gridDim.x) * WPB + wv;
  if (chunk_id >= nchunks) return;
  const SchedChunk ch = sched[chunk_id];
  const int s_lo = ch.s0, ns = ch.s1 - ch.s0;
  const bool staged = ns <= SCAP;
  if (staged) {
    // all of a lane's loads issued before any is stored: one round trip
    constexpr int KS = SCAP / kWave;
    int cv[KS], mv[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int i = lane + k * kWave;
      if (i < ns) {
        cv[k] = code[s_lo + i];
        mv[k] = cols[s_lo + i];
      }
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int i = lane + k * kWave;
      if (i < ns) {
        code_s[wv][i] = cv[k];
        col_s[wv][i] = mv[k];
      }
    }
  }
  wave_lds_sync();
  const int n = ch.r0 + lane;
  const bool active = n < ch.r1;
  double qn[C], acc[NG];
  for (int k = 0; k < NG; ++k) acc[k] = 0.0;
  const double inv = active ? 1.0 / dualvol[n] : 0.0;
  for (int j = 0; j < C; ++j) qn[j] = active ? q[C * (int64_t)n + j] : 0.0;
  const int s1 = active ? row_ptr[n + 1] : 0;
  for (int s = active ? row_ptr[n] : 0; s < s1; s += U) {
    int cu[U], mu[U];
    for (int u = 0; u < U; ++u) {
      const int t = s + u;
      cu[u] = t < s1 ? (staged ? code_s[wv][t - s_lo] : code[t]) : -1;
      mu[u] = t < s1 ? (staged ? col_s[wv][t - s_lo] : cols[t]) : n;
    }
    double av[U][3], qm[U][C];
    for (int u = 0; u < U; ++u) {
      if (cu[u] < 0) continue;  // the diagonal slot, or past the row
      const int64_t e = cu[u] >> 1;
      for (int i = 0; i < 3; ++i) av[u][i] = areav[3 * e + i];
      for (int j = 0; j < C; ++j) qm[u][j] = q[C * (int64_t)mu[u] + j];
    }
    for (int u = 0; u < U; ++u) {
      if (cu[u] < 0) continue;
      const double sgn = (cu[u] & 1) ? -1.0 : 1.0;  // n = R: minus, n = L: plus
      for (int j = 0; j < C; ++j) {
        const double qip = 0.5 * (qn[j] + qm[u][j]);
        for (int i = 0; i < 3; ++i) acc[3 * j + i] += sgn * ((av[u][i] * qip) * inv);
      }
    }
  }
  // the wave's rows are consecutive nodes: stage their gradients in LDS and
  // write them as one contiguous range (a lane's own NG values would be
  // strided stores)
  if (active)
    for (int k = 0; k < NG; ++k) out_s[wv][NG * lane + k] = acc[k];
  wave_lds_sync();
  const int nv = (ch.r1 - ch.r0) * NG;
  double* dst = grad + (int64_t)ch.r0 * NG;
  if (accumulate)
    for (int i = lane; i < nv; i += kWave) dst[i] += out_s[wv][i];
  else
    for (int i = lane; i < nv; i += kWave) dst[i] = out_s[wv][i];
}

}  // namespace

}  // namespace nw

using namespace nw;

extern "C" {

int nw_compute_mdot(nw_ctx* c, const nw_mdot_opts* o)
{
  NW_REQUIRE(c && o, NW_E_ARG, "null argument");
  NW_REQUIRE(c->nnode > 0 && c->nedge >= 0, NW_E_STATE, "mesh not set");
  const double *coords, *vel, *dpdx, *rho, *p, *udiag, *areav;
  int rc = 0;
  if ((rc = require_field(c, NW_F_COORDINATES, &coords)) ||
      (rc = require_field(c, NW_F_VELOCITY, &vel)) ||
      (rc = require_field(c, NW_F_DPDX, &dpdx)) ||
      (rc = require_field(c, NW_F_DENSITY, &rho)) ||
      (rc = require_field(c, NW_F_PRESSURE, &p)) ||
      (rc = require_field(c, NW_F_UDIAG, &udiag)) ||
      (rc = require_field(c, NW_E_AREA_VECTOR, &areav)))
    return rc;
  double* mdot = nullptr;
  if ((rc = ensure_field(c, NW_E_MASS_FLOW_RATE, &mdot))) return rc;
  if (c->nedge == 0) return NW_OK;
  MdotArgs a{coords, vel, dpdx, rho, p, udiag, areav, c->nnode, c->nedge,
             o->noc, o->interp_together, nullptr, nullptr, {0.0, 0.0, 0.0}};
  if (o->balanced_buoyancy) {
    if ((rc = require_field(c, NW_F_BUOYANCY_SOURCE, &a.src)) ||
        (rc = require_field(c, NW_F_BUOYANCY_SOURCE_MASK, &a.mask)))
      return rc;
    for (int d = 0; d < 3; ++d) a.grav[d] = o->gravity[d];
  }
  // with the edge graph built, the edges go in its schedule's order (rows
  // of a chunk are neighbours in space, so the two ends' node data is
  // shared by nearby lanes); otherwise in their own order
  const int32_t* order = nullptr;
  if (c->g.row_ptr && c->g.nnode == c->nnode) {
    if ((rc = ensure_schedule(c))) return rc;
    order = c->g.edge_order;
  }
  // XCD-aware block order: each XCD takes one contiguous stretch
  const unsigned nb = (grid_for(c->nedge, kBlock) + 7) / 8 * 8;
  if (o->balanced_buoyancy)
    k_mdot<true><<<nb, kBlock, 0, c->stream>>>(
      a, reinterpret_cast<const int2*>(c->edges), order, mdot);
  else
    k_mdot<false><<<nb, kBlock, 0, c->stream>>>(
      a, reinterpret_cast<const int2*>(c->edges), order, mdot);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

int nw_nodal_grad(nw_ctx* c, int fin, int fout, int accumulate)
{
  NW_REQUIRE(c, NW_E_ARG, "null context");
  NW_REQUIRE(fin >= 0 && fin < NW_NUM_FIELDS && fout >= 0 && fout < NW_NUM_FIELDS &&
               !field_is_edge(fin) && !field_is_edge(fout) && fin != fout,
             NW_E_ARG, "nodal gradient needs two distinct node fields");
  const int C = field_components(fin);
  NW_REQUIRE(C == 1 || C == 3, NW_E_ARG, "input field must have 1 or 3 components");
  NW_REQUIRE(field_components(fout) == 3 * C, NW_E_ARG,
             "output field must have %d components", 3 * C);
  const Graph& g = c->g;
  NW_REQUIRE(g.row_ptr && g.nnode == c->nnode, NW_E_STATE,
             "nodal gradient needs the edge graph (nw_linsys_build_edge_graph)");
  const double *q, *areav, *dualvol;
  int rc = 0;
  if ((rc = require_field(c, fin, &q)) ||
      (rc = require_field(c, NW_E_AREA_VECTOR, &areav)) ||
      (rc = require_field(c, NW_F_DUAL_VOLUME, &dualvol)))
    return rc;
  double* grad = nullptr;
  if ((rc = ensure_field(c, fout, &grad))) return rc;
  if ((rc = ensure_schedule(c))) return rc;
  if (g.nchunks == 0) return NW_OK;
  constexpr int64_t wpb = kBlock / kWave;
  const unsigned nb = (unsigned)(((g.nchunks + wpb - 1) / wpb + 7) / 8 * 8);
  const bool small = g.chunk_slots_p99 <= kGradSlotsSmall;
#define NW_GRAD(CC, SC)                                                       \
  k_nodal_grad<CC, 4, SC><<<nb, kBlock, 0, c->stream>>>(                      \
    g.row_ptr, g.code, g.cols, g.sched, (int)g.nchunks, q, areav, dualvol,    \
    grad, c->nnode, c->nedge, accumulate)
  if (C == 1 && small)
    NW_GRAD(1, kGradSlotsSmall);
  else if (C == 1)
    NW_GRAD(1, kGradSlotsLarge);
  else if (small)
    NW_GRAD(3, kGradSlotsSmall);
  else
    NW_GRAD(3, kGradSlotsLarge);
#undef NW_GRAD
  NW_LAUNCH_CHECK();
  return NW_OK;
}

}  // extern "C"
