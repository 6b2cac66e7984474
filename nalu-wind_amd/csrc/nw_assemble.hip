// nw_assemble.hip -- the interior edge solver algorithms on gfx950.
//
// Layout.  The linear system is a block CSR over nodes: row n holds the
// slots of n and of its edge neighbours, sorted by column (the sorted Tpetra
// graph).  Every slot carries a code: the diagonal slot of row n is
// -(n+1); the slot (n, m) reached through edge e = (L, R) is 2e + side with
// side = 0 when n == L and 1 when n == R.  Node and edge fields are stored
// node-major per field, the caller's own layout: component d of node n is
// f[ncomp * n + d], so a neighbour's vector or gradient is one contiguous
// record (a few vector loads) and consecutive nodes are still consecutive.
//
// Kernel shape ("row owner").  G lanes own one row (G = 1 for the 1-dof
// systems, 2 for momentum), and the lanes of an instruction sit on
// consecutive rows.  Each lane loads the row node's fields, then walks
// every G-th slot of the row: for each off-diagonal slot it gathers the
// neighbour's fields (the slot's column) and the edge's fields, evaluates
// the edge flux in registers exactly as the reference lambda does for that
// edge, puts the (row, neighbour) block of the edge LHS into the slot, and
// adds the edge's (row, row) block and rhs(row) contributions to register
// accumulators.  The G partial sums are combined by a fixed butterfly and
// the diagonal block and rhs are written once.  A wave's rows are
// consecutive, so their slots are one contiguous range: the wave stages
// that range's codes, columns and values in LDS and moves them to and from
// HBM as whole cache lines.  Every edge is evaluated twice (once from each
// end), each matrix entry is written once, no atomics are used, and the
// result is bitwise reproducible.
//
// Rows are taken in chunks of 64 consecutive rows from a schedule ordered
// for locality (the row order, a plane-by-plane sweep of strips, or the
// Morton code of the chunk centroids, see build_schedule), and workgroups
// are renumbered so that each XCD works on one contiguous stretch of that
// schedule: rows processed together are neighbours in space and their
// shared node and edge data stays in the XCD's L2.
//
// The physics restates the reference edge kernels:
//   ScalarEdgeSolverAlg.C:87-205, MomentumEdgeSolverAlg.C:107-312,
//   ContinuityEdgeSolverAlg.C:111-194 (src/edge_kernels/).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "nw_internal.h"

namespace nw {

namespace {

struct PecletDev {
  int form;
  double hybrid, c1, c2;
  __device__ double operator()(double pe) const
  {
    // src/PecletFunction.C:41-45 (classic), 68-71 (tanh)
    if (form == 0) {
      const double m = hybrid * pe;
      return m * m / (5.0 + m * m);
    }
    return 0.50 * (1.0 + tanh((pe - c1) / c2));
  }
};

__device__ inline double van_leer(double dqm, double dqp, double eps)
{
  // include/edge_kernels/EdgeKernelUtils.h:18-24.  No contraction: fused
  // into fma(dqm, dqp, |dqm dqp|) the numerator of slopes of opposite sign
  // is the product's rounding error instead of 0, which the denominator
  // (dqm + dqp)^2 + eps turns into a limiter of order one when the slopes
  // nearly cancel.
#pragma clang fp contract(off)
  const double p = dqm * dqp;
  return (2.0 * (p + fabs(p))) / ((dqm + dqp) * (dqm + dqp) + eps);
}

// ------------------------------------------------------------------ physics
// Each physics type has a Node record of the fields one node contributes,
// load(n, rec), and edge(e, side, L, R, off, acc) which computes, for the
// row n = node[side] of edge e:
//   off[NB]  the (n, other) block of the edge LHS, row-major ndof x ndof
//   acc[NV]  the (n, n) block contribution (NB values) then rhs(n) (NDOF)

// s, m: buoyancy_source and its mask, loaded only with the balanced
// buoyancy force (unused registers are dropped otherwise)
struct ContNode {
  double x[3], u[3], g[3], rho, p, ud, s[3], m;
};

// Launch shape of each physics (measured on the 256^3 continuity, 128^3
// momentum and 160^3 mixed-mesh workloads): kLanes lanes per row (the 1-dof
// systems choose 1 or 2 from the graph, see launch_rows), kUnroll slots in
// flight per lane with one lane per row and kUnroll2 with two, non-temporal
// streaming, the LDS slot capacities per wave, the occupancy the register
// allocation must allow, and whether the persistent kernel
// (k_assemble_persistent) runs it.
// BUOY: the balanced buoyancy force (ContinuityEdgeSolverAlg.C:147-152,
// 168-173), a separate instantiation so that the plain kernel reads no
// extra field.
template <bool BUOY>
struct ContinuityPhysT {
  static constexpr int NDOF = 1;
  static constexpr int kLanes = 1, kUnroll = 2, kUnroll2 = 1, kMinWaves = 2;
  static constexpr bool kNonTemporal = false, kPersistent = false;
  static constexpr int kSlotsSmall = 512, kSlotsLarge = 1024;
  using Node = ContNode;
  const double *coords, *vel, *dpdx, *rho, *p, *udiag, *areav;
  int64_t N, E;
  double nocFac, tauScale, interpTogether, incompressible;
  const double *src, *mask;
  double grav[3];

  __device__ void load(int n, Node& r) const
  {
    for (int d = 0; d < 3; ++d) {
      r.x[d] = coords[3 * (int64_t)n + d];
      r.u[d] = vel[3 * (int64_t)n + d];
      r.g[d] = dpdx[3 * (int64_t)n + d];
    }
    r.rho = rho[n];
    r.p = p[n];
    r.ud = udiag[n];
    if constexpr (BUOY) {
      for (int d = 0; d < 3; ++d) r.s[d] = src[3 * (int64_t)n + d];
      r.m = mask[n];
    }
  }

  struct Edge {
    double av[3];
  };

  __device__ void load_edge(int e, Edge& r) const
  {
    for (int d = 0; d < 3; ++d) r.av[d] = areav[3 * (int64_t)e + d];
  }

  __device__ void edge(const Edge& er, int side, const Node& L, const Node& R,
                       double* off, double* acc) const
  {
    const double om_interpTogether = 1.0 - interpTogether;
    const double om_incomp = 1.0 - incompressible;
    const double* av = er.av;
    const double projTimeScale = 0.5 * (1.0 / L.ud + 1.0 / R.ud);
    const double rhoIp = 0.5 * (L.rho + R.rho);
    const double denScale = (1.0 / rhoIp) * incompressible + om_incomp;
    double dx[3];
    double axdx = 0.0, asq = 0.0;
    for (int d = 0; d < 3; ++d) {
      dx[d] = R.x[d] - L.x[d];
      asq += av[d] * av[d];
      axdx += av[d] * dx[d];
    }
    const double inv_axdx = 1.0 / axdx;
    double tmdot = -projTimeScale * (R.p - L.p) * asq * inv_axdx;
    if constexpr (BUOY) {
      const double masked_weights = 0.5 * (L.m + R.m);
      for (int d = 0; d < 3; ++d)
        tmdot += projTimeScale * av[d] * grav[d] * rhoIp * masked_weights;
    }
    for (int d = 0; d < 3; ++d) {
      const double kxj = av[d] - asq * inv_axdx * dx[d];
      const double rhoUjIp = 0.5 * (R.rho * R.u[d] + L.rho * L.u[d]);
      const double ujIp = 0.5 * (R.u[d] + L.u[d]);
      double GjIp = 0.5 * (R.g[d] / R.ud + L.g[d] / L.ud);
      if constexpr (BUOY)
        GjIp -= 0.5 * ((R.m * R.s[d]) / R.ud + (L.m * L.s[d]) / L.ud);
      tmdot += (interpTogether * rhoUjIp + om_interpTogether * rhoIp * ujIp +
                GjIp) * av[d] -
               kxj * GjIp * nocFac;
    }
    tmdot /= tauScale;
    tmdot *= denScale;
    const double lhsfac = -asq * inv_axdx * projTimeScale * denScale / tauScale;
    off[0] = lhsfac;   // lhs(0,1) == lhs(1,0)
    acc[0] = -lhsfac;  // lhs(0,0) == lhs(1,1)
    acc[1] = side ? tmdot : -tmdot;
  }
};

struct ScalNode {
  double x[3], u[3], g[3], q, rho, dflux;
};

struct ScalarPhys {
  static constexpr int NDOF = 1;
  static constexpr int kLanes = 1, kUnroll = 1, kUnroll2 = 1, kMinWaves = 2;
  static constexpr bool kNonTemporal = false, kPersistent = false;
  static constexpr int kSlotsSmall = 512, kSlotsLarge = 1024;
  using Node = ScalNode;
  const double *coords, *vrtm, *q, *dqdx, *rho, *dflux, *areav, *mdotf;
  int64_t N, E;
  double alpha, alphaUpw, hoUpwind, relaxFac;
  int useLimiter;
  PecletDev pec;

  __device__ void load(int n, Node& r) const
  {
    for (int d = 0; d < 3; ++d) {
      r.x[d] = coords[3 * (int64_t)n + d];
      r.u[d] = vrtm[3 * (int64_t)n + d];
      r.g[d] = dqdx[3 * (int64_t)n + d];
    }
    r.q = q[n];
    r.rho = rho[n];
    r.dflux = dflux[n];
  }

  struct Edge {
    double av[3], mdot;
  };

  __device__ void load_edge(int e, Edge& r) const
  {
    for (int d = 0; d < 3; ++d) r.av[d] = areav[3 * (int64_t)e + d];
    r.mdot = mdotf[e];
  }

  __device__ void edge(const Edge& er, int side, const Node& L, const Node& R,
                       double* off, double* acc) const
  {
    const double eps = 1.0e-16;
    const double om_alpha = 1.0 - alpha, om_alphaUpw = 1.0 - alphaUpw;
    const double* av = er.av;
    const double mdot = er.mdot;
    const double qNp1L = L.q, qNp1R = R.q;
    const double viscIp = 0.5 * (L.dflux + R.dflux);
    const double diffIp = 0.5 * (L.dflux / L.rho + R.dflux / R.rho);
    double dx[3];
    double axdx = 0.0, asq = 0.0, udotx = 0.0;
    for (int d = 0; d < 3; ++d) {
      dx[d] = R.x[d] - L.x[d];
      asq += av[d] * av[d];
      axdx += av[d] * dx[d];
      udotx += 0.5 * dx[d] * (R.u[d] + L.u[d]);
    }
    const double inv_axdx = 1.0 / axdx;
    double dqL = 0.0, dqR = 0.0, nonOrth = 0.0;
    for (int d = 0; d < 3; ++d) {
      dqL += 0.5 * dx[d] * L.g[d];
      dqR += 0.5 * dx[d] * R.g[d];
      const double kxj = av[d] - asq * inv_axdx * dx[d];
      nonOrth += -viscIp * kxj * 0.5 * (R.g[d] + L.g[d]);
    }
    const double pecnum = fabs(udotx) / (diffIp + eps);
    const double pecfac = pec(pecnum);
    const double om_pecfac = 1.0 - pecfac;
    double limitL = 1.0, limitR = 1.0;
    if (useLimiter) {
      const double dq = qNp1R - qNp1L;
      limitL = van_leer(4.0 * dqL - dq, dq, eps);
      limitR = van_leer(4.0 * dqR - dq, dq, eps);
    }
    const double qIpL = qNp1L + dqL * hoUpwind * limitL;
    const double qIpR = qNp1R - dqR * hoUpwind * limitR;
    const double lhsfac = -viscIp * asq * inv_axdx;
    const double diffFlux = lhsfac * (qNp1R - qNp1L) + nonOrth;
    double l00 = -lhsfac / relaxFac, l01 = lhsfac, l10 = lhsfac,
           l11 = -lhsfac / relaxFac;
    double r0 = -diffFlux, r1 = diffFlux;
    const double qIp = 0.5 * (qNp1R + qNp1L);
    const double qUpw = (mdot > 0) ? (alphaUpw * qIpL + om_alphaUpw * qIp)
                                   : (alphaUpw * qIpR + om_alphaUpw * qIp);
    const double qHatL = (alpha * qIpL + om_alpha * qIp);
    const double qHatR = (alpha * qIpR + om_alpha * qIp);
    const double qCds = 0.5 * (qHatL + qHatR);
    const double adv_flux = mdot * (pecfac * qUpw + om_pecfac * qCds);
    r0 -= adv_flux;
    r1 += adv_flux;
    double alhsfac = 0.5 * (mdot + fabs(mdot)) * pecfac * alphaUpw +
                     0.5 * alpha * om_pecfac * mdot;
    l00 += alhsfac / relaxFac;
    l10 -= alhsfac;
    alhsfac = 0.5 * (mdot - fabs(mdot)) * pecfac * alphaUpw +
              0.5 * alpha * om_pecfac * mdot;
    l11 -= alhsfac / relaxFac;
    l01 += alhsfac;
    alhsfac = 0.5 * mdot * (pecfac * om_alphaUpw + om_pecfac * om_alpha);
    l00 += alhsfac / relaxFac;
    l01 += alhsfac;
    l10 -= alhsfac;
    l11 -= alhsfac / relaxFac;
    off[0] = side ? l10 : l01;
    acc[0] = side ? l11 : l00;
    acc[1] = side ? r1 : r0;
  }
};

// Node and edge records of the momentum edge; with PEC the Peclet factor
// of the edge is computed in the kernel (MomentumEdgePecletAlg fused in),
// which needs the density of both ends and the edge's id.
template <bool PEC>
struct MomNodeT {
  double x[3], u[3], g[9], visc, wall;
};
template <>
struct MomNodeT<true> {
  double x[3], u[3], g[9], visc, wall, rho;
};
template <bool PEC>
struct MomEdgeT {
  double av[3], mdot, pecfac;
};
template <>
struct MomEdgeT<true> {
  double av[3], mdot;
  int e;
};

template <bool PEC>
struct MomentumPhysT {
  static constexpr int NDOF = 3;
  static constexpr int kLanes = 2, kUnroll = 2, kUnroll2 = 2, kMinWaves = 2;
  // persistent: 0.577 vs 0.643 ms on the 128^3 box; with the Peclet factor
  // in the kernel (18 dwords of spill, 33 persistent) it is slower (0.95 vs
  // 0.76 ms per step)
  static constexpr bool kNonTemporal = true, kPersistent = !PEC;
  // 32 rows of 7 slots (Hex8 edge graph); 72 bytes per slot of LDS
  static constexpr int kSlotsSmall = 224, kSlotsLarge = 224;
  using Node = MomNodeT<PEC>;
  using Edge = MomEdgeT<PEC>;
  const double *coords, *vel, *dudx, *visc, *areav, *mdotf, *pecfacf, *wall;
  int64_t N, E;
  double alpha, alphaUpw, hoUpwind, relaxFacU, includeDivU;
  int useLimiter;
  // PEC only: density in, the edge Peclet number and factor out (written
  // once, by the L end's row)
  const double* rho;
  double *pecnum_out, *pecfac_out;
  PecletDev pec;

  __device__ void load(int n, Node& r) const
  {
    for (int d = 0; d < 3; ++d) {
      r.x[d] = coords[3 * (int64_t)n + d];
      r.u[d] = vel[3 * (int64_t)n + d];
    }
    for (int k = 0; k < 9; ++k) r.g[k] = dudx[9 * (int64_t)n + k];
    r.visc = visc[n];
    r.wall = wall[n];
    if constexpr (PEC) r.rho = rho[n];
  }

  __device__ void load_edge(int e, Edge& r) const
  {
    for (int d = 0; d < 3; ++d) r.av[d] = areav[3 * (int64_t)e + d];
    r.mdot = mdotf[e];
    if constexpr (PEC)
      r.e = e;
    else
      r.pecfac = pecfacf[e];
  }

  // MomentumEdgePecletAlg::execute (src/edge_kernels/MomentumEdgePecletAlg.C:
  // 74-101), the same operations in the same order as k_momentum_peclet
  __device__ double peclet_factor(const Edge& er, int side, const Node& L,
                                  const Node& R) const
  {
    if constexpr (PEC) {
      const double eps = 1.0e-16;
      const double diffIp = 0.5 * (L.visc / L.rho + R.visc / R.rho);
      double udotx = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double dxj = R.x[d] - L.x[d];
        udotx += 0.5 * dxj * (R.u[d] + L.u[d]);
      }
      const double pn = fabs(udotx) / (diffIp + eps);
      const double pf = pec(pn);
      if (side == 0) {
        pecnum_out[er.e] = pn;
        pecfac_out[er.e] = pf;
      }
      return pf;
    } else {
      return er.pecfac;
    }
  }

  __device__ void edge(const Edge& er, int side, const Node& L, const Node& R,
                       double* off, double* acc) const
  {
    // first, so that the densities it reads are dead before the blocks
    const double pecfac = peclet_factor(er, side, L, R);
    const double eps = 1.0e-16;
    const double om_alpha = 1.0 - alpha, om_alphaUpw = 1.0 - alphaUpw;
    const double* av = er.av;
    const double mdot = er.mdot;
    const double viscIp = 0.5 * (L.visc + R.visc);
    double dx[3];
    double axdx = 0.0, asq = 0.0;
    for (int d = 0; d < 3; ++d) {
      dx[d] = R.x[d] - L.x[d];
      asq += av[d] * av[d];
      axdx += av[d] * dx[d];
    }
    const double inv_axdx = 1.0 / axdx;
    double duL[3], duR[3];
    for (int i = 0; i < 3; ++i) {
      duL[i] = 0.0;
      duR[i] = 0.0;
      for (int j = 0; j < 3; ++j) {
        const double dxj = 0.5 * dx[j];
        duL[i] += dxj * L.g[3 * i + j];
        duR[i] += dxj * R.g[3 * i + j];
      }
    }
    double limitL[3] = {1.0, 1.0, 1.0}, limitR[3] = {1.0, 1.0, 1.0};
    if (useLimiter) {
      for (int d = 0; d < 3; ++d) {
        const double du = R.u[d] - L.u[d];
        limitL[d] = van_leer(4.0 * duL[d] - du, du, eps);
        limitR[d] = van_leer(4.0 * duR[d] - du, du, eps);
      }
    }
    const double om_pecfac = 1.0 - pecfac;
    double uIpL[3], uIpR[3];
    for (int d = 0; d < 3; ++d) {
      uIpL[d] = L.u[d] + duL[d] * hoUpwind * limitL[d] * 1.0;
      uIpR[d] = R.u[d] - duR[d] * hoUpwind * limitR[d] * 1.0;
    }
    double duidxj[3][3];
    for (int i = 0; i < 3; ++i) {
      const double dui = R.u[i] - L.u[i];
      double gjuidx = 0.0;
      for (int j = 0; j < 3; ++j)
        gjuidx += 0.5 * (R.g[3 * i + j] + L.g[3 * i + j]) * dx[j];
      for (int j = 0; j < 3; ++j) {
        const double gjui = 0.5 * (R.g[3 * i + j] + L.g[3 * i + j]);
        duidxj[i][j] = gjui + (dui - gjuidx) * av[j] * inv_axdx;
      }
    }
    const double dlhsfac = -viscIp * asq * inv_axdx;
    const double mask = fmin(L.wall, R.wall);
    const double alhsL = 0.5 * (mdot + fabs(mdot)) * pecfac * alphaUpw +
                         0.5 * alpha * om_pecfac * mdot;
    const double alhsR = 0.5 * (mdot - fabs(mdot)) * pecfac * alphaUpw +
                         0.5 * alpha * om_pecfac * mdot;
    const double alhsC =
      0.5 * mdot * (pecfac * om_alphaUpw + om_pecfac * om_alpha);
    const double r = relaxFacU;
    double divU = 0.0;
    for (int j = 0; j < 3; ++j) divU += duidxj[j][j];
    for (int i = 0; i < 3; ++i) {
      const double uiIp = 0.5 * (R.u[i] + L.u[i]);
      const double uiUpw = (mdot > 0.0)
                             ? (alphaUpw * uIpL[i] + om_alphaUpw * uiIp)
                             : (alphaUpw * uIpR[i] + om_alphaUpw * uiIp);
      const double uiHatL = (alpha * uIpL[i] + om_alpha * uiIp);
      const double uiHatR = (alpha * uIpR[i] + om_alpha * uiIp);
      const double uiCds = 0.5 * (uiHatL + uiHatR);
      const double adv_flux = mdot * (pecfac * uiUpw + om_pecfac * uiCds);
      double diff_flux = divU;
      diff_flux *= 2.0 / 3.0 * viscIp * av[i] * includeDivU;
      for (int j = 0; j < 3; ++j)
        diff_flux += -viscIp * (duidxj[i][j] + duidxj[j][i]) * av[j];
      const double total_flux = adv_flux + diff_flux * mask;
      // this side's rows: (rowL,rowL)/(rowL,rowR) or (rowR,rowR)/(rowR,rowL)
      double dii, oii;
      if (side == 0) {
        dii = alhsL / r;
        dii += alhsC / r;
        dii -= dlhsfac / r;
        oii = alhsR;
        oii += alhsC;
        oii += dlhsfac;
        acc[9 + i] = -total_flux;
      } else {
        dii = -(alhsR / r);
        dii -= alhsC / r;
        dii -= dlhsfac / r;
        oii = -alhsL;
        oii -= alhsC;
        oii += dlhsfac;
        acc[9 + i] = total_flux;
      }
      for (int j = 0; j < 3; ++j) {
        const double lhsfacNS = -viscIp * av[i] * av[j] * inv_axdx;
        if (j == i) {
          acc[3 * i + j] = dii - lhsfacNS / r;
          off[3 * i + j] = oii + lhsfacNS;
        } else {
          acc[3 * i + j] = -(lhsfacNS / r);
          off[3 * i + j] = lhsfacNS;
        }
      }
    }
  }
};

// ------------------------------------------------------------------ kernel

// Walks the off-diagonal slots of `row` (slots [s0, s1) less the diagonal
// slot sd): the row's k-th off-diagonal slot goes to lane k % G, so the G
// lanes of a row share its edges evenly.  For each slot s evaluates the
// edge, hands the (row, neighbour) block to put(s, off) and adds the edge's
// (row, row) block and rhs(row) contributions to acc.
// code_at(s) and col_at(s) read the slot's code and column.  U slots are
// taken per step: their neighbour and edge gathers are issued together
// before any of them is evaluated, and they are evaluated and summed in
// slot order, so U changes only the overlap of the loads.
template <class P, int G, int U, class Code, class Col, class Put>
__device__ inline void walk_row(const P& phys, const typename P::Node& me,
                                int s0, int s1, int sd, int g, Code code_at,
                                Col col_at, Put put, double* acc)
{
  constexpr int ND = P::NDOF, NB = ND * ND, NV = NB + ND;
  constexpr bool kCoRead = !(ND == 1 && G == 2);
  typename P::Node ot[U];
  typename P::Edge er[U];
  const int noff = s1 - s0 - 1;
#pragma clang loop unroll(disable)
  for (int k = g; k < noff; k += G * U) {
    // the u loops are unrolled so that ot[] and er[] stay in registers
    int sl[U], c[U], m[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kk = k + u * G;
      sl[u] = s0 + kk + (s0 + kk >= sd ? 1 : 0);
      if constexpr (kCoRead) {
        // the column read with the code (neither waits for the other)
        ok[u] = kk < noff;  // else past the row's end (true for u = 0)
        c[u] = ok[u] ? code_at(sl[u]) : 0;
        m[u] = ok[u] ? col_at(sl[u]) : 0;
      } else {
        // two lanes per 1-dof row (uneven mixed-mesh rows) measured faster
        // reading the column after the code, behind a test of the code
        c[u] = kk < noff ? code_at(sl[u]) : -1;
        ok[u] = c[u] >= 0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      if constexpr (!kCoRead) m[u] = col_at(sl[u]);
      phys.load(m[u], ot[u]);
      phys.load_edge(c[u] >> 1, er[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      double off[NB], con[NV];
      if (c[u] & 1)
        phys.edge(er[u], 1, ot[u], me, off, con);
      else
        phys.edge(er[u], 0, me, ot[u], off, con);
      put(sl[u], off);
      for (int v = 0; v < NV; ++v) acc[v] += con[v];
    }
  }
}

// Sums acc over the G lanes of a row by a fixed butterfly (all lanes of the
// wave take part), so the result does not depend on scheduling.
template <int G, int NV>
__device__ inline void row_butterfly(double* acc)
{
  for (int o = G / 2; o > 0; o >>= 1)
    for (int v = 0; v < NV; ++v) acc[v] += __shfl_xor(acc[v], o, G);
}

// Writes the (row, row) block and rhs(row) of a row, straight to HBM.
template <int ND>
__device__ inline void put_diag_rhs(double* __restrict__ dp,
                                    double* __restrict__ rp, const double* acc,
                                    int accumulate)
{
  constexpr int NB = ND * ND;
  if (accumulate) {
    for (int k = 0; k < NB; ++k) dp[k] += acc[k];
    for (int d = 0; d < ND; ++d) rp[d] += acc[NB + d];
  } else {
    for (int k = 0; k < NB; ++k) dp[k] = acc[k];
    for (int d = 0; d < ND; ++d) rp[d] = acc[NB + d];
  }
}

// Stages the ns <= SCAP codes and columns of a wave's slot range into LDS:
// every lane issues all of its (up to SCAP/64) loads of each array before it
// writes any of them, so the staging costs one memory round trip, not one
// per 64 slots (a loop that stores each loaded pair before loading the
// next waits for every load in turn).
template <int SCAP, bool NT>
__device__ inline void stage_slots(const int32_t* __restrict__ code,
                                   const int32_t* __restrict__ cols, int ns,
                                   int lane, int* code_s, int* col_s)
{
  constexpr int KS = (SCAP + kWave - 1) / kWave;
  int cv[KS], mv[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int i = lane + k * kWave;
    if (i < ns) {
      if (NT) {
        cv[k] = __builtin_nontemporal_load(code + i);
        mv[k] = __builtin_nontemporal_load(cols + i);
      } else {
        cv[k] = code[i];
        mv[k] = cols[i];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int i = lane + k * kWave;
    if (i < ns) {
      code_s[i] = cv[k];
      col_s[i] = mv[k];
    }
  }
}

// One task of the assembly: the rows of one wave.  G lanes own one row and
// a wave takes 64/G consecutive rows of a scheduled chunk (G waves cover the
// chunk, `half` says which); lane g of a row walks the row's off-diagonal
// slots g, g+G, ... (walk_row) and the G partial sums of the diagonal block
// and rhs are combined by a fixed butterfly.  The lanes of one instruction sit on
// consecutive rows, so their neighbour and edge gathers go in one stencil
// direction and stay compact.
//
// The slots of a wave's rows form one contiguous range [s_lo, s_hi).  The
// wave loads the range's codes and columns into LDS with coalesced loads,
// puts every evaluated block, the diagonal included, into an LDS copy of
// the range, and writes the range to HBM at the end as whole cache lines:
// no partially written lines linger in L2, where they would evict the node
// and edge data that neighbouring rows reuse.  A wave whose range holds
// more than SCAP slots writes straight to HBM instead.  NT marks the
// streamed code/column reads and value writes non-temporal.  The LDS
// arrays are the wave's own; a following task of the same wave writes its
// values only after its staging barrier, behind this task's reads.
//
// Rows flagged in `skip` (the skipped rows of the hypre backend, may be
// null) are evaluated but contribute nothing: their entries are written as
// 0 (or left as they are when accumulating).  When `xdiag` is set, the
// (0, 0) entry of every row's summed (row, row) block is added to
// xdiag[row] (NGPApplyCoeff::extract_diagonal), skipped rows included.
template <class P, int G, int U, int SCAP, bool NT>
__device__ inline void assemble_task(
  const P& phys, const int32_t* __restrict__ row_ptr,
  const int32_t* __restrict__ code, const int32_t* __restrict__ cols,
  const int32_t* __restrict__ diag, const SchedChunk& chunk, int half,
  double* __restrict__ vals, double* __restrict__ rhs, int accumulate,
  const uint8_t* __restrict__ skip, double* __restrict__ xdiag,
  int* code_s, int* col_s, double* val_s)
{
  constexpr int ND = P::NDOF, NB = ND * ND, NV = NB + ND;
  constexpr int RPW = kWave / G;
  static_assert(G == 1 || G == 2, "one or two waves per chunk");
  const int lane = threadIdx.x & (kWave - 1);
  const int rw0 = chunk.r0 + half * RPW;
  const int nr = min(RPW, chunk.r1 - rw0);
  if (nr <= 0) return;
  const int row = rw0 + lane / G;
  const int g = lane & (G - 1);
  const bool active = lane / G < nr;
  const int s_lo = G == 1 ? chunk.s0 : (half ? chunk.smid : chunk.s0);
  const int s_hi = G == 1 ? chunk.s1 : (half ? chunk.s1 : chunk.smid);
  const int ns = s_hi - s_lo;
  // the row's slot range, diagonal slot and own node fields do not depend
  // on the staged codes: load them now, so that they arrive with the
  // staging loads
  int r_s0 = 0, r_s1 = 0, r_diag = 0;
  bool skipped = false;
  typename P::Node me;
  if (active) {
    r_s0 = row_ptr[row];
    r_s1 = row_ptr[row + 1];
    r_diag = diag[row];
    if (skip) skipped = skip[row] != 0;
    phys.load(row, me);
  }
  double acc[NV];
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  if (ns > SCAP) {
    if (active)
      walk_row<P, G, U>(
        phys, me, r_s0, r_s1, r_diag, g, [&](int s) { return code[s]; },
        [&](int s) { return cols[s]; },
        [&](int s, const double* off) {
          double* vp = vals + (int64_t)s * NB;
          if (accumulate) {
            if (!skipped)
              for (int k = 0; k < NB; ++k) vp[k] += off[k];
          } else {
            for (int k = 0; k < NB; ++k) vp[k] = skipped ? 0.0 : off[k];
          }
        },
        acc);
    row_butterfly<G, NV>(acc);
    if (active && g == 0) {
      if (xdiag) xdiag[row] += acc[0];
      if (skipped)
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
      put_diag_rhs<ND>(vals + (int64_t)r_diag * NB, rhs + (int64_t)row * ND,
                       acc, accumulate);
    }
    return;
  }
  stage_slots<SCAP, NT>(code + s_lo, cols + s_lo, ns, lane, code_s, col_s);
  wave_lds_sync();
  if (active)
    walk_row<P, G, U>(
      phys, me, r_s0, r_s1, r_diag, g, [&](int s) { return code_s[s - s_lo]; },
      [&](int s) { return col_s[s - s_lo]; },
      [&](int s, const double* off) {
        double* vp = &val_s[(s - s_lo) * NB];
        for (int k = 0; k < NB; ++k) vp[k] = skipped ? 0.0 : off[k];
      },
      acc);
  row_butterfly<G, NV>(acc);
  if (active && g == 0) {
    if (xdiag) xdiag[row] += acc[0];
    if (skipped)
      for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    double* dp = &val_s[(r_diag - s_lo) * NB];
    for (int k = 0; k < NB; ++k) dp[k] = acc[k];
    double* rp = rhs + (int64_t)row * ND;
    if (accumulate) {
      for (int d = 0; d < ND; ++d) rp[d] += acc[NB + d];
    } else {
      for (int d = 0; d < ND; ++d) rp[d] = acc[NB + d];
    }
  }
  wave_lds_sync();
  double* dst = vals + (int64_t)s_lo * NB;
  for (int i = lane; i < ns * NB; i += kWave) {
    double v = val_s[i];
    if (accumulate) v += dst[i];
    if (NT)
      __builtin_nontemporal_store(v, dst + i);
    else
      dst[i] = v;
  }
}

// The assembly kernel: one task per wave, G waves per chunk.
template <class P, int G, int U, int SCAP, bool NT>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(P::kMinWaves))) void k_assemble(
  P phys, const int32_t* __restrict__ row_ptr,
  const int32_t* __restrict__ code, const int32_t* __restrict__ cols,
  const int32_t* __restrict__ diag, const SchedChunk* __restrict__ sched,
  int nchunks, double* __restrict__ vals, double* __restrict__ rhs,
  int accumulate, const uint8_t* __restrict__ skip, double* __restrict__ xdiag)
{
  constexpr int NB = P::NDOF * P::NDOF;
  constexpr int WPB = kBlock / kWave;
  __shared__ int code_s[WPB][SCAP];
  __shared__ int col_s[WPB][SCAP];
  __shared__ double val_s[WPB][SCAP * NB];
  const int wv = (int)(threadIdx.x / kWave);
  const int gw = xcd_block((int)gridDim.x) * WPB + wv;
  const int chunk_id = gw / G;
  if (chunk_id >= nchunks) return;
  const SchedChunk chunk = sched[chunk_id];
  assemble_task<P, G, U, SCAP, NT>(phys, row_ptr, code, cols, diag, chunk,
                                   gw % G, vals, rhs, accumulate, skip, xdiag,
                                   code_s[wv], col_s[wv], val_s[wv]);
}

// Persistent form: a grid of resident waves walks the tasks t = w, w + W,
// ... (W the grid's wave count, w the XCD-ordered wave id, task t = half
// t % G of chunk t / G), so the 8 XCDs work on neighbouring tasks at any
// time (giving each XCD its own contiguous eighth of the tasks, as
// k_assemble does, measured 2 % slower).  A wave reads the next task's
// 32-byte schedule record with a vector load (lanes 0-7, one dword each)
// issued before the current task's gathers, so it arrives with them: after
// a wave's first task no task waits for its schedule record.  Every wave
// leaves the loop when its next task is past the end.
template <class P, int G, int U, int SCAP, bool NT>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(P::kMinWaves))) void k_assemble_persistent(
  P phys, const int32_t* __restrict__ row_ptr,
  const int32_t* __restrict__ code, const int32_t* __restrict__ cols,
  const int32_t* __restrict__ diag, const SchedChunk* __restrict__ sched,
  int nchunks, double* __restrict__ vals, double* __restrict__ rhs,
  int accumulate, const uint8_t* __restrict__ skip, double* __restrict__ xdiag)
{
  constexpr int NB = P::NDOF * P::NDOF;
  constexpr int WPB = kBlock / kWave;
  __shared__ int code_s[WPB][SCAP];
  __shared__ int col_s[WPB][SCAP];
  __shared__ double val_s[WPB][SCAP * NB];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = (int)(threadIdx.x / kWave);
  const int nw = (int)gridDim.x * WPB;
  const int ntasks = nchunks * G;
  int t = xcd_block((int)gridDim.x) * WPB + wv;
  if (t >= ntasks) return;
  const int32_t* rec = reinterpret_cast<const int32_t*>(sched);
  int cur = lane < 8 ? rec[(int64_t)(t / G) * 8 + lane] : 0;
  for (;;) {
    const int tn = t + nw;
    const int nxt = (tn < ntasks && lane < 8) ? rec[(int64_t)(tn / G) * 8 + lane] : 0;
    SchedChunk chunk;
    chunk.r0 = __builtin_amdgcn_readlane(cur, 0);
    chunk.r1 = __builtin_amdgcn_readlane(cur, 1);
    chunk.s0 = __builtin_amdgcn_readlane(cur, 2);
    chunk.smid = __builtin_amdgcn_readlane(cur, 3);
    chunk.s1 = __builtin_amdgcn_readlane(cur, 4);
    assemble_task<P, G, U, SCAP, NT>(phys, row_ptr, code, cols, diag, chunk,
                                     t % G, vals, rhs, accumulate, skip, xdiag,
                                     code_s[wv], col_s[wv], val_s[wv]);
    if (tn >= ntasks) break;
    t = tn;
    cur = nxt;
  }
}

// MomentumEdgePecletAlg::execute (src/edge_kernels/MomentumEdgePecletAlg.C:
// 74-101): one edge per lane, in the row schedule's edge order when the
// graph is built (else in edge order), blocks in XCD-aware order.
__global__ __launch_bounds__(kBlock) void k_momentum_peclet(
  int64_t nedge, int64_t N, const int2* __restrict__ enodes,
  const int32_t* __restrict__ order, const double* __restrict__ coords,
  const double* __restrict__ vel, const double* __restrict__ rho,
  const double* __restrict__ visc, PecletDev pec, double* __restrict__ pecnum,
  double* __restrict__ pecfac)
{
  const int64_t i = (int64_t)xcd_block((int)gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= nedge) return;
  const int64_t e = order ? (int64_t)order[i] : i;
  const double eps = 1.0e-16;
  const int2 lr = enodes[e];
  const int nL = lr.x, nR = lr.y;
  const double diffIp = 0.5 * (visc[nL] / rho[nL] + visc[nR] / rho[nR]);
  double udotx = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double dxj = coords[3 * (int64_t)nR + d] - coords[3 * (int64_t)nL + d];
    udotx += 0.5 * dxj * (vel[3 * (int64_t)nR + d] + vel[3 * (int64_t)nL + d]);
  }
  const double pn = fabs(udotx) / (diffIp + eps);
  pecnum[e] = pn;
  pecfac[e] = pec(pn);
}

uint64_t spread3(uint64_t v)
{
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

// Centroid and extent of every chunk, and the bounding box of the nodes.
struct ChunkGeom {
  std::vector<double> cen[3], ext[3];
  double lo[3], hi[3];
};

ChunkGeom chunk_geometry(const std::vector<double>& x, int64_t N,
                         const std::vector<int2>& ch)
{
  ChunkGeom g;
  const int64_t nch = (int64_t)ch.size();
  for (int d = 0; d < 3; ++d) {
    g.lo[d] = x[d];
    g.hi[d] = x[d];
    for (int64_t n = 1; n < N; ++n) {
      g.lo[d] = std::min(g.lo[d], x[3 * n + d]);
      g.hi[d] = std::max(g.hi[d], x[3 * n + d]);
    }
    g.cen[d].resize(nch);
    g.ext[d].resize(nch);
    for (int64_t i = 0; i < nch; ++i) {
      double m = 0.0, a = x[3 * (int64_t)ch[i].x + d], b = a;
      for (int r = ch[i].x; r < ch[i].y; ++r) {
        const double v = x[3 * (int64_t)r + d];
        m += v;
        a = std::min(a, v);
        b = std::max(b, v);
      }
      g.cen[d][i] = m / (ch[i].y - ch[i].x);
      g.ext[d][i] = b - a;
    }
  }
  return g;
}

std::vector<int2> morton_order(const ChunkGeom& geo, const std::vector<int2>& ch)
{
  const int64_t nch = (int64_t)ch.size();
  std::vector<std::pair<uint64_t, int64_t>> key(nch);
  for (int64_t i = 0; i < nch; ++i) {
    uint64_t k = 0;
    for (int d = 0; d < 3; ++d) {
      const double span = geo.hi[d] > geo.lo[d] ? geo.hi[d] - geo.lo[d] : 1.0;
      const double t = (geo.cen[d][i] - geo.lo[d]) / span;
      const uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, t * 2097151.0));
      k |= spread3(q) << d;
    }
    key[i] = {k, i};
  }
  std::sort(key.begin(), key.end());
  std::vector<int2> o(nch);
  for (int64_t i = 0; i < nch; ++i) o[i] = ch[key[i].second];
  return o;
}

// Sweep order.  The chunks' own axis (the one along which they extend
// furthest) is the inner axis; of the other two the longer is swept in
// planes, and the remaining one is cut into strips of equal chunk count,
// about kPlaneChunks chunks of a strip per plane.  A plane of a strip is a
// run of chunks (sorted along the sweep axis) whose centroids lie within
// half a mesh spacing h = (box volume / nodes)^(1/3) of the run's first:
// on a grid this is one grid plane however the spacing is graded, and it
// never cuts a grid plane in two.  The schedule runs strip by
// strip, plane by plane, ordered inside a plane along the strip axis then
// the inner axis.  Neighbours in the sweep direction are then one
// plane-tile (~kPlaneChunks chunks) apart, neighbours along the strip axis
// a few chunks, so the node and edge data a plane shares with the next is
// still in L2 when it is read again; each XCD's contiguous share of the
// schedule is a strip or a part of one.
std::vector<int2> sweep_order(const ChunkGeom& geo, int64_t N, const std::vector<int2>& ch)
{
  constexpr double kPlaneChunks = 128.0;
  const int64_t nch = (int64_t)ch.size();
  double span[3], mext[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < 3; ++d) {
    span[d] = geo.hi[d] - geo.lo[d];
    for (int64_t i = 0; i < nch; ++i) mext[d] += geo.ext[d][i];
  }
  const int ai = (int)(std::max_element(mext, mext + 3) - mext);
  const int o1 = (ai + 1) % 3, o2 = (ai + 2) % 3;
  const int as = span[o1] > span[o2] ? o1 : o2, at = as == o1 ? o2 : o1;
  const double vol = std::max(span[0], 1e-300) * std::max(span[1], 1e-300) *
                     std::max(span[2], 1e-300);
  const double h = std::cbrt(vol / (double)N);
  const int64_t nplanes = std::max<int64_t>(1, std::llround(span[as] / h));
  const int64_t nstrips = std::max<int64_t>(
    1, std::llround((double)nch / (double)nplanes / kPlaneChunks));
  // strips: equal chunk counts along the strip axis
  std::vector<int64_t> idx(nch), strip(nch);
  for (int64_t i = 0; i < nch; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    return geo.cen[at][a] < geo.cen[at][b] || (geo.cen[at][a] == geo.cen[at][b] && a < b);
  });
  for (int64_t r = 0; r < nch; ++r) strip[idx[r]] = r * nstrips / nch;
  std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    if (strip[a] != strip[b]) return strip[a] < strip[b];
    return geo.cen[as][a] < geo.cen[as][b] || (geo.cen[as][a] == geo.cen[as][b] && a < b);
  });
  std::vector<int64_t> plane(nch);
  for (int64_t r = 0, p = 0; r < nch; ++r) {
    const int64_t i = idx[r];
    if (r > 0 && (strip[i] != strip[idx[r - 1]] ||
                  geo.cen[as][i] - geo.cen[as][idx[p]] > 0.5 * h))
      p = r;
    plane[i] = p;  // the rank of the plane's first chunk
  }
  std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    if (strip[a] != strip[b]) return strip[a] < strip[b];
    if (plane[a] != plane[b]) return plane[a] < plane[b];
    if (geo.cen[at][a] != geo.cen[at][b]) return geo.cen[at][a] < geo.cen[at][b];
    if (geo.cen[ai][a] != geo.cen[ai][b]) return geo.cen[ai][a] < geo.cen[ai][b];
    return a < b;
  });
  std::vector<int2> o(nch);
  for (int64_t i = 0; i < nch; ++i) o[i] = ch[idx[i]];
  return o;
}

// Fraction of (row, neighbour) pairs whose chunks lie more than kWindow
// positions apart in the schedule `ch`, over every 16th row: the pairs whose
// node data is unlikely to be still in L2 when the second row is assembled
// (kWindow ~ the chunks one XCD keeps in flight).
double far_pair_fraction(const Graph& g, const std::vector<int2>& ch)
{
  constexpr int64_t kWindow = 512;
  const int64_t nch = (int64_t)ch.size();
  std::vector<int64_t> pos(nch);
  for (int64_t p = 0; p < nch; ++p) pos[ch[p].x / kRowsPerWave] = p;
  int64_t far = 0, tot = 0;
  for (int64_t r = 0; r < g.nnode; r += 16) {
    const int64_t pr = pos[r / kRowsPerWave];
    for (int32_t s = g.h_row_ptr[r]; s < g.h_row_ptr[r + 1]; ++s) {
      const int64_t pc = pos[g.h_cols[s] / kRowsPerWave];
      far += std::llabs(pr - pc) > kWindow;
      ++tot;
    }
  }
  return tot ? (double)far / (double)tot : 0.0;
}

// p-quantile of the slot counts of consecutive row groups of `rows` rows
int64_t slot_quantile(const Graph& g, int64_t rows, double q)
{
  const int64_t ng = (g.nnode + rows - 1) / rows;
  if (ng == 0) return 0;
  std::vector<int64_t> v(ng);
  for (int64_t i = 0; i < ng; ++i)
    v[i] = g.h_row_ptr[std::min<int64_t>(g.nnode, (i + 1) * rows)] -
           g.h_row_ptr[i * rows];
  const int64_t k = std::min<int64_t>(ng - 1, (int64_t)(q * (double)ng));
  std::nth_element(v.begin(), v.begin() + k, v.end());
  return v[k];
}

// Chunks of <= kRowsPerWave consecutive rows in one of three orders,
// whichever leaves the fewest neighbour pairs far apart in the schedule:
//   natural  the row order: compact for bandwidth-reducing (RCM)
//            numberings, whose level sets are local;
//   sweep    strips along one axis, each swept plane by plane along another
//            (see sweep_order): for lexicographically numbered grids, whose
//            chunks are segments of grid lines;
//   morton   the Morton code of the chunk centroids.
// NW_ROW_ORDER=natural|sweep|morton forces one.
int build_schedule(Ctx* c)
{
  Graph& g = c->g;
  const int64_t N = g.nnode;
  const int64_t nch = (N + kRowsPerWave - 1) / kRowsPerWave;
  std::vector<int2> ch(nch);
  for (int64_t i = 0; i < nch; ++i)
    ch[i] = make_int2((int)(i * kRowsPerWave),
                      (int)std::min<int64_t>(N, (i + 1) * kRowsPerWave));
  const char* env = getenv("NW_ROW_ORDER");
  const std::string force = env ? std::string(env) : std::string();
  if (force != "natural" && c->field[NW_F_COORDINATES] && nch > 1) {
    std::vector<double> x(3 * N);  // x[3 n + d]
    NW_HIP(hipMemcpyAsync(x.data(), c->field[NW_F_COORDINATES],
                          sizeof(double) * 3 * N, hipMemcpyDeviceToHost,
                          c->stream));
    NW_HIP(hipStreamSynchronize(c->stream));
    const ChunkGeom geo = chunk_geometry(x, N, ch);
    std::vector<int2> mo = morton_order(geo, ch), sw = sweep_order(geo, N, ch);
    if (force == "morton") {
      ch.swap(mo);
    } else if (force == "sweep") {
      ch.swap(sw);
    } else {
      const double fn = far_pair_fraction(g, ch), fs = far_pair_fraction(g, sw),
                   fm = far_pair_fraction(g, mo);
      if (fs < fn && fs <= fm)
        ch.swap(sw);
      else if (fm < fn && fm < fs)
        ch.swap(mo);
    }
  }
  // the edges in the same order: by chunk, row, then the row's edges as
  // node[0] (each edge once), for the edge-parallel kernels
  {
    const int64_t E = c->nedge;
    const int32_t* ed = c->h_edges.data();
    std::vector<int32_t> start(N + 1, 0), idx(std::max<int64_t>(E, 1));
    for (int64_t e = 0; e < E; ++e) start[ed[2 * e] + 1]++;
    for (int64_t n = 0; n < N; ++n) start[n + 1] += start[n];
    {
      std::vector<int32_t> pos(start.begin(), start.end() - 1);
      for (int64_t e = 0; e < E; ++e) idx[pos[ed[2 * e]]++] = (int32_t)e;
    }
    std::vector<int32_t> order;
    order.reserve(E);
    for (int64_t i = 0; i < nch; ++i)
      for (int r = ch[i].x; r < ch[i].y; ++r)
        for (int32_t k = start[r]; k < start[r + 1]; ++k) order.push_back(idx[k]);
    if (!g.edge_order)
      NW_HIP(hipMalloc(&g.edge_order, sizeof(int32_t) * std::max<int64_t>(E, 1)));
    NW_HIP(hipMemcpyAsync(g.edge_order, order.data(), sizeof(int32_t) * E,
                          hipMemcpyHostToDevice, c->stream));
    NW_HIP(hipStreamSynchronize(c->stream));
  }
  std::vector<SchedChunk> rec(nch);
  for (int64_t i = 0; i < nch; ++i) {
    const int r0 = ch[i].x, r1 = ch[i].y;
    rec[i] = SchedChunk{r0, r1, g.h_row_ptr[r0],
                        g.h_row_ptr[std::min(r1, r0 + kRowsPerWave / 2)],
                        g.h_row_ptr[r1], {0, 0, 0}};
  }
  if (!g.sched)
    NW_HIP(hipMalloc(&g.sched, sizeof(SchedChunk) * std::max<int64_t>(nch, 1)));
  NW_HIP(hipMemcpyAsync(g.sched, rec.data(), sizeof(SchedChunk) * nch,
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  g.nchunks = nch;
  // the slot counts that 99 % of the chunks (and of the half chunks) stay
  // within: the kernels size their LDS staging for them, the rest take the
  // direct path
  g.chunk_slots_p99 = slot_quantile(g, kRowsPerWave, 0.99);
  g.half_chunk_slots_p99 = slot_quantile(g, kRowsPerWave / 2, 0.99);
  g.sched_valid = true;
  return NW_OK;
}

template <class P, int G, int SCAP>
void launch_assemble(Ctx* c, const P& phys, int accumulate, double* xdiag)
{
  constexpr int kU = G == 1 ? P::kUnroll : P::kUnroll2;
  const Graph& g = c->g;
  const uint8_t* skip = active_skip_mask(g);
  const int64_t wpb = kBlock / kWave;
  int64_t blocks = (g.nchunks * G + wpb - 1) / wpb;
  if constexpr (P::kPersistent) {
    // as many blocks as the device holds at once (queried once; contexts
    // on other host threads may ask at the same time)
    static std::atomic<int> cached{0};
    int resident = cached.load(std::memory_order_relaxed);
    if (resident == 0) {
      int ncu = 0, per_cu = 0;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                c->device) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_assemble_persistent<P, G, kU, SCAP, P::kNonTemporal>,
            kBlock, 0) != hipSuccess)
        ncu = per_cu = 0;
      resident = ncu * per_cu > 0 ? ncu * per_cu : -1;
      cached.store(resident, std::memory_order_relaxed);
    }
    if (resident > 0) {
      // a multiple of the XCD count, all resident (blocks never wait for
      // each other, so the grid's size is a matter of speed only)
      blocks = std::min<int64_t>(blocks, std::max(8, resident / 8 * 8));
      blocks = (blocks + 7) / 8 * 8;
      k_assemble_persistent<P, G, kU, SCAP, P::kNonTemporal>
        <<<(unsigned)blocks, kBlock, 0, c->stream>>>(
          phys, g.row_ptr, g.code, g.cols, g.diag, g.sched, (int)g.nchunks,
          g.vals, g.rhs, accumulate, skip, xdiag);
      return;
    }
  }
  blocks = (blocks + 7) / 8 * 8;  // a multiple of the XCD count
  k_assemble<P, G, kU, SCAP, P::kNonTemporal>
    <<<(unsigned)blocks, kBlock, 0, c->stream>>>(phys, g.row_ptr, g.code, g.cols,
                                                 g.diag, g.sched, (int)g.nchunks,
                                                 g.vals, g.rhs, accumulate, skip,
                                                 xdiag);
}

template <class P>
int launch_rows(Ctx* c, const P& phys, int accumulate, double* xdiag = nullptr)
{
  Graph& g = c->g;
  if (!g.sched_valid) {
    int rc = build_schedule(c);
    if (rc) return rc;
  }
  if (g.nchunks == 0) return NW_OK;
  if constexpr (P::NDOF == 1) {
    // one lane per row when 99 % of the 64-row chunks fit kSlotsSmall slots
    // (Hex8 graphs); else two lanes per row, which also balances rows of
    // uneven length (mixed meshes), with the LDS the half chunks need
    if (g.chunk_slots_p99 <= P::kSlotsSmall)
      launch_assemble<P, 1, P::kSlotsSmall>(c, phys, accumulate, xdiag);
    else if (g.half_chunk_slots_p99 <= P::kSlotsSmall)
      launch_assemble<P, 2, P::kSlotsSmall>(c, phys, accumulate, xdiag);
    else
      launch_assemble<P, 2, P::kSlotsLarge>(c, phys, accumulate, xdiag);
  } else {
    launch_assemble<P, P::kLanes, P::kSlotsSmall>(c, phys, accumulate, xdiag);
  }
  NW_LAUNCH_CHECK();
  return NW_OK;
}

#define NW_GET(f, var)                        \
  const double* var = nullptr;                \
  do {                                        \
    int rc_ = require_field(c, (f), &(var));  \
    if (rc_) return rc_;                      \
  } while (0)

}  // namespace

int ensure_schedule(Ctx* c) { return c->g.sched_valid ? NW_OK : build_schedule(c); }

}  // namespace nw

using namespace nw;

extern "C" {

int nw_assemble_continuity(nw_ctx* c, const nw_continuity_opts* o)
{
  NW_REQUIRE(c && o, NW_E_ARG, "null argument");
  NW_REQUIRE(c->g.ndof == 1, NW_E_STATE, "continuity needs a 1-dof graph");
  NW_REQUIRE(o->gamma1 != 0.0 && o->dt != 0.0, NW_E_ARG, "dt/gamma1 zero");
  NW_GET(NW_F_COORDINATES, coords);
  NW_GET(NW_F_VELOCITY, vel);
  NW_GET(NW_F_DPDX, dpdx);
  NW_GET(NW_F_DENSITY, rho);
  NW_GET(NW_F_PRESSURE, p);
  NW_GET(NW_F_UDIAG, udiag);
  NW_GET(NW_E_AREA_VECTOR, areav);
  if (o->balanced_buoyancy) {
    NW_GET(NW_F_BUOYANCY_SOURCE, src);
    NW_GET(NW_F_BUOYANCY_SOURCE_MASK, mask);
    ContinuityPhysT<true> P{coords, vel, dpdx, rho, p, udiag, areav, c->nnode,
                            c->nedge, o->noc, o->dt / o->gamma1,
                            o->interp_together, o->incompressible, src, mask,
                            {o->gravity[0], o->gravity[1], o->gravity[2]}};
    return check_system_finite(c, launch_rows(c, P, o->accumulate),
                               "ContinuityEdgeSolverAlg");
  }
  ContinuityPhysT<false> P{coords, vel, dpdx, rho, p, udiag, areav, c->nnode,
                           c->nedge, o->noc, o->dt / o->gamma1,
                           o->interp_together, o->incompressible, nullptr,
                           nullptr, {0.0, 0.0, 0.0}};
  return check_system_finite(c, launch_rows(c, P, o->accumulate),
                             "ContinuityEdgeSolverAlg");
}

int nw_assemble_scalar(nw_ctx* c, const nw_scalar_opts* o)
{
  NW_REQUIRE(c && o, NW_E_ARG, "null argument");
  NW_REQUIRE(c->g.ndof == 1, NW_E_STATE, "scalar needs a 1-dof graph");
  NW_GET(NW_F_COORDINATES, coords);
  NW_GET(NW_F_VELOCITY, vel);
  NW_GET(NW_F_SCALAR_Q, q);
  NW_GET(NW_F_DQDX, dqdx);
  NW_GET(NW_F_DENSITY, rho);
  NW_GET(NW_F_DIFF_FLUX_COEFF, dflux);
  NW_GET(NW_E_AREA_VECTOR, areav);
  NW_GET(NW_E_MASS_FLOW_RATE, mdot);
  ScalarPhys P{coords, vel, q, dqdx, rho, dflux, areav, mdot, c->nnode, c->nedge,
               o->alpha, o->alpha_upw, o->upw, o->relax, o->use_limiter,
               PecletDev{o->peclet.form, o->peclet.hybrid, o->peclet.c1,
                         o->peclet.c2}};
  return check_system_finite(c, launch_rows(c, P, o->accumulate),
                             "ScalarEdgeSolverAlg");
}

int nw_assemble_momentum(nw_ctx* c, const nw_momentum_opts* o)
{
  NW_REQUIRE(c && o, NW_E_ARG, "null argument");
  NW_REQUIRE(c->g.ndof == 3, NW_E_STATE, "momentum needs a 3-dof graph");
  NW_GET(NW_F_COORDINATES, coords);
  NW_GET(NW_F_VELOCITY, vel);
  NW_GET(NW_F_DUDX, dudx);
  NW_GET(NW_F_VISCOSITY, visc);
  NW_GET(NW_F_WALL_MASK, wall);
  NW_GET(NW_E_AREA_VECTOR, areav);
  NW_GET(NW_E_MASS_FLOW_RATE, mdot);
  NW_GET(NW_E_PECLET_FACTOR, pecfac);
  MomentumPhysT<false> P{coords, vel, dudx, visc, areav, mdot, pecfac, wall,
                         c->nnode, c->nedge, o->alpha, o->alpha_upw, o->upw,
                         o->relax, o->include_divU, o->use_limiter,
                         nullptr, nullptr, nullptr, PecletDev{}};
  double* xdiag = nullptr;
  if (o->extract_diagonal) {
    int rc = ensure_field(c, NW_F_UDIAG, &xdiag);
    if (rc) return rc;
  }
  return check_system_finite(c, launch_rows(c, P, o->accumulate, xdiag),
                             "MomentumEdgeSolverAlg");
}

int nw_assemble_momentum_with_peclet(nw_ctx* c, const nw_momentum_opts* o,
                                     const nw_peclet* pf)
{
  NW_REQUIRE(c && o && pf, NW_E_ARG, "null argument");
  NW_REQUIRE(c->g.ndof == 3, NW_E_STATE, "momentum needs a 3-dof graph");
  NW_GET(NW_F_COORDINATES, coords);
  NW_GET(NW_F_VELOCITY, vel);
  NW_GET(NW_F_DUDX, dudx);
  NW_GET(NW_F_VISCOSITY, visc);
  NW_GET(NW_F_WALL_MASK, wall);
  NW_GET(NW_F_DENSITY, rho);
  NW_GET(NW_E_AREA_VECTOR, areav);
  NW_GET(NW_E_MASS_FLOW_RATE, mdot);
  double *pn = nullptr, *pfac = nullptr;
  int rc = ensure_field(c, NW_E_PECLET_NUMBER, &pn);
  if (rc) return rc;
  rc = ensure_field(c, NW_E_PECLET_FACTOR, &pfac);
  if (rc) return rc;
  MomentumPhysT<true> P{coords, vel, dudx, visc, areav, mdot, nullptr, wall,
                        c->nnode, c->nedge, o->alpha, o->alpha_upw, o->upw,
                        o->relax, o->include_divU, o->use_limiter,
                        rho, pn, pfac,
                        PecletDev{pf->form, pf->hybrid, pf->c1, pf->c2}};
  double* xdiag = nullptr;
  if (o->extract_diagonal) {
    rc = ensure_field(c, NW_F_UDIAG, &xdiag);
    if (rc) return rc;
  }
  return check_system_finite(c, launch_rows(c, P, o->accumulate, xdiag),
                             "MomentumEdgeSolverAlg");
}

int nw_momentum_edge_peclet(nw_ctx* c, const nw_peclet* pf)
{
  NW_REQUIRE(c && pf, NW_E_ARG, "null argument");
  NW_REQUIRE(c->nnode > 0, NW_E_STATE, "mesh not set");
  NW_GET(NW_F_COORDINATES, coords);
  NW_GET(NW_F_VELOCITY, vel);
  NW_GET(NW_F_DENSITY, rho);
  NW_GET(NW_F_VISCOSITY, visc);
  double *pn = nullptr, *pfac = nullptr;
  int rc = ensure_field(c, NW_E_PECLET_NUMBER, &pn);
  if (rc) return rc;
  rc = ensure_field(c, NW_E_PECLET_FACTOR, &pfac);
  if (rc) return rc;
  if (c->nedge == 0) return NW_OK;
  // with the edge graph built, the edges go in its schedule's order and
  // each XCD takes one contiguous stretch of it (as the row kernels), so
  // the node data of neighbouring edges is shared in one L2
  const int32_t* order = nullptr;
  if (c->g.row_ptr && c->g.nnode == c->nnode) {
    if ((rc = ensure_schedule(c))) return rc;
    order = c->g.edge_order;
  }
  const unsigned nb = (grid_for(c->nedge, kBlock) + 7) / 8 * 8;
  k_momentum_peclet<<<nb, kBlock, 0, c->stream>>>(
    c->nedge, c->nnode, reinterpret_cast<const int2*>(c->edges), order, coords,
    vel, rho, visc, PecletDev{pf->form, pf->hybrid, pf->c1, pf->c2}, pn, pfac);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

}  // extern "C"
