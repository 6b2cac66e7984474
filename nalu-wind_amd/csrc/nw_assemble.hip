// nw_assemble.hip -- the interior edge solver algorithms on gfx950.
//
// Layout.  The linear system is a block CSR over nodes: row n holds the
// slots of n and of its edge neighbours, sorted by column (the sorted Tpetra
// graph).  Every slot carries a code: the diagonal slot of row n is
// -(n+1); the slot (n, m) reached through edge e = (L, R) is 2e + side with
// side = 0 when n == L and 1 when n == R.
//
// Kernel shape ("row owner, slot parallel").  A wavefront owns 64
// consecutive rows and walks their slots 64 at a time, one slot per lane:
//   * an off-diagonal slot (n, m) recomputes its edge's flux in registers
//     from the edge's area vector and both nodes' fields, exactly as the
//     reference lambda does for that edge, and stores the (n, m) block of
//     the edge's LHS straight into its own slot (coalesced: slots are
//     consecutive in memory);
//   * the edge's contributions to the diagonal block of n and to rhs(n)
//     are parked in LDS, and the last lane of each row sums the row's slots
//     in column order and writes the diagonal block and the rhs once.
// Each edge is therefore evaluated twice (once from each end), the matrix
// and rhs are written exactly once, and no atomics are used: results are
// bitwise reproducible run to run.  Rows may span chunks; the partial sums
// are carried in registers (wave-uniform) to the next chunk.
//
// The physics of each slot restates the reference edge kernels:
//   ScalarEdgeSolverAlg.C:87-205, MomentumEdgeSolverAlg.C:107-312,
//   ContinuityEdgeSolverAlg.C:111-194 (src/edge_kernels/).
#include <cmath>

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
  // include/edge_kernels/EdgeKernelUtils.h:18-24
  return (2.0 * (dqm * dqp + fabs(dqm * dqp))) /
         ((dqm + dqp) * (dqm + dqp) + eps);
}

// ------------------------------------------------------------------ physics
// Each physics type computes, for slot (row n = node[side] of edge e):
//   off[NB]  the (n, other) block of the edge LHS, row-major ndof x ndof
//   acc[NV]  the (n, n) block contribution (NB values) then rhs(n) (NDOF)

struct ContinuityPhys {
  static constexpr int NDOF = 1;
  const double *coords, *vel, *dpdx, *rho, *p, *udiag, *areav;
  double nocFac, tauScale, interpTogether, incompressible;

  __device__ void operator()(int e, int side, int nL, int nR, double* off,
                             double* acc) const
  {
    const double om_interpTogether = 1.0 - interpTogether;
    const double om_incomp = 1.0 - incompressible;
    const double av[3] = {areav[3 * e], areav[3 * e + 1], areav[3 * e + 2]};
    const double pressureL = p[nL], pressureR = p[nR];
    const double densityL = rho[nL], densityR = rho[nR];
    const double udiagL = udiag[nL], udiagR = udiag[nR];
    const double projTimeScale = 0.5 * (1.0 / udiagL + 1.0 / udiagR);
    const double rhoIp = 0.5 * (densityL + densityR);
    const double denScale = (1.0 / rhoIp) * incompressible + om_incomp;
    double dx[3];
    double axdx = 0.0, asq = 0.0;
    for (int d = 0; d < 3; ++d) {
      dx[d] = coords[3 * nR + d] - coords[3 * nL + d];
      asq += av[d] * av[d];
      axdx += av[d] * dx[d];
    }
    const double inv_axdx = 1.0 / axdx;
    double tmdot = -projTimeScale * (pressureR - pressureL) * asq * inv_axdx;
    for (int d = 0; d < 3; ++d) {
      const double kxj = av[d] - asq * inv_axdx * dx[d];
      const double uR = vel[3 * nR + d], uL = vel[3 * nL + d];
      const double rhoUjIp = 0.5 * (densityR * uR + densityL * uL);
      const double ujIp = 0.5 * (uR + uL);
      const double GjIp =
        0.5 * (dpdx[3 * nR + d] / udiagR + dpdx[3 * nL + d] / udiagL);
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

struct ScalarPhys {
  static constexpr int NDOF = 1;
  const double *coords, *vrtm, *q, *dqdx, *rho, *dflux, *areav, *mdotf;
  double alpha, alphaUpw, hoUpwind, relaxFac;
  int useLimiter;
  PecletDev pec;

  __device__ void operator()(int e, int side, int nL, int nR, double* off,
                             double* acc) const
  {
    const double eps = 1.0e-16;
    const double om_alpha = 1.0 - alpha, om_alphaUpw = 1.0 - alphaUpw;
    const double av[3] = {areav[3 * e], areav[3 * e + 1], areav[3 * e + 2]};
    const double mdot = mdotf[e];
    const double densityL = rho[nL], densityR = rho[nR];
    const double qNp1L = q[nL], qNp1R = q[nR];
    const double viscosityL = dflux[nL], viscosityR = dflux[nR];
    const double viscIp = 0.5 * (viscosityL + viscosityR);
    const double diffIp = 0.5 * (viscosityL / densityL + viscosityR / densityR);
    double dx[3];
    double axdx = 0.0, asq = 0.0, udotx = 0.0;
    for (int d = 0; d < 3; ++d) {
      dx[d] = coords[3 * nR + d] - coords[3 * nL + d];
      asq += av[d] * av[d];
      axdx += av[d] * dx[d];
      udotx += 0.5 * dx[d] * (vrtm[3 * nR + d] + vrtm[3 * nL + d]);
    }
    const double inv_axdx = 1.0 / axdx;
    double dqL = 0.0, dqR = 0.0, nonOrth = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double gL = dqdx[3 * nL + d], gR = dqdx[3 * nR + d];
      dqL += 0.5 * dx[d] * gL;
      dqR += 0.5 * dx[d] * gR;
      const double kxj = av[d] - asq * inv_axdx * dx[d];
      nonOrth += -viscIp * kxj * 0.5 * (gR + gL);
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

struct MomentumPhys {
  static constexpr int NDOF = 3;
  const double *coords, *vel, *dudx, *visc, *areav, *mdotf, *pecfacf, *wall;
  double alpha, alphaUpw, hoUpwind, relaxFacU, includeDivU;
  int useLimiter;

  __device__ void operator()(int e, int side, int nL, int nR, double* off,
                             double* acc) const
  {
    const double eps = 1.0e-16;
    const double om_alpha = 1.0 - alpha, om_alphaUpw = 1.0 - alphaUpw;
    const double av[3] = {areav[3 * e], areav[3 * e + 1], areav[3 * e + 2]};
    const double mdot = mdotf[e];
    const double viscIp = 0.5 * (visc[nL] + visc[nR]);
    double dx[3], uL[3], uR[3];
    double axdx = 0.0, asq = 0.0;
    for (int d = 0; d < 3; ++d) {
      dx[d] = coords[3 * nR + d] - coords[3 * nL + d];
      uL[d] = vel[3 * nL + d];
      uR[d] = vel[3 * nR + d];
      asq += av[d] * av[d];
      axdx += av[d] * dx[d];
    }
    const double inv_axdx = 1.0 / axdx;
    double gL[9], gR[9];
    for (int k = 0; k < 9; ++k) {
      gL[k] = dudx[9 * (int64_t)nL + k];
      gR[k] = dudx[9 * (int64_t)nR + k];
    }
    double duL[3], duR[3];
    for (int i = 0; i < 3; ++i) {
      duL[i] = 0.0;
      duR[i] = 0.0;
      for (int j = 0; j < 3; ++j) {
        const double dxj = 0.5 * dx[j];
        duL[i] += dxj * gL[3 * i + j];
        duR[i] += dxj * gR[3 * i + j];
      }
    }
    double limitL[3] = {1.0, 1.0, 1.0}, limitR[3] = {1.0, 1.0, 1.0};
    if (useLimiter) {
      for (int d = 0; d < 3; ++d) {
        const double du = uR[d] - uL[d];
        limitL[d] = van_leer(4.0 * duL[d] - du, du, eps);
        limitR[d] = van_leer(4.0 * duR[d] - du, du, eps);
      }
    }
    const double pecfac = pecfacf[e];
    const double om_pecfac = 1.0 - pecfac;
    double uIpL[3], uIpR[3];
    for (int d = 0; d < 3; ++d) {
      uIpL[d] = uL[d] + duL[d] * hoUpwind * limitL[d] * 1.0;
      uIpR[d] = uR[d] - duR[d] * hoUpwind * limitR[d] * 1.0;
    }
    double duidxj[3][3];
    for (int i = 0; i < 3; ++i) {
      const double dui = uR[i] - uL[i];
      double gjuidx = 0.0;
      for (int j = 0; j < 3; ++j)
        gjuidx += 0.5 * (gR[3 * i + j] + gL[3 * i + j]) * dx[j];
      for (int j = 0; j < 3; ++j) {
        const double gjui = 0.5 * (gR[3 * i + j] + gL[3 * i + j]);
        duidxj[i][j] = gjui + (dui - gjuidx) * av[j] * inv_axdx;
      }
    }
    const double dlhsfac = -viscIp * asq * inv_axdx;
    const double mask = fmin(wall[nL], wall[nR]);
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
      const double uiIp = 0.5 * (uR[i] + uL[i]);
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
      // rows of this side: (rowL,rowL)/(rowL,rowR) or (rowR,rowR)/(rowR,rowL)
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

template <class P>
__global__ __launch_bounds__(kBlock) void k_assemble_rows(
  P phys, const int32_t* __restrict__ row_ptr,
  const int32_t* __restrict__ code, const int2* __restrict__ enodes,
  double* __restrict__ vals, double* __restrict__ rhs, int nrows,
  int accumulate)
{
  constexpr int ND = P::NDOF, NB = ND * ND, NV = NB + ND;
  __shared__ double sh[kBlock / kWave][NV][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  const int gw = blockIdx.x * (kBlock / kWave) + wv;
  const int r0 = gw * kRowsPerWave;
  if (r0 >= nrows) return;
  const int r1 = min(r0 + kRowsPerWave, nrows);
  const int s0 = row_ptr[r0], s1 = row_ptr[r1];

  double carry[NV];
  for (int v = 0; v < NV; ++v) carry[v] = 0.0;
  int crow = -1, cdslot = -1;

  for (int base = s0; base < s1; base += kWave) {
    const int s = base + lane;
    const bool act = s < s1;
    int row = INT_MAX;
    bool isdiag = false;
    double acc[NV];
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    if (act) {
      const int c = code[s];
      if (c < 0) {
        row = -c - 1;
        isdiag = true;
      } else {
        const int e = c >> 1, side = c & 1;
        const int2 lr = enodes[e];
        row = side ? lr.y : lr.x;
        double off[NB];
        phys(e, side, lr.x, lr.y, off, acc);
        double* vp = vals + (int64_t)s * NB;
        if (accumulate) {
          for (int k = 0; k < NB; ++k) vp[k] += off[k];
        } else {
          for (int k = 0; k < NB; ++k) vp[k] = off[k];
        }
      }
    }
    for (int v = 0; v < NV; ++v) sh[wv][v][lane] = acc[v];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const int prow = __shfl_up(row, 1);
    const int nrow = __shfl_down(row, 1);
    const bool head = act && (lane == 0 || prow != row);
    const bool tail = act && (lane == kWave - 1 || nrow != row);
    const uint64_t heads = __ballot(head);
    const uint64_t diags = __ballot(isdiag);

    double sum[NV];
    for (int v = 0; v < NV; ++v) sum[v] = 0.0;
    int dslot = -1;
    bool open = false;  // row continues into the next chunk
    if (tail) {
      const uint64_t upto =
        (lane == kWave - 1) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
      const uint64_t hm = heads & upto;
      const int start = 63 - __builtin_clzll(hm);
      const bool cont = (start == 0) && (row == crow);
      if (cont)
        for (int v = 0; v < NV; ++v) sum[v] = carry[v];
      for (int t = start; t <= lane; ++t)
        for (int v = 0; v < NV; ++v) sum[v] += sh[wv][v][t];
      const uint64_t dm = diags & upto & ~((1ull << start) - 1ull);
      dslot = dm ? base + __builtin_ctzll(dm) : (cont ? cdslot : -1);
      open = (lane == kWave - 1) && (s + 1 < s1) && (row_ptr[row + 1] > s + 1);
      if (!open && dslot >= 0) {
        double* dp = vals + (int64_t)dslot * NB;
        double* rp = rhs + (int64_t)row * ND;
        if (accumulate) {
          for (int k = 0; k < NB; ++k) dp[k] += sum[k];
          for (int d = 0; d < ND; ++d) rp[d] += sum[NB + d];
        } else {
          for (int k = 0; k < NB; ++k) dp[k] = sum[k];
          for (int d = 0; d < ND; ++d) rp[d] = sum[NB + d];
        }
      }
    }
    // carry the open row of lane 63 (wave-uniform)
    const int open63 = __shfl((int)open, kWave - 1);
    if (open63) {
      crow = __shfl(row, kWave - 1);
      cdslot = __shfl(dslot, kWave - 1);
      for (int v = 0; v < NV; ++v) carry[v] = __shfl(sum[v], kWave - 1);
    } else {
      crow = -1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// MomentumEdgePecletAlg::execute (src/edge_kernels/MomentumEdgePecletAlg.C:
// 74-101): one edge per lane, edge-ordered (coalesced) writes.
__global__ __launch_bounds__(kBlock) void k_momentum_peclet(
  int64_t nedge, const int2* __restrict__ enodes,
  const double* __restrict__ coords, const double* __restrict__ vel,
  const double* __restrict__ rho, const double* __restrict__ visc,
  PecletDev pec, double* __restrict__ pecnum, double* __restrict__ pecfac)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nedge) return;
  const double eps = 1.0e-16;
  const int2 lr = enodes[e];
  const int nL = lr.x, nR = lr.y;
  const double diffIp = 0.5 * (visc[nL] / rho[nL] + visc[nR] / rho[nR]);
  double udotx = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double dxj = coords[3 * nR + d] - coords[3 * nL + d];
    udotx += 0.5 * dxj * (vel[3 * nR + d] + vel[3 * nL + d]);
  }
  const double pn = fabs(udotx) / (diffIp + eps);
  pecnum[e] = pn;
  pecfac[e] = pec(pn);
}

template <class P>
int launch_rows(Ctx* c, const P& phys, int accumulate)
{
  Graph& g = c->g;
  const int64_t nwaves = (g.nnode + kRowsPerWave - 1) / kRowsPerWave;
  const unsigned blocks = (unsigned)((nwaves + 3) / 4);
  if (blocks == 0) return NW_OK;
  k_assemble_rows<P><<<blocks, kBlock, 0, c->stream>>>(
    phys, g.row_ptr, g.code, reinterpret_cast<const int2*>(c->edges), g.vals,
    g.rhs, (int)g.nnode, accumulate);
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
  ContinuityPhys P{coords, vel, dpdx, rho, p, udiag, areav, o->noc,
                   o->dt / o->gamma1, o->interp_together, o->incompressible};
  return launch_rows(c, P, o->accumulate);
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
  ScalarPhys P{coords, vel, q, dqdx, rho, dflux, areav, mdot,
               o->alpha, o->alpha_upw, o->upw, o->relax, o->use_limiter,
               PecletDev{o->peclet.form, o->peclet.hybrid, o->peclet.c1,
                         o->peclet.c2}};
  return launch_rows(c, P, o->accumulate);
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
  MomentumPhys P{coords, vel, dudx, visc, areav, mdot, pecfac, wall,
                 o->alpha, o->alpha_upw, o->upw, o->relax, o->include_divU,
                 o->use_limiter};
  return launch_rows(c, P, o->accumulate);
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
  k_momentum_peclet<<<grid_for(c->nedge, kBlock), kBlock, 0, c->stream>>>(
    c->nedge, reinterpret_cast<const int2*>(c->edges), coords, vel, rho, visc,
    PecletDev{pf->form, pf->hybrid, pf->c1, pf->c2}, pn, pfac);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

}  // extern "C"
