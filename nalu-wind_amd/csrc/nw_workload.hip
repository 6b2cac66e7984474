// nw_workload.hip -- device generator of the synthetic velocity used by the
// pressure-solve workloads: a divergence-free sum of Fourier modes
// u(x, t) = sum_m b_m sin(k_m . x + phi_m + w_m t), b_m orthogonal to k_m.
// Evaluated where the node coordinates already live, so a time series of
// pressure systems needs no host-to-device traffic per step.
#include <cmath>

#include "nalu_amd_mesh.h"
#include "nw_internal.h"

namespace nw {
namespace {

constexpr int kMaxModes = 32;

struct Modes {
  int n;
  double k[kMaxModes][3], b[kMaxModes][3], ph[kMaxModes];
};

__global__ void k_fourier_velocity(const double* __restrict__ x,
                                   double* __restrict__ u, int64_t n, Modes m)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  for (int q = 0; q < m.n; ++q) {
    const double s = sin(m.k[q][0] * x0 + m.k[q][1] * x1 + m.k[q][2] * x2 + m.ph[q]);
    u0 += s * m.b[q][0];
    u1 += s * m.b[q][1];
    u2 += s * m.b[q][2];
  }
  u[3 * i] = u0;
  u[3 * i + 1] = u1;
  u[3 * i + 2] = u2;
}

}  // namespace
}  // namespace nw

using namespace nw;

extern "C" int nw_fourier_velocity(nw_ctx* c, int nmodes, const double* k,
                                   const double* b, const double* phase,
                                   const double* omega, double t)
{
  NW_REQUIRE(c && k && b && phase && omega, NW_E_ARG, "null argument");
  NW_REQUIRE(nmodes >= 0 && nmodes <= kMaxModes, NW_E_ARG,
             "nmodes must be in [0, %d]", kMaxModes);
  NW_REQUIRE(c->nnode > 0, NW_E_STATE, "mesh not set");
  NW_HIP(hipSetDevice(c->device));
  Modes m;
  m.n = nmodes;
  for (int q = 0; q < nmodes; ++q) {
    for (int d = 0; d < 3; ++d) {
      m.k[q][d] = k[3 * q + d];
      m.b[q][d] = b[3 * q + d];
    }
    m.ph[q] = phase[q] + omega[q] * t;
  }
  const double* x = nullptr;
  int rc = require_field(c, NW_F_COORDINATES, &x);
  if (rc) return rc;
  double* u = nullptr;
  rc = ensure_field(c, NW_F_VELOCITY, &u);
  if (rc) return rc;
  k_fourier_velocity<<<grid_for(c->nnode, kBlock), kBlock, 0, c->stream>>>(
    x, u, c->nnode, m);
  NW_LAUNCH_CHECK();
  return NW_OK;
}
