// nw_update.hip -- the field updates of the projection step around the
// pressure solve: EquationSystem::solution_update (src/EquationSystem.C:
// 556-573, nalu_ngp::field_axpby) and LowMachEquationSystem::
// project_nodal_velocity (src/LowMachEquationSystem.C:882-939).  Both are
// streaming node loops (one element per lane, consecutive lanes on
// consecutive values), bound by HBM bandwidth.
#include "nw_internal.h"

namespace nw {
namespace {

// field = a * delta + b * field
__global__ __launch_bounds__(kBlock) void k_axpby(double a,
                                                  const double* __restrict__ delta,
                                                  double b, double* __restrict__ field,
                                                  int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) field[i] = a * delta[i] + b * field[i];
}

// u -= (g_new - g_old) / (rho Udiag), one lane per node component
__global__ __launch_bounds__(kBlock) void k_project(
  const double* __restrict__ gnew, const double* __restrict__ gold,
  const double* __restrict__ rho, const double* __restrict__ udiag,
  double* __restrict__ u, int64_t n3)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const int64_t node = i / 3;
  const double fac = 1.0 / (rho[node] * udiag[node]);
  u[i] -= fac * (gnew[i] - gold[i]);
}

}  // namespace
}  // namespace nw

using namespace nw;

extern "C" {

int nw_solution_update(nw_ctx* c, double delta_frac, int delta, double field_frac,
                       int field)
{
  NW_REQUIRE(c, NW_E_ARG, "null context");
  NW_REQUIRE(field >= 0 && field < NW_NUM_FIELDS && field != delta, NW_E_ARG,
             "bad field ids (%d, %d)", delta, field);
  double* f = nullptr;
  int rc = ensure_field(c, field, &f);
  if (rc) return rc;
  const int64_t n = field_count(c, field);
  const double* d = nullptr;
  if (delta == NW_LINSYS_SOLUTION) {
    NW_REQUIRE(c->g.ndof && !field_is_edge(field) &&
                 c->g.ndof == field_components(field) && c->g.nnode == c->nnode,
               NW_E_STATE, "the linear system does not match field %d", field);
    d = c->g.sln;
  } else {
    NW_REQUIRE(delta >= 0 && delta < NW_NUM_FIELDS &&
                 field_is_edge(delta) == field_is_edge(field) &&
                 field_components(delta) == field_components(field),
               NW_E_ARG, "fields %d and %d differ in shape", delta, field);
    if ((rc = require_field(c, delta, &d))) return rc;
  }
  if (n == 0) return NW_OK;
  k_axpby<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(delta_frac, d, field_frac,
                                                        f, n);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

int nw_project_nodal_velocity(nw_ctx* c)
{
  NW_REQUIRE(c && c->nnode > 0, NW_E_STATE, "mesh not set");
  const double *rho, *udiag;
  double *u, *dpdx, *utmp;
  int rc = 0;
  if ((rc = require_field(c, NW_F_DENSITY, &rho)) ||
      (rc = require_field(c, NW_F_UDIAG, &udiag)) ||
      (rc = ensure_field(c, NW_F_VELOCITY, &u)) ||
      (rc = ensure_field(c, NW_F_DPDX, &dpdx)) ||
      (rc = ensure_field(c, NW_F_UTMP, &utmp)))
    return rc;
  const int64_t n3 = 3 * c->nnode;
  // save off dp/dx (src/LowMachEquationSystem.C:901-908)
  NW_HIP(hipMemcpyAsync(utmp, dpdx, sizeof(double) * n3, hipMemcpyDeviceToDevice,
                        c->stream));
  // the gradient of the new pressure (:913)
  if ((rc = nw_nodal_grad(c, NW_F_PRESSURE, NW_F_DPDX, 0))) return rc;
  // u^{n+1} = u^{k+1} - (Gp^{N+1} - Gp^N) / (rho Udiag) (:921-939)
  k_project<<<grid_for(n3, kBlock), kBlock, 0, c->stream>>>(dpdx, utmp, rho, udiag,
                                                          u, n3);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

}  // extern "C"
