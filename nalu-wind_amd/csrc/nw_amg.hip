// nw_amg.hip -- smoothed-aggregation AMG preconditioner on gfx950: the
// hierarchy built on the host (host/nw_amg_setup.cpp) is uploaded once and
// applied as a V-cycle with Chebyshev smoothing (degree 2, eigenvalue ratio
// 20 on D^-1 A, zero initial guess), restriction R = P^T and prolongation P
// as SpMVs, and a dense inverse on the coarsest level.  The V-cycle is a
// fixed linear operator, so it serves as a right preconditioner for GMRES.
// It runs in the solver's storage type: double for the FP64 method, float
// for the FP32-inner method (the FP32 copies are made at upload).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nw_amg.h"
#include "nw_sparse.h"

namespace nw {

namespace {

__global__ void k_dense_mv(const double* __restrict__ M, int n,
                           const double* __restrict__ x, double* __restrict__ y)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; ++j) s += M[(int64_t)i * n + j] * x[j];
  y[i] = s;
}

__global__ void k_dense_mv_f(const float* __restrict__ M, int n,
                             const float* __restrict__ x, float* __restrict__ y)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; ++j) s += (double)M[(int64_t)i * n + j] * (double)x[j];
  y[i] = (float)s;
}

template <typename T>
__global__ void k_axpy_t(T* __restrict__ x, const T* __restrict__ t, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += t[i];
}

template <typename T>
__global__ void k_cast(const double* __restrict__ a, T* __restrict__ b, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}

}  // namespace

template <typename T>
struct DevCsr {
  int n = 0;
  int G = 8;
  int32_t* rp = nullptr;
  int32_t* ci = nullptr;
  T* v = nullptr;
  bool owns = true;
};

template <typename T>
struct DevLevel {
  DevCsr<T> A, P, R;
  T* dinv = nullptr;
  T *x = nullptr, *b = nullptr, *r = nullptr, *t = nullptr;
  T *cd = nullptr, *cz = nullptr, *cz2 = nullptr;
  double lmax = 0;
  int n = 0;
};

template <typename T>
struct DevHier {
  std::vector<DevLevel<T>> lv;
  T* cinv = nullptr;
  int nc = 0;
  bool ready = false;
};

struct AmgDevice {
  DevHier<double> d;
  DevHier<float> f;
  std::vector<amg::Level> hostlv;  // kept for the FP32 upload on demand
  std::vector<double> coarse_inv;
  int nlev = 0;
  double setup_seconds = 0;
  int cheb_degree = 2;
  double cheb_ratio = 20.0;
  int64_t fine_nnz = 0;
};

template <typename T>
static void free_csr(DevCsr<T>& m)
{
  if (m.owns) {
    (void)hipFree(m.rp);
    (void)hipFree(m.ci);
  }
  (void)hipFree(m.v);
  m = DevCsr<T>();
}

template <typename T>
static void free_hier(DevHier<T>& h)
{
  for (auto& L : h.lv) {
    free_csr(L.A);
    free_csr(L.P);
    free_csr(L.R);
    for (T* p : {L.dinv, L.x, L.b, L.r, L.t, L.cd, L.cz, L.cz2}) (void)hipFree(p);
  }
  h.lv.clear();
  (void)hipFree(h.cinv);
  h.cinv = nullptr;
  h.ready = false;
}

void amg_free(AmgDevice* a)
{
  if (!a) return;
  free_hier(a->d);
  free_hier(a->f);
  delete a;
}

template <typename T>
static int upload_csr(const amg::Csr& h, DevCsr<T>& d, hipStream_t st)
{
  NW_REQUIRE(h.nnz() < INT32_MAX, NW_E_ARG, "AMG level too large");
  std::vector<int32_t> rp(h.rp.begin(), h.rp.end());
  std::vector<T> v(h.v.begin(), h.v.end());
  d.n = (int)h.nrows;
  d.G = lanes_per_row((double)h.nnz() / std::max<int64_t>(h.nrows, 1));
  NW_HIP(hipMalloc(&d.rp, sizeof(int32_t) * rp.size()));
  NW_HIP(hipMalloc(&d.ci, sizeof(int32_t) * std::max<int64_t>(h.nnz(), 1)));
  NW_HIP(hipMalloc(&d.v, sizeof(T) * std::max<int64_t>(h.nnz(), 1)));
  NW_HIP(hipMemcpyAsync(d.rp, rp.data(), sizeof(int32_t) * rp.size(),
                        hipMemcpyHostToDevice, st));
  NW_HIP(hipMemcpyAsync(d.ci, h.ci.data(), sizeof(int32_t) * h.nnz(),
                        hipMemcpyHostToDevice, st));
  NW_HIP(hipMemcpyAsync(d.v, v.data(), sizeof(T) * h.nnz(),
                        hipMemcpyHostToDevice, st));
  NW_HIP(hipStreamSynchronize(st));
  return NW_OK;
}

template <typename T>
static int upload_hier(Ctx* c, AmgDevice* a, DevHier<T>& H)
{
  hipStream_t st = c->stream;
  const int nl = (int)a->hostlv.size();
  H.lv.resize(nl);
  for (int l = 0; l < nl; ++l) {
    const amg::Level& h = a->hostlv[l];
    DevLevel<T>& L = H.lv[l];
    L.n = (int)(l == 0 ? c->g.nnode : h.A.nrows);
    L.lmax = h.lmax;
    if (l == 0) {
      // the fine operator is the linear system's own matrix
      L.A.n = L.n;
      L.A.G = lanes_per_row((double)c->g.nnzb / std::max<int64_t>(L.n, 1));
      L.A.rp = c->g.row_ptr;
      L.A.ci = c->g.cols;
      L.A.owns = false;
      NW_HIP(hipMalloc(&L.A.v, sizeof(T) * c->g.nnzb));
    } else {
      int rc = upload_csr(h.A, L.A, st);
      if (rc) return rc;
    }
    std::vector<T> dinv(h.dinv.begin(), h.dinv.end());
    NW_HIP(hipMalloc(&L.dinv, sizeof(T) * L.n));
    NW_HIP(hipMemcpyAsync(L.dinv, dinv.data(), sizeof(T) * L.n,
                          hipMemcpyHostToDevice, st));
    for (T** p : {&L.x, &L.b, &L.r, &L.t, &L.cd, &L.cz, &L.cz2})
      NW_HIP(hipMalloc(p, sizeof(T) * std::max(L.n, 1)));
    if (l + 1 < nl) {
      int rc = upload_csr(h.P, L.P, st);
      if (rc) return rc;
      rc = upload_csr(h.R, L.R, st);
      if (rc) return rc;
    }
    NW_HIP(hipStreamSynchronize(st));
  }
  H.nc = (int)a->hostlv.back().A.nrows;
  std::vector<T> ci(a->coarse_inv.begin(), a->coarse_inv.end());
  NW_HIP(hipMalloc(&H.cinv, sizeof(T) * std::max<size_t>(ci.size(), 1)));
  NW_HIP(hipMemcpyAsync(H.cinv, ci.data(), sizeof(T) * ci.size(),
                        hipMemcpyHostToDevice, st));
  NW_HIP(hipStreamSynchronize(st));
  H.ready = true;
  return NW_OK;
}

int amg_setup(Ctx* c, AmgDevice** out)
{
  Graph& g = c->g;
  NW_REQUIRE(g.ndof == 1, NW_E_STATE, "AMG needs a 1-dof system");
  amg::Csr A;
  A.nrows = A.ncols = g.nnode;
  A.rp.assign(g.h_row_ptr.begin(), g.h_row_ptr.end());
  A.ci = g.h_cols;
  A.v.resize(g.nnzb);
  NW_HIP(hipMemcpyAsync(A.v.data(), g.vals, sizeof(double) * g.nnzb,
                        hipMemcpyDeviceToHost, c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  amg::Options o;
  amg::Hierarchy h;
  const int rc = amg::build(A, o, h);
  NW_REQUIRE(rc == 0, NW_E_ARG, "AMG setup failed (%d): coarse level too "
             "large or singular", rc);
  AmgDevice* a = new AmgDevice();
  a->setup_seconds = h.setup_seconds;
  a->cheb_degree = o.cheb_degree;
  a->cheb_ratio = o.cheb_ratio;
  a->nlev = (int)h.levels.size();
  a->fine_nnz = g.nnzb;
  a->coarse_inv = std::move(h.coarse_inv);
  a->hostlv = std::move(h.levels);
  a->hostlv[0].A = amg::Csr();  // the fine matrix lives on the device
  *out = a;
  return NW_OK;
}

int amg_levels(const AmgDevice* a) { return a ? a->nlev : 0; }
double amg_setup_seconds(const AmgDevice* a) { return a ? a->setup_seconds : 0; }
int64_t amg_level_rows(const AmgDevice* a, int l)
{
  return (a && l < a->nlev) ? (l == 0 ? -1 : a->hostlv[l].A.nrows) : 0;
}

template <typename T>
static DevHier<T>& hier(AmgDevice* a);
template <>
DevHier<double>& hier<double>(AmgDevice* a)
{
  return a->d;
}
template <>
DevHier<float>& hier<float>(AmgDevice* a)
{
  return a->f;
}

template <typename T>
int amg_prepare(Ctx* c, AmgDevice* a)
{
  DevHier<T>& H = hier<T>(a);
  if (!H.ready) {
    int rc = upload_hier<T>(c, a, H);
    if (rc) return rc;
  }
  // the fine level follows the current matrix values
  k_cast<T><<<grid_for(c->g.nnzb, kBlock), kBlock, 0, c->stream>>>(
    c->g.vals, H.lv[0].A.v, c->g.nnzb);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

// z = p(D^-1 A) D^-1 b on level L (zero initial guess, degree k)
template <typename T>
static int cheb0(AmgDevice* a, DevLevel<T>& L, const T* b, T** zout,
                 hipStream_t st)
{
  const double lmax = 1.1 * L.lmax, lmin = lmax / a->cheb_ratio;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  k_cheb_init<T><<<grid_for(L.n, kBlock), kBlock, 0, st>>>(b, L.dinv, L.cd, L.cz,
                                                          L.n, (T)(1.0 / theta));
  NW_LAUNCH_CHECK();
  T* z = L.cz;
  T* z2 = L.cz2;
  for (int it = 1; it < a->cheb_degree; ++it) {
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    NW_ROW_LAUNCH(k_cheb_step, T, L.A.G, L.n, st, L.A.rp, L.A.ci, L.A.v, z, b,
                  L.dinv, L.cd, z2, L.n, (T)(rho_new * rho),
                  (T)(2.0 * rho_new / delta));
    NW_LAUNCH_CHECK();
    std::swap(z, z2);
    rho = rho_new;
  }
  *zout = z;
  return NW_OK;
}

template <typename T>
static int vcycle(AmgDevice* a, DevHier<T>& H, int l, const T* b, T* x,
                  hipStream_t st)
{
  DevLevel<T>& L = H.lv[l];
  const int last = (int)H.lv.size() - 1;
  if (l == last) {
    if constexpr (sizeof(T) == 8)
      k_dense_mv<<<grid_for(H.nc, 64), 64, 0, st>>>(H.cinv, H.nc, b, x);
    else
      k_dense_mv_f<<<grid_for(H.nc, 64), 64, 0, st>>>(H.cinv, H.nc, b, x);
    NW_LAUNCH_CHECK();
    return NW_OK;
  }
  DevLevel<T>& C = H.lv[l + 1];
  T* z = nullptr;
  int rc = cheb0<T>(a, L, b, &z, st);  // pre-smooth from zero
  if (rc) return rc;
  NW_HIP(hipMemcpyAsync(x, z, sizeof(T) * L.n, hipMemcpyDeviceToDevice, st));
  NW_ROW_LAUNCH(k_resid, T, L.A.G, L.n, st, L.A.rp, L.A.ci, L.A.v, x, b, L.r, L.n);
  NW_LAUNCH_CHECK();
  NW_ROW_LAUNCH(k_spmv, T, L.R.G, L.R.n, st, L.R.rp, L.R.ci, L.R.v, L.r, C.b,
                L.R.n, (T)1);
  NW_LAUNCH_CHECK();
  rc = vcycle<T>(a, H, l + 1, C.b, C.x, st);
  if (rc) return rc;
  NW_ROW_LAUNCH(k_spmv_add, T, L.P.G, L.P.n, st, L.P.rp, L.P.ci, L.P.v, C.x, x,
                L.P.n);
  NW_LAUNCH_CHECK();
  NW_ROW_LAUNCH(k_resid, T, L.A.G, L.n, st, L.A.rp, L.A.ci, L.A.v, x, b, L.r, L.n);
  NW_LAUNCH_CHECK();
  rc = cheb0<T>(a, L, L.r, &z, st);  // post-smooth the correction
  if (rc) return rc;
  k_axpy_t<T><<<grid_for(L.n, kBlock), kBlock, 0, st>>>(x, z, L.n);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

template <typename T>
int amg_apply(AmgDevice* a, const T* b, T* x, hipStream_t st)
{
  return vcycle<T>(a, hier<T>(a), 0, b, x, st);
}

template int amg_prepare<double>(Ctx*, AmgDevice*);
template int amg_prepare<float>(Ctx*, AmgDevice*);
template int amg_apply<double>(AmgDevice*, const double*, double*, hipStream_t);
template int amg_apply<float>(AmgDevice*, const float*, float*, hipStream_t);

}  // namespace nw
