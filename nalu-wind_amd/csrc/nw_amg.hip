// nw_amg.hip -- smoothed-aggregation AMG preconditioner on gfx950: the
// hierarchy built on the host (host/nw_amg_setup.cpp) is uploaded once and
// applied as a V-cycle with Chebyshev smoothing (degree 2, eigenvalue ratio
// 20 on D^-1 A, zero initial guess), restriction R = P^T and prolongation P
// as SpMVs, and a dense inverse on the coarsest level.  The V-cycle is a
// fixed linear operator, so it serves as a right preconditioner for GMRES.
// It runs in the solver's storage type: double for the FP64 method, float
// for the FP32-inner method (the FP32 copy is narrowed from the FP64 one on
// the device and shares its structure arrays).
#include <algorithm>
#include <chrono>
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


}  // namespace

// A level operator, in sliced ELL form (sp != nullptr: the layout of Sell,
// one lane per row) when its rows are short, else in CSR with G lanes per
// row (rp).  The fine-level A uses the graph's sliced ELL structure; the
// others own theirs.
template <typename T>
struct DevMat {
  int n = 0;  // rows
  int G = 8;
  const int64_t* sp = nullptr;
  const int32_t* rp = nullptr;
  const int32_t* ci = nullptr;
  const int32_t* dia_w = nullptr;  // offset and delta forms of the fine
  const int32_t* dia_off = nullptr;  // level's slices
  const int8_t* d16 = nullptr;
  const int16_t* cd = nullptr;
  T* v = nullptr;
  int64_t nv = 0;  // stored values (padding included)
  SellView view() const { return {sp, ci, dia_w, dia_off, d16, cd}; }
  int64_t* own_sp = nullptr;
  int32_t* own_rp = nullptr;
  int32_t* own_ci = nullptr;
};

// Rows up to this many entries on average are applied in sliced ELL form
// (measured at 256^3: the 7-entry fine matrix and the prolongators gain
// 1.5-1.9x over CSR, and the 20-24-entry operators of levels 1-2 cut the
// whole solve by 5-10 %; the 36-200-entry coarser operators lose).
constexpr double kSellMaxAvgRow = 25.0;

template <typename T>
static void mat_spmv(const DevMat<T>& M, const T* x, T* y, T alpha,
                     hipStream_t st)
{
  if (M.sp)
    k_sell_spmv<T><<<grid_for(M.n, kBlock), kBlock, 0, st>>>(M.view(), M.v, x, y, M.n,
                                                             alpha);
  else
    NW_ROW_LAUNCH(k_spmv, T, M.G, M.n, st, M.rp, M.ci, M.v, x, y, M.n, alpha);
}

template <typename T>
static void mat_spmv_add(const DevMat<T>& M, const T* x, T* y, hipStream_t st)
{
  if (M.sp)
    k_sell_spmv_add<T><<<grid_for(M.n, kBlock), kBlock, 0, st>>>(M.view(), M.v, x, y,
                                                                 M.n);
  else
    NW_ROW_LAUNCH(k_spmv_add, T, M.G, M.n, st, M.rp, M.ci, M.v, x, y, M.n);
}

template <typename T>
static void mat_resid(const DevMat<T>& M, const T* x, const T* b, T* r,
                      hipStream_t st)
{
  if (M.sp)
    k_sell_resid<T><<<grid_for(M.n, kBlock), kBlock, 0, st>>>(M.view(), M.v, x, b, r,
                                                              M.n);
  else
    NW_ROW_LAUNCH(k_resid, T, M.G, M.n, st, M.rp, M.ci, M.v, x, b, r, M.n);
}

// zo = z + d after the step, or with add zo += z + d
template <typename T>
static void mat_cheb_step(const DevMat<T>& M, const T* z, const T* v,
                          const T* dinv, T* d, T* zo, T c1, T c2, bool add,
                          hipStream_t st)
{
  if (M.sp && add)
    k_sell_cheb_step<T, true><<<grid_for(M.n, kBlock), kBlock, 0, st>>>(
      M.view(), M.v, z, v, dinv, d, zo, M.n, c1, c2);
  else if (M.sp)
    k_sell_cheb_step<T><<<grid_for(M.n, kBlock), kBlock, 0, st>>>(
      M.view(), M.v, z, v, dinv, d, zo, M.n, c1, c2);
  else if (add)
    NW_ROW_LAUNCH_ADD(k_cheb_step, T, M.G, M.n, st, M.rp, M.ci, M.v, z, v,
                      dinv, d, zo, M.n, c1, c2);
  else
    NW_ROW_LAUNCH(k_cheb_step, T, M.G, M.n, st, M.rp, M.ci, M.v, z, v, dinv,
                  d, zo, M.n, c1, c2);
}

// r = b - A x, then d = z = dinv r inv_theta
template <typename T>
static void mat_resid_cheb_init(const DevMat<T>& M, const T* x, const T* b,
                                T* r, const T* dinv, T* d, T* z, T inv_theta,
                                hipStream_t st)
{
  if (M.sp)
    k_sell_resid_cheb_init<T><<<grid_for(M.n, kBlock), kBlock, 0, st>>>(
      M.view(), M.v, x, b, r, dinv, d, z, M.n, inv_theta);
  else
    NW_ROW_LAUNCH(k_resid_cheb_init, T, M.G, M.n, st, M.rp, M.ci, M.v, x, b,
                  r, dinv, d, z, M.n, inv_theta);
}

template <typename T>
struct DevLevel {
  DevMat<T> A, P, R;
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
  DevHier<float> f;  // made from d on the device when first asked for
  std::vector<int64_t> rows, nnz;  // per level
  int nlev = 0;
  double setup_seconds = 0;
  int cheb_degree = 2;
  double cheb_ratio = 20.0;
};

template <typename T>
static void free_mat(DevMat<T>& m)
{
  (void)hipFree(m.own_sp);
  (void)hipFree(m.own_rp);
  (void)hipFree(m.own_ci);
  (void)hipFree(m.v);
  m = DevMat<T>();
}

template <typename T>
static void free_hier(DevHier<T>& h)
{
  for (auto& L : h.lv) {
    free_mat(L.A);
    free_mat(L.P);
    free_mat(L.R);
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
  free_hier(a->f);  // shares structure arrays with d
  free_hier(a->d);
  delete a;
}

namespace {

// device buffers released when the scope ends
struct DevScratch {
  std::vector<void*> p;
  template <typename T>
  int alloc(T** q, int64_t n)
  {
    *q = nullptr;
    if (hipMalloc(q, sizeof(T) * std::max<int64_t>(n, 1)) != hipSuccess) return NW_E_ALLOC;
    p.push_back(*q);
    return NW_OK;
  }
  ~DevScratch()
  {
    for (void* q : p) (void)hipFree(q);
  }
};

template <typename T>
__global__ void k_narrow(const double* __restrict__ in, T* __restrict__ out, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)in[i];
}

// The power iteration's start vector: x_i = 0.5 + u_i, u_i in [0, 1) a
// splitmix64 hash of i (independent of the launch)
__global__ void k_pow_start(double* __restrict__ x, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = (uint64_t)i + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  x[i] = 0.5 + (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_pow_div(double* __restrict__ x, double d, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] /= d;
}

// y = D^-1 y (y holding A x) and, per block in a fixed tree order, the
// partial sums of x.y and y.y; with dinv == nullptr only y.y of y as it is
__global__ __launch_bounds__(kBlock) void k_pow_scale(double* __restrict__ y,
                                                     const double* __restrict__ dinv,
                                                     const double* __restrict__ x, int n,
                                                     double* __restrict__ part)
{
  __shared__ double sxy[kBlock], syy[kBlock];
  const int t = threadIdx.x, i = blockIdx.x * kBlock + t;
  double a = 0.0, b = 0.0;
  if (i < n) {
    double v = y[i];
    if (dinv) {
      v *= dinv[i];
      y[i] = v;
      a = x[i] * v;
    }
    b = v * v;
  }
  sxy[t] = a;
  syy[t] = b;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (t < o) {
      sxy[t] += sxy[t + o];
      syy[t] += syy[t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[2 * blockIdx.x] = sxy[0];
    part[2 * blockIdx.x + 1] = syy[0];
  }
}

// The fine level's filtered operator F of the smoothed-aggregation setup
// (host/nw_amg_setup.cpp, coarsen): entry (i, j != i) is kept when
// a_ij^2 > theta^2 |a_ii a_jj| and neither row is a Dirichlet row (no
// nonzero off-diagonal); the others are lumped into the diagonal.  Per
// row: a_ii and the Dirichlet flag.
__global__ void k_filter_diag(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                              const double* __restrict__ a, int n, double* __restrict__ d,
                              int8_t* __restrict__ dir)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double di = 0.0;
  int8_t only_diag = 1;
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    if (ci[k] == i)
      di = a[k];
    else if (a[k] != 0.0)
      only_diag = 0;
  }
  d[i] = di;
  dir[i] = only_diag;
}

__device__ inline bool filter_keep(int i, int j, double aij, const double* d,
                                   const int8_t* dir, double th2)
{
  return j == i || (!dir[i] && !dir[j] && aij * aij > th2 * fabs(d[i] * d[j]));
}

// F's diagonal a_ii + (sum of the row's dropped entries, in row order) and
// its inverse (0 for a zero diagonal)
__global__ void k_filter_lump(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                              const double* __restrict__ a, int n, const double* __restrict__ d,
                              const int8_t* __restrict__ dir, double th2,
                              double* __restrict__ fd, double* __restrict__ fdinv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double lump = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; ++k)
    if (!filter_keep(i, ci[k], a[k], d, dir, th2)) lump += a[k];
  const double f = d[i] + lump;
  fd[i] = f;
  fdinv[i] = f != 0.0 ? 1.0 / f : 0.0;
}

// y = F x over the kept entries in row order (the diagonal as F's)
__global__ void k_filter_spmv(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                              const double* __restrict__ a, int n, const double* __restrict__ d,
                              const int8_t* __restrict__ dir, double th2,
                              const double* __restrict__ fd, const double* __restrict__ x,
                              double* __restrict__ y)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    const int j = ci[k];
    if (j == i)
      s += fd[i] * x[j];
    else if (filter_keep(i, j, a[k], d, dir, th2))
      s += a[k] * x[j];
  }
  y[i] = s;
}

}  // namespace

// lambda_max(D^-1 A) by `iters` power-iteration steps from a fixed start
// vector (the hashed one, or the host array `start`; spmv(x, y) forms
// y = A x; dinv = 1 / diag): each step normalises
// x, forms y = D^-1 A x and the Rayleigh quotient x.y with the next norm
// y.y; block partial sums are added on the host in block order, so the
// estimate is reproducible.  x and y are n-vectors of work space.
template <typename M>
static int device_lmax(int n, double* x, double* y, const double* dinv, const M& spmv,
                       int iters, hipStream_t st, double* out,
                       const double* start = nullptr)
{
  const unsigned nb = grid_for(n, kBlock);
  DevScratch tmp;
  double* part = nullptr;
  NW_REQUIRE(tmp.alloc(&part, 2 * (int64_t)nb) == NW_OK, NW_E_ALLOC,
             "AMG eigenvalue estimate: device allocation failed");
  std::vector<double> hp(2 * (size_t)nb);
  auto sums = [&](double& xy, double& yy) -> int {
    NW_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(),
                          hipMemcpyDeviceToHost, st));
    NW_HIP(hipStreamSynchronize(st));
    xy = yy = 0.0;
    for (unsigned b = 0; b < nb; ++b) {
      xy += hp[2 * b];
      yy += hp[2 * b + 1];
    }
    return NW_OK;
  };
  if (start)
    NW_HIP(hipMemcpyAsync(x, start, sizeof(double) * n, hipMemcpyHostToDevice, st));
  else
    k_pow_start<<<nb, kBlock, 0, st>>>(x, n);
  k_pow_scale<<<nb, kBlock, 0, st>>>(x, nullptr, nullptr, n, part);
  NW_LAUNCH_CHECK();
  double lam = 1.0, nx2 = 0.0, unused = 0.0;
  int rc = sums(unused, nx2);
  if (rc) return rc;
  for (int it = 0; it < iters; ++it) {
    const double nx = std::sqrt(nx2);
    if (!(nx > 0)) break;
    k_pow_div<<<nb, kBlock, 0, st>>>(x, nx, n);
    spmv(x, y);
    k_pow_scale<<<nb, kBlock, 0, st>>>(y, dinv, x, n, part);
    NW_LAUNCH_CHECK();
    rc = sums(lam, nx2);
    if (rc) return rc;
    std::swap(x, y);
  }
  *out = lam;
  return NW_OK;
}

// Uploads a host CSR matrix.  Rows of more than kSellMaxAvgRow entries on
// average stay in CSR (G lanes per row); shorter ones are packed into sliced
// ELL on the device from the uploaded CSR, with padding positions repeating
// the row's last column (value 0), so that they stay inside the columns of
// a rectangular P or R.
static int upload_mat(const amg::Csr& h, DevMat<double>& d, hipStream_t st)
{
  NW_REQUIRE(h.nrows >= 1 && h.nrows < INT32_MAX && h.ncols >= 1 && h.nnz() < INT32_MAX,
             NW_E_ARG, "AMG level too large or empty");
  const int64_t n = h.nrows, nnz = h.nnz();
  const double avg = (double)nnz / (double)n;
  const std::vector<int32_t> rp(h.rp.begin(), h.rp.end());
  d.n = (int)n;
  if (avg > kSellMaxAvgRow) {
    d.G = lanes_per_row(avg);
    d.nv = nnz;
    NW_HIP(hipMalloc(&d.own_rp, sizeof(int32_t) * rp.size()));
    NW_HIP(hipMalloc(&d.own_ci, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    NW_HIP(hipMalloc(&d.v, sizeof(double) * std::max<int64_t>(nnz, 1)));
    d.rp = d.own_rp;
    d.ci = d.own_ci;
    NW_HIP(hipMemcpyAsync(d.own_rp, rp.data(), sizeof(int32_t) * rp.size(),
                          hipMemcpyHostToDevice, st));
    NW_HIP(hipMemcpyAsync(d.own_ci, h.ci.data(), sizeof(int32_t) * nnz,
                          hipMemcpyHostToDevice, st));
    NW_HIP(hipMemcpyAsync(d.v, h.v.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, st));
    NW_HIP(hipStreamSynchronize(st));
    return NW_OK;
  }
  DevScratch tmp;
  int32_t *drp = nullptr, *dci = nullptr;
  double* dv = nullptr;
  int rc = tmp.alloc(&drp, n + 1);
  if (!rc) rc = tmp.alloc(&dci, nnz);
  if (!rc) rc = tmp.alloc(&dv, nnz);
  NW_REQUIRE(rc == NW_OK, NW_E_ALLOC, "AMG level upload: device allocation failed");
  NW_HIP(hipMemcpyAsync(drp, rp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
  NW_HIP(hipMemcpyAsync(dci, h.ci.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
  NW_HIP(hipMemcpyAsync(dv, h.v.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, st));
  Sell s;
  rc = sell_build(drp, dci, (int)n, s, st, true);
  d.own_sp = s.sp;  // owned by d from here on, also on failure
  d.own_ci = s.ci;
  tmp.p.push_back(s.slot);
  if (rc) return rc;
  d.sp = s.sp;
  d.ci = s.ci;
  d.nv = s.len;
  NW_HIP(hipMalloc(&d.v, sizeof(double) * std::max<int64_t>(s.len, 1)));
  if (s.len > 0) {
    k_sell_values<double><<<grid_for(s.len, kBlock), kBlock, 0, st>>>(s.slot, s.ci, dv,
                                                                      nullptr, d.v, s.len);
    NW_LAUNCH_CHECK();
  }
  NW_HIP(hipStreamSynchronize(st));
  return NW_OK;
}

// The FP32 copy of an FP64 level matrix: the same structure arrays (not
// owned), values narrowed on the device.
static int narrow_mat(const DevMat<double>& s, DevMat<float>& d, hipStream_t st)
{
  d.n = s.n;
  d.G = s.G;
  d.sp = s.sp;
  d.rp = s.rp;
  d.ci = s.ci;
  d.nv = s.nv;
  NW_HIP(hipMalloc(&d.v, sizeof(float) * std::max<int64_t>(s.nv, 1)));
  if (s.nv > 0) {
    k_narrow<float><<<grid_for(s.nv, kBlock), kBlock, 0, st>>>(s.v, d.v, s.nv);
    NW_LAUNCH_CHECK();
  }
  return NW_OK;
}

template <typename T>
static int alloc_level_vectors(DevLevel<T>& L)
{
  for (T** p : {&L.x, &L.b, &L.r, &L.t, &L.cd, &L.cz, &L.cz2})
    NW_HIP(hipMalloc(p, sizeof(T) * std::max(L.n, 1)));
  return NW_OK;
}

// The FP64 hierarchy on the device from the host one.
static int upload_hier(Ctx* c, const amg::Hierarchy& h, DevHier<double>& H)
{
  hipStream_t st = c->stream;
  const int nl = (int)h.levels.size();
  H.lv.resize(nl);
  for (int l = 0; l < nl; ++l) {
    const amg::Level& hl = h.levels[l];
    DevLevel<double>& L = H.lv[l];
    L.n = (int)hl.A.nrows;
    L.lmax = hl.lmax;
    if (l == 0) {
      // the fine operator is the linear system's own matrix, in the
      // graph's sliced ELL order (values refreshed at every solve)
      NW_REQUIRE(c->g.sell.ci && L.n == c->g.nnode, NW_E_STATE,
                 "sliced ELL form not built");
      L.A.n = L.n;
      L.A.sp = c->g.sell.sp;
      L.A.ci = c->g.sell.ci;
      L.A.dia_w = c->g.sell.dia_w;
      L.A.dia_off = c->g.sell.dia_off;
      L.A.d16 = c->g.sell.d16;
      L.A.cd = c->g.sell.cd;
      L.A.nv = c->g.sell.len;
      NW_HIP(hipMalloc(&L.A.v, sizeof(double) * std::max<int64_t>(c->g.sell.len, 1)));
    } else {
      int rc = upload_mat(hl.A, L.A, st);
      if (rc) return rc;
    }
    NW_HIP(hipMalloc(&L.dinv, sizeof(double) * std::max(L.n, 1)));
    NW_HIP(hipMemcpyAsync(L.dinv, hl.dinv.data(), sizeof(double) * L.n,
                          hipMemcpyHostToDevice, st));
    int rc = alloc_level_vectors(L);
    if (!rc && l + 1 < nl) rc = upload_mat(hl.P, L.P, st);
    if (!rc && l + 1 < nl) rc = upload_mat(hl.R, L.R, st);
    if (rc) return rc;
  }
  H.nc = (int)h.levels.back().A.nrows;
  NW_HIP(hipMalloc(&H.cinv, sizeof(double) * std::max<size_t>(h.coarse_inv.size(), 1)));
  NW_HIP(hipMemcpyAsync(H.cinv, h.coarse_inv.data(), sizeof(double) * h.coarse_inv.size(),
                        hipMemcpyHostToDevice, st));
  NW_HIP(hipStreamSynchronize(st));
  H.ready = true;
  return NW_OK;
}

// The FP32 hierarchy from the FP64 one, on the device.
static int narrow_hier(Ctx* c, const DevHier<double>& D, DevHier<float>& H)
{
  hipStream_t st = c->stream;
  const int nl = (int)D.lv.size();
  H.lv.resize(nl);
  for (int l = 0; l < nl; ++l) {
    const DevLevel<double>& S = D.lv[l];
    DevLevel<float>& L = H.lv[l];
    L.n = S.n;
    L.lmax = S.lmax;
    if (l == 0) {
      L.A.n = S.A.n;
      L.A.sp = S.A.sp;
      L.A.ci = S.A.ci;
      L.A.dia_w = S.A.dia_w;
      L.A.dia_off = S.A.dia_off;
      L.A.d16 = S.A.d16;
      L.A.cd = S.A.cd;
      L.A.nv = S.A.nv;
      NW_HIP(hipMalloc(&L.A.v, sizeof(float) * std::max<int64_t>(S.A.nv, 1)));
    } else {
      int rc = narrow_mat(S.A, L.A, st);
      if (rc) return rc;
    }
    NW_HIP(hipMalloc(&L.dinv, sizeof(float) * std::max(L.n, 1)));
    k_narrow<float><<<grid_for(L.n, kBlock), kBlock, 0, st>>>(S.dinv, L.dinv, L.n);
    NW_LAUNCH_CHECK();
    int rc = alloc_level_vectors(L);
    if (!rc && l + 1 < nl) rc = narrow_mat(S.P, L.P, st);
    if (!rc && l + 1 < nl) rc = narrow_mat(S.R, L.R, st);
    if (rc) return rc;
  }
  H.nc = D.nc;
  const int64_t nci = (int64_t)D.nc * D.nc;
  NW_HIP(hipMalloc(&H.cinv, sizeof(float) * std::max<int64_t>(nci, 1)));
  if (nci > 0) {
    k_narrow<float><<<grid_for(nci, kBlock), kBlock, 0, st>>>(D.cinv, H.cinv, nci);
    NW_LAUNCH_CHECK();
  }
  NW_HIP(hipStreamSynchronize(st));
  H.ready = true;
  return NW_OK;
}

int amg_setup(Ctx* c, AmgDevice** out)
{
  const auto t0 = std::chrono::steady_clock::now();
  Graph& g = c->g;
  NW_REQUIRE(g.ndof == 1, NW_E_STATE, "AMG needs a 1-dof system");
  amg::Csr A = amg::csr_with_structure(g.nnode, g.h_row_ptr.data(), g.h_cols.data());
  NW_HIP(hipMemcpyAsync(A.v.data(), g.vals, sizeof(double) * g.nnzb,
                        hipMemcpyDeviceToHost, c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  amg::Options o;
  o.smoother_lmax = false;  // estimated on the device after the upload
  {
    // the fine prolongator's damping estimate, lambda_max(D_F^-1 F), on
    // the device from the matrix there
    const int n = (int)g.nnode;
    const unsigned nb = grid_for(n, kBlock);
    DevScratch tmp;
    double *d = nullptr, *fd = nullptr, *fdinv = nullptr, *x = nullptr, *y = nullptr;
    int8_t* dir = nullptr;
    int arc = tmp.alloc(&d, n);
    if (!arc) arc = tmp.alloc(&fd, n);
    if (!arc) arc = tmp.alloc(&fdinv, n);
    if (!arc) arc = tmp.alloc(&x, n);
    if (!arc) arc = tmp.alloc(&y, n);
    if (!arc) arc = tmp.alloc(&dir, n);
    NW_REQUIRE(arc == NW_OK, NW_E_ALLOC, "AMG setup: device allocation failed");
    const double th2 = o.drop_tol * o.drop_tol;
    k_filter_diag<<<nb, kBlock, 0, c->stream>>>(g.row_ptr, g.cols, g.vals, n, d, dir);
    k_filter_lump<<<nb, kBlock, 0, c->stream>>>(g.row_ptr, g.cols, g.vals, n, d, dir, th2, fd,
                                                fdinv);
    NW_LAUNCH_CHECK();
    double lf = 0.0;
    // from the host setup's start vector, so that the estimate is the one
    // the host would make, up to the order of the dot-product sums
    const std::vector<double> x0 = amg::power_start(n);
    const int prc = device_lmax(n, x, y, fdinv, [&](const double* xv, double* yv) {
      k_filter_spmv<<<nb, kBlock, 0, c->stream>>>(g.row_ptr, g.cols, g.vals, n, d, dir, th2, fd,
                                                  xv, yv);
    }, o.power_iters, c->stream, &lf, x0.data());
    if (prc) return prc;
    o.filtered_lmax0 = lf;
  }
  amg::Hierarchy h;
  const int rc = amg::build(std::move(A), o, h);
  NW_REQUIRE(rc == 0, NW_E_ARG, "AMG setup failed (%d): coarse level too "
             "large or singular", rc);
  AmgDevice* a = new AmgDevice();
  a->cheb_degree = o.cheb_degree;
  a->cheb_ratio = o.cheb_ratio;
  a->nlev = (int)h.levels.size();
  for (const auto& L : h.levels) {
    a->rows.push_back(L.A.nrows);
    a->nnz.push_back(L.A.nnz());
  }
  int urc = upload_hier(c, h, a->d);
  // the smoothers' lambda_max(D^-1 A) on the device, every level but the
  // coarsest (solved directly); the fine level's values first
  if (!urc) {
    DevLevel<double>& L0 = a->d.lv[0];
    const Sell& sl = g.sell;
    k_sell_values<double><<<grid_for(sl.len, kBlock), kBlock, 0, c->stream>>>(
      sl.slot, sl.ci, g.vals, nullptr, L0.A.v, sl.len);
    if (hipGetLastError() != hipSuccess) urc = NW_E_HIP;
  }
  for (int l = 0; !urc && l + 1 < a->nlev; ++l) {
    DevLevel<double>& L = a->d.lv[l];
    urc = device_lmax(L.n, L.x, L.t, L.dinv, [&](const double* x, double* y) {
      mat_spmv<double>(L.A, x, y, 1.0, c->stream);
    }, o.power_iters, c->stream, &L.lmax);
  }
  if (urc) {
    amg_free(a);
    return urc;
  }
  a->setup_seconds =
    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = a;
  return NW_OK;
}

int amg_levels(const AmgDevice* a) { return a ? a->nlev : 0; }
double amg_setup_seconds(const AmgDevice* a) { return a ? a->setup_seconds : 0; }
int amg_level_info(const AmgDevice* a, int l, int64_t* rows, int64_t* nnz)
{
  if (!a || l < 0 || l >= a->nlev) return NW_E_ARG;
  *rows = a->rows[l];
  *nnz = a->nnz[l];
  return NW_OK;
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
int amg_prepare(Ctx* c, AmgDevice* a, double* seconds)
{
  DevHier<T>& H = hier<T>(a);
  NW_REQUIRE(a->d.ready, NW_E_STATE, "AMG hierarchy not set up");
  if (seconds) *seconds = 0.0;
  if (!H.ready) {
    if constexpr (sizeof(T) == 4) {
      const auto t0 = std::chrono::steady_clock::now();
      int rc = narrow_hier(c, a->d, a->f);
      if (rc) return rc;
      if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  }
  // the fine level follows the current matrix values (sliced ELL order)
  const Sell& sl = c->g.sell;
  NW_REQUIRE(sl.ci && H.lv[0].A.ci == sl.ci, NW_E_STATE,
             "AMG fine level needs the graph's sliced ELL form");
  k_sell_values<T><<<grid_for(sl.len, kBlock), kBlock, 0, c->stream>>>(
    sl.slot, sl.ci, c->g.vals, nullptr, H.lv[0].A.v, sl.len);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

// L.r = b - A x on level L
template <typename T>
static void level_resid(DevLevel<T>& L, const T* x, const T* b, hipStream_t st)
{
  mat_resid<T>(L.A, x, b, L.r, st);
}

// Chebyshev interval centre theta = (lmax + lmin) / 2 of level L
static double cheb_theta(const AmgDevice* a, double lmax_est)
{
  const double lmax = 1.1 * lmax_est, lmin = lmax / a->cheb_ratio;
  return 0.5 * (lmax + lmin);
}

// z = p(D^-1 A) D^-1 b on level L (zero initial guess, degree k), written
// to out, or with add added to out.  With init_done the first term
// d = z = D^-1 b / theta is already in L.cd and L.cz (computed with the
// residual b, mat_resid_cheb_init).
template <typename T>
static int cheb0(AmgDevice* a, DevLevel<T>& L, const T* b, T* out, bool add,
                 bool init_done, hipStream_t st)
{
  const double lmax = 1.1 * L.lmax, lmin = lmax / a->cheb_ratio;
  const double theta = cheb_theta(a, L.lmax), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  const int k = a->cheb_degree;
  if (!init_done) {
    T* z0 = (k == 1 && !add) ? out : L.cz;
    k_cheb_init<T><<<grid_for(L.n, kBlock), kBlock, 0, st>>>(
      b, L.dinv, L.cd, z0, L.n, (T)(1.0 / theta));
    NW_LAUNCH_CHECK();
  }
  NW_REQUIRE(k > 1 || !(add == false && init_done), NW_E_STATE,
             "degree-1 smoother with a fused start");
  if (k == 1) {
    if (add) {
      k_axpy_t<T><<<grid_for(L.n, kBlock), kBlock, 0, st>>>(out, L.cz, L.n);
      NW_LAUNCH_CHECK();
    }
    return NW_OK;
  }
  T* z = L.cz;
  T* z2 = L.cz2;
  for (int it = 1; it < k; ++it) {
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    const bool last = it == k - 1;
    mat_cheb_step<T>(L.A, z, b, L.dinv, L.cd, last ? out : z2,
                     (T)(rho_new * rho), (T)(2.0 * rho_new / delta),
                     last && add, st);
    NW_LAUNCH_CHECK();
    std::swap(z, z2);
    rho = rho_new;
  }
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
  int rc = cheb0<T>(a, L, b, x, false, false, st);  // pre-smooth from zero
  if (rc) return rc;
  level_resid<T>(L, x, b, st);
  NW_LAUNCH_CHECK();
  mat_spmv<T>(L.R, L.r, C.b, (T)1, st);
  NW_LAUNCH_CHECK();
  rc = vcycle<T>(a, H, l + 1, C.b, C.x, st);
  if (rc) return rc;
  mat_spmv_add<T>(L.P, C.x, x, st);
  NW_LAUNCH_CHECK();
  // post-smooth the correction of the new residual, added to x
  mat_resid_cheb_init<T>(L.A, x, b, L.r, L.dinv, L.cd, L.cz,
                         (T)(1.0 / cheb_theta(a, L.lmax)), st);
  NW_LAUNCH_CHECK();
  return cheb0<T>(a, L, L.r, x, true, true, st);
}

template <typename T>
int amg_apply(AmgDevice* a, const T* b, T* x, hipStream_t st)
{
  return vcycle<T>(a, hier<T>(a), 0, b, x, st);
}

template int amg_prepare<double>(Ctx*, AmgDevice*, double*);
template int amg_prepare<float>(Ctx*, AmgDevice*, double*);
template int amg_apply<double>(AmgDevice*, const double*, double*, hipStream_t);
template int amg_apply<float>(AmgDevice*, const float*, float*, hipStream_t);

}  // namespace nw
