// nw_krylov.hip -- pressure-Poisson Krylov solve on gfx950.
//
// Replaces TpetraLinearSystem::solve -> TpetraLinearSolver::solve (Belos
// GMRES with an Ifpack2 preconditioner, src/LinearSolver.C:209-254, settings
// src/LinearSolverConfig.C:37-140): restarted GMRES(m), right
// preconditioning, classical Gram-Schmidt with reorthogonalisation (the
// reference configures Belos' iterated CGS, "ICGS"), x0 = 0 by default,
// convergence on ||b - A x||_2 / ||b||_2 evaluated in FP64.
//
// Orthogonalisation is CGS2 with the second pass delayed by one step
// (Swirydowicz, Langou, Ananthan, Yang, Thomas, "Low synchronization
// Gram-Schmidt and GMRES algorithms", NLAA 2021): per Arnoldi step one
// fused dot pass over the basis ([Q, y]^T [y, z]) and one fused update pass
// that both finishes q_j = (y - Q s)/rho and forms the next candidate, with
// one host synchronisation.  The Hessenberg column of step j-1 is corrected
// with the reorthogonalisation coefficients, so every column is final one
// step after it is first formed.
//
// Every restart cycle begins from the FP64 residual r = b - A x and ends
// with x += M^-1 Q y in FP64, so a cycle is one iterative-refinement step.
// The cycle (SpMV, Gram-Schmidt, preconditioner) is templated on its
// storage type: double gives the plain FP64 method; float gives the
// FP32-inner / FP64-refined method over an FP32 copy of the matrix.  Dot
// products accumulate in FP64 and are reduced in a fixed order, so results
// are reproducible run to run.
//
// Preconditioners (right): Jacobi, applied by scaling the matrix columns
// once per solve (A D^-1); Chebyshev of degree k on D^-1 A over
// [lmax/ratio, lmax] (Saad, Iterative Methods for Sparse Linear Systems,
// 2nd ed., Alg. 12.1), lmax from a seeded power iteration.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "nw_internal.h"
#include "nw_sparse.h"
#include "nw_amg.h"

namespace nw {

namespace {

constexpr int kMaxBasis = 64;  // kspace + 1 <= 64
constexpr int kTile = 8;       // basis vectors per dot tile
// row blocks of the fused dot kernels: 2048 keeps 8 waves per CU busy even
// for a short tile (512 measured 2 % / 4 % slower for the FP64 / FP32
// solves at 256^3, whose dots were latency-bound; 4096-8192 no better)
constexpr int kRowBlocks = 2048;
constexpr int kDotVals = 2 * kTile + 3;  // per tile: Q.y, Q.z, y.y, y.z, z.z

__device__ inline double wave_sum(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum NA per-thread accumulators over the block; thread k < NA writes
// out[k].
template <int NA>
__device__ inline void block_reduce_store(double (&acc)[NA],
                                          double* __restrict__ out)
{
  __shared__ double sh[kBlock / kWave][NA];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) sh[wv][k] = s;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NA; k += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) s += sh[w][k];
    out[k] = s;
  }
}

// out[v] = sum_b partial[b * nv + v] over nb blocks, one workgroup per value,
// fixed-order tree (reproducible).
__global__ __launch_bounds__(kBlock) void k_reduce(
  const double* __restrict__ partial, int nb, int nv, double* __restrict__ out)
{
  const int v = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    s += partial[(int64_t)b * nv + v];
  double a[1] = {s};
  __shared__ double tot[1];
  block_reduce_store<1>(a, tot);
  if (threadIdx.x == 0) out[v] = tot[0];
}

// Fused Arnoldi dots.  Block b = rowblock * ntiles + tile (tiles of one row
// block are adjacent, so the y/z rows they share are re-read from L2).
// Tile t covers basis vectors Q_k, k in [8t, 8t+8) and k < nq, and writes
// partial[b][0..7] = Q_k.y, [8..15] = Q_k.z; tile 0 also writes [16] = y.y,
// [17] = y.z, [18] = z.z.  z may be null (then its dots are zero).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_arnoldi_dots(
  const T* __restrict__ V, int64_t n, int nq, const T* __restrict__ y,
  const T* __restrict__ z, int ntiles, double* __restrict__ partial)
{
  const int tile = blockIdx.x % ntiles;
  const int rb = blockIdx.x / ntiles;
  const int nrb = gridDim.x / ntiles;
  const int k0 = tile * kTile;
  const int kc = min(kTile, nq - k0);
  double acc[kDotVals];
#pragma unroll
  for (int k = 0; k < kDotVals; ++k) acc[k] = 0.0;
  const int64_t chunk = (n + nrb - 1) / nrb;
  const int64_t i0 = (int64_t)rb * chunk;
  const int64_t i1 = min(n, i0 + chunk);
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double yi = (double)y[i];
    const double zi = z ? (double)z[i] : 0.0;
#pragma unroll
    for (int k = 0; k < kTile; ++k)
      if (k < kc) {
        const double q = (double)V[(int64_t)(k0 + k) * n + i];
        acc[k] += q * yi;
        acc[kTile + k] += q * zi;
      }
    if (tile == 0) {
      acc[2 * kTile] += yi * yi;
      acc[2 * kTile + 1] += yi * zi;
      acc[2 * kTile + 2] += zi * zi;
    }
  }
  block_reduce_store<kDotVals>(acc, partial + (int64_t)blockIdx.x * kDotVals);
}

// Fused Arnoldi update: with Q_k = V[k] (k < nq), y = V[nq], z = V[nq+1]:
//   V[nq]   = a y - sum_k b_k Q_k          (the finished basis vector)
//   V[nq+1] = c z - e y - sum_k d_k Q_k    (the next candidate)
// coef = {a, c, e, b_0..b_{nq-1}, d_0..d_{nq-1}}.  With has_z == 0 only the
// first line is applied.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_arnoldi_update(
  T* __restrict__ V, int64_t n, int nq, const double* __restrict__ coef,
  int has_z)
{
  __shared__ double cb[kMaxBasis], cd[kMaxBasis];
  for (int k = threadIdx.x; k < nq; k += blockDim.x) {
    cb[k] = coef[3 + k];
    cd[k] = coef[3 + nq + k];
  }
  __syncthreads();
  const double a = coef[0], c = coef[1], e = coef[2];
  T* __restrict__ Y = V + (int64_t)nq * n;
  T* __restrict__ Z = Y + n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double yi = (double)Y[i];
    double q = a * yi;
    double w = has_z ? c * (double)Z[i] - e * yi : 0.0;
    int k = 0;
    for (; k + 4 <= nq; k += 4) {
      const double v0 = (double)V[(int64_t)(k + 0) * n + i];
      const double v1 = (double)V[(int64_t)(k + 1) * n + i];
      const double v2 = (double)V[(int64_t)(k + 2) * n + i];
      const double v3 = (double)V[(int64_t)(k + 3) * n + i];
      q -= cb[k] * v0 + cb[k + 1] * v1 + cb[k + 2] * v2 + cb[k + 3] * v3;
      w -= cd[k] * v0 + cd[k + 1] * v1 + cd[k + 2] * v2 + cd[k + 3] * v3;
    }
    for (; k < nq; ++k) {
      const double v0 = (double)V[(int64_t)k * n + i];
      q -= cb[k] * v0;
      w -= cd[k] * v0;
    }
    Y[i] = (T)q;
    if (has_z) Z[i] = (T)w;
  }
}

// u = sum_k coef[k] V[k] (k < nk), stored as T
template <typename T>
__global__ __launch_bounds__(kBlock) void k_combine(
  const T* __restrict__ V, int64_t n, int nk, const double* __restrict__ coef,
  T* __restrict__ u)
{
  __shared__ double cc[kMaxBasis];
  for (int k = threadIdx.x; k < nk; k += blockDim.x) cc[k] = coef[k];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < nk; ++k) s += cc[k] * (double)V[(int64_t)k * n + i];
    u[i] = (T)s;
  }
}

// x += (dinv ? dinv * u : u), FP64 accumulation of the correction
template <typename T>
__global__ void k_axpy_prec(double* __restrict__ x, const T* __restrict__ u,
                            const double* __restrict__ dinv, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ui = (double)u[i];
  x[i] += dinv ? dinv[i] * ui : ui;
}

// r = b - A x (FP64) with per-block partial ||r||^2; G lanes per row
template <int G>
__global__ __launch_bounds__(kBlock) void k_residual(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const double* __restrict__ va, const double* __restrict__ x,
  const double* __restrict__ b, double* __restrict__ r, int n,
  double* __restrict__ partial)
{
  double acc[1] = {0.0};
  for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x; t0 < (int64_t)n * G;
       t0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = t0 + threadIdx.x;
    const int row = (int)(t / G);
    const int g = threadIdx.x & (G - 1);
    double s = row < n ? csr_dot<double, G>(rp, ci, va, x, row, g) : 0.0;
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    if (row < n && g == 0) {
      const double ri = b[row] - s;
      r[row] = ri;
      acc[0] += ri * ri;
    }
  }
  block_reduce_store<1>(acc, partial + blockIdx.x);
}

__global__ void k_diag_inv(const int32_t* __restrict__ diag,
                           const double* __restrict__ va,
                           double* __restrict__ dinv, int n,
                           int* __restrict__ bad)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = va[diag[i]];
  if (!(d != 0.0) || !isfinite(d)) atomicOr(bad, 1);
  dinv[i] = 1.0 / d;
}

template <typename T>
__global__ void k_convert(const double* __restrict__ a, T* __restrict__ b,
                          double scale, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)(a[i] * scale);
}

__global__ void k_seed_vec(double* __restrict__ x, int64_t n, uint64_t seed)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = seed + 0x9e3779b97f4a7c15ULL * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  x[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) + 0.5;
}

__global__ void k_vmul(double* __restrict__ y, const double* __restrict__ d,
                       int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] *= d[i];
}

// *flag = 1 if a and b differ anywhere (flag cleared by the caller)
__global__ void k_differs(const double* __restrict__ a,
                          const double* __restrict__ b, int64_t n,
                          int* __restrict__ flag)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && a[i] != b[i]) *flag = 1;
}


unsigned stream_grid(int64_t n)
{
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(grid_for(n, kBlock), 4096));
}

// width[k] = longest row of slice k (a wave per slice)
__global__ __launch_bounds__(kBlock) void k_sell_width(
  const int32_t* __restrict__ rp, int n, int nslices, int32_t* __restrict__ width)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int len = t < n ? rp[t + 1] - rp[t] : 0;
  for (int o = kWave / 2; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0 && (t >> 6) < nslices) width[t >> 6] = len;
}

// The offset form of each slice (a wave per slice): the sorted union of
// its rows' column offsets (column - row), taken one smallest offset at a
// time over the wave, when it has at most kDiaMax values and row + offset
// lies in [0, n) for every row of the slice; dia_w[k] = its size, or 0.
__global__ __launch_bounds__(kBlock) void k_dia_pattern(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, int n, int nslices,
  int32_t* __restrict__ dia_w, int32_t* __restrict__ dia_off)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 6;
  if (k >= nslices) return;  // whole waves
  const int lane = t & 63;
  const bool valid = t < n;
  int pos = valid ? rp[t] : 0;
  const int end = valid ? rp[t + 1] : 0;
  int cnt = 0;
  bool ok = true;
  for (;;) {
    const int mine = pos < end ? ci[pos] - t : INT_MAX;
    int cand = mine;
    for (int o = kWave / 2; o > 0; o >>= 1) cand = min(cand, __shfl_xor(cand, o, kWave));
    if (cand == INT_MAX) break;  // the same on every lane
    if (cnt == kDiaMax) {
      ok = false;
      break;
    }
    const int64_t col = (int64_t)t + cand;
    if (__ballot(valid && (col < 0 || col >= n)) != 0) {
      ok = false;
      break;
    }
    if (lane == 0) dia_off[(int64_t)k * kDiaMax + cnt] = cand;
    ++cnt;
    if (mine == cand) ++pos;
  }
  if (lane == 0) dia_w[k] = ok ? cnt : 0;
}

// columns and CSR slots of every sliced-ELL position (a lane per row);
// padding positions take the row's own index as their column (square
// matrices) or, with pad_last, the row's last column (column 0 for an
// empty row), which stays inside any matrix's columns.  A slice in offset
// form puts each row's entry with offset dia_off[j] at position j.
__global__ __launch_bounds__(kBlock) void k_sell_struct(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, int n,
  const int64_t* __restrict__ sp, int nslices, int32_t* __restrict__ sci,
  int32_t* __restrict__ sslot, bool pad_last, const int32_t* __restrict__ dia_w,
  const int32_t* __restrict__ dia_off)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 6;
  if (k >= nslices) return;
  const int64_t b = sp[k] + (t & 63);
  const int w = (int)((sp[k + 1] - sp[k]) >> 6);
  const int r0 = t < n ? rp[t] : 0;
  const int len = t < n ? rp[t + 1] - r0 : 0;
  if (dia_w && dia_w[k] > 0) {
    int q = r0;  // the row's entries, in column (= offset) order
    for (int j = 0; j < w; ++j) {
      const int64_t p = b + 64 * (int64_t)j;
      const int off = dia_off[(int64_t)k * kDiaMax + j];
      if (t < n) {
        sci[p] = t + off;
        if (q < r0 + len && ci[q] - t == off) {
          sslot[p] = q;
          ++q;
        } else {
          sslot[p] = -1;
        }
      } else {
        sci[p] = 0;
        sslot[p] = -1;
      }
    }
    return;
  }
  for (int j = 0; j < w; ++j) {
    const int64_t p = b + 64 * (int64_t)j;
    if (j < len) {
      sci[p] = ci[r0 + j];
      sslot[p] = r0 + j;
    } else {
      sci[p] = pad_last ? (len > 0 ? ci[r0 + len - 1] : 0) : (t < n ? t : 0);
      sslot[p] = -1;
    }
  }
}

// The delta form of each slice not in offset form (a wave per slice):
// when every column of its rows lies within 32767 of the row, cd[p] =
// column - row and d16[k] = 1; else d16[k] = 0.
__global__ __launch_bounds__(kBlock) void k_sell_delta16(
  const int64_t* __restrict__ sp, const int32_t* __restrict__ sci, int n, int nslices,
  const int32_t* __restrict__ dia_w, int16_t* __restrict__ cd, int8_t* __restrict__ d16)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 6;
  if (k >= nslices) return;  // whole waves
  const int64_t b = sp[k] + (t & 63);
  const int w = (int)((sp[k + 1] - sp[k]) >> 6);
  bool fits = true;
  for (int j = 0; j < w && t < n; ++j) {
    const int d = sci[b + 64 * (int64_t)j] - t;
    fits = fits && d >= -32767 && d <= 32767;
  }
  const bool ok = (dia_w == nullptr || dia_w[k] == 0) && __ballot(!fits) == 0;
  for (int j = 0; j < w; ++j) {
    const int64_t p = b + 64 * (int64_t)j;
    cd[p] = ok && t < n ? (int16_t)(sci[p] - t) : (int16_t)0;
  }
  if ((t & 63) == 0) d16[k] = ok ? 1 : 0;
}

}  // namespace

int sell_build(const int32_t* rp, const int32_t* ci, int n, Sell& s,
               hipStream_t st, bool pad_last, bool offsets)
{
  // the callers pass a new structure or one a failed build left part way
  for (void* p : {(void*)s.sp, (void*)s.ci, (void*)s.slot, (void*)s.dia_w, (void*)s.dia_off,
                  (void*)s.d16, (void*)s.cd})
    (void)hipFree(p);
  s = Sell();
  s.n = n;
  s.nslices = (n + kWave - 1) / kWave;
  const int ns = s.nslices;
  const unsigned grid = grid_for((int64_t)ns * kWave, kBlock);
  int32_t* dw = nullptr;
  NW_HIP(hipMalloc(&dw, sizeof(int32_t) * std::max(ns, 1)));
  k_sell_width<<<grid, kBlock, 0, st>>>(rp, n, ns, dw);
  std::vector<int32_t> w(ns), dia(offsets ? ns : 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && offsets) {
    e = hipMalloc(&s.dia_w, sizeof(int32_t) * std::max(ns, 1));
    if (e == hipSuccess)
      e = hipMalloc(&s.dia_off, sizeof(int32_t) * (size_t)std::max(ns, 1) * kDiaMax);
    if (e == hipSuccess) {
      k_dia_pattern<<<grid, kBlock, 0, st>>>(rp, ci, n, ns, s.dia_w, s.dia_off);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(dia.data(), s.dia_w, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(w.data(), dw, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(dw);
  NW_HIP(e);
  std::vector<int64_t> sp(ns + 1, 0);
  s.dia_slices = 0;
  for (int k = 0; k < ns; ++k) {
    const int wk = offsets && dia[k] > 0 ? dia[k] : w[k];  // the union holds every row
    s.dia_slices += offsets && dia[k] > 0;
    sp[k + 1] = sp[k] + (int64_t)kWave * wk;
  }
  if (offsets && s.dia_slices == 0) {  // nothing to gain: the plain form
    (void)hipFree(s.dia_w);
    (void)hipFree(s.dia_off);
    s.dia_w = s.dia_off = nullptr;
  }
  s.len = sp[ns];
  NW_HIP(hipMalloc(&s.sp, sizeof(int64_t) * (ns + 1)));
  NW_HIP(hipMalloc(&s.ci, sizeof(int32_t) * std::max<int64_t>(s.len, 1)));
  NW_HIP(hipMalloc(&s.slot, sizeof(int32_t) * std::max<int64_t>(s.len, 1)));
  NW_HIP(hipMemcpyAsync(s.sp, sp.data(), sizeof(int64_t) * (ns + 1),
                        hipMemcpyHostToDevice, st));
  k_sell_struct<<<grid, kBlock, 0, st>>>(rp, ci, n, s.sp, ns, s.ci, s.slot, pad_last,
                                         s.dia_w, s.dia_off);
  NW_LAUNCH_CHECK();
  if (offsets) {
    NW_HIP(hipMalloc(&s.d16, sizeof(int8_t) * std::max(ns, 1)));
    NW_HIP(hipMalloc(&s.cd, sizeof(int16_t) * std::max<int64_t>(s.len, 1)));
    k_sell_delta16<<<grid, kBlock, 0, st>>>(s.sp, s.ci, n, ns, s.dia_w, s.cd, s.d16);
    NW_LAUNCH_CHECK();
    std::vector<int8_t> d16(ns);
    NW_HIP(hipMemcpyAsync(d16.data(), s.d16, ns, hipMemcpyDeviceToHost, st));
    NW_HIP(hipStreamSynchronize(st));
    s.d16_slices = 0;
    for (int k = 0; k < ns; ++k) s.d16_slices += d16[k];
    if (s.d16_slices == 0) {  // nothing to gain
      (void)hipFree(s.d16);
      (void)hipFree(s.cd);
      s.d16 = nullptr;
      s.cd = nullptr;
    }
  }
  NW_HIP(hipStreamSynchronize(st));
  return NW_OK;
}

// ---------------------------------------------------------------- workspace

template <typename T>
struct Space {
  T* V = nullptr;      // (m+1) x n basis
  T* mat = nullptr;    // matrix values in T (Jacobi: columns scaled)
  T* dinv = nullptr;   // Chebyshev only
  T *cz = nullptr, *cd = nullptr, *cz2 = nullptr, *u = nullptr;
};

struct Krylov {
  int64_t n = 0, nnz = 0;
  int m = 0;
  double* dinv = nullptr;  // FP64 1/diag
  double* r = nullptr;     // FP64 residual
  double* partial = nullptr;
  double* red = nullptr;   // reduced results (device)
  double* coef = nullptr;  // coefficients (device)
  double* pe = nullptr;    // power iteration work
  double* xg = nullptr;    // learned guess
  int* bad = nullptr;
  double* h_pinned = nullptr;  // host pinned scratch
  Space<double> d;
  Space<float> f;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

constexpr int kPartialCap = kRowBlocks * kTile * kDotVals;  // doubles
constexpr int kRedCap = kTile * kDotVals + 8;

template <typename T>
static void free_space(Space<T>& s)
{
  (void)hipFree(s.V);
  (void)hipFree(s.mat);
  (void)hipFree(s.dinv);
  (void)hipFree(s.cz);
  (void)hipFree(s.cd);
  (void)hipFree(s.cz2);
  (void)hipFree(s.u);
  s = Space<T>();
}

void free_krylov(Krylov* k)
{
  if (!k) return;
  (void)hipFree(k->dinv);
  (void)hipFree(k->r);
  (void)hipFree(k->partial);
  (void)hipFree(k->red);
  (void)hipFree(k->coef);
  (void)hipFree(k->pe);
  (void)hipFree(k->xg);
  (void)hipFree(k->bad);
  (void)hipHostFree(k->h_pinned);
  free_space(k->d);
  free_space(k->f);
  if (k->ev0) (void)hipEventDestroy(k->ev0);
  if (k->ev1) (void)hipEventDestroy(k->ev1);
  delete k;
}

template <typename T>
static int alloc_space(Space<T>& s, int64_t n, int64_t nnz, int m, bool cheb)
{
  if (!s.V) NW_HIP(hipMalloc(&s.V, sizeof(T) * n * (m + 1)));
  if (!s.mat) NW_HIP(hipMalloc(&s.mat, sizeof(T) * nnz));  // sliced ELL order
  if (!s.u) NW_HIP(hipMalloc(&s.u, sizeof(T) * n));
  if (cheb) {
    if (!s.dinv) NW_HIP(hipMalloc(&s.dinv, sizeof(T) * n));
    if (!s.cz) NW_HIP(hipMalloc(&s.cz, sizeof(T) * n));
    if (!s.cd) NW_HIP(hipMalloc(&s.cd, sizeof(T) * n));
    if (!s.cz2) NW_HIP(hipMalloc(&s.cz2, sizeof(T) * n));
  }
  return NW_OK;
}

// The solver state of one call.
struct Solve {
  Ctx* c;
  Krylov* k;
  const nw_solver_cfg* cfg;
  int G;  // lanes per row
  double lmax = 0, lmin = 0;
  int iters = 0;
  hipStream_t st = nullptr;

  // reduce nb x nv partials to nv values and copy them to the host
  int reduce(int nb, int nv, double* host)
  {
    k_reduce<<<nv, kBlock, 0, st>>>(k->partial, nb, nv, k->red);
    NW_LAUNCH_CHECK();
    NW_HIP(hipMemcpyAsync(k->h_pinned, k->red, sizeof(double) * nv,
                          hipMemcpyDeviceToHost, st));
    NW_HIP(hipStreamSynchronize(st));
    std::memcpy(host, k->h_pinned, sizeof(double) * nv);
    return NW_OK;
  }

  // y = alpha A x with A in sliced ELL order (the space's matrix copy)
  template <typename T>
  int spmv(const T* va, const T* x, T* y, T alpha)
  {
    const int n = (int)k->n;
    k_sell_spmv<T><<<grid_for(n, kBlock), kBlock, 0, st>>>(
      sell_view(c->g.sell), va, x, y, n, alpha);
    NW_LAUNCH_CHECK();
    return NW_OK;
  }

  // y = alpha A x with A the assembled CSR values
  int spmv_csr(const double* va, const double* x, double* y, double alpha)
  {
    const int n = (int)k->n;
    NW_ROW_LAUNCH(k_spmv, double, G, n, st, c->g.row_ptr, c->g.cols, va, x, y,
                  n, alpha);
    NW_LAUNCH_CHECK();
    return NW_OK;
  }

  template <typename T>
  int cheb_step(const T* va, const T* z, const T* v, const T* dinv, T* d, T* zo,
                T c1, T c2)
  {
    const int n = (int)k->n;
    k_sell_cheb_step<T><<<grid_for(n, kBlock), kBlock, 0, st>>>(
      sell_view(c->g.sell), va, z, v, dinv, d, zo, n, c1, c2);
    NW_LAUNCH_CHECK();
    return NW_OK;
  }

  // z = p(D^-1 A) D^-1 v, the Chebyshev polynomial (Jacobi is folded into
  // the matrix columns instead when it is the preconditioner)
  template <typename T>
  int cheb_apply(Space<T>& s, const T* v, T** zout)
  {
    const int n = (int)k->n;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    k_cheb_init<T><<<grid_for(n, kBlock), kBlock, 0, st>>>(v, s.dinv, s.cd, s.cz,
                                                          n, (T)(1.0 / theta));
    NW_LAUNCH_CHECK();
    T* z = s.cz;
    T* z2 = s.cz2;
    for (int it = 1; it < cfg->cheb_degree; ++it) {
      const double rho_new = 1.0 / (2.0 * sigma - rho);
      int rc = cheb_step<T>(s.mat, z, v, s.dinv, s.cd, z2, (T)(rho_new * rho),
                            (T)(2.0 * rho_new / delta));
      if (rc) return rc;
      std::swap(z, z2);
      rho = rho_new;
    }
    *zout = z;
    return NW_OK;
  }

  // W = A M^-1 v
  template <typename T>
  int apply_op(Space<T>& s, const T* v, T* W)
  {
    if (cfg->precond == NW_PREC_AMG) {
      int rc = amg_apply<T>(c->amg, v, s.cz, st);
      if (rc) return rc;
      return spmv<T>(s.mat, s.cz, W, (T)1);
    }
    if (cfg->precond == NW_PREC_CHEBYSHEV) {
      T* z = nullptr;
      int rc = cheb_apply<T>(s, v, &z);
      if (rc) return rc;
      return spmv<T>(s.mat, z, W, (T)1);
    }
    return spmv<T>(s.mat, v, W, (T)1);
  }

  // Q = V[0..nq), y = V[nq], z = V[nq+1] (if has_z):
  // qy[k] = Q_k.y, qz[k] = Q_k.z, yy, yz
  template <typename T>
  int arnoldi_dots(const T* V, int nq, bool has_z, double* qy, double* qz,
                   double* yy, double* yz)
  {
    const int64_t n = k->n;
    const int ntiles = std::max(1, (nq + kTile - 1) / kTile);
    const int nb = kRowBlocks * ntiles;
    k_arnoldi_dots<T><<<nb, kBlock, 0, st>>>(
      V, n, nq, V + (int64_t)nq * n, has_z ? V + (int64_t)(nq + 1) * n : nullptr,
      ntiles, k->partial);
    NW_LAUNCH_CHECK();
    // reduce per (tile, value) over the row blocks: partial layout is
    // [rowblock][tile][value] = [rowblock][ntiles * kDotVals]
    std::vector<double> r(ntiles * kDotVals);
    int rc = reduce(kRowBlocks, ntiles * kDotVals, r.data());
    if (rc) return rc;
    for (int q = 0; q < nq; ++q) {
      const int t = q / kTile, kk = q % kTile;
      qy[q] = r[t * kDotVals + kk];
      qz[q] = r[t * kDotVals + kTile + kk];
    }
    *yy = r[2 * kTile];
    *yz = r[2 * kTile + 1];
    return NW_OK;
  }

  template <typename T>
  int arnoldi_update(T* V, int nq, const std::vector<double>& coef, bool has_z)
  {
    NW_HIP(hipMemcpyAsync(k->coef, coef.data(), sizeof(double) * coef.size(),
                          hipMemcpyHostToDevice, st));
    k_arnoldi_update<T><<<stream_grid(k->n), kBlock, 0, st>>>(
      V, k->n, nq, k->coef, has_z ? 1 : 0);
    NW_LAUNCH_CHECK();
    return NW_OK;
  }

  template <typename T>
  int combine(const T* V, int nk, const double* coef, T* u)
  {
    NW_HIP(hipMemcpyAsync(k->coef, coef, sizeof(double) * nk,
                          hipMemcpyHostToDevice, st));
    k_combine<T><<<stream_grid(k->n), kBlock, 0, st>>>(V, k->n, nk, k->coef, u);
    NW_LAUNCH_CHECK();
    return NW_OK;
  }

  // r = b - A x in FP64; returns ||r||
  int residual(const double* x, double* rnorm)
  {
    const int n = (int)k->n;
    // one pass over the rows, at most kPartialCap partials
    const int nb = (int)std::min<int64_t>(grid_for((int64_t)n * G, kBlock),
                                          kPartialCap);
    switch (G) {
    case 1:
      k_residual<1><<<nb, kBlock, 0, st>>>(c->g.row_ptr, c->g.cols, c->g.vals, x, c->g.rhs, k->r, n, k->partial);
      break;
    case 2:
      k_residual<2><<<nb, kBlock, 0, st>>>(c->g.row_ptr, c->g.cols, c->g.vals, x, c->g.rhs, k->r, n, k->partial);
      break;
    case 4:
      k_residual<4><<<nb, kBlock, 0, st>>>(c->g.row_ptr, c->g.cols, c->g.vals, x, c->g.rhs, k->r, n, k->partial);
      break;
    case 8:
      k_residual<8><<<nb, kBlock, 0, st>>>(c->g.row_ptr, c->g.cols, c->g.vals, x, c->g.rhs, k->r, n, k->partial);
      break;
    case 16:
      k_residual<16><<<nb, kBlock, 0, st>>>(c->g.row_ptr, c->g.cols, c->g.vals, x, c->g.rhs, k->r, n, k->partial);
      break;
    default:
      k_residual<32><<<nb, kBlock, 0, st>>>(c->g.row_ptr, c->g.cols, c->g.vals, x, c->g.rhs, k->r, n, k->partial);
    }
    NW_LAUNCH_CHECK();
    double r2 = 0;
    int rc = reduce(nb, 1, &r2);
    *rnorm = std::sqrt(r2);
    return rc;
  }

  // ||v||^2 and v.w of FP64 vectors (w may equal v)
  int dot64(const double* v, const double* w, double* vw)
  {
    k_arnoldi_dots<double><<<kRowBlocks, kBlock, 0, st>>>(
      v, k->n, 0, v, w, 1, k->partial);
    NW_LAUNCH_CHECK();
    std::vector<double> r(kDotVals);
    int rc = reduce(kRowBlocks, kDotVals, r.data());
    *vw = r[2 * kTile + 1];
    return rc;
  }

  // One GMRES(m) cycle in type T on A d = r (r = k->r, FP64, ||r|| = beta),
  // then x += M^-1 Q y in FP64.  Stops early when the Arnoldi residual
  // estimate reaches stop_abs.
  template <typename T>
  int cycle(Space<T>& s, double* x, double beta, double stop_abs, int max_it)
  {
    const int64_t n = k->n;
    const int m = std::min(cfg->kspace, max_it);
    // Hr: unrotated Hessenberg (the Arnoldi relation A M^-1 Q = Q H);
    // R: the same columns after the Givens rotations (least squares)
    std::vector<double> Hr((size_t)(m + 1) * m, 0.0), R((size_t)(m + 1) * m, 0.0);
    std::vector<double> cs(m, 1.0), sn(m, 0.0), g(m + 2, 0.0), sv(kMaxBasis),
      tv(kMaxBasis), hprev(kMaxBasis, 0.0), Hs(kMaxBasis), hj(kMaxBasis),
      cv(kMaxBasis), yls(kMaxBasis), coef;
    auto H = [&](int i, int j) -> double& { return Hr[(size_t)i * m + j]; };
    auto Rm = [&](int i, int j) -> double& { return R[(size_t)i * m + j]; };
    T* V = s.V;
    k_convert<T><<<grid_for(n, kBlock), kBlock, 0, st>>>(k->r, V, 1.0 / beta, n);
    NW_LAUNCH_CHECK();
    g[0] = beta;
    double rho_prev = 1.0;
    int kdim = 0;
    for (int j = 0; j <= m; ++j) {
      const bool grow = j < m;
      int rc;
      if (grow) {
        rc = apply_op<T>(s, V + (int64_t)j * n, V + (int64_t)(j + 1) * n);
        if (rc) return rc;
      }
      double yy = 0, yz = 0;
      rc = arnoldi_dots<T>(V, j, grow, sv.data(), tv.data(), &yy, &yz);
      if (rc) return rc;
      if (!(yy > 0) || !std::isfinite(yy)) break;
      const double sigma = 1.0 / std::sqrt(yy);
      double ss = 0.0;
      for (int q = 0; q < j; ++q) {
        sv[q] *= sigma;  // s = Q^T u, u = sigma y
        ss += sv[q] * sv[q];
      }
      const double rho = ss < 1.0 ? std::sqrt(1.0 - ss) : 0.0;
      if (j >= 1) {
        // finish column j-1: A M^-1 q_{j-1} = Q (hprev + nu s) + nu rho q_j
        const int col = j - 1;
        const double nu = (1.0 / sigma) / rho_prev;
        for (int q = 0; q < j; ++q) H(q, col) = hprev[q] + nu * sv[q];
        H(j, col) = nu * rho;
        for (int q = 0; q <= j; ++q) Rm(q, col) = H(q, col);
        for (int q = 0; q < col; ++q) {
          const double a = Rm(q, col), b = Rm(q + 1, col);
          Rm(q, col) = cs[q] * a + sn[q] * b;
          Rm(q + 1, col) = -sn[q] * a + cs[q] * b;
        }
        const double a = Rm(col, col), b = Rm(col + 1, col);
        const double rr = std::sqrt(a * a + b * b);
        cs[col] = rr > 0 ? a / rr : 1.0;
        sn[col] = rr > 0 ? b / rr : 0.0;
        Rm(col, col) = rr;
        Rm(col + 1, col) = 0.0;
        g[col + 1] = -sn[col] * g[col];
        g[col] = cs[col] * g[col];
        ++iters;
        kdim = j;
        if (std::fabs(g[col + 1]) <= stop_abs) break;
      }
      if (!grow || !(rho > 1e-12)) break;  // end of cycle / lucky breakdown
      // tentative column j from the dots of z = A M^-1 u, z = sigma z'
      double st_ = 0.0;
      for (int q = 0; q < j; ++q) {
        tv[q] *= sigma;  // t = Q^T z
        st_ += sv[q] * tv[q];
      }
      const double uz = sigma * sigma * yz;
      for (int q = 0; q <= j; ++q) {  // Hs = H[0..j][0..j-1] s
        double a = 0.0;
        for (int l = std::max(0, q - 1); l < j; ++l) a += H(q, l) * sv[l];
        Hs[q] = a;
      }
      for (int q = 0; q < j; ++q) hj[q] = (tv[q] - Hs[q]) / rho;
      hj[j] = ((uz - st_) / rho - Hs[j]) / rho;
      for (int q = 0; q <= j; ++q) cv[q] = Hs[q] + rho * hj[q];
      const double e = cv[j] / rho;
      // V[j] <- q_j = (sigma/rho) y - sum (s_q/rho) Q_q
      // V[j+1] <- sigma z' - e sigma y - sum (c_q - e s_q) Q_q
      coef.assign(3 + 2 * j, 0.0);
      coef[0] = sigma / rho;
      coef[1] = sigma;
      coef[2] = e * sigma;
      for (int q = 0; q < j; ++q) {
        coef[3 + q] = sv[q] / rho;
        coef[3 + j + q] = cv[q] - e * sv[q];
      }
      rc = arnoldi_update<T>(V, j, coef, true);
      if (rc) return rc;
      for (int q = 0; q <= j; ++q) hprev[q] = hj[q];
      rho_prev = rho;
    }
    if (kdim == 0) return NW_OK;
    // least squares R y = g, u = Q y, x += M^-1 u
    for (int i = kdim - 1; i >= 0; --i) {
      double t = g[i];
      for (int l = i + 1; l < kdim; ++l) t -= Rm(i, l) * yls[l];
      yls[i] = t / Rm(i, i);
    }
    int rc = combine<T>(V, kdim, yls.data(), s.u);
    if (rc) return rc;
    const T* u = s.u;
    if (cfg->precond == NW_PREC_CHEBYSHEV) {
      T* z = nullptr;
      rc = cheb_apply<T>(s, s.u, &z);
      if (rc) return rc;
      u = z;
    } else if (cfg->precond == NW_PREC_AMG) {
      rc = amg_apply<T>(c->amg, s.u, s.cz, st);
      if (rc) return rc;
      u = s.cz;
    }
    k_axpy_prec<T><<<grid_for(n, kBlock), kBlock, 0, st>>>(
      x, u, cfg->precond == NW_PREC_JACOBI ? k->dinv : nullptr, n);
    NW_LAUNCH_CHECK();
    return NW_OK;
  }

  // lambda_max of D^-1 A by power iteration from a seeded vector (FP64)
  int power_lambda(double* lam)
  {
    const int64_t n = k->n;
    double* xa = k->pe;
    double* ya = k->pe + n;
    k_seed_vec<<<grid_for(n, kBlock), kBlock, 0, st>>>(xa, n, 12345);
    NW_LAUNCH_CHECK();
    double l = 0.0;
    for (int it = 0; it < 30; ++it) {
      double xx = 0;
      int rc = dot64(xa, xa, &xx);
      if (rc) return rc;
      const double inv = 1.0 / std::sqrt(xx);
      rc = spmv_csr(c->g.vals, xa, ya, inv);
      if (rc) return rc;
      k_vmul<<<grid_for(n, kBlock), kBlock, 0, st>>>(ya, k->dinv, n);
      NW_LAUNCH_CHECK();
      double xy = 0;
      rc = dot64(xa, ya, &xy);
      if (rc) return rc;
      l = xy * inv;
      std::swap(xa, ya);
    }
    *lam = l;
    return NW_OK;
  }
};

}  // namespace nw

using namespace nw;

extern "C" {

int nw_linsys_amg_level(nw_ctx* c, int level, int64_t* rows, int64_t* nnz)
{
  NW_REQUIRE(c && rows && nnz, NW_E_ARG, "null argument");
  NW_REQUIRE(c->amg, NW_E_ARG, "no AMG hierarchy (run an AMG solve first)");
  NW_REQUIRE(amg_level_info(c->amg, level, rows, nnz) == NW_OK, NW_E_ARG,
             "level %d out of range", level);
  return NW_OK;
}

int nw_linsys_spmv(nw_ctx* c, const double* d_x, double* d_y)
{
  NW_REQUIRE(c && c->g.ndof == 1 && d_x && d_y, NW_E_STATE,
             "spmv needs a built 1-dof system");
  const int n = (int)c->g.nnode;
  const int G = lanes_per_row((double)c->g.nnzb / std::max<int64_t>(n, 1));
  NW_ROW_LAUNCH(k_spmv, double, G, n, c->stream, c->g.row_ptr, c->g.cols,
                c->g.vals, d_x, d_y, n, 1.0);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

int nw_linsys_solve(nw_ctx* c, const nw_solver_cfg* cfg, nw_solve_stats* out)
{
  NW_REQUIRE(c && cfg && out, NW_E_ARG, "null argument");
  NW_REQUIRE(c->g.ndof == 1, NW_E_STATE, "solve needs a built 1-dof system");
  NW_REQUIRE(cfg->kspace >= 1 && cfg->kspace < kMaxBasis, NW_E_ARG,
             "kspace must be in [1, %d)", kMaxBasis);
  NW_REQUIRE(cfg->tol > 0 && cfg->max_iters >= 0, NW_E_ARG, "bad tolerance");
  NW_REQUIRE(cfg->precond >= NW_PREC_NONE && cfg->precond <= NW_PREC_AMG,
             NW_E_ARG, "bad preconditioner");
  NW_REQUIRE(cfg->precond != NW_PREC_CHEBYSHEV ||
               (cfg->cheb_degree >= 1 && cfg->cheb_ratio > 1.0),
             NW_E_ARG, "bad Chebyshev parameters");
  NW_REQUIRE(cfg->mode == NW_SOLVE_FP64 || cfg->mode == NW_SOLVE_FP32_REFINE,
             NW_E_ARG, "bad solve mode");
  NW_REQUIRE(cfg->x0 == NW_X0_ZERO || cfg->x0 == NW_X0_SOLUTION ||
               cfg->x0 == NW_X0_LEARNED,
             NW_E_ARG, "bad x0 mode %d", cfg->x0);
  NW_REQUIRE(cfg->x0 != NW_X0_LEARNED || c->mpnn, NW_E_STATE,
             "x0 = learned needs a network (nw_linsys_set_learned_guess)");
  std::memset(out, 0, sizeof(*out));
  NW_HIP(hipSetDevice(c->device));
  const int64_t n = c->g.nnode, nnz = c->g.nnzb;
  const bool cheb = cfg->precond == NW_PREC_CHEBYSHEV;
  const bool amgp = cfg->precond == NW_PREC_AMG;
  const bool f32 = cfg->mode == NW_SOLVE_FP32_REFINE;
  if (c->kry && (c->kry->n != n || c->kry->m < cfg->kspace)) {
    free_krylov(c->kry);
    c->kry = nullptr;
  }
  if (!c->kry) {
    c->kry = new Krylov();
    c->kry->n = n;
    c->kry->nnz = nnz;
    c->kry->m = cfg->kspace;
  }
  Krylov* k = c->kry;
  if (!k->dinv) {
    NW_HIP(hipMalloc(&k->dinv, sizeof(double) * n));
    NW_HIP(hipMalloc(&k->r, sizeof(double) * n));
    NW_HIP(hipMalloc(&k->partial, sizeof(double) * kPartialCap));
    NW_HIP(hipMalloc(&k->red, sizeof(double) * kRedCap));
    NW_HIP(hipMalloc(&k->coef, sizeof(double) * (3 + 2 * kMaxBasis)));
    NW_HIP(hipMalloc(&k->bad, sizeof(int)));
    NW_HIP(hipHostMalloc(&k->h_pinned, sizeof(double) * kRedCap));
    NW_HIP(hipEventCreate(&k->ev0));
    NW_HIP(hipEventCreate(&k->ev1));
  }
  if (!c->g.sell.ci) {
    int rc0 = sell_build(c->g.row_ptr, c->g.cols, (int)n, c->g.sell, c->stream,
                         false, true);
    if (rc0) return rc0;
  }
  const int64_t slen = c->g.sell.len;
  int rc = f32 ? alloc_space(k->f, n, slen, k->m, cheb || amgp)
               : alloc_space(k->d, n, slen, k->m, cheb || amgp);
  if (rc) return rc;
  if (cheb && !k->pe) NW_HIP(hipMalloc(&k->pe, sizeof(double) * 2 * n));

  hipStream_t st = c->stream;
  if (amgp && (!c->amg || !cfg->reuse_preconditioner)) {
    amg_free(c->amg);
    c->amg = nullptr;
    rc = amg_setup(c, &c->amg);
    if (rc) return rc;
    out->setup_ms = 1e3 * amg_setup_seconds(c->amg);
  }
  if (amgp) {
    double narrow_s = 0.0;  // the FP32 hierarchy, made on first use
    rc = f32 ? amg_prepare<float>(c, c->amg, &narrow_s) : amg_prepare<double>(c, c->amg);
    if (rc) return rc;
    out->setup_ms += 1e3 * narrow_s;
    out->precond_levels = amg_levels(c->amg);
  }
  NW_HIP(hipEventRecord(k->ev0, st));
  Solve S{c, k, cfg, lanes_per_row((double)nnz / std::max<int64_t>(n, 1))};
  S.st = st;
  // preconditioner setup: 1/diag, the matrix in the cycle type
  NW_HIP(hipMemsetAsync(k->bad, 0, sizeof(int), st));
  k_diag_inv<<<grid_for(n, kBlock), kBlock, 0, st>>>(c->g.diag, c->g.vals,
                                                    k->dinv, (int)n, k->bad);
  NW_LAUNCH_CHECK();
  int bad = 0;
  NW_HIP(hipMemcpyAsync(&bad, k->bad, sizeof(int), hipMemcpyDeviceToHost, st));
  NW_HIP(hipStreamSynchronize(st));
  NW_REQUIRE(!bad, NW_E_NONFINITE, "zero or non-finite diagonal in the matrix");
  const double* colscale = (cfg->precond == NW_PREC_JACOBI) ? k->dinv : nullptr;
  const Sell& sl = c->g.sell;
  if (f32) {
    k_sell_values<float><<<grid_for(sl.len, kBlock), kBlock, 0, st>>>(
      sl.slot, sl.ci, c->g.vals, colscale, k->f.mat, sl.len);
    if (cheb)
      k_convert<float><<<grid_for(n, kBlock), kBlock, 0, st>>>(k->dinv, k->f.dinv, 1.0, n);
  } else {
    k_sell_values<double><<<grid_for(sl.len, kBlock), kBlock, 0, st>>>(
      sl.slot, sl.ci, c->g.vals, colscale, k->d.mat, sl.len);
    if (cheb)
      k_convert<double><<<grid_for(n, kBlock), kBlock, 0, st>>>(k->dinv, k->d.dinv, 1.0, n);
  }
  NW_LAUNCH_CHECK();
  if (cheb) {
    double lam = 0;
    rc = S.power_lambda(&lam);
    if (rc) return rc;
    S.lmax = 1.1 * lam;
    S.lmin = S.lmax / cfg->cheb_ratio;
  }
  double* x = c->g.sln;
  if (cfg->x0 == NW_X0_ZERO) NW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
  // History of starting guesses (the learned guess's inputs and recorded
  // training pairs): xh[0..nh-1] are the guesses of the latest (up to
  // kHist) solves that started from the solution vector with distinct
  // guesses, newest first.  A solve from x0 = 0, or from the same guess as
  // the last one (another method on the same step), leaves the history as
  // it is.  The buffers rotate; only the new guess is copied.
  if (cfg->x0 != NW_X0_ZERO) {
    bool shift = true;
    if (c->nh > 0) {
      NW_HIP(hipMemsetAsync(k->bad, 0, sizeof(int), st));
      k_differs<<<grid_for(n, kBlock), kBlock, 0, st>>>(x, c->xh[0], n, k->bad);
      NW_LAUNCH_CHECK();
      int differs = 0;
      NW_HIP(hipMemcpyAsync(&differs, k->bad, sizeof(int),
                            hipMemcpyDeviceToHost, st));
      NW_HIP(hipStreamSynchronize(st));
      shift = differs != 0;
    }
    if (shift) {
      double* last = c->xh[kHist - 1];
      for (int j = kHist - 1; j > 0; --j) c->xh[j] = c->xh[j - 1];
      if (!last) NW_HIP(hipMalloc(&last, sizeof(double) * n));
      c->xh[0] = last;
      NW_HIP(hipMemcpyAsync(c->xh[0], x, sizeof(double) * n,
                            hipMemcpyDeviceToDevice, st));
      c->nh = std::min(c->nh + 1, kHist);
    }
  }
  double bb = 0;
  rc = S.dot64(c->g.rhs, c->g.rhs, &bb);
  if (rc) return rc;
  out->bnorm = std::sqrt(bb);
  const double bnorm = out->bnorm > 0 ? out->bnorm : 1.0;
  double rnorm = 0;
  rc = S.residual(x, &rnorm);
  if (rc) return rc;
  out->guess = cfg->x0;
  if (cfg->x0 == NW_X0_LEARNED) {
    // the network's prediction from (A, b, x0); kept only if it lowers the
    // residual, otherwise GMRES starts from x0 (the previous-step guess)
    if (!k->xg) NW_HIP(hipMalloc(&k->xg, sizeof(double) * n));
    rc = mpnn_predict_device(c->mpnn, c->xh, c->nh, k->xg);
    out->guess_ms = mpnn_last_ms(c->mpnn);
    double rl = 0;
    if (rc == NW_OK) rc = S.residual(k->xg, &rl);
    if (rc != NW_OK && rc != NW_E_NONFINITE) return rc;
    if (rc == NW_OK && std::isfinite(rl) && rl < rnorm) {
      NW_HIP(hipMemcpyAsync(x, k->xg, sizeof(double) * n,
                            hipMemcpyDeviceToDevice, st));
      rnorm = rl;
    } else {
      out->guess = NW_X0_SOLUTION;
      rc = S.residual(x, &rnorm);  // the cycle starts from the residual in k->r
      if (rc) return rc;
    }
    rc = NW_OK;
  }
  out->x0_rel = rnorm / bnorm;
  while (rnorm > cfg->tol * bnorm && S.iters < cfg->max_iters &&
         std::isfinite(rnorm)) {
    const double stop_abs = cfg->tol * bnorm;
    const int budget = cfg->max_iters - S.iters;
    const int before = S.iters;
    if (f32) {
      const double inner = std::max(cfg->inner_tol, 0.0) * rnorm;
      rc = S.cycle<float>(k->f, x, rnorm, std::max(stop_abs, inner), budget);
    } else {
      rc = S.cycle<double>(k->d, x, rnorm, stop_abs, budget);
    }
    if (rc) return rc;
    ++out->refinements;
    rc = S.residual(x, &rnorm);
    if (rc) return rc;
    if (S.iters == before) break;  // no progress possible (breakdown)
  }
  NW_HIP(hipEventRecord(k->ev1, st));
  NW_HIP(hipEventSynchronize(k->ev1));
  float ms = 0;
  NW_HIP(hipEventElapsedTime(&ms, k->ev0, k->ev1));
  out->ms = ms;
  out->iters = S.iters;
  out->reorthogonalizations = S.iters;
  out->final_abs = rnorm;
  out->final_rel = rnorm / bnorm;
  out->converged = rnorm <= cfg->tol * bnorm;
  if (!std::isfinite(rnorm)) {
    set_error("solve diverged (non-finite residual)");
    return NW_E_NONFINITE;
  }
  if (!out->converged) {
    set_error("GMRES stopped at %d iterations, relative residual %.3e",
              out->iters, out->final_rel);
    return NW_E_NOCONV;
  }
  return NW_OK;
}

}  // extern "C"
