// nw_sparse.h -- sparse kernels shared by the Krylov solver and the AMG
// preconditioner: SpMV, residual and the fused Chebyshev step in two
// layouts.  Sliced ELL (one lane per row, independent loads) serves short
// rows: the fine matrix and the prolongators.  CSR with G lanes per row
// serves long rows (coarse AMG operators, restrictions) and the assembled
// matrix as it is (power iteration, nw_linsys_spmv).
#pragma once

#include "nw_internal.h"

namespace nw {

// Lanes per row for CSR rows of `avg` entries: each lane takes kCsrUnroll
// entries per step as independent loads, so a row of average length is
// covered by about one step (measured on the 20-entry AMG coarse operator
// at 256^3: 8 lanes x 4 loads beat 16 lanes x 1 load by ~1.6x).
constexpr int kCsrUnroll = 4;
inline int lanes_per_row(double avg)
{
  int g = 1;
  while (g < 32 && g * kCsrUnroll < avg) g *= 2;
  return g;
}

// Partial sum of row `row` over the entries g, g+G, ... (lane g of G),
// kCsrUnroll entries per step loaded before any is used; padding entries
// repeat the row's last column with value 0.  Rows are contiguous, so the
// G lanes of a row read consecutive entries.
template <typename T, int G>
__device__ inline T csr_dot(const int32_t* __restrict__ rp,
                            const int32_t* __restrict__ ci,
                            const T* __restrict__ va, const T* __restrict__ x,
                            int row, int g)
{
  const int k0 = rp[row], k1 = rp[row + 1];
  T s = 0;
  for (int k = k0 + g; k < k1; k += G * kCsrUnroll) {
    int c[kCsrUnroll];
    T v[kCsrUnroll];
#pragma unroll
    for (int u = 0; u < kCsrUnroll; ++u) {
      const int kk = k + u * G;
      const bool ok = kk < k1;
      c[u] = ci[ok ? kk : k1 - 1];
      v[u] = ok ? va[kk] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kCsrUnroll; ++u) s += v[u] * x[c[u]];
  }
  return s;
}

// y = alpha * (A x); G lanes per row.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_spmv(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const T* __restrict__ va, const T* __restrict__ x, T* __restrict__ y, int n,
  T alpha)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int g = threadIdx.x & (G - 1);
  T s = row < n ? csr_dot<T, G>(rp, ci, va, x, row, g) : (T)0;
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < n && g == 0) y[row] = alpha * s;
}

// d = dinv * v * inv_theta; z = d
template <typename T>
__global__ void k_cheb_init(const T* __restrict__ v, const T* __restrict__ dinv,
                            T* __restrict__ d, T* __restrict__ z, int n,
                            T inv_theta)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T di = dinv[i] * v[i] * inv_theta;
  d[i] = di;
  z[i] = di;
}

// Chebyshev step: r = v - A z; d = c1 d + c2 dinv r; zo = z + d (with ADD,
// zo += z + d: the last step of a smoother that corrects zo).
template <typename T, int G, bool ADD = false>
__global__ __launch_bounds__(kBlock) void k_cheb_step(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const T* __restrict__ va, const T* __restrict__ z, const T* __restrict__ v,
  const T* __restrict__ dinv, T* __restrict__ d, T* __restrict__ zo, int n,
  T c1, T c2)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int g = threadIdx.x & (G - 1);
  T s = row < n ? csr_dot<T, G>(rp, ci, va, z, row, g) : (T)0;
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < n && g == 0) {
    const T dn = c1 * d[row] + c2 * dinv[row] * (v[row] - s);
    d[row] = dn;
    if (ADD)
      zo[row] += z[row] + dn;
    else
      zo[row] = z[row] + dn;
  }
}

// y += A x (prolongation-correction); G lanes per row.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_spmv_add(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const T* __restrict__ va, const T* __restrict__ x, T* __restrict__ y, int n)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int g = threadIdx.x & (G - 1);
  T s = row < n ? csr_dot<T, G>(rp, ci, va, x, row, g) : (T)0;
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < n && g == 0) y[row] += s;
}

// r = b - A x; G lanes per row.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_resid(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const T* __restrict__ va, const T* __restrict__ x, const T* __restrict__ b,
  T* __restrict__ r, int n)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int g = threadIdx.x & (G - 1);
  T s = row < n ? csr_dot<T, G>(rp, ci, va, x, row, g) : (T)0;
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < n && g == 0) r[row] = b[row] - s;
}

// r = b - A x, and the start of a Chebyshev smoother on r:
// d = z = dinv r inv_theta (k_resid followed by k_cheb_init).
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_resid_cheb_init(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const T* __restrict__ va, const T* __restrict__ x, const T* __restrict__ b,
  T* __restrict__ r, const T* __restrict__ dinv, T* __restrict__ d,
  T* __restrict__ z, int n, T inv_theta)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int g = threadIdx.x & (G - 1);
  T s = row < n ? csr_dot<T, G>(rp, ci, va, x, row, g) : (T)0;
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < n && g == 0) {
    const T ri = b[row] - s;
    r[row] = ri;
    const T di = dinv[row] * ri * inv_theta;
    d[row] = di;
    z[row] = di;
  }
}

// ------------------------------------------------------------ sliced ELL
// One lane per row, a wave per slice: for every j the 64 lanes read 64
// consecutive values and columns, and all loads of a row are independent
// (issued eight at a time), so a wave keeps a whole row's gathers in
// flight instead of walking a dependent chain.  Padding reads x at a valid
// column with value 0.  The row sum runs over j in order.  A slice in
// offset form (A.dia_w[k] > 0) takes its columns as row + the slice's
// offsets, which are the same for the whole wave (scalar loads), and reads
// no column array; a slice in delta form (A.d16[k]) reads 16-bit column -
// row deltas instead of the 32-bit columns.
template <typename T>
__device__ inline T sell_dot(const SellView& A, const T* __restrict__ va,
                             const T* __restrict__ x, int row)
{
  // rows of one slice are one wave: the slice index is wave-uniform
  const int k = __builtin_amdgcn_readfirstlane(row >> 6);
  const int64_t b = A.sp[k] + (row & 63);
  const int w = (int)((A.sp[k + 1] - A.sp[k]) >> 6);
  T s = 0;
  if (A.dia_w && A.dia_w[k] > 0) {
    const int32_t* off = A.dia_off + (int64_t)k * kDiaMax;
    for (int j0 = 0; j0 < w; j0 += 8) {
      int c[8];
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = min(j0 + u, w - 1);
        c[u] = row + off[j];
        v[u] = j0 + u < w ? va[b + 64 * (int64_t)j] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u] * x[c[u]];
    }
    return s;
  }
  if (A.d16 && A.d16[k]) {
    for (int j0 = 0; j0 < w; j0 += 8) {
      int c[8];
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = min(j0 + u, w - 1);
        c[u] = row + A.cd[b + 64 * (int64_t)j];
        v[u] = j0 + u < w ? va[b + 64 * (int64_t)j] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u] * x[c[u]];
    }
    return s;
  }
  for (int j0 = 0; j0 < w; j0 += 8) {
    int c[8];
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = min(j0 + u, w - 1);
      c[u] = A.ci[b + 64 * (int64_t)j];
      v[u] = j0 + u < w ? va[b + 64 * (int64_t)j] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u] * x[c[u]];
  }
  return s;
}

// y = alpha * (A x)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sell_spmv(
  SellView A, const T* __restrict__ va, const T* __restrict__ x, T* __restrict__ y, int n,
  T alpha)
{
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  y[row] = alpha * sell_dot(A, va, x, row);
}

// y += A x
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sell_spmv_add(
  SellView A, const T* __restrict__ va, const T* __restrict__ x, T* __restrict__ y, int n)
{
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  y[row] += sell_dot(A, va, x, row);
}

// Chebyshev step: r = v - A z; d = c1 d + c2 dinv r; zo = z + d (with ADD,
// zo += z + d).
template <typename T, bool ADD = false>
__global__ __launch_bounds__(kBlock) void k_sell_cheb_step(
  SellView A, const T* __restrict__ va, const T* __restrict__ z, const T* __restrict__ v,
  const T* __restrict__ dinv, T* __restrict__ d, T* __restrict__ zo, int n,
  T c1, T c2)
{
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const T s = sell_dot(A, va, z, row);
  const T dn = c1 * d[row] + c2 * dinv[row] * (v[row] - s);
  d[row] = dn;
  if (ADD)
    zo[row] += z[row] + dn;
  else
    zo[row] = z[row] + dn;
}

// r = b - A x
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sell_resid(
  SellView A, const T* __restrict__ va, const T* __restrict__ x, const T* __restrict__ b,
  T* __restrict__ r, int n)
{
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  r[row] = b[row] - sell_dot(A, va, x, row);
}

// r = b - A x and d = z = dinv r inv_theta (k_sell_resid + k_cheb_init)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sell_resid_cheb_init(
  SellView A, const T* __restrict__ va, const T* __restrict__ x, const T* __restrict__ b,
  T* __restrict__ r, const T* __restrict__ dinv, T* __restrict__ d,
  T* __restrict__ z, int n, T inv_theta)
{
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const T ri = b[row] - sell_dot(A, va, x, row);
  r[row] = ri;
  const T di = dinv[row] * ri * inv_theta;
  d[row] = di;
  z[row] = di;
}

// out[p] = T(a[slot[p]] * (colscale ? colscale[ci[p]] : 1)), 0 for padding
template <typename T>
__global__ void k_sell_values(const int32_t* __restrict__ slot,
                              const int32_t* __restrict__ ci,
                              const double* __restrict__ a,
                              const double* __restrict__ colscale,
                              T* __restrict__ out, int64_t len)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  const int32_t k = slot[p];
  out[p] = k < 0 ? (T)0 : (T)(colscale ? a[k] * colscale[ci[p]] : a[k]);
}

// Builds the sliced ELL structure of the CSR graph (rp, ci; device arrays)
// with n rows into s, which must be new or left by a failed build (its
// arrays are released first).  Padding positions read the row's own index as column
// (square matrices), or with pad_last the row's last column, which keeps
// them inside the columns of a rectangular matrix.  With offsets (square
// matrices only) the slices that qualify are stored in offset form, and
// the others whose columns lie within 32767 of their rows get 16-bit
// deltas.
int sell_build(const int32_t* rp, const int32_t* ci, int n, Sell& s,
               hipStream_t st, bool pad_last = false, bool offsets = false);

inline SellView sell_view(const Sell& s)
{
  return {s.sp, s.ci, s.dia_w, s.dia_off, s.d16, s.cd};
}

// launch helper: dispatch a G-templated row kernel by lanes-per-row
// (kernel<T, G>; NW_ROW_LAUNCH_ADD: kernel<T, G, true>)
#define NW_ROW_LAUNCH_ADD(kernel, T, G, n, stream, ...)                      \
  do {                                                                       \
    const unsigned nw_blocks_ = grid_for((int64_t)(n) * (G), kBlock);        \
    switch (G) {                                                             \
    case 1:                                                                  \
      kernel<T, 1, true><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);    \
      break;                                                                 \
    case 2:                                                                  \
      kernel<T, 2, true><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);    \
      break;                                                                 \
    case 4:                                                                  \
      kernel<T, 4, true><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);    \
      break;                                                                 \
    case 8:                                                                  \
      kernel<T, 8, true><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);    \
      break;                                                                 \
    case 16:                                                                 \
      kernel<T, 16, true><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);   \
      break;                                                                 \
    default:                                                                 \
      kernel<T, 32, true><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);   \
    }                                                                        \
  } while (0)

#define NW_ROW_LAUNCH(kernel, T, G, n, stream, ...)                          \
  do {                                                                       \
    const unsigned nw_blocks_ = grid_for((int64_t)(n) * (G), kBlock);        \
    switch (G) {                                                             \
    case 1:                                                                  \
      kernel<T, 1><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);          \
      break;                                                                 \
    case 2:                                                                  \
      kernel<T, 2><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);          \
      break;                                                                 \
    case 4:                                                                  \
      kernel<T, 4><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);          \
      break;                                                                 \
    case 8:                                                                  \
      kernel<T, 8><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);          \
      break;                                                                 \
    case 16:                                                                 \
      kernel<T, 16><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);         \
      break;                                                                 \
    default:                                                                 \
      kernel<T, 32><<<nw_blocks_, kBlock, 0, stream>>>(__VA_ARGS__);         \
    }                                                                        \
  } while (0)

}  // namespace nw
