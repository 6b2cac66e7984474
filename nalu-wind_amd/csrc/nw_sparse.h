// nw_sparse.h -- CSR kernels shared by the Krylov solver and the AMG
// preconditioner: SpMV with G lanes per row (rows are contiguous, so the
// lanes of a row read consecutive entries), the fused Chebyshev step and
// its zero-start initialisation.
#pragma once

#include "nw_internal.h"

namespace nw {

// Lanes per row for rows of `avg` entries: about two entries per lane
// (measured on the 7-entry rows of the hex pressure system: 4 lanes beat 8
// by 4-12 %, 2 lanes are 1.5x slower, 1 lane 2.7x).
inline int lanes_per_row(double avg)
{
  return avg <= 9.0 ? 4 : (avg <= 18.0 ? 8 : (avg <= 36.0 ? 16 : 32));
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
  T s = 0;
  if (row < n) {
    for (int k = rp[row] + g; k < rp[row + 1]; k += G) s += va[k] * x[ci[k]];
  }
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < n && g == 0) y[row] = alpha * s;
}

// Chebyshev step: r = v - A z; d = c1 d + c2 dinv r; zo = z + d.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_cheb_step(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const T* __restrict__ va, const T* __restrict__ z, const T* __restrict__ v,
  const T* __restrict__ dinv, T* __restrict__ d, T* __restrict__ zo, int n,
  T c1, T c2)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int g = threadIdx.x & (G - 1);
  T s = 0;
  if (row < n) {
    for (int k = rp[row] + g; k < rp[row + 1]; k += G) s += va[k] * z[ci[k]];
  }
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < n && g == 0) {
    const T dn = c1 * d[row] + c2 * dinv[row] * (v[row] - s);
    d[row] = dn;
    zo[row] = z[row] + dn;
  }
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


// y += A x (prolongation-correction); G lanes per row.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_spmv_add(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const T* __restrict__ va, const T* __restrict__ x, T* __restrict__ y, int n)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int g = threadIdx.x & (G - 1);
  T s = 0;
  if (row < n) {
    for (int k = rp[row] + g; k < rp[row + 1]; k += G) s += va[k] * x[ci[k]];
  }
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
  T s = 0;
  if (row < n) {
    for (int k = rp[row] + g; k < rp[row + 1]; k += G) s += va[k] * x[ci[k]];
  }
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < n && g == 0) r[row] = b[row] - s;
}

// launch helper: dispatch a G-templated row kernel by lanes-per-row
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
