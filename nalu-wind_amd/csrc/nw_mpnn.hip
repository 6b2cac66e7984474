// nw_mpnn.hip -- learned initial guess for the pressure-Poisson solve.
//
// A message-passing network on the graph of the assembled 1-dof matrix
// predicts the change of the solution from the previous-step guess x0:
//
//   history h_0 .. h_{nh-1}: the guesses of the latest nh <= 5 solves,
//                   newest first (h_0 = x0, the previous step's solution)
//   xb      = polynomial extrapolation in time of the newest min(nh, 3):
//             3 h_0 - 3 h_1 + h_2 (2 h_0 - h_1, h_0 for shorter histories)
//   features f_n = (r/s_r, log|a_nn| - mean, dirichlet,
//                   (h_0 - 3 h_1 + 3 h_2 - h_3)/s_d,
//                   (h_0 - 4 h_1 + 6 h_2 - 4 h_3 + h_4)/s_d),  r = b - A xb,
//                   the third and fourth backward differences of the history
//                   (0 when it is shorter), s_* the RMS of each vector, s_d
//                   of h_0 - h_1; differences only, so what is learned on
//                   earlier steps carries over
//   h0      = relu(W_in f + b_in)
//   m_l     = sum_{m != n} w_nm h_{l-1,m},  w_nm = sqrt|a_nm a_mn| / sqrt|a_nn a_mm|
//   h_l     = relu(W_self,l h_{l-1} + W_nbr,l m_l + b_l),   l = 1..L
//   a_k,n   = w_out,k . h_L + b_out,k                   k = 1, 2
//   delta_n = a_1,n f_3,n + a_2,n f_4,n   (the two history features)
//   x       = xb + s_o delta,  s_o = s_d, or s_e = rms(r / diag) without
//             history
// so the network learns per-node coefficients of the higher history
// differences on top of the quadratic extrapolation: a = (1, 0) is the cubic
// and a = (1, 1) the quartic extrapolation, other values damped or per-node
// blends.  The predictor's output vanishes with the history and has no
// offset of its own (an offset would be the smoothest, slowest error for
// GMRES).  The output head starts at zero: an untrained network gives the
// quadratic extrapolation 3 x0 - 3 xp + xpp, and training learns from the
// recorded solves how far to extrapolate (per node, from the local
// features).  Rows without
// off-diagonal entries (Dirichlet rows) are decoupled: their prediction is
// b_n / a_nn exactly, so no boundary error enters the guess.
//
// The reference has no counterpart (its pressure solve starts from x0 = 0,
// src/LinearSolver.C:209-254); the network only changes where GMRES starts,
// and GMRES to the same tolerance still decides the answer.  When the
// prediction has a larger residual than x0 (or is not finite) the solve falls
// back to x0, the previous-step guess.
//
// Training pairs are recorded from solves on the device: the features of
// (A, b, x0, xp) and the target (x - xb)/s_o in FP32 SoA buffers.  Training is
// mean-squared error with Adam; forward and backward run in the kernels
// below (FP32 storage and arithmetic, FP64 reductions in a fixed order, so a
// gradient is reproducible run to run), the Adam update on the host over the
// FP64 master parameters.
//
// Layouts: node activations are node-major [n][H] float rows, so a
// neighbour's whole row is one contiguous 256-byte read; features are SoA
// [5][n].  Parameters are the flat vector W_in[H][5], b_in[H], per layer
// W_self[H][H], W_nbr[H][H], b[H], then w_out[2][H], b_out[2]; the forward reads a
// transposed copy (k-major) so that the H outputs of one input are
// consecutive and wave-uniform.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "nw_internal.h"

namespace nw {

namespace {

constexpr int kF = 5;           // node features
constexpr int kTileN = 32;      // nodes per staging tile (weight gradients)
constexpr int kGradBlocks = 512;
constexpr int kStats = 6;       // feature statistics

__device__ inline double mp_wave_sum(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <int NA>
__device__ inline void mp_block_store(double (&acc)[NA], double* __restrict__ out)
{
  __shared__ double sh[kBlock / kWave][NA];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const double s = mp_wave_sum(acc[k]);
    if (lane == 0) sh[wv][k] = s;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NA; k += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) s += sh[w][k];
    out[k] = s;
  }
}

// out[v] = sum_b partial[b * nv + v], one workgroup per value, fixed order.
__global__ __launch_bounds__(kBlock) void k_mp_reduce(
  const double* __restrict__ partial, int nb, int nv, double* __restrict__ out)
{
  const int v = blockIdx.x;
  double a[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    a[0] += partial[(int64_t)b * nv + v];
  __shared__ double tot[1];
  mp_block_store<1>(a, tot);
  if (threadIdx.x == 0) out[v] = tot[0];
}

__device__ inline float relu(float z) { return z > 0.f ? z : 0.f; }

// eslot[code] = s for every off-diagonal slot (code = 2e + side): the
// two slots (L, R) and (R, L) of edge e
__global__ void k_edge_slots(const int32_t* __restrict__ code,
                             int32_t* __restrict__ eslot, int64_t nnz)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const int32_t c = code[s];
  if (c >= 0) eslot[c] = (int32_t)s;
}

// w = sqrt|a_LR a_RL| / sqrt|a_LL a_RR| for edge e = (L, R), written to
// both of its slots (one thread per edge; the diagonal slots keep 0).
__global__ void k_edge_weights(const int2* __restrict__ edges,
                               const int2* __restrict__ eslot,
                               const double* __restrict__ va,
                               const int32_t* __restrict__ diag,
                               float* __restrict__ w, int64_t nedge)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nedge) return;
  const int2 lr = edges[e];
  const int2 sl = eslot[e];
  if (sl.x < 0 || sl.y < 0) return;  // not in the graph
  const double dl = fabs(va[diag[lr.x]]), dr = fabs(va[diag[lr.y]]);
  const double a = fabs(va[sl.x]) * fabs(va[sl.y]);
  const float we = (dl > 0 && dr > 0) ? (float)(sqrt(a) / sqrt(dl * dr)) : 0.f;
  w[sl.x] = we;
  w[sl.y] = we;
}

// The history of guesses, newest first (h[j] valid for j < nh).
struct Hist {
  const double* h[kHist];
  int nh;
};

// xb = polynomial extrapolation in time of the newest min(nh, 3) guesses.
__global__ void k_base_guess(Hist H, double* __restrict__ xb, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h0 = H.h[0][i];
  xb[i] = H.nh >= 3 ? 3.0 * h0 - 3.0 * H.h[1][i] + H.h[2][i]
        : H.nh == 2 ? 2.0 * h0 - H.h[1][i]
                    : h0;
}

// r = b - A xb and the sums of b^2, h_0^2, r^2, (r/|d|)^2, log|d|,
// (h_0 - h_1)^2.
__global__ __launch_bounds__(kBlock) void k_feat_stats(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const double* __restrict__ va, const int32_t* __restrict__ diag,
  const double* __restrict__ b, const double* __restrict__ xb, Hist H,
  double* __restrict__ r, int n, double* __restrict__ partial)
{
  double acc[kStats] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    double ax = 0.0;
    for (int s = rp[i]; s < rp[i + 1]; ++s) ax += va[s] * xb[ci[s]];
    const double ri = b[i] - ax;
    r[i] = ri;
    const double d = fabs(va[diag[i]]);
    const double e = d > 0 ? ri / d : 0.0;
    const double h0 = H.h[0][i];
    acc[0] += b[i] * b[i];
    acc[1] += h0 * h0;
    acc[2] += ri * ri;
    acc[3] += e * e;
    acc[4] += log(d > 0 ? d : 1.0);
    const double dx = H.nh >= 2 ? h0 - H.h[1][i] : 0.0;
    acc[5] += dx * dx;
  }
  mp_block_store<kStats>(acc, partial + (int64_t)blockIdx.x * kStats);
}

struct FeatScale {
  double isr, mlog, isd, iso;
};

// Features (SoA) and, when xs is given, the target (xs - xb)/s_o.
__global__ void k_features(const int32_t* __restrict__ rp,
                           const double* __restrict__ va,
                           const int32_t* __restrict__ ci,
                           const int32_t* __restrict__ diag,
                           const double* __restrict__ b, Hist H,
                           const double* __restrict__ xb,
                           const double* __restrict__ r,
                           const double* __restrict__ xs, FeatScale sc,
                           float* __restrict__ feat, float* __restrict__ target,
                           int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = fabs(va[diag[i]]);
  int dir = 1;
  for (int s = rp[i]; s < rp[i + 1]; ++s)
    if (ci[s] != i && va[s] != 0.0) dir = 0;
  feat[i] = (float)(r[i] * sc.isr);
  feat[(int64_t)n + i] = (float)(log(d > 0 ? d : 1.0) - sc.mlog);
  feat[2 * (int64_t)n + i] = (float)dir;
  double f3 = 0.0, f4 = 0.0;
  if (H.nh >= 4) {
    const double h0 = H.h[0][i], h1 = H.h[1][i], h2 = H.h[2][i], h3 = H.h[3][i];
    f3 = (h0 - 3.0 * h1 + 3.0 * h2 - h3) * sc.isd;
    if (H.nh >= 5)
      f4 = (h0 - 4.0 * h1 + 6.0 * h2 - 4.0 * h3 + H.h[4][i]) * sc.isd;
  }
  feat[3 * (int64_t)n + i] = (float)f3;
  feat[4 * (int64_t)n + i] = (float)f4;
  if (xs) target[i] = (float)((xs[i] - xb[i]) * sc.iso);
}

template <int H>
__global__ __launch_bounds__(kBlock) void k_input_layer(
  const float* __restrict__ feat, const float* __restrict__ WinT,
  const float* __restrict__ bin, float* __restrict__ h0, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f[kF];
#pragma unroll
  for (int k = 0; k < kF; ++k) f[k] = feat[(int64_t)k * n + i];
  float4* out = reinterpret_cast<float4*>(h0 + (int64_t)i * H);
#pragma unroll
  for (int j4 = 0; j4 < H / 4; ++j4) {
    float z[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = 4 * j4 + q;
      float a = bin[j];
#pragma unroll
      for (int k = 0; k < kF; ++k) a = fmaf(WinT[k * H + j], f[k], a);
      z[q] = relu(a);
    }
    out[j4] = make_float4(z[0], z[1], z[2], z[3]);
  }
}

// m[i] = sum_s w_s h[col_s] over the row's slots: one thread per (node,
// 4-feature chunk), so the 16 threads of a node read a neighbour's 256-byte
// row together.
template <int H>
__global__ void k_aggregate(const float* __restrict__ h,
                            const int32_t* __restrict__ rp,
                            const int32_t* __restrict__ ci,
                            const float* __restrict__ w, float* __restrict__ m,
                            int n)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * (H / 4)) return;
  const int i = (int)(t / (H / 4)), k4 = (int)(t % (H / 4));
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int s = rp[i]; s < rp[i + 1]; ++s) {
    const float ws = w[s];
    if (ws != 0.f) {
      const float4 hc = reinterpret_cast<const float4*>(h + (int64_t)ci[s] * H)[k4];
      a.x = fmaf(ws, hc.x, a.x);
      a.y = fmaf(ws, hc.y, a.y);
      a.z = fmaf(ws, hc.z, a.z);
      a.w = fmaf(ws, hc.w, a.w);
    }
  }
  reinterpret_cast<float4*>(m + (int64_t)i * H)[k4] = a;
}

// delta = a1 f3 + a2 f4 with a_k = w_out,k . h + b_out,k
template <int H>
__global__ void k_output(const float* __restrict__ hL,
                         const float* __restrict__ wout, float b1, float b2,
                         const float* __restrict__ f3,
                         const float* __restrict__ f4,
                         float* __restrict__ delta, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* h = reinterpret_cast<const float4*>(hL + (int64_t)i * H);
  float a1 = b1, a2 = b2;
#pragma unroll
  for (int j4 = 0; j4 < H / 4; ++j4) {
    const float4 v = h[j4];
    const float hv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a1 = fmaf(wout[4 * j4 + q], hv[q], a1);
      a2 = fmaf(wout[H + 4 * j4 + q], hv[q], a2);
    }
  }
  delta[i] = a1 * f3[i] + a2 * f4[i];
}

// ------------------------------------------------- bfloat16 inference path
// bfloat16 = the upper 16 bits of an FP32 value, rounded to nearest even.
__host__ __device__ inline uint16_t f2bf(float f)
{
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline float bf2f(uint32_t b16)
{
  return __uint_as_float(b16 << 16);
}

__device__ inline uint32_t pack_bf2(float lo, float hi)
{
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTileNodes = 16;  // nodes per wave tile of the fused layer

// h0 = relu(W_in f + b_in) stored in bfloat16 (one thread per node)
template <int H>
__global__ __launch_bounds__(kBlock) void k_input_bf16(
  const float* __restrict__ feat, const float* __restrict__ WinT,
  const float* __restrict__ bin, uint16_t* __restrict__ h0, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f[kF];
#pragma unroll
  for (int k = 0; k < kF; ++k) f[k] = feat[(int64_t)k * n + i];
  uint4* out = reinterpret_cast<uint4*>(h0 + (int64_t)i * H);
#pragma unroll
  for (int j8 = 0; j8 < H / 8; ++j8) {
    float z[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int j = 8 * j8 + q;
      float a = bin[j];
#pragma unroll
      for (int k = 0; k < kF; ++k) a = fmaf(WinT[k * H + j], f[k], a);
      z[q] = relu(a);
    }
    out[j8] = make_uint4(pack_bf2(z[0], z[1]), pack_bf2(z[2], z[3]),
                         pack_bf2(z[4], z[5]), pack_bf2(z[6], z[7]));
  }
}

// One hidden layer, fused: for a tile of 16 nodes a wave
//   1. gathers the messages m_n = sum_s w_s h[col_s] in FP32 (H/8 lanes per
//      node, each on 8 consecutive features, so a neighbour's 2H-byte row
//      is read by consecutive lanes) and puts [h_n, bf16(m_n)] in LDS;
//   2. multiplies the [16 x 2H] tile by [W_self W_nbr]^T (bfloat16, in the
//      B-operand layout of v_mfma_f32_16x16x16_bf16, staged once per block
//      in LDS, so that registers leave room for four waves per SIMD) on
//      the matrix cores, FP32 accumulation starting from the bias;
//   3. LAST = false: relu, bfloat16, written back through LDS as whole
//      rows; LAST = true: relu in FP32 and the output head
//      delta = (w_out,1 . h + b_out,1) f3 + (w_out,2 . h + b_out,2) f4,
//      the 16 columns of a tile summed by a fixed butterfly.
// MFMA 16x16x16 operand layout: lane l holds A[l%16][4(l/16)..+3] and
// B[4(l/16)..+3][l%16]; the result D[4(l/16)+t][l%16], t = 0..3.
// The grid is persistent: wave w takes tiles w, w + waves, ...
template <int H, bool LAST, int UNR>
__global__ __launch_bounds__(kBlock) void k_layer_bf16(
  const uint16_t* __restrict__ hin, const int32_t* __restrict__ rp,
  const int32_t* __restrict__ ci, const float* __restrict__ w,
  const bf16x4* __restrict__ wfrag, const float* __restrict__ bias,
  uint16_t* __restrict__ hout, const float* __restrict__ wout, float b1,
  float b2, const float* __restrict__ f3, const float* __restrict__ f4,
  float* __restrict__ delta, int n)
{
  constexpr int NT = H / 16, KS = 2 * H / 16;
  constexpr int LPN = H / 8;         // lanes per node row
  constexpr int NPR = kWave / LPN;   // nodes per gather round
  constexpr int TN = kTileNodes;
  constexpr int AS = 2 * H + 8;      // LDS row stride (bf16), padded
  constexpr int WPB = kBlock / kWave;
  __shared__ __attribute__((aligned(16))) uint16_t sA[WPB][TN * AS];
  __shared__ bf16x4 sW[NT * KS * kWave];
  static_assert((TN * AS * 2) % 16 == 0, "16-byte aligned rows per wave");
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = (int)(threadIdx.x / kWave);
  const int col = lane & 15, grp = lane >> 4;
  for (int k = (int)threadIdx.x; k < NT * KS * kWave; k += (int)blockDim.x)
    sW[k] = wfrag[k];
  __syncthreads();
  float bs[NT], wo1[NT], wo2[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    bs[nt] = bias[nt * 16 + col];
    wo1[nt] = LAST ? wout[nt * 16 + col] : 0.f;
    wo2[nt] = LAST ? wout[H + nt * 16 + col] : 0.f;
  }
  uint16_t* A = sA[wv];
  const int sub = lane % LPN;
  const int ntiles = (n + TN - 1) / TN;
  const int nwaves = (int)gridDim.x * WPB;
  for (int t = (int)blockIdx.x * WPB + wv; t < ntiles; t += nwaves) {
    const int i0 = t * TN;
    // 1. gather
#pragma unroll 1
    for (int r = 0; r < TN / NPR; ++r) {
      const int q = r * NPR + lane / LPN;
      const int i = i0 + q;
      uint4 own = make_uint4(0u, 0u, 0u, 0u);
      float a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = 0.f;
      if (i < n) {
        const uint4* h4 = reinterpret_cast<const uint4*>(hin);
        own = h4[(int64_t)i * LPN + sub];
        const int s1 = rp[i + 1];
        // UNR slots at a time: their columns, weights and neighbour rows
        // are independent loads; the sum runs in slot order
        for (int s0 = rp[i]; s0 < s1; s0 += UNR) {
          int cc[UNR];
          float ww[UNR];
#pragma unroll
          for (int u = 0; u < UNR; ++u) {
            const bool ok = s0 + u < s1;
            cc[u] = ok ? ci[s0 + u] : i;
            ww[u] = ok ? w[s0 + u] : 0.f;
          }
          uint4 hv[UNR];
#pragma unroll
          for (int u = 0; u < UNR; ++u) hv[u] = h4[(int64_t)cc[u] * LPN + sub];
#pragma unroll
          for (int u = 0; u < UNR; ++u) {
            const uint32_t hw[4] = {hv[u].x, hv[u].y, hv[u].z, hv[u].w};
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              a[2 * v] = fmaf(ww[u], bf2f(hw[v] & 0xFFFFu), a[2 * v]);
              a[2 * v + 1] = fmaf(ww[u], bf2f(hw[v] >> 16), a[2 * v + 1]);
            }
          }
        }
      }
      uint4* row = reinterpret_cast<uint4*>(A + q * AS);
      row[sub] = own;
      row[LPN + sub] = make_uint4(pack_bf2(a[0], a[1]), pack_bf2(a[2], a[3]),
                                  pack_bf2(a[4], a[5]), pack_bf2(a[6], a[7]));
    }
    wave_lds_sync();
    // 2./3. matrix cores, then the epilogue
#pragma unroll
    for (int mt = 0; mt < TN / 16; ++mt) {
      f32x4 acc[NT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[nt] = (f32x4){bs[nt], bs[nt], bs[nt], bs[nt]};
      const uint16_t* ar = A + (mt * 16 + col) * AS + 4 * grp;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const bf16x4 av = *reinterpret_cast<const bf16x4*>(ar + ks * 16);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
            av, sW[(nt * KS + ks) * kWave + lane], acc[nt], 0, 0, 0);
      }
      if (!LAST) {
        // this M tile's rows of A are no longer read: reuse them for the
        // output rows (first H columns)
        wave_lds_sync();
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
          for (int tt = 0; tt < 4; ++tt)
            A[(mt * 16 + 4 * grp + tt) * AS + nt * 16 + col] =
              f2bf(fmaxf(acc[nt][tt], 0.f));
      } else {
        float p1[4], p2[4];
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
          p1[tt] = 0.f;
          p2[tt] = 0.f;
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) {
            const float hv = fmaxf(acc[nt][tt], 0.f);
            p1[tt] = fmaf(wo1[nt], hv, p1[tt]);
            p2[tt] = fmaf(wo2[nt], hv, p2[tt]);
          }
        }
#pragma unroll
        for (int o = 1; o < 16; o <<= 1)
#pragma unroll
          for (int tt = 0; tt < 4; ++tt) {
            p1[tt] += __shfl_xor(p1[tt], o, 16);
            p2[tt] += __shfl_xor(p2[tt], o, 16);
          }
        if (col == 0) {
#pragma unroll
          for (int tt = 0; tt < 4; ++tt) {
            const int i = i0 + mt * 16 + 4 * grp + tt;
            if (i < n) delta[i] = (p1[tt] + b1) * f3[i] + (p2[tt] + b2) * f4[i];
          }
        }
      }
    }
    wave_lds_sync();
    if (!LAST) {
#pragma unroll 1
      for (int r = 0; r < TN / NPR; ++r) {
        const int q = r * NPR + lane / LPN;
        const int i = i0 + q;
        if (i < n)
          reinterpret_cast<uint4*>(hout + (int64_t)i * H)[sub] =
            reinterpret_cast<const uint4*>(A + q * AS)[sub];
      }
      wave_lds_sync();
    }
  }
}

// x = xb + s_o delta; Dirichlet rows (feature 2 set) get b / a_nn
__global__ void k_apply_guess(const double* __restrict__ xb,
                              const float* __restrict__ delta,
                              const float* __restrict__ dir, double so,
                              const double* __restrict__ b,
                              const double* __restrict__ va,
                              const int32_t* __restrict__ diag,
                              double* __restrict__ x, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (dir[i] != 0.f) {
    x[i] = b[i] / va[diag[i]];
  } else {
    x[i] = delta ? xb[i] + so * (double)delta[i] : xb[i];  // null: a zero head
  }
}

// dd2[i] = coef (delta - target) (f3, f4), the loss gradient at the two
// output coefficients; partial[blk * nv + off] = sum (delta - t)^2
__global__ __launch_bounds__(kBlock) void k_loss(
  const float* __restrict__ delta, const float* __restrict__ target,
  const float* __restrict__ f3, const float* __restrict__ f4, float coef,
  float* __restrict__ dd2, int n, double* __restrict__ partial, int nv,
  int off)
{
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const float e = delta[i] - target[i];
    const float d = coef * e;
    dd2[2 * (int64_t)i] = d * f3[i];
    dd2[2 * (int64_t)i + 1] = d * f4[i];
    acc[0] += (double)e * (double)e;
  }
  __shared__ double tot[1];
  mp_block_store<1>(acc, tot);
  if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * nv + off] = tot[0];
}

// Residual loss, forward: u = delta - t off the Dirichlet rows (dir set),
// v = D^-1 A u; vd = D^-1 v; partial[blk * nv + off] = sum v^2.  One lane
// per row, FP64 sums.
__global__ __launch_bounds__(kBlock) void k_res_fwd(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const double* __restrict__ va, const int32_t* __restrict__ diag,
  const float* __restrict__ delta, const float* __restrict__ target,
  const float* __restrict__ dir, float* __restrict__ vd, int n,
  double* __restrict__ partial, int nv, int off)
{
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    double au = 0.0;
    for (int s = rp[i]; s < rp[i + 1]; ++s) {
      const int j = ci[s];
      const double u = dir[j] != 0.f ? 0.0 : (double)delta[j] - (double)target[j];
      au += va[s] * u;
    }
    const double di = va[diag[i]];
    const double v = au / di;
    acc[0] += v * v;
    vd[i] = (float)(v / di);
  }
  __shared__ double tot[1];
  mp_block_store<1>(acc, tot);
  if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * nv + off] = tot[0];
}

// Residual loss, backward: g = coef A^T vd (0 on Dirichlet rows) and
// dd2[i] = g_i (f3_i, f4_i).  (A^T vd)_i sums a_ji vd_j over the row's
// slots (i, j); a_ji is the edge's other slot (eslot), the diagonal its own.
__global__ __launch_bounds__(kBlock) void k_res_back(
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const int32_t* __restrict__ code, const double* __restrict__ va,
  const int32_t* __restrict__ eslot, const float* __restrict__ vd,
  const float* __restrict__ dir, const float* __restrict__ f3,
  const float* __restrict__ f4, float coef, float* __restrict__ dd2, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double g = 0.0;
  if (dir[i] == 0.f) {
    for (int s = rp[i]; s < rp[i + 1]; ++s) {
      const int c = code[s];
      const int t = c < 0 ? s : eslot[c ^ 1];
      if (t >= 0) g += va[t] * (double)vd[ci[s]];
    }
  }
  const float d = coef * (float)g;
  dd2[2 * (int64_t)i] = d * f3[i];
  dd2[2 * (int64_t)i + 1] = d * f4[i];
}

// gz[i][j] = (hL[i][j] > 0) ? dd2[i] . (wout[0][j], wout[1][j]) : 0
template <int H>
__global__ void k_output_back(const float* __restrict__ hL,
                              const float* __restrict__ dd2,
                              const float* __restrict__ wout,
                              float* __restrict__ gz, int n)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * (H / 4)) return;
  const int i = (int)(t / (H / 4)), j4 = (int)(t % (H / 4));
  const float4 h = reinterpret_cast<const float4*>(hL + (int64_t)i * H)[j4];
  const float d1 = dd2[2 * (int64_t)i], d2 = dd2[2 * (int64_t)i + 1];
  const float* w1 = wout + 4 * j4;
  const float* w2 = wout + H + 4 * j4;
  float4 g;
  g.x = h.x > 0.f ? d1 * w1[0] + d2 * w2[0] : 0.f;
  g.y = h.y > 0.f ? d1 * w1[1] + d2 * w2[1] : 0.f;
  g.z = h.z > 0.f ? d1 * w1[2] + d2 * w2[2] : 0.f;
  g.w = h.w > 0.f ? d1 * w1[3] + d2 * w2[3] : 0.f;
  reinterpret_cast<float4*>(gz + (int64_t)i * H)[j4] = g;
}

// gz_prev[i] = mask(h_prev[i] > 0) * (gs[i] + sum_s w_s q[col_s]); the
// aggregation is the forward's (the edge weights are symmetric), so the
// transpose needs no scatter.  gz_prev may alias gs.
template <int H>
__global__ void k_back_aggregate(const float* gs, const float* __restrict__ q,
                                 const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci,
                                 const float* __restrict__ w,
                                 const float* __restrict__ hp, float* gz_prev,
                                 int n)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * (H / 4)) return;
  const int i = (int)(t / (H / 4)), k4 = (int)(t % (H / 4));
  float4 g = reinterpret_cast<const float4*>(gs + (int64_t)i * H)[k4];
  for (int s = rp[i]; s < rp[i + 1]; ++s) {
    const float ws = w[s];
    if (ws != 0.f) {
      const float4 qc = reinterpret_cast<const float4*>(q + (int64_t)ci[s] * H)[k4];
      g.x = fmaf(ws, qc.x, g.x);
      g.y = fmaf(ws, qc.y, g.y);
      g.z = fmaf(ws, qc.z, g.z);
      g.w = fmaf(ws, qc.w, g.w);
    }
  }
  const float4 h = reinterpret_cast<const float4*>(hp + (int64_t)i * H)[k4];
  g.x = h.x > 0.f ? g.x : 0.f;
  g.y = h.y > 0.f ? g.y : 0.f;
  g.z = h.z > 0.f ? g.z : 0.f;
  g.w = h.w > 0.f ? g.w : 0.f;
  reinterpret_cast<float4*>(gz_prev + (int64_t)i * H)[k4] = g;
}

enum { kModeLayer = 0, kModeInput = 1, kModeOutput = 2 };

template <int H, int MODE>
struct WgradShape {
  static constexpr int JA = MODE == kModeOutput ? 2 : H;
  static constexpr int KB =
    MODE == kModeLayer ? 2 * H + 1 : (MODE == kModeInput ? kF + 1 : H + 1);
  static constexpr int NOUT = JA * KB;
};

// Weight gradients as node sums of outer products A_i B_i^T (the layer
// gradients, the large ones, have their own kernel k_wgrad_layer_mfma):
//   input : A = gz0 [n][H], B = [features | 1]
//   output: A = dd2 [n][2], B = [h_L | 1]
// Block b sums a contiguous node range in tiles of kTileN staged in LDS
// (FP32 products within a tile, FP64 across tiles) and writes
// partial[b * nv + off + o], o = j * KB + k.
template <int H, int MODE>
__global__ __launch_bounds__(kBlock) void k_wgrad(
  const float* __restrict__ A, const float* __restrict__ B,
  const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
  const float* __restrict__ w, int n, double* __restrict__ partial, int nv,
  int off)
{
  using S = WgradShape<H, MODE>;
  constexpr int JA = S::JA, KB = S::KB, NOUT = S::NOUT;
  constexpr int PER = (NOUT + kBlock - 1) / kBlock;
  __shared__ float sA[kTileN][JA];
  __shared__ float sB[kTileN][KB];
  double acc[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) acc[q] = 0.0;
  const int per_block = ((n + gridDim.x - 1) / gridDim.x + kTileN - 1) / kTileN * kTileN;
  const int i0 = min(n, (int)blockIdx.x * per_block);
  const int i1 = min(n, i0 + per_block);
  for (int t0 = i0; t0 < i1; t0 += kTileN) {
    for (int idx = threadIdx.x; idx < kTileN * JA; idx += kBlock) {
      const int t = idx / JA, j = idx % JA, node = t0 + t;
      sA[t][j] = node < i1 ? A[(int64_t)node * JA + j] : 0.f;
    }
    for (int idx = threadIdx.x; idx < kTileN * KB; idx += kBlock) {
      const int t = idx / KB, k = idx % KB, node = t0 + t;
      float v = 0.f;
      if (node < i1) {
        if (MODE == kModeLayer) {
          v = 0.f;  // layer gradients: k_wgrad_layer_mfma
        } else if (MODE == kModeInput) {
          v = k < kF ? B[(int64_t)k * n + node] : 1.f;
        } else {
          v = k < H ? B[(int64_t)node * H + k] : 1.f;
        }
      }
      sB[t][k] = v;
    }
    __syncthreads();
    float tacc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) tacc[q] = 0.f;
    for (int t = 0; t < kTileN; ++t) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int o = threadIdx.x + q * kBlock;
        if (o < NOUT) tacc[q] = fmaf(sA[t][o / KB], sB[t][o % KB], tacc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] += (double)tacc[q];
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int o = threadIdx.x + q * kBlock;
    if (o < NOUT) partial[(int64_t)blockIdx.x * nv + off + o] = acc[q];
  }
}

// ------------------------------------------------- dense parts on MFMA
// The dense products of the FP32 training pass run on the matrix cores as
// v_mfma_f32_16x16x4_f32 (FP32 operands and accumulation; operand layout
// of a 16 x 4 A and 4 x 16 B: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; the 16 x 16 result leaves D[4 (l / 16) + t][l % 16]
// in item t of lane l).  Tiles of 16 nodes per wave; the weights are
// staged once per block in LDS in the B-operand layout (one float per
// lane per fragment, built on the host with the parameters).

// hn[i] = relu(b + W_self h_i + W_nbr m_i): A = the tile's [h | m] rows
// (LDS, row stride 2H + 2 floats, so that the 32 lanes of a half-wave read
// 32 different banks), B = [W_self W_nbr]^T fragments
// wfr[(nt KS + ks) 64 + l] = W[16 nt + l % 16][4 ks + l / 16] over [h, m].
template <int H>
__global__ __launch_bounds__(kBlock) void k_dense_mfma(
  const float* __restrict__ hp, const float* __restrict__ mp,
  const float* __restrict__ wfr, const float* __restrict__ bias,
  float* __restrict__ hn, int n)
{
  constexpr int NT = H / 16, KS = 2 * H / 4, AS = 2 * H + 2;
  constexpr int WPB = kBlock / kWave, C4 = 2 * H / 4;
  __shared__ float sW[NT * KS * kWave];
  __shared__ float sA[WPB][16 * AS];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = (int)(threadIdx.x / kWave);
  const int col = lane & 15, grp = lane >> 4;
  for (int k = (int)threadIdx.x; k < NT * KS * kWave; k += (int)blockDim.x)
    sW[k] = wfr[k];
  __syncthreads();
  float bs[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) bs[nt] = bias[nt * 16 + col];
  float* A = sA[wv];
  const int ntiles = (n + 15) / 16;
  const int nwaves = (int)gridDim.x * WPB;
  // the next tile's rows are loaded into registers while the matrix cores
  // work on the current one
  constexpr int PF = 16 * C4 / kWave;
  float4 pv[PF];
  auto load_tile = [&](int i0) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int q = lane + u * kWave, r = q / C4, c4 = q % C4, i = i0 + r;
      pv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (i < n)
        pv[u] = c4 < H / 4
                  ? reinterpret_cast<const float4*>(hp + (int64_t)i * H)[c4]
                  : reinterpret_cast<const float4*>(mp + (int64_t)i * H)[c4 - H / 4];
    }
  };
  int t = (int)blockIdx.x * WPB + wv;
  if (t < ntiles) load_tile(t * 16);
  for (; t < ntiles; t += nwaves) {
    const int i0 = t * 16;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int q = lane + u * kWave, r = q / C4, c4 = q % C4;
      float* d = A + r * AS + 4 * c4;
      d[0] = pv[u].x;
      d[1] = pv[u].y;
      d[2] = pv[u].z;
      d[3] = pv[u].w;
    }
    wave_lds_sync();
    if (t + nwaves < ntiles) load_tile((t + nwaves) * 16);
    f32x4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = (f32x4){bs[nt], bs[nt], bs[nt], bs[nt]};
    const float* ar = A + col * AS + grp;
#pragma unroll 8
    for (int ks = 0; ks < KS; ++ks) {
      const float a = ar[4 * ks];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
          a, sW[(nt * KS + ks) * kWave + lane], acc[nt], 0, 0, 0);
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int tt = 0; tt < 4; ++tt) {
        const int i = i0 + 4 * grp + tt;
        if (i < n) hn[(int64_t)i * H + nt * 16 + col] = relu(acc[nt][tt]);
      }
    wave_lds_sync();  // the next tile's staging overwrites A
  }
}

// gs[i] = W_self^T gz_i, q[i] = W_nbr^T gz_i: A = the tile's gz rows (LDS,
// stride H + 2), B fragments wbk[(nt KS + ks) 64 + l] =
// W[4 ks + l / 16][16 nt + l % 16] over [W_self | W_nbr] (output columns
// 16 nt + l % 16 < H go to gs, the rest to q).
template <int H>
__global__ __launch_bounds__(kBlock) void k_back_dense_mfma(
  const float* __restrict__ gz, const float* __restrict__ wbk,
  float* __restrict__ gs, float* __restrict__ q, int n)
{
  constexpr int NT = 2 * H / 16, KS = H / 4, AS = H + 2;
  constexpr int WPB = kBlock / kWave, C4 = H / 4;
  __shared__ float sW[NT * KS * kWave];
  __shared__ float sA[WPB][16 * AS];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = (int)(threadIdx.x / kWave);
  const int col = lane & 15, grp = lane >> 4;
  for (int k = (int)threadIdx.x; k < NT * KS * kWave; k += (int)blockDim.x)
    sW[k] = wbk[k];
  __syncthreads();
  float* A = sA[wv];
  const int ntiles = (n + 15) / 16;
  const int nwaves = (int)gridDim.x * WPB;
  // the next tile's rows are loaded into registers during the current one
  constexpr int PF = 16 * C4 / kWave;
  float4 pv[PF];
  auto load_tile = [&](int i0) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int p = lane + u * kWave, r = p / C4, c4 = p % C4, i = i0 + r;
      pv[u] = i < n ? reinterpret_cast<const float4*>(gz + (int64_t)i * H)[c4]
                    : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  int t = (int)blockIdx.x * WPB + wv;
  if (t < ntiles) load_tile(t * 16);
  for (; t < ntiles; t += nwaves) {
    const int i0 = t * 16;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int p = lane + u * kWave, r = p / C4, c4 = p % C4;
      float* d = A + r * AS + 4 * c4;
      d[0] = pv[u].x;
      d[1] = pv[u].y;
      d[2] = pv[u].z;
      d[3] = pv[u].w;
    }
    wave_lds_sync();
    if (t + nwaves < ntiles) load_tile((t + nwaves) * 16);
    f32x4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const float* ar = A + col * AS + grp;
#pragma unroll 4
    for (int ks = 0; ks < KS; ++ks) {
      const float a = ar[4 * ks];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
          a, sW[(nt * KS + ks) * kWave + lane], acc[nt], 0, 0, 0);
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int tt = 0; tt < 4; ++tt) {
        const int i = i0 + 4 * grp + tt;
        const int k = nt * 16 + col;
        if (i < n) (k < H ? gs : q)[(int64_t)i * H + (k % H)] = acc[nt][tt];
      }
    wave_lds_sync();
  }
}

// Layer weight gradients G[j][k] = sum_i gz[i][j] [h_prev | m][i][k] and
// gb[j] = sum_i gz[i][j] on the matrix cores: wave w takes the rows
// j in [16 w, 16 w + 16) (H = 32: waves 0, 1) and all 2H / 16 column tiles;
// a K step is 4 nodes of the block's LDS tile of kTileN nodes (row
// strides padded to 16 mod 32 floats: a half-wave reads 32 banks).  FP32
// within a tile, FP64 across tiles (a fixed order: the result is bitwise
// reproducible); written as partial[b][off + j (2H+1) + k], bias at k = 2H.
template <int H>
__global__ __launch_bounds__(kBlock) void k_wgrad_layer_mfma(
  const float* __restrict__ gz, const float* __restrict__ hp,
  const float* __restrict__ mp, int n, double* __restrict__ partial, int nv,
  int off)
{
  constexpr int NT = 2 * H / 16, SA = H + 16, SB = 2 * H + 16;
  __shared__ float sA[kTileN][SA];
  __shared__ float sB[kTileN][SB];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = (int)(threadIdx.x / kWave);
  const int col = lane & 15, grp = lane >> 4;
  const bool own = wv * 16 < H;
  double accd[NT][4], accb = 0.0;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) accd[nt][tt] = 0.0;
  const int per_block = ((n + gridDim.x - 1) / gridDim.x + kTileN - 1) / kTileN * kTileN;
  const int i0 = min(n, (int)blockIdx.x * per_block);
  const int i1 = min(n, i0 + per_block);
  // the next tile's rows are loaded into registers while the matrix cores
  // work on the current one (PF floats of each array per thread)
  constexpr int PF = kTileN * H / kBlock;
  float ra[PF], rh[PF], rm[PF];
  auto load_tile = [&](int t0) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int idx = (int)threadIdx.x + u * kBlock;
      const int node = t0 + idx / H, k = idx % H;
      const bool in = node < i1;
      ra[u] = in ? gz[(int64_t)node * H + k] : 0.f;
      rh[u] = in ? hp[(int64_t)node * H + k] : 0.f;
      rm[u] = in ? mp[(int64_t)node * H + k] : 0.f;
    }
  };
  if (i0 < i1) load_tile(i0);
  for (int t0 = i0; t0 < i1; t0 += kTileN) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int idx = (int)threadIdx.x + u * kBlock;
      const int t = idx / H, k = idx % H;
      sA[t][k] = ra[u];
      sB[t][k] = rh[u];
      sB[t][H + k] = rm[u];
    }
    __syncthreads();
    if (t0 + kTileN < i1) load_tile(t0 + kTileN);
    if (own) {
      f32x4 acc[NT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[nt] = (f32x4){0.f, 0.f, 0.f, 0.f};
      float tb = 0.f;
#pragma unroll
      for (int ks = 0; ks < kTileN / 4; ++ks) {
        const float a = sA[4 * ks + grp][16 * wv + col];
        tb += a;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
            a, sB[4 * ks + grp][16 * nt + col], acc[nt], 0, 0, 0);
      }
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) accd[nt][tt] += (double)acc[nt][tt];
      accb += (double)tb;
    }
    __syncthreads();
  }
  if (!own) return;
  // the four node groups' bias sums, in a fixed order
  accb += __shfl_xor(accb, 16);
  accb += __shfl_xor(accb, 32);
  double* out = partial + (int64_t)blockIdx.x * nv + off;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      const int j = 16 * wv + 4 * grp + tt;
      out[j * (2 * H + 1) + 16 * nt + col] = accd[nt][tt];
    }
  if (grp == 0) out[(16 * wv + col) * (2 * H + 1) + 2 * H] = accb;
}

}  // namespace

// ---------------------------------------------------------------- model

struct Sample {
  float* feat = nullptr;    // [5][n]
  float* target = nullptr;  // [n]
};

struct Mpnn {
  Ctx* c = nullptr;
  int H = 0, L = 0;
  int64_t n = 0, nnz = 0;
  int64_t np = 0;
  std::vector<double> P;   // FP64 master parameters
  std::vector<double> am, av;  // Adam moments
  int64_t adam_t = 0;
  float* dP = nullptr;     // device parameters, flat layout
  float* dPT = nullptr;    // W_in^T (the input layers read it k-major)
  float* w = nullptr;      // [nnz] slot weights
  std::vector<float*> h;   // L+1 activations [n][H]
  std::vector<float*> msg; // messages m_l [n][H] (L when training, else 1)
  float *gs = nullptr, *gq = nullptr, *gz = nullptr;
  float *feat = nullptr, *target = nullptr, *delta = nullptr, *dd = nullptr;
  double* r = nullptr;     // FP64 residual of x0
  double* partial = nullptr;
  double* red = nullptr;
  int nv = 0;              // values in one gradient partial row
  std::vector<int64_t> gmap;  // reduced value -> parameter index (-1: loss)
  std::vector<Sample> samples;
  double so = 1.0;         // output scale of the last features
  double* xb = nullptr;    // [n] base guess of the last features
  double last_ms = 0.0;
  // the output head (w_out, b_out) is all zero: the prediction is the base
  // guess exactly, and the message-passing layers need not run
  bool head_zero = true;
  int precision = NW_MPNN_BF16;  // arithmetic of the prediction
  int loss_kind = NW_MPNN_LOSS_RESIDUAL;  // training loss
  float* vd = nullptr;        // [n] residual loss: D^-2 A u
  uint16_t* hb[2] = {nullptr, nullptr};  // bfloat16 activations [n][H]
  uint16_t* wfrag = nullptr;  // per layer [H/16][2H/16][64 lanes][4] bf16
  float* wf32 = nullptr;      // per layer [H/16][2H/4][64] FP32 B fragments
  float* wb32 = nullptr;      // per layer [2H/16][H/4][64] (backward)
  int32_t* eslot = nullptr;   // [nedge][2] the slots (L, R), (R, L) of an edge
  uint64_t eslot_gen = 0;     // generation of the graph eslot was built for
  int64_t eslot_nedge = -1;
};

void mpnn_detach(Ctx* c, Mpnn* m);

}  // namespace nw

struct nw_mpnn : nw::Mpnn {};

namespace nw {
namespace {

struct Off {
  int64_t win, bin, l0, ls, wout, bout, n;
  Off(int H, int L)
  {
    win = 0;
    bin = (int64_t)H * kF;
    l0 = bin + H;
    ls = 2 * (int64_t)H * H + H;
    wout = l0 + L * ls;
    bout = wout + 2 * H;
    n = bout + 2;
  }
  int64_t ws(int l) const { return l0 + l * ls; }
  int64_t wn(int l, int H) const { return ws(l) + (int64_t)H * H; }
  int64_t b(int l, int H) const { return ws(l) + 2 * (int64_t)H * H; }
};

int upload_params(Mpnn* m)
{
  const int H = m->H, L = m->L;
  Off o(H, L);
  std::vector<float> f(m->np), t((size_t)H * kF);
  for (int64_t i = 0; i < m->np; ++i) f[i] = (float)m->P[i];
  m->head_zero = true;
  for (int64_t i = o.wout; i < o.bout + 2; ++i) m->head_zero = m->head_zero && m->P[i] == 0.0;
  for (int j = 0; j < H; ++j)
    for (int k = 0; k < kF; ++k) t[(size_t)k * H + j] = f[o.win + j * kF + k];
  // bfloat16 weights of the fused layers in the MFMA B-operand layout:
  // fragment (nt, ks), lane l, item t holds W^T[k][j] with
  // k = 16 ks + 4 (l / 16) + t over [h, m] (W_self then W_nbr) and
  // j = 16 nt + l % 16 the output feature
  const int NT = H / 16, KS = 2 * H / 16;
  std::vector<uint16_t> wb((size_t)L * NT * KS * kWave * 4);
  for (int l = 0; l < L; ++l)
    for (int nt = 0; nt < NT; ++nt)
      for (int ks = 0; ks < KS; ++ks)
        for (int ln = 0; ln < kWave; ++ln)
          for (int tt = 0; tt < 4; ++tt) {
            const int k = 16 * ks + 4 * (ln / 16) + tt, j = 16 * nt + ln % 16;
            const float v = k < H ? f[o.ws(l) + (int64_t)j * H + k]
                                  : f[o.wn(l, H) + (int64_t)j * H + (k - H)];
            wb[((((size_t)l * NT + nt) * KS + ks) * kWave + ln) * 4 + tt] = f2bf(v);
          }
  // FP32 fragments of the training pass (v_mfma_f32_16x16x4_f32 B
  // operands, one float per lane): forward wf[l][nt][ks][lane] =
  // W[16 nt + lane % 16][4 ks + lane / 16] over [W_self | W_nbr] (k < 2H),
  // backward wk[l][nt][ks][lane] = W[4 ks + lane / 16][16 nt + lane % 16]
  // over the output columns [W_self | W_nbr] (n < 2H)
  const size_t per = 2 * (size_t)H * H;
  std::vector<float> wf((size_t)L * per), wk((size_t)L * per);
  auto wcat = [&](int l, int j, int k) {  // [W_self | W_nbr][j][k], k < 2H
    return k < H ? f[o.ws(l) + (int64_t)j * H + k]
                 : f[o.wn(l, H) + (int64_t)j * H + (k - H)];
  };
  for (int l = 0; l < L; ++l) {
    const int FNT = H / 16, FKS = 2 * H / 4, BNT = 2 * H / 16, BKS = H / 4;
    for (int nt = 0; nt < FNT; ++nt)
      for (int ks = 0; ks < FKS; ++ks)
        for (int ln = 0; ln < kWave; ++ln)
          wf[l * per + ((size_t)nt * FKS + ks) * kWave + ln] =
            wcat(l, 16 * nt + ln % 16, 4 * ks + ln / 16);
    for (int nt = 0; nt < BNT; ++nt)
      for (int ks = 0; ks < BKS; ++ks)
        for (int ln = 0; ln < kWave; ++ln)
          wk[l * per + ((size_t)nt * BKS + ks) * kWave + ln] =
            wcat(l, 4 * ks + ln / 16, 16 * nt + ln % 16);
  }
  if (!m->wf32) NW_HIP(hipMalloc(&m->wf32, sizeof(float) * wf.size()));
  if (!m->wb32) NW_HIP(hipMalloc(&m->wb32, sizeof(float) * wk.size()));
  NW_HIP(hipMemcpyAsync(m->wf32, wf.data(), sizeof(float) * wf.size(),
                        hipMemcpyHostToDevice, m->c->stream));
  NW_HIP(hipMemcpyAsync(m->wb32, wk.data(), sizeof(float) * wk.size(),
                        hipMemcpyHostToDevice, m->c->stream));
  if (!m->wfrag) NW_HIP(hipMalloc(&m->wfrag, sizeof(uint16_t) * wb.size()));
  NW_HIP(hipMemcpyAsync(m->dP, f.data(), sizeof(float) * m->np,
                        hipMemcpyHostToDevice, m->c->stream));
  NW_HIP(hipMemcpyAsync(m->dPT, t.data(), sizeof(float) * t.size(),
                        hipMemcpyHostToDevice, m->c->stream));
  NW_HIP(hipMemcpyAsync(m->wfrag, wb.data(), sizeof(uint16_t) * wb.size(),
                        hipMemcpyHostToDevice, m->c->stream));
  NW_HIP(hipStreamSynchronize(m->c->stream));
  return NW_OK;
}

int check_graph(Mpnn* m)
{
  const Graph& g = m->c->g;
  NW_REQUIRE(g.ndof == 1 && g.nnode > 0, NW_E_STATE,
             "the learned guess needs a built 1-dof system");
  NW_REQUIRE(g.nnode == m->n && g.nnzb == m->nnz, NW_E_STATE,
             "the system changed size since the network was created");
  return NW_OK;
}

int ensure_buffers(Mpnn* m, bool training)
{
  const int64_t n = m->n;
  if (!m->w) {
    NW_HIP(hipMalloc(&m->w, sizeof(float) * m->nnz));
    NW_HIP(hipMalloc(&m->feat, sizeof(float) * kF * n));
    NW_HIP(hipMalloc(&m->target, sizeof(float) * n));
    NW_HIP(hipMalloc(&m->delta, sizeof(float) * n));
    NW_HIP(hipMalloc(&m->dd, sizeof(float) * 2 * n));
    NW_HIP(hipMalloc(&m->r, sizeof(double) * n));
    NW_HIP(hipMalloc(&m->xb, sizeof(double) * n));
    NW_HIP(hipMalloc(&m->partial, sizeof(double) * (size_t)kGradBlocks * m->nv));
    NW_HIP(hipMalloc(&m->red, sizeof(double) * m->nv));
    m->h.assign(2, nullptr);
    for (auto& p : m->h) NW_HIP(hipMalloc(&p, sizeof(float) * n * m->H));
    m->msg.assign(1, nullptr);
    NW_HIP(hipMalloc(&m->msg[0], sizeof(float) * n * m->H));
  }
  if (training && (int)m->msg.size() < m->L) {
    while ((int)m->msg.size() < m->L) {
      float* p = nullptr;
      NW_HIP(hipMalloc(&p, sizeof(float) * n * m->H));
      m->msg.push_back(p);
    }
  }
  if (training && (int)m->h.size() < m->L + 1) {
    while ((int)m->h.size() < m->L + 1) {
      float* p = nullptr;
      NW_HIP(hipMalloc(&p, sizeof(float) * n * m->H));
      m->h.push_back(p);
    }
  }
  if (!m->vd) NW_HIP(hipMalloc(&m->vd, sizeof(float) * n));
  if (training && !m->gs) {
    NW_HIP(hipMalloc(&m->gs, sizeof(float) * n * m->H));
    NW_HIP(hipMalloc(&m->gq, sizeof(float) * n * m->H));
    NW_HIP(hipMalloc(&m->gz, sizeof(float) * n * m->H));
  }
  return NW_OK;
}

int slot_weights(Mpnn* m)
{
  Ctx* c = m->c;
  const Graph& g = c->g;
  hipStream_t st = c->stream;
  const int64_t ne = std::max<int64_t>(c->nedge, 0);
  if (m->eslot && (m->eslot_gen != g.gen || m->eslot_nedge != ne)) {
    // the graph was rebuilt (same sizes, check_graph passed): the slots of
    // the edges are rebuilt; the node-sized buffers keep their size
    (void)hipFree(m->eslot);
    m->eslot = nullptr;
  }
  if (!m->eslot) {
    // the two slots of every edge, built once per network; the diagonal
    // (and any slot not reached through an edge) keeps weight 0
    NW_HIP(hipMalloc(&m->eslot, sizeof(int32_t) * std::max<int64_t>(2 * ne, 1)));
    NW_HIP(hipMemsetAsync(m->eslot, 0xFF, sizeof(int32_t) * std::max<int64_t>(2 * ne, 1), st));
    NW_HIP(hipMemsetAsync(m->w, 0, sizeof(float) * m->nnz, st));
    k_edge_slots<<<grid_for(m->nnz, kBlock), kBlock, 0, st>>>(g.code, m->eslot, m->nnz);
    NW_LAUNCH_CHECK();
    m->eslot_gen = g.gen;
    m->eslot_nedge = ne;
  }
  if (ne == 0) return NW_OK;
  k_edge_weights<<<grid_for(ne, kBlock), kBlock, 0, st>>>(
    reinterpret_cast<const int2*>(c->edges), reinterpret_cast<const int2*>(m->eslot),
    g.vals, g.diag, m->w, ne);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

// features of (A, b, history) into feat; target from xs when given; the
// base guess in m->xb
int features(Mpnn* m, const double* const* hist, int nh, const double* xs,
             float* feat, float* target)
{
  Ctx* c = m->c;
  const Graph& g = c->g;
  const int n = (int)m->n;
  hipStream_t st = c->stream;
  NW_REQUIRE(nh >= 1 && nh <= kHist && hist[0], NW_E_STATE,
             "the learned guess needs at least one earlier guess");
  Hist H{};
  H.nh = nh;
  for (int j = 0; j < kHist; ++j) H.h[j] = j < nh ? hist[j] : hist[0];
  k_base_guess<<<grid_for(n, kBlock), kBlock, 0, st>>>(H, m->xb, n);
  k_feat_stats<<<kGradBlocks, kBlock, 0, st>>>(g.row_ptr, g.cols, g.vals,
                                              g.diag, g.rhs, m->xb, H, m->r, n,
                                              m->partial);
  NW_LAUNCH_CHECK();
  k_mp_reduce<<<kStats, kBlock, 0, st>>>(m->partial, kGradBlocks, kStats, m->red);
  NW_LAUNCH_CHECK();
  double s[kStats];
  NW_HIP(hipMemcpyAsync(s, m->red, sizeof(s), hipMemcpyDeviceToHost, st));
  NW_HIP(hipStreamSynchronize(st));
  auto rms = [n](double ss) {
    const double v = std::sqrt(ss / (double)std::max(n, 1));
    return v > 0 ? v : 1.0;
  };
  FeatScale sc;
  sc.isr = 1.0 / rms(s[2]);
  sc.mlog = s[4] / (double)std::max(n, 1);
  sc.isd = 1.0 / rms(s[5]);
  m->so = s[5] > 0 ? rms(s[5]) : rms(s[3]);
  sc.iso = 1.0 / m->so;
  NW_REQUIRE(std::isfinite(s[0] + s[1] + s[2] + s[3] + s[4] + s[5]),
             NW_E_NONFINITE, "non-finite input to the learned guess");
  k_features<<<grid_for(n, kBlock), kBlock, 0, st>>>(g.row_ptr, g.vals, g.cols,
                                                    g.diag, g.rhs, H, m->xb,
                                                    m->r, xs, sc, feat, target, n);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

// blocks of the tile kernels: one wave per 16-node tile up to 8 resident
// blocks per CU's worth (the waves then stride over the tiles)
inline unsigned dense_grid(int n)
{
  const int64_t tiles = (n + 15) / 16;
  const int64_t blocks = (tiles + kBlock / kWave - 1) / (kBlock / kWave);
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks, 256 * 8));
}

template <int H>
int forward_h(Mpnn* m, const float* feat, bool keep)
{
  Ctx* c = m->c;
  const Graph& g = c->g;
  const int n = (int)m->n, L = m->L;
  hipStream_t st = c->stream;
  const Off o(H, L);
  const float* WT = m->dPT;
  k_input_layer<H><<<grid_for(n, kBlock), kBlock, 0, st>>>(
    feat, WT, m->dP + o.bin, m->h[0], n);
  NW_LAUNCH_CHECK();
  int cur = 0;
  const unsigned gq = (unsigned)(((int64_t)n * (H / 4) + kBlock - 1) / kBlock);
  for (int l = 0; l < L; ++l) {
    const int nxt = keep ? l + 1 : 1 - cur;
    float* msg = m->msg[keep ? l : 0];
    k_aggregate<H><<<gq, kBlock, 0, st>>>(m->h[cur], g.row_ptr, g.cols, m->w,
                                          msg, n);
    NW_LAUNCH_CHECK();
    k_dense_mfma<H><<<dense_grid(n), kBlock, 0, st>>>(
      m->h[cur], msg, m->wf32 + (size_t)l * 2 * H * H, m->dP + o.b(l, H),
      m->h[nxt], n);
    NW_LAUNCH_CHECK();
    cur = nxt;
  }
  k_output<H><<<grid_for(n, kBlock), kBlock, 0, st>>>(
    m->h[cur], m->dP + o.wout, (float)m->P[o.bout], (float)m->P[o.bout + 1],
    feat + 3 * (int64_t)n, feat + 4 * (int64_t)n, m->delta, n);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

int forward(Mpnn* m, const float* feat, bool keep)
{
  return m->H == 64 ? forward_h<64>(m, feat, keep) : forward_h<32>(m, feat, keep);
}

// The bfloat16 prediction forward: delta from the features.
template <int H>
int forward_bf16_h(Mpnn* m, const float* feat)
{
  Ctx* c = m->c;
  const Graph& g = c->g;
  const int n = (int)m->n, L = m->L;
  hipStream_t st = c->stream;
  const Off o(H, L);
  for (auto& p : m->hb)
    if (!p) NW_HIP(hipMalloc(&p, sizeof(uint16_t) * (size_t)n * H));
  k_input_bf16<H><<<grid_for(n, kBlock), kBlock, 0, st>>>(
    feat, m->dPT, m->dP + o.bin, m->hb[0], n);
  NW_LAUNCH_CHECK();
  const int64_t ntiles = ((int64_t)n + kTileNodes - 1) / kTileNodes;
  const int wpb = kBlock / kWave;
  const unsigned blocks =
    (unsigned)std::max<int64_t>(1, std::min<int64_t>((ntiles + wpb - 1) / wpb, 2048));
  const size_t fl = (size_t)(H / 16) * (2 * H / 16) * kWave;  // fragments per layer
  const bf16x4* wf = reinterpret_cast<const bf16x4*>(m->wfrag);
  const float b1 = (float)m->P[o.bout], b2 = (float)m->P[o.bout + 1];
  const float* f3 = feat + 3 * (int64_t)n;
  const float* f4 = feat + 4 * (int64_t)n;
  int cur = 0;
  for (int l = 0; l < L; ++l) {
#define NW_LAYER(LASTV)                                                        \
  k_layer_bf16<H, LASTV, 8><<<blocks, kBlock, 0, st>>>(                        \
    m->hb[cur], g.row_ptr, g.cols, m->w, wf + l * fl, m->dP + o.b(l, H),       \
    LASTV ? nullptr : m->hb[1 - cur], m->dP + o.wout, b1, b2, f3, f4,          \
    m->delta, n)
    if (l + 1 < L)
      NW_LAYER(false);
    else
      NW_LAYER(true);
#undef NW_LAYER
    NW_LAUNCH_CHECK();
    cur = 1 - cur;
  }
  return NW_OK;
}

int forward_predict(Mpnn* m, const float* feat)
{
  if (m->precision == NW_MPNN_BF16)
    return m->H == 64 ? forward_bf16_h<64>(m, feat) : forward_bf16_h<32>(m, feat);
  return forward(m, feat, false);
}

// Loss and gradient of one sample accumulated into grad (scaled by 1/batch).
template <int H>
int loss_grad_h(Mpnn* m, const Sample& smp, double scale, double* loss,
                std::vector<double>& grad)
{
  Ctx* c = m->c;
  const Graph& g = c->g;
  const int n = (int)m->n, L = m->L, nv = m->nv;
  hipStream_t st = c->stream;
  const Off o(H, L);
  int rc = forward_h<H>(m, smp.feat, true);
  if (rc) return rc;
  const int off_loss = nv - 1;
  if (m->loss_kind == NW_MPNN_LOSS_RESIDUAL) {
    k_res_fwd<<<kGradBlocks, kBlock, 0, st>>>(
      g.row_ptr, g.cols, g.vals, g.diag, m->delta, smp.target,
      smp.feat + 2 * (int64_t)n, m->vd, n, m->partial, nv, off_loss);
    NW_LAUNCH_CHECK();
    k_res_back<<<grid_for(n, kBlock), kBlock, 0, st>>>(
      g.row_ptr, g.cols, g.code, g.vals, m->eslot, m->vd,
      smp.feat + 2 * (int64_t)n, smp.feat + 3 * (int64_t)n,
      smp.feat + 4 * (int64_t)n, (float)(2.0 * scale / n), m->dd, n);
    NW_LAUNCH_CHECK();
  } else {
    k_loss<<<kGradBlocks, kBlock, 0, st>>>(
      m->delta, smp.target, smp.feat + 3 * (int64_t)n, smp.feat + 4 * (int64_t)n,
      (float)(2.0 * scale / n), m->dd, n, m->partial, nv, off_loss);
    NW_LAUNCH_CHECK();
  }
  // value offsets: output block, then layers L-1..0, then input block
  int off = 0;
  k_wgrad<H, kModeOutput><<<kGradBlocks, kBlock, 0, st>>>(
    m->dd, m->h[L], g.row_ptr, g.cols, m->w, n, m->partial, nv, off);
  NW_LAUNCH_CHECK();
  off += WgradShape<H, kModeOutput>::NOUT;
  const unsigned gq = (unsigned)(((int64_t)n * (H / 4) + kBlock - 1) / kBlock);
  k_output_back<H><<<gq, kBlock, 0, st>>>(m->h[L], m->dd, m->dP + o.wout, m->gz, n);
  NW_LAUNCH_CHECK();
  float* gz = m->gz;
  float* gs = m->gs;
  for (int l = L - 1; l >= 0; --l) {
    k_wgrad_layer_mfma<H><<<kGradBlocks, kBlock, 0, st>>>(
      gz, m->h[l], m->msg[l], n, m->partial, nv, off);
    NW_LAUNCH_CHECK();
    off += WgradShape<H, kModeLayer>::NOUT;
    k_back_dense_mfma<H><<<dense_grid(n), kBlock, 0, st>>>(
      gz, m->wb32 + (size_t)l * 2 * H * H, gs, m->gq, n);
    NW_LAUNCH_CHECK();
    k_back_aggregate<H><<<gq, kBlock, 0, st>>>(gs, m->gq, g.row_ptr, g.cols,
                                               m->w, m->h[l], gs, n);
    NW_LAUNCH_CHECK();
    std::swap(gz, gs);
  }
  k_wgrad<H, kModeInput><<<kGradBlocks, kBlock, 0, st>>>(
    gz, smp.feat, g.row_ptr, g.cols, m->w, n, m->partial, nv, off);
  NW_LAUNCH_CHECK();
  k_mp_reduce<<<nv, kBlock, 0, st>>>(m->partial, kGradBlocks, nv, m->red);
  NW_LAUNCH_CHECK();
  std::vector<double> red(nv);
  NW_HIP(hipMemcpyAsync(red.data(), m->red, sizeof(double) * nv,
                        hipMemcpyDeviceToHost, st));
  NW_HIP(hipStreamSynchronize(st));
  for (int v = 0; v < nv - 1; ++v) grad[m->gmap[v]] += red[v];
  *loss = red[off_loss] / n;
  NW_REQUIRE(std::isfinite(*loss), NW_E_NONFINITE, "non-finite training loss");
  return NW_OK;
}

int loss_grad(Mpnn* m, const Sample& s, double scale, double* loss,
              std::vector<double>& grad)
{
  return m->H == 64 ? loss_grad_h<64>(m, s, scale, loss, grad)
                    : loss_grad_h<32>(m, s, scale, loss, grad);
}

template <int H>
void build_gmap(Mpnn* m)
{
  const int L = m->L;
  const Off o(H, L);
  m->gmap.clear();
  using SO = WgradShape<H, kModeOutput>;
  for (int j = 0; j < SO::JA; ++j)
    for (int k = 0; k < SO::KB; ++k)
      m->gmap.push_back(k < H ? o.wout + (int64_t)j * H + k : o.bout + j);
  using SL = WgradShape<H, kModeLayer>;
  for (int l = L - 1; l >= 0; --l)
    for (int j = 0; j < SL::JA; ++j)
      for (int k = 0; k < SL::KB; ++k)
        m->gmap.push_back(k < H ? o.ws(l) + (int64_t)j * H + k
                          : k < 2 * H ? o.wn(l, H) + (int64_t)j * H + (k - H)
                                      : o.b(l, H) + j);
  using SI = WgradShape<H, kModeInput>;
  for (int j = 0; j < SI::JA; ++j)
    for (int k = 0; k < SI::KB; ++k)
      m->gmap.push_back(k < kF ? o.win + (int64_t)j * kF + k : o.bin + j);
  m->gmap.push_back(-1);  // loss
  m->nv = (int)m->gmap.size();
}

void free_sample(Sample& s)
{
  (void)hipFree(s.feat);
  (void)hipFree(s.target);
  s = Sample();
}

}  // namespace

void mpnn_detach(Ctx* c, Mpnn* m)
{
  if (c && c->mpnn == m) c->mpnn = nullptr;
}

// Prediction for the solver: x = xb + s_o delta from the current system
// and the history of guesses (device vectors, newest first).
int mpnn_predict_device(Mpnn* m, const double* const* hist, int nh, double* x)
{
  const auto t0 = std::chrono::steady_clock::now();
  int rc = check_graph(m);
  if (rc) return rc;
  rc = ensure_buffers(m, false);
  if (rc) return rc;
  rc = slot_weights(m);
  if (rc) return rc;
  rc = features(m, hist, nh, nullptr, m->feat, nullptr);
  if (rc) return rc;
  // a zero output head multiplies the layers' output by zero: the guess is
  // the base guess, without running them
  if (!m->head_zero) {
    rc = forward_predict(m, m->feat);
    if (rc) return rc;
  }
  const Graph& g = m->c->g;
  k_apply_guess<<<grid_for(m->n, kBlock), kBlock, 0, m->c->stream>>>(
    m->xb, m->head_zero ? nullptr : m->delta, m->feat + 2 * m->n, m->so, g.rhs, g.vals,
    g.diag, x, (int)m->n);
  NW_LAUNCH_CHECK();
  NW_HIP(hipStreamSynchronize(m->c->stream));
  m->last_ms = 1e3 * std::chrono::duration<double>(
                       std::chrono::steady_clock::now() - t0).count();
  return NW_OK;
}

double mpnn_last_ms(const Mpnn* m) { return m->last_ms; }

}  // namespace nw

using namespace nw;

extern "C" {

int64_t nw_mpnn_nparams(int hidden, int layers)
{
  if (hidden <= 0 || layers < 0) return -1;
  return Off(hidden, layers).n;
}

int nw_mpnn_create(nw_ctx* c, int hidden, int layers, uint64_t seed,
                   nw_mpnn** out)
{
  NW_REQUIRE(c && out, NW_E_ARG, "null argument");
  NW_REQUIRE(hidden == 32 || hidden == 64, NW_E_ARG,
             "hidden width must be 32 or 64 (got %d)", hidden);
  NW_REQUIRE(layers >= 1 && layers <= 8, NW_E_ARG,
             "message-passing layers must be in [1, 8] (got %d)", layers);
  NW_REQUIRE(c->g.ndof == 1 && c->g.nnode > 0, NW_E_STATE,
             "the learned guess needs a built 1-dof system");
  NW_HIP(hipSetDevice(c->device));
  nw_mpnn* m = new nw_mpnn();
  m->c = c;
  m->H = hidden;
  m->L = layers;
  m->n = c->g.nnode;
  m->nnz = c->g.nnzb;
  Off o(hidden, layers);
  m->np = o.n;
  m->P.assign(m->np, 0.0);
  m->am.assign(m->np, 0.0);
  m->av.assign(m->np, 0.0);
  // He-uniform initialisation from a seeded generator; biases zero
  std::mt19937_64 gen(seed);
  auto fill = [&](int64_t at, int64_t cnt, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    std::uniform_real_distribution<double> u(-a, a);
    for (int64_t i = 0; i < cnt; ++i) m->P[at + i] = u(gen);
  };
  fill(o.win, (int64_t)hidden * kF, kF);
  for (int l = 0; l < layers; ++l) {
    fill(o.ws(l), (int64_t)hidden * hidden, 2 * hidden);
    fill(o.wn(l, hidden), (int64_t)hidden * hidden, 2 * hidden);
  }
  // output head: w_out = 0 and b_out = 0, so the untrained network
  // predicts the base guess, the quadratic extrapolation 3 x0 - 3 xp + xpp;
  // training learns (per node) how much of the third and fourth history
  // differences to add
  if (hidden == 64) build_gmap<64>(m);
  else build_gmap<32>(m);
  int rc = NW_OK;
  if (hipMalloc(&m->dP, sizeof(float) * m->np) != hipSuccess ||
      hipMalloc(&m->dPT, sizeof(float) * (size_t)hidden * kF) != hipSuccess) {
    set_error("device allocation of the network parameters failed");
    rc = NW_E_ALLOC;
  }
  if (!rc) rc = upload_params(m);
  if (rc) {
    nw_mpnn_destroy(m);
    return rc;
  }
  *out = m;
  return NW_OK;
}

void nw_mpnn_destroy(nw_mpnn* m)
{
  if (!m) return;
  (void)hipSetDevice(m->c->device);
  (void)hipStreamSynchronize(m->c->stream);
  mpnn_detach(m->c, m);
  for (auto& s : m->samples) free_sample(s);
  for (auto p : m->h) (void)hipFree(p);
  for (auto p : m->msg) (void)hipFree(p);
  (void)hipFree(m->dP);
  (void)hipFree(m->dPT);
  (void)hipFree(m->w);
  (void)hipFree(m->gs);
  (void)hipFree(m->gq);
  (void)hipFree(m->gz);
  (void)hipFree(m->feat);
  (void)hipFree(m->target);
  (void)hipFree(m->delta);
  (void)hipFree(m->dd);
  (void)hipFree(m->r);
  (void)hipFree(m->xb);
  (void)hipFree(m->partial);
  (void)hipFree(m->red);
  for (auto p : m->hb) (void)hipFree(p);
  (void)hipFree(m->wfrag);
  (void)hipFree(m->wf32);
  (void)hipFree(m->wb32);
  (void)hipFree(m->eslot);
  (void)hipFree(m->vd);
  delete m;
}

int nw_mpnn_set_params(nw_mpnn* m, const double* p, int64_t n)
{
  NW_REQUIRE(m && p, NW_E_ARG, "null argument");
  NW_REQUIRE(n == m->np, NW_E_ARG, "expected %lld parameters, got %lld",
             (long long)m->np, (long long)n);
  for (int64_t i = 0; i < n; ++i)
    NW_REQUIRE(std::isfinite(p[i]), NW_E_NONFINITE, "non-finite parameter %lld",
               (long long)i);
  NW_HIP(hipSetDevice(m->c->device));
  m->P.assign(p, p + n);
  std::fill(m->am.begin(), m->am.end(), 0.0);
  std::fill(m->av.begin(), m->av.end(), 0.0);
  m->adam_t = 0;
  return upload_params(m);
}

int nw_mpnn_get_params(const nw_mpnn* m, double* p, int64_t n)
{
  NW_REQUIRE(m && p, NW_E_ARG, "null argument");
  NW_REQUIRE(n == m->np, NW_E_ARG, "expected %lld parameters, got %lld",
             (long long)m->np, (long long)n);
  std::memcpy(p, m->P.data(), sizeof(double) * n);
  return NW_OK;
}

int nw_mpnn_record(nw_mpnn* m)
{
  NW_REQUIRE(m, NW_E_ARG, "null argument");
  int rc = check_graph(m);
  if (rc) return rc;
  NW_REQUIRE(m->c->nh > 0, NW_E_STATE,
             "no solve has run on this system yet (nothing to record)");
  NW_HIP(hipSetDevice(m->c->device));
  rc = ensure_buffers(m, false);
  if (rc) return rc;
  Sample s;
  if (hipMalloc(&s.feat, sizeof(float) * kF * m->n) != hipSuccess ||
      hipMalloc(&s.target, sizeof(float) * m->n) != hipSuccess) {
    free_sample(s);
    set_error("device allocation of a training sample failed");
    return NW_E_ALLOC;
  }
  rc = features(m, m->c->xh, m->c->nh, m->c->g.sln, s.feat, s.target);
  if (rc) {
    free_sample(s);
    return rc;
  }
  NW_HIP(hipStreamSynchronize(m->c->stream));
  m->samples.push_back(s);
  return NW_OK;
}

int nw_mpnn_samples(const nw_mpnn* m) { return m ? (int)m->samples.size() : -1; }

int nw_mpnn_clear_samples(nw_mpnn* m)
{
  NW_REQUIRE(m, NW_E_ARG, "null argument");
  NW_HIP(hipSetDevice(m->c->device));
  NW_HIP(hipStreamSynchronize(m->c->stream));
  for (auto& s : m->samples) free_sample(s);
  m->samples.clear();
  return NW_OK;
}

int nw_mpnn_loss_grad(nw_mpnn* m, int sample, double* loss, double* grad)
{
  NW_REQUIRE(m && loss, NW_E_ARG, "null argument");
  NW_REQUIRE(sample >= 0 && sample < (int)m->samples.size(), NW_E_ARG,
             "sample %d out of range [0, %d)", sample, (int)m->samples.size());
  int rc = check_graph(m);
  if (rc) return rc;
  NW_HIP(hipSetDevice(m->c->device));
  rc = ensure_buffers(m, true);
  if (rc) return rc;
  rc = slot_weights(m);
  if (rc) return rc;
  std::vector<double> g(m->np, 0.0);
  rc = loss_grad(m, m->samples[sample], 1.0, loss, g);
  if (rc) return rc;
  if (grad) std::memcpy(grad, g.data(), sizeof(double) * m->np);
  return NW_OK;
}

int nw_mpnn_train(nw_mpnn* m, const nw_mpnn_train_cfg* cfg, double* last_loss)
{
  NW_REQUIRE(m && cfg, NW_E_ARG, "null argument");
  NW_REQUIRE(!m->samples.empty(), NW_E_STATE, "no recorded training samples");
  NW_REQUIRE(cfg->epochs >= 0 && cfg->batch >= 1 && cfg->lr > 0 &&
               cfg->beta1 >= 0 && cfg->beta1 < 1 && cfg->beta2 >= 0 &&
               cfg->beta2 < 1 && cfg->eps > 0 && cfg->holdout >= 0,
             NW_E_ARG, "bad training configuration");
  NW_REQUIRE(cfg->holdout < (int)m->samples.size(), NW_E_ARG,
             "holdout %d leaves no training samples (%d recorded)",
             cfg->holdout, (int)m->samples.size());
  int rc = check_graph(m);
  if (rc) return rc;
  NW_HIP(hipSetDevice(m->c->device));
  rc = ensure_buffers(m, true);
  if (rc) return rc;
  rc = slot_weights(m);
  if (rc) return rc;
  // the last `holdout` samples (the most recent steps) validate; the
  // parameters with the lowest validation loss, the initial ones included,
  // are kept
  const int nv = cfg->holdout;
  const int ns = (int)m->samples.size() - nv;
  std::vector<int> order(ns);
  for (int i = 0; i < ns; ++i) order[i] = i;
  std::mt19937_64 gen(cfg->seed);
  std::vector<double> g(m->np), best;
  double best_val = 0.0;
  auto validate = [&](double* out) -> int {
    double s = 0.0;
    std::vector<double> gg(m->np);
    for (int q = 0; q < nv; ++q) {
      double l = 0.0;
      int r = loss_grad(m, m->samples[ns + q], 1.0, &l, gg);
      if (r) return r;
      s += l;
    }
    *out = s / std::max(nv, 1);
    return NW_OK;
  };
  if (nv > 0) {
    rc = validate(&best_val);
    if (rc) return rc;
    best = m->P;
  }
  double epoch_loss = 0.0;
  for (int ep = 0; ep < cfg->epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), gen);
    epoch_loss = 0.0;
    for (int b0 = 0; b0 < ns; b0 += cfg->batch) {
      const int nb = std::min(cfg->batch, ns - b0);
      std::fill(g.begin(), g.end(), 0.0);
      for (int q = 0; q < nb; ++q) {
        double l = 0.0;
        rc = loss_grad(m, m->samples[order[b0 + q]], 1.0 / nb, &l, g);
        if (rc) return rc;
        epoch_loss += l;
      }
      // Adam (Kingma & Ba 2015) on the FP64 master parameters
      ++m->adam_t;
      const double c1 = 1.0 - std::pow(cfg->beta1, (double)m->adam_t);
      const double c2 = 1.0 - std::pow(cfg->beta2, (double)m->adam_t);
      for (int64_t i = 0; i < m->np; ++i) {
        m->am[i] = cfg->beta1 * m->am[i] + (1.0 - cfg->beta1) * g[i];
        m->av[i] = cfg->beta2 * m->av[i] + (1.0 - cfg->beta2) * g[i] * g[i];
        m->P[i] -= cfg->lr * (m->am[i] / c1) / (std::sqrt(m->av[i] / c2) + cfg->eps);
      }
      rc = upload_params(m);
      if (rc) return rc;
    }
    epoch_loss /= ns;
    if (nv > 0) {
      double v = 0.0;
      rc = validate(&v);
      if (rc) return rc;
      if (v < best_val) {
        best_val = v;
        best = m->P;
      }
    }
  }
  if (nv > 0) {
    m->P = best;
    rc = upload_params(m);
    if (rc) return rc;
  }
  if (last_loss) *last_loss = nv > 0 ? best_val : epoch_loss;
  return NW_OK;
}

int nw_mpnn_predict(nw_mpnn* m, double* x_host, double* scale)
{
  NW_REQUIRE(m && x_host, NW_E_ARG, "null argument");
  NW_HIP(hipSetDevice(m->c->device));
  const int64_t n = m->n;
  double* x = nullptr;
  NW_HIP(hipMalloc(&x, sizeof(double) * n));
  // the next step's guess: the solution vector, then the guesses the
  // latest solves started from
  const double* hist[kHist];
  hist[0] = m->c->g.sln;
  int nh = 1;
  for (int j = 0; j < m->c->nh && nh < kHist; ++j) hist[nh++] = m->c->xh[j];
  int rc = mpnn_predict_device(m, hist, nh, x);
  if (!rc) {
    if (hipMemcpy(x_host, x, sizeof(double) * n, hipMemcpyDeviceToHost) !=
        hipSuccess) {
      set_error("copy of the prediction failed");
      rc = NW_E_HIP;
    }
  }
  (void)hipFree(x);
  if (scale) *scale = m->so;
  return rc;
}

int nw_mpnn_set_precision(nw_mpnn* m, int precision)
{
  NW_REQUIRE(m, NW_E_ARG, "null argument");
  NW_REQUIRE(precision == NW_MPNN_FP32 || precision == NW_MPNN_BF16, NW_E_ARG,
             "bad precision %d", precision);
  m->precision = precision;
  return NW_OK;
}

int nw_mpnn_set_loss(nw_mpnn* m, int kind)
{
  NW_REQUIRE(m, NW_E_ARG, "null argument");
  NW_REQUIRE(kind == NW_MPNN_LOSS_SOLUTION || kind == NW_MPNN_LOSS_RESIDUAL,
             NW_E_ARG, "bad loss kind %d", kind);
  m->loss_kind = kind;
  return NW_OK;
}

int nw_linsys_set_learned_guess(nw_ctx* c, nw_mpnn* m)
{
  NW_REQUIRE(c, NW_E_ARG, "null context");
  NW_REQUIRE(!m || m->c == c, NW_E_ARG,
             "the network was created on another context");
  c->mpnn = m;
  return NW_OK;
}

}  // extern "C"
