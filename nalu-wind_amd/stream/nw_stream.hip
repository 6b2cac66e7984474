// nw_stream.hip -- a STREAM-style device copy of known size: the measured
// HBM rate of the GPU the kernels run on, the achievable denominator the
// benchmark reports beside the nominal 8 TB/s.  Not part of the edge path
// and not linked into libnaluwind_amd; built as libnw_stream.so.
//
// Each lane moves U 16-byte elements whose loads are all issued before the
// first store (U loads in flight per lane), lanes of a wave on consecutive
// elements, blocks on consecutive 16 KiB tiles.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kThreads = 256;
constexpr int kU = 4;
typedef float v4f __attribute__((ext_vector_type(4)));  // one 16-byte element

template <bool NT>
__global__ __launch_bounds__(kThreads) void k_copy(const v4f* __restrict__ src,
                                                   v4f* __restrict__ dst, int64_t n)
{
  const int64_t base = (int64_t)blockIdx.x * kThreads * kU + threadIdx.x;
  v4f v[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int64_t i = base + (int64_t)u * kThreads;
    if (i < n) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int64_t i = base + (int64_t)u * kThreads;
    if (i < n) {
      if (NT)
        __builtin_nontemporal_store(v[u], dst + i);
      else
        dst[i] = v[u];
    }
  }
}

}  // namespace

// Copies nbytes (a multiple of 16, both buffers 16-byte aligned device
// memory) from src to dst `reps` times on `stream` after one untimed copy;
// *ms = the average time of one copy by HIP events on that stream.
// nontemporal = 1 streams loads and stores past the caches.  0 on success,
// -1 on a bad argument, -3 on a HIP error.
extern "C" int nw_stream_copy(void* dst, const void* src, int64_t nbytes, int reps,
                              int nontemporal, void* stream, double* ms)
{
  if (!dst || !src || !ms || nbytes <= 0 || nbytes % 16 || reps <= 0 ||
      ((uintptr_t)dst | (uintptr_t)src) % 16)
    return -1;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = nbytes / 16;
  const int64_t per = (int64_t)kThreads * kU;
  const unsigned blocks = (unsigned)((n + per - 1) / per);
  auto launch = [&]() {
    if (nontemporal)
      k_copy<true><<<blocks, kThreads, 0, s>>>((const v4f*)src, (v4f*)dst, n);
    else
      k_copy<false><<<blocks, kThreads, 0, s>>>((const v4f*)src, (v4f*)dst, n);
  };
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return -3;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return -3;
  }
  launch();
  (void)hipEventRecord(a, s);
  for (int r = 0; r < reps; ++r) launch();
  (void)hipEventRecord(b, s);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess) return -3;
  *ms = (double)t / reps;
  return 0;
}
