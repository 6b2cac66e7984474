// nw_sum_into.hip -- CoeffApplier::operator() / LinearSystem::sumInto on
// device arrays: blocks of nent nodes with their dense lhs and rhs, as a
// device algorithm (the boundary-face kernels, FixPressure on the device)
// hands them over, added into the block-CSR system without a host round
// trip and without atomics.
//
// Every scalar contribution becomes a (target, value) pair: target = the
// value index slot * ndof^2 + di * ndof + dj of the (row, column) block
// entry (the slot found by a binary search of the sorted row), or
// nvals + row * ndof + di for the rhs.  A stable radix sort by target keeps
// the contributions of one target in block order; one lane per run of equal
// targets sums the run from zero in that order and adds the sum to the
// target.  The result is bitwise that of the host nw_linsys_sum_into (which
// sums the same runs in the same order) and does not depend on scheduling.
// A column absent from the graph, or a node id out of range, is an error
// and nothing of the call is applied; contributions to skipped rows (hypre
// semantics, while the check is on) are dropped.
//
// Scratch: the pairs, their sorted copies, the sort's temporary storage
// and a status word, in a context buffer that grows (and is allocated) only
// when a call needs more than any call before.
//
// nw_linsys_sum_into_device_noalloc runs the same kernels with that scratch
// in a fixed device array (`scratch`, kNoAllocBytes per GPU, held until the
// process ends) and allocates nothing, for callers that cannot allocate
// device memory.  Its address on each device is taken once, with
// hipGetSymbolAddress, when the first context on that device is created.
#include <cstring>
#include <mutex>

#include <rocprim/device/device_radix_sort.hpp>

#include "nw_internal.h"

namespace nw {
namespace {

constexpr uint32_t kDropped = 0xFFFFFFFFu;  // sorts last, never applied

// The fixed scratch of the allocation-free entry point: 32 MiB per device
// (about 900 thousand scalar contributions per call).  Every call on a
// device shares it; calls that overlap on one device (two contexts on two
// streams) race, so they must be ordered by the caller.
constexpr size_t kNoAllocBytes = (size_t)32 << 20;
static __device__ __align__(256) char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`

// status bits
constexpr int kBadNode = 1, kNoColumn = 2;

__global__ void k_status_clear(int* status) { *status = 0; }

// One lane per (block, i, j) node pair of the input.
__global__ __launch_bounds__(kBlock) void k_sum_emit(
  int nblocks, int nent, int ndof, const int32_t* __restrict__ nodes,
  const double* __restrict__ lhs, const double* __restrict__ rhs,
  const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
  const uint8_t* __restrict__ skip, int64_t nnode, uint32_t nvals,
  uint32_t* __restrict__ key, double* __restrict__ val, int* status)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t npair = (int64_t)nblocks * nent * nent;
  if (t >= npair) return;
  const int j = (int)(t % nent);
  const int i = (int)((t / nent) % nent);
  const int64_t b = t / ((int64_t)nent * nent);
  const int nb = ndof * ndof, nr = nent * ndof;
  const int32_t row = nodes[b * nent + i], col = nodes[b * nent + j];
  const double* L = lhs + b * (int64_t)nr * nr;
  const double* R = rhs + b * (int64_t)nr;
  int64_t slot = -1;
  bool drop = false;
  if (row < 0 || row >= nnode || col < 0 || col >= nnode) {
    atomicOr(status, kBadNode);
    drop = true;
  } else if (skip && skip[row]) {
    drop = true;  // HypreLinSysCoeffApplier::sum_into_1DoF skips the row
  } else {
    int32_t lo = row_ptr[row], hi = row_ptr[row + 1];
    while (lo < hi) {  // first column >= col
      const int32_t mid = (lo + hi) >> 1;
      if (cols[mid] < col)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo < row_ptr[row + 1] && cols[lo] == col) {
      slot = lo;
    } else {
      atomicOr(status, kNoColumn);
      drop = true;
    }
  }
  const int64_t k0 = t * nb;
  for (int di = 0; di < ndof; ++di)
    for (int dj = 0; dj < ndof; ++dj) {
      const int64_t k = k0 + di * ndof + dj;
      key[k] = drop ? kDropped : (uint32_t)(slot * nb + di * ndof + dj);
      val[k] = L[(int64_t)(i * ndof + di) * nr + j * ndof + dj];
    }
  if (j == 0) {
    const int64_t r0 = npair * nb + (b * nent + i) * ndof;
    for (int di = 0; di < ndof; ++di) {
      key[r0 + di] = drop ? kDropped : nvals + (uint32_t)((int64_t)row * ndof + di);
      val[r0 + di] = R[i * ndof + di];
    }
  }
}

// One lane per sorted position; the first of each run of equal targets
// sums the run in order and adds it to the target.
__global__ __launch_bounds__(kBlock) void k_sum_apply(
  const uint32_t* __restrict__ key, const double* __restrict__ val, int64_t n,
  uint32_t nvals, double* __restrict__ vals, double* __restrict__ rhs)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = key[i];
  if (k == kDropped || (i > 0 && key[i - 1] == k)) return;
  double s = 0.0;
  for (int64_t m = i; m < n && key[m] == k; ++m) s += val[m];
  if (k < nvals)
    vals[k] += s;
  else
    rhs[k - nvals] += s;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The scratch of one call: status word, pairs, sorted pairs, sort storage.
struct SumLayout {
  int64_t n = 0;  // scalar contributions
  size_t off_key_in = 0, off_key_out = 0, off_val_in = 0, off_val_out = 0,
         off_tmp = 0, tmp_bytes = 0, total = 0;
};

int sum_layout(int64_t n, hipStream_t st, SumLayout* l)
{
  l->n = n;
  size_t o = 256;  // the status word
  l->off_key_in = o;
  o = align256(o + sizeof(uint32_t) * n);
  l->off_key_out = o;
  o = align256(o + sizeof(uint32_t) * n);
  l->off_val_in = o;
  o = align256(o + sizeof(double) * n);
  l->off_val_out = o;
  o = align256(o + sizeof(double) * n);
  l->off_tmp = o;
  size_t tb = 0;
  NW_HIP(rocprim::radix_sort_pairs(nullptr, tb, (const uint32_t*)nullptr,
                                   (uint32_t*)nullptr, (const double*)nullptr,
                                   (double*)nullptr, (unsigned)n, 0, 32, st));
  l->tmp_bytes = tb;
  l->total = align256(o + tb);
  return NW_OK;
}

// The call, with its scratch at `base` (at least layout.total bytes).
int sum_into_device(Ctx* c, int nblocks, int nent, const int32_t* d_nodes,
                    const double* d_lhs, const double* d_rhs, char* base,
                    const SumLayout& l)
{
  Graph& g = c->g;
  const int nd = g.ndof;
  hipStream_t st = c->stream;
  int* status = reinterpret_cast<int*>(base);
  uint32_t* key_in = reinterpret_cast<uint32_t*>(base + l.off_key_in);
  uint32_t* key_out = reinterpret_cast<uint32_t*>(base + l.off_key_out);
  double* val_in = reinterpret_cast<double*>(base + l.off_val_in);
  double* val_out = reinterpret_cast<double*>(base + l.off_val_out);
  const uint32_t nvals = (uint32_t)(g.nnzb * nd * nd);
  const int64_t npair = (int64_t)nblocks * nent * nent;
  k_status_clear<<<1, 1, 0, st>>>(status);
  k_sum_emit<<<grid_for(npair, kBlock), kBlock, 0, st>>>(
    nblocks, nent, nd, d_nodes, d_lhs, d_rhs, g.row_ptr, g.cols,
    active_skip_mask(g), g.nnode, nvals, key_in, val_in, status);
  NW_LAUNCH_CHECK();
  int h_status = 0;
  NW_HIP(hipMemcpyAsync(&h_status, status, sizeof(int), hipMemcpyDeviceToHost, st));
  NW_HIP(hipStreamSynchronize(st));
  NW_REQUIRE(!(h_status & kBadNode), NW_E_ARG, "sumInto: node id out of range");
  NW_REQUIRE(!(h_status & kNoColumn), NW_E_NOGRAPHCOL,
             "sumInto: a column is absent from its row of the graph");
  size_t tb = l.tmp_bytes;
  NW_HIP(rocprim::radix_sort_pairs(base + l.off_tmp, tb, key_in, key_out, val_in,
                                   val_out, (unsigned)l.n, 0, 32, st));
  k_sum_apply<<<grid_for(l.n, kBlock), kBlock, 0, st>>>(key_out, val_out, l.n, nvals,
                                                      g.vals, g.rhs);
  NW_LAUNCH_CHECK();
  return NW_OK;
}

int check_args(Ctx* c, int nblocks, int nent, const void* nodes, const void* lhs,
               const void* rhs, int64_t* n)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  NW_REQUIRE(nblocks >= 0 && nent > 0 && nent <= 64 && nodes && lhs && rhs,
             NW_E_ARG, "bad sumInto arguments");
  const Graph& g = c->g;
  const int64_t nd = g.ndof;
  NW_REQUIRE(g.nnzb * nd * nd + g.nnode * nd < (int64_t)kDropped, NW_E_ARG,
             "system too large for 32-bit targets");
  *n = (int64_t)nblocks * nent * nd * (nent * nd + 1);
  NW_REQUIRE(*n < (int64_t)1 << 31, NW_E_ARG, "too many contributions in one call");
  return NW_OK;
}

}  // namespace

int noalloc_scratch_init(int device)
{
  NW_REQUIRE(device >= 0 && device < kMaxDevices, NW_E_ARG,
             "device %d beyond the fixed-scratch table", device);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return NW_OK;
  // the caller made `device` current; the array's address differs per device
  void* ptr = nullptr;
  NW_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  NW_REQUIRE(ptr != nullptr, NW_E_HIP, "no device address for the fixed scratch");
  g_scratch_ptr[device] = ptr;
  return NW_OK;
}

void free_sum_scratch(Ctx* c)
{
  (void)hipFree(c->sum_buf);
  c->sum_buf = nullptr;
  c->sum_cap = 0;
}

}  // namespace nw

using namespace nw;

extern "C" {

int nw_linsys_sum_into_device(nw_ctx* c, int nblocks, int nent,
                              const int32_t* d_nodes, const double* d_lhs,
                              const double* d_rhs)
{
  int64_t n = 0;
  int rc = check_args(c, nblocks, nent, d_nodes, d_lhs, d_rhs, &n);
  if (rc) return rc;
  if (n == 0) return NW_OK;
  SumLayout l;
  if ((rc = sum_layout(n, c->stream, &l))) return rc;
  if (l.total > c->sum_cap) {
    // grows, never shrinks: repeated calls of up to this size allocate nothing
    NW_HIP(hipStreamSynchronize(c->stream));
    free_sum_scratch(c);
    const size_t cap = l.total + l.total / 2;
    NW_HIP(hipMalloc(&c->sum_buf, cap));
    c->sum_cap = cap;
  }
  return sum_into_device(c, nblocks, nent, d_nodes, d_lhs, d_rhs,
                         static_cast<char*>(c->sum_buf), l);
}

int nw_linsys_sum_into_device_noalloc(nw_ctx* c, int nblocks, int nent,
                                      const int32_t* d_nodes, const double* d_lhs,
                                      const double* d_rhs)
{
  int64_t n = 0;
  int rc = check_args(c, nblocks, nent, d_nodes, d_lhs, d_rhs, &n);
  if (rc) return rc;
  if (n == 0) return NW_OK;
  SumLayout l;
  if ((rc = sum_layout(n, c->stream, &l))) return rc;
  NW_REQUIRE(l.total <= kNoAllocBytes, NW_E_ARG,
             "sumInto (no allocation): %lld contributions need %zu bytes of "
             "scratch, more than the fixed %zu",
             (long long)n, l.total, kNoAllocBytes);
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    base = g_scratch_ptr[c->device];
  }
  NW_REQUIRE(base != nullptr, NW_E_STATE,
             "fixed scratch of device %d not set up", c->device);
  return sum_into_device(c, nblocks, nent, d_nodes, d_lhs, d_rhs,
                         static_cast<char*>(base), l);
}

}  // extern "C"
