// nw_internal.h -- context layout and helpers shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "nalu_amd.h"

namespace nw {

void set_error(const char* fmt, ...);

#define NW_HIP(call)                                                      \
  do {                                                                    \
    hipError_t nw_e_ = (call);                                            \
    if (nw_e_ != hipSuccess) {                                            \
      ::nw::set_error("%s failed: %s (%s:%d)", #call,                     \
                      hipGetErrorString(nw_e_), __FILE__, __LINE__);      \
      return NW_E_HIP;                                                    \
    }                                                                     \
  } while (0)

#define NW_LAUNCH_CHECK()                                                 \
  do {                                                                    \
    hipError_t nw_e_ = hipGetLastError();                                 \
    if (nw_e_ != hipSuccess) {                                            \
      ::nw::set_error("kernel launch failed: %s (%s:%d)",                 \
                      hipGetErrorString(nw_e_), __FILE__, __LINE__);      \
      return NW_E_HIP;                                                    \
    }                                                                     \
  } while (0)

#define NW_REQUIRE(cond, code, ...)                                       \
  do {                                                                    \
    if (!(cond)) {                                                        \
      ::nw::set_error(__VA_ARGS__);                                       \
      return code;                                                        \
    }                                                                     \
  } while (0)

// Rows handled by one wavefront in the slot-parallel row kernels.
constexpr int kRowsPerWave = 64;
constexpr int kWave = 64;
constexpr int kBlock = 256;

// Orders this wave's LDS accesses: LDS operations of one wave complete in
// issue order, so only the compiler has to be kept from moving them.
__device__ inline void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// XCD-aware block order: the hardware deals workgroups round-robin to the
// 8 XCDs; renumbering them as below hands each XCD one contiguous stretch
// of a locality-ordered chunk list, so neighbouring rows share an L2.
// nblocks_padded must be a multiple of 8.
__device__ inline int xcd_block(int nblocks_padded)
{
  const int per = nblocks_padded / 8;
  return (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
}

// Slot code of a CSR (block) slot: diagonal slot of row n -> -(n+1);
// off-diagonal slot reached through edge e from its node[side] -> 2e+side.
__host__ __device__ inline int32_t diag_code(int32_t n) { return -(n + 1); }

// One chunk of a row schedule: rows [r0, r1) and the row pointer at the
// chunk's start, at its middle row min(r1, r0 + kRowsPerWave / 2) and at its
// end, so that a wave finds its slot range in the one 32-byte record it
// reads first instead of in a second, dependent load of the row pointer.
struct alignas(32) SchedChunk {
  int32_t r0, r1, s0, smid, s1, pad[3];
};

// Sliced ELL copy of a CSR structure: slices of 64 consecutive rows (one
// wave each); entry j of row 64k + l is stored at sp[k] + 64 j + l, the
// slice's width being its longest row; shorter rows (and the rows past n in
// the last slice) are padded with column = the row itself (0 past n) and
// value 0.  slot[p] is the CSR entry stored at position p, -1 for padding.
//
// Offset form (square matrices whose slices repeat a stencil, as a
// structured mesh's do): a slice whose rows' column offsets (column - row)
// form a union of at most kDiaMax values, all inside the matrix for every
// row, stores position j of each row at the slice's j-th offset
// (dia_off[k kDiaMax + j], dia_w[k] of them; value 0 where the row has no
// entry there), so that a product computes its columns as row + offset
// instead of loading them.  ci still holds those columns, for readers
// that do not use the offsets.  dia_w[k] = 0 for slices in the plain form.
//
// Delta form (square matrices numbered for a narrow band, e.g. by reverse
// Cuthill-McKee): a slice not in offset form whose columns all lie within
// 32767 of their rows also keeps cd[p] = column - row as 16-bit integers
// (d16[k] = 1), read instead of ci.
constexpr int kDiaMax = 16;
struct Sell {
  int n = 0, nslices = 0;
  int64_t len = 0;              // positions, padding included
  int64_t* sp = nullptr;        // [nslices + 1]
  int32_t* ci = nullptr;        // [len]
  int32_t* slot = nullptr;      // [len]
  int32_t* dia_w = nullptr;     // [nslices], null when not built
  int32_t* dia_off = nullptr;   // [nslices][kDiaMax]
  int64_t dia_slices = 0;       // slices in offset form
  int8_t* d16 = nullptr;        // [nslices], null when not built
  int16_t* cd = nullptr;        // [len] column - row on slices with d16[k]
  int64_t d16_slices = 0;       // slices in delta form
};

// What the sliced ELL products read of a structure.
struct SellView {
  const int64_t* sp;
  const int32_t* ci;
  const int32_t* dia_w;    // null: no slice in offset form
  const int32_t* dia_off;
  const int8_t* d16;       // null: no slice in delta form
  const int16_t* cd;
};

struct Graph {
  // process-wide unique number of this build of the graph (0: none), so
  // that objects derived from a graph notice a rebuild even when the new
  // arrays reuse the old addresses
  uint64_t gen = 0;
  int ndof = 0;
  int64_t nnode = 0;
  int64_t nnzb = 0;           // block slots (one per node pair)
  int32_t* row_ptr = nullptr; // [nnode+1]
  int32_t* cols = nullptr;    // [nnzb]
  int32_t* code = nullptr;    // [nnzb]
  int32_t* diag = nullptr;    // [nnode] slot of the diagonal block
  double* vals = nullptr;     // [nnzb][ndof][ndof]
  double* rhs = nullptr;      // [nnode][ndof]
  double* sln = nullptr;      // [nnode][ndof]
  std::vector<int32_t> h_row_ptr, h_cols;
  // Row chunks (<= kRowsPerWave consecutive rows each) in the order the
  // row kernels visit them: a locality order (the row order, a sweep or a
  // Morton order of the chunk centroids), so that chunks processed
  // together are neighbours in space.
  SchedChunk* sched = nullptr;
  // [nedge] the edges in the schedule's order: for each chunk, each row,
  // the edges whose node[0] the row is (edge-parallel kernels)
  int32_t* edge_order = nullptr;
  int64_t nchunks = 0;
  int64_t chunk_slots_p99 = 0;       // slots 99 % of the chunks stay within
  int64_t half_chunk_slots_p99 = 0;  // the same for 32-row half chunks
  Sell sell;  // sliced ELL form of the 1-dof graph (built by the first solve)
  bool sched_valid = false;
  // Skipped rows of the hypre backend (nw_linsys_set_skipped_rows): [nnode]
  // flags, the number flagged, and whether contributions to them are
  // currently dropped (on after zeroSystem, off after a row reset, as the
  // reference's checkSkippedRows_ flag)
  uint8_t* skip = nullptr;
  std::vector<uint8_t> h_skip;
  int64_t nskip = 0;
  bool skip_check = true;
};

// The skip flags the scatter must honour now, or null.
inline const uint8_t* active_skip_mask(const Graph& g)
{
  return (g.skip && g.nskip > 0 && g.skip_check) ? g.skip : nullptr;
}

constexpr int kHist = 5;  // guesses kept for the learned initial guess

struct Krylov;     // solver workspace (nw_krylov.hip)
struct AmgDevice;  // AMG hierarchy (nw_amg.hip)
struct Mpnn;       // learned initial guess (nw_mpnn.hip)

// Elements of the mesh and the incidence lists of the device geometry
// (nw_mesh_elements, nw_compute_geometry); all device arrays.
struct Geom {
  int64_t nelem = 0, nhex = 0;
  int32_t* type = nullptr;   // [nelem] 0 Hex8, 1 Tet4, 2 Wed6, 3 Pyr5
  int32_t* conn = nullptr;   // [nelem][8]
  int32_t* e_ptr = nullptr;  // [nedge + 1]
  int32_t* e_inc = nullptr;  // 2 (12 el + s) + flip, element order per edge
  int32_t* n_ptr = nullptr;  // [nnode + 1]
  int32_t* n_inc = nullptr;  // 8 el + a, element order per node
  // per-element scratch of nw_compute_geometry (sub-control-surface area
  // vectors and sub-control volumes), allocated with the elements so that a
  // call (e.g. after mesh motion) allocates nothing
  double* scs = nullptr;
  double* scv = nullptr;
};
void geom_free(Geom* g);

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t nnode = -1;
  int64_t nedge = -1;
  int32_t* edges = nullptr;  // [nedge][2] device
  std::vector<int32_t> h_edges;
  double* field[NW_NUM_FIELDS] = {};
  Graph g;
  Krylov* kry = nullptr;
  AmgDevice* amg = nullptr;
  // [kHist][nnode]: the guesses the latest solves started from, newest
  // first (xh[0]: the last solve's), nh of them valid
  double* xh[kHist] = {};
  int nh = 0;
  Mpnn* mpnn = nullptr;      // network for NW_X0_LEARNED (not owned)
  Geom* geom = nullptr;      // elements for the device geometry
  void* sum_buf = nullptr;   // scratch of nw_linsys_sum_into_device
  size_t sum_cap = 0;        // its size in bytes (grows, never shrinks)
  bool check_finite = false; // nw_ctx_set_check_finite
  int64_t* d_bad = nullptr;  // first non-finite value index found
};

// After an assembly (`who` names it): when the context's finite check is
// on, scans the system's values and rhs and returns NW_E_NONFINITE, naming
// the first non-finite point row, if any entry is NaN or Inf; else rc.
int check_system_finite(Ctx* c, int rc, const char* who);

void free_sum_scratch(Ctx* c);
// the fixed scratch of nw_linsys_sum_into_device_noalloc: takes its device
// address on `device` (current) once per process
int noalloc_scratch_init(int device);

int field_components(int f);
bool field_is_edge(int f);
int64_t field_count(const Ctx* c, int f);  // doubles in the field
// Device pointer of a field, allocated (zero-filled) on first use.
int ensure_field(Ctx* c, int f, double** out);
int require_field(Ctx* c, int f, const double** out);
void free_krylov(Krylov* k);
void amg_free(AmgDevice* a);
// Builds the graph's locality-ordered chunk schedule (g.sched, g.nchunks)
// if it is not current.
int ensure_schedule(Ctx* c);
// Learned guess predicted from the context's system and the history of
// guesses hist[0..nh-1] (newest first, device vectors of nnode doubles);
// wall time in mpnn_last_ms.
int mpnn_predict_device(Mpnn* m, const double* const* hist, int nh, double* x);
// Frees the context's history of guesses.
void free_history(Ctx* c);
double mpnn_last_ms(const Mpnn* m);

inline unsigned grid_for(int64_t n, int block) {
  return (unsigned)((n + block - 1) / block);
}

}  // namespace nw

struct nw_ctx : nw::Ctx {};
