// nw_context.hip -- context lifecycle, mesh and field storage, the edge
// graph (TpetraLinearSystem::buildEdgeToNodeGraph / finalizeLinearSystem
// analogue), and the small row operations of the linear system:
// zeroSystem, sumInto, resetRows, applyDirichletBCs, FixPressureAtNode.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>

#include "nw_internal.h"

namespace nw {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int field_components(int f)
{
  switch (f) {
  case NW_F_COORDINATES:
  case NW_F_VELOCITY:
  case NW_F_DPDX:
  case NW_F_DQDX:
  case NW_F_BC_VALUE:
  case NW_F_UTMP:
  case NW_F_BUOYANCY_SOURCE:
  case NW_E_AREA_VECTOR:
    return 3;
  case NW_F_DUDX:
    return 9;
  default:
    return (f >= 0 && f < NW_NUM_FIELDS) ? 1 : 0;
  }
}

bool field_is_edge(int f)
{
  return f == NW_E_AREA_VECTOR || f == NW_E_MASS_FLOW_RATE ||
         f == NW_E_PECLET_FACTOR || f == NW_E_PECLET_NUMBER;
}

int64_t field_count(const Ctx* c, int f)
{
  return (field_is_edge(f) ? c->nedge : c->nnode) * field_components(f);
}

int ensure_field(Ctx* c, int f, double** out)
{
  NW_REQUIRE(f >= 0 && f < NW_NUM_FIELDS, NW_E_ARG, "unknown field id %d", f);
  NW_REQUIRE(c->nnode >= 0, NW_E_STATE, "mesh not set");
  if (!c->field[f]) {
    const int64_t n = field_count(c, f);
    NW_HIP(hipMalloc(&c->field[f], sizeof(double) * std::max<int64_t>(n, 1)));
    NW_HIP(hipMemsetAsync(c->field[f], 0, sizeof(double) * n, c->stream));
  }
  *out = c->field[f];
  return NW_OK;
}

int require_field(Ctx* c, int f, const double** out)
{
  double* p = nullptr;
  int rc = ensure_field(c, f, &p);
  *out = p;
  return rc;
}

int check_system_finite(Ctx* c, int rc, const char* who);

void free_history(Ctx* c)
{
  for (auto& p : c->xh) {
    (void)hipFree(p);
    p = nullptr;
  }
  c->nh = 0;
}

static void free_graph(Graph& g)
{
  (void)hipFree(g.row_ptr);
  (void)hipFree(g.cols);
  (void)hipFree(g.code);
  (void)hipFree(g.diag);
  (void)hipFree(g.vals);
  (void)hipFree(g.rhs);
  (void)hipFree(g.sln);
  (void)hipFree(g.sched);
  (void)hipFree(g.edge_order);
  (void)hipFree(g.sell.sp);
  (void)hipFree(g.sell.ci);
  (void)hipFree(g.sell.slot);
  (void)hipFree(g.sell.dia_w);
  (void)hipFree(g.sell.dia_off);
  (void)hipFree(g.sell.d16);
  (void)hipFree(g.sell.cd);
  (void)hipFree(g.skip);
  g = Graph();
}

// ------------------------------------------------------------- kernels

__global__ void k_fill(double* __restrict__ p, double v, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Smallest index i < n with a non-finite p[i], as min into *first (which
// starts at n): one atomicMin per wave that found one.
__global__ void k_first_nonfinite(const double* __restrict__ p, int64_t n,
                                  unsigned long long* first)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool bad = i < n && !__builtin_isfinite(p[i]);
  if (bad) atomicMin(first, (unsigned long long)i);
}

// Point-row reset for the nodes listed: every entry of point row
// (node, d), d in [begin, end), becomes diag_value on the diagonal and 0
// elsewhere; rhs becomes rhs_value (or bc - sol when bc != nullptr).
__global__ void k_reset_rows(const int32_t* __restrict__ nodes, int64_t nn,
                             int begin, int end, int ndof,
                             const int32_t* __restrict__ row_ptr,
                             const int32_t* __restrict__ cols,
                             double* __restrict__ vals,
                             double* __restrict__ rhs, double diag_value,
                             double rhs_value, const double* __restrict__ bc,
                             const double* __restrict__ sol, int fnc)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nd = end - begin;
  if (t >= nn * nd) return;
  const int32_t n = nodes[t / nd];
  const int d = begin + (int)(t % nd);
  const int nb = ndof * ndof;
  for (int32_t s = row_ptr[n]; s < row_ptr[n + 1]; ++s) {
    const bool self = cols[s] == n;
    for (int c = 0; c < ndof; ++c)
      vals[(int64_t)s * nb + d * ndof + c] = (self && c == d) ? diag_value : 0.0;
  }
  rhs[(int64_t)n * ndof + d] =
    bc ? (bc[(int64_t)n * fnc + d] - sol[(int64_t)n * fnc + d]) : rhs_value;
}

// Segments of sorted (target, value) contributions, each summed in input
// order and added to its target: deterministic sumInto of many blocks.
__global__ void k_add_segments(const int64_t* __restrict__ seg,
                               const int64_t* __restrict__ target,
                               const double* __restrict__ value, int64_t nseg,
                               double* __restrict__ vals,
                               double* __restrict__ rhs)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  double s = 0.0;
  for (int64_t k = seg[i]; k < seg[i + 1]; ++k) s += value[k];
  const int64_t t = target[seg[i]];
  if (t >= 0)
    vals[t] += s;
  else
    rhs[-t - 1] += s;
}

__global__ void k_set_u64x2(unsigned long long* p, unsigned long long a,
                            unsigned long long b)
{
  p[0] = a;
  p[1] = b;
}

int check_system_finite(Ctx* c, int rc, const char* who)
{
  if (rc || !c->check_finite || !c->g.vals) return rc;
  const Graph& g = c->g;
  const int64_t nv = g.nnzb * g.ndof * g.ndof, nr = g.nnode * g.ndof;
  auto* first = reinterpret_cast<unsigned long long*>(c->d_bad);
  // first[0]: first non-finite value (nv if none), first[1]: rhs (nr)
  k_set_u64x2<<<1, 1, 0, c->stream>>>(first, (unsigned long long)nv,
                                      (unsigned long long)nr);
  if (nv > 0)
    k_first_nonfinite<<<grid_for(nv, 256), 256, 0, c->stream>>>(g.vals, nv, first);
  if (nr > 0)
    k_first_nonfinite<<<grid_for(nr, 256), 256, 0, c->stream>>>(g.rhs, nr, first + 1);
  NW_LAUNCH_CHECK();
  unsigned long long h[2] = {0, 0};
  NW_HIP(hipMemcpyAsync(h, first, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  if (h[0] < (unsigned long long)nv) {
    // the slot's row: the last row whose first slot is <= slot
    const int64_t slot = (int64_t)h[0] / (g.ndof * g.ndof);
    std::vector<int32_t> rp(g.nnode + 1);
    NW_HIP(hipMemcpy(rp.data(), g.row_ptr, sizeof(int32_t) * rp.size(),
                     hipMemcpyDeviceToHost));
    const int64_t row = std::upper_bound(rp.begin(), rp.end(), (int32_t)slot) -
                        rp.begin() - 1;
    set_error("%s: non-finite matrix entry in the row of node %lld (an input "
              "field holds NaN or Inf, or an edge has a . dx = 0)", who,
              (long long)row);
    return NW_E_NONFINITE;
  }
  if (h[1] < (unsigned long long)nr) {
    set_error("%s: non-finite rhs entry at node %lld (an input field holds NaN "
              "or Inf, or an edge has a . dx = 0)", who, (long long)(h[1] / g.ndof));
    return NW_E_NONFINITE;
  }
  return NW_OK;
}

}  // namespace nw

using namespace nw;

extern "C" {

const char* nw_last_error(void) { return g_last_error.c_str(); }

int nw_field_components(int field) { return field_components(field); }

int nw_ctx_create(int device, nw_ctx** out)
{
  NW_REQUIRE(out, NW_E_ARG, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  NW_HIP(hipGetDeviceCount(&ndev));
  NW_REQUIRE(device >= 0 && device < ndev, NW_E_ARG,
             "device %d not present (%d devices)", device, ndev);
  NW_HIP(hipSetDevice(device));
  int rc = noalloc_scratch_init(device);
  if (rc) return rc;
  nw_ctx* c = new nw_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) !=
      hipSuccess) {
    delete c;
    set_error("hipStreamCreate failed");
    return NW_E_HIP;
  }
  *out = c;
  return NW_OK;
}

int nw_ctx_destroy(nw_ctx* c)
{
  if (!c) return NW_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_krylov(c->kry);
  amg_free(c->amg);
  free_history(c);
  free_graph(c->g);
  geom_free(c->geom);
  free_sum_scratch(c);
  (void)hipFree(c->d_bad);
  for (auto& f : c->field) (void)hipFree(f);
  (void)hipFree(c->edges);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return NW_OK;
}

int nw_ctx_set_check_finite(nw_ctx* c, int on)
{
  NW_REQUIRE(c, NW_E_ARG, "null context");
  if (on && !c->d_bad) {
    NW_HIP(hipSetDevice(c->device));
    NW_HIP(hipMalloc(&c->d_bad, 2 * sizeof(int64_t)));
  }
  c->check_finite = on != 0;
  return NW_OK;
}

int nw_ctx_stream(nw_ctx* c, void** stream)
{
  NW_REQUIRE(c && stream, NW_E_ARG, "null argument");
  *stream = (void*)c->stream;
  return NW_OK;
}

int nw_ctx_synchronize(nw_ctx* c)
{
  NW_REQUIRE(c, NW_E_ARG, "null context");
  NW_HIP(hipStreamSynchronize(c->stream));
  return NW_OK;
}

int nw_mesh_set(nw_ctx* c, int64_t nnode, int64_t nedge,
                const int32_t* edge_nodes)
{
  NW_REQUIRE(c && nnode > 0 && nedge >= 0 && (edge_nodes || nedge == 0),
             NW_E_ARG, "bad mesh arguments");
  NW_REQUIRE(nnode < INT32_MAX && 2 * nedge < INT32_MAX, NW_E_ARG,
             "mesh too large for 32-bit local ids");
  for (int64_t e = 0; e < nedge; ++e) {
    const int32_t a = edge_nodes[2 * e], b = edge_nodes[2 * e + 1];
    NW_REQUIRE(a >= 0 && a < nnode && b >= 0 && b < nnode && a != b,
               NW_E_ARG, "edge %lld has bad nodes (%d, %d)", (long long)e, a,
               b);
  }
  NW_HIP(hipSetDevice(c->device));
  NW_HIP(hipStreamSynchronize(c->stream));
  free_krylov(c->kry);
  c->kry = nullptr;
  amg_free(c->amg);
  c->amg = nullptr;
  free_history(c);
  free_graph(c->g);
  geom_free(c->geom);
  c->geom = nullptr;
  for (auto& f : c->field) {
    (void)hipFree(f);
    f = nullptr;
  }
  (void)hipFree(c->edges);
  c->edges = nullptr;
  c->nnode = nnode;
  c->nedge = nedge;
  c->h_edges.assign(edge_nodes, edge_nodes + 2 * nedge);
  NW_HIP(hipMalloc(&c->edges, sizeof(int32_t) * std::max<int64_t>(2 * nedge, 1)));
  NW_HIP(hipMemcpyAsync(c->edges, edge_nodes, sizeof(int32_t) * 2 * nedge,
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  return NW_OK;
}

int nw_field_device_ptr(nw_ctx* c, int field, double** dptr)
{
  NW_REQUIRE(c && dptr, NW_E_ARG, "null argument");
  return ensure_field(c, field, dptr);
}

int nw_mesh_coordinates_changed(nw_ctx* c)
{
  NW_REQUIRE(c, NW_E_ARG, "null context");
  c->g.sched_valid = false;
  return NW_OK;
}

int nw_field_upload(nw_ctx* c, int field, const double* host, int64_t count)
{
  NW_REQUIRE(c && host, NW_E_ARG, "null argument");
  double* d = nullptr;
  int rc = ensure_field(c, field, &d);
  if (rc) return rc;
  NW_REQUIRE(count == field_count(c, field), NW_E_ARG,
             "field %d expects %lld doubles, got %lld", field,
             (long long)field_count(c, field), (long long)count);
  // the device keeps the host's node-major layout ([n][ncomp])
  NW_HIP(hipMemcpyAsync(d, host, sizeof(double) * count, hipMemcpyHostToDevice,
                        c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  if (field == NW_F_COORDINATES) c->g.sched_valid = false;
  return NW_OK;
}

int nw_field_download(nw_ctx* c, int field, double* host, int64_t count)
{
  NW_REQUIRE(c && host, NW_E_ARG, "null argument");
  double* d = nullptr;
  int rc = ensure_field(c, field, &d);
  if (rc) return rc;
  NW_REQUIRE(count == field_count(c, field), NW_E_ARG,
             "field %d holds %lld doubles, asked %lld", field,
             (long long)field_count(c, field), (long long)count);
  NW_HIP(hipMemcpyAsync(host, d, sizeof(double) * count, hipMemcpyDeviceToHost,
                        c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  return NW_OK;
}

int nw_field_fill(nw_ctx* c, int field, double value)
{
  NW_REQUIRE(c, NW_E_ARG, "null context");
  double* d = nullptr;
  int rc = ensure_field(c, field, &d);
  if (rc) return rc;
  const int64_t n = field_count(c, field);
  if (n > 0) {
    k_fill<<<grid_for(n, 256), 256, 0, c->stream>>>(d, value, n);
    NW_LAUNCH_CHECK();
  }
  return NW_OK;
}

int nw_linsys_build_edge_graph(nw_ctx* c, int ndof)
{
  NW_REQUIRE(c, NW_E_ARG, "null context");
  NW_REQUIRE(ndof == 1 || ndof == 3, NW_E_ARG, "ndof must be 1 or 3");
  NW_REQUIRE(c->nnode > 0, NW_E_STATE, "mesh not set");
  NW_HIP(hipSetDevice(c->device));
  free_krylov(c->kry);
  c->kry = nullptr;
  amg_free(c->amg);
  c->amg = nullptr;
  free_history(c);
  free_graph(c->g);
  Graph& g = c->g;
  const int64_t N = c->nnode, E = c->nedge;
  const int32_t* ed = c->h_edges.data();
  // rows: self + edge neighbours, sorted by column (the sorted Tpetra graph)
  std::vector<int32_t> deg(N + 1, 1);
  for (int64_t e = 0; e < E; ++e) {
    deg[ed[2 * e]]++;
    deg[ed[2 * e + 1]]++;
  }
  g.h_row_ptr.assign(N + 1, 0);
  for (int64_t n = 0; n < N; ++n) g.h_row_ptr[n + 1] = g.h_row_ptr[n] + deg[n];
  const int64_t nnzb = g.h_row_ptr[N];
  NW_REQUIRE(nnzb < INT32_MAX, NW_E_ARG, "graph too large");
  std::vector<int64_t> key(nnzb);  // (col << 32) | code-bias
  std::vector<int32_t> fill(N);
  for (int64_t n = 0; n < N; ++n) {
    fill[n] = g.h_row_ptr[n] + 1;
    key[g.h_row_ptr[n]] =
      ((int64_t)n << 32) | (uint32_t)diag_code((int32_t)n);
  }
  for (int64_t e = 0; e < E; ++e) {
    const int32_t a = ed[2 * e], b = ed[2 * e + 1];
    key[fill[a]++] = ((int64_t)b << 32) | (uint32_t)(2 * e + 0);
    key[fill[b]++] = ((int64_t)a << 32) | (uint32_t)(2 * e + 1);
  }
  for (int64_t n = 0; n < N; ++n)
    std::sort(key.begin() + g.h_row_ptr[n], key.begin() + g.h_row_ptr[n + 1]);
  g.h_cols.resize(nnzb);
  std::vector<int32_t> code(nnzb), diag(N);
  for (int64_t n = 0; n < N; ++n)
    for (int64_t s = g.h_row_ptr[n]; s < g.h_row_ptr[n + 1]; ++s) {
      g.h_cols[s] = (int32_t)(key[s] >> 32);
      code[s] = (int32_t)(uint32_t)(key[s] & 0xffffffffu);
      if (s > g.h_row_ptr[n]) {
        NW_REQUIRE(g.h_cols[s] != g.h_cols[s - 1], NW_E_ARG,
                   "duplicate edge between nodes %lld and %d", (long long)n,
                   g.h_cols[s]);
      }
      if (g.h_cols[s] == n) diag[n] = (int32_t)s;
    }
  key.clear();
  key.shrink_to_fit();
  static std::atomic<uint64_t> generation{0};
  g.gen = ++generation;
  g.ndof = ndof;
  g.nnode = N;
  g.nnzb = nnzb;
  const int64_t nb = (int64_t)ndof * ndof;
  NW_HIP(hipMalloc(&g.row_ptr, sizeof(int32_t) * (N + 1)));
  NW_HIP(hipMalloc(&g.cols, sizeof(int32_t) * nnzb));
  NW_HIP(hipMalloc(&g.code, sizeof(int32_t) * nnzb));
  NW_HIP(hipMalloc(&g.diag, sizeof(int32_t) * N));
  NW_HIP(hipMalloc(&g.vals, sizeof(double) * nnzb * nb));
  NW_HIP(hipMalloc(&g.rhs, sizeof(double) * N * ndof));
  NW_HIP(hipMalloc(&g.sln, sizeof(double) * N * ndof));
  NW_HIP(hipMemcpyAsync(g.row_ptr, g.h_row_ptr.data(), sizeof(int32_t) * (N + 1),
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipMemcpyAsync(g.cols, g.h_cols.data(), sizeof(int32_t) * nnzb,
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipMemcpyAsync(g.code, code.data(), sizeof(int32_t) * nnzb,
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipMemcpyAsync(g.diag, diag.data(), sizeof(int32_t) * N,
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipMemsetAsync(g.vals, 0, sizeof(double) * nnzb * nb, c->stream));
  NW_HIP(hipMemsetAsync(g.rhs, 0, sizeof(double) * N * ndof, c->stream));
  NW_HIP(hipMemsetAsync(g.sln, 0, sizeof(double) * N * ndof, c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  return NW_OK;
}

int nw_linsys_zero(nw_ctx* c)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  Graph& g = c->g;
  const int64_t nv = g.nnzb * g.ndof * g.ndof, nr = g.nnode * g.ndof;
  k_fill<<<grid_for(nv, 256), 256, 0, c->stream>>>(g.vals, 0.0, nv);
  k_fill<<<grid_for(nr, 256), 256, 0, c->stream>>>(g.rhs, 0.0, nr);
  k_fill<<<grid_for(nr, 256), 256, 0, c->stream>>>(g.sln, 0.0, nr);
  NW_LAUNCH_CHECK();
  g.skip_check = true;  // HypreLinearSystem::resetCoeffApplierData
  return NW_OK;
}

int nw_linsys_set_skipped_rows(nw_ctx* c, int64_t nnodes, const int32_t* nodes)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  NW_REQUIRE(nnodes >= 0 && (nodes || nnodes == 0), NW_E_ARG, "bad node list");
  Graph& g = c->g;
  std::vector<uint8_t> m(g.nnode, 0);
  for (int64_t i = 0; i < nnodes; ++i) {
    NW_REQUIRE(nodes[i] >= 0 && nodes[i] < g.nnode, NW_E_ARG, "bad node id %d",
               nodes[i]);
    m[nodes[i]] = 1;
  }
  if (!g.skip) NW_HIP(hipMalloc(&g.skip, std::max<int64_t>(g.nnode, 1)));
  NW_HIP(hipMemcpyAsync(g.skip, m.data(), g.nnode, hipMemcpyHostToDevice,
                        c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  g.nskip = 0;
  for (uint8_t v : m) g.nskip += v;
  g.h_skip.swap(m);
  g.skip_check = true;
  return NW_OK;
}

int nw_linsys_load_complete(nw_ctx* c)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  return NW_OK;
}

int nw_linsys_sizes(nw_ctx* c, int64_t* nrows, int64_t* nnz)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  if (nrows) *nrows = c->g.nnode * c->g.ndof;
  if (nnz) *nnz = c->g.nnzb * c->g.ndof * c->g.ndof;
  return NW_OK;
}

int nw_linsys_device_arrays(nw_ctx* c, int32_t** row_ptr, int32_t** cols,
                            double** vals, double** rhs)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  if (row_ptr) *row_ptr = c->g.row_ptr;
  if (cols) *cols = c->g.cols;
  if (vals) *vals = c->g.vals;
  if (rhs) *rhs = c->g.rhs;
  return NW_OK;
}

int nw_linsys_download(nw_ctx* c, int64_t* row_ptr, int32_t* cols,
                       double* vals, double* rhs)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  Graph& g = c->g;
  const int nd = g.ndof, nb = nd * nd;
  NW_HIP(hipStreamSynchronize(c->stream));
  std::vector<double> bv;
  if (vals) {
    bv.resize(g.nnzb * nb);
    NW_HIP(hipMemcpy(bv.data(), g.vals, sizeof(double) * bv.size(),
                     hipMemcpyDeviceToHost));
  }
  if (rhs)
    NW_HIP(hipMemcpy(rhs, g.rhs, sizeof(double) * g.nnode * nd,
                     hipMemcpyDeviceToHost));
  int64_t k = 0;
  if (row_ptr) row_ptr[0] = 0;
  for (int64_t n = 0; n < g.nnode; ++n) {
    for (int d = 0; d < nd; ++d) {
      for (int32_t s = g.h_row_ptr[n]; s < g.h_row_ptr[n + 1]; ++s)
        for (int cc = 0; cc < nd; ++cc) {
          if (cols) cols[k] = g.h_cols[s] * nd + cc;
          if (vals) vals[k] = bv[(int64_t)s * nb + d * nd + cc];
          ++k;
        }
      if (row_ptr) row_ptr[n * nd + d + 1] = k;
    }
  }
  return NW_OK;
}

// Slot of (row, col) in the block graph, -1 if absent.
static int64_t find_slot(const Graph& g, int32_t row, int32_t col)
{
  const int32_t* b = g.h_cols.data() + g.h_row_ptr[row];
  const int32_t* e = g.h_cols.data() + g.h_row_ptr[row + 1];
  const int32_t* it = std::lower_bound(b, e, col);
  return (it != e && *it == col) ? (int64_t)(it - g.h_cols.data()) : -1;
}

int nw_linsys_sum_into(nw_ctx* c, int nblocks, int nent, const int32_t* nodes,
                       const double* lhs, const double* rhs)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  NW_REQUIRE(nblocks >= 0 && nent > 0 && nodes && lhs && rhs, NW_E_ARG,
             "bad sumInto arguments");
  Graph& g = c->g;
  const int nd = g.ndof, nb = nd * nd, nr = nent * nd;
  // contributions as (target, value): target >= 0 is a value index,
  // target < 0 is rhs index -(t+1)
  std::vector<std::pair<int64_t, double>> contrib;
  contrib.reserve((size_t)nblocks * (nr * nr + nr));
  // rows of skipped nodes take nothing while the skip check is on
  // (HypreLinSysCoeffApplier::sum_into_1DoF, src/HypreLinearSystem.C:2190-2195)
  const bool skipping = active_skip_mask(g) != nullptr;
  for (int b = 0; b < nblocks; ++b) {
    const int32_t* nn = nodes + (int64_t)b * nent;
    const double* L = lhs + (int64_t)b * nr * nr;
    const double* R = rhs + (int64_t)b * nr;
    for (int i = 0; i < nent; ++i)
      NW_REQUIRE(nn[i] >= 0 && nn[i] < g.nnode, NW_E_ARG, "bad node id %d",
                 nn[i]);
    for (int i = 0; i < nent; ++i) {
      if (skipping && g.h_skip[nn[i]]) continue;
      for (int di = 0; di < nd; ++di) {
        const int row = i * nd + di;
        contrib.push_back({-((int64_t)nn[i] * nd + di) - 1, R[row]});
        for (int j = 0; j < nent; ++j) {
          const int64_t s = find_slot(g, nn[i], nn[j]);
          NW_REQUIRE(s >= 0, NW_E_NOGRAPHCOL,
                     "sumInto column %d absent from row %d of the graph",
                     nn[j], nn[i]);
          for (int dj = 0; dj < nd; ++dj)
            contrib.push_back(
              {s * nb + di * nd + dj, L[(int64_t)row * nr + j * nd + dj]});
        }
      }
    }
  }
  if (contrib.empty()) return NW_OK;
  std::stable_sort(contrib.begin(), contrib.end(),
                   [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<int64_t> seg, target(contrib.size());
  std::vector<double> value(contrib.size());
  for (size_t k = 0; k < contrib.size(); ++k) {
    target[k] = contrib[k].first;
    value[k] = contrib[k].second;
    if (k == 0 || contrib[k].first != contrib[k - 1].first) seg.push_back(k);
  }
  seg.push_back((int64_t)contrib.size());
  const int64_t nseg = (int64_t)seg.size() - 1;
  int64_t *d_seg = nullptr, *d_t = nullptr;
  double* d_v = nullptr;
  NW_HIP(hipMalloc(&d_seg, sizeof(int64_t) * seg.size()));
  NW_HIP(hipMalloc(&d_t, sizeof(int64_t) * target.size()));
  NW_HIP(hipMalloc(&d_v, sizeof(double) * value.size()));
  NW_HIP(hipMemcpyAsync(d_seg, seg.data(), sizeof(int64_t) * seg.size(),
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipMemcpyAsync(d_t, target.data(), sizeof(int64_t) * target.size(),
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipMemcpyAsync(d_v, value.data(), sizeof(double) * value.size(),
                        hipMemcpyHostToDevice, c->stream));
  k_add_segments<<<grid_for(nseg, 256), 256, 0, c->stream>>>(
    d_seg, d_t, d_v, nseg, g.vals, g.rhs);
  hipError_t le = hipGetLastError();
  NW_HIP(hipStreamSynchronize(c->stream));
  (void)hipFree(d_seg);
  (void)hipFree(d_t);
  (void)hipFree(d_v);
  NW_REQUIRE(le == hipSuccess, NW_E_HIP, "sumInto kernel: %s",
             hipGetErrorString(le));
  return NW_OK;
}

static int reset_rows_impl(nw_ctx* c, int64_t nnodes, const int32_t* nodes,
                           int begin, int end, double diag_value,
                           double rhs_value, const double* bc,
                           const double* sol, int fnc)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  NW_REQUIRE(nnodes >= 0 && (nodes || nnodes == 0), NW_E_ARG, "bad node list");
  NW_REQUIRE(begin >= 0 && end <= c->g.ndof && begin <= end, NW_E_ARG,
             "dof range [%d,%d) outside [0,%d)", begin, end, c->g.ndof);
  if (nnodes == 0 || begin == end) return NW_OK;
  for (int64_t i = 0; i < nnodes; ++i)
    NW_REQUIRE(nodes[i] >= 0 && nodes[i] < c->g.nnode, NW_E_ARG,
               "bad node id %d", nodes[i]);
  int32_t* d_nodes = nullptr;
  NW_HIP(hipMalloc(&d_nodes, sizeof(int32_t) * nnodes));
  NW_HIP(hipMemcpyAsync(d_nodes, nodes, sizeof(int32_t) * nnodes,
                        hipMemcpyHostToDevice, c->stream));
  const int64_t nt = nnodes * (end - begin);
  k_reset_rows<<<grid_for(nt, 256), 256, 0, c->stream>>>(
    d_nodes, nnodes, begin, end, c->g.ndof, c->g.row_ptr, c->g.cols, c->g.vals,
    c->g.rhs, diag_value, rhs_value, bc, sol, fnc);
  hipError_t le = hipGetLastError();
  NW_HIP(hipStreamSynchronize(c->stream));
  (void)hipFree(d_nodes);
  NW_REQUIRE(le == hipSuccess, NW_E_HIP, "reset kernel: %s",
             hipGetErrorString(le));
  return NW_OK;
}

int nw_linsys_reset_rows(nw_ctx* c, int64_t nnodes, const int32_t* nodes,
                         int begin, int end, double diag_value,
                         double rhs_residual)
{
  const int rc = reset_rows_impl(c, nnodes, nodes, begin, end, diag_value,
                                 rhs_residual, nullptr, nullptr, 1);
  // a row reset lifts the skip check until the next zeroSystem
  // (HypreLinSysCoeffApplier::resetRows, src/HypreLinearSystem.C:2326)
  if (rc == NW_OK) c->g.skip_check = false;
  return rc;
}

int nw_linsys_apply_dirichlet(nw_ctx* c, int64_t nnodes, const int32_t* nodes,
                              int sol_field, int bc_field, int begin, int end)
{
  NW_REQUIRE(c && c->g.ndof, NW_E_STATE, "linear system not built");
  NW_REQUIRE(!field_is_edge(sol_field) && !field_is_edge(bc_field) &&
               field_components(sol_field) == field_components(bc_field) &&
               end <= field_components(sol_field),
             NW_E_ARG, "solution/bc fields must be node fields of equal width");
  const double *sol = nullptr, *bc = nullptr;
  int rc = require_field(c, sol_field, &sol);
  if (rc) return rc;
  rc = require_field(c, bc_field, &bc);
  if (rc) return rc;
  return reset_rows_impl(c, nnodes, nodes, begin, end, 1.0, 0.0, bc, sol,
                         field_components(sol_field));
}

int nw_fix_pressure_at_node(nw_ctx* c, int32_t node, double pref)
{
  NW_REQUIRE(c && c->g.ndof == 1, NW_E_STATE,
             "fix-pressure needs a built 1-dof system");
  NW_REQUIRE(node >= 0 && node < c->g.nnode, NW_E_ARG, "bad node id %d", node);
  int rc = nw_linsys_reset_rows(c, 1, &node, 0, 1, 0.0, 0.0);
  if (rc) return rc;
  const double* p = nullptr;
  rc = require_field(c, NW_F_PRESSURE, &p);
  if (rc) return rc;
  double pn = 0.0;
  NW_HIP(hipMemcpyAsync(&pn, p + node, sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  const double l = 1.0, r = pref - pn;
  return nw_linsys_sum_into(c, 1, 1, &node, &l, &r);
}

int nw_linsys_solution_upload(nw_ctx* c, const double* host)
{
  NW_REQUIRE(c && c->g.ndof && host, NW_E_STATE, "linear system not built");
  NW_HIP(hipMemcpyAsync(c->g.sln, host, sizeof(double) * c->g.nnode * c->g.ndof,
                        hipMemcpyHostToDevice, c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  return NW_OK;
}

int nw_linsys_solution_download(nw_ctx* c, double* host)
{
  NW_REQUIRE(c && c->g.ndof && host, NW_E_STATE, "linear system not built");
  NW_HIP(hipMemcpyAsync(host, c->g.sln, sizeof(double) * c->g.nnode * c->g.ndof,
                        hipMemcpyDeviceToHost, c->stream));
  NW_HIP(hipStreamSynchronize(c->stream));
  return NW_OK;
}

int nw_linsys_solution_device_ptr(nw_ctx* c, double** dptr)
{
  NW_REQUIRE(c && c->g.ndof && dptr, NW_E_STATE, "linear system not built");
  *dptr = c->g.sln;
  return NW_OK;
}

}  // extern "C"
