// nw_exodus.cpp -- on-disk mesh input and output for the edge path: Exodus II
// files stored in the netCDF classic format (CDF-1) or its 64-bit-offset
// form (CDF-2).  The reader takes what an edge realm needs from a mesh
// database -- coordinates, element blocks of Hex8 / Tet4 / Wed6 / Pyr5, node
// sets, side sets and nodal variables of a time step -- and the writer
// produces the same content with one element block per topology.
//
// The netCDF layer follows the published classic-format grammar: a header
// (magic "CDF" + version byte, record count, dimension, global attribute and
// variable lists, names and values padded to 4 bytes, all big-endian), then
// the fixed-size variables at their stored offsets and the record variables
// interleaved record by record.  netCDF-4 (HDF5) and CDF-5 files are reported
// as unsupported rather than guessed at.  The Exodus layer uses the Exodus II
// variable names (coordx/coordy/coordz or coord, connect<b> with its
// elem_type attribute, eb_prop1, node_ns<s>, elem_ss<s>/side_ss<s>,
// name_nod_var, vals_nod_var<v>, time_whole) and 1-based node and element
// numbers on disk; everything handed to the caller is 0-based.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "nalu_amd.h"
#include "nalu_amd_mesh.h"

namespace nw {
void set_error(const char* fmt, ...);
}

namespace {

enum NcType { kByte = 1, kChar = 2, kShort = 3, kInt = 4, kFloat = 5, kDouble = 6 };
constexpr uint32_t kTagDim = 0x0A, kTagVar = 0x0B, kTagAtt = 0x0C;
constexpr uint64_t kStreaming = 0xFFFFFFFFu;

int type_size(int t)
{
  switch (t) {
    case kByte: case kChar: return 1;
    case kShort: return 2;
    case kInt: case kFloat: return 4;
    case kDouble: return 8;
  }
  return 0;
}

uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

uint32_t get32(const uint8_t* p)
{
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}
uint64_t get64(const uint8_t* p) { return (uint64_t)get32(p) << 32 | get32(p + 4); }
void put32(uint8_t* p, uint32_t v)
{
  p[0] = v >> 24; p[1] = v >> 16; p[2] = v >> 8; p[3] = v;
}
void put64(uint8_t* p, uint64_t v) { put32(p, v >> 32); put32(p + 4, (uint32_t)v); }

struct NcAtt {
  std::string name;
  int type = 0;
  uint64_t n = 0;
  std::vector<uint8_t> raw;  // big-endian values, unpadded
  std::string text() const
  {
    std::string s(raw.begin(), raw.end());
    return s.substr(0, s.find('\0'));
  }
};
struct NcDim {
  std::string name;
  uint64_t len = 0;  // 0 = the record dimension
};
struct NcVar {
  std::string name;
  std::vector<int> dims;
  std::vector<NcAtt> atts;
  int type = 0;
  uint64_t vsize = 0, begin = 0;
  bool rec = false;
  const NcAtt* att(const char* n) const
  {
    for (const auto& a : atts)
      if (a.name == n) return &a;
    return nullptr;
  }
};

// Big-endian values of type `t` at src -> T.
template <class T>
void decode(int t, const uint8_t* src, T* dst, int64_t n)
{
#pragma omp parallel for schedule(static) if (n > (1 << 16))
  for (int64_t i = 0; i < n; ++i) {
    switch (t) {
      case kByte: dst[i] = (T)(int8_t)src[i]; break;
      case kChar: dst[i] = (T)src[i]; break;
      case kShort: dst[i] = (T)(int16_t)(src[2 * i] << 8 | src[2 * i + 1]); break;
      case kInt: dst[i] = (T)(int32_t)get32(src + 4 * i); break;
      case kFloat: {
        const uint32_t u = get32(src + 4 * i);
        float f;
        std::memcpy(&f, &u, 4);
        dst[i] = (T)f;
        break;
      }
      default: {
        const uint64_t u = get64(src + 8 * i);
        double d;
        std::memcpy(&d, &u, 8);
        dst[i] = (T)d;
      }
    }
  }
}

// Bounds-checked reader over the header bytes.
struct Cursor {
  const uint8_t* p;
  size_t n, at = 0;
  bool short_read = false;
  bool need(uint64_t k)
  {
    if (short_read || at + k > n) {
      short_read = true;
      return false;
    }
    return true;
  }
  uint32_t u32()
  {
    if (!need(4)) return 0;
    at += 4;
    return get32(p + at - 4);
  }
  uint64_t u64()
  {
    if (!need(8)) return 0;
    at += 8;
    return get64(p + at - 8);
  }
  std::string name()
  {
    const uint32_t len = u32();
    if (!need(pad4(len))) return {};
    std::string s((const char*)p + at, len);
    at += pad4(len);
    return s;
  }
};

struct NcFile {
  int fd = -1;
  int version = 0;
  uint64_t numrecs = 0, recsize = 0, file_size = 0;
  std::vector<NcDim> dims;
  std::vector<NcAtt> gatts;
  std::vector<NcVar> vars;
  std::string path;

  ~NcFile()
  {
    if (fd >= 0) ::close(fd);
  }

  const NcVar* var(const std::string& name) const
  {
    for (const auto& v : vars)
      if (v.name == name) return &v;
    return nullptr;
  }
  uint64_t dim_len(const std::string& name) const
  {
    for (const auto& d : dims)
      if (d.name == name) return d.len;
    return 0;
  }
  // values of one record of a record variable, or of the whole variable
  uint64_t count(const NcVar& v) const
  {
    uint64_t c = 1;
    for (size_t k = v.rec ? 1 : 0; k < v.dims.size(); ++k) c *= dims[v.dims[k]].len;
    return c;
  }

  int read_raw(uint64_t off, uint64_t bytes, void* dst) const
  {
    uint8_t* d = (uint8_t*)dst;
    while (bytes) {
      const ssize_t r = ::pread(fd, d, std::min<uint64_t>(bytes, 1u << 30), (off_t)off);
      if (r <= 0) {
        nw::set_error("%s: read of %llu bytes at offset %llu failed (%s)", path.c_str(),
                      (unsigned long long)bytes, (unsigned long long)off,
                      r < 0 ? std::strerror(errno) : "end of file");
        return NW_E_ARG;
      }
      d += r;
      off += (uint64_t)r;
      bytes -= (uint64_t)r;
    }
    return NW_OK;
  }

  // Values [first, first + n) of variable v (record `rec` of a record
  // variable) converted to T.
  template <class T>
  int read(const NcVar& v, uint64_t rec, uint64_t first, uint64_t n, T* out) const
  {
    const int ts = type_size(v.type);
    if (v.rec && rec >= numrecs) {
      nw::set_error("%s: variable %s has %llu records, record %llu asked for", path.c_str(),
                    v.name.c_str(), (unsigned long long)numrecs, (unsigned long long)rec);
      return NW_E_ARG;
    }
    if (first + n > count(v)) {
      nw::set_error("%s: variable %s holds %llu values, %llu asked for", path.c_str(),
                    v.name.c_str(), (unsigned long long)count(v),
                    (unsigned long long)(first + n));
      return NW_E_ARG;
    }
    uint64_t off = v.begin + (v.rec ? rec * recsize : 0) + first * ts;
    const uint64_t chunk = 1u << 22;
    std::vector<uint8_t> buf((size_t)std::min<uint64_t>(n, chunk) * ts);
    for (uint64_t i = 0; i < n; i += chunk) {
      const uint64_t m = std::min(chunk, n - i);
      if (off + m * ts > file_size) {
        nw::set_error("%s: variable %s runs past the end of the file (truncated?)",
                      path.c_str(), v.name.c_str());
        return NW_E_ARG;
      }
      int rc = read_raw(off, m * ts, buf.data());
      if (rc) return rc;
      decode(v.type, buf.data(), out + i, (int64_t)m);
      off += m * ts;
    }
    return NW_OK;
  }

  // Rows of a [n][len] character variable as strings.
  int strings(const NcVar& v, std::vector<std::string>& out) const
  {
    if (v.type != kChar || v.dims.size() != 2 || v.rec) {
      nw::set_error("%s: %s is not a table of names", path.c_str(), v.name.c_str());
      return NW_E_ARG;
    }
    const uint64_t rows = dims[v.dims[0]].len, len = dims[v.dims[1]].len;
    std::vector<uint8_t> raw(rows * len);
    if (rows * len > 0) {
      int rc = read_raw(v.begin, rows * len, raw.data());
      if (rc) return rc;
    }
    out.resize(rows);
    for (uint64_t r = 0; r < rows; ++r) {
      std::string s((const char*)raw.data() + r * len, len);
      out[r] = s.substr(0, s.find('\0'));
    }
    return NW_OK;
  }

  int parse(const uint8_t* p, size_t n);
  int open(const char* file);
};

int NcFile::parse(const uint8_t* p, size_t n)
{
  Cursor c{p, n};
  const bool off64 = version == 2;
  numrecs = c.u32();
  auto att_list = [&](std::vector<NcAtt>& atts) -> bool {
    const uint32_t tag = c.u32(), cnt = c.u32();
    if (tag == 0 && cnt == 0) return true;
    if (tag != kTagAtt) return false;
    if (!c.need(12ull * cnt)) return true;  // each entry takes 12 bytes or more
    atts.resize(cnt);
    for (auto& a : atts) {
      a.name = c.name();
      a.type = (int)c.u32();
      a.n = c.u32();
      const int ts = type_size(a.type);
      if (!ts) return false;
      if (!c.need(pad4(a.n * ts))) return true;  // short read: caller rereads
      a.raw.assign(p + c.at, p + c.at + a.n * ts);
      c.at += pad4(a.n * ts);
    }
    return true;
  };
  // dimensions
  uint32_t tag = c.u32(), cnt = c.u32();
  bool ok = (tag == 0 && cnt == 0) || tag == kTagDim;
  if (ok && tag == kTagDim && c.need(8ull * cnt)) {
    dims.resize(cnt);
    for (auto& d : dims) {
      d.name = c.name();
      d.len = c.u32();
    }
  }
  ok = ok && att_list(gatts);
  tag = c.u32();
  cnt = c.u32();
  ok = ok && ((tag == 0 && cnt == 0) || tag == kTagVar);
  if (ok && tag == kTagVar && c.need(24ull * cnt)) {
    vars.resize(cnt);
    for (auto& v : vars) {
      v.name = c.name();
      const uint32_t nd = c.u32();
      if (!c.need(4ull * nd)) break;
      v.dims.resize(nd);
      for (auto& d : v.dims) {
        d = (int)c.u32();
        if (!c.short_read && (d < 0 || (size_t)d >= dims.size())) ok = false;
      }
      if (!ok || !att_list(v.atts)) {
        ok = false;
        break;
      }
      v.type = (int)c.u32();
      v.vsize = c.u32();
      v.begin = off64 ? c.u64() : c.u32();
      if (c.short_read) break;
      if (!type_size(v.type)) ok = false;
      v.rec = !v.dims.empty() && dims[v.dims[0]].len == 0;
      // only the first dimension may be the record dimension
      for (size_t k = 1; k < v.dims.size(); ++k)
        if (dims[v.dims[k]].len == 0) ok = false;
    }
  }
  if (c.short_read) return 1;  // the header goes on past the bytes read
  if (!ok) {
    nw::set_error("%s: malformed netCDF header", path.c_str());
    return NW_E_ARG;
  }
  if (numrecs == kStreaming) {
    nw::set_error("%s: netCDF file in streaming mode (record count not set)", path.c_str());
    return NW_E_ARG;
  }
  // one record: every record variable's slice, each padded to 4 bytes
  // unless there is only one record variable
  int nrec = 0;
  for (const auto& v : vars) nrec += v.rec;
  recsize = 0;
  for (const auto& v : vars)
    if (v.rec) {
      const uint64_t b = count(v) * type_size(v.type);
      recsize += nrec == 1 ? b : pad4(b);
    }
  return NW_OK;
}

int NcFile::open(const char* file)
{
  path = file;
  fd = ::open(file, O_RDONLY);
  if (fd < 0) {
    nw::set_error("%s: cannot open (%s)", file, std::strerror(errno));
    return NW_E_ARG;
  }
  struct stat st;
  if (fstat(fd, &st) != 0) {
    nw::set_error("%s: cannot stat (%s)", file, std::strerror(errno));
    return NW_E_ARG;
  }
  file_size = (uint64_t)st.st_size;
  uint8_t magic[4] = {0, 0, 0, 0};
  if (file_size < 4 || read_raw(0, 4, magic)) {
    nw::set_error("%s: not a netCDF file (too short)", file);
    return NW_E_ARG;
  }
  if (magic[0] == 0x89 && magic[1] == 'H' && magic[2] == 'D' && magic[3] == 'F') {
    nw::set_error("%s: netCDF-4 (HDF5) file; only the netCDF classic formats (CDF-1, "
                  "CDF-2) are read: convert it with 'nccopy -k 64-bit-offset'", file);
    return NW_E_ARG;
  }
  if (magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F') {
    nw::set_error("%s: not a netCDF file", file);
    return NW_E_ARG;
  }
  if (magic[3] != 1 && magic[3] != 2) {
    nw::set_error("%s: netCDF format version %d; only the classic formats (CDF-1, CDF-2) "
                  "are read: convert it with 'nccopy -k 64-bit-offset'", file, magic[3]);
    return NW_E_ARG;
  }
  version = magic[3];
  for (uint64_t want = 1 << 16;; want *= 4) {
    const uint64_t n = std::min<uint64_t>(want, file_size) - 4;
    std::vector<uint8_t> hdr(n);
    int rc = read_raw(4, n, hdr.data());
    if (rc) return rc;
    dims.clear();
    gatts.clear();
    vars.clear();
    rc = parse(hdr.data(), hdr.size());
    if (rc <= 0) return rc;
    if (n + 4 == file_size) {
      nw::set_error("%s: malformed netCDF header (it runs past the end of the file)", file);
      return NW_E_ARG;
    }
  }
}

// ---- Exodus II ----------------------------------------------------------

// Faces of each topology by side ordinal, ordered so that the right-hand
// rule gives the outward normal (the Exodus II side numbering; the
// sideNodeOrdinals_ tables of include/master_element/Hex8CVFEM.h:395-401,
// Tet4CVFEM.h:258-263, Wed6CVFEM.h:259-265, Pyr5CVFEM.h:291-297).
const int kNumSides[4] = {6, 4, 5, 5};
const int kSideNodes[4][6][5] = {
  {{4, 0, 1, 5, 4}, {4, 1, 2, 6, 5}, {4, 2, 3, 7, 6}, {4, 0, 4, 7, 3}, {4, 0, 3, 2, 1},
   {4, 4, 5, 6, 7}},
  {{3, 0, 1, 3}, {3, 1, 2, 3}, {3, 0, 3, 2}, {3, 0, 2, 1}},
  {{4, 0, 1, 4, 3}, {4, 1, 2, 5, 4}, {4, 0, 3, 5, 2}, {3, 0, 2, 1}, {3, 3, 4, 5}},
  {{3, 0, 1, 4}, {3, 1, 2, 4}, {3, 2, 3, 4}, {3, 0, 4, 3}, {4, 0, 3, 2, 1}}};
const int kNpe[4] = {8, 4, 6, 5};
const char* const kElemType[4] = {"HEX8", "TETRA4", "WEDGE6", "PYRAMID5"};
const char* const kBlockName[4] = {"block_hex8", "block_tet4", "block_wedge6",
                                   "block_pyramid5"};

// Exodus element type name + nodes per element -> topology (-1: not one
// of the four linear 3-D topologies of the edge path).
int topo_of(std::string t, uint64_t npe)
{
  for (auto& ch : t) ch = (char)std::toupper((unsigned char)ch);
  auto starts = [&](const char* s) { return t.compare(0, std::strlen(s), s) == 0; };
  if (starts("HEX") && npe == 8) return 0;
  if (starts("TET") && npe == 4) return 1;
  if (starts("WEDGE") && npe == 6) return 2;
  if (starts("PYR") && npe == 5) return 3;
  return -1;
}

struct Block {
  int32_t id = 0;
  std::string name, etype;
  int topo = -1;
  uint64_t npe = 0;
  int64_t n = 0, first = 0;
  const NcVar* conn = nullptr;
};
struct Set {
  int32_t id = 0;
  std::string name;
  int64_t n = 0;
  const NcVar* a = nullptr;  // node_ns<s> / elem_ss<s>
  const NcVar* b = nullptr;  // side_ss<s>
};

}  // namespace

struct nw_exo {
  NcFile nc;
  int64_t nnode = 0, nelem = 0;
  std::vector<Block> blocks;
  std::vector<Set> nsets, ssets;
  std::vector<std::string> nodvars;
  const NcVar* time = nullptr;
};

namespace {

int read_ids(const NcFile& nc, const char* prop, int64_t n, std::vector<int32_t>& ids)
{
  ids.resize(n);
  for (int64_t i = 0; i < n; ++i) ids[i] = (int32_t)(i + 1);
  const NcVar* v = nc.var(prop);
  if (!v || n == 0) return NW_OK;
  return nc.read<int32_t>(*v, 0, 0, n, ids.data());
}

int read_names(const NcFile& nc, const char* var, int64_t n, std::vector<std::string>& names)
{
  names.assign(n, std::string());
  const NcVar* v = nc.var(var);
  if (!v || n == 0) return NW_OK;
  std::vector<std::string> s;
  int rc = nc.strings(*v, s);
  if (rc) return rc;
  for (int64_t i = 0; i < n && i < (int64_t)s.size(); ++i) names[i] = s[i];
  return NW_OK;
}

int parse_exodus(nw_exo& x)
{
  NcFile& nc = x.nc;
  const char* f = nc.path.c_str();
  if (nc.dim_len("num_dim") != 3) {
    nw::set_error("%s: not a 3-D Exodus II mesh (num_dim = %llu)", f,
                  (unsigned long long)nc.dim_len("num_dim"));
    return NW_E_ARG;
  }
  x.nnode = (int64_t)nc.dim_len("num_nodes");
  x.nelem = (int64_t)nc.dim_len("num_elem");
  if (x.nnode >= INT32_MAX || x.nelem >= INT32_MAX) {
    nw::set_error("%s: %lld nodes / %lld elements exceed 32-bit ids", f,
                  (long long)x.nnode, (long long)x.nelem);
    return NW_E_ARG;
  }
  char nm[64];
  // element blocks
  const int64_t nblk = (int64_t)nc.dim_len("num_el_blk");
  std::vector<int32_t> ids;
  std::vector<std::string> names;
  int rc = read_ids(nc, "eb_prop1", nblk, ids);
  if (!rc) rc = read_names(nc, "eb_names", nblk, names);
  if (rc) return rc;
  int64_t first = 0;
  for (int64_t b = 0; b < nblk; ++b) {
    Block blk;
    blk.id = ids[b];
    blk.name = names[b];
    std::snprintf(nm, sizeof nm, "num_el_in_blk%lld", (long long)b + 1);
    blk.n = (int64_t)nc.dim_len(nm);
    std::snprintf(nm, sizeof nm, "num_nod_per_el%lld", (long long)b + 1);
    blk.npe = nc.dim_len(nm);
    std::snprintf(nm, sizeof nm, "connect%lld", (long long)b + 1);
    blk.conn = nc.var(nm);
    blk.first = first;
    if (blk.n > 0) {
      if (!blk.conn || blk.conn->rec || blk.conn->dims.size() != 2 ||
          nc.count(*blk.conn) != (uint64_t)blk.n * blk.npe) {
        nw::set_error("%s: element block %d: no connectivity %s of %lld x %llu", f, blk.id,
                      nm, (long long)blk.n, (unsigned long long)blk.npe);
        return NW_E_ARG;
      }
      const NcAtt* a = blk.conn->att("elem_type");
      blk.etype = a ? a->text() : std::string();
      blk.topo = topo_of(blk.etype, blk.npe);
    }
    first += blk.n;
    x.blocks.push_back(blk);
  }
  if (first != x.nelem) {
    nw::set_error("%s: element blocks hold %lld elements, num_elem is %lld", f,
                  (long long)first, (long long)x.nelem);
    return NW_E_ARG;
  }
  // node sets and side sets
  for (int kind = 0; kind < 2; ++kind) {
    const int64_t ns = (int64_t)nc.dim_len(kind ? "num_side_sets" : "num_node_sets");
    rc = read_ids(nc, kind ? "ss_prop1" : "ns_prop1", ns, ids);
    if (!rc) rc = read_names(nc, kind ? "ss_names" : "ns_names", ns, names);
    if (rc) return rc;
    for (int64_t s = 0; s < ns; ++s) {
      Set set;
      set.id = ids[s];
      set.name = names[s];
      std::snprintf(nm, sizeof nm, kind ? "num_side_ss%lld" : "num_nod_ns%lld", (long long)s + 1);
      set.n = (int64_t)nc.dim_len(nm);
      std::snprintf(nm, sizeof nm, kind ? "elem_ss%lld" : "node_ns%lld", (long long)s + 1);
      set.a = nc.var(nm);
      if (kind) {
        std::snprintf(nm, sizeof nm, "side_ss%lld", (long long)s + 1);
        set.b = nc.var(nm);
      }
      if (set.n > 0 && (!set.a || set.a->rec || nc.count(*set.a) != (uint64_t)set.n ||
                        (kind && (!set.b || set.b->rec || nc.count(*set.b) != (uint64_t)set.n)))) {
        nw::set_error("%s: %s set %d: its entries are missing or of the wrong size", f,
                      kind ? "side" : "node", set.id);
        return NW_E_ARG;
      }
      (kind ? x.ssets : x.nsets).push_back(set);
    }
  }
  // nodal variables
  const int64_t nv = (int64_t)nc.dim_len("num_nod_var");
  rc = read_names(nc, "name_nod_var", nv, x.nodvars);
  if (rc) return rc;
  x.time = nc.var("time_whole");
  return NW_OK;
}

const char* kind_name(int kind)
{
  return kind == NW_EXO_BLOCK ? "element block" : kind == NW_EXO_NODE_SET ? "node set"
         : kind == NW_EXO_SIDE_SET ? "side set" : "nodal variable";
}

// ---- netCDF writer ------------------------------------------------------

struct WVar {
  std::string name;
  std::vector<int> dims;
  std::vector<NcAtt> atts;
  int type = kInt;
  // writes values [first, first + n) big-endian at dst
  std::function<void(uint8_t*, uint64_t, uint64_t)> fill;
};

NcAtt text_att(const char* name, const std::string& s)
{
  NcAtt a;
  a.name = name;
  a.type = kChar;
  a.n = s.size();
  a.raw.assign(s.begin(), s.end());
  return a;
}
NcAtt int_att(const char* name, int32_t v)
{
  NcAtt a;
  a.name = name;
  a.type = kInt;
  a.n = 1;
  a.raw.resize(4);
  put32(a.raw.data(), (uint32_t)v);
  return a;
}
NcAtt float_att(const char* name, float v)
{
  NcAtt a;
  a.name = name;
  a.type = kFloat;
  a.n = 1;
  a.raw.resize(4);
  uint32_t u;
  std::memcpy(&u, &v, 4);
  put32(a.raw.data(), u);
  return a;
}

struct Writer {
  std::vector<NcDim> dims;
  std::vector<NcAtt> gatts;
  std::vector<WVar> vars;
  uint64_t numrecs = 0;

  int dim(const std::string& name, uint64_t len)
  {
    dims.push_back({name, len});
    return (int)dims.size() - 1;
  }
  bool is_rec(const WVar& v) const { return !v.dims.empty() && dims[v.dims[0]].len == 0; }
  uint64_t count(const WVar& v) const
  {
    uint64_t c = 1;
    for (size_t k = is_rec(v) ? 1 : 0; k < v.dims.size(); ++k) c *= dims[v.dims[k]].len;
    return c;
  }

  static void name(std::vector<uint8_t>& h, const std::string& s)
  {
    const size_t at = h.size();
    h.resize(at + 4 + pad4(s.size()), 0);
    put32(&h[at], (uint32_t)s.size());
    std::memcpy(&h[at + 4], s.data(), s.size());
  }
  static void u32(std::vector<uint8_t>& h, uint32_t v)
  {
    h.resize(h.size() + 4);
    put32(&h[h.size() - 4], v);
  }
  static void atts(std::vector<uint8_t>& h, const std::vector<NcAtt>& a)
  {
    u32(h, a.empty() ? 0 : kTagAtt);
    u32(h, (uint32_t)a.size());
    for (const auto& x : a) {
      name(h, x.name);
      u32(h, (uint32_t)x.type);
      u32(h, (uint32_t)x.n);
      const size_t at = h.size();
      h.resize(at + pad4(x.raw.size()), 0);
      std::memcpy(&h[at], x.raw.data(), x.raw.size());
    }
  }
  std::vector<uint8_t> header(const std::vector<uint64_t>& begin) const
  {
    std::vector<uint8_t> h = {'C', 'D', 'F', 2};
    u32(h, (uint32_t)numrecs);
    u32(h, dims.empty() ? 0 : kTagDim);
    u32(h, (uint32_t)dims.size());
    for (const auto& d : dims) {
      name(h, d.name);
      u32(h, (uint32_t)d.len);
    }
    atts(h, gatts);
    u32(h, vars.empty() ? 0 : kTagVar);
    u32(h, (uint32_t)vars.size());
    for (size_t i = 0; i < vars.size(); ++i) {
      const WVar& v = vars[i];
      name(h, v.name);
      u32(h, (uint32_t)v.dims.size());
      for (int d : v.dims) u32(h, (uint32_t)d);
      atts(h, v.atts);
      u32(h, (uint32_t)v.type);
      const uint64_t vs = pad4(count(v) * type_size(v.type));
      u32(h, vs > 0xFFFFFFFCu ? 0xFFFFFFFFu : (uint32_t)vs);
      h.resize(h.size() + 8);
      put64(&h[h.size() - 8], begin[i]);
    }
    return h;
  }

  int write(const char* path) const
  {
    // fixed-size variables first, in order, then the record section
    std::vector<uint64_t> begin(vars.size(), 0);
    const uint64_t hsize = header(begin).size();
    uint64_t off = hsize;
    int nrec = 0;
    for (size_t i = 0; i < vars.size(); ++i)
      if (!is_rec(vars[i])) {
        const uint64_t b = count(vars[i]) * type_size(vars[i].type);
        if (b > 0xFFFFFFFCu && i + 1 != vars.size()) {
          nw::set_error("%s: variable %s (%llu bytes) is too large for a CDF-2 file", path,
                        vars[i].name.c_str(), (unsigned long long)b);
          return NW_E_ARG;
        }
        begin[i] = off;
        off += pad4(b);
      } else {
        ++nrec;
      }
    for (size_t i = 0; i < vars.size(); ++i)
      if (is_rec(vars[i])) {
        begin[i] = off;
        const uint64_t b = count(vars[i]) * type_size(vars[i].type);
        off += nrec == 1 ? b : pad4(b);
      }
    FILE* fp = std::fopen(path, "wb");
    if (!fp) {
      nw::set_error("%s: cannot create (%s)", path, std::strerror(errno));
      return NW_E_ARG;
    }
    std::vector<uint8_t> h = header(begin);
    bool ok = std::fwrite(h.data(), 1, h.size(), fp) == h.size();
    std::vector<uint8_t> buf;
    auto emit = [&](const WVar& v, bool pad) {
      const int ts = type_size(v.type);
      const uint64_t n = count(v), chunk = 1u << 20;
      for (uint64_t i = 0; ok && i < n; i += chunk) {
        const uint64_t m = std::min(chunk, n - i);
        buf.assign(m * ts, 0);
        v.fill(buf.data(), i, m);
        ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
      }
      const uint64_t tail = pad ? pad4(n * ts) - n * ts : 0;
      const uint8_t zero[4] = {0, 0, 0, 0};
      if (ok && tail) ok = std::fwrite(zero, 1, tail, fp) == tail;
    };
    for (const auto& v : vars)
      if (!is_rec(v)) emit(v, true);
    for (uint64_t r = 0; r < numrecs; ++r)
      for (const auto& v : vars)
        if (is_rec(v)) emit(v, nrec != 1);
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) {
      nw::set_error("%s: write failed (%s)", path, std::strerror(errno));
      return NW_E_ARG;
    }
    return NW_OK;
  }
};

void fill_i32(uint8_t* dst, uint64_t n, const std::function<int32_t(uint64_t)>& f,
              uint64_t first)
{
#pragma omp parallel for schedule(static) if (n > (1 << 16))
  for (int64_t i = 0; i < (int64_t)n; ++i) put32(dst + 4 * i, (uint32_t)f(first + i));
}
void fill_f64(uint8_t* dst, uint64_t n, const std::function<double(uint64_t)>& f,
              uint64_t first)
{
#pragma omp parallel for schedule(static) if (n > (1 << 16))
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const double d = f(first + i);
    uint64_t u;
    std::memcpy(&u, &d, 8);
    put64(dst + 8 * i, u);
  }
}

}  // namespace

// ---- C ABI -----------------------------------------------------------------

int nw_exo_open(const char* path, nw_exo** exo)
{
  if (!path || !exo) return NW_E_ARG;
  *exo = nullptr;
  nw_exo* x = new nw_exo;
  int rc = x->nc.open(path);
  if (!rc) rc = parse_exodus(*x);
  if (rc) {
    delete x;
    return rc;
  }
  *exo = x;
  return NW_OK;
}

void nw_exo_close(nw_exo* exo) { delete exo; }

int nw_exo_sizes(const nw_exo* x, nw_exo_info* info)
{
  if (!x || !info) return NW_E_ARG;
  info->nnode = x->nnode;
  info->nelem = x->nelem;
  info->nblock = (int32_t)x->blocks.size();
  info->nnodeset = (int32_t)x->nsets.size();
  info->nsideset = (int32_t)x->ssets.size();
  info->nnodvar = (int32_t)x->nodvars.size();
  info->ntime = x->time ? (int32_t)x->nc.numrecs : 0;
  info->format = x->nc.version;
  return NW_OK;
}

int nw_exo_entity(const nw_exo* x, int kind, int32_t index, int32_t* id, int64_t* count,
                  int32_t* topo, char* name, int32_t name_cap)
{
  if (!x) return NW_E_ARG;
  const int64_t n = kind == NW_EXO_BLOCK ? (int64_t)x->blocks.size()
                    : kind == NW_EXO_NODE_SET ? (int64_t)x->nsets.size()
                    : kind == NW_EXO_SIDE_SET ? (int64_t)x->ssets.size()
                    : kind == NW_EXO_NODAL_VAR ? (int64_t)x->nodvars.size() : -1;
  if (n < 0) {
    nw::set_error("nw_exo_entity: unknown kind %d", kind);
    return NW_E_ARG;
  }
  if (index < 0 || index >= n) {
    nw::set_error("nw_exo_entity: %s %d of %lld", kind_name(kind), index, (long long)n);
    return NW_E_ARG;
  }
  int32_t i = index + 1, t = -1;
  int64_t c = x->nnode;
  std::string s;
  if (kind == NW_EXO_BLOCK) {
    const Block& b = x->blocks[index];
    i = b.id, c = b.n, t = b.topo, s = b.name;
  } else if (kind == NW_EXO_NODE_SET || kind == NW_EXO_SIDE_SET) {
    const Set& e = (kind == NW_EXO_NODE_SET ? x->nsets : x->ssets)[index];
    i = e.id, c = e.n, s = e.name;
  } else {
    s = x->nodvars[index];
  }
  if (id) *id = i;
  if (count) *count = c;
  if (topo) *topo = t;
  if (name && name_cap > 0) {
    std::strncpy(name, s.c_str(), (size_t)name_cap - 1);
    name[name_cap - 1] = '\0';
  }
  return NW_OK;
}

int nw_exo_find(const nw_exo* x, int kind, const char* name, int32_t* index)
{
  if (!x || !name || !index) return NW_E_ARG;
  std::vector<std::string> names;
  if (kind == NW_EXO_BLOCK)
    for (const auto& b : x->blocks) names.push_back(b.name);
  else if (kind == NW_EXO_NODE_SET || kind == NW_EXO_SIDE_SET)
    for (const auto& e : kind == NW_EXO_NODE_SET ? x->nsets : x->ssets) names.push_back(e.name);
  else if (kind == NW_EXO_NODAL_VAR)
    names = x->nodvars;
  else
    return NW_E_ARG;
  for (size_t i = 0; i < names.size(); ++i)
    if (names[i] == name) {
      *index = (int32_t)i;
      return NW_OK;
    }
  nw::set_error("%s: no %s named '%s'", x->nc.path.c_str(), kind_name(kind), name);
  return NW_E_ARG;
}

int nw_exo_read_mesh(const nw_exo* x, double* coords, int32_t* type, int32_t* conn,
                     int32_t* block)
{
  if (!x || !coords || !type || !conn) return NW_E_ARG;
  const NcFile& nc = x->nc;
  const char* f = nc.path.c_str();
  const int64_t nn = x->nnode;
  // coordinates: coordx/coordy/coordz, or one coord [num_dim][num_nodes]
  std::vector<double> c1((size_t)nn);
  const NcVar* whole = nc.var("coord");
  for (int d = 0; d < 3; ++d) {
    const char* nm = d == 0 ? "coordx" : d == 1 ? "coordy" : "coordz";
    const NcVar* v = nc.var(nm);
    int rc;
    if (v && !v->rec && nc.count(*v) == (uint64_t)nn)
      rc = nc.read<double>(*v, 0, 0, nn, c1.data());
    else if (whole && !whole->rec && nc.count(*whole) == 3 * (uint64_t)nn)
      rc = nc.read<double>(*whole, 0, (uint64_t)d * nn, nn, c1.data());
    else {
      nw::set_error("%s: no node coordinates (coordx/coordy/coordz or coord)", f);
      return NW_E_ARG;
    }
    if (rc) return rc;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nn; ++i) coords[3 * i + d] = c1[i];
  }
  // element blocks, in file order
  std::vector<int32_t> cb;
  for (size_t b = 0; b < x->blocks.size(); ++b) {
    const Block& blk = x->blocks[b];
    if (blk.n == 0) continue;
    if (blk.topo < 0) {
      nw::set_error("%s: element block %d ('%s') has elements of type '%s' with %llu nodes; "
                    "the edge path takes Hex8, Tet4, Wed6 and Pyr5", f, blk.id,
                    blk.name.c_str(), blk.etype.c_str(), (unsigned long long)blk.npe);
      return NW_E_ARG;
    }
    cb.resize((size_t)blk.n * blk.npe);
    int rc = nc.read<int32_t>(*blk.conn, 0, 0, cb.size(), cb.data());
    if (rc) return rc;
    int64_t bad = -1;
#pragma omp parallel for schedule(static) reduction(max : bad)
    for (int64_t e = 0; e < blk.n; ++e) {
      const int64_t g = blk.first + e;
      type[g] = blk.topo;
      if (block) block[g] = (int32_t)b;
      for (int k = 0; k < 8; ++k) {
        int32_t v = -1;
        if (k < (int)blk.npe) {
          v = cb[e * blk.npe + k] - 1;
          if (v < 0 || v >= nn) bad = std::max(bad, e);
        }
        conn[8 * g + k] = v;
      }
    }
    if (bad >= 0) {
      nw::set_error("%s: element %lld of block %d names a node outside 1..%lld", f,
                    (long long)(bad + 1), blk.id, (long long)nn);
      return NW_E_ARG;
    }
  }
  return NW_OK;
}

int nw_exo_node_set(const nw_exo* x, int32_t index, int32_t* nodes)
{
  if (!x || !nodes || index < 0 || index >= (int32_t)x->nsets.size()) return NW_E_ARG;
  const Set& s = x->nsets[index];
  if (s.n == 0) return NW_OK;
  int rc = x->nc.read<int32_t>(*s.a, 0, 0, s.n, nodes);
  if (rc) return rc;
  for (int64_t i = 0; i < s.n; ++i) {
    if (nodes[i] < 1 || nodes[i] > x->nnode) {
      nw::set_error("%s: node set %d names node %d outside 1..%lld", x->nc.path.c_str(), s.id,
                    nodes[i], (long long)x->nnode);
      return NW_E_ARG;
    }
    nodes[i] -= 1;
  }
  return NW_OK;
}

int nw_exo_side_set(const nw_exo* x, int32_t index, int32_t* elem, int32_t* side)
{
  if (!x || !elem || !side || index < 0 || index >= (int32_t)x->ssets.size()) return NW_E_ARG;
  const Set& s = x->ssets[index];
  if (s.n == 0) return NW_OK;
  int rc = x->nc.read<int32_t>(*s.a, 0, 0, s.n, elem);
  if (!rc) rc = x->nc.read<int32_t>(*s.b, 0, 0, s.n, side);
  if (rc) return rc;
  for (int64_t i = 0; i < s.n; ++i) {
    const int32_t e = elem[i] - 1;
    int topo = -1;
    if (e >= 0 && e < x->nelem)
      for (const auto& b : x->blocks)
        if (e >= b.first && e < b.first + b.n) topo = b.topo;
    if (topo < 0 || side[i] < 1 || side[i] > kNumSides[topo]) {
      nw::set_error("%s: side set %d entry %lld (element %d, side %d) is not a side of an "
                    "element of the mesh", x->nc.path.c_str(), s.id, (long long)i, elem[i],
                    side[i]);
      return NW_E_ARG;
    }
    elem[i] = e;
    side[i] -= 1;
  }
  return NW_OK;
}

int nw_exo_read_nodal_var(const nw_exo* x, int32_t index, int32_t step, double* values,
                          double* time)
{
  if (!x || !values || index < 0 || index >= (int32_t)x->nodvars.size() || step < 0)
    return NW_E_ARG;
  const NcFile& nc = x->nc;
  char nm[64];
  std::snprintf(nm, sizeof nm, "vals_nod_var%d", index + 1);
  const NcVar* v = nc.var(nm);
  int rc;
  if (v && v->rec && nc.count(*v) == (uint64_t)x->nnode) {
    rc = nc.read<double>(*v, (uint64_t)step, 0, x->nnode, values);
  } else if ((v = nc.var("vals_nod_var")) && v->rec &&
             nc.count(*v) == (uint64_t)x->nnode * x->nodvars.size()) {
    rc = nc.read<double>(*v, (uint64_t)step, (uint64_t)index * x->nnode, x->nnode, values);
  } else {
    nw::set_error("%s: no values of nodal variable '%s'", nc.path.c_str(),
                  x->nodvars[index].c_str());
    return NW_E_ARG;
  }
  if (!rc && time) {
    *time = 0.0;
    if (x->time) rc = nc.read<double>(*x->time, (uint64_t)step, 0, 1, time);
  }
  return rc;
}

int nw_side_nodes(int32_t type, int32_t side, int32_t* nodes)
{
  if (type < 0 || type > 3 || side < 0 || side >= kNumSides[type] || !nodes) return NW_E_ARG;
  const int* s = kSideNodes[type][side];
  for (int k = 0; k < s[0]; ++k) nodes[k] = s[k + 1];
  return s[0];
}

int nw_sides_to_nodes(int64_t nelem, const int32_t* type, const int32_t* conn, int64_t nnode,
                      int64_t nside, const int32_t* elem, const int32_t* side, int32_t* nodes,
                      int64_t* count)
{
  if (nelem < 0 || nnode < 0 || nside < 0 || !count || (nside && (!type || !conn || !elem || !side)))
    return NW_E_ARG;
  std::vector<uint8_t> mark((size_t)nnode, 0);
  for (int64_t i = 0; i < nside; ++i) {
    const int64_t e = elem[i];
    if (e < 0 || e >= nelem || type[e] < 0 || type[e] > 3 || side[i] < 0 ||
        side[i] >= kNumSides[type[e]]) {
      nw::set_error("nw_sides_to_nodes: entry %lld (element %lld, side %d) is not a side of "
                    "an element of the mesh", (long long)i, (long long)e, side[i]);
      return NW_E_ARG;
    }
    const int* s = kSideNodes[type[e]][side[i]];
    for (int k = 1; k <= s[0]; ++k) {
      const int32_t v = conn[8 * e + s[k]];
      if (v < 0 || v >= nnode) {
        nw::set_error("nw_sides_to_nodes: element %lld names node %d of %lld", (long long)e, v,
                      (long long)nnode);
        return NW_E_ARG;
      }
      mark[v] = 1;
    }
  }
  int64_t c = 0;
  for (int64_t v = 0; v < nnode; ++v)
    if (mark[v]) {
      if (nodes) nodes[c] = (int32_t)v;
      ++c;
    }
  *count = c;
  return NW_OK;
}

int nw_exo_write(const char* path, const char* title, int64_t nnode, const double* coords,
                 int64_t nelem, const int32_t* type, const int32_t* conn,
                 const nw_exo_sets* node_sets, const nw_exo_sets* side_sets, int32_t nvar,
                 const char* const* var_names, const double* var_values, double time,
                 int64_t* elem_order)
{
  if (!path || nnode <= 0 || nnode >= INT32_MAX || !coords || nelem < 0 ||
      nelem >= INT32_MAX || (nelem && (!type || !conn)) || nvar < 0 ||
      (nvar && (!var_names || !var_values))) {
    nw::set_error("nw_exo_write: bad arguments");
    return NW_E_ARG;
  }
  // elements grouped by topology (stable): one block per topology present
  std::vector<int64_t> order((size_t)nelem), pos((size_t)nelem);
  int64_t cnt[4] = {0, 0, 0, 0};
  for (int64_t e = 0; e < nelem; ++e) {
    if (type[e] < 0 || type[e] > 3) {
      nw::set_error("nw_exo_write: element %lld has type %d", (long long)e, type[e]);
      return NW_E_ARG;
    }
    for (int k = 0; k < kNpe[type[e]]; ++k)
      if (conn[8 * e + k] < 0 || conn[8 * e + k] >= nnode) {
        nw::set_error("nw_exo_write: element %lld names node %d of %lld", (long long)e,
                      conn[8 * e + k], (long long)nnode);
        return NW_E_ARG;
      }
    ++cnt[type[e]];
  }
  int64_t start[5] = {0, 0, 0, 0, 0};
  for (int t = 0; t < 4; ++t) start[t + 1] = start[t] + cnt[t];
  {
    int64_t at[4] = {start[0], start[1], start[2], start[3]};
    for (int64_t e = 0; e < nelem; ++e) {
      order[at[type[e]]] = e;
      pos[e] = at[type[e]]++;
    }
  }
  auto check_sets = [&](const nw_exo_sets* s, bool sides) -> int {
    const char* what = sides ? "side" : "node";
    if (!s || s->count == 0) return NW_OK;
    if (s->count < 0 || !s->offsets || !s->entries || (sides && !s->sides)) {
      nw::set_error("nw_exo_write: %s sets: count %d without offsets/entries%s", what,
                    s->count, sides ? "/sides" : "");
      return NW_E_ARG;
    }
    for (int32_t i = 0; i < s->count; ++i)
      if (s->offsets[i] < 0 || s->offsets[i + 1] < s->offsets[i]) {
        nw::set_error("nw_exo_write: %s set %d: offsets not ascending", what, i);
        return NW_E_ARG;
      }
    for (int64_t j = s->offsets[0]; j < s->offsets[s->count]; ++j) {
      const int64_t v = s->entries[j];
      if (v < 0 || v >= (sides ? nelem : nnode) ||
          (sides && (s->sides[j] < 0 || s->sides[j] >= kNumSides[type[v]]))) {
        nw::set_error("nw_exo_write: %s set entry %lld is out of range", what, (long long)j);
        return NW_E_ARG;
      }
    }
    return NW_OK;
  };
  int rc = check_sets(node_sets, false);
  if (!rc) rc = check_sets(side_sets, true);
  if (rc) return rc;

  Writer w;
  w.numrecs = nvar > 0 ? 1 : 0;
  w.gatts.push_back(float_att("api_version", 8.03f));
  w.gatts.push_back(float_att("version", 8.03f));
  w.gatts.push_back(int_att("floating_point_word_size", 8));
  w.gatts.push_back(int_att("file_size", 1));
  w.gatts.push_back(int_att("maximum_name_length", 32));
  w.gatts.push_back(int_att("int64_status", 0));
  w.gatts.push_back(text_att("title", title && *title ? title : "mesh"));
  const int d_time = w.dim("time_step", 0);
  const int d_len_name = w.dim("len_name", 33);
  const int d_dim = w.dim("num_dim", 3);
  const int d_nodes = w.dim("num_nodes", (uint64_t)nnode);
  int nblk = 0;
  for (int t = 0; t < 4; ++t) nblk += cnt[t] > 0;
  const int d_elem = nelem ? w.dim("num_elem", (uint64_t)nelem) : -1;
  const int d_blk = nblk ? w.dim("num_el_blk", (uint64_t)nblk) : -1;
  const int nns = node_sets ? node_sets->count : 0, nss = side_sets ? side_sets->count : 0;
  const int d_ns = nns ? w.dim("num_node_sets", (uint64_t)nns) : -1;
  const int d_ss = nss ? w.dim("num_side_sets", (uint64_t)nss) : -1;
  const int d_nvar = nvar ? w.dim("num_nod_var", (uint64_t)nvar) : -1;

  auto ints = [](std::function<int32_t(uint64_t)> f) {
    return [f](uint8_t* dst, uint64_t first, uint64_t n) { fill_i32(dst, n, f, first); };
  };
  auto reals = [](std::function<double(uint64_t)> f) {
    return [f](uint8_t* dst, uint64_t first, uint64_t n) { fill_f64(dst, n, f, first); };
  };
  auto names = [](std::vector<std::string> s) {
    return [s](uint8_t* dst, uint64_t first, uint64_t n) {
      for (uint64_t i = first; i < first + n; ++i) {
        const std::string& x = s[i / 33];
        const size_t k = i % 33;
        dst[i - first] = k < x.size() && k < 32 ? (uint8_t)x[k] : 0;
      }
    };
  };
  auto add = [&](const std::string& name, std::vector<int> d, int type,
                 std::function<void(uint8_t*, uint64_t, uint64_t)> fill,
                 std::vector<NcAtt> atts = {}) {
    WVar v;
    v.name = name;
    v.dims = std::move(d);
    v.type = type;
    v.fill = std::move(fill);
    v.atts = std::move(atts);
    w.vars.push_back(std::move(v));
  };

  add("time_whole", {d_time}, kDouble, reals([time](uint64_t) { return time; }));
  std::vector<int> btopo;
  for (int t = 0; t < 4; ++t)
    if (cnt[t]) btopo.push_back(t);
  if (nblk) {
    add("eb_status", {d_blk}, kInt, ints([](uint64_t) { return 1; }));
    add("eb_prop1", {d_blk}, kInt, ints([](uint64_t i) { return (int32_t)i + 1; }),
        {text_att("name", "ID")});
    std::vector<std::string> bn;
    for (int t : btopo) bn.push_back(kBlockName[t]);
    add("eb_names", {d_blk, d_len_name}, kChar, names(bn));
  }
  auto set_vars = [&](const nw_exo_sets* s, bool sides, int dcount) {
    const char* pre = sides ? "ss" : "ns";
    add(std::string(pre) + "_status", {dcount}, kInt, ints([](uint64_t) { return 1; }));
    add(std::string(pre) + "_prop1", {dcount}, kInt,
        ints([s](uint64_t i) { return s->ids ? s->ids[i] : (int32_t)i + 1; }),
        {text_att("name", "ID")});
    std::vector<std::string> sn;
    for (int32_t i = 0; i < s->count; ++i)
      sn.push_back(s->names && s->names[i] ? s->names[i] : "");
    add(std::string(pre) + "_names", {dcount, d_len_name}, kChar, names(sn));
  };
  if (nns) set_vars(node_sets, false, d_ns);
  if (nss) set_vars(side_sets, true, d_ss);
  for (int d = 0; d < 3; ++d)
    add(d == 0 ? "coordx" : d == 1 ? "coordy" : "coordz", {d_nodes}, kDouble,
        reals([coords, d](uint64_t i) { return coords[3 * i + d]; }));
  add("coor_names", {d_dim, d_len_name}, kChar, names({"x", "y", "z"}));
  (void)d_elem;
  for (size_t b = 0; b < btopo.size(); ++b) {
    const int t = btopo[b];
    char nm[64];
    std::snprintf(nm, sizeof nm, "num_el_in_blk%zu", b + 1);
    const int de = w.dim(nm, (uint64_t)cnt[t]);
    std::snprintf(nm, sizeof nm, "num_nod_per_el%zu", b + 1);
    const int dn = w.dim(nm, (uint64_t)kNpe[t]);
    std::snprintf(nm, sizeof nm, "connect%zu", b + 1);
    const int64_t s0 = start[t], npe = kNpe[t];
    const int64_t* ord = order.data();
    add(nm, {de, dn}, kInt,
        ints([=](uint64_t i) { return conn[8 * ord[s0 + (int64_t)i / npe] + (int64_t)i % npe] + 1; }),
        {text_att("elem_type", kElemType[t])});
  }
  for (int kind = 0; kind < 2; ++kind) {
    const nw_exo_sets* s = kind ? side_sets : node_sets;
    for (int32_t i = 0; s && i < s->count; ++i) {
      char nm[64];
      const int64_t o = s->offsets[i], n = s->offsets[i + 1] - o;
      if (n == 0) continue;  // netCDF has no empty fixed dimension; the set reads as empty
      std::snprintf(nm, sizeof nm, kind ? "num_side_ss%d" : "num_nod_ns%d", i + 1);
      const int dn = w.dim(nm, (uint64_t)n);
      const int32_t* ent = s->entries + o;
      if (!kind) {
        std::snprintf(nm, sizeof nm, "node_ns%d", i + 1);
        add(nm, {dn}, kInt, ints([ent](uint64_t j) { return ent[j] + 1; }));
      } else {
        const int64_t* p = pos.data();
        const int32_t* sd = s->sides + o;
        std::snprintf(nm, sizeof nm, "elem_ss%d", i + 1);
        add(nm, {dn}, kInt, ints([ent, p](uint64_t j) { return (int32_t)p[ent[j]] + 1; }));
        std::snprintf(nm, sizeof nm, "side_ss%d", i + 1);
        add(nm, {dn}, kInt, ints([sd](uint64_t j) { return sd[j] + 1; }));
      }
    }
  }
  if (nvar) {
    std::vector<std::string> vn;
    for (int32_t v = 0; v < nvar; ++v) vn.push_back(var_names[v] ? var_names[v] : "");
    add("name_nod_var", {d_nvar, d_len_name}, kChar, names(vn));
    for (int32_t v = 0; v < nvar; ++v) {
      char nm[64];
      std::snprintf(nm, sizeof nm, "vals_nod_var%d", v + 1);
      const double* val = var_values + (size_t)v * nnode;
      add(nm, {d_time, d_nodes}, kDouble, reals([val](uint64_t i) { return val[i]; }));
    }
  }
  rc = w.write(path);
  if (rc) return rc;
  if (elem_order) std::copy(order.begin(), order.end(), elem_order);
  return NW_OK;
}
