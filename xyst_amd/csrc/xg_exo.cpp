// ExodusII mesh and nodal-field files in the NetCDF classic format (CDF-1
// and CDF-2, the format of every mesh and result file the reference's
// regression tests hold), read and written without the SEACAS/NetCDF
// libraries.  Host code only.
//
// Reader: coordinates, element blocks (tetrahedra, triangles), side sets as
// boundary triangles, node-variable names, time values and node variables,
// mirroring ExodusIIMeshReader (IO/ExodusIIMeshReader.hpp:61-136).
// Writer: coordinates, a triangle and a tetrahedron block, side sets naming
// triangle elements, node-variable names and per-step node variables, as
// ExodusIIMeshWriter writes them (IO/ExodusIIMeshWriter.cpp:85-221:
// ex_put_init, ex_put_coord, blocks, side sets, time values, node scalars).
//
// NetCDF classic layout: a big-endian header (magic "CDF" + version byte,
// record count, dimension list, global attributes, variable list with each
// variable's dimensions, attributes, type, padded size and data offset:
// 32-bit offsets in CDF-1, 64-bit in CDF-2), then the fixed-size variables,
// then the records.  Record r of a time-dependent variable lies at
// begin + r * recsize, recsize being the sum of the padded per-record sizes
// of all record variables (the unpadded size when there is only one).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "xg_internal.hpp"

namespace {

enum : int { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum : uint32_t { TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C };

int type_size(int t) {
  switch (t) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    default: return 0;
  }
}

int64_t pad4(int64_t n) { return (n + 3) & ~int64_t(3); }

struct NcAtt {
  std::string name;
  int type = 0;
  int64_t n = 0;
  std::vector<uint8_t> raw;   // big-endian values as stored
};

struct NcVar {
  std::string name;
  std::vector<int> dims;
  std::vector<NcAtt> atts;
  int type = 0;
  int64_t vsize = 0;   // padded bytes (one record's worth for record vars)
  int64_t begin = 0;
  bool rec = false;
  int64_t count = 0;   // values per record (record vars) or in total
};

struct NcDim {
  std::string name;
  int64_t len = 0;     // 0: the record (unlimited) dimension
};

// ------------------------------------------------------------ byte order --

uint32_t be32(const uint8_t* p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}
uint64_t be64(const uint8_t* p) { return (uint64_t)be32(p) << 32 | be32(p + 4); }
void put32(std::vector<uint8_t>& b, uint32_t v) {
  for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}
void put64(std::vector<uint8_t>& b, uint64_t v) {
  put32(b, (uint32_t)(v >> 32));
  put32(b, (uint32_t)v);
}
void put_name(std::vector<uint8_t>& b, const std::string& s) {
  put32(b, (uint32_t)s.size());
  b.insert(b.end(), s.begin(), s.end());
  while (b.size() % 4) b.push_back(0);
}

// big-endian array of n values of type t -> doubles / int64
bool decode_f64(int t, const uint8_t* p, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    switch (t) {
      case NC_DOUBLE: { uint64_t u = be64(p + 8 * i); std::memcpy(&out[i], &u, 8); break; }
      case NC_FLOAT: { uint32_t u = be32(p + 4 * i); float f; std::memcpy(&f, &u, 4); out[i] = f; break; }
      case NC_INT: out[i] = (double)(int32_t)be32(p + 4 * i); break;
      case NC_SHORT: out[i] = (double)(int16_t)((uint16_t)p[2 * i] << 8 | p[2 * i + 1]); break;
      case NC_BYTE: out[i] = (double)(int8_t)p[i]; break;
      default: return false;
    }
  }
  return true;
}
bool decode_i64(int t, const uint8_t* p, int64_t n, int64_t* out) {
  for (int64_t i = 0; i < n; ++i) {
    switch (t) {
      case NC_INT: out[i] = (int32_t)be32(p + 4 * i); break;
      case NC_SHORT: out[i] = (int16_t)((uint16_t)p[2 * i] << 8 | p[2 * i + 1]); break;
      case NC_BYTE: out[i] = (int8_t)p[i]; break;
      default: return false;
    }
  }
  return true;
}

}  // namespace

// An open ExodusII file (reader, appender or writer).
struct xg_exo {
  std::string path, err;
  std::FILE* fp = nullptr;
  bool writable = false;
  int version = 0;               // 1 (CDF-1) or 2 (CDF-2)
  int64_t numrecs = 0, recsize = 0;
  std::vector<NcDim> dims;
  std::vector<NcAtt> gatts;
  std::vector<NcVar> vars;
  std::unordered_map<std::string, int> vidx;
  std::map<std::string, int64_t> dimlen;
  ~xg_exo() {
    if (fp) std::fclose(fp);
  }
};

namespace {

int exo_fail(xg_exo* h, int code, const std::string& msg) {
  h->err = msg + (h->path.empty() ? "" : " (" + h->path + ")");
  return code;
}

// ------------------------------------------------------- header parsing --

struct Cursor {
  const std::vector<uint8_t>& b;
  size_t p = 0;
  bool ok = true;
  explicit Cursor(const std::vector<uint8_t>& buf) : b(buf) {}
  bool need(size_t n) {
    if (!ok || p + n > b.size()) ok = false;
    return ok;
  }
  uint32_t u32() {
    if (!need(4)) return 0;
    uint32_t v = be32(&b[p]);
    p += 4;
    return v;
  }
  uint64_t u64() {
    if (!need(8)) return 0;
    uint64_t v = be64(&b[p]);
    p += 8;
    return v;
  }
  std::string name() {
    uint32_t n = u32();
    if (n > (1u << 20) || !need(pad4(n))) { ok = false; return std::string(); }
    std::string s((const char*)&b[p], n);
    p += pad4(n);
    return s;
  }
};

bool parse_atts(Cursor& c, std::vector<NcAtt>& atts) {
  uint32_t tag = c.u32(), n = c.u32();
  if (!c.ok) return false;
  if (tag == 0 && n == 0) return true;
  if (tag != TAG_ATT || n > (1u << 20)) return false;
  atts.resize(n);
  for (auto& a : atts) {
    a.name = c.name();
    a.type = (int)c.u32();
    a.n = c.u32();
    const int ts = type_size(a.type);
    if (!c.ok || ts == 0) return false;
    const int64_t nb = a.n * ts;
    if (!c.need(pad4(nb))) return false;
    a.raw.assign(c.b.begin() + c.p, c.b.begin() + c.p + nb);
    c.p += pad4(nb);
  }
  return c.ok;
}

// reads and parses the header; the header of any file is far smaller than
// the first 64 KiB plus whatever the variable list needs, so the buffer grows
// until the parse completes
int read_header(xg_exo* h) {
  std::vector<uint8_t> buf;
  size_t want = 1 << 16;
  std::fseek(h->fp, 0, SEEK_END);
  const long fsize = std::ftell(h->fp);
  if (fsize < 8) return exo_fail(h, XG_EIO, "exodus: file too short for a NetCDF header");
  for (;;) {
    const size_t n = std::min((size_t)fsize, want);
    buf.resize(n);
    std::fseek(h->fp, 0, SEEK_SET);
    if (std::fread(buf.data(), 1, n, h->fp) != n)
      return exo_fail(h, XG_EIO, "exodus: read error");
    if (buf[0] != 'C' || buf[1] != 'D' || buf[2] != 'F' || (buf[3] != 1 && buf[3] != 2))
      return exo_fail(h, XG_EIO,
                      "exodus: not a NetCDF classic (CDF-1/CDF-2) file; NetCDF-4/HDF5 and "
                      "CDF-5 files are not supported");
    h->version = buf[3];
    Cursor c(buf);
    c.p = 4;
    h->numrecs = c.u32();
    bool good = c.ok;
    // dimensions
    h->dims.clear();
    uint32_t tag = c.u32(), nd = c.u32();
    if (good && c.ok && !(tag == 0 && nd == 0)) {
      if (tag != TAG_DIM || nd > (1u << 20)) good = false;
      for (uint32_t i = 0; good && i < nd; ++i) {
        NcDim d;
        d.name = c.name();
        d.len = c.u32();
        h->dims.push_back(d);
        good = c.ok;
      }
    }
    h->gatts.clear();
    good = good && c.ok && parse_atts(c, h->gatts);
    h->vars.clear();
    tag = c.u32();
    uint32_t nv = c.u32();
    if (good && c.ok && !(tag == 0 && nv == 0)) {
      if (tag != TAG_VAR || nv > (1u << 20)) good = false;
      for (uint32_t i = 0; good && i < nv; ++i) {
        NcVar v;
        v.name = c.name();
        uint32_t nvd = c.u32();
        if (nvd > 64) { good = false; break; }
        for (uint32_t k = 0; k < nvd; ++k) v.dims.push_back((int)c.u32());
        good = c.ok && parse_atts(c, v.atts);
        v.type = (int)c.u32();
        v.vsize = c.u32();
        v.begin = h->version == 2 ? (int64_t)c.u64() : (int64_t)c.u32();
        good = good && c.ok;
        h->vars.push_back(v);
      }
    }
    if (good && c.ok) break;
    if (n == (size_t)fsize)
      return exo_fail(h, XG_EIO, "exodus: malformed or truncated NetCDF header");
    want *= 4;
  }
  // derived: record flags, value counts, recsize
  h->vidx.clear();
  h->dimlen.clear();
  for (auto& d : h->dims) h->dimlen[d.name] = d.len;
  h->recsize = 0;
  int nrec = 0;
  for (size_t i = 0; i < h->vars.size(); ++i) {
    NcVar& v = h->vars[i];
    if (type_size(v.type) == 0)
      return exo_fail(h, XG_EIO, "exodus: variable " + v.name + " has an unknown type");
    v.count = 1;
    for (size_t k = 0; k < v.dims.size(); ++k) {
      const int d = v.dims[k];
      if (d < 0 || d >= (int)h->dims.size())
        return exo_fail(h, XG_EIO, "exodus: variable " + v.name + " names a bad dimension");
      if (h->dims[d].len == 0) {
        if (k != 0)
          return exo_fail(h, XG_EIO, "exodus: record dimension not first in " + v.name);
        v.rec = true;
      } else {
        v.count *= h->dims[d].len;
      }
    }
    if (v.rec) {
      h->recsize += pad4(v.count * type_size(v.type));
      ++nrec;
    }
    h->vidx[v.name] = (int)i;
  }
  if (nrec == 1)
    for (auto& v : h->vars)
      if (v.rec) h->recsize = v.count * type_size(v.type);
  if (h->numrecs == 0xFFFFFFFFu) {   // streaming: count the records present
    int64_t end0 = 0;
    for (auto& v : h->vars)
      if (v.rec) end0 = end0 ? std::min(end0, v.begin) : v.begin;
    h->numrecs = h->recsize ? (fsize - end0) / h->recsize : 0;
  }
  // every fixed variable must lie inside the file
  for (auto& v : h->vars)
    if (!v.rec && v.begin + v.count * type_size(v.type) > fsize)
      return exo_fail(h, XG_EIO, "exodus: variable " + v.name + " extends past the end of file");
  return XG_OK;
}

const NcVar* find_var(const xg_exo* h, const std::string& name) {
  auto it = h->vidx.find(name);
  return it == h->vidx.end() ? nullptr : &h->vars[it->second];
}

int64_t dim(const xg_exo* h, const std::string& name) {
  auto it = h->dimlen.find(name);
  return it == h->dimlen.end() ? -1 : it->second;
}

std::string att_text(const NcVar& v, const char* name) {
  for (auto& a : v.atts)
    if (a.name == name && a.type == NC_CHAR) {
      std::string s(a.raw.begin(), a.raw.end());
      while (!s.empty() && (s.back() == '\0' || s.back() == ' ')) s.pop_back();
      return s;
    }
  return std::string();
}

int read_raw(xg_exo* h, const NcVar& v, int64_t rec, std::vector<uint8_t>& raw) {
  const int64_t nb = v.count * type_size(v.type);
  int64_t off = v.begin;
  if (v.rec) {
    if (rec < 0 || rec >= h->numrecs)
      return exo_fail(h, XG_EINVAL, "exodus: step out of range for " + v.name);
    off += rec * h->recsize;
  }
  raw.resize(nb);
  if (nb == 0) return XG_OK;
  if (std::fseek(h->fp, (long)off, SEEK_SET) != 0 ||
      std::fread(raw.data(), 1, nb, h->fp) != (size_t)nb)
    return exo_fail(h, XG_EIO, "exodus: short read of " + v.name);
  return XG_OK;
}

int read_f64(xg_exo* h, const std::string& name, int64_t rec, std::vector<double>& out) {
  const NcVar* v = find_var(h, name);
  if (!v) return exo_fail(h, XG_EIO, "exodus: no variable " + name);
  std::vector<uint8_t> raw;
  int rc = read_raw(h, *v, rec, raw);
  if (rc) return rc;
  out.resize(v->count);
  if (!decode_f64(v->type, raw.data(), v->count, out.data()))
    return exo_fail(h, XG_EIO, "exodus: variable " + name + " is not numeric");
  return XG_OK;
}

int read_i64(xg_exo* h, const std::string& name, std::vector<int64_t>& out) {
  const NcVar* v = find_var(h, name);
  if (!v) return exo_fail(h, XG_EIO, "exodus: no variable " + name);
  std::vector<uint8_t> raw;
  int rc = read_raw(h, *v, 0, raw);
  if (rc) return rc;
  out.resize(v->count);
  if (!decode_i64(v->type, raw.data(), v->count, out.data()))
    return exo_fail(h, XG_EIO, "exodus: variable " + name + " is not an integer variable");
  return XG_OK;
}

// element blocks: (first element id, count, nodes per element, element type)
struct Blk {
  int64_t first = 0, n = 0, nnode = 0;
  std::string type;
};

int blocks(xg_exo* h, std::vector<Blk>& b) {
  b.clear();
  const int64_t nblk = dim(h, "num_el_blk");
  int64_t first = 0;
  for (int64_t k = 1; k <= nblk; ++k) {
    const std::string s = std::to_string(k);
    const NcVar* v = find_var(h, "connect" + s);
    Blk B;
    B.first = first;
    B.n = std::max<int64_t>(0, dim(h, "num_el_in_blk" + s));
    B.nnode = std::max<int64_t>(0, dim(h, "num_nod_per_el" + s));
    if (v) B.type = att_text(*v, "elem_type");
    first += B.n;
    b.push_back(B);
  }
  return XG_OK;
}

bool is_tet(const Blk& B) {
  return B.nnode == 4 && (B.type.empty() || B.type.compare(0, 3, "TET") == 0);
}
bool is_tri(const Blk& B) {
  return B.nnode == 3 && (B.type.empty() || B.type.compare(0, 3, "TRI") == 0);
}

// connectivity of all blocks of one kind, 0-based
int gather_blocks(xg_exo* h, bool tets, std::vector<int64_t>& conn) {
  std::vector<Blk> b;
  blocks(h, b);
  conn.clear();
  for (size_t k = 0; k < b.size(); ++k) {
    if (tets ? !is_tet(b[k]) : !is_tri(b[k])) continue;
    std::vector<int64_t> c;
    int rc = read_i64(h, "connect" + std::to_string(k + 1), c);
    if (rc) return rc;
    // the callers size their buffers from the block dimensions
    if ((int64_t)c.size() != b[k].n * b[k].nnode)
      return exo_fail(h, XG_EIO, "exodus: connect" + std::to_string(k + 1) +
                                     " size disagrees with its block dimensions");
    for (int64_t& x : c) x -= 1;
    conn.insert(conn.end(), c.begin(), c.end());
  }
  return XG_OK;
}

// ExodusII side numbering of a TETRA4 (1-based side -> local nodes)
constexpr int kTetSide[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};

// ----------------------------------------------------------------- writer --

struct WVar {
  std::string name;
  std::vector<std::string> dims;
  int type;
  std::vector<std::pair<std::string, std::string>> text_atts;
};

}  // namespace

extern "C" {

const char* xg_exo_error(const xg_exo* h) { return h ? h->err.c_str() : "null exodus handle"; }

void xg_exo_close(xg_exo* h) { delete h; }

int xg_exo_open(const char* path, int append, xg_exo** out) {
  if (!out) return XG_EINVAL;
  *out = nullptr;
  xg_exo* h = new (std::nothrow) xg_exo;
  if (!h) return XG_EINVAL;
  *out = h;
  if (!path) return exo_fail(h, XG_EINVAL, "exodus: null path");
  h->path = path;
  h->fp = std::fopen(path, append ? "r+b" : "rb");
  if (!h->fp) return exo_fail(h, XG_EIO, "exodus: cannot open file");
  h->writable = append != 0;
  return read_header(h);
}

int xg_exo_sizes(xg_exo* h, int64_t* npoin, int64_t* ntet, int64_t* ntri,
                 int* nsideset, int* nnodevar, int64_t* nstep) {
  if (!h) return XG_EINVAL;
  std::vector<Blk> b;
  blocks(h, b);
  int64_t nt = 0, nf = 0;
  for (auto& B : b) {
    if (is_tet(B)) nt += B.n;
    else if (is_tri(B)) nf += B.n;
  }
  if (npoin) *npoin = std::max<int64_t>(0, dim(h, "num_nodes"));
  if (ntet) *ntet = nt;
  if (ntri) *ntri = nf;
  if (nsideset) *nsideset = (int)std::max<int64_t>(0, dim(h, "num_side_sets"));
  if (nnodevar) *nnodevar = (int)std::max<int64_t>(0, dim(h, "num_nod_var"));
  if (nstep) *nstep = h->numrecs;
  return XG_OK;
}

int xg_exo_coords(xg_exo* h, double* x, double* y, double* z) {
  if (!h || !x || !y || !z) return XG_EINVAL;
  const int64_t n = dim(h, "num_nodes");
  if (n < 0) return exo_fail(h, XG_EIO, "exodus: no num_nodes dimension");
  double* dst[3] = {x, y, z};
  if (find_var(h, "coordx")) {
    const char* nm[3] = {"coordx", "coordy", "coordz"};
    for (int j = 0; j < 3; ++j) {
      std::vector<double> v;
      int rc = read_f64(h, nm[j], 0, v);
      if (rc) return rc;
      if ((int64_t)v.size() != n) return exo_fail(h, XG_EIO, "exodus: coordinate size");
      std::memcpy(dst[j], v.data(), n * sizeof(double));
    }
    return XG_OK;
  }
  std::vector<double> v;   // older files: coord(num_dim, num_nodes)
  int rc = read_f64(h, "coord", 0, v);
  if (rc) return rc;
  if ((int64_t)v.size() != 3 * n) return exo_fail(h, XG_EIO, "exodus: expected 3-d coordinates");
  for (int j = 0; j < 3; ++j) std::memcpy(dst[j], v.data() + j * n, n * sizeof(double));
  return XG_OK;
}

int xg_exo_tets(xg_exo* h, int64_t* inpoel) {
  if (!h || !inpoel) return XG_EINVAL;
  std::vector<int64_t> c;
  int rc = gather_blocks(h, true, c);
  if (rc) return rc;
  std::memcpy(inpoel, c.data(), c.size() * sizeof(int64_t));
  return XG_OK;
}

int xg_exo_tris(xg_exo* h, int64_t* triinpoel) {
  if (!h || !triinpoel) return XG_EINVAL;
  std::vector<int64_t> c;
  int rc = gather_blocks(h, false, c);
  if (rc) return rc;
  std::memcpy(triinpoel, c.data(), c.size() * sizeof(int64_t));
  return XG_OK;
}

int xg_exo_sideset(xg_exo* h, int k, int* id, int64_t* nface) {
  if (!h || k < 0) return XG_EINVAL;
  const int64_t ns = dim(h, "num_side_sets");
  if (k >= ns) return exo_fail(h, XG_EINVAL, "exodus: side set index out of range");
  if (id) {
    std::vector<int64_t> ids;
    if (find_var(h, "ss_prop1")) {
      int rc = read_i64(h, "ss_prop1", ids);
      if (rc) return rc;
      *id = (int)ids[k];
    } else {
      *id = k + 1;
    }
  }
  if (nface) *nface = std::max<int64_t>(0, dim(h, "num_side_ss" + std::to_string(k + 1)));
  return XG_OK;
}

int xg_exo_sideset_faces(xg_exo* h, int k, int64_t* tri) {
  if (!h || !tri || k < 0) return XG_EINVAL;
  int64_t nf = 0;
  int rc = xg_exo_sideset(h, k, nullptr, &nf);
  if (rc) return rc;
  if (nf == 0) return XG_OK;
  const std::string s = std::to_string(k + 1);
  std::vector<int64_t> elem, side;
  if ((rc = read_i64(h, "elem_ss" + s, elem)) || (rc = read_i64(h, "side_ss" + s, side)))
    return rc;
  std::vector<Blk> b;
  blocks(h, b);
  std::vector<std::vector<int64_t>> conn(b.size());
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t e = elem[f] - 1;   // global 0-based element id
    size_t kb = 0;
    while (kb < b.size() && !(e >= b[kb].first && e < b[kb].first + b[kb].n)) ++kb;
    if (kb == b.size()) return exo_fail(h, XG_EIO, "exodus: side set element out of range");
    if (conn[kb].empty()) {
      rc = read_i64(h, "connect" + std::to_string(kb + 1), conn[kb]);
      if (rc) return rc;
    }
    const int64_t* en = conn[kb].data() + (e - b[kb].first) * b[kb].nnode;
    if (is_tri(b[kb])) {
      for (int a = 0; a < 3; ++a) tri[f * 3 + a] = en[a] - 1;
    } else if (is_tet(b[kb])) {
      const int64_t sd = side[f];
      if (sd < 1 || sd > 4) return exo_fail(h, XG_EIO, "exodus: tetrahedron side out of range");
      for (int a = 0; a < 3; ++a) tri[f * 3 + a] = en[kTetSide[sd - 1][a]] - 1;
    } else {
      return exo_fail(h, XG_EUNSUPPORTED,
                      "exodus: side set on an element that is neither a triangle nor a "
                      "tetrahedron");
    }
  }
  return XG_OK;
}

int xg_exo_node_var_name(xg_exo* h, int var, char* buf, int len) {
  if (!h || !buf || len <= 0 || var < 0) return XG_EINVAL;
  const int64_t nv = dim(h, "num_nod_var");
  if (var >= nv) return exo_fail(h, XG_EINVAL, "exodus: node variable out of range");
  const NcVar* v = find_var(h, "name_nod_var");
  if (!v || v->type != NC_CHAR || v->dims.size() != 2)
    return exo_fail(h, XG_EIO, "exodus: no node variable names");
  std::vector<uint8_t> raw;
  int rc = read_raw(h, *v, 0, raw);
  if (rc) return rc;
  const int64_t w = h->dims[v->dims[1]].len;
  std::string s((const char*)raw.data() + var * w, w);
  s = s.substr(0, s.find('\0'));
  while (!s.empty() && s.back() == ' ') s.pop_back();
  std::snprintf(buf, len, "%s", s.c_str());
  return XG_OK;
}

int xg_exo_times(xg_exo* h, double* t) {
  if (!h || !t) return XG_EINVAL;
  const NcVar* v = find_var(h, "time_whole");
  if (!v) return h->numrecs ? exo_fail(h, XG_EIO, "exodus: no time_whole") : XG_OK;
  for (int64_t r = 0; r < h->numrecs; ++r) {
    std::vector<double> x;
    int rc = read_f64(h, "time_whole", r, x);
    if (rc) return rc;
    t[r] = x.at(0);
  }
  return XG_OK;
}

int xg_exo_read_node_var(xg_exo* h, int64_t step, int var, double* out) {
  if (!h || !out || var < 0) return XG_EINVAL;
  if (var >= dim(h, "num_nod_var"))
    return exo_fail(h, XG_EINVAL, "exodus: node variable out of range");
  std::vector<double> x;
  int rc = read_f64(h, "vals_nod_var" + std::to_string(var + 1), step, x);
  if (rc) return rc;
  std::memcpy(out, x.data(), x.size() * sizeof(double));
  return XG_OK;
}

int xg_exo_create(const char* path, int64_t npoin, const double* x, const double* y,
                  const double* z, int64_t ntet, const int64_t* inpoel, int64_t ntri,
                  const int64_t* triinpoel, int nsideset, const int* ss_id,
                  const int64_t* ss_off, const int64_t* ss_face, int nnodevar,
                  const char* const* names, xg_exo** out) {
  if (!out) return XG_EINVAL;
  *out = nullptr;
  xg_exo* h = new (std::nothrow) xg_exo;
  if (!h) return XG_EINVAL;
  *out = h;
  if (!path) return exo_fail(h, XG_EINVAL, "exodus: null path");
  h->path = path;
  if (npoin <= 0 || npoin > INT32_MAX || ntet < 0 || ntri < 0 || nsideset < 0 ||
      nnodevar < 0 || !x || !y || !z || (ntet && !inpoel) || (ntri && !triinpoel) ||
      (nsideset && (!ss_id || !ss_off || !ss_face)) || (nnodevar && !names))
    return exo_fail(h, XG_EINVAL, "exodus: bad arguments to create");
  if (ntet + ntri > INT32_MAX) return exo_fail(h, XG_EINVAL, "exodus: too many elements");
  for (int64_t i = 0; i < 4 * ntet; ++i)
    if (inpoel[i] < 0 || inpoel[i] >= npoin)
      return exo_fail(h, XG_EINVAL, "exodus: tetrahedron node id out of range");
  for (int64_t i = 0; i < 3 * ntri; ++i)
    if (triinpoel[i] < 0 || triinpoel[i] >= npoin)
      return exo_fail(h, XG_EINVAL, "exodus: triangle node id out of range");
  for (int s = 0; s < nsideset; ++s) {
    if (ss_off[s + 1] < ss_off[s]) return exo_fail(h, XG_EINVAL, "exodus: side set offsets");
    for (int64_t k = ss_off[s]; k < ss_off[s + 1]; ++k)
      if (ss_face[k] < 0 || ss_face[k] >= ntri)
        return exo_fail(h, XG_EINVAL, "exodus: side set face out of range");
  }
  // dimensions (in definition order)
  std::vector<NcDim> D;
  auto add_dim = [&](const std::string& n, int64_t len) { D.push_back({n, len}); };
  const int nblk = (ntri ? 1 : 0) + (ntet ? 1 : 0);
  add_dim("len_string", 33);
  add_dim("len_line", 81);
  add_dim("four", 4);
  add_dim("len_name", 33);
  add_dim("time_step", 0);
  add_dim("num_dim", 3);
  add_dim("num_nodes", npoin);
  add_dim("num_elem", std::max<int64_t>(1, ntet + ntri));
  add_dim("num_el_blk", std::max(1, nblk));
  if (nsideset) add_dim("num_side_sets", nsideset);
  int b = 0;
  if (ntri) {
    ++b;
    add_dim("num_el_in_blk" + std::to_string(b), ntri);
    add_dim("num_nod_per_el" + std::to_string(b), 3);
  }
  if (ntet) {
    ++b;
    add_dim("num_el_in_blk" + std::to_string(b), ntet);
    add_dim("num_nod_per_el" + std::to_string(b), 4);
  }
  for (int s = 0; s < nsideset; ++s)
    if (ss_off[s + 1] > ss_off[s])
      add_dim("num_side_ss" + std::to_string(s + 1), ss_off[s + 1] - ss_off[s]);
  if (nnodevar) add_dim("num_nod_var", nnodevar);
  std::map<std::string, int> did;
  for (size_t i = 0; i < D.size(); ++i) did[D[i].name] = (int)i;
  // variables
  std::vector<WVar> V;
  V.push_back({"time_whole", {"time_step"}, NC_DOUBLE, {}});
  V.push_back({"eb_status", {"num_el_blk"}, NC_INT, {}});
  V.push_back({"eb_prop1", {"num_el_blk"}, NC_INT, {{"name", "ID"}}});
  if (nsideset) {
    V.push_back({"ss_status", {"num_side_sets"}, NC_INT, {}});
    V.push_back({"ss_prop1", {"num_side_sets"}, NC_INT, {{"name", "ID"}}});
  }
  V.push_back({"coordx", {"num_nodes"}, NC_DOUBLE, {}});
  V.push_back({"coordy", {"num_nodes"}, NC_DOUBLE, {}});
  V.push_back({"coordz", {"num_nodes"}, NC_DOUBLE, {}});
  V.push_back({"eb_names", {"num_el_blk", "len_name"}, NC_CHAR, {}});
  V.push_back({"coor_names", {"num_dim", "len_name"}, NC_CHAR, {}});
  b = 0;
  if (ntri) {
    ++b;
    const std::string s = std::to_string(b);
    V.push_back({"connect" + s, {"num_el_in_blk" + s, "num_nod_per_el" + s}, NC_INT,
                 {{"elem_type", "TRI3"}}});
  }
  if (ntet) {
    ++b;
    const std::string s = std::to_string(b);
    V.push_back({"connect" + s, {"num_el_in_blk" + s, "num_nod_per_el" + s}, NC_INT,
                 {{"elem_type", "TETRA4"}}});
  }
  for (int s = 0; s < nsideset; ++s)
    if (ss_off[s + 1] > ss_off[s]) {
      const std::string k = std::to_string(s + 1);
      V.push_back({"elem_ss" + k, {"num_side_ss" + k}, NC_INT, {}});
      V.push_back({"side_ss" + k, {"num_side_ss" + k}, NC_INT, {}});
    }
  if (nnodevar) {
    V.push_back({"name_nod_var", {"num_nod_var", "len_name"}, NC_CHAR, {}});
    for (int v = 0; v < nnodevar; ++v)
      V.push_back({"vals_nod_var" + std::to_string(v + 1), {"time_step", "num_nodes"},
                   NC_DOUBLE, {}});
  }
  // header with sizes and offsets (CDF-2: 64-bit offsets)
  h->version = 2;
  h->dims = D;
  auto header = [&](const std::vector<int64_t>& begin) {
    std::vector<uint8_t> hb = {'C', 'D', 'F', 2};
    put32(hb, 0);
    put32(hb, TAG_DIM);
    put32(hb, (uint32_t)D.size());
    for (auto& d : D) {
      put_name(hb, d.name);
      put32(hb, (uint32_t)d.len);
    }
    // global attributes as ex_put_init leaves them
    struct G { const char* n; int t; double v; const char* s; };
    const G ga[] = {{"api_version", NC_FLOAT, 8.21, nullptr},
                    {"version", NC_FLOAT, 8.21, nullptr},
                    {"floating_point_word_size", NC_INT, 8, nullptr},
                    {"file_size", NC_INT, 1, nullptr},
                    {"maximum_name_length", NC_INT, 32, nullptr},
                    {"int64_status", NC_INT, 0, nullptr},
                    {"title", NC_CHAR, 0, "Written by xyst_amd"}};
    put32(hb, TAG_ATT);
    put32(hb, (uint32_t)(sizeof(ga) / sizeof(ga[0])));
    for (auto& a : ga) {
      put_name(hb, a.n);
      put32(hb, (uint32_t)a.t);
      if (a.t == NC_CHAR) {
        put_name(hb, a.s);
      } else {
        put32(hb, 1);
        if (a.t == NC_FLOAT) {
          float f = (float)a.v;
          uint32_t u;
          std::memcpy(&u, &f, 4);
          put32(hb, u);
        } else {
          put32(hb, (uint32_t)(int32_t)a.v);
        }
      }
    }
    put32(hb, TAG_VAR);
    put32(hb, (uint32_t)V.size());
    for (size_t i = 0; i < V.size(); ++i) {
      const WVar& v = V[i];
      put_name(hb, v.name);
      put32(hb, (uint32_t)v.dims.size());
      for (auto& dn : v.dims) put32(hb, (uint32_t)did.at(dn));
      if (v.text_atts.empty()) {
        put32(hb, 0);
        put32(hb, 0);
      } else {
        put32(hb, TAG_ATT);
        put32(hb, (uint32_t)v.text_atts.size());
        for (auto& a : v.text_atts) {
          put_name(hb, a.first);
          put32(hb, NC_CHAR);
          put_name(hb, a.second);
        }
      }
      put32(hb, (uint32_t)v.type);
      int64_t cnt = 1;
      for (auto& dn : v.dims)
        if (D[did.at(dn)].len) cnt *= D[did.at(dn)].len;
      put32(hb, (uint32_t)std::min<int64_t>(pad4(cnt * type_size(v.type)), 0xFFFFFFFFll));
      put64(hb, (uint64_t)begin[i]);
    }
    return hb;
  };
  std::vector<int64_t> begin(V.size(), 0);
  const size_t hsize = header(begin).size();
  int64_t off = (int64_t)hsize;
  for (size_t i = 0; i < V.size(); ++i)   // fixed variables first
    if (V[i].dims.empty() || V[i].dims[0] != "time_step") {
      begin[i] = off;
      int64_t cnt = 1;
      for (auto& dn : V[i].dims) cnt *= D[did.at(dn)].len;
      off += pad4(cnt * type_size(V[i].type));
    }
  for (size_t i = 0; i < V.size(); ++i)   // then the record variables
    if (!V[i].dims.empty() && V[i].dims[0] == "time_step") {
      begin[i] = off;
      int64_t cnt = 1;
      for (size_t k = 1; k < V[i].dims.size(); ++k) cnt *= D[did.at(V[i].dims[k])].len;
      off += pad4(cnt * type_size(V[i].type));
    }
  const std::vector<uint8_t> hb = header(begin);
  h->fp = std::fopen(path, "w+b");
  if (!h->fp) return exo_fail(h, XG_EIO, "exodus: cannot create file");
  h->writable = true;
  // fixed data, in variable order
  std::vector<uint8_t> body;
  auto ints = [&](const std::vector<int64_t>& a) {
    for (int64_t v : a) put32(body, (uint32_t)(int32_t)v);
  };
  auto doubles = [&](const double* a, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
      uint64_t u;
      std::memcpy(&u, &a[i], 8);
      put64(body, u);
    }
  };
  auto chars = [&](const std::vector<std::string>& rows, int64_t w) {
    for (auto& r : rows) {
      std::string s = r.substr(0, (size_t)w - 1);
      s.resize((size_t)w, '\0');
      body.insert(body.end(), s.begin(), s.end());
    }
  };
  std::vector<int64_t> tmp;
  for (size_t i = 0; i < V.size(); ++i) {
    const WVar& v = V[i];
    if (!v.dims.empty() && v.dims[0] == "time_step") continue;
    const size_t start = body.size();
    if (v.name == "eb_status") {
      ints(std::vector<int64_t>(std::max(1, nblk), 1));
    } else if (v.name == "eb_prop1") {
      tmp.clear();
      for (int k = 0; k < std::max(1, nblk); ++k) tmp.push_back(k + 1);
      ints(tmp);
    } else if (v.name == "ss_status") {
      tmp.clear();
      for (int s = 0; s < nsideset; ++s) tmp.push_back(ss_off[s + 1] > ss_off[s] ? 1 : 0);
      ints(tmp);
    } else if (v.name == "ss_prop1") {
      tmp.assign(ss_id, ss_id + nsideset);
      ints(tmp);
    } else if (v.name == "coordx") {
      doubles(x, npoin);
    } else if (v.name == "coordy") {
      doubles(y, npoin);
    } else if (v.name == "coordz") {
      doubles(z, npoin);
    } else if (v.name == "eb_names") {
      chars(std::vector<std::string>(std::max(1, nblk), std::string()), 33);
    } else if (v.name == "coor_names") {
      chars({"x", "y", "z"}, 33);
    } else if (v.name.compare(0, 7, "connect") == 0) {
      const bool tri = v.text_atts[0].second == "TRI3";
      const int64_t n = tri ? 3 * ntri : 4 * ntet;
      const int64_t* c = tri ? triinpoel : inpoel;
      tmp.assign(c, c + n);
      for (auto& q : tmp) q += 1;
      ints(tmp);
    } else if (v.name.compare(0, 7, "elem_ss") == 0 || v.name.compare(0, 7, "side_ss") == 0) {
      const int s = std::stoi(v.name.substr(7)) - 1;
      tmp.clear();
      for (int64_t k = ss_off[s]; k < ss_off[s + 1]; ++k)
        tmp.push_back(v.name[0] == 'e' ? ss_face[k] + 1 : 1);   // triangle elements, side 1
      ints(tmp);
    } else if (v.name == "name_nod_var") {
      std::vector<std::string> rows;
      for (int k = 0; k < nnodevar; ++k) rows.push_back(names[k] ? names[k] : "");
      chars(rows, 33);
    }
    while (body.size() % 4) body.push_back(0);
    if ((int64_t)(hsize + start) != begin[i])
      return exo_fail(h, XG_EINVAL, "exodus: internal layout error at " + v.name);
  }
  if (std::fwrite(hb.data(), 1, hb.size(), h->fp) != hb.size() ||
      std::fwrite(body.data(), 1, body.size(), h->fp) != body.size() ||
      std::fflush(h->fp) != 0)
    return exo_fail(h, XG_EIO, "exodus: write error");
  std::fclose(h->fp);
  h->fp = std::fopen(path, "r+b");
  if (!h->fp) return exo_fail(h, XG_EIO, "exodus: cannot reopen file");
  return read_header(h);
}

// writes the record slab of one variable at record `step`, appending empty
// (zero) records up to it; updates the record count in the header
static int put_record(xg_exo* h, const NcVar& v, int64_t step, const std::vector<uint8_t>& raw) {
  if (!h->writable) return exo_fail(h, XG_EINVAL, "exodus: file not opened for writing");
  if (step < 0 || step > h->numrecs)
    return exo_fail(h, XG_EINVAL, "exodus: steps must be written in order (step <= count)");
  if (step == h->numrecs) {   // a new record, zero-filled
    int64_t end0 = -1;
    for (auto& w : h->vars)
      if (w.rec) end0 = end0 < 0 ? w.begin : std::min(end0, w.begin);
    std::vector<uint8_t> zero(h->recsize, 0);
    if (std::fseek(h->fp, (long)(end0 + step * h->recsize), SEEK_SET) != 0 ||
        std::fwrite(zero.data(), 1, zero.size(), h->fp) != zero.size())
      return exo_fail(h, XG_EIO, "exodus: write error");
    h->numrecs = step + 1;
    std::vector<uint8_t> nr;
    put32(nr, (uint32_t)h->numrecs);
    if (std::fseek(h->fp, 4, SEEK_SET) != 0 || std::fwrite(nr.data(), 1, 4, h->fp) != 4)
      return exo_fail(h, XG_EIO, "exodus: write error");
  }
  if (std::fseek(h->fp, (long)(v.begin + step * h->recsize), SEEK_SET) != 0 ||
      std::fwrite(raw.data(), 1, raw.size(), h->fp) != raw.size() || std::fflush(h->fp) != 0)
    return exo_fail(h, XG_EIO, "exodus: write error");
  return XG_OK;
}

int xg_exo_write_time(xg_exo* h, int64_t step, double time) {
  if (!h) return XG_EINVAL;
  const NcVar* v = find_var(h, "time_whole");
  if (!v || !v->rec || v->type != NC_DOUBLE) return exo_fail(h, XG_EIO, "exodus: no time_whole");
  std::vector<uint8_t> raw;
  uint64_t u;
  std::memcpy(&u, &time, 8);
  put64(raw, u);
  return put_record(h, *v, step, raw);
}

int xg_exo_write_node_var(xg_exo* h, int64_t step, int var, const double* vals) {
  if (!h || !vals || var < 0) return XG_EINVAL;
  const NcVar* v = find_var(h, "vals_nod_var" + std::to_string(var + 1));
  if (!v || !v->rec || v->type != NC_DOUBLE)
    return exo_fail(h, XG_EINVAL, "exodus: node variable out of range");
  std::vector<uint8_t> raw;
  raw.reserve(v->count * 8);
  for (int64_t i = 0; i < v->count; ++i) {
    uint64_t u;
    std::memcpy(&u, &vals[i], 8);
    put64(raw, u);
  }
  return put_record(h, *v, step, raw);
}

}  // extern "C"
