// Field files and result-file names of the reference (data_ops.f90,
// function_ops.f90:4590-4674, initialize.f90:49-70), host code only.
//
// Field files are NetCDF classic files written and read by this file's own
// code (no NetCDF library): dimensions x, y, z; double variables Ux, Uy, Uz
// with C-order dimensions (z, y, x), so x varies fastest as in the
// reference's Fortran arrays (save_field_R3toR3_ncdf, data_ops.f90:390-447).
// The classic format (CDF-1) is written while every offset fits 32 bits, the
// 64-bit-offset format (CDF-2) beyond; when one variable exceeds the 4 GiB a
// fixed-size CDF-2 variable may hold (1024^3), z is the record (unlimited)
// dimension, as the reference's parallel writer defines it.  The reader takes
// either format and either layout (read_field_R3toR3_ncdf, :1198-1397).
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/maxlq.h"

namespace {

constexpr uint32_t NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12, NC_DOUBLE = 6;
const char* const VARS[3] = {"Ux", "Uy", "Uz"};

struct Out {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void name(const char* s) {
    const uint32_t n = (uint32_t)std::strlen(s);
    u32(n);
    for (uint32_t i = 0; i < n; ++i) b.push_back((unsigned char)s[i]);
    while (b.size() % 4) b.push_back(0);
  }
};

// big-endian copies of doubles (the format's byte order)
void put_be(const double* src, size_t n, unsigned char* dst) {
  for (size_t i = 0; i < n; ++i) {
    uint64_t v;
    std::memcpy(&v, src + i, 8);
    for (int k = 0; k < 8; ++k) dst[8 * i + k] = (unsigned char)(v >> (56 - 8 * k));
  }
}
void get_be(const unsigned char* src, size_t n, double* dst) {
  for (size_t i = 0; i < n; ++i) {
    uint64_t v = 0;
    for (int k = 0; k < 8; ++k) v = (v << 8) | src[8 * i + k];
    std::memcpy(dst + i, &v, 8);
  }
}

struct FileCloser {
  void operator()(FILE* f) const { if (f) std::fclose(f); }
};
using File = std::unique_ptr<FILE, FileCloser>;

// header of the three-variable file; begin offsets are patched in by the caller
std::vector<unsigned char> header(int version, bool rec, uint32_t nx, uint32_t ny, uint32_t nz,
                                  const uint64_t begin[3], size_t* begin_pos) {
  Out o;
  o.b = {'C', 'D', 'F', (unsigned char)version};
  o.u32(rec ? nz : 0);                       // numrecs
  o.u32(NC_DIMENSION);
  o.u32(3);
  o.name("x"); o.u32(nx);
  o.name("y"); o.u32(ny);
  o.name("z"); o.u32(rec ? 0 : nz);          // 0: the record dimension
  o.u32(0); o.u32(0);                        // no global attributes
  o.u32(NC_VARIABLE);
  o.u32(3);
  const uint64_t vbytes = (uint64_t)nx * ny * (rec ? 1 : nz) * 8;
  for (int v = 0; v < 3; ++v) {
    o.name(VARS[v]);
    o.u32(3);
    o.u32(2); o.u32(1); o.u32(0);            // dims (z, y, x): x fastest
    o.u32(0); o.u32(0);                      // no attributes
    o.u32(NC_DOUBLE);
    o.u32(vbytes >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)vbytes);
    begin_pos[v] = o.b.size();
    if (version == 1) o.u32((uint32_t)begin[v]);
    else o.u64(begin[v]);
  }
  return o.b;
}

struct In {
  const unsigned char* p;
  size_t n, i = 0;
  bool ok = true;
  uint32_t u32() {
    if (i + 4 > n) { ok = false; return 0; }
    uint32_t v = ((uint32_t)p[i] << 24) | ((uint32_t)p[i + 1] << 16) | ((uint32_t)p[i + 2] << 8) | p[i + 3];
    i += 4;
    return v;
  }
  uint64_t u64() {
    const uint64_t hi = u32();
    return (hi << 32) | u32();
  }
  std::string name() {
    const uint32_t len = u32();
    if (!ok || i + len > n) { ok = false; return ""; }
    std::string s((const char*)p + i, len);
    i += (len + 3) / 4 * 4;
    return s;
  }
  void skip_atts() {   // attribute list: ABSENT or a list of (name, type, values)
    const uint32_t tag = u32(), cnt = u32();
    if (tag == 0) return;
    if (tag != NC_ATTRIBUTE) { ok = false; return; }
    static const int width[7] = {0, 1, 1, 2, 4, 4, 8};
    for (uint32_t a = 0; a < cnt && ok; ++a) {
      name();
      const uint32_t ty = u32(), nel = u32();
      if (ty < 1 || ty > 6) { ok = false; return; }
      i += ((size_t)nel * width[ty] + 3) / 4 * 4;
    }
  }
};

std::string qtext(double q) {   // setLebesgueQtext, initialize.f90:49-70
  char buf[64];
  const double qi = std::nearbyint(q);
  if (std::fabs(qi - q) > 2.0e-16 || q >= 100.0 - 2.0e-16) std::snprintf(buf, sizeof buf, "%.2E", q);
  else std::snprintf(buf, sizeof buf, "%d", (int)qi);
  return buf;
}

}  // namespace

extern "C" {

int mlq_save_field_nc(const char* path, int n1, int n2, int n3, const double* field) {
  if (!path || !field || n1 < 1 || n2 < 1 || n3 < 1) return MLQ_EINVAL;
  const uint64_t plane = (uint64_t)n1 * n2, vbytes = plane * n3 * 8;
  // fixed-size variables while each holds < 4 GiB, else z as the record
  // dimension (MLQ_NC_LAYOUT=record / fixed selects one for any size)
  const char* lay = std::getenv("MLQ_NC_LAYOUT");
  bool rec = vbytes >= 0xFFFFFFFCull;
  if (lay && std::strcmp(lay, "record") == 0) rec = true;
  if (lay && std::strcmp(lay, "fixed") == 0 && vbytes < 0xFFFFFFFCull) rec = false;
  uint64_t begin[3] = {0, 0, 0};
  size_t bpos[3];
  int version = 1;
  std::vector<unsigned char> h = header(2, rec, n1, n2, n3, begin, bpos);   // CDF-2 header size
  const uint64_t h64 = h.size(), h32 = h64 - 3 * 4;                         // CDF-1: 4-byte offsets
  if (h32 + 3 * vbytes < 0x7FFFFFFFull && !rec) version = 1;
  else version = 2;
  const char* cdf = std::getenv("MLQ_NC_CDF");   // "2": the 64-bit-offset format for any size
  if (cdf && cdf[0] == '2') version = 2;
  const uint64_t hsize = version == 1 ? h32 : h64;
  for (int v = 0; v < 3; ++v) begin[v] = hsize + (rec ? plane * 8 : vbytes) * v;
  h = header(version, rec, n1, n2, n3, begin, bpos);
  File f(std::fopen(path, "wb"));
  if (!f) return MLQ_EINVAL;
  if (std::fwrite(h.data(), 1, h.size(), f.get()) != h.size()) return MLQ_EINVAL;
  std::vector<unsigned char> buf(plane * 8);
  if (!rec) {
    for (int v = 0; v < 3; ++v)
      for (int z = 0; z < n3; ++z) {
        put_be(field + ((uint64_t)v * n3 + z) * plane, plane, buf.data());
        if (std::fwrite(buf.data(), 1, buf.size(), f.get()) != buf.size()) return MLQ_EINVAL;
      }
  } else {
    for (int z = 0; z < n3; ++z)          // record z: Ux, Uy, Uz planes
      for (int v = 0; v < 3; ++v) {
        put_be(field + ((uint64_t)v * n3 + z) * plane, plane, buf.data());
        if (std::fwrite(buf.data(), 1, buf.size(), f.get()) != buf.size()) return MLQ_EINVAL;
      }
  }
  return std::fflush(f.get()) == 0 ? MLQ_OK : MLQ_EINVAL;
}

int mlq_read_field_nc(const char* path, int n1, int n2, int n3, double* field) {
  if (!path || !field) return MLQ_EINVAL;
  File f(std::fopen(path, "rb"));
  if (!f) return MLQ_EINVAL;
  std::vector<unsigned char> h(1 << 16);
  const size_t got = std::fread(h.data(), 1, h.size(), f.get());
  In in{h.data(), got};
  if (got < 8 || h[0] != 'C' || h[1] != 'D' || h[2] != 'F' || (h[3] != 1 && h[3] != 2))
    return MLQ_EINVAL;
  const int version = h[3];
  in.i = 4;
  uint32_t numrecs = in.u32();
  if (in.u32() != NC_DIMENSION) return MLQ_EINVAL;
  const uint32_t ndim = in.u32();
  std::vector<std::string> dn(ndim);
  std::vector<uint32_t> dl(ndim);
  int recdim = -1;
  for (uint32_t d = 0; d < ndim && in.ok; ++d) {
    dn[d] = in.name();
    dl[d] = in.u32();
    if (dl[d] == 0) recdim = (int)d;
  }
  in.skip_atts();
  const uint32_t vtag = in.u32(), nvar = in.u32();
  if (!in.ok || vtag != NC_VARIABLE) return MLQ_EINVAL;
  struct Var { std::string name; std::vector<uint32_t> dims; uint32_t type; uint64_t begin; };
  std::vector<Var> vars(nvar);
  uint64_t recsize = 0;
  for (uint32_t v = 0; v < nvar && in.ok; ++v) {
    vars[v].name = in.name();
    const uint32_t nd = in.u32();
    for (uint32_t k = 0; k < nd && in.ok; ++k) vars[v].dims.push_back(in.u32());
    in.skip_atts();
    vars[v].type = in.u32();
    const uint32_t vsize = in.u32();
    vars[v].begin = version == 1 ? in.u32() : in.u64();
    if (!vars[v].dims.empty() && (int)vars[v].dims[0] == recdim) recsize += vsize;
  }
  if (!in.ok) return MLQ_EINVAL;
  const uint64_t plane = (uint64_t)n1 * n2;
  std::vector<unsigned char> buf(plane * 8);
  for (int c = 0; c < 3; ++c) {
    const Var* var = nullptr;
    for (auto& v : vars) if (v.name == VARS[c]) var = &v;
    if (!var || var->type != NC_DOUBLE || var->dims.size() != 3) return MLQ_EINVAL;
    // dims (z, y, x) with the grid's sizes (z may be the record dimension)
    const uint32_t dz = var->dims[0], dy = var->dims[1], dx = var->dims[2];
    if (dz >= ndim || dy >= ndim || dx >= ndim) return MLQ_EINVAL;
    if (dn[dx] != "x" || dn[dy] != "y" || dn[dz] != "z") return MLQ_EINVAL;
    const bool rec = (int)dz == recdim;
    const uint64_t nz = rec ? numrecs : dl[dz];
    if (dl[dx] != (uint32_t)n1 || dl[dy] != (uint32_t)n2 || nz != (uint64_t)n3) return MLQ_EINVAL;
    for (int z = 0; z < n3; ++z) {
      const uint64_t off = var->begin + (rec ? (uint64_t)z * recsize : (uint64_t)z * plane * 8);
      if (fseeko(f.get(), (off_t)off, SEEK_SET) != 0) return MLQ_EINVAL;
      if (std::fread(buf.data(), 1, buf.size(), f.get()) != buf.size()) return MLQ_EINVAL;
      get_be(buf.data(), plane, field + ((uint64_t)c * n3 + z) * plane);
    }
  }
  return MLQ_OK;
}

int mlq_result_filename(char* buf, int buflen, const char* dir, const char* field_name, double q,
                        int n, int b_iter, int iter) {
  if (!buf || buflen < 1 || !field_name || n < 1 || b_iter < 0 || b_iter > 999 || iter > 99999)
    return MLQ_EINVAL;
  char tail[96];
  if (iter < 0) std::snprintf(tail, sizeof tail, "_B%03d_iterend.nc", b_iter);
  else std::snprintf(tail, sizeof tail, "_B%03d_iter%05d.nc", b_iter, iter);
  const std::string s = std::string(dir ? dir : "") + field_name + "_q" + qtext(q) + "_n" +
                        std::to_string(n) + tail;
  if ((int)s.size() + 1 > buflen) return MLQ_EINVAL;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return MLQ_OK;
}

int mlq_q_text(double q, char* buf, int buflen) {
  const std::string s = qtext(q);
  if (!buf || (int)s.size() + 1 > buflen) return MLQ_EINVAL;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return MLQ_OK;
}

int mlq_parse_result_filename(const char* name, double* q, int* resol, int* b_iter_offset,
                              int* optim_iter_offset) {
  if (!name || !q || !resol || !b_iter_offset || !optim_iter_offset) return MLQ_EINVAL;
  const std::string s(name);
  // the text after the first occurrence of `key` up to the next `stop`
  auto field = [&](const char* key, const char* stop, std::string& out) {
    const size_t a = s.find(key);
    if (a == std::string::npos) return false;
    const std::string t = s.substr(a + std::strlen(key));
    const size_t b = t.find(stop);
    if (b == std::string::npos) return false;
    out = t.substr(0, b);
    return true;
  };
  std::string tq, tn, tb, ti;
  if (!field("_q", "_", tq) || !field("_n", "_", tn) || !field("_B", "_", tb) ||
      !field("_iter", ".nc", ti))
    return MLQ_EINVAL;
  char* e = nullptr;
  errno = 0;
  const double qv = std::strtod(tq.c_str(), &e);
  if (e == tq.c_str() || errno) return MLQ_EINVAL;
  const long nv = std::strtol(tn.c_str(), &e, 10);
  if (e == tn.c_str()) return MLQ_EINVAL;
  const double bv = std::strtod(tb.c_str(), &e);   // read as a real, then assigned to an integer
  if (e == tb.c_str()) return MLQ_EINVAL;
  int boff = (int)bv, ioff = 0;
  if (ti.find("end") == std::string::npos) {
    boff -= 1;                                       // resume inside B solve number boff + 1
    ioff = (int)std::strtol(ti.c_str(), &e, 10);
    if (e == ti.c_str()) return MLQ_EINVAL;
  }
  *q = qv;
  *resol = (int)nv;
  *b_iter_offset = boff;
  *optim_iter_offset = ioff;
  return MLQ_OK;
}

}  // extern "C"
