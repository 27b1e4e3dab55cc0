// Flow files in the MAT 7.3 layout (HDF5 behind a 512-byte MATLAB userblock), the
// format of source/Fortran/readWriteMat.F90 (matOpen 'w7.3', matPutVariable of t and
// U,V,W,R,E), with the NaN convention of exportFlow.F90 / importFlow.F90:24-40.
// Host code only; built against libhdf5 into its own library (libgatt_flowio.so).
#include <hdf5.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <limits>
#include <string>
#include <vector>

#include "../../include/gatt_dns.h"
#include "../../include/gatt_flow_io.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

const char* const kNames[5] = {"U", "V", "W", "R", "E"};

// dataset dimensions of an nx x ny x nz MATLAB array: HDF5 lists them slowest
// first, and MATLAB drops a trailing singleton z
int dims_of(int32_t nx, int32_t ny, int32_t nz, hsize_t d[3]) {
  if (nz > 1) {
    d[0] = (hsize_t)nz;
    d[1] = (hsize_t)ny;
    d[2] = (hsize_t)nx;
    return 3;
  }
  d[0] = (hsize_t)ny;
  d[1] = (hsize_t)nx;
  return 2;
}

bool put_class(hid_t obj) {
  // MATLAB_class: fixed-length ASCII string of 6 characters, null-terminated padding
  hid_t st = H5Tcopy(H5T_C_S1);
  H5Tset_size(st, 6);
  H5Tset_strpad(st, H5T_STR_NULLTERM);
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t at = H5Acreate2(obj, "MATLAB_class", st, sp, H5P_DEFAULT, H5P_DEFAULT);
  bool ok = at >= 0 && H5Awrite(at, st, "double") >= 0;
  if (at >= 0) H5Aclose(at);
  H5Sclose(sp);
  H5Tclose(st);
  return ok;
}

bool put_var(hid_t file, const char* name, int rank, const hsize_t* dims, const double* data) {
  hid_t sp = H5Screate_simple(rank, dims, nullptr);
  hid_t ds = H5Dcreate2(file, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  bool ok = ds >= 0 && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0 && put_class(ds);
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  return ok;
}

// the 128-byte MAT-file header at the start of the userblock: 116 bytes of text,
// 8 bytes of subsystem offset, version 0x0200 and the endian indicator "IM"
bool put_header(const char* path) {
  char hdr[128];
  std::memset(hdr, ' ', sizeof(hdr));
  char date[64] = "";
  const std::time_t now = std::time(nullptr);
  std::strftime(date, sizeof(date), "%a %b %d %H:%M:%S %Y", std::gmtime(&now));
  char text[160];
  std::snprintf(text, sizeof(text), "MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: %s HDF5 schema 1.00 .", date);
  std::memcpy(hdr, text, std::min<size_t>(std::strlen(text), 116));
  std::memset(hdr + 116, 0, 8);
  hdr[124] = 0x00;
  hdr[125] = 0x02;
  hdr[126] = 'I';
  hdr[127] = 'M';
  std::FILE* f = std::fopen(path, "r+b");
  if (!f) return false;
  const bool ok = std::fwrite(hdr, 1, sizeof(hdr), f) == sizeof(hdr);
  return std::fclose(f) == 0 && ok;
}

bool get_var(hid_t file, const char* name, int32_t nx, int32_t ny, int32_t nz, double* out, bool* present) {
  *present = H5Lexists(file, name, H5P_DEFAULT) > 0;
  if (!*present) return true;
  hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
  if (ds < 0) return false;
  hid_t sp = H5Dget_space(ds);
  hsize_t d[8] = {0};
  const int rank = H5Sget_simple_extent_ndims(sp);
  bool ok = rank > 0 && rank <= 8 && H5Sget_simple_extent_dims(sp, d, nullptr) == rank;
  if (ok) {
    // accept the MATLAB shapes (nz,ny,nx) / (ny,nx) and singleton dimensions around them
    std::vector<hsize_t> keep;
    for (int i = 0; i < rank; ++i)
      if (d[i] != 1) keep.push_back(d[i]);
    std::vector<hsize_t> want;
    if (nz > 1) want.push_back((hsize_t)nz);
    if (ny > 1) want.push_back((hsize_t)ny);
    if (nx > 1) want.push_back((hsize_t)nx);
    ok = keep == want;
  }
  if (ok) ok = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0;
  H5Sclose(sp);
  H5Dclose(ds);
  return ok;
}

// one MAT 7.3 file of named nx*ny*nz arrays (optionally preceded by the 1x1 t), each
// field NaN where inside_wall is set (exportFlow.F90 / exportMeanFlow.F90)
int write_vars(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* t, int nvars,
               const char* const* names, const double* const* q, const uint8_t* inside_wall) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const size_t np = (size_t)nx * ny * nz;
  hsize_t d[3];
  const int rank = dims_of(nx, ny, nz, d);
  hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
  H5Pset_userblock(fcpl, 512);
  hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT);
  H5Pclose(fcpl);
  if (file < 0) return fail(GATT_E_INVALID, std::string("cannot create ") + path);
  const hsize_t d1[2] = {1, 1};
  bool ok = !t || put_var(file, "t", 2, d1, t);
  std::vector<double> tmp;
  for (int f = 0; ok && f < nvars; ++f) {
    const double* src = q[f];
    if (inside_wall) {
      tmp.assign(src, src + np);
      for (size_t n = 0; n < np; ++n)
        if (inside_wall[n]) tmp[n] = std::numeric_limits<double>::quiet_NaN();
      src = tmp.data();
    }
    ok = put_var(file, names[f], rank, d, src);
  }
  if (H5Fclose(file) < 0) ok = false;
  if (!ok) return fail(GATT_E_INVALID, std::string("writing ") + path + " failed");
  if (!put_header(path)) return fail(GATT_E_INVALID, std::string("cannot write the MAT header of ") + path);
  return GATT_OK;
}

// a double or complex double array of MATLAB size dims[0] x dims[1] x ... (column-major data);
// complex values are the compound {real, imag} MATLAB stores in a 7.3 file
bool put_array(hid_t file, const char* name, int rank, const int64_t* dims, const double* re, const double* im) {
  std::vector<hsize_t> d(rank);
  size_t n = 1;
  for (int i = 0; i < rank; ++i) {
    d[rank - 1 - i] = (hsize_t)dims[i];
    n *= (size_t)dims[i];
  }
  if (!im) return put_var(file, name, rank, d.data(), re);
  hid_t ct = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
  H5Tinsert(ct, "real", 0, H5T_NATIVE_DOUBLE);
  H5Tinsert(ct, "imag", sizeof(double), H5T_NATIVE_DOUBLE);
  hid_t ft = H5Tcreate(H5T_COMPOUND, 16);
  H5Tinsert(ft, "real", 0, H5T_IEEE_F64LE);
  H5Tinsert(ft, "imag", 8, H5T_IEEE_F64LE);
  std::vector<double> buf(2 * n);
  for (size_t k = 0; k < n; ++k) {
    buf[2 * k] = re[k];
    buf[2 * k + 1] = im[k];
  }
  hid_t sp = H5Screate_simple(rank, d.data(), nullptr);
  hid_t ds = H5Dcreate2(file, name, ft, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  bool ok = ds >= 0 && H5Dwrite(ds, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) >= 0 && put_class(ds);
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  H5Tclose(ft);
  H5Tclose(ct);
  return ok;
}

// ---- MAT v5 / v7 (MATLAB's default save format; v7 compresses each variable with zlib) ----
// Level 5 MAT-file layout: a 128-byte header, then data elements of an 8-byte tag
// (type, byte count; a "small" element packs both into 4 bytes and its data into the
// next 4) padded to 8 bytes.  A variable is a miMATRIX element holding array flags,
// dimensions, name and the real part, whose numeric type may be narrower than double.

bool mat5_tag(const uint8_t* p, size_t left, uint32_t* type, uint32_t* nbytes, const uint8_t** data, size_t* adv) {
  if (left < 8) return false;
  uint32_t t, n;
  std::memcpy(&t, p, 4);
  std::memcpy(&n, p + 4, 4);
  if (t >> 16) {            // small data element
    *type = t & 0xffff;
    *nbytes = t >> 16;
    *data = p + 4;
    *adv = 8;
    return *nbytes <= 4;
  }
  *type = t;
  *nbytes = n;
  *data = p + 8;
  *adv = 8 + ((size_t)n + 7) / 8 * 8;
  if (t == 15) *adv = 8 + (size_t)n;     // miCOMPRESSED is not padded
  // padding past the end of a truncated buffer: the caller's next step lands on the end
  if (*adv > left) *adv = left;
  return (size_t)n <= left - 8;
}

bool mat5_number(uint32_t type, const uint8_t* p, size_t i, double* v) {
  switch (type) {
    case 1: { int8_t x; std::memcpy(&x, p + i, 1); *v = x; return true; }
    case 2: { uint8_t x; std::memcpy(&x, p + i, 1); *v = x; return true; }
    case 3: { int16_t x; std::memcpy(&x, p + 2 * i, 2); *v = x; return true; }
    case 4: { uint16_t x; std::memcpy(&x, p + 2 * i, 2); *v = x; return true; }
    case 5: { int32_t x; std::memcpy(&x, p + 4 * i, 4); *v = x; return true; }
    case 6: { uint32_t x; std::memcpy(&x, p + 4 * i, 4); *v = x; return true; }
    case 7: { float x; std::memcpy(&x, p + 4 * i, 4); *v = x; return true; }
    case 9: { std::memcpy(v, p + 8 * i, 8); return true; }
    case 12: { int64_t x; std::memcpy(&x, p + 8 * i, 8); *v = (double)x; return true; }
    case 13: { uint64_t x; std::memcpy(&x, p + 8 * i, 8); *v = (double)x; return true; }
    default: return false;
  }
}

size_t mat5_width(uint32_t type) {
  switch (type) {
    case 1: case 2: return 1;
    case 3: case 4: return 2;
    case 5: case 6: case 7: return 4;
    case 9: case 12: case 13: return 8;
    default: return 0;
  }
}

// one miMATRIX element: 1 = found the variable (sizes checked, values in out), 0 = another
// variable, -1 = malformed or sizes that do not match
int mat5_matrix(const uint8_t* p, size_t n, const char* name, size_t want, double* out) {
  uint32_t type, nb;
  const uint8_t* d;
  size_t adv;
  if (!mat5_tag(p, n, &type, &nb, &d, &adv) || type != 6 || nb < 8) return -1;   // array flags
  uint32_t flags;
  std::memcpy(&flags, d, 4);
  const bool complex_ = (flags >> 11) & 1;
  p += adv; n -= adv;
  if (!mat5_tag(p, n, &type, &nb, &d, &adv) || type != 5) return -1;             // dimensions
  size_t count = 1;
  for (uint32_t i = 0; i < nb / 4; ++i) {
    int32_t x;
    std::memcpy(&x, d + 4 * i, 4);
    count *= (size_t)(x < 0 ? 0 : x);
  }
  p += adv; n -= adv;
  if (!mat5_tag(p, n, &type, &nb, &d, &adv) || (type != 1 && type != 2)) return -1;   // name
  const std::string vname((const char*)d, nb);
  p += adv; n -= adv;
  if (vname != name) return 0;
  if ((flags & 0xff) != 6 || complex_ || count != want) return -1;               // mxDOUBLE_CLASS, real
  if (!mat5_tag(p, n, &type, &nb, &d, &adv)) return -1;
  const size_t w = mat5_width(type);
  if (!w || nb != w * count) return -1;
  for (size_t i = 0; i < count; ++i)
    if (!mat5_number(type, d, i, out + i)) return -1;
  return 1;
}

// The double arrays names[0..nv) of a v5/v7 MAT file, the file read and each element
// inflated once: found[v] = 1 for each variable present; returns 0, or -1 when the file
// is malformed or a named variable is not a real double array of want[v] values.
int mat5_read_vars(const char* path, int nv, const char* const* names, const size_t* want, double* const* out,
                   int* found) {
  for (int v = 0; v < nv; ++v) found[v] = 0;
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return -1;
  std::vector<uint8_t> buf;
  uint8_t chunk[65536];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  std::fclose(f);
  if (buf.size() < 128 || buf[126] != 'I' || buf[127] != 'M') return -1;         // little-endian files only
  auto match = [&](const uint8_t* d, uint32_t nb) {
    for (int v = 0; v < nv; ++v) {
      const int r = mat5_matrix(d, nb, names[v], want[v], out[v]);
      if (r < 0) return -1;
      if (r == 1) {
        found[v] = 1;
        return 0;
      }
    }
    return 0;
  };
  size_t pos = 128;
  while (pos + 8 <= buf.size()) {
    uint32_t type, nb;
    const uint8_t* d;
    size_t adv;
    if (!mat5_tag(buf.data() + pos, buf.size() - pos, &type, &nb, &d, &adv)) return -1;
    int r = 0;
    if (type == 14) {
      r = match(d, nb);
    } else if (type == 15) {
      // inflate the element; its first 8 bytes are the inner miMATRIX tag
      std::vector<uint8_t> inner(1 << 16);
      z_stream zs{};
      if (inflateInit(&zs) != Z_OK) return -1;
      zs.next_in = const_cast<Bytef*>(d);
      zs.avail_in = nb;
      size_t have = 0;
      int zr = Z_OK;
      while (zr == Z_OK) {
        if (have == inner.size()) inner.resize(inner.size() * 2);
        zs.next_out = inner.data() + have;
        zs.avail_out = (uInt)(inner.size() - have);
        zr = inflate(&zs, Z_NO_FLUSH);
        have = inner.size() - zs.avail_out;
      }
      inflateEnd(&zs);
      if (zr != Z_STREAM_END) return -1;
      uint32_t it, inb;
      const uint8_t* idata;
      size_t iadv;
      if (!mat5_tag(inner.data(), have, &it, &inb, &idata, &iadv)) return -1;
      if (it == 14) r = match(idata, inb);
    }
    if (r != 0) return r;
    if (adv == 0) break;
    pos += adv;
  }
  return 0;
}

// the double array `name` of a v5/v7 MAT file: 1 found, 0 absent, -1 malformed / wrong size
int mat5_read(const char* path, const char* name, size_t want, double* out) {
  int found = 0;
  const int r = mat5_read_vars(path, 1, &name, &want, &out, &found);
  return r < 0 ? -1 : found;
}

bool check_sizes(int32_t nx, int32_t ny, int32_t nz) { return nx >= 1 && ny >= 1 && nz >= 1; }

// ---- Fortran edit descriptors of the log line ------------------------------------------
// The processor-dependent parts follow the Fortran compiler of this image (amdflang):
// no '+' on positive numbers, "NaN" / "Inf" / "-Inf" right-justified, a field too narrow
// for its value (or an exponent wider than Ee allows) is w asterisks.

std::string right(const std::string& s, int w) {
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}

std::string special(double x, int w) {
  if (std::isnan(x)) return right("NaN", w);
  return right(x < 0 ? "-Inf" : "Inf", w);
}

// ESw.dEe
std::string fmt_es(double x, int w, int d, int e) {
  if (!std::isfinite(x)) return special(x, w);
  char b[64];
  std::snprintf(b, sizeof(b), "%.*e", d, x);
  char* ep = std::strchr(b, 'e');
  const int ex = std::atoi(ep + 1);
  *ep = 0;
  int lim = 1;
  for (int i = 0; i < e; ++i) lim *= 10;
  if (std::abs(ex) >= lim) return std::string(w, '*');
  char eb[16];
  std::snprintf(eb, sizeof(eb), "E%c%0*d", ex < 0 ? '-' : '+', e, std::abs(ex));
  return right(std::string(b) + eb, w);
}

// Fw.d (the optional zero before the decimal point is dropped when the field needs the room)
std::string fmt_f(double x, int w, int d) {
  if (!std::isfinite(x)) return special(x, w);
  char b[512];
  std::snprintf(b, sizeof(b), "%.*f", d, x);
  std::string s(b);
  if ((int)s.size() == w + 1) {
    if (s.compare(0, 2, "0.") == 0) s.erase(0, 1);
    else if (s.compare(0, 3, "-0.") == 0) s.erase(1, 1);
  }
  return right(s, w);
}

std::string fmt_i(long long n, int w) { return right(std::to_string(n), w); }

int put_text(const std::string& s, char* buf, int64_t size) {
  if (!buf || size < (int64_t)s.size() + 1) return fail(GATT_E_INVALID, "log buffer too small");
  std::memcpy(buf, s.data(), s.size());
  buf[s.size()] = 0;
  return (int)s.size();
}

}  // namespace

extern "C" {

const char* gatt_flow_last_error(void) { return g_err.c_str(); }

int gatt_flow_write(const char* path, int32_t nx, int32_t ny, int32_t nz, double t, const double* const* q,
                    const uint8_t* inside_wall) {
  if (!path || !q) return fail(GATT_E_INVALID, "null argument");
  if (!check_sizes(nx, ny, nz)) return fail(GATT_E_INVALID, "grid sizes must be positive");
  for (int f = 0; f < 5; ++f)
    if (!q[f]) return fail(GATT_E_INVALID, "null field pointer");
  return write_vars(path, nx, ny, nz, &t, 5, kNames, q, inside_wall);
}

int gatt_flow_write_deriv(const char* path, int32_t nx, int32_t ny, int32_t nz, double t, const double* const* q,
                          const double* const* qdot, const uint8_t* inside_wall) {
  if (!path || !q || !qdot) return fail(GATT_E_INVALID, "null argument");
  if (!check_sizes(nx, ny, nz)) return fail(GATT_E_INVALID, "grid sizes must be positive");
  static const char* const names[10] = {"U", "V", "W", "R", "E", "Udot", "Vdot", "Wdot", "Rdot", "Edot"};
  const double* all[10];
  for (int f = 0; f < 5; ++f) {
    all[f] = q[f];
    all[5 + f] = qdot[f];
    if (!q[f] || !qdot[f]) return fail(GATT_E_INVALID, "null field pointer");
  }
  return write_vars(path, nx, ny, nz, &t, 10, names, all, inside_wall);
}

int gatt_meanflow_write(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* const* q,
                        const uint8_t* inside_wall) {
  if (!path || !q) return fail(GATT_E_INVALID, "null argument");
  if (!check_sizes(nx, ny, nz)) return fail(GATT_E_INVALID, "grid sizes must be positive");
  for (int f = 0; f < 5; ++f)
    if (!q[f]) return fail(GATT_E_INVALID, "null field pointer");
  return write_vars(path, nx, ny, nz, nullptr, 5, kNames, q, inside_wall);
}

int gatt_sfd_write(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* sfd_x) {
  if (!path || !sfd_x) return fail(GATT_E_INVALID, "null argument");
  if (!check_sizes(nx, ny, nz)) return fail(GATT_E_INVALID, "grid sizes must be positive");
  static const char* const names[1] = {"SFD_X"};
  return write_vars(path, nx, ny, nz, nullptr, 1, names, &sfd_x, nullptr);
}

int gatt_sfd_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* sfd_x) {
  if (!path || !sfd_x) return fail(GATT_E_INVALID, "null argument");
  if (!check_sizes(nx, ny, nz)) return fail(GATT_E_INVALID, "grid sizes must be positive");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  if (H5Fis_hdf5(path) <= 0) {
    // MATLAB's default save format (preprocessing.m:104 saves without '-v7.3')
    const int r = mat5_read(path, "SFD_X", (size_t)nx * ny * nz, sfd_x);
    if (r == 1) return GATT_OK;
    return fail(GATT_E_INVALID, std::string(path) + (r == 0 ? ": no SFD_X" : ": not a readable MAT file, or SFD_X "
                                                               "is not a real double array of the grid's size"));
  }
  hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return fail(GATT_E_INVALID, std::string("cannot open ") + path + " as a MAT 7.3 (HDF5) file");
  bool present = false;
  const bool ok = get_var(file, "SFD_X", nx, ny, nz, sfd_x, &present) && present;
  H5Fclose(file);
  if (!ok) return fail(GATT_E_INVALID, std::string(path) + ": no SFD_X of the grid's sizes");
  return GATT_OK;
}

int gatt_meanflow_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* const* q,
                       const uint8_t* inside_wall) {
  if (!path || !q) return fail(GATT_E_INVALID, "null argument");
  if (!check_sizes(nx, ny, nz)) return fail(GATT_E_INVALID, "grid sizes must be positive");
  for (int f = 0; f < 5; ++f)
    if (!q[f]) return fail(GATT_E_INVALID, "null field pointer");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return fail(GATT_E_INVALID, std::string("cannot open ") + path + " as a MAT 7.3 (HDF5) file");
  const size_t np = (size_t)nx * ny * nz;
  bool present = false;
  bool ok = true;
  for (int f = 0; ok && f < 5; ++f) {
    ok = get_var(file, kNames[f], nx, ny, nz, q[f], &present);
    if (ok && !present) {
      if (f != 2) ok = false;
      else std::memset(q[f], 0, np * sizeof(double));
    }
  }
  H5Fclose(file);
  if (!ok) return fail(GATT_E_INVALID, std::string(path) + ": missing variable or sizes that do not match the grid");
  // importMeanFlow.F90: the flow's wall points (or, without the flow's mask, the NaN points)
  for (size_t n = 0; n < np; ++n) {
    const bool wall = inside_wall ? inside_wall[n] != 0 : std::isnan(q[0][n]);
    if (wall) {
      q[0][n] = 0;
      q[1][n] = 0;
      q[2][n] = 0;
      q[3][n] = 1;
      q[4][n] = 1;
    }
  }
  return GATT_OK;
}

int gatt_mat_write(const char* path, int32_t nvars, const char* const* names, const int32_t* ranks,
                   const int64_t* dims, const double* const* re, const double* const* im) {
  if (!path || nvars < 1 || !names || !ranks || !dims || !re) return fail(GATT_E_INVALID, "null argument");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
  H5Pset_userblock(fcpl, 512);
  hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT);
  H5Pclose(fcpl);
  if (file < 0) return fail(GATT_E_INVALID, std::string("cannot create ") + path);
  bool ok = true;
  const int64_t* d = dims;
  for (int32_t v = 0; ok && v < nvars; ++v) {
    if (!names[v] || !re[v] || ranks[v] < 2 || ranks[v] > 8) {
      ok = false;
      break;
    }
    for (int i = 0; i < ranks[v]; ++i)
      if (d[i] < 0) ok = false;
    if (ok) ok = put_array(file, names[v], ranks[v], d, re[v], im ? im[v] : nullptr);
    d += ranks[v];
  }
  if (H5Fclose(file) < 0) ok = false;
  if (!ok) return fail(GATT_E_INVALID, std::string("writing ") + path + " failed (names, ranks 2..8, sizes)");
  if (!put_header(path)) return fail(GATT_E_INVALID, std::string("cannot write the MAT header of ") + path);
  return GATT_OK;
}

int gatt_log_header(int32_t ntracked, char* buf, int64_t size) {
  if (ntracked < 0) return fail(GATT_E_INVALID, "negative tracked-point count");
  std::string s = "Save number\tIteration\tSimulation time\tdt       \tCFL      \tU change\tV change\tW change\tR change\tE change";
  for (int i = 1; i <= ntracked; ++i) {
    const std::string k = std::to_string(i);
    for (const char* v : {"U", "V", "W", "R", "E"}) s += std::string("\t") + v + k + "            ";
  }
  s += "\n";
  return put_text(s, buf, size);
}

int gatt_log_line(int32_t nsave, int32_t tstep, double t, double dt, double cfl, const double* max_change,
                  int32_t ntracked, const double* tracked, double tracked_norm, char* buf, int64_t size) {
  if (!max_change || (ntracked > 0 && !tracked) || ntracked < 0) return fail(GATT_E_INVALID, "null argument");
  std::string s = fmt_i(nsave, 10) + "\t" + fmt_i(tstep, 10) + "\t" + fmt_es(t, 13, 8, 1) + "\t" + fmt_es(dt, 10, 5, 1) +
                  "\t" + fmt_f(cfl, 8, 5);
  for (int k = 0; k < 5; ++k) s += "\t" + fmt_es(max_change[k], 10, 4, 2);
  for (int i = 0; i < ntracked; ++i) {
    double v[5];
    for (int k = 0; k < 5; ++k) v[k] = tracked[5 * i + k];
    if (tracked_norm != 0) {
      v[3] = v[3] - 1;
      v[4] = v[4] / tracked_norm - 1;
    }
    for (int k = 0; k < 5; ++k) s += v[k] == 0 ? std::string("\t0") : "\t" + fmt_es(v[k], 15, 8, 2);
  }
  s += " \n";
  return put_text(s, buf, size);
}

int gatt_flow_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* t, double* const* q,
                   uint8_t* inside_wall) {
  if (!path || !q || !t) return fail(GATT_E_INVALID, "null argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(GATT_E_INVALID, "grid sizes must be positive");
  for (int f = 0; f < 5; ++f)
    if (!q[f]) return fail(GATT_E_INVALID, "null field pointer");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const size_t np = (size_t)nx * ny * nz;
  bool ok = true;
  if (H5Fis_hdf5(path) <= 0) {
    // MATLAB's default save format: the averaged flow files of runDNSautomean.m:80 are
    // saved without '-v7.3' (level 5 MAT file, variables possibly zlib-compressed)
    const char* names[6] = {"t", kNames[0], kNames[1], kNames[2], kNames[3], kNames[4]};
    const size_t want[6] = {1, np, np, np, np, np};
    double* outs[6] = {t, q[0], q[1], q[2], q[3], q[4]};
    int found[6];
    if (mat5_read_vars(path, 6, names, want, outs, found) < 0)
      return fail(GATT_E_INVALID, std::string(path) + ": not a readable MAT file, or a flow variable is not a real "
                                                      "double array of the grid's size");
    for (int v = 0; v < 6; ++v)
      if (!found[v]) {
        if (v == 3) std::memset(q[2], 0, np * sizeof(double));
        else ok = false;
      }
  } else {
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return fail(GATT_E_INVALID, std::string("cannot open ") + path + " as a MAT 7.3 (HDF5) file");
    bool present = false;
    ok = get_var(file, "t", 1, 1, 1, t, &present) && present;
    for (int f = 0; ok && f < 5; ++f) {
      ok = get_var(file, kNames[f], nx, ny, nz, q[f], &present);
      if (ok && !present) {
        if (f != 2) ok = false;
        else std::memset(q[f], 0, np * sizeof(double));
      }
    }
    H5Fclose(file);
  }
  if (!ok) return fail(GATT_E_INVALID, std::string(path) + ": missing variable or sizes that do not match the grid");
  // importFlow.F90:26-40
  for (size_t n = 0; n < np; ++n) {
    const bool wall = std::isnan(q[0][n]);
    if (inside_wall) inside_wall[n] = wall ? 1 : 0;
    if (wall) {
      q[0][n] = 0;
      q[1][n] = 0;
      q[2][n] = 0;
      q[3][n] = 1;
      q[4][n] = 1;
    }
  }
  return GATT_OK;
}

}  // extern "C"
