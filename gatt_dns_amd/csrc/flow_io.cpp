// Flow files in the MAT 7.3 layout (HDF5 behind a 512-byte MATLAB userblock), the
// format of source/Fortran/readWriteMat.F90 (matOpen 'w7.3', matPutVariable of t and
// U,V,W,R,E), with the NaN convention of exportFlow.F90 / importFlow.F90:24-40.
// Host code only; built against libhdf5 into its own library (libgatt_flowio.so).
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
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

}  // namespace

extern "C" {

const char* gatt_flow_last_error(void) { return g_err.c_str(); }

int gatt_flow_write(const char* path, int32_t nx, int32_t ny, int32_t nz, double t, const double* const* q,
                    const uint8_t* inside_wall) {
  if (!path || !q) return fail(GATT_E_INVALID, "null argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(GATT_E_INVALID, "grid sizes must be positive");
  for (int f = 0; f < 5; ++f)
    if (!q[f]) return fail(GATT_E_INVALID, "null field pointer");
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
  bool ok = put_var(file, "t", 2, d1, &t);
  std::vector<double> tmp;
  for (int f = 0; ok && f < 5; ++f) {
    const double* src = q[f];
    if (inside_wall) {
      // exportFlow.F90: NaN in every field inside the walls
      tmp.assign(src, src + np);
      for (size_t n = 0; n < np; ++n)
        if (inside_wall[n]) tmp[n] = std::numeric_limits<double>::quiet_NaN();
      src = tmp.data();
    }
    ok = put_var(file, kNames[f], rank, d, src);
  }
  if (H5Fclose(file) < 0) ok = false;
  if (!ok) return fail(GATT_E_INVALID, std::string("writing ") + path + " failed");
  if (!put_header(path)) return fail(GATT_E_INVALID, std::string("cannot write the MAT header of ") + path);
  return GATT_OK;
}

int gatt_flow_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* t, double* const* q,
                   uint8_t* inside_wall) {
  if (!path || !q || !t) return fail(GATT_E_INVALID, "null argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(GATT_E_INVALID, "grid sizes must be positive");
  for (int f = 0; f < 5; ++f)
    if (!q[f]) return fail(GATT_E_INVALID, "null field pointer");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return fail(GATT_E_INVALID, std::string("cannot open ") + path + " as a MAT 7.3 (HDF5) file");
  const size_t np = (size_t)nx * ny * nz;
  bool present = false;
  bool ok = get_var(file, "t", 1, 1, 1, t, &present) && present;
  for (int f = 0; ok && f < 5; ++f) {
    ok = get_var(file, kNames[f], nx, ny, nz, q[f], &present);
    if (ok && !present) {
      if (f != 2) ok = false;
      else std::memset(q[f], 0, np * sizeof(double));
    }
  }
  H5Fclose(file);
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
