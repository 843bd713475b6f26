// HDF5 result files in DUST's layout (host code, HDF5 C API).
//
// The reference writes its results and restart state with the HDF5 Fortran
// API through src/io/mod_hdf5_io.f90: IEEE F64LE and STD I32LE datasets,
// scalars in scalar dataspaces, fixed-length space-padded strings, and the
// root attributes of sim_param%save_param.  A Fortran array x(3,n) is stored
// with the dimensions reversed, so the file holds the C array [n][3]: the
// interleaved (x1,y1,z1,x2,...) layout that the rest of this library uses
// stores unchanged.  Paths name nested groups ("ParticleWake/WakePoints");
// missing intermediate groups are created.
//
// libdust_h5.so is separate from libdust_gpu.so so that the GPU library does
// not depend on an HDF5 installation.
#include <hdf5.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dust_h5.h"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// HDF5 prints its error stack by default; errors are reported as status codes.
void quiet() {
  static bool done = false;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    done = true;
  }
}

// Closes an HDF5 identifier of any kind at scope exit.
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  Id(const Id &) = delete;
  Id &operator=(const Id &) = delete;
  bool ok() const { return id >= 0; }
};

hid_t file_type(int type) {
  return type == DUST_H5_F64 ? H5T_IEEE_F64LE : H5T_STD_I32LE;
}
hid_t mem_type(int type) { return type == DUST_H5_F64 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_INT32; }

// true when every group on the way to `path` and the last link exist
bool exists(hid_t f, const char *path) {
  std::string p(path);
  size_t pos = 0;
  while (true) {
    size_t q = p.find('/', pos);
    std::string pre = p.substr(0, q);
    if (!pre.empty() && pre != "/") {
      if (H5Lexists(f, pre.c_str(), H5P_DEFAULT) <= 0) return false;
    }
    if (q == std::string::npos) return true;
    pos = q + 1;
  }
}

hid_t link_props() {
  hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
  if (lcpl >= 0) H5Pset_create_intermediate_group(lcpl, 1);
  return lcpl;
}

int check_type(int type) {
  if (type != DUST_H5_F64 && type != DUST_H5_I32)
    return fail(DUST_H5_E_ARG, "unknown element type %d", type);
  return 0;
}

int check_dims(int rank, const int64_t *dims, std::vector<hsize_t> &h) {
  if (rank < 0 || rank > DUST_H5_MAX_RANK) return fail(DUST_H5_E_ARG, "rank %d out of range", rank);
  if (rank > 0 && !dims) return fail(DUST_H5_E_ARG, "null dimensions");
  h.resize(rank);
  for (int i = 0; i < rank; ++i) {
    if (dims[i] < 0) return fail(DUST_H5_E_ARG, "negative dimension");
    h[i] = (hsize_t)dims[i];
  }
  return 0;
}

}  // namespace

extern "C" {

const char *dust_h5_last_error(void) { return g_err; }

int dust_h5_create(const char *path, int64_t *file) {
  quiet();
  if (!path || !file) return fail(DUST_H5_E_ARG, "null argument");
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail(DUST_H5_E_IO, "cannot create %s", path);
  *file = (int64_t)f;
  return 0;
}

int dust_h5_open(const char *path, int64_t *file) {
  quiet();
  if (!path || !file) return fail(DUST_H5_E_ARG, "null argument");
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail(DUST_H5_E_IO, "cannot open %s", path);
  *file = (int64_t)f;
  return 0;
}

int dust_h5_close(int64_t file) {
  if (H5Fclose((hid_t)file) < 0) return fail(DUST_H5_E_IO, "close failed");
  return 0;
}

int dust_h5_write(int64_t file, const char *name, int type, int rank, const int64_t *dims,
                  const void *data) {
  quiet();
  if (!name) return fail(DUST_H5_E_ARG, "null name");
  if (int s = check_type(type)) return s;
  std::vector<hsize_t> h;
  if (int s = check_dims(rank, dims, h)) return s;
  size_t count = 1;
  for (hsize_t d : h) count *= d;
  if (count && !data) return fail(DUST_H5_E_ARG, "null data for %s", name);
  if (exists((hid_t)file, name)) return fail(DUST_H5_E_ARG, "%s exists already", name);
  Id space(rank ? H5Screate_simple(rank, h.data(), nullptr) : H5Screate(H5S_SCALAR), H5Sclose);
  Id lcpl(link_props(), H5Pclose);
  if (!space.ok() || !lcpl.ok()) return fail(DUST_H5_E_IO, "dataspace for %s", name);
  Id d(H5Dcreate2((hid_t)file, name, file_type(type), space.id, lcpl.id, H5P_DEFAULT,
                  H5P_DEFAULT),
       H5Dclose);
  if (!d.ok()) return fail(DUST_H5_E_IO, "cannot create dataset %s", name);
  if (count && H5Dwrite(d.id, mem_type(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(DUST_H5_E_IO, "cannot write %s", name);
  return 0;
}

int dust_h5_write_string(int64_t file, const char *name, const char *s) {
  quiet();
  if (!name || !s) return fail(DUST_H5_E_ARG, "null argument");
  if (exists((hid_t)file, name)) return fail(DUST_H5_E_ARG, "%s exists already", name);
  const size_t len = std::strlen(s) ? std::strlen(s) : 1;
  Id t(H5Tcopy(H5T_C_S1), H5Tclose);
  if (!t.ok() || H5Tset_size(t.id, len) < 0 || H5Tset_strpad(t.id, H5T_STR_SPACEPAD) < 0)
    return fail(DUST_H5_E_IO, "string type for %s", name);
  std::string buf(s);
  buf.resize(len, ' ');
  Id space(H5Screate(H5S_SCALAR), H5Sclose);
  Id lcpl(link_props(), H5Pclose);
  Id d(H5Dcreate2((hid_t)file, name, t.id, space.id, lcpl.id, H5P_DEFAULT, H5P_DEFAULT),
       H5Dclose);
  if (!d.ok()) return fail(DUST_H5_E_IO, "cannot create dataset %s", name);
  if (H5Dwrite(d.id, t.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
    return fail(DUST_H5_E_IO, "cannot write %s", name);
  return 0;
}

int dust_h5_write_attr(int64_t file, const char *name, int type, int rank, const int64_t *dims,
                       const void *data) {
  quiet();
  if (!name || !data) return fail(DUST_H5_E_ARG, "null argument");
  if (int s = check_type(type)) return s;
  std::vector<hsize_t> h;
  if (int s = check_dims(rank, dims, h)) return s;
  if (H5Aexists((hid_t)file, name) > 0) return fail(DUST_H5_E_ARG, "attribute %s exists", name);
  Id space(rank ? H5Screate_simple(rank, h.data(), nullptr) : H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2((hid_t)file, name, file_type(type), space.id, H5P_DEFAULT, H5P_DEFAULT),
       H5Aclose);
  if (!a.ok() || H5Awrite(a.id, mem_type(type), data) < 0)
    return fail(DUST_H5_E_IO, "cannot write attribute %s", name);
  return 0;
}

int dust_h5_write_attr_string(int64_t file, const char *name, const char *s, int64_t len) {
  quiet();
  if (!name || !s) return fail(DUST_H5_E_ARG, "null argument");
  const size_t n = len > 0 ? (size_t)len : (std::strlen(s) ? std::strlen(s) : 1);
  if (H5Aexists((hid_t)file, name) > 0) return fail(DUST_H5_E_ARG, "attribute %s exists", name);
  Id t(H5Tcopy(H5T_C_S1), H5Tclose);
  if (!t.ok() || H5Tset_size(t.id, n) < 0 || H5Tset_strpad(t.id, H5T_STR_SPACEPAD) < 0)
    return fail(DUST_H5_E_IO, "string type for %s", name);
  std::string buf(s);
  buf.resize(n, ' ');
  Id space(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2((hid_t)file, name, t.id, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Awrite(a.id, t.id, buf.data()) < 0)
    return fail(DUST_H5_E_IO, "cannot write attribute %s", name);
  return 0;
}

int dust_h5_exists(int64_t file, const char *name, int *out) {
  quiet();
  if (!name || !out) return fail(DUST_H5_E_ARG, "null argument");
  *out = exists((hid_t)file, name) ? 1 : 0;
  return 0;
}

int dust_h5_info(int64_t file, const char *name, int *type, int *rank, int64_t *dims) {
  quiet();
  if (!name || !type || !rank || !dims) return fail(DUST_H5_E_ARG, "null argument");
  if (!exists((hid_t)file, name)) return fail(DUST_H5_E_ARG, "no dataset %s", name);
  Id d(H5Dopen2((hid_t)file, name, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(DUST_H5_E_ARG, "%s is not a dataset", name);
  Id t(H5Dget_type(d.id), H5Tclose);
  Id s(H5Dget_space(d.id), H5Sclose);
  const H5T_class_t c = H5Tget_class(t.id);
  if (c == H5T_FLOAT) *type = DUST_H5_F64;
  else if (c == H5T_INTEGER) *type = DUST_H5_I32;
  else if (c == H5T_STRING) *type = DUST_H5_STR;
  else return fail(DUST_H5_E_ARG, "%s has an unsupported element type", name);
  const int r = H5Sget_simple_extent_ndims(s.id);
  if (r < 0 || r > DUST_H5_MAX_RANK) return fail(DUST_H5_E_ARG, "%s has rank %d", name, r);
  hsize_t h[DUST_H5_MAX_RANK];
  H5Sget_simple_extent_dims(s.id, h, nullptr);
  *rank = r;
  for (int i = 0; i < r; ++i) dims[i] = (int64_t)h[i];
  if (c == H5T_STRING) {  // the string length as the only dimension
    *rank = 1;
    dims[0] = (int64_t)H5Tget_size(t.id);
  }
  return 0;
}

int dust_h5_read(int64_t file, const char *name, int type, void *data, int64_t count) {
  quiet();
  if (!name) return fail(DUST_H5_E_ARG, "null name");
  if (int s = check_type(type)) return s;
  int ft, rank;
  int64_t dims[DUST_H5_MAX_RANK];
  if (int s = dust_h5_info(file, name, &ft, &rank, dims)) return s;
  if (ft == DUST_H5_STR) return fail(DUST_H5_E_ARG, "%s is a string", name);
  int64_t n = 1;
  for (int i = 0; i < rank; ++i) n *= dims[i];
  if (n != count) return fail(DUST_H5_E_ARG, "%s holds %lld values, %lld expected", name,
                              (long long)n, (long long)count);
  if (n && !data) return fail(DUST_H5_E_ARG, "null data");
  Id d(H5Dopen2((hid_t)file, name, H5P_DEFAULT), H5Dclose);
  if (n && H5Dread(d.id, mem_type(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(DUST_H5_E_IO, "cannot read %s", name);
  return 0;
}

int dust_h5_read_string(int64_t file, const char *name, char *out, int64_t cap) {
  quiet();
  if (!name || !out || cap < 1) return fail(DUST_H5_E_ARG, "null argument");
  int ft, rank;
  int64_t dims[DUST_H5_MAX_RANK];
  if (int s = dust_h5_info(file, name, &ft, &rank, dims)) return s;
  if (ft != DUST_H5_STR) return fail(DUST_H5_E_ARG, "%s is not a string", name);
  if (dims[0] + 1 > cap) return fail(DUST_H5_E_ARG, "string %s longer than the buffer", name);
  Id d(H5Dopen2((hid_t)file, name, H5P_DEFAULT), H5Dclose);
  Id t(H5Dget_type(d.id), H5Tclose);
  std::vector<char> buf((size_t)dims[0] + 1, 0);
  if (H5Dread(d.id, t.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
    return fail(DUST_H5_E_IO, "cannot read %s", name);
  size_t n = (size_t)dims[0];
  while (n > 0 && (buf[n - 1] == ' ' || buf[n - 1] == '\0')) --n;  // space padding
  std::memcpy(out, buf.data(), n);
  out[n] = '\0';
  return 0;
}

}  // extern "C"
