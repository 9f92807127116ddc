// HOPr mesh file access (HDF5 C API) for libself_amd_io.so: the attribute and dataset reads
// of Read_HOPr_Mesh2D (src/SELF_Mesh.f90:852-1016).  Host code only.
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/self_amd_io.h"

namespace {

constexpr int kOk = 0, kBadArgument = 1, kFileError = 5;
thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

// RAII handles; HDF5's own error stack printing is turned off, failures are reported here
struct H5Handle {
  hid_t id = H5I_INVALID_HID;
  herr_t (*close)(hid_t) = nullptr;
  H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H5Handle() {
    if (id >= 0 && close) close(id);
  }
  H5Handle(const H5Handle &) = delete;
  H5Handle &operator=(const H5Handle &) = delete;
  bool ok() const { return id >= 0; }
};

hid_t open_file(const char *path) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  return H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
}

// reads a whole dataset as mem_type into out (count elements)
int read_dataset(const char *path, const char *name, hid_t memType, void *out, int64_t count, const char *who) {
  if (!path || !name || !out || count < 0) return fail(kBadArgument, "%s: null argument or negative count", who);
  H5Handle f(open_file(path), H5Fclose);
  if (!f.ok()) return fail(kFileError, "%s: cannot open %s", who, path);
  H5Handle d(H5Dopen2(f.id, name, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(kFileError, "%s: no dataset %s in %s", who, name, path);
  H5Handle sp(H5Dget_space(d.id), H5Sclose);
  const hssize_t n = sp.ok() ? H5Sget_simple_extent_npoints(sp.id) : -1;
  if (n < 0) return fail(kFileError, "%s: cannot size dataset %s", who, name);
  if ((int64_t)n != count)
    return fail(kBadArgument, "%s: dataset %s has %lld elements, buffer %lld", who, name, (long long)n,
                (long long)count);
  if (n > 0 && H5Dread(d.id, memType, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
    return fail(kFileError, "%s: reading dataset %s failed", who, name);
  return kOk;
}

}  // namespace

extern "C" {

const char *selfamd_io_last_error(void) { return g_error.c_str(); }

int selfamd_hopr_attribute_int(const char *path, const char *name, int32_t *value) {
  const char *who = "hopr_attribute_int";
  if (!path || !name || !value) return fail(kBadArgument, "%s: null argument", who);
  H5Handle f(open_file(path), H5Fclose);
  if (!f.ok()) return fail(kFileError, "%s: cannot open %s", who, path);
  H5Handle a(H5Aopen_by_name(f.id, "/", name, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return fail(kFileError, "%s: no attribute %s in %s", who, name, path);
  H5Handle sp(H5Aget_space(a.id), H5Sclose);
  if (!sp.ok() || H5Sget_simple_extent_npoints(sp.id) != 1)
    return fail(kFileError, "%s: attribute %s is not a single value", who, name);
  if (H5Aread(a.id, H5T_NATIVE_INT32, value) < 0) return fail(kFileError, "%s: reading attribute %s failed", who, name);
  return kOk;
}

int selfamd_hopr_dataset_dims(const char *path, const char *name, int32_t *rank, int64_t *dims) {
  const char *who = "hopr_dataset_dims";
  if (!path || !name || !rank || !dims) return fail(kBadArgument, "%s: null argument", who);
  H5Handle f(open_file(path), H5Fclose);
  if (!f.ok()) return fail(kFileError, "%s: cannot open %s", who, path);
  H5Handle d(H5Dopen2(f.id, name, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(kFileError, "%s: no dataset %s in %s", who, name, path);
  H5Handle sp(H5Dget_space(d.id), H5Sclose);
  const int r = sp.ok() ? H5Sget_simple_extent_ndims(sp.id) : -1;
  if (r < 1 || r > 2) return fail(kFileError, "%s: dataset %s has rank %d (1 or 2 expected)", who, name, r);
  hsize_t h[2] = {0, 0};
  H5Sget_simple_extent_dims(sp.id, h, nullptr);
  *rank = r;
  for (int k = 0; k < r; k++) dims[k] = (int64_t)h[k];
  return kOk;
}

int selfamd_hopr_read_int(const char *path, const char *name, int32_t *out, int64_t count) {
  return read_dataset(path, name, H5T_NATIVE_INT32, out, count, "hopr_read_int");
}

int selfamd_hopr_read_double(const char *path, const char *name, double *out, int64_t count) {
  return read_dataset(path, name, H5T_NATIVE_DOUBLE, out, count, "hopr_read_double");
}

}  // extern "C"
