// HDF5 file access for libself_amd_io.so (host code only): the attribute and dataset reads of
// Read_HOPr_Mesh2D (src/SELF_Mesh.f90:852-1016) and the model pickup files of WriteModel /
// ReadModel (Write_Model2D / Read_Model2D, src/SELF_Model2D.f90:861-1099).
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include <vector>

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
  if (r < 1 || r > 5) return fail(kFileError, "%s: dataset %s has rank %d (1 to 5 expected)", who, name, r);
  hsize_t h[5] = {0, 0, 0, 0, 0};
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

// ---- model pickup files ----

namespace {

// writes a double dataset of the given row-major dims (creating it)
bool write_double(hid_t file, const char *name, int rank, const hsize_t *dims, const double *data) {
  H5Handle sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  if (!sp.ok()) return false;
  H5Handle d(H5Dcreate2(file, name, H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  return d.ok() && H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

bool make_group(hid_t file, const char *name) {
  H5Handle g(H5Gcreate2(file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  return g.ok();
}

}  // namespace

int selfamd_pickup_write(const char *path, int N, int nVar, int nEl, const double *xi, const double *weights,
                         const double *dgMatrix, const double *dMatrix, const double *bMatrix, const double *solution,
                         const double *solutionBoundary, const double *x, const double *xBoundary) {
  const char *who = "pickup_write";
  if (!path || !xi || !weights || !dgMatrix || !dMatrix || !bMatrix || !solution || !solutionBoundary || !x ||
      !xBoundary)
    return fail(kBadArgument, "%s: null argument", who);
  if (N < 1 || nVar < 1 || nEl < 1) return fail(kBadArgument, "%s: bad sizes N=%d nVar=%d nEl=%d", who, N, nVar, nEl);
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  H5Handle f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(kFileError, "%s: cannot create %s", who, path);
  const hsize_t P = N + 1, V = nVar, E = nEl;
  // Fortran arrays appear in the file with their dimensions reversed (row-major view)
  const hsize_t d1[1] = {P}, dPP[2] = {P, P}, dB[2] = {2, P};
  const hsize_t dSol[4] = {E, V, P, P}, dSolB[4] = {E, 4, V, P};
  const hsize_t dX[5] = {E, 1, P, P, 2}, dXB[5] = {E, 4, 1, P, 2};
  {
    // the attribute N that Read_Model2D requires (src/SELF_Model2D.f90:1078)
    H5Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
    H5Handle a(H5Acreate2(f.id, "N", H5T_STD_I32LE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    const int32_t n32 = N;
    if (!a.ok() || H5Awrite(a.id, H5T_NATIVE_INT32, &n32) < 0) return fail(kFileError, "%s: writing attribute N failed", who);
  }
  bool ok = make_group(f.id, "/quadrature") && write_double(f.id, "/quadrature/xi", 1, d1, xi) &&
            write_double(f.id, "/quadrature/weights", 1, d1, weights) &&
            write_double(f.id, "/quadrature/dgmatrix", 2, dPP, dgMatrix) &&
            write_double(f.id, "/quadrature/dmatrix", 2, dPP, dMatrix) &&
            write_double(f.id, "/quadrature/bmatrix", 2, dB, bMatrix) && make_group(f.id, "/state") &&
            make_group(f.id, "/state/interior") && make_group(f.id, "/state/boundary") && make_group(f.id, "/mesh") &&
            make_group(f.id, "/mesh/interior") && make_group(f.id, "/mesh/boundary") &&
            write_double(f.id, "/state/interior/solution", 4, dSol, solution) &&
            write_double(f.id, "/state/boundary/solution", 4, dSolB, solutionBoundary) &&
            write_double(f.id, "/mesh/interior/x", 5, dX, x) && write_double(f.id, "/mesh/boundary/x", 5, dXB, xBoundary);
  if (!ok) return fail(kFileError, "%s: writing %s failed", who, path);
  return kOk;
}

int selfamd_pickup_write_registers(const char *path, int N, int nVar, int nEl, const double *fluxDivergence,
                                   const double *flux, const double *fluxBoundary, const double *solutionGradient,
                                   const double *solutionGradientBoundary) {
  const char *who = "pickup_write_registers";
  if (!path) return fail(kBadArgument, "%s: null path", who);
  if (N < 1 || nVar < 1 || nEl < 1) return fail(kBadArgument, "%s: bad sizes N=%d nVar=%d nEl=%d", who, N, nVar, nEl);
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  H5Handle f(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(kFileError, "%s: cannot open %s", who, path);
  const hsize_t P = N + 1, V = nVar, E = nEl;
  const hsize_t dSol[4] = {E, V, P, P}, dVec[5] = {E, V, P, P, 2}, dVecB[5] = {E, 4, V, P, 2};
  const size_t nVecB = (size_t)E * 4 * V * P * 2, nVec = (size_t)E * V * P * P * 2;
  std::vector<double> zeros;
  auto or_zeros = [&](const double *a, size_t n) {
    if (a) return a;
    if (zeros.size() < n) zeros.assign(n, 0.0);
    return (const double *)zeros.data();
  };
  bool ok = write_double(f.id, "/state/interior/fluxDivergence", 4, dSol, or_zeros(fluxDivergence, nVec / 2));
  ok = ok && write_double(f.id, "/state/interior/flux", 5, dVec, or_zeros(flux, nVec));
  ok = ok && write_double(f.id, "/state/boundary/flux", 5, dVecB, or_zeros(fluxBoundary, nVecB));
  ok = ok && write_double(f.id, "/state/interior/solutionGradient", 5, dVec, or_zeros(solutionGradient, nVec));
  ok = ok && write_double(f.id, "/state/boundary/solutionGradient", 5, dVecB,
                          or_zeros(solutionGradientBoundary, nVecB));
  if (!ok) return fail(kFileError, "%s: writing %s failed", who, path);
  return kOk;
}

int selfamd_pickup_read(const char *path, int N, int nVar, int nEl, double *solution) {
  const char *who = "pickup_read";
  if (!path || !solution) return fail(kBadArgument, "%s: null argument", who);
  int32_t nFile = -1;
  // (a pickup written by the reference itself has no N: Read_Model2D cannot read it either)
  const int rc = selfamd_hopr_attribute_int(path, "N", &nFile);
  if (rc) return rc;
  if (nFile != N)
    return fail(kBadArgument, "%s: solution polynomial degree %d does not match the file's %d", who, N, nFile);
  int32_t rank = 0;
  int64_t dims[5] = {0, 0, 0, 0, 0};
  const int rd = selfamd_hopr_dataset_dims(path, "/state/interior/solution", &rank, dims);
  if (rd) return rd;
  if (rank != 4 || dims[0] != nEl || dims[1] != nVar || dims[2] != N + 1 || dims[3] != N + 1)
    return fail(kBadArgument, "%s: /state/interior/solution is (%lld, %lld, %lld, %lld), expected (%d, %d, %d, %d)", who,
                (long long)dims[0], (long long)dims[1], (long long)dims[2], (long long)dims[3], nEl, nVar, N + 1, N + 1);
  return read_dataset(path, "/state/interior/solution", H5T_NATIVE_DOUBLE, solution,
                      (int64_t)nEl * nVar * (N + 1) * (N + 1), who);
}

}  // extern "C"
