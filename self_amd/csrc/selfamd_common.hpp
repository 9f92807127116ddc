// Common definitions for the self_amd HIP library (gfx950 / MI355X).
//
// Layout contract (shared by every entry point): the reference's Fortran column-major
// shapes, zero-based, with P = N+1:
//   interior(0:N,0:N,nVar,nEl)        SC(i,j,v,e)   = i + P*(j + P*(v + nVar*e))
//   vector interior(2,0:N,0:N,...)    VE(d,i,j,v,e) = d + 2*SC(i,j,v,e)
//   tensor interior(2,2,0:N,0:N,...)  TE(r,c,...)   = r + 2*c + 4*SC(...)
//   boundary(0:N,nVar,4,nEl)          SCB(i,v,s,e)  = i + P*(v + nVar*(s + 4*e))
//   sideInfo(5,4,nEl) int32           SI(k,s,e)     = k + 5*(s + 4*e)
// These shapes are already element-major: one element's nodes of one variable are
// P*P contiguous doubles, so a wavefront that owns an element issues fully coalesced
// 512-byte loads per variable at N = 7.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

namespace selfamd {

// Error handling: every C entry point returns 0 on success or a positive code, and
// keeps a readable message for selfamd_last_error().
enum Status : int {
  kOk = 0,
  kBadArgument = 1,
  kUnsupported = 2,
  kHipError = 3,
  kNotAffine = 4,
};

void set_error(const char *fmt, ...);
const char *last_error();

#define SELFAMD_HIP(call)                                                                  \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess) {                                                                \
      ::selfamd::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                           __LINE__);                                                      \
      return ::selfamd::kHipError;                                                         \
    }                                                                                      \
  } while (0)

#define SELFAMD_CHECK_LAUNCH()                                                             \
  do {                                                                                     \
    hipError_t _e = hipGetLastError();                                                     \
    if (_e != hipSuccess) {                                                                \
      ::selfamd::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),     \
                           __FILE__, __LINE__);                                            \
      return ::selfamd::kHipError;                                                         \
    }                                                                                      \
  } while (0)

#define SELFAMD_REQUIRE(cond, ...)                                                         \
  do {                                                                                     \
    if (!(cond)) {                                                                         \
      ::selfamd::set_error(__VA_ARGS__);                                                   \
      return ::selfamd::kBadArgument;                                                      \
    }                                                                                      \
  } while (0)

// Largest polynomial degree the per-operator kernels accept; the fused stage kernels
// are instantiated for N in {1,...,7}.
constexpr int kMaxN = 15;

inline __host__ __device__ size_t sc_idx(int i, int j, int v, int e, int P, int nVar) {
  return (size_t)i + P * ((size_t)j + P * ((size_t)v + (size_t)nVar * e));
}
inline __host__ __device__ size_t scb_idx(int i, int v, int s, int e, int P, int nVar) {
  return (size_t)i + P * ((size_t)v + nVar * ((size_t)s + 4 * (size_t)e));
}

// Per-element record of an affine element (model engine):
//   dsdx(4) J(1) dt(1, batched runs) 1/J(1) pad(1) nHat_x(4 sides) nHat_y(4) nScale(4)
constexpr int kGeomStride = 20;

inline unsigned grid1d(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

}  // namespace selfamd
