// Device memory and synchronisation plumbing for hosts that bind the C ABI without a HIP
// runtime of their own (Python via ctypes, Fortran via iso_c_binding).  Keeping these in the
// library means a process loads exactly one HIP runtime: the one libself_amd.so links.
#include "selfamd_common.hpp"
#include "../../include/self_amd.h"

using namespace selfamd;

extern "C" {

int selfamd_device_count(int *count) {
  SELFAMD_REQUIRE(count, "device_count: null argument");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) *count = 0;
  return kOk;
}

int selfamd_device_name(int device, char *name, int len) {
  SELFAMD_REQUIRE(name && len > 0, "device_name: null argument");
  hipDeviceProp_t prop;
  SELFAMD_HIP(hipGetDeviceProperties(&prop, device));
  snprintf(name, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return kOk;
}

int selfamd_malloc(void **ptr, size_t bytes) {
  SELFAMD_REQUIRE(ptr, "malloc: null argument");
  *ptr = nullptr;
  SELFAMD_HIP(hipMalloc(ptr, bytes ? bytes : 8));
  return kOk;
}

int selfamd_free(void *ptr) {
  if (ptr) SELFAMD_HIP(hipFree(ptr));
  return kOk;
}

int selfamd_memcpy(void *dst, const void *src, size_t bytes, int kind) {
  SELFAMD_REQUIRE(kind >= 1 && kind <= 3, "memcpy: kind %d (1 = H2D, 2 = D2H, 3 = D2D)", kind);
  if (!bytes) return kOk;
  SELFAMD_REQUIRE(dst && src, "memcpy: null pointer");
  const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  SELFAMD_HIP(hipMemcpy(dst, src, bytes, k));
  return kOk;
}

int selfamd_memset(void *ptr, int value, size_t bytes) {
  if (!bytes) return kOk;
  SELFAMD_REQUIRE(ptr, "memset: null pointer");
  SELFAMD_HIP(hipMemset(ptr, value, bytes));
  return kOk;
}

int selfamd_time_copy_f64(const double *src, double *dst, size_t n, int reps, double *elapsedMs) {
  SELFAMD_REQUIRE(elapsedMs && reps > 0, "time_copy_f64: bad arguments");
  hipStream_t st;
  SELFAMD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  SELFAMD_HIP(hipEventCreate(&e0));
  SELFAMD_HIP(hipEventCreate(&e1));
  int rc = selfamd_copy_f64(src, dst, n, st);  // warm-up
  if (rc) return rc;
  SELFAMD_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps && rc == 0; r++) rc = selfamd_copy_f64(src, dst, n, st);
  if (rc) return rc;
  SELFAMD_HIP(hipEventRecord(e1, st));
  SELFAMD_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  SELFAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
  *elapsedMs = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SELFAMD_HIP(hipStreamDestroy(st));
  return kOk;
}

int selfamd_device_synchronize(void) {
  SELFAMD_HIP(hipDeviceSynchronize());
  return kOk;
}

}  // extern "C"
