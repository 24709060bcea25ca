// Host-only stand-ins for the few HIP runtime calls the memory manager
// makes, so that its pool policies run in a CPU test: device memory is
// plain host memory of a fixed-size device (1 GiB) and every call succeeds
// unless memory runs out.  TEST INFRASTRUCTURE ONLY.
#include <hip/hip_runtime_api.h>
#include <stdlib.h>

static int g_device = 0;
static size_t g_allocated = 0;
static const size_t kDeviceBytes = (size_t)1 << 30;

extern "C" {
hipError_t hipGetDevice(int* d) {
  *d = g_device;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) {
  g_device = d;
  return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b) {
  *total_b = kDeviceBytes;
  *free_b = kDeviceBytes - g_allocated;
  return hipSuccess;
}
hipError_t hipMalloc(void** p, size_t n) {
  if (g_allocated + n > kDeviceBytes) return hipErrorOutOfMemory;
  *p = malloc(n + 16);
  ((size_t*)*p)[0] = n;
  *p = (char*)*p + 16;
  g_allocated += n;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  char* b = (char*)p - 16;
  g_allocated -= ((size_t*)b)[0];
  free(b);
  return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
}
