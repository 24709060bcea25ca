// noalloc.hip -- the device scratch of the allocation-free batched raycast
// (tvk_render_batch_noalloc): one job table and its counters in a file-scope
// device array, so that a launch needs no device memory from the library
// beyond what the caller passes in.
//
// Each GPU the library runs on holds its own copy of the array (64 KiB)
// until the process ends.  Host code never uses the array's name as an
// address: at set-up (HipRenderer::Initialize), once per device with that
// device current, hipGetSymbolAddress gives the device's address of its
// copy, which is kept per device and used for every kernel argument and
// copy.  Every call on a device shares that device's copy; calls that
// overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.hpp"
#include "tvk_common.hpp"

namespace tvk {

constexpr size_t kNoAllocScratchBytes = 64 << 10;
constexpr size_t kNoAllocCountersOffset = (sizeof(JobTable) + 255) & ~(size_t)255;
static_assert(kNoAllocCountersOffset + kCounterWords * sizeof(unsigned long long) <= kNoAllocScratchBytes,
              "the scratch holds one job table of kMaxJobs jobs and its counters");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (g_ptr[device]) return hipSuccess;
  int cur = -1;
  if (hipError_t e = hipGetDevice(&cur)) return e;
  if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current
  void* p = nullptr;
  if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) return e;
  g_ptr[device] = p;
  return hipSuccess;
}

hipError_t noalloc_scratch(int device, JobTable** table, unsigned long long** counters) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[device]) return hipErrorNotInitialized;
  *table = (JobTable*)g_ptr[device];
  *counters = (unsigned long long*)((unsigned char*)g_ptr[device] + kNoAllocCountersOffset);
  return hipSuccess;
}

}  // namespace tvk
