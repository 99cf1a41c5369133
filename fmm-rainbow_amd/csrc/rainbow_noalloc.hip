// The FP64 direct sum without allocation: rainbow_hip_direct_sum_f64_noalloc,
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid it).
//
// rainbow_hip_direct_sum_f64 works in the handle's staging buffer, which
// grows (hipMalloc) with the number of targets.  This variant runs the same
// kernels on the same scratch layout (target list, each target's slot, the
// partial sums of the body chunks, the result) inside `scratch`, a
// file-scope device array of kScratchBytes that every GPU the library runs
// on holds until the process ends, so its results equal that entry point's
// bit for bit.  It takes at most kMaxTargets targets.  Every call on one
// GPU uses that GPU's one array: calls that overlap on a GPU race (the
// call waits for its own stream before returning, so calls made one after
// another from any number of handles are safe).  Nothing in the array has
// to start at zero: the slot list is filled before it is searched, the
// partial sums and the result are written in full.

#include <hip/hip_runtime.h>

#include <mutex>

#include "rainbow_fmm.h"
#include "rainbow_launch.h"
#include "rainbow_runtime.h"

namespace rainbow {
namespace {

constexpr uint32_t kMaxTargets = 4096u;
constexpr size_t kScratchBytes = 17u << 20;  // 17 MiB
static_assert(direct_scratch_bytes(kMaxTargets) <= kScratchBytes, "scratch holds the largest layout");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device
// with that device current.  In host code the array's name is only a
// stand-in: kernels and copies get this address.
int scratch_on(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail(RAINBOW_E_INVALID, "device %d out of range", device);
  std::lock_guard<std::mutex> guard(g_lock);
  if (g_scratch[device] == nullptr) {
    int current = 0;
    hipError_t e = hipGetDevice(&current);
    if (e == hipSuccess && current != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (e != hipSuccess || p == nullptr)
      return fail(RAINBOW_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                  hipGetErrorString(e));
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return RAINBOW_OK;
}

}  // namespace
}  // namespace rainbow

using rainbow::fail;

extern "C" int rainbow_hip_direct_sum_f64_noalloc(rainbow_hip_t* h, const uint32_t* targets, uint32_t ntargets,
                                                  double* out) {
  if (h == nullptr) return fail(RAINBOW_E_INVALID, "null handle");
  if (ntargets > rainbow::kMaxTargets)
    return fail(RAINBOW_E_INVALID, "%u targets: the allocation-free direct sum takes at most %u", ntargets,
                rainbow::kMaxTargets);
  rainbow::Engine& e = h->batch.engine;
  int rc = e.check_direct_args(0u, targets, ntargets, out);
  if (rc != RAINBOW_OK || ntargets == 0u) return rc;
  void* ptr = nullptr;
  if ((rc = rainbow::scratch_on(e.device, &ptr)) != RAINBOW_OK) return rc;
  return e.direct_sum_f64_in(0u, targets, ntargets, out, ptr);
}
