// Host-side launch wrappers of the step kernels (defined with the kernels).
#pragma once

#include <hip/hip_runtime.h>

#include "rainbow_device.h"

namespace rainbow {

// Optional event pair recorded by the kernel dispatch itself (start and end
// of the kernel's execution, without the gap before it starts).
struct Timing {
  hipEvent_t start;
  hipEvent_t stop;
};

hipError_t launch_clear(const StepArgs& a, uint32_t nblocks, hipStream_t stream, const Timing* t = nullptr);
hipError_t launch_deposit(const StepArgs& a, uint32_t nblocks, hipStream_t stream, const Timing* t = nullptr);
hipError_t launch_upward(const StepArgs& a, uint32_t nblocks, hipStream_t stream, const Timing* t = nullptr);
hipError_t launch_top(const StepArgs& a, hipStream_t stream, const Timing* t = nullptr);
// Tree pass of levels lo..hi (2 <= lo <= hi <= 6) in one launch.
hipError_t launch_tree(const StepArgs& a, uint32_t lo, uint32_t hi, uint32_t nblocks, hipStream_t stream,
                       const Timing* t = nullptr);
hipError_t launch_integrate(const StepArgs& a, uint32_t nblocks, hipStream_t stream, const Timing* t = nullptr);
// Owned Morton ranges of the deposit blocks, from the sorted keys.
hipError_t launch_own(const StepArgs& a, uint32_t ndeposit, const uint32_t* sorted, uint2* own, hipStream_t stream);
hipError_t launch_sort_keys(const StepArgs& a, uint32_t nblocks, uint32_t* keys, uint32_t* vals,
                            hipStream_t stream);
hipError_t launch_reorder(const StepArgs& a, uint32_t nblocks, const uint32_t* vals, float4* tpos, float4* tvel,
                          uint32_t* tperm, hipStream_t stream);
// Radix sort of (key, value) pairs over the low end_bit key bits; with
// temp == nullptr writes the scratch size to temp_bytes.
hipError_t sort_pairs(void* temp, size_t& temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                      const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, uint32_t end_bit, hipStream_t stream);
hipError_t launch_accel(const StepArgs& a, uint32_t nblocks, uint32_t sys, float4* out,
                        hipStream_t stream);
hipError_t launch_diag(const StepArgs& a, uint32_t nblocks, uint32_t sys, double* partial,
                       double* out, hipStream_t stream);

}  // namespace rainbow
