// Host-side launch wrappers of the step kernels (defined with the kernels).
#pragma once

#include <hip/hip_runtime.h>

#include "rainbow_device.h"

namespace rainbow {

hipError_t launch_clear(const StepArgs& a, uint32_t nblocks, hipStream_t stream);
hipError_t launch_deposit(const StepArgs& a, uint32_t nblocks, hipStream_t stream);
hipError_t launch_upward(const StepArgs& a, uint32_t nblocks, hipStream_t stream);
hipError_t launch_top(const StepArgs& a, hipStream_t stream);
hipError_t launch_tree(const StepArgs& a, uint32_t level, uint32_t nblocks, hipStream_t stream);
hipError_t launch_integrate(const StepArgs& a, uint32_t nblocks, hipStream_t stream);
hipError_t launch_accel(const StepArgs& a, uint32_t nblocks, uint32_t sys, float4* out,
                        hipStream_t stream);
hipError_t launch_diag(const StepArgs& a, uint32_t nblocks, uint32_t sys, double* partial,
                       double* out, hipStream_t stream);

}  // namespace rainbow
