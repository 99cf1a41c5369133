// Calibration program for HBM byte counters: one plain float4 copy kernel of
// a known size.  Under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) the
// counter value of a copy_kernel dispatch divided by the bytes it must move
// (the argument, in MiB, each way) is the factor by which the counter over-
// or under-reports HBM traffic on this device; the step kernels' counters
// are divided by that factor.
//
//   bin/hbm_copy [MiB]      (default 512)
//
// Prints the bytes moved and the rate of the timed launches.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ a, float4* __restrict__ b,
                                                   size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

#define CHECK(x)                                                         \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));      \
      return 1;                                                          \
    }                                                                    \
  } while (0)

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 512u;
  if (mib == 0 || mib > 65536) {
    std::fprintf(stderr, "size must be 1..65536 MiB\n");
    return 2;
  }
  const size_t bytes = mib << 20;
  const size_t n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 1, bytes));
  CHECK(hipMemset(b, 0, bytes));
  const dim3 grid(8192), block(256);
  for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(copy_kernel, grid, block, 0, 0, a, b, n);
  CHECK(hipGetLastError());
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const int reps = 10;
  CHECK(hipEventRecord(e0));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copy_kernel, grid, block, 0, 0, a, b, n);
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  std::printf("copy_kernel: %zu bytes read + %zu bytes written per launch (%zu KiB each), %d launches, "
              "%.4f ms per launch, %.1f GB/s\n",
              bytes, bytes, bytes >> 10, 3 + reps, ms / reps, 2.0 * bytes * reps / (ms * 1e-3) / 1e9);
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return 0;
}
