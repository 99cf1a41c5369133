// Exhaustive device check of the kernels' square-root forms (rainbow_math.h)
// against the compiler's correctly rounded sqrtf and 1 / sqrtf, over all
// 2^32 float bit patterns:
//   - inside [2^-94, FLT_MAX] the unguarded Newton forms must give the same
//     bits as sqrtf(x) and 1.0f / sqrtf(x);
//   - over all inputs the guarded sqrt_rn / rsqrt_rn must give the same bits
//     (any NaN matches any NaN);
//   - over all inputs the one-instruction saturating conversion to_i32_sat
//     must equal the reference form to_i32_sat_ref.
// Prints the mismatch counts; exit status 0 iff every count is zero.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "rainbow_math.h"

namespace {

__device__ __forceinline__ bool same(float a, float b) {
  return __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
}

// cnt[0]: unguarded sqrt in the domain, [1]: unguarded rsqrt in the domain,
// [2]: guarded sqrt anywhere, [3]: guarded rsqrt anywhere,
// [4]: unguarded rsqrt outside the domain (informational), [5]: inputs seen,
// [6]: i32 conversion anywhere.
__global__ __launch_bounds__(256) void k_check(unsigned long long* cnt) {
  unsigned long long c[7] = {0, 0, 0, 0, 0, 0, 0};
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t u = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < (1ull << 32); u += stride) {
    const float x = __uint_as_float(static_cast<uint32_t>(u));
    const float s_ref = sqrtf(x);
    const float q_ref = 1.0f / sqrtf(x);
    if (rainbow::fast_root_domain(x)) {
      c[0] += __float_as_uint(rainbow::sqrt_fast_unguarded(x)) != __float_as_uint(s_ref);
      c[1] += __float_as_uint(rainbow::rsqrt_fast_unguarded(x)) != __float_as_uint(q_ref);
    } else {
      c[4] += !same(rainbow::rsqrt_fast_unguarded(x), q_ref);
    }
    c[2] += !same(rainbow::sqrt_rn(x), s_ref);
    c[3] += !same(rainbow::rsqrt_rn(x), q_ref);
    c[6] += rainbow::to_i32_sat(x) != rainbow::to_i32_sat_ref(x);
    c[5] += 1u;
  }
  for (int k = 0; k < 7; ++k)
    if (c[k]) atomicAdd(&cnt[k], c[k]);
}

}  // namespace

int main() {
  unsigned long long* d = nullptr;
  if (hipMalloc(&d, 7 * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(d, 0, 7 * sizeof(unsigned long long)) != hipSuccess) {
    std::fprintf(stderr, "math_check: device allocation failed\n");
    return 2;
  }
  hipLaunchKernelGGL(k_check, dim3(16384), dim3(256), 0, nullptr, d);
  unsigned long long h[7] = {0, 0, 0, 0, 0, 0, 0};
  if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
    std::fprintf(stderr, "math_check: kernel failed\n");
    return 2;
  }
  (void)hipFree(d);
  std::printf("inputs %llu\n", h[5]);
  std::printf("sqrt  unguarded, [2^-94, FLT_MAX]: %llu mismatches\n", h[0]);
  std::printf("rsqrt unguarded, [2^-94, FLT_MAX]: %llu mismatches\n", h[1]);
  std::printf("sqrt_rn,  all inputs: %llu mismatches\n", h[2]);
  std::printf("rsqrt_rn, all inputs: %llu mismatches\n", h[3]);
  std::printf("rsqrt unguarded outside the domain: %llu mismatches (the guard handles these)\n", h[4]);
  std::printf("i32 conversion, all inputs: %llu mismatches\n", h[6]);
  return (h[5] == (1ull << 32) && h[0] == 0 && h[1] == 0 && h[2] == 0 && h[3] == 0 && h[6] == 0) ? 0 : 1;
}
