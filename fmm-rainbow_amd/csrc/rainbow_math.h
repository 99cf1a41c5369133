// Correctly rounded f32 square root and inverse square root for the kernels,
// in fewer instructions than the compiler's general expansions.
//
// The kernels follow the reference's f32 expressions operation by operation:
// inverseSqrt(x) is taken as 1 / sqrt(x) with both operations correctly
// rounded (RN(1 / RN(sqrt(x)))), and length() as RN(sqrt(...)).  HIP's
// sqrtf and '/' give exactly that, but their expansions carry scaling and
// special-case handling for the whole float range (about 27 VALU
// instructions for 1 / sqrtf(x)).  Inside the domain [2^-94, FLT_MAX] the
// forms below give the same bits:
//   y = rsq(x)                       hardware estimate of 1/sqrt(x)
//   s = s0 + (x - s0^2) * y/2        s0 = x*y; one Newton step on sqrt, FMA
//   q = r + (1 - s r) r              r = rcp(s); one Newton step on 1/s, FMA
// s equals RN(sqrt(x)) and q equals RN(1 / s) for every float x in the
// domain: bin/math_check compares them with sqrtf and 1 / sqrtf over all
// 2^32 inputs on the device.  Outside the domain (zero, tiny, infinite or
// NaN arguments) a wave that holds such a lane takes the compiler's full
// expansion for those lanes, so the functions equal the reference forms
// for every input.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace rainbow {

constexpr float kFastRootMin = 0x1p-94f;

__device__ __forceinline__ bool fast_root_domain(float x) { return x >= kFastRootMin && x <= FLT_MAX; }

// s = RN(sqrt(x)) and y = rsq(x) for x in the domain.
__device__ __forceinline__ float sqrt_newton(float x, float& y) {
  y = __builtin_amdgcn_rsqf(x);
  const float s0 = x * y;
  const float h = 0.5f * y;
  return __builtin_fmaf(__builtin_fmaf(-s0, s0, x), h, s0);
}

// RN(1 / s) for s = RN(sqrt(x)), x in the domain.
__device__ __forceinline__ float recip_newton(float s) {
  const float r = __builtin_amdgcn_rcpf(s);
  return __builtin_fmaf(__builtin_fmaf(-s, r, 1.0f), r, r);
}

// The full expansions for the lanes outside the domain, kept out of line so
// that the compiler cannot evaluate them speculatively for every lane.
__device__ __attribute__((noinline)) float sqrt_slow(float x) { return sqrtf(x); }
__device__ __attribute__((noinline)) float rsqrt_slow(float x) { return 1.0f / sqrtf(x); }

// Equals sqrtf(x) for every float x.
__device__ __forceinline__ float sqrt_rn(float x) {
  float y;
  float s = sqrt_newton(x, y);
  const bool ok = fast_root_domain(x);
  if (!__all(ok)) {
    if (!ok) s = sqrt_slow(x);
  }
  return s;
}

// Equals 1.0f / sqrtf(x) for every float x.
__device__ __forceinline__ float rsqrt_rn(float x) {
  float y;
  float q = recip_newton(sqrt_newton(x, y));
  const bool ok = fast_root_domain(x);
  if (!__all(ok)) {
    if (!ok) q = rsqrt_slow(x);
  }
  return q;
}

// WGSL i32(f32) as the reference evaluates it for the fixed-point sums:
// truncation toward zero, saturated at the i32 range, NaN to 0.
__device__ __forceinline__ int32_t to_i32_sat_ref(float x) {
  if (!(x == x)) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x < -2147483648.0f) return static_cast<int32_t>(0x80000000u);
  return static_cast<int32_t>(x);
}

// The same in one instruction: the hardware conversion truncates, clamps
// out-of-range values to the i32 range and turns NaN into 0, so it equals
// to_i32_sat_ref for every float (bin/math_check compares them over all
// 2^32 inputs), where the C++ cast's range checks cost a branchy sequence.
__device__ __forceinline__ int32_t to_i32_sat(float x) {
  int32_t r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Unguarded forms, for the checker (exact only inside the domain).
__device__ __forceinline__ float sqrt_fast_unguarded(float x) {
  float y;
  return sqrt_newton(x, y);
}
__device__ __forceinline__ float rsqrt_fast_unguarded(float x) {
  float y;
  return recip_newton(sqrt_newton(x, y));
}

}  // namespace rainbow
