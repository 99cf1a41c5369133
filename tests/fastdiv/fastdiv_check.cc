// Exhaustive host check of the kernels' division-by-constant forms.
//
// The integrate kernels replace a correctly rounded x / d by
//   q0 = x * r;  rem = fma(-q0, d, x);  q = fma(rem, r, q0)   (r = RN(1 / d))
// only where the value that is kept afterwards is the same for every float x.
// q can differ from x / d in the last bit for some x; what is checked here is
// the quantity the kernel actually uses after the division:
//   unit  : clamp(q, -0.99999, 0.99999) * 0.5 + 0.5   (position -> unit cube)
//   position: the same for every float bit pattern, infinities and NaN
//           included, with the first estimate x * r clamped to [-2, 2] as
//           the kernel does (fmax/fmin return the non-NaN operand)
//   clamp01: clamp(q, 0, 1)                            (ramps and weights)
// and two ramps of the integrate step in the exact form the kernel uses,
// where the dividend is first clamped to a range that leaves the clamped
// quotient unchanged (so no overflow or underflow reaches the correction):
//   excitation: clamp((e - 0.30) / (0.70 - 0.30), 0, 1) for every float e,
//               dividend clamped to [-1, 1]
//   density   : clamp((float(c) - 16) / 72, 0, 1) for every uint32 c,
//               dividend clamped to [-1, 73]
// Usage: fastdiv_check <d> <unit|clamp01|excitation|density>; prints the
// mismatch count of the final quantity over all 2^32 inputs (float bit
// patterns, or counts) and exits 0 iff it is zero.
// Built with -ffp-contract=off so each line rounds as written.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {
float from_bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
uint32_t to_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
float clampf(float v, float lo, float hi) { return std::fmin(std::fmax(v, lo), hi); }
}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) { std::fprintf(stderr, "usage: %s <d> <unit|position|clamp01|excitation|density>\n", argv[0]); return 2; }
  const float d = std::strtof(argv[1], nullptr);
  const bool position = std::strcmp(argv[2], "position") == 0;
  const bool unit = position || std::strcmp(argv[2], "unit") == 0;
  const bool excitation = std::strcmp(argv[2], "excitation") == 0;
  const bool density = std::strcmp(argv[2], "density") == 0;
  if (!unit && !excitation && !density && std::strcmp(argv[2], "clamp01") != 0) return 2;
  volatile float dv = d;
  const float r = 1.0f / dv;
  unsigned long long raw = 0, kept = 0;
  if (excitation || density) {
    // d must be the divisor the kernel uses: RN(0.70f - 0.30f) or 72.
#pragma omp parallel for reduction(+ : kept) schedule(static, 1 << 20)
    for (long long i = 0; i < (1ll << 32); ++i) {
      const float x = excitation ? from_bits(static_cast<uint32_t>(i)) - 0.30f
                                 : static_cast<float>(static_cast<uint32_t>(i)) - 16.0f;
      const float exact = clampf(x / dv, 0.0f, 1.0f);
      const float xc = excitation ? clampf(x, -1.0f, 1.0f) : clampf(x, -1.0f, 73.0f);
      const float q0 = xc * r;
      const float q = std::fmaf(std::fmaf(-q0, d, xc), r, q0);
      const float fast = clampf(q, 0.0f, 1.0f);
      if (to_bits(fast) != to_bits(exact)) ++kept;
    }
    std::printf("d=%.9g r=%.9g %s: kept values differing %llu\n", d, r, argv[2], kept);
    return kept == 0 ? 0 : 1;
  }
#pragma omp parallel for reduction(+ : raw, kept) schedule(static, 1 << 20)
  for (long long i = 0; i < (1ll << 32); ++i) {
    const float xi = from_bits(static_cast<uint32_t>(i));
    if (!position && !std::isfinite(xi)) continue;
    const float exact = xi / dv;
    const float x = xi;
    const float q0 = position ? clampf(x * r, -2.0f, 2.0f) : x * r;
    const float rem = std::fmaf(-q0, d, x);
    const float q = std::fmaf(rem, r, q0);
    if (to_bits(q) == to_bits(exact)) continue;
    ++raw;
    const float a = unit ? clampf(exact, -0.99999f, 0.99999f) * 0.5f + 0.5f : clampf(exact, 0.0f, 1.0f);
    const float b = unit ? clampf(q, -0.99999f, 0.99999f) * 0.5f + 0.5f : clampf(q, 0.0f, 1.0f);
    if (to_bits(a) != to_bits(b)) ++kept;
  }
  std::printf("d=%.9g r=%.9g %s: quotients differing %llu, kept values differing %llu\n", d, r,
              argv[2], raw, kept);
  return kept == 0 ? 0 : 1;
}
