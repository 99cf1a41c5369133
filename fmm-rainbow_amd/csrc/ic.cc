// Seeded initial conditions on the host: uniform ball, Plummer sphere and
// Plummer clumps.  The reference seeds its own galaxy on the GPU
// (initialize, src/rainbow_fmm.wgsl:232-374); these generators produce the
// synthetic inputs of the benchmark workloads instead.  Generation is in f64
// from a SplitMix64 stream and rounded to f32 at the end; all bodies have
// the reference's equal mass 1/N (src/rainbow_fmm.wgsl:182).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "rainbow_fmm.h"

namespace rainbow {
int fail(int code, const char* fmt, ...);  // sets the thread's error message (rainbow_runtime.hip)
}

namespace {

constexpr double kPi = 3.14159265358979323846;

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // Uniform in [0, 1).
  double u01() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
  // Uniform in (0, 1].
  double u01_open0() { return static_cast<double>((next() >> 11) + 1u) * (1.0 / 9007199254740992.0); }
};

struct Body {
  double x[3];
  double v[3];
  double hue;
};

void isotropic(SplitMix64& rng, double len, double out[3]) {
  const double z = 2.0 * rng.u01() - 1.0;
  const double phi = 2.0 * kPi * rng.u01();
  const double ring = std::sqrt(std::max(0.0, 1.0 - z * z));
  out[0] = len * ring * std::cos(phi);
  out[1] = len * ring * std::sin(phi);
  out[2] = len * z;
}

// Plummer body of mass fraction-sampled radius (Aarseth, Henon & Wielen 1974),
// truncated at rmax, with isotropic velocity from q^2 (1-q^2)^(7/2).
void plummer_body(SplitMix64& rng, double a, double rmax, double gm, Body& b) {
  double r;
  do {
    const double u = rng.u01_open0();
    r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
  } while (!(r < rmax));
  isotropic(rng, r, b.x);
  double q;
  for (;;) {
    q = rng.u01();
    const double g = 0.1 * rng.u01();
    if (g < q * q * std::pow(1.0 - q * q, 3.5)) break;
  }
  const double ve = std::sqrt(2.0 * gm / std::sqrt(r * r + a * a));
  isotropic(rng, q * ve, b.v);
}

void recenter(std::vector<Body>& bodies, size_t first, size_t last, const double centre[3]) {
  double mx[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
  const double inv = 1.0 / static_cast<double>(last - first);
  for (size_t i = first; i < last; ++i)
    for (int k = 0; k < 3; ++k) {
      mx[k] += bodies[i].x[k] * inv;
      mv[k] += bodies[i].v[k] * inv;
    }
  for (size_t i = first; i < last; ++i)
    for (int k = 0; k < 3; ++k) {
      bodies[i].x[k] = bodies[i].x[k] - mx[k] + centre[k];
      bodies[i].v[k] -= mv[k];
    }
}

void write_out(const std::vector<Body>& bodies, void* out) {
  float* f = static_cast<float*>(out);
  for (size_t i = 0; i < bodies.size(); ++i) {
    f[8 * i + 0] = static_cast<float>(bodies[i].x[0]);
    f[8 * i + 1] = static_cast<float>(bodies[i].x[1]);
    f[8 * i + 2] = static_cast<float>(bodies[i].x[2]);
    f[8 * i + 3] = static_cast<float>(bodies[i].hue);
    f[8 * i + 4] = static_cast<float>(bodies[i].v[0]);
    f[8 * i + 5] = static_cast<float>(bodies[i].v[1]);
    f[8 * i + 6] = static_cast<float>(bodies[i].v[2]);
    f[8 * i + 7] = 0.0f;
  }
}

}  // namespace

extern "C" {

int rainbow_ic_uniform_ball(uint32_t n, uint64_t seed, double radius, void* out) {
  if (out == nullptr || n == 0u || !(radius > 0.0))
    return rainbow::fail(RAINBOW_E_INVALID, "uniform ball: need an output buffer, n >= 1 and radius > 0");
  SplitMix64 rng(seed);
  std::vector<Body> bodies(n);
  for (uint32_t i = 0; i < n; ++i) {
    const double r = radius * std::cbrt(rng.u01());
    isotropic(rng, r, bodies[i].x);
    bodies[i].v[0] = bodies[i].v[1] = bodies[i].v[2] = 0.0;
    bodies[i].hue = rng.u01();
  }
  write_out(bodies, out);
  return RAINBOW_OK;
}

int rainbow_ic_plummer(uint32_t n, uint64_t seed, double a, double rmax, double gravity, void* out) {
  if (out == nullptr || n == 0u || !(a > 0.0) || !(rmax > 0.0))
    return rainbow::fail(RAINBOW_E_INVALID, "Plummer sphere: need an output buffer, n >= 1, a > 0 and rmax > 0");
  SplitMix64 rng(seed);
  std::vector<Body> bodies(n);
  for (uint32_t i = 0; i < n; ++i) {
    plummer_body(rng, a, rmax, gravity * 1.0, bodies[i]);
    bodies[i].hue = rng.u01();
  }
  const double centre[3] = {0.0, 0.0, 0.0};
  recenter(bodies, 0, n, centre);
  write_out(bodies, out);
  return RAINBOW_OK;
}

int rainbow_ic_clumps(uint32_t n, uint64_t seed, uint32_t clumps, double a, double rcut_over_a,
                      double gravity, void* out) {
  if (out == nullptr || n == 0u || clumps == 0u || clumps > n || !(a > 0.0))
    return rainbow::fail(RAINBOW_E_INVALID, "clumps: need an output buffer, 1 <= clumps <= n and a > 0");
  SplitMix64 rng(seed);
  std::vector<double> w(clumps), centre(3u * clumps);
  double wsum = 0.0;
  for (uint32_t c = 0; c < clumps; ++c) {
    w[c] = std::pow(10.0, 2.0 * rng.u01());
    wsum += w[c];
    for (int k = 0; k < 3; ++k) centre[3u * c + k] = rng.u01() - 0.5;
  }
  std::vector<uint32_t> count(clumps);
  uint32_t assigned = 0, largest = 0;
  for (uint32_t c = 0; c < clumps; ++c) {
    count[c] = static_cast<uint32_t>(std::floor(static_cast<double>(n) * w[c] / wsum));
    if (count[c] == 0u) count[c] = 1u;
    assigned += count[c];
    if (w[c] > w[largest]) largest = c;
  }
  if (assigned > n) {
    // Only when many clumps were rounded up to one body; take the excess
    // from the largest clump.
    count[largest] -= assigned - n;
  } else {
    count[largest] += n - assigned;
  }
  std::vector<Body> bodies(n);
  size_t i = 0;
  for (uint32_t c = 0; c < clumps; ++c) {
    const size_t first = i;
    const double mc = static_cast<double>(count[c]) / static_cast<double>(n);
    for (uint32_t k = 0; k < count[c]; ++k, ++i) {
      plummer_body(rng, a, rcut_over_a * a, gravity * mc, bodies[i]);
      bodies[i].hue = rng.u01();
    }
    recenter(bodies, first, i, &centre[3u * c]);
  }
  write_out(bodies, out);
  return RAINBOW_OK;
}

}  // extern "C"
