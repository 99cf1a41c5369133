// Headless harness for the gfx950 FMM step, the counterpart of the
// reference's native/main.rs: plans a system, seeds it, prints the plan, the
// fields/moments FNV-1a digests of one solve without integrate, per-phase
// GPU times, the amortised time of a full step and a finiteness check over
// a particle slice -- the same report lines, read side by side.
//
//   rainbow_native [particles (default 2^20)] [steps (default 40)]
//                  [galaxy|plummer|file:<path>]
//
// As the reference harness (native/main.rs:423-434), the default scene is
// the galaxy seeded on the device (initialize, src/rainbow_fmm.wgsl:232-374);
// `plummer` seeds a Plummer sphere on the host instead (a = 0.1, cut at
// 1.5, seed 1); `file:<path>` reads `particles` reference-layout bodies
// (32 bytes each) from a file.  Physics constants: native/main.rs:57-82.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rainbow_fmm.h"

static int check(int rc, const char* what) {
  if (rc != RAINBOW_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, rainbow_last_error());
    std::exit(1);
  }
  return rc;
}

int main(int argc, char** argv) {
  const uint32_t count = argc > 1 ? static_cast<uint32_t>(std::strtoul(argv[1], nullptr, 10)) : (1u << 20);
  const uint32_t repeat = argc > 2 ? static_cast<uint32_t>(std::strtoul(argv[2], nullptr, 10)) : 40u;
  const bool plummer = argc > 3 && std::strcmp(argv[3], "plummer") == 0;
  const char* file = argc > 3 && std::strncmp(argv[3], "file:", 5) == 0 ? argv[3] + 5 : nullptr;
  if (count == 0u || repeat == 0u ||
      (argc > 3 && !plummer && file == nullptr && std::strcmp(argv[3], "galaxy") != 0)) {
    std::fprintf(stderr, "usage: rainbow_native [particles] [steps] [galaxy|plummer|file:<path>]\n");
    return 2;
  }
  std::vector<float> from_file;
  if (file != nullptr) {
    FILE* f = std::fopen(file, "rb");
    from_file.resize(8ull * count);
    const size_t got = f != nullptr ? std::fread(from_file.data(), 32, count, f) : 0u;
    if (f != nullptr) std::fclose(f);
    if (got != count) {
      std::fprintf(stderr, "%s: expected %u particles of 32 bytes\n", file, count);
      return 2;
    }
  }

  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
    std::fprintf(stderr, "no HIP device\n");
    return 1;
  }
  std::printf("adapter: %s (HIP, %s)\n", prop.name[0] != '\0' ? prop.name : "AMD GPU", prop.gcnArchName);

  uint8_t slots[4096];
  const uint32_t depth = rainbow_build_plans(count, slots);
  rainbow_params_t params;
  std::memcpy(&params, slots + depth * RAINBOW_PARAM_STRIDE, sizeof(params));
  std::printf("count=%u depth=%u cells=%u scale=%g\n", count, depth, params.total_cells,
              static_cast<double>(params.control_settings[2]));

  rainbow_hip_t* h = nullptr;
  check(rainbow_hip_create(count, &params, &h), "create");
  if (plummer) {
    std::vector<float> parts(8ull * count);
    check(rainbow_ic_plummer(count, 1u, 0.1, 1.5, params.gravity, parts.data()), "seed");
    check(rainbow_hip_upload_particles(h, parts.data(), parts.size() * sizeof(float)), "upload");
  } else if (file != nullptr) {
    check(rainbow_hip_upload_particles(h, from_file.data(), from_file.size() * sizeof(float)), "upload");
  } else {
    check(rainbow_hip_seed_galaxy(h), "seed galaxy");
  }
  std::printf("scene: %s\n", plummer ? "Plummer sphere (host seed)"
                                     : (file != nullptr ? "particles from file" : "galaxy (device seed, initialize)"));

  // Content digest of one field solve (native/main.rs:436-454).
  check(rainbow_hip_solve_fields(h), "solve");
  uint32_t fh = 0, mh = 0;
  check(rainbow_hip_digest(h, RAINBOW_FIELDS, &fh), "digest");
  check(rainbow_hip_digest(h, RAINBOW_MOMENTS, &mh), "digest");
  std::printf("fieldsHash=%08x momentsHash=%08x\n", fh, mh);

  // Warm up (native/main.rs:456-464).
  check(rainbow_hip_step(h, 300u), "warm-up");
  check(rainbow_hip_synchronize(h), "sync");

  // Per-phase GPU time of one step, events around each kernel.
  float phase[16];
  const int nphase = rainbow_hip_phase_ms(h, 5u, phase, 16u);
  if (nphase < 0) check(nphase, "phase timing");
  const char* fixed[3] = {"deposit", "upward", "top"};
  uint32_t lo_hi[16];
  const int ntree = rainbow_hip_tree_launches(h, lo_hi, 16u);
  if (ntree < 0) check(ntree, "tree launches");
  double total = 0.0;
  for (int i = 0; i < nphase; ++i) {
    char name[32];
    if (i < 3)
      std::snprintf(name, sizeof(name), "%s", fixed[i]);
    else if (i < 3 + ntree) {
      // lo | form << 8 (0: a run of levels, 1: M2L sums, 2: fields from them,
      // 3: M2L sums with the top levels), hi.
      const uint32_t lo = lo_hi[2 * (i - 3)] & 0xFFu, form = lo_hi[2 * (i - 3)] >> 8, hi = lo_hi[2 * (i - 3) + 1];
      const char* pre = form == 1u ? "tree_m2l_" : (form == 2u ? "tree_finish_" : (form == 3u ? "tree_m2l_top_" : "tree_"));
      if (lo == hi)
        std::snprintf(name, sizeof(name), "%sl%u", pre, lo);
      else
        std::snprintf(name, sizeof(name), "%sl%u-%u", pre, lo, hi);
    } else if (i == nphase - 2)
      std::snprintf(name, sizeof(name), "integrate");
    else
      std::snprintf(name, sizeof(name), "sort");
    std::printf("  ts %-20s %.4f ms\n", name, static_cast<double>(phase[i]));
    total += phase[i];
  }
  std::printf("  ts %-20s %.4f ms\n", "TOTAL", total);

  // Amortised time of a full step (native/main.rs:522-534): graph replays.
  float ms = 0.0f;
  check(rainbow_hip_time_steps(h, repeat, &ms), "time");
  std::printf("full solve: %.4f ms/step  (%u steps)  %.4g particle-steps/s\n", ms / repeat, repeat,
              static_cast<double>(count) * repeat / (ms * 1e-3));

  // Finiteness over a particle slice (native/main.rs:536-549).
  std::vector<float> out(8ull * count);
  check(rainbow_hip_download(h, RAINBOW_PARTICLES, out.data(), out.size() * sizeof(float)), "download");
  const size_t probe = std::min<size_t>(count, 8192u) * 8u;
  size_t non_finite = 0;
  float max_abs = 0.0f;
  for (size_t i = 0; i < probe; ++i) {
    if (std::isfinite(out[i]))
      max_abs = std::max(max_abs, std::fabs(out[i]));
    else
      ++non_finite;
  }
  std::printf("non-finite=%zu maxAbs=%.6f\n", non_finite, static_cast<double>(max_abs));
  rainbow_hip_destroy(h);
  return 0;
}
