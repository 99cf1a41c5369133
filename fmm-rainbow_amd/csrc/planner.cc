// Planner ABI: octree sizing, fixed-point scale, work bound, Params image,
// plan table and FNV-1a digest.  Pure functions; no allocation, no errors.
//
// Semantics follow src/lib.rs:10-157 (planner and its extern "C" exports)
// and native/main.rs:38-128, 152-163 (Params serialisation, plan table,
// digest).  Constants are the reference's (src/lib.rs:10-23).

#include <cmath>
#include <cstring>

#include "rainbow_fmm.h"
#include "rainbow_plan.h"

namespace rainbow {

uint32_t level_cells(uint32_t level) { return 1u << (3u * level); }

uint32_t level_offset(uint32_t level) { return (level_cells(level) - 1u) / 7u; }

uint32_t total_cells(uint32_t depth) { return level_offset(depth + 1u); }

// src/lib.rs:45-52: smallest depth in [4, 6] with 8^depth >= ceil(n / 8).
uint32_t depth_for_particles(uint32_t particles) {
  const uint32_t target_cells =
      particles / kTargetParticlesPerLeaf +
      (particles % kTargetParticlesPerLeaf != 0u ? 1u : 0u);
  uint32_t depth = kMinDepth;
  while (depth < kMaxDepth && level_cells(depth) < target_cells) depth += 1u;
  return depth;
}

// src/lib.rs:59-69: 2^floor(log2(0.82 (2^31 - 1) / max(n * 1.6f, 1))),
// evaluated in f64 and rounded to f32.  WORLD_HALF_EXTENT is the f32 1.6
// widened to f64, as in the reference.
float fixed_point_scale(uint32_t particles) {
  const double headroom = 0.82;
  const double worst =
      static_cast<double>(particles) * static_cast<double>(kWorldHalfExtent);
  const double max_scale = headroom * 2147483647.0 / (worst > 1.0 ? worst : 1.0);
  const double exponent = std::floor(std::log2(max_scale));
  return static_cast<float>(std::pow(2.0, exponent));
}

uint32_t clamp_particle_count(uint32_t requested) {
  if (requested < kMinParticles) return kMinParticles;
  if (requested > kMaxParticles) return kMaxParticles;
  return requested;
}

// src/lib.rs:83-86.
uint64_t linear_work_bound(uint32_t particles, uint32_t depth) {
  const uint64_t cells = total_cells(depth);
  return static_cast<uint64_t>(particles) * 3u + cells * (189u + 8u + 2u);
}

static uint32_t clamp_depth(uint32_t depth) {
  return depth < kMinDepth ? kMinDepth : (depth > kMaxDepth ? kMaxDepth : depth);
}

Plan make_plan(uint32_t particles) {
  Plan p;
  p.count = particles;
  p.depth = depth_for_particles(particles);
  p.cells = total_cells(p.depth);
  p.scale = fixed_point_scale(particles);
  return p;
}

}  // namespace rainbow

using namespace rainbow;

static const char kAlgorithm[] =
    "Sparse p=1 Cartesian FMM / BLDTT-inspired adaptive hierarchy";

// The kernel source compiled into this library, embedded at build time
// (the counterpart of `include_str!("rainbow_fmm.wgsl")`, src/lib.rs:7).
extern "C" const uint8_t rainbow_kernel_source_begin[];
extern "C" const uint8_t rainbow_kernel_source_end[];

extern "C" {

uint32_t rainbow_api_version(void) { return 1u; }

const uint8_t* rainbow_shader_ptr(void) { return rainbow_kernel_source_begin; }

size_t rainbow_shader_len(void) {
  return static_cast<size_t>(rainbow_kernel_source_end -
                             rainbow_kernel_source_begin);
}

const uint8_t* rainbow_algorithm_ptr(void) {
  return reinterpret_cast<const uint8_t*>(kAlgorithm);
}

size_t rainbow_algorithm_len(void) { return sizeof(kAlgorithm) - 1u; }

uint32_t rainbow_depth_for_particles(uint32_t particles) {
  return depth_for_particles(clamp_particle_count(particles));
}

uint32_t rainbow_level_cells(uint32_t level) {
  return level <= kMaxDepth ? level_cells(level) : 0u;
}

uint32_t rainbow_level_offset(uint32_t level) {
  return level <= kMaxDepth + 1u ? level_offset(level) : 0u;
}

uint32_t rainbow_total_cells(uint32_t depth) {
  return total_cells(clamp_depth(depth));
}

uint32_t rainbow_clamp_particle_count(uint32_t requested) {
  return clamp_particle_count(requested);
}

float rainbow_fixed_point_scale(uint32_t particles) {
  return fixed_point_scale(clamp_particle_count(particles));
}

uint32_t rainbow_linear_work_lo(uint32_t particles, uint32_t depth) {
  return static_cast<uint32_t>(
      linear_work_bound(clamp_particle_count(particles), clamp_depth(depth)));
}

uint32_t rainbow_plan_depth(uint32_t particles) {
  return depth_for_particles(particles);
}

float rainbow_plan_fixed_point_scale(uint32_t particles) {
  return fixed_point_scale(particles);
}

uint64_t rainbow_plan_linear_work(uint32_t particles, uint32_t depth) {
  return linear_work_bound(particles, depth);
}

// app.js:42-43 (BASE_DT, REFERENCE_FRAME_MS) and :716-719: JavaScript
// arithmetic in double, stored as an f32 by DataView.setFloat32.
float rainbow_frame_dt(double frame_interval_ms) {
  const double base_dt = 0.0009;
  const double reference_frame_ms = 1000.0 / 60.0;
  const double interval = frame_interval_ms < 100.0 ? frame_interval_ms : 100.0;
  return static_cast<float>(base_dt * (interval / reference_frame_ms));
}

void rainbow_write_params(uint32_t count, uint32_t depth, uint32_t level,
                          uint32_t level_offset_, uint32_t level_dim,
                          uint32_t parent_offset, uint32_t parent_dim,
                          uint32_t cells, float scale, rainbow_params_t* out) {
  if (out == nullptr) return;
  std::memset(out, 0, sizeof(*out));
  out->particle_count = count;
  out->depth = depth;
  out->level = level;
  out->level_offset = level_offset_;
  out->level_dim = level_dim;
  out->parent_offset = parent_offset;
  out->parent_dim = parent_dim;
  out->total_cells = cells;
  // Harness constants, native/main.rs:57-82.
  out->dt = 0.0009f;
  out->gravity = 0.34f;
  out->softening = 0.012f;
  out->point_size = 5.0f;
  out->time = 0.0f;
  out->aspect = 1.0f;
  out->exposure = 0.84f;
  out->speed = 1.0f;
  out->control_settings[0] = 0.7f;
  out->control_settings[1] = 0.22f;
  out->control_settings[2] = scale;
  out->control_settings[3] = 0.0f;
}

uint32_t rainbow_build_plans(uint32_t count, uint8_t out_slots[4096]) {
  const Plan p = make_plan(count);
  if (out_slots == nullptr) return p.depth;
  std::memset(out_slots, 0, RAINBOW_PARAM_SLOTS * RAINBOW_PARAM_STRIDE);
  for (uint32_t l = 0; l <= p.depth; ++l) {
    const uint32_t parent = l > 0u ? l - 1u : 0u;
    rainbow_params_t prm;
    rainbow_write_params(count, p.depth, l, level_offset(l), 1u << l,
                         level_offset(parent), 1u << parent, p.cells, p.scale,
                         &prm);
    std::memcpy(out_slots + l * RAINBOW_PARAM_STRIDE, &prm, sizeof(prm));
  }
  for (uint32_t l = 0; l < p.depth; ++l) {
    rainbow_params_t prm;
    rainbow_write_params(count, p.depth, l, level_offset(l), 1u << l,
                         level_offset(l + 1u), 1u << (l + 1u), p.cells, p.scale,
                         &prm);
    std::memcpy(out_slots + (8u + l) * RAINBOW_PARAM_STRIDE, &prm, sizeof(prm));
  }
  return p.depth;
}

uint32_t rainbow_fnv1a(const void* data, size_t bytes) {
  const uint8_t* b = static_cast<const uint8_t*>(data);
  uint32_t h = 2166136261u;
  for (size_t i = 0; i < bytes; ++i) {
    h ^= b[i];
    h *= 16777619u;
  }
  return h;
}

}  // extern "C"

static_assert(sizeof(rainbow_params_t) == RAINBOW_PARAM_BYTES,
              "Params must be 128 bytes (src/rainbow_fmm.wgsl:18-50)");
static_assert(offsetof(rainbow_params_t, dt) == 32, "f32 block at byte 32");
static_assert(offsetof(rainbow_params_t, control_point) == 80, "vec4 at 80");
static_assert(offsetof(rainbow_params_t, control_axis_mode) == 96, "vec4 at 96");
static_assert(offsetof(rainbow_params_t, control_settings) == 112, "vec4 at 112");
