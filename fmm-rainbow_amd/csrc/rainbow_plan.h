// Planner constants and functions shared by the host runtime and the planner
// ABI.  Values are the reference's (src/lib.rs:10-23,
// src/rainbow_fmm.wgsl:9-16).
#pragma once

#include <cstdint>

namespace rainbow {

constexpr uint32_t kMinDepth = 4u;
constexpr uint32_t kMaxDepth = 6u;
constexpr uint32_t kMinParticles = 1u << 14;
constexpr uint32_t kMaxParticles = 1u << 23;
constexpr uint32_t kTargetParticlesPerLeaf = 8u;
constexpr float kWorldHalfExtent = 1.60f;
constexpr uint32_t kAdaptiveLeafCapacity = 16u;

struct Plan {
  uint32_t count = 0;
  uint32_t depth = 0;
  uint32_t cells = 0;
  float scale = 0.0f;
};

uint32_t level_cells(uint32_t level);
uint32_t level_offset(uint32_t level);
uint32_t total_cells(uint32_t depth);
uint32_t depth_for_particles(uint32_t particles);
float fixed_point_scale(uint32_t particles);
uint32_t clamp_particle_count(uint32_t requested);
uint64_t linear_work_bound(uint32_t particles, uint32_t depth);
// Unclamped plan, as the native harness builds it (native/main.rs:89-92).
Plan make_plan(uint32_t particles);

}  // namespace rainbow
