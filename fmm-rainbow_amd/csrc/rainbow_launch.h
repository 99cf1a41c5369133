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

hipError_t launch_deposit(const StepArgs& a, uint32_t nblocks, hipStream_t stream, const Timing* t = nullptr);
hipError_t launch_upward(const StepArgs& a, uint32_t nblocks, hipStream_t stream, const Timing* t = nullptr);
hipError_t launch_top(const StepArgs& a, hipStream_t stream, const Timing* t = nullptr);
// Tree pass launches over levels lo..hi (2 <= lo <= hi <= 6):
//   kTreePassRun     M2L, L2L and near of the levels, parents' fields from
//                    memory below lo and recomputed from lo on;
//   kTreePassM2l     the M2L sums of the targets below each system's finest
//                    level (into its m2l buffer);
//   kTreePassFinish  after a kTreePassM2l launch over lo..hi-1: every
//                    target's field from those sums (the finest level
//                    evaluates its own), L2L and near.
constexpr uint32_t kTreePassRun = 0u, kTreePassM2l = 1u, kTreePassFinish = 2u;
// Block shapes of a step (results never depend on them; the autotuner
// times them):
//   integrate_threads  width of an integrate+deposit block over its
//                      1024-body deposit range: 1024 (one body per thread)
//                      or 512 (two per thread);
//   tree_chunks        the tree items' pair tile: chunks of 8 targets x 8
//                      sources evaluated per round, 8 (two per wave, 6
//                      blocks per CU), 4 (one per wave, half the terms'
//                      LDS, 8 blocks per CU) or 1 (wave items: each wave
//                      walks items of its own, one chunk at a time, with no
//                      block barrier; unfused launches only, a fused run
//                      of levels takes 8), or 0: per launch, wave items for
//                      a launch of at least kWaveItemsMin items (measured
//                      in the last recorded step), 8 otherwise.
struct Shape {
  uint32_t integrate_threads = 1024u;
  uint32_t tree_chunks = 8u;
};
inline bool shape_valid(const Shape& s) {
  return (s.integrate_threads == 1024u || s.integrate_threads == 512u) &&
         (s.tree_chunks == 8u || s.tree_chunks == 4u || s.tree_chunks == 1u || s.tree_chunks == 0u);
}
// Wave items pay off once a launch holds many rounds of resident blocks'
// items (on MI355X: batches of 262144-body systems run faster with block
// items up to 32 systems, about 20k items per step, and the 256-system
// batch of 18.6M bodies faster with wave items on its levels of 12.7k and
// 22.4k items).
constexpr uint32_t kWaveItemsMin = 8192u;
// Blocks of a single-level tree launch resident on the device at once.
uint32_t tree_single_level_capacity(bool compensated, uint32_t chunks);
// lead > 0 (kTreePassM2l, one system): the launch also runs the top
// levels (block 0) and the items of levels 2..depth-3 (blocks 1..lead-1);
// lo..hi are then the levels whose parents upward lists.  chunks: the pair
// tile (Shape::tree_chunks).
hipError_t launch_tree(const StepArgs& a, uint32_t kind, uint32_t lo, uint32_t hi, uint32_t nblocks, uint32_t lead,
                       uint32_t chunks, hipStream_t stream, const Timing* t = nullptr);
// Integrate plus the next step's deposit (into finest buffer parity ^ 1),
// one block of `threads` (Shape::integrate_threads) per deposit range.
hipError_t launch_integrate_deposit(const StepArgs& a, uint32_t ndeposit, uint32_t threads, hipStream_t stream,
                                    const Timing* t = nullptr);
// The counting sort's scatter into the staging arrays (after upward and top
// of a step run with a.sorting set; that step's integrate reads them).
hipError_t launch_sort(const StepArgs& a, uint32_t nintegrate, hipStream_t stream);
// The same scatter over given deposit ranges (1024 bodies each): the systems
// that joined, when only they sort.
hipError_t launch_sort_ranges(const StepArgs& a, const BodyBlock* deposit_ranges, uint32_t nranges,
                              hipStream_t stream);
// FP64 direct sum at `nt` target bodies given by caller index.  It works in
// one scratch buffer laid out by direct_layout(nt): the target list (the
// caller copies it in), each target's slot, the partial sums of the body
// chunks (direct_chunks(nt) of them, a fixed rule so that every caller gets
// the same summation order) and the result out[4 t ..] = {ax, ay, az, phi}.
constexpr uint32_t kDirectMaxTargetsPerPass = 4096u;  // targets staged in LDS per pass
constexpr uint32_t kDirectMaxChunks = 128u;
constexpr uint32_t kDirectBlocks = 2048u;  // target tiles x body chunks
constexpr uint32_t direct_chunks(uint32_t nt) {
  const uint32_t tiles = (nt + 255u) / 256u;
  const uint32_t c = tiles == 0u ? kDirectMaxChunks : kDirectBlocks / tiles;
  return c < 1u ? 1u : (c > kDirectMaxChunks ? kDirectMaxChunks : c);
}
struct DirectLayout {
  size_t targets, slots, partial, out, bytes;
  uint32_t nchunks;
};
constexpr size_t direct_align(size_t x) { return (x + 255u) & ~static_cast<size_t>(255u); }
constexpr DirectLayout direct_layout(uint32_t nt) {
  const uint32_t nchunks = direct_chunks(nt);
  const size_t slots = direct_align(4ull * nt);
  const size_t out = slots + direct_align(4ull * nt);
  const size_t partial = out + direct_align(32ull * nt);
  return DirectLayout{0u, slots, partial, out, partial + direct_align(32ull * nt * nchunks), nchunks};
}
constexpr size_t direct_scratch_bytes(uint32_t nt) { return direct_layout(nt).bytes; }
hipError_t launch_direct_f64(const SysDesc& d, uint32_t nt, void* scratch, hipStream_t stream);
// The reference's default scene (initialize, src/rainbow_fmm.wgsl:232-374)
// into a system's bodies, in caller order.
hipError_t launch_seed_galaxy(const SysDesc& d, hipStream_t stream);
// Interleaved particles (32 B each, caller order) to the pos/vel/perm
// streams of a system, and back.
hipError_t launch_split_particles(const SysDesc& d, const float4* in, hipStream_t stream);
hipError_t launch_join_particles(const SysDesc& d, float4* out, hipStream_t stream);
hipError_t launch_accel(const StepArgs& a, uint32_t nblocks, uint32_t sys, float4* out,
                        hipStream_t stream);
hipError_t launch_diag(const StepArgs& a, uint32_t nblocks, uint32_t sys, double* partial,
                       double* out, hipStream_t stream);

}  // namespace rainbow
