// gfx950 kernels of the p=1 Cartesian uniform-octree FMM step.
//
// One step runs these passes over a batch of systems (a plain solve is a
// batch of one), launched in order on one stream and captured in HIP
// graphs:
//
//   upward     M2M: one block per 8x8x8 finest-cell subtree merges the
//              straggler sums of the last deposit, sums three levels in LDS,
//              writes each cell's source record (centre of mass and G m),
//              lists the internal cells (count > 16), zeroes the fields of
//              cells that stopped being M2L targets and the finest-level
//              buffer of the next deposit
//   top        the levels above the subtrees and the fields of level 1,
//              one block per system (for one system it runs inside the
//              first tree launch)
//   tree       for levels 2..depth, M2L, L2L and the near field of every
//              M2L target (a non-empty child of an internal cell), four
//              waves per internal parent; in runs of levels, or in two
//              launches (the M2L sums above the finest level, then every
//              level's fields from them)
//   integrate  adaptive descent, local expansion, own-leaf monopole,
//     +deposit symplectic Euler kick-drift and the soft absorbing boundary,
//              then the next step's P2M of the new positions: finest-cell
//              binning and fixed-point sums pre-aggregated per block in an
//              LDS hash table; a block stores the cells of its own Morton
//              range and adds the others into a straggler accumulator with
//              integer atomics (exact in any order)
// A solve's first deposit runs alone (k_deposit).  A sorting step (every few
// steps) also reorders the bodies by Morton code: a scatter into staging
// arrays after top, from which that step's integrate reads them.
//
// The reference runs the same physics as 2d+5 dispatches
// (native/main.rs:411-421): clear_cells, deposit, upward per level, m2l_all,
// downward per level, near_all, integrate (src/rainbow_fmm.wgsl:394-771).
//
// Floating point follows the reference expressions operation for operation
// (src/rainbow_fmm.wgsl:100-203, 457-771), with correctly rounded division
// and square root and no contraction (-ffp-contract=off), so the fields and
// particles equal a sequential f32 evaluation of the WGSL bit for bit.
// Sums keep the reference order: M2L visits sources pz, py, px, cz, cy, cx
// (src/rainbow_fmm.wgsl:485-517), then the L2L term is added
// (src/rainbow_fmm.wgsl:610), then the near sources dz, dy, dx
// (src/rainbow_fmm.wgsl:551-570).  Empty sources are skipped: the reference
// adds +0 for them, and a running sum that starts at +0 is never -0, so
// adding +0 never changes it.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <atomic>

#include "rainbow_device.h"
#include "rainbow_launch.h"
#include "rainbow_math.h"

namespace rainbow {
namespace {

constexpr float kWorldHalfExtent = 1.60f;
constexpr int kLeafCapacity = 16;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;
// Rank flag: the body's index counts from the end of its cell's bodies
// deposited by the owning block.
constexpr uint32_t kRankExtra = 0x80000000u;

// Global-memory view of a buffer pointer: accesses through it are global
// (not flat) loads, stores and atomics, addressed from a scalar base.  The
// host pass of this file (which only type-checks the device functions) sees
// plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <class T>
using gptr = T*;
#endif
template <class T>
__device__ __forceinline__ gptr<T> G(T* p) {
  return (gptr<T>)p;
}

// An empty asm taking uniform values in SGPRs: each value is loaded (and
// waited for) before this point.  Holding a descriptor's words before the
// branch on its live flag makes them load in one round trip; otherwise the
// compiler loads the flag alone and the other words after the branch.
__device__ __forceinline__ void in_sgpr(uint32_t x) { asm volatile("" ::"s"(x)); }
__device__ __forceinline__ void in_sgpr(float x) { asm volatile("" ::"s"(x)); }
__device__ __forceinline__ void in_sgpr(const void* p) { asm volatile("" ::"s"(p)); }

__device__ __forceinline__ uint32_t level_off(uint32_t level) {
  return ((1u << (3u * level)) - 1u) / 7u;
}

__device__ __forceinline__ uint32_t lin(uint32_t x, uint32_t y, uint32_t z, uint32_t dim) {
  return x + dim * (y + dim * z);
}

// Inverse of lin for dim = 2^lg (shifts, no integer division).
// Byte flags of the finest level (the target/non-empty flags in `active`
// and `xflag`) are stored subtree-major: the 8^3 cells under each level-(D-3)
// cell contiguously, x fastest inside.  Upward's waves (one thread per cell
// of a subtree) then read 64 contiguous bytes, not eight 8-byte pieces of
// rows.
__device__ __forceinline__ uint32_t flag_index(uint32_t x, uint32_t y, uint32_t z, uint32_t depth) {
  const uint32_t rd = depth - 3u;
  const uint32_t root = (x >> 3) + ((y >> 3) << rd) + ((z >> 3) << (2u * rd));
  return (root << 9) | (x & 7u) | ((y & 7u) << 3) | ((z & 7u) << 6);
}

__device__ __forceinline__ void unlin(uint32_t i, uint32_t lg, uint32_t& x, uint32_t& y, uint32_t& z) {
  const uint32_t m = (1u << lg) - 1u;
  x = i & m;
  y = (i >> lg) & m;
  z = i >> (2u * lg);
}

// 1 / dim for dim a power of two (exact), so x * inv_pow2(dim) equals the
// correctly rounded x / dim for every x: both round the same real number.
__device__ __forceinline__ float inv_pow2(uint32_t dim) {
  return __uint_as_float((127u - (31u - static_cast<uint32_t>(__clz(dim)))) << 23);
}

// cell_center, src/rainbow_fmm.wgsl:136-142 (one axis); 2 / dim is exact.
__device__ __forceinline__ float center1(uint32_t c, uint32_t dim) {
  return ((static_cast<float>(c) + 0.5f) * (2.0f * inv_pow2(dim)) - 1.0f) * kWorldHalfExtent;
}

// position_cell, src/rainbow_fmm.wgsl:144-152, split in the level-independent
// part clamp(p / 1.6) * 0.5 + 0.5 and the per-level floor(u * dim).  The
// product by a power of two is exact, so this equals the reference at every
// level.  p / 1.6f is formed as q = p * 0.625 (= RN(1 / 1.6f)) corrected once
// by its exact remainder: the clamped result equals the one from the
// correctly rounded quotient for every float p (all 2^32 inputs are checked
// on the host by the tests), at a third of a division's instructions.
// x / d for a constant d, r = RN(1 / d): the product x * r corrected once by
// its exact remainder (three instructions instead of a division).  Used only
// where the value kept afterwards is shown equal to the correctly rounded
// quotient's for every input.
__device__ __forceinline__ float div_corrected(float x, float d, float r) {
  const float q0 = x * r;
  return __builtin_fmaf(__builtin_fmaf(-q0, d, x), r, q0);
}

// The first estimate p * 0.625 is clamped to [-2, 2] before the correction:
// that leaves the clamped result of every finite p unchanged and keeps
// infinities out of the remainder (inf - inf), so +inf lands in the last
// cell and -inf and NaN (quieted by the product; fmax/fmin then return the
// other operand) in the first, as with the division.
__device__ __forceinline__ float unit_coord(float p) {
  constexpr float kInv = 0.625f;  // RN(1 / 1.6f)
  const float q0 = fminf(fmaxf(p * kInv, -2.0f), 2.0f);
  float n = __builtin_fmaf(__builtin_fmaf(-q0, kWorldHalfExtent, p), kInv, q0);
  n = fminf(fmaxf(n, -0.99999f), 0.99999f);
  return n * 0.5f + 0.5f;
}

__device__ __forceinline__ uint32_t cell1(float u, uint32_t dim) {
  return static_cast<uint32_t>(floorf(u * static_cast<float>(dim)));
}

// WGSL i32(f32) (truncation, saturated at the i32 range) is to_i32_sat of
// rainbow_math.h.

__device__ __forceinline__ float clampf(float x, float lo, float hi) {
  return fminf(fmaxf(x, lo), hi);
}

__device__ __forceinline__ void load_field(gptr<const float4> fields, uint32_t cell, float4& a, float4& j0,
                                           float4& j1) {
  a = fields[3u * cell + 0u];
  j0 = fields[3u * cell + 1u];
  j1 = fields[3u * cell + 2u];
}

// The finest level of the step's parity (see SysDesc::finest_alt).
// Whether this step sorts system d: every system on a sorting step, a
// system that just joined a running set on its first step.
__device__ __forceinline__ bool sorts(const StepArgs& a, const SysDesc& d) {
  return a.sorting != 0u || d.sort_now != 0u;
}

__device__ __forceinline__ gptr<int4> finest_of(const SysDesc& d, uint32_t parity) {
  return parity ? G(d.finest_alt) : G(d.moments) + level_off(d.depth);
}

// Body count of a cell (level, index within the level) at a parity.
__device__ __forceinline__ int cell_count(const SysDesc& d, uint32_t parity, uint32_t level, uint32_t idx) {
  return level == d.depth ? finest_of(d, parity)[idx].x : G(d.moments)[level_off(level) + idx].x;
}

// Interleave the low 10 bits of v with two zero bits between each.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3FFu;
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// ------------------------------------------------------------------------
// deposit (P2M), src/rainbow_fmm.wgsl:408-422
// ------------------------------------------------------------------------

// LDS of a deposit block: the open-addressing hash of its cells.
// Inclusive prefix sum over the 64 lanes of a wave (wrapping u32 adds): four
// DPP row shifts sum within each row of 16 lanes, then the row broadcasts of
// lanes 15 and 31 carry the rows below into the rows above.
__device__ __forceinline__ uint32_t wave_prefix_sum(uint32_t x) {
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x111, 0xF, 0xF, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x112, 0xF, 0xF, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x114, 0xF, 0xF, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x118, 0xF, 0xF, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x142, 0xA, 0xF, false));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x143, 0xC, 0xF, false));
  return x;
}

struct DepositLds {
  uint32_t keys[kDepositSlots];  // after the flush: rank base per slot
  int4 vals[kDepositSlots];      // {count, sum_x, sum_y, sum_z} per slot
  uint32_t used[kDepositBodies]; // the slots that received a key
  uint32_t nused;
};

// Deposit of one block's bodies first + k * kThreads + thread (positions
// p[k], k < kDepositBodies / kThreads) into the finest buffer `parity` of
// system d; writes the bodies' ranks when `ranks`.  All threads of the
// block call it.
// Empties the deposit table (every thread of the block calls it; the
// deposit's first barrier orders it before the inserts).
template <uint32_t kThreads>
__device__ __forceinline__ void deposit_clear(DepositLds& L) {
  for (uint32_t s = threadIdx.x; s < kDepositSlots; s += kThreads) {
    L.keys[s] = kEmptyKey;
    L.vals[s] = make_int4(0, 0, 0, 0);
  }
  if (threadIdx.x == 0u) L.nused = 0u;
}

// kCleared: the caller emptied the table earlier (deposit_clear).
template <uint32_t kThreads, bool kCleared = false>
__device__ void deposit_block(const StepArgs& a, const SysDesc& d, uint32_t first,
                              const float4 (&p)[kDepositBodies / kThreads], uint32_t parity, bool ranks,
                              DepositLds& L, const uint2* own_lds = nullptr) {
  constexpr uint32_t kPer = kDepositBodies / kThreads;
  uint32_t* const keys = L.keys;
  int4* const vals = L.vals;
  if (!kCleared) deposit_clear<kThreads>(L);
  __syncthreads();
  const uint32_t dim = 1u << d.depth;
  const uint32_t lane = threadIdx.x & 63u;
  // Per body of this thread: its hash slot and its index among the block's
  // bodies of that cell (used when the step sorts).
  uint32_t my_slot[kPer], my_rank[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t i = first + k * kThreads + threadIdx.x;
    const bool valid = i < d.n;
    uint32_t key = kEmptyKey;
    int v[4] = {0, 0, 0, 0};
    if (valid) {
      const float4 q = p[k];
      key = lin(cell1(unit_coord(q.x), dim), cell1(unit_coord(q.y), dim), cell1(unit_coord(q.z), dim), dim);
      v[0] = 1;
      v[1] = to_i32_sat(rintf(q.x * d.scale));
      v[2] = to_i32_sat(rintf(q.y * d.scale));
      v[3] = to_i32_sat(rintf(q.z * d.scale));
    }
    // Bodies are kept in Morton order, so equal keys form runs across the
    // wave.  A run's sums are the differences of the wave's inclusive prefix
    // sums at its last lane and before its first lane (wrapping i32 sums,
    // exact in any order); its last lane adds them to the table.  Its count
    // is its length (lanes without a body form runs of their own key).
    const uint32_t prev_key = __shfl_up(key, 1u);
    const unsigned long long heads = __ballot(lane == 0u || key != prev_key);
    const uint32_t start = 63u - static_cast<uint32_t>(__clzll(heads & ((2ull << lane) - 1ull)));
    {
      uint32_t incl[4];
#pragma unroll
      for (int c = 1; c < 4; ++c) incl[c] = wave_prefix_sum(static_cast<uint32_t>(v[c]));
      const int src = static_cast<int>(start == 0u ? 0u : start - 1u);
#pragma unroll
      for (int c = 1; c < 4; ++c) {
        const uint32_t before = static_cast<uint32_t>(__shfl(static_cast<int>(incl[c]), src));
        v[c] = static_cast<int>(incl[c] - (start == 0u ? 0u : before));
      }
      v[0] = valid ? static_cast<int>(lane - start + 1u) : 0;
    }
    // Heads after this lane; the lane before the next head ends the run.
    const unsigned long long later = lane == 63u ? 0ull : (heads >> (lane + 1u));
    const bool tail = lane == 63u || (later & 1ull) != 0ull;
    uint32_t h = 0u, run_base = 0u;
    if (valid && tail) {
      // Open addressing; at most kDepositBodies distinct keys in a table of
      // 2 kDepositBodies slots, so an empty slot always exists.
      // The lane that places a key lists its slot for the flush.
      h = (key * 2654435761u) >> (32u - kDepositSlotBits);
      for (;;) {
        const uint32_t prev = atomicCAS(&keys[h], kEmptyKey, key);
        if (prev == kEmptyKey) {
          L.used[atomicAdd(&L.nused, 1u)] = h;
          break;
        }
        if (prev == key) break;
        h = (h + 1u) & (kDepositSlots - 1u);
      }
      run_base = static_cast<uint32_t>(atomicAdd(&vals[h].x, v[0]));
      atomicAdd(&vals[h].y, v[1]);
      atomicAdd(&vals[h].z, v[2]);
      atomicAdd(&vals[h].w, v[3]);
    }
    // The run's last lane holds its slot and its first index in the slot.
    const uint32_t tail_lane = later == 0ull ? 63u : lane + static_cast<uint32_t>(__ffsll(static_cast<long long>(later))) - 1u;
    my_slot[k] = static_cast<uint32_t>(__shfl(static_cast<int>(h), static_cast<int>(tail_lane)));
    my_rank[k] = static_cast<uint32_t>(__shfl(static_cast<int>(run_base), static_cast<int>(tail_lane))) + (lane - start);
  }
  __syncthreads();
  // Cells in this block's Morton range belong to it alone: their sums are
  // stored.  Any other cell's sums go to `extra` with atomics (bodies that
  // moved since the last sort, and cells shared with other blocks);
  // upward adds `extra` into the finest level.
  const uint2 own = own_lds != nullptr ? *own_lds : G(d.own)[first / kDepositBodies];
  const gptr<int4> finest = finest_of(d, parity);
  const uint32_t nused = L.nused;
  for (uint32_t u = threadIdx.x; u < nused; u += kThreads) {
    const uint32_t s = L.used[u];
    const uint32_t key = keys[s];
    const int4 sums = vals[s];
    uint32_t x, y, z;
    unlin(key, d.depth, x, y, z);
    const uint32_t morton = spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
    if (morton >= own.x && morton < own.y) {
      finest[key] = sums;
      keys[s] = 0u;
      continue;
    }
    // count and sum_x share one 64-bit atomic: the count never carries out
    // of its 32 bits, and the high word wraps as an i32 sum.  The old count
    // places this block's bodies after those already in `extra`.
    const unsigned long long packed = static_cast<unsigned long long>(static_cast<uint32_t>(sums.x)) |
                                      (static_cast<unsigned long long>(static_cast<uint32_t>(sums.y)) << 32);
    unsigned long long* const ce = reinterpret_cast<unsigned long long*>(G(d.extra) + key);
    G(d.xflag)[flag_index(x, y, z, d.depth)] = 1u;
    if (ranks) {
      const unsigned long long old = atomicAdd(ce, packed);
      keys[s] = kRankExtra | static_cast<uint32_t>(old);
    } else {
      (void)atomicAdd(ce, packed);  // no return value needed: the hardware need not wait for it
    }
    atomicAdd(reinterpret_cast<int*>(G(d.extra) + key) + 2, sums.z);
    atomicAdd(reinterpret_cast<int*>(G(d.extra) + key) + 3, sums.w);
  }
  if (!ranks) return;
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t i = first + k * kThreads + threadIdx.x;
    if (i < d.n) G(d.rank)[i] = keys[my_slot[k]] + my_rank[k];
  }
}

// Workgroups are dealt to the 8 XCDs in turn (workgroup b to XCD b mod 8),
// each XCD with its own L2.  Giving XCD x a contiguous eighth of the body
// tiles (tiles are in Morton order within each system) keeps the tree
// cells its tiles read, and the finest cells they deposit into, in its L2.
// A bijection of [0, n): the tiles of XCD x start at x q + min(x, r),
// q = n / 8, r = n mod 8.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  const uint32_t x = b & 7u, k = b >> 3, q = n >> 3, r = n & 7u;
  return x * q + min(x, r) + k;
}

// Zeroes the work-list counters of the step whose upward follows (their
// last reader, the previous tree pass, has finished).
__device__ __forceinline__ void reset_lists(const StepArgs& a) {
  if (blockIdx.x == 0u)
    for (uint32_t i = threadIdx.x; i < kLevels * kStripes; i += blockDim.x) G(a.list_count)[i] = 0u;
}

// Standalone deposit (a solve's first pass; a step normally deposits at the
// end of the previous step's integrate).
__global__ __launch_bounds__(kDepositThreads) void k_deposit(StepArgs a) {
  __shared__ DepositLds L;
  reset_lists(a);
  const BodyBlock blk = a.deposit_blocks[xcd_tile(blockIdx.x, gridDim.x)];
  const SysDesc d = a.sys[blk.sys];
  if (!d.live) return;
  float4 p[kDepositBodies / kDepositThreads];
#pragma unroll
  for (uint32_t k = 0; k < kDepositBodies / kDepositThreads; ++k) {
    const uint32_t i = blk.first + k * kDepositThreads + threadIdx.x;
    p[k] = i < d.n ? G(d.pos)[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  deposit_block<kDepositThreads>(a, d, blk.first, p, a.parity, sorts(a, d), L);
}

// Morton code of the finest cell of a position.
__device__ __forceinline__ uint32_t morton_of(const float4 p, uint32_t dim) {
  return spread3(cell1(unit_coord(p.x), dim)) | (spread3(cell1(unit_coord(p.y), dim)) << 1) |
         (spread3(cell1(unit_coord(p.z), dim)) << 2);
}

// Inverse of spread3 on 3k-bit codes: every third bit of v.
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
  v &= 0x49249249u;
  v = (v ^ (v >> 2)) & 0xC30C30C3u;
  v = (v ^ (v >> 4)) & 0x0F00F00Fu;
  v = (v ^ (v >> 8)) & 0xFF0000FFu;
  v = (v ^ (v >> 16)) & 0x000003FFu;
  return v;
}

// Scatter of the counting sort (bodies in Morton order of their finest
// cell): body i goes to slot sstart[subtree] + cstart[cell] + its index in
// the cell, in the staging buffers.
// kThreads bodies per block from `blocks`: the integrate ranges (every
// system, 256 bodies each) on a sorting step, or the deposit ranges of the
// systems that joined (1024 each) when only they sort.
template <uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void k_scatter(StepArgs a, const BodyBlock* blocks) {
  const BodyBlock blk = blocks[xcd_tile(blockIdx.x, gridDim.x)];
  const SysDesc d = a.sys[blk.sys];
  if (!d.live || !sorts(a, d)) return;
  const uint32_t i = blk.first + threadIdx.x;
  if (i >= d.n) return;
  const float4 p = G(d.pos)[i];
  const uint32_t dim = 1u << d.depth;
  const uint32_t cx = cell1(unit_coord(p.x), dim), cy = cell1(unit_coord(p.y), dim), cz = cell1(unit_coord(p.z), dim);
  const uint32_t m = spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2);
  const uint32_t r = G(d.rank)[i];
  // Bodies flagged in the rank follow the cell's bodies stored by its
  // owning block.
  const gptr<uint32_t> cs = (r & kRankExtra) ? G(d.xstart) : G(d.cstart);
  const uint32_t dst = G(d.sstart)[m >> 9] + cs[m] + (r & ~kRankExtra);
  if (dst >= d.n) return;  // unreachable when the counts are consistent
  G(d.tpos)[dst] = p;
  G(d.tvel)[dst] = G(d.vel)[i];
  G(d.tperm)[dst] = G(d.perm)[i];
}

// ------------------------------------------------------------------------
// upward (M2M), src/rainbow_fmm.wgsl:424-455, plus per-cell bookkeeping
// ------------------------------------------------------------------------

__device__ __forceinline__ int4 sum8(const int4* src, uint32_t cx, uint32_t cy, uint32_t cz,
                                     uint32_t side) {
  uint32_t count = 0u;
  uint32_t sx = 0u, sy = 0u, sz = 0u;  // wrapping i32 sums
#pragma unroll
  for (uint32_t dz = 0; dz < 2u; ++dz)
#pragma unroll
    for (uint32_t dy = 0; dy < 2u; ++dy)
#pragma unroll
      for (uint32_t dx = 0; dx < 2u; ++dx) {
        const int4 m = src[lin(2u * cx + dx, 2u * cy + dy, 2u * cz + dz, side)];
        count += static_cast<uint32_t>(m.x);
        sx += static_cast<uint32_t>(m.y);
        sy += static_cast<uint32_t>(m.z);
        sz += static_cast<uint32_t>(m.w);
      }
  return make_int4(static_cast<int>(count), static_cast<int>(sx), static_cast<int>(sy),
                   static_cast<int>(sz));
}

// One pair's ten field terms (pair_field, src/rainbow_fmm.wgsl:170-203).
// The source record {com, G m} holds com = f32(sum) * (1 / (scale
// f32(count))) and gm = G (f32(count) / f32(N)), formed once per cell with
// the same operations the reference repeats per pair.
// kFast: the caller guarantees r2 in [2^-94, FLT_MAX] (finite coordinates
// of magnitude <= 2^60 and eps2 in [2^-94, 2^100]), where the unguarded
// Newton form equals 1.0f / sqrtf(r2).
template <bool kFast = false>
__device__ __forceinline__ void pair_terms(const float4 src, float tx, float ty, float tz, float eps2,
                                           float out[10]) {
  const float rx = src.x - tx;
  const float ry = src.y - ty;
  const float rz = src.z - tz;
  const float r2 = ((rx * rx + ry * ry) + rz * rz) + eps2;
  const float inv_r = kFast ? rsqrt_fast_unguarded(r2) : rsqrt_rn(r2);  // = 1.0f / sqrtf(r2)
  const float inv_r3 = inv_r * inv_r * inv_r;
  const float inv_r5 = inv_r3 * inv_r * inv_r;
  const float gm = src.w;
  const float diagonal = -gm * inv_r3;
  const float tidal = 3.0f * gm * inv_r5;
  out[0] = gm * rx * inv_r3;
  out[1] = gm * ry * inv_r3;
  out[2] = gm * rz * inv_r3;
  out[3] = -gm * inv_r;
  out[4] = diagonal + tidal * rx * rx;
  out[5] = tidal * rx * ry;
  out[6] = tidal * rx * rz;
  out[7] = diagonal + tidal * ry * ry;
  out[8] = tidal * ry * rz;
  out[9] = diagonal + tidal * rz * rz;
}

// One running sum: plain f32, or Neumaier-compensated (the precision form).
template <bool kComp>
struct Chain {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float t) {
    if (!kComp) {
      s = s + t;
    } else {
      const float u = s + t;
      if (fabsf(s) >= fabsf(t))
        c = c + ((s - u) + t);
      else
        c = c + ((t - u) + s);
      s = u;
    }
  }
  __device__ __forceinline__ float value() const { return kComp ? s + c : s; }
  __device__ __forceinline__ void reset(float x) {
    s = x;
    c = 0.0f;
  }
};

// Source record of a cell: {com, G m}, zero when empty
// (moment_com_read and the mass/gm lines of pair_field,
// src/rainbow_fmm.wgsl:163-168, 182-183).
__device__ __forceinline__ float4 source_record(const int4 m, const SysDesc& d) {
  const uint32_t count = static_cast<uint32_t>(m.x);
  if (count == 0u) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float inv = 1.0f / (d.scale * static_cast<float>(count));
  const float mass = static_cast<float>(count) / static_cast<float>(d.n);
  return make_float4(static_cast<float>(m.y) * inv, static_cast<float>(m.z) * inv,
                     static_cast<float>(m.w) * inv, d.gravity * mass);
}

// A cell of level >= 1 is an M2L target when it is non-empty and its parent
// holds more than 16 bodies (src/rainbow_fmm.wgsl:472-476, 593-595).  Every
// field outside the current target set is zero; a cell that leaves the set
// has its field cleared here, before the tree passes run.
// `prev` is the cell's flag before this step (loaded by the caller ahead of
// time, off the critical path).
__device__ __forceinline__ void update_active(const SysDesc& d, uint32_t cell, int count, int parent_count,
                                              uint8_t prev) {
  const uint8_t now = (count != 0 && parent_count > kLeafCapacity) ? 1u : 0u;
  if (prev != now) {
    if (prev) {
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      G(d.fields)[3u * cell + 0u] = z;
      G(d.fields)[3u * cell + 1u] = z;
      G(d.fields)[3u * cell + 2u] = z;
    }
    G(d.active)[cell] = now;
  }
}

// Finest cells: bit 0 as update_active, bit 1 = non-empty now (the next
// step's upward reads it to skip the cells empty in both solves).
// `cell`: the finest cell's global index (its field); `flag`: its flag's
// place in `active` (subtree-major, see flag_index).
__device__ __forceinline__ void update_finest_flags(const SysDesc& d, uint32_t cell, uint32_t flag, int count,
                                                    int parent_count, uint8_t prev) {
  const uint8_t target = (count != 0 && parent_count > kLeafCapacity) ? 1u : 0u;
  const uint8_t now = target | (count != 0 ? 2u : 0u);
  if (prev != now) {
    if ((prev & 1u) && !target) {
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      G(d.fields)[3u * cell + 0u] = z;
      G(d.fields)[3u * cell + 1u] = z;
      G(d.fields)[3u * cell + 2u] = z;
    }
    G(d.active)[flag] = now;
  }
}

// Internal cells (count > 16) are collected per level into kStripes striped
// lists; a block reserves its items of one level with a single atomic on
// the counter of its stripe, so no counter sees more than a few atomics.
struct PushStage {
  uint32_t n[4];
  uint32_t base[4];
  uint32_t cap[4];  // the stripe's slots and first slot per staged level,
  uint32_t off[4];  // loaded when the block starts
};

// Threads 0..nlev-1: load the stripe's capacity and offset of each staged
// level (constant between configurations), off the critical path.
__device__ __forceinline__ void prefetch_stripes(const StepArgs& a, PushStage& ps, const uint32_t levels[4],
                                                 uint32_t nlev, uint32_t stripe, uint32_t t) {
  if (t < nlev) {
    const uint32_t si = levels[t] * kStripes + stripe;
    ps.cap[t] = G(a.stripe_cap)[si];
    ps.off[t] = G(a.stripe_off)[si];
  }
}

__device__ __forceinline__ uint32_t stage_item(PushStage& ps, uint32_t k) { return atomicAdd(&ps.n[k], 1u); }

// Thread 0: reserve each staged level's items in stripe `stripe`.
__device__ __forceinline__ void reserve_items(const StepArgs& a, PushStage& ps, const uint32_t levels[4],
                                              uint32_t nlev, uint32_t stripe) {
  for (uint32_t k = 0; k < nlev; ++k)
    ps.base[k] = ps.n[k] ? atomicAdd(&G(a.list_count)[levels[k] * kStripes + stripe], ps.n[k]) : 0u;
}

// Stores an item of staged level k at its reserved slot (items beyond the
// stripe's capacity are dropped; the planner sizes stripes so none is).
__device__ __forceinline__ void write_item(const StepArgs& a, const PushStage& ps, uint32_t k, uint32_t slot,
                                           uint32_t item) {
  if (slot < ps.cap[k]) G(a.lists)[ps.off[k] + slot] = item;
}

// upward (M2M), src/rainbow_fmm.wgsl:424-455: one block per 8x8x8 finest
// cells of a subtree rooted at level r = depth-3 sums levels depth-1..r in
// LDS, writes them with their source records, updates the target flags of
// the levels whose parents it holds, and lists the internal cells.
// kOne: the step holds one system, whose subtree table lists its 8^r roots
// in Morton order: the block's root follows from its index, and the system
// descriptor is loaded at entry (no dependent load of the table entry).
template <bool kOne>
__global__ __launch_bounds__(kUpwardThreads) void k_upward(StepArgs a) {
  static_assert(kUpwardThreads == 512u, "one thread per finest cell of the subtree");
  // Levels of the subtree in one LDS array: 8^3 finest cells at 0, then
  // 64, 8 and 1 cells of the three levels above.
  __shared__ int4 lv[512 + 64 + 8 + 1];
  __shared__ PushStage ps;
  __shared__ uint32_t wsum8[8];
  int4* const lvl0 = lv;
  int4* const lvl1 = lv + 512;
  int4* const lvl2 = lv + 576;
  int4* const lvl3 = lv + 584;
  const uint32_t sb = xcd_tile(blockIdx.x, gridDim.x);
  SubtreeBlock blk;
  SysDesc d;
  if constexpr (kOne) {
    d = a.sys[0];
    blk.sys = 0u;
    blk.root = lin(compact3(sb), compact3(sb >> 1), compact3(sb >> 2), 1u << (d.depth - 3u));
  } else {
    blk = a.subtree_blocks[sb];
    d = a.sys[blk.sys];
  }
  in_sgpr(d.live);
  in_sgpr(d.depth);
  in_sgpr(d.n);
  in_sgpr(d.scale);
  in_sgpr(d.gravity);
  in_sgpr(d.moments);
  in_sgpr(d.active);
  in_sgpr(d.xflag);
  in_sgpr(d.extra);
  in_sgpr(d.finest_alt);
  in_sgpr(d.source);
  if (!d.live) return;
  const uint32_t t = threadIdx.x;
  const uint32_t D = d.depth;
  const uint32_t r = D - 3u;
  uint32_t rx, ry, rz;
  unlin(blk.root, r, rx, ry, rz);
  const gptr<int4> M = G(d.moments);
  if (t < 4u) ps.n[t] = 0u;
  const uint32_t stripe = sb % kStripes;  // the stripe the table gave this subtree
  const uint32_t lev_of[4] = {D - 1u, D - 2u, r, 0u};
  // The stripe's capacity and list offset per staged level (constant between
  // configurations): uniform loads kept in registers until the items are
  // written, in flight with the cells' loads.
  uint32_t scap[3], soff[3];
#pragma unroll
  for (uint32_t k = 0; k < 3u; ++k) {
    const uint32_t si = lev_of[k] * kStripes + stripe;
    scap[k] = G(a.stripe_cap)[si];
    soff[k] = G(a.stripe_off)[si];
  }

  // This thread's cells: finest cell t (x fastest, rows of 8 contiguous
  // moments), level D-1 cell t (t < 64), level D-2 cell t - 64 (64 <= t <
  // 72), the root (t == 72).  Their target flags before this step are
  // loaded with the finest cells (the flag updates below then wait on
  // nothing).
  const uint32_t ix = t & 7u, iy = (t >> 3) & 7u, iz = t >> 6;
  const uint32_t fcell = lin(8u * rx + ix, 8u * ry + iy, 8u * rz + iz, 1u << D);
  const uint32_t fflag = (blk.root << 9) | t;  // flag_index of the cell
  const uint8_t prev_f = G(d.active)[level_off(D) + fflag];
  // The flag of this thread's cell of level D-1 (t < 64) or D-2 (t < 72).
  uint8_t prev_l = 0u;
  if (t < 72u) {
    const uint32_t u = t - 64u;
    const uint32_t idx = t < 64u ? level_off(D - 1u) + lin(4u * rx + (t & 3u), 4u * ry + ((t >> 2) & 3u),
                                                            4u * rz + (t >> 4), 1u << (D - 1u))
                                 : level_off(D - 2u) + lin(2u * rx + (u & 1u), 2u * ry + ((u >> 1) & 1u),
                                                            2u * rz + (u >> 2), 1u << (D - 2u));
    prev_l = G(d.active)[idx];
  }
  // Finest level (this step's buffer) plus the sums deposited into `extra`
  // (wrapping i32 adds, exact in any order), written back when non-zero;
  // `extra` and the other buffer's cells are zeroed for the next step's
  // deposit.
  const gptr<int4> F = finest_of(d, a.parity);
  const gptr<int4> next = finest_of(d, a.parity ^ 1u);
  int4 m0 = F[fcell];
  const uint32_t owned = static_cast<uint32_t>(m0.x);
  if (G(d.xflag)[fflag] != 0u) {
    const int4 e = G(d.extra)[fcell];
    m0.x = static_cast<int>(static_cast<uint32_t>(m0.x) + static_cast<uint32_t>(e.x));
    m0.y = static_cast<int>(static_cast<uint32_t>(m0.y) + static_cast<uint32_t>(e.y));
    m0.z = static_cast<int>(static_cast<uint32_t>(m0.z) + static_cast<uint32_t>(e.z));
    m0.w = static_cast<int>(static_cast<uint32_t>(m0.w) + static_cast<uint32_t>(e.w));
    F[fcell] = m0;
    G(d.extra)[fcell] = make_int4(0, 0, 0, 0);
    G(d.xflag)[fflag] = 0u;
  }
  // The other buffer holds the last solve's counts: non-zero only where the
  // cell's flag says it was non-empty (bit 1), so only those are zeroed.
  if (prev_f & 2u) next[fcell] = make_int4(0, 0, 0, 0);
  lvl0[t] = m0;
  __syncthreads();
  if (sorts(a, d)) {
    // Offsets of the subtree's cells in Morton order (the subtree's 512
    // finest cells are the contiguous Morton range root << 9 ..): thread u
    // takes Morton cell u; wave prefix sums, then the waves' totals.
    const uint32_t mi = compact3(t) + 8u * (compact3(t >> 1) + 8u * compact3(t >> 2));
    const uint32_t cnt = static_cast<uint32_t>(lvl0[mi].x);
    const uint32_t incl = wave_prefix_sum(cnt);
    if ((t & 63u) == 63u) wsum8[t >> 6] = incl;
    __syncthreads();
    uint32_t base = 0u;
    for (uint32_t w = 0; w < (t >> 6); ++w) base += wsum8[w];
    const uint32_t start = base + incl - cnt;
    const uint32_t root_m = spread3(rx) | (spread3(ry) << 1) | (spread3(rz) << 2);
    const uint32_t g = (root_m << 9) + t;
    // The owned count of Morton cell t lives with the thread of its lin index.
    __shared__ uint32_t owned_s[512];
    owned_s[t] = owned;
    __syncthreads();
    G(d.cstart)[g] = start;
    G(d.xstart)[g] = start + owned_s[mi];
  }
  if (t < 64u) {
    const uint32_t cx = t & 3u, cy = (t >> 2) & 3u, cz = t >> 4;
    const int4 s = sum8(lvl0, cx, cy, cz, 8u);
    lvl1[t] = s;
    M[level_off(D - 1u) + lin(4u * rx + cx, 4u * ry + cy, 4u * rz + cz, 1u << (D - 1u))] = s;
  }
  __syncthreads();
  if (t < 8u) {
    const uint32_t cx = t & 1u, cy = (t >> 1) & 1u, cz = t >> 2;
    const int4 s = sum8(lvl1, cx, cy, cz, 4u);
    lvl2[t] = s;
    M[level_off(D - 2u) + lin(2u * rx + cx, 2u * ry + cy, 2u * rz + cz, 1u << (D - 2u))] = s;
  }
  __syncthreads();
  if (t == 0u) {
    const int4 s = sum8(lvl2, 0u, 0u, 0u, 2u);
    lvl3[0] = s;
    M[level_off(r) + blk.root] = s;
  }
  __syncthreads();

  // Internal cells (levels D-1, D-2 and the root) are staged first, so that
  // thread 0's list reservation is in flight while the records and flags
  // are stored.
  uint32_t my_item = 0u, my_slot = 0u, my_level = 0u, local = 0u, lcell = 0u;
  int4 lm = make_int4(0, 0, 0, 0);
  int lparent = 0;
  uint8_t lprev = 0u;
  bool mine = false;
  if (t < 64u) {
    const uint32_t cx = t & 3u, cy = (t >> 2) & 3u, cz = t >> 4;
    local = lin(4u * rx + cx, 4u * ry + cy, 4u * rz + cz, 1u << (D - 1u));
    lcell = level_off(D - 1u) + local;
    lm = lvl1[t];
    lparent = lvl2[lin(cx >> 1, cy >> 1, cz >> 1, 2u)].x;
    lprev = prev_l;
    my_level = 0u;
    mine = true;
  } else if (t < 72u) {
    const uint32_t u = t - 64u;
    const uint32_t cx = u & 1u, cy = (u >> 1) & 1u, cz = u >> 2;
    local = lin(2u * rx + cx, 2u * ry + cy, 2u * rz + cz, 1u << (D - 2u));
    lcell = level_off(D - 2u) + local;
    lm = lvl2[u];
    lparent = lvl3[0].x;
    lprev = prev_l;
    my_level = 1u;
    mine = true;
  } else if (t == 72u) {
    local = blk.root;
    lcell = level_off(r) + blk.root;
    lm = lvl3[0];
    my_level = 2u;
    mine = true;
  }
  const bool pushes = mine && lm.x > kLeafCapacity;
  if (pushes) {
    my_item = pack_item(blk.sys, local);
    my_slot = stage_item(ps, my_level);
  }
  __syncthreads();
  if (t == 0u) reserve_items(a, ps, lev_of, 3u, stripe);

  {
    // A cell empty now and in the last solve keeps its zero record.
    const int4 m = lvl0[t];
    if (m.x != 0 || (prev_f & 2u)) G(d.source)[level_off(D) + fcell] = source_record(m, d);
    update_finest_flags(d, level_off(D) + fcell, level_off(D) + fflag, m.x, lvl1[lin(ix >> 1, iy >> 1, iz >> 1, 4u)].x,
                        prev_f);
  }
  if (mine) {
    G(d.source)[lcell] = source_record(lm, d);
    // The root's flag (level r) is set by the pass above the subtrees.
    if (my_level < 2u) update_active(d, lcell, lm.x, lparent, lprev);
  }
  __syncthreads();
  if (pushes) {
    const uint32_t slot = ps.base[my_level] + my_slot;
    const uint32_t cap = my_level == 0u ? scap[0] : (my_level == 1u ? scap[1] : scap[2]);
    const uint32_t off = my_level == 0u ? soff[0] : (my_level == 1u ? soff[1] : soff[2]);
    if (slot < cap) G(a.lists)[off + slot] = my_item;
  }
}

// LDS of the top levels' pass.
struct TopLds {
  int4 lv[512 + 64 + 8 + 1];
  PushStage ps;
  uint32_t wsum8[8];
  uint8_t prev[1u + 8u + 64u + 512u];
};

// The levels above the subtrees of system s, by one block of kThreads
// threads: levels r-1..0 from the 8^r subtree roots, their source records
// and lists, the target flags of levels 1..r, the subtree offsets of a
// sorting step, and the fields of level 1.
template <bool kComp, uint32_t kThreads>
__device__ void top_levels(const StepArgs& a, uint32_t s, TopLds& L) {
  int4* const lv = L.lv;
  PushStage& ps = L.ps;
  uint8_t* const prev = L.prev;
  const SysDesc d = a.sys[s];
  if (!d.live) return;
  const uint32_t t = threadIdx.x;
  const uint32_t r = d.depth - 3u;
  const gptr<int4> M = G(d.moments);
  auto buf = [&](uint32_t j) -> int4* {
    return lv + (j == 0u ? 0u : (j == 1u ? 512u : (j == 2u ? 576u : 584u)));
  };
  if (t < 4u) ps.n[t] = 0u;
  const uint32_t stripe = s % kStripes;
  const uint32_t levels[4] = {0u, 1u, 2u, 3u};
  prefetch_stripes(a, ps, levels, r, stripe, t);
  // Level r (8^r <= 512 cells) into buf(0), then levels r-1, r-2, ... into
  // buf(1), buf(2), buf(3).  The target flags of levels 0..r before this
  // step are loaded alongside.
  const uint32_t nr = 1u << (3u * r);
  for (uint32_t i = t; i < nr; i += kThreads) buf(0)[i] = M[level_off(r) + i];
  for (uint32_t c = t; c < level_off(r + 1u); c += kThreads) prev[c] = G(d.active)[c];
  __syncthreads();
  for (uint32_t j = 1u; j <= r; ++j) {
    const uint32_t k = r - j;
    const uint32_t dk = 1u << k;
    if (t < dk * dk * dk) {
      uint32_t cx, cy, cz;
      unlin(t, k, cx, cy, cz);
      const int4 v = sum8(buf(j - 1u), cx, cy, cz, 2u * dk);
      buf(j)[t] = v;
      M[level_off(k) + t] = v;
    }
    __syncthreads();
  }
  if (sorts(a, d)) {
    // Offsets of the subtrees (level r cells, Morton order) in the system:
    // 512 values (the absent ones zero), kThreads per pass; wave prefix
    // sums, then the waves' totals.
    const uint32_t dr = 1u << r;
    uint32_t* const wsum8 = L.wsum8;
    constexpr uint32_t kPasses = (512u + kThreads - 1u) / kThreads;
    uint32_t own[kPasses], incl[kPasses];
#pragma unroll
    for (uint32_t k = 0; k < kPasses; ++k) {
      const uint32_t e = k * kThreads + t;
      own[k] = 0u;
      incl[k] = 0u;
      if (e < 512u) {
        own[k] = e < nr ? static_cast<uint32_t>(buf(0)[lin(compact3(e), compact3(e >> 1), compact3(e >> 2), dr)].x)
                        : 0u;
        incl[k] = wave_prefix_sum(own[k]);
        if ((e & 63u) == 63u) wsum8[e >> 6] = incl[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kPasses; ++k) {
      const uint32_t e = k * kThreads + t;
      if (e < nr) {
        uint32_t base = 0u;
        for (uint32_t w = 0; w < (e >> 6); ++w) base += wsum8[w];
        G(d.sstart)[e] = base + incl[k] - own[k];
      }
    }
  }
  // Internal cells of levels 0..r-1 are staged first, so that thread 0's
  // list reservation is in flight while the records and flags are stored.
  uint32_t items[4], slots[4], levs[4];
  uint32_t nmine = 0u;
  for (uint32_t j = 1u; j <= r; ++j) {
    const uint32_t k = r - j;
    const uint32_t dk = 1u << k;
    const uint32_t ncell = dk * dk * dk;
    for (uint32_t i = t; i < ncell; i += kThreads) {
      if (buf(j)[i].x > kLeafCapacity && nmine < 4u) {
        items[nmine] = pack_item(s, i);
        levs[nmine] = k;
        slots[nmine] = stage_item(ps, k);
        ++nmine;
      }
    }
  }
  __syncthreads();
  if (t == 0u) {
    if (a.nsys == 1u) {
      // One system: levels below r are listed by this block alone, so the
      // counts are stored and the slots start at 0 (no atomics to wait on).
      for (uint32_t k = 0; k < r; ++k) {
        G(a.list_count)[k * kStripes + stripe] = ps.n[k];
        ps.base[k] = 0u;
      }
    } else {
      reserve_items(a, ps, levels, r, stripe);
    }
  }
  // Levels 0..r-1: source records; levels 1..r: flags; one cell per thread.
  for (uint32_t cell = t; cell < level_off(r + 1u); cell += kThreads) {
    uint32_t k = 0u;
    while (cell >= level_off(k + 1u)) ++k;
    const uint32_t j = r - k, dk = 1u << k, i = cell - level_off(k);
    const int4 m = buf(j)[i];
    if (k < r) G(d.source)[cell] = source_record(m, d);
    if (k >= 1u) {
      uint32_t cx, cy, cz;
      unlin(i, k, cx, cy, cz);
      update_active(d, cell, m.x, buf(j + 1u)[lin(cx >> 1, cy >> 1, cz >> 1, dk >> 1)].x, prev[cell]);
    }
  }
  __syncthreads();
  for (uint32_t q = 0; q < nmine; ++q) write_item(a, ps, levs[q], ps.base[levs[q]] + slots[q], items[q]);

  // Level 1 (tree pass of level 1, src/rainbow_fmm.wgsl:457-611): the root's
  // children have no well-separated source (every level-1 cell is a
  // neighbour) and inherit the root's zero field, so a level-1 target's
  // field is +0 plus, for a leaf, its near field: the other level-1 cells
  // in dz, dy, dx order.
  if (t < 8u) {
    const int4* l1 = buf(r - 1u);
    const int root = buf(r)[0].x;
    const int count = l1[t].x;
    if (count != 0 && root > kLeafCapacity) {
      const uint32_t tx = t & 1u, ty = (t >> 1) & 1u, tz = t >> 2;
      const float cx = center1(tx, 2u), cy = center1(ty, 2u), cz = center1(tz, 2u);
      Chain<kComp> ch[10];
      if (count <= kLeafCapacity) {  // level 1 is never the finest (depth >= 4)
        const float near_epsilon = fmaxf(d.softening, (0.72f * kWorldHalfExtent) / 2.0f);
        const float eps2 = near_epsilon * near_epsilon;
        for (uint32_t zz = 0; zz < 2u; ++zz)
          for (uint32_t yy = 0; yy < 2u; ++yy)
            for (uint32_t xx = 0; xx < 2u; ++xx) {
              if (xx == tx && yy == ty && zz == tz) continue;
              const float4 sr = source_record(l1[lin(xx, yy, zz, 2u)], d);
              if (sr.w == 0.0f) continue;
              float v[10];
              pair_terms(sr, cx, cy, cz, eps2, v);
#pragma unroll
              for (int k = 0; k < 10; ++k) ch[k].add(v[k]);
            }
      }
      const gptr<float4> f = G(d.fields) + 3u * (level_off(1u) + t);
      f[0] = make_float4(ch[0].value(), ch[1].value(), ch[2].value(), ch[3].value());
      f[1] = make_float4(ch[4].value(), ch[5].value(), ch[6].value(), ch[7].value());
      f[2] = make_float4(ch[8].value(), ch[9].value(), 0.0f, 0.0f);
    }
  }
}

// One block per system.
template <bool kComp>
__global__ __launch_bounds__(kTopThreads) void k_top(StepArgs a) {
  __shared__ TopLds L;
  top_levels<kComp, kTopThreads>(a, blockIdx.x, L);
}

// ------------------------------------------------------------------------
// tree(lo..hi): M2L (src/rainbow_fmm.wgsl:457-519), L2L (:585-611) and near
// (:521-572) of the targets of levels lo..hi, one 256-thread block per
// internal parent.  The parent's 6x6x6 source window is staged in LDS with
// one round of loads; the window is walked one source-parent block (8
// sources) at a time: each of the four waves evaluates one block's 8 x 8
// (target, source) pair fields, then one lane per (target, component)
// chain adds the terms in the reference order.
//
// A launch over several levels has no dependency between its blocks: an
// item of level l > lo needs its parent's field, which another block of the
// same launch writes, so the block recomputes the fields of the parent and
// of its ancestors down to level lo (single targets, the same operations in
// the same order, so the values are identical) from the field of the
// ancestor at level lo - 1, written by an earlier launch.
// ------------------------------------------------------------------------

// The L2L term of one field component (downward, src/rainbow_fmm.wgsl:
// 598-610): the parent's field {a, phi, J} shifted by (dx, dy, dz) from the
// parent's centre to the child's.
__device__ __forceinline__ float l2l_term(uint32_t k, float4 pa, float4 pj0, float4 pj1, float dx, float dy,
                                          float dz) {
  switch (k) {
    case 0: return pa.x + ((pj0.x * dx + pj0.y * dy) + pj0.z * dz);
    case 1: return pa.y + ((pj0.y * dx + pj0.w * dy) + pj1.x * dz);
    case 2: return pa.z + ((pj0.z * dx + pj1.x * dy) + pj1.y * dz);
    case 3: return pa.w - ((pa.x * dx + pa.y * dy) + pa.z * dz);
    case 4: return pj0.x;
    case 5: return pj0.y;
    case 6: return pj0.z;
    case 7: return pj0.w;
    case 8: return pj1.x;
    default: return pj1.y;
  }
}

// Stages the source records of cells 2P-2 .. 2P+3 per axis of level `level`
// (P = the parent's coordinates); cells outside the level read as empty.
// Returns false when a source this thread staged has a coordinate that is
// not finite or exceeds 2^60 in magnitude.
// kStride threads stage the window, thread `tid` every kStride-th record.
template <uint32_t kStride = kTreeThreads>
__device__ __forceinline__ bool stage_window(float4* win, const SysDesc& d, uint32_t level, uint32_t px,
                                             uint32_t py, uint32_t pz, uint32_t tid) {
  bool bounded = true;
  const uint32_t dim = 1u << level;
  const int dim_i = static_cast<int>(dim);
  const gptr<float4> src = G(d.source) + level_off(level);
  for (uint32_t w = tid; w < 216u; w += kStride) {
    const uint32_t wz = w / 36u, wy = (w / 6u) % 6u, wx = w % 6u;
    const int qx = 2 * static_cast<int>(px) - 2 + static_cast<int>(wx);
    const int qy = 2 * static_cast<int>(py) - 2 + static_cast<int>(wy);
    const int qz = 2 * static_cast<int>(pz) - 2 + static_cast<int>(wz);
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (qx >= 0 && qy >= 0 && qz >= 0 && qx < dim_i && qy < dim_i && qz < dim_i)
      v = src[lin(static_cast<uint32_t>(qx), static_cast<uint32_t>(qy), static_cast<uint32_t>(qz), dim)];
    win[w] = v;
    bounded = bounded && fabsf(v.x) <= 0x1p60f && fabsf(v.y) <= 0x1p60f && fabsf(v.z) <= 0x1p60f;
  }
  return bounded;
}

// Moment of cell (x, y, z) of level l < depth from its 8 children (wrapping
// i32 sums, as upward and the top levels form it), for a launch that runs
// before the level's own moments are written.
__device__ __forceinline__ int4 children_moment(const SysDesc& d, uint32_t l, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t cdim = 2u << l;
  const gptr<int4> M = G(d.moments) + level_off(l + 1u);
  uint32_t s0 = 0u, s1 = 0u, s2 = 0u, s3 = 0u;
#pragma unroll
  for (uint32_t c = 0; c < 8u; ++c) {
    const int4 m = M[lin(2u * x + (c & 1u), 2u * y + ((c >> 1) & 1u), 2u * z + (c >> 2), cdim)];
    s0 += static_cast<uint32_t>(m.x);
    s1 += static_cast<uint32_t>(m.y);
    s2 += static_cast<uint32_t>(m.z);
    s3 += static_cast<uint32_t>(m.w);
  }
  return make_int4(static_cast<int>(s0), static_cast<int>(s1), static_cast<int>(s2), static_cast<int>(s3));
}

// stage_window with the sources' records formed from their children's
// moments (the same integer sums, so the same records).
template <uint32_t kStride = kTreeThreads>
__device__ __forceinline__ bool stage_window_from_children(float4* win, const SysDesc& d, uint32_t level, uint32_t px,
                                                           uint32_t py, uint32_t pz, uint32_t tid) {
  bool bounded = true;
  const int dim_i = static_cast<int>(1u << level);
  for (uint32_t w = tid; w < 216u; w += kStride) {
    const uint32_t wz = w / 36u, wy = (w / 6u) % 6u, wx = w % 6u;
    const int qx = 2 * static_cast<int>(px) - 2 + static_cast<int>(wx);
    const int qy = 2 * static_cast<int>(py) - 2 + static_cast<int>(wy);
    const int qz = 2 * static_cast<int>(pz) - 2 + static_cast<int>(wz);
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (qx >= 0 && qy >= 0 && qz >= 0 && qx < dim_i && qy < dim_i && qz < dim_i)
      v = source_record(children_moment(d, level, static_cast<uint32_t>(qx), static_cast<uint32_t>(qy),
                                        static_cast<uint32_t>(qz)),
                        d);
    win[w] = v;
    bounded = bounded && fabsf(v.x) <= 0x1p60f && fabsf(v.y) <= 0x1p60f && fabsf(v.z) <= 0x1p60f;
  }
  return bounded;
}

// LDS of the ancestor path per ancestor: its 6^3 window and its 10 x 216
// pair terms (rows padded to 220 floats so the chains' rows start in
// different LDS banks).
constexpr uint32_t kAncRow = 220u;
constexpr uint32_t kAncestorBytes = 216u * 16u + 10u * kAncRow * 4u;
// LDS of one item: the window and a round's 10 x 64 terms of each of its
// kCpr chunks of 8 x 8 pairs (the tree kernel's pair tile: 8 chunks per
// round, two per wave, or 4, one per wave).
constexpr uint32_t item_bytes(uint32_t cpr) { return 216u * 16u + cpr * 10u * 64u * 4u; }

// Fields of the internal ancestors A_lo .. A_{level-1} of an item of
// `level` whose parent P = A_{level-1} has coordinates (px, py, pz), from
// the field of A_{lo-1} in memory.  The field of A_j is its M2L sum over
// its parent's neighbours' children in the reference order (source parent
// pz, py, px, then child cz, cy, cx), then + the L2L term of A_{j-1}'s
// field.  The M2L sums of all ancestors are independent and run in
// parallel; only the L2L additions run in order.  Terms of cells adjacent
// to the target, empty cells and cells outside the level are zero and
// leave the sums unchanged (a running sum that starts at +0 never becomes
// -0, so adding +0 is the identity).  Writes the parent's field to pf[0..9].
// All threads of the block call it; lds holds (level - lo) * kAncestorBytes.
template <bool kComp>
__device__ void ancestor_fields(const SysDesc& d, uint32_t lo, uint32_t level, uint32_t px, uint32_t py,
                                uint32_t pz, unsigned char* lds, float* pf, uint32_t tid) {
  const uint32_t na = level - lo;
  float4* awin = reinterpret_cast<float4*>(lds);               // [na][216]
  float* st = reinterpret_cast<float*>(lds + na * 216u * 16u);  // [na][10][kAncRow]
  // One round of loads: every ancestor's window, and the field of A_{lo-1}.
  bool bounded = true;
  for (uint32_t w = tid; w < na * 216u; w += kTreeThreads) {
    const uint32_t i = w / 216u, ww = w - i * 216u;
    const uint32_t sh = level - 1u - (lo + i);  // A_{lo+i} = P >> sh
    const uint32_t al = lo + i;
    const uint32_t dim = 1u << al;
    const int dim_i = static_cast<int>(dim);
    const uint32_t qpx = (px >> sh) >> 1, qpy = (py >> sh) >> 1, qpz = (pz >> sh) >> 1;
    const uint32_t wz = ww / 36u, wy = (ww / 6u) % 6u, wx = ww % 6u;
    const int qx = 2 * static_cast<int>(qpx) - 2 + static_cast<int>(wx);
    const int qy = 2 * static_cast<int>(qpy) - 2 + static_cast<int>(wy);
    const int qz = 2 * static_cast<int>(qpz) - 2 + static_cast<int>(wz);
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (qx >= 0 && qy >= 0 && qz >= 0 && qx < dim_i && qy < dim_i && qz < dim_i)
      v = G(d.source)[level_off(al) +
                   lin(static_cast<uint32_t>(qx), static_cast<uint32_t>(qy), static_cast<uint32_t>(qz), dim)];
    awin[w] = v;
    bounded = bounded && fabsf(v.x) <= 0x1p60f && fabsf(v.y) <= 0x1p60f && fabsf(v.z) <= 0x1p60f;
  }
  if (tid < 10u) {
    const uint32_t up = level - lo;  // P >> up = A_{lo-1}
    const uint32_t adim = 1u << (lo - 1u);
    pf[tid] = reinterpret_cast<const float*>(
        G(d.fields) + 3u * (level_off(lo - 1u) + lin(px >> up, py >> up, pz >> up, adim)))[tid];
  }
  // Bounded sources (and eps2 = (0.096 / dim)^2 in [2^-94, 1] for every
  // level) keep each r2 in the unguarded inverse square root's domain; lanes
  // without a pair then evaluate a zero-mass source (terms of +-0).
  const bool fast = __syncthreads_and(bounded) != 0;
  for (uint32_t w = tid; w < na * 216u; w += kTreeThreads) {
    const uint32_t i = w / 216u, u = w - i * 216u;
    const uint32_t sh = level - 1u - (lo + i);
    const uint32_t dim = 1u << (lo + i);
    const uint32_t cx = px >> sh, cy = py >> sh, cz = pz >> sh;
    const float cell_epsilon = (0.06f * kWorldHalfExtent) * inv_pow2(dim);
    const float eps2 = cell_epsilon * cell_epsilon;
    const uint32_t q = u >> 3, c = u & 7u;
    const uint32_t wsx = 2u * (q % 3u) + (c & 1u);
    const uint32_t wsy = 2u * ((q / 3u) % 3u) + ((c >> 1) & 1u);
    const uint32_t wsz = 2u * (q / 9u) + (c >> 2);
    const uint32_t wtx = 2u + (cx & 1u), wty = 2u + (cy & 1u), wtz = 2u + (cz & 1u);
    const bool near = (wsx + 1u - wtx) <= 2u && (wsy + 1u - wty) <= 2u && (wsz + 1u - wtz) <= 2u;
    float t[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    float4 sr = awin[i * 216u + wsx + 6u * (wsy + 6u * wsz)];
    if (fast) {
      if (near) sr.w = 0.0f;
      if (__any(sr.w != 0.0f))
        pair_terms<true>(sr, center1(cx, dim), center1(cy, dim), center1(cz, dim), eps2, t);
    } else if (!near && sr.w != 0.0f) {
      pair_terms<false>(sr, center1(cx, dim), center1(cy, dim), center1(cz, dim), eps2, t);
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) st[(i * 10u + k) * kAncRow + u] = t[k];
  }
  __syncthreads();
  // Chains: thread i * 10 + k sums component k of ancestor i.
  const uint32_t ci = tid / 10u, ck = tid - 10u * ci;
  Chain<kComp> ch;
  if (ci < na) {
    // 216 terms = 54 float4 (rows start 16-byte aligned: 220 * 4 = 880).
    const float4* row = reinterpret_cast<const float4*>(st + (ci * 10u + ck) * kAncRow);
#pragma unroll 3
    for (uint32_t u = 0; u < 54u; ++u) {
      const float4 v = row[u];
      ch.add(v.x);
      ch.add(v.y);
      ch.add(v.z);
      ch.add(v.w);
    }
  }
  // L2L in order, from A_lo up to the parent.
  for (uint32_t i = 0; i < na; ++i) {
    float v = 0.0f;
    if (ci == i) {
      const uint32_t sh = level - 1u - (lo + i);
      const uint32_t dim = 1u << (lo + i);
      const uint32_t cx = px >> sh, cy = py >> sh, cz = pz >> sh;
      const float4 pa = make_float4(pf[0], pf[1], pf[2], pf[3]);
      const float4 pj0 = make_float4(pf[4], pf[5], pf[6], pf[7]);
      const float4 pj1 = make_float4(pf[8], pf[9], 0.0f, 0.0f);
      const float dx = center1(cx, dim) - center1(cx >> 1, dim >> 1);
      const float dy = center1(cy, dim) - center1(cy >> 1, dim >> 1);
      const float dz = center1(cz, dim) - center1(cz >> 1, dim >> 1);
      ch.reset(ch.value() + l2l_term(ck, pa, pj0, pj1, dx, dy, dz));
      v = ch.value();
    }
    __syncthreads();
    if (ci == i) pf[ck] = v;
    __syncthreads();
  }
}

// Window offsets (2 qx + 12 qy + 72 qz) of the source-parent blocks q < 27
// that wave w evaluates, q = r0 + w * per-wave + j, packed one byte per
// (round r0 / chunks-per-round, j).
struct TreeBlockOffsets {
  uint64_t pack[kTreeThreads / 64u];
};
template <uint32_t kCpr>
constexpr TreeBlockOffsets make_tree_block_offsets() {
  TreeBlockOffsets t{};
  constexpr uint32_t waves = kTreeThreads / 64u, per_wave = kCpr / waves;
  for (uint32_t w = 0; w < waves; ++w)
    for (uint32_t r0 = 0; r0 < 32u; r0 += kCpr)
      for (uint32_t j = 0; j < per_wave; ++j) {
        const uint32_t q = r0 + w * per_wave + j;
        if (q < 27u) {
          const uint64_t off = 2u * (q % 3u) + 12u * ((q / 3u) % 3u) + 72u * (q / 9u);
          t.pack[w] |= off << (8u * (r0 / kCpr * per_wave + j));
        }
      }
  return t;
}
template <uint32_t kCpr>
__device__ __forceinline__ uint64_t tree_block_offsets(uint32_t wave) {
  static_assert(32u / kCpr * (kCpr / (kTreeThreads / 64u)) <= 8u, "a wave's block offsets fit in 64 bits");
  constexpr TreeBlockOffsets kT = make_tree_block_offsets<kCpr>();
  static_assert(kTreeThreads / 64u == 4u, "four waves per tree block");
  return wave == 0u ? kT.pack[0] : wave == 1u ? kT.pack[1] : wave == 2u ? kT.pack[2] : kT.pack[3];
}

// LDS word of the term of pair (target tt, source cc) for component k in
// chunk slot u: rows of 8 sources, component-major ([u][k][tt][cc]), so a
// component's 64 pair lanes store 64 consecutive words; each row's two
// halves swap places when bit 2 ^ bit 3 of its chain thread (8 k + tt) is
// set, so that any 8 (and any 16) consecutive chain threads read their
// 16-byte halves from distinct banks.
__device__ __forceinline__ uint32_t term_index(uint32_t u, uint32_t k, uint32_t tt, uint32_t cc) {
  const uint32_t reader = 8u * k + tt;
  const uint32_t swap = ((reader >> 2) ^ (reader >> 3)) & 1u;
  return u * 640u + reader * 8u + 4u * ((cc >> 2) ^ swap) + (cc & 3u);
}

// Forms of the tree kernel:
//   kTreeSingle  one level; the parent's field from memory.
//   kTreeFused   levels lo..hi; the parent's field recomputed along the
//                ancestor path (ancestor_fields) when its level is >= lo.
//   kTreeM2l     levels lo..hi, items below their system's finest level
//                only: each target's M2L sum, into the system's m2l buffer.
//   kTreeFinish  levels lo..hi after a kTreeM2l launch over the levels below
//                the finest: the parent's field from the level-1 field and
//                the ancestors' M2L sums (L2L additions in order, one wave),
//                then per target its M2L sum (evaluated here at the finest
//                level, read from the m2l buffer above it), + the L2L term,
//                + the near field of leaves.  The M2L sums do not depend on
//                the parents' fields, so the first launch has no order
//                between its items and nothing is recomputed.
constexpr uint32_t kTreeSingle = 0u, kTreeFused = 1u, kTreeM2l = 2u, kTreeFinish = 3u;

// The L2L additions of the ancestor path from the M2L sums (kTreeFinish):
// F(A_1) from memory, F(A_j) = M2L(A_j) + L2L term of F(A_{j-1}) for j =
// 2..level-1, the same operations as ancestor_fields.  Run by one whole
// wave, lane k < 10 on component k: path_loads issues the loads (before the
// item's other loads, so that their latencies overlap), path_field adds.
struct PathSums {
  float f;                // component of F(A_1)
  float m[kLevels - 3u];  // components of the M2L sums of A_2 .. A_5
};

__device__ __forceinline__ PathSums path_loads(const SysDesc& d, uint32_t level, uint32_t px, uint32_t py, uint32_t pz,
                                               uint32_t lane) {
  PathSums p;
  const uint32_t k = lane < 10u ? lane : 9u;
  const uint32_t up = level - 2u;  // P >> up = A_1
  p.f = reinterpret_cast<gptr<float>>(G(d.fields) + 3u * (level_off(1u) + lin(px >> up, py >> up, pz >> up, 2u)))[k];
#pragma unroll
  for (uint32_t j = 2u; j < kLevels - 1u; ++j) {
    p.m[j - 2u] = 0.0f;
    if (j < level) {
      const uint32_t sh = level - 1u - j;
      p.m[j - 2u] =
          reinterpret_cast<gptr<float>>(G(d.m2l) + 3u * (level_off(j) + lin(px >> sh, py >> sh, pz >> sh, 1u << j)))[k];
    }
  }
  return p;
}

// Component `lane` (< 10) of the parent's field.
__device__ __forceinline__ float path_field(const PathSums& p, uint32_t level, uint32_t px, uint32_t py, uint32_t pz,
                                            uint32_t lane) {
  const uint32_t k = lane < 10u ? lane : 9u;
  float f = p.f;
  auto at = [&](int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(f), l)); };
#pragma unroll
  for (uint32_t j = 2u; j < kLevels - 1u; ++j) {
    if (j < level) {
      const float4 pa = make_float4(at(0), at(1), at(2), at(3));
      const float4 pj0 = make_float4(at(4), at(5), at(6), at(7));
      const float4 pj1 = make_float4(at(8), at(9), 0.0f, 0.0f);
      const uint32_t sh = level - 1u - j;
      const uint32_t dim = 1u << j;
      const uint32_t cx = px >> sh, cy = py >> sh, cz = pz >> sh;
      const float dx = center1(cx, dim) - center1(cx >> 1, dim >> 1);
      const float dy = center1(cy, dim) - center1(cy >> 1, dim >> 1);
      const float dz = center1(cz, dim) - center1(cz >> 1, dim >> 1);
      f = p.m[j - 2u] + l2l_term(k, pa, pj0, pj1, dx, dy, dz);
    }
  }
  return f;
}

// Orders a wave's LDS accesses across its lanes: the compiler moves no
// memory access over it (a wave's LDS instructions then execute in program
// order, so a lane reads what other lanes of its wave wrote before it).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of a finish item above its system's finest level run by one wave: the
// 4^3 cells around its targets (their near sources) and one chunk's terms.
struct ShortWaveLds {
  float4 win[64];
  float terms[640];
};

// A finish item above its system's finest level (every target's M2L sum is
// in the m2l buffer), by one wave: the parent's field from the level-1
// field and the ancestors' M2L sums, then per target its M2L sum + the L2L
// term + the near field of a leaf target, the same operations and order as
// k_tree's finish form (near sources dz, dy, dx; the centre skipped).
// Chain lanes as in k_tree_w: lane l sums component l / 8 of target l mod
// 8, lanes 0..15 also component 8 + l / 8.
template <bool kComp>
__device__ void finish_short_wave(const StepArgs& a, const SysDesc& d, uint32_t level, uint32_t pidx, uint32_t lane,
                                  ShortWaveLds& W) {
  const uint32_t dim = 1u << level, pdim = dim >> 1u, off = level_off(level);
  uint32_t px, py, pz;
  unlin(pidx, level - 1u, px, py, pz);
  const PathSums path = path_loads(d, level, px, py, pz, lane);
  const uint32_t ct = lane & 7u, ka = lane >> 3;
  const bool has_b = lane < 16u;
  const uint32_t kb = has_b ? 8u + (lane >> 3) : 9u;
  const uint32_t cx = 2u * px + (ct & 1u), cy = 2u * py + ((ct >> 1) & 1u), cz = 2u * pz + (ct >> 2);
  const uint32_t cg = off + lin(cx, cy, cz, dim);
  const gptr<float> msum = reinterpret_cast<gptr<float>>(G(d.m2l) + 3u * cg);
  const float m2l_a = msum[ka], m2l_b = msum[kb];
  const uint32_t tt = lane >> 3, cc = lane & 7u;
  const uint32_t tx = 2u * px + (tt & 1u), ty = 2u * py + ((tt >> 1) & 1u), tz = 2u * pz + (tt >> 2);
  const int tcount = cell_count(d, a.parity, level, lin(tx, ty, tz, dim));
  // Near window: cell (2P - 1 + (lane & 3), 2P - 1 + (lane >> 2 & 3),
  // 2P - 1 + (lane >> 4)); cells outside the level are empty.
  bool bounded = true;
  {
    const int dim_i = static_cast<int>(dim);
    const int qx = 2 * static_cast<int>(px) - 1 + static_cast<int>(lane & 3u);
    const int qy = 2 * static_cast<int>(py) - 1 + static_cast<int>((lane >> 2) & 3u);
    const int qz = 2 * static_cast<int>(pz) - 1 + static_cast<int>(lane >> 4);
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (qx >= 0 && qy >= 0 && qz >= 0 && qx < dim_i && qy < dim_i && qz < dim_i)
      v = G(d.source)[off + lin(static_cast<uint32_t>(qx), static_cast<uint32_t>(qy), static_cast<uint32_t>(qz), dim)];
    wave_sync();  // the last item's window reads come first
    W.win[lane] = v;
    bounded = fabsf(v.x) <= 0x1p60f && fabsf(v.y) <= 0x1p60f && fabsf(v.z) <= 0x1p60f;
  }
  // The chains start from the M2L sum + the L2L term of the parent's field.
  const float f = path_field(path, level, px, py, pz, lane);
  auto at = [&](int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(f), l)); };
  const float4 pa = make_float4(at(0), at(1), at(2), at(3));
  const float4 pj0 = make_float4(at(4), at(5), at(6), at(7));
  const float4 pj1 = make_float4(at(8), at(9), 0.0f, 0.0f);
  const float dx = center1(cx, dim) - center1(px, pdim);
  const float dy = center1(cy, dim) - center1(py, pdim);
  const float dz = center1(cz, dim) - center1(pz, pdim);
  Chain<kComp> cha, chb;
  cha.reset(m2l_a + l2l_term(ka, pa, pj0, pj1, dx, dy, dz));
  chb.reset(m2l_b + l2l_term(kb, pa, pj0, pj1, dx, dy, dz));
  const bool leaf = tcount != 0 && tcount <= kLeafCapacity;
  if (__any(leaf)) {
    const float near_epsilon = fmaxf(d.softening, (0.72f * kWorldHalfExtent) * inv_pow2(dim));
    const float eps2 = near_epsilon * near_epsilon;
    const bool fast = __all(bounded) && eps2 >= 0x1p-94f && eps2 <= 0x1p100f;
    const float tcx = center1(tx, dim), tcy = center1(ty, dim), tcz = center1(tz, dim);
    wave_sync();
#pragma unroll 1
    for (uint32_t q = 0; q < 4u; ++q) {
      const uint32_t idx = q * 8u + cc;
      uint32_t widx = 0u;
      bool active = false;
      if (idx < 27u) {
        widx = ((tt & 1u) + idx % 3u) + 4u * (((tt >> 1) & 1u) + (idx / 3u) % 3u) + 16u * ((tt >> 2) + idx / 9u);
        active = leaf && idx != 13u;
      }
      float t[10];
      if (fast) {
        float4 sr = W.win[widx];
        if (!active) sr.w = 0.0f;
        pair_terms<true>(sr, tcx, tcy, tcz, eps2, t);
      } else {
#pragma unroll
        for (int k = 0; k < 10; ++k) t[k] = 0.0f;
        const float4 sr = W.win[widx];
        if (active && sr.w != 0.0f) pair_terms<false>(sr, tcx, tcy, tcz, eps2, t);
      }
      wave_sync();  // the last chunk's term reads come first
#pragma unroll
      for (uint32_t k = 0; k < 10u; ++k) W.terms[term_index(0u, k, tt, cc)] = t[k];
      wave_sync();
      {
        const float4 v0 = *reinterpret_cast<const float4*>(&W.terms[term_index(0u, ka, ct, 0u)]);
        const float4 v1 = *reinterpret_cast<const float4*>(&W.terms[term_index(0u, ka, ct, 4u)]);
        cha.add(v0.x);
        cha.add(v0.y);
        cha.add(v0.z);
        cha.add(v0.w);
        cha.add(v1.x);
        cha.add(v1.y);
        cha.add(v1.z);
        cha.add(v1.w);
      }
      if (has_b) {
        const float4 v0 = *reinterpret_cast<const float4*>(&W.terms[term_index(0u, kb, ct, 0u)]);
        const float4 v1 = *reinterpret_cast<const float4*>(&W.terms[term_index(0u, kb, ct, 4u)]);
        chb.add(v0.x);
        chb.add(v0.y);
        chb.add(v0.z);
        chb.add(v0.w);
        chb.add(v1.x);
        chb.add(v1.y);
        chb.add(v1.z);
        chb.add(v1.w);
      }
    }
  }
  if (__shfl(tcount, static_cast<int>(ct << 3)) != 0) {
    const gptr<float> fo = reinterpret_cast<gptr<float>>(G(d.fields) + 3u * cg);
    fo[ka] = cha.value();
    if (has_b) {
      fo[kb] = chb.value();
      if (kb == 9u) {
        fo[10] = 0.0f;
        fo[11] = 0.0f;
      }
    }
  }
}

// kCpr: chunks of 8 x 8 pairs per round (the pair tile).  Occupancy: with
// 8 chunks per round 6 waves per SIMD for the unfused forms (their LDS
// allows 6 blocks per CU), 5 for fused launches, which carry the ancestor
// path; with 4 (half the terms' LDS) 8 for the unfused forms.
constexpr uint32_t tree_occupancy(uint32_t kind, uint32_t cpr) {
  return kind == kTreeFused ? 5u : (cpr == 4u ? 8u : 6u);
}
template <bool kComp, uint32_t kKind, uint32_t kCpr>
__global__ __launch_bounds__(kTreeThreads, tree_occupancy(kKind, kCpr)) void k_tree(StepArgs a, uint32_t lo,
                                                                                    uint32_t hi, uint32_t lead) {
  static_assert(kCpr == 4u || kCpr == 8u, "4 or 8 chunks per round");
  constexpr bool kFused = kKind == kTreeFused;
  constexpr uint32_t kWaves = kTreeThreads / 64u;
  constexpr uint32_t kChunksPerRound = kCpr;
  // Single level: static LDS.  Fused: one dynamic region of
  // max(item_bytes(kCpr), (hi - lo) * kAncestorBytes) bytes, used by the
  // ancestor path and then by the item.
  __shared__ float4 win_s[kFused ? 1 : 216];
  // The M2L form's block 0 may run the top levels in this array instead.
  constexpr uint32_t kTermFloats = kChunksPerRound * 10u * 64u;
  constexpr uint32_t kTopFloats = (static_cast<uint32_t>(sizeof(TopLds)) + 3u) / 4u;
  __shared__ __attribute__((aligned(16))) float
      terms_s[kFused ? 1u : (kKind == kTreeM2l && kTopFloats > kTermFloats ? kTopFloats : kTermFloats)];
  extern __shared__ float4 dyn[];
  __shared__ float pf[12];
  __shared__ uint32_t votes[kTreeThreads / 64u];
  float4* win = kFused ? dyn : win_s;
  float* terms = kFused ? reinterpret_cast<float*>(dyn + 216) : terms_s;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = tid >> 6;

  // The M2L form of a one-system step can carry the top levels (lead > 0):
  // block 0 runs the top levels' pass, blocks 1..lead-1 take the items of
  // levels 2..r (r = depth - 3; their parents, at levels 1..r-1, are the
  // ones that pass lists) directly, one parent each, and the other blocks
  // the listed items of levels lo..hi (parents pushed by upward).
  if constexpr (kKind == kTreeM2l) {
    static_assert(sizeof(TopLds) <= sizeof(terms_s), "the top levels' LDS fits the terms array");
    if (lead != 0u && blockIdx.x == 0u) {
      top_levels<kComp, kTreeThreads>(a, 0u, *reinterpret_cast<TopLds*>(terms_s));
      return;
    }
  }
  const bool direct = kKind == kTreeM2l && blockIdx.x < lead;
  const uint32_t first_block = blockIdx.x - (kKind == kTreeM2l ? lead : 0u);
  const uint32_t nlisted = gridDim.x - (kKind == kTreeM2l ? lead : 0u);
  // System 0's descriptor is loaded at entry, with the work table: for a
  // one-system step it is the items' descriptor, so the first item need not
  // load it after its list entry (a launch always has a system 0).
  SysDesc d = a.sys[0];
  uint32_t cur_sys = a.nsys == 1u ? 0u : 0xFFFFFFFFu;

  // Items of levels hi, hi-1, ..., lo in that order.  The work table holds,
  // per level and stripe, the inclusive prefix of the stripe counts and the
  // stripe's list offset; it is read from memory once per block, so locating
  // an item costs LDS reads only.
  __shared__ uint32_t tab_incl[kLevels][kStripes];
  __shared__ uint32_t tab_off[kLevels][kStripes];
  if (kKind == kTreeSingle) lo = hi;
  if (tid < (hi - lo + 1u) * kStripes) {
    const uint32_t level = lo + tid / kStripes, j = tid % kStripes, e = (level - 1u) * kStripes + j;
    uint32_t incl = min(G(a.list_count)[e], G(a.stripe_cap)[e]);
    const uint32_t off = G(a.stripe_off)[e];
    // Prefix over the 32 stripes of a level: one aligned half-wave.
#pragma unroll
    for (uint32_t o = 1u; o < kStripes; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, kStripes);
      if (j >= o) incl += v;
    }
    tab_incl[level][j] = incl;
    tab_off[level][j] = off;
  }
  __syncthreads();
  uint32_t total = 0u;
  for (uint32_t level = lo; level <= hi; ++level) total += tab_incl[level][kStripes - 1u];
  // List position and level of item `it` (< total).
  auto locate = [&](uint32_t it, uint32_t& level) -> uint32_t {
    uint32_t rem = it;
    for (level = hi;; --level) {
      const uint32_t tl = tab_incl[level][kStripes - 1u];
      if (rem < tl || level == lo) break;
      rem -= tl;
    }
    const bool below = lane < kStripes && tab_incl[level][lane & (kStripes - 1u)] <= rem;
    const uint32_t j = static_cast<uint32_t>(__popcll(__ballot(below)));
    const uint32_t base = j == 0u ? 0u : tab_incl[level][j - 1u];
    return tab_off[level][j] + (rem - base);
  };

  // Pair lane: target tt (child of the parent), source cc of the chunk.
  const uint32_t tt = lane >> 3, cc = lane & 7u;
  const uint32_t wtx = 2u + (tt & 1u), wty = 2u + ((tt >> 1) & 1u), wtz = 2u + (tt >> 2);
  // Fixed per lane: the window offset of source child cc within a source
  // parent's 2x2x2 block, the window index of target tt, and bit q set for
  // the source parents q (pz, py, px order) whose child cc is adjacent to
  // target tt (per axis: the middle parent always; the low one when the
  // child is its high cell and the target the low cell of the middle, the
  // high one conversely).
  const uint32_t lane_src = (cc & 1u) + 6u * ((cc >> 1) & 1u) + 36u * (cc >> 2);
  const uint32_t lane_tgt = wtx + 6u * (wty + 6u * wtz);
  uint32_t near_q = 0u;
  {
    auto adjacent = [](uint32_t c, uint32_t w) -> uint32_t {
      return 2u | (c == 1u && w == 2u ? 1u : 0u) | (c == 0u && w == 3u ? 4u : 0u);
    };
    const uint32_t mx = adjacent(cc & 1u, wtx), my = adjacent((cc >> 1) & 1u, wty), mz = adjacent(cc >> 2, wtz);
    // Outer product of the axis masks: the x mask repeated at bit 3 qy for
    // each qy in my, that row repeated at bit 9 qz for each qz in mz (the
    // copies do not overlap, so the products carry nothing).
    const uint32_t ys = (my & 1u) | ((my & 2u) << 2) | ((my & 4u) << 4);
    const uint32_t zs = (mz & 1u) | ((mz & 2u) << 8) | ((mz & 4u) << 16);
    near_q = (mx * ys) * zs;
  }
  const uint32_t wave_s = __builtin_amdgcn_readfirstlane(wave);
  // A wave's chunks are the same for every item: q = r0 + wave * per-wave +
  // j; the window offsets of their source-parent blocks come packed from a
  // table built at compile time (one byte per round and j).
  constexpr uint32_t kPerWave = kChunksPerRound / kWaves;
  const uint64_t block_off = tree_block_offsets<kCpr>(wave_s);
  // Chain threads 0..79: target tid % 8, component tid / 8.
  const bool chain = tid < 80u;
  const uint32_t ct = tid & 7u, ck = tid >> 3;

  // Each item's list entry is loaded one item ahead; the system descriptor
  // is reloaded only when the system changes.
  uint32_t next_level = lo, next_item = 0u;
  uint32_t it0 = first_block, stride = nlisted, items = total;
  if constexpr (kKind == kTreeFinish && kTermFloats * 4u >= kWaves * sizeof(ShortWaveLds)) {
    // One system: the items of the highest level (its finest) are long (M2L
    // and near) and taken one per block; the blocks after them take the
    // items above the finest level (L2L and near only) four at a time, one
    // per wave, so that the launch needs about one round of resident blocks.
    // When the items need more than one round, the blocks are split in
    // proportion to the work (a long item about three groups of four short
    // ones), the long items still at most one per block.
    if (a.nsys == 1u && hi > lo && gridDim.x >= 2u) {
      const uint32_t nlong = tab_incl[hi][kStripes - 1u];
      const uint32_t ngroups = (total - nlong + kWaves - 1u) / kWaves;
      const uint64_t wlong = 3ull * nlong;
      const uint32_t share = static_cast<uint32_t>(uint64_t{gridDim.x} * wlong / (wlong + ngroups + 1ull));
      const uint32_t lblocks = ngroups == 0u ? gridDim.x : min(nlong, min(max(share, 1u), gridDim.x - 1u));
      if (blockIdx.x >= lblocks) {
        const SysDesc& d0 = d;
        if (!d0.live || d0.depth != hi) return;
        ShortWaveLds& W = reinterpret_cast<ShortWaveLds*>(terms_s)[wave_s];
        const uint32_t sblocks = gridDim.x - lblocks;
        for (uint32_t g = blockIdx.x - lblocks; g < ngroups; g += sblocks) {
          const uint32_t it = nlong + g * kWaves + wave_s;
          if (it >= total) continue;
          uint32_t level = lo;
          const uint32_t item = __builtin_amdgcn_readfirstlane(G(a.lists)[locate(it, level)]);
          level = __builtin_amdgcn_readfirstlane(level);
          if ((item >> 16) != 0u) continue;
          finish_short_wave<kComp>(a, d0, level, item & 0xFFFFu, lane, W);
        }
        return;
      }
      stride = lblocks;
      items = nlong;
    }
  }
  if (direct) {
    // Block 1 + k: parent k of level 1, then k - 8 of level 2, ...
    uint32_t k = blockIdx.x - 1u, pl = 1u;
    while (k >= (1u << (3u * pl))) {
      k -= 1u << (3u * pl);
      ++pl;
    }
    next_level = pl + 1u;
    next_item = pack_item(0u, k);
    it0 = 0u;
    stride = 1u;
    items = 1u;
  } else if (first_block < total) {
    next_item = G(a.lists)[locate(first_block, next_level)];
  }
  for (uint32_t it = it0; it < items; it += stride) {
    const uint32_t level = __builtin_amdgcn_readfirstlane(next_level);
    const uint32_t item = __builtin_amdgcn_readfirstlane(next_item);
    if (!direct && it + stride < items) next_item = G(a.lists)[locate(it + stride, next_level)];
    const uint32_t s = item >> 16;
    if (s >= a.nsys) continue;
    const uint32_t pidx = item & 0xFFFFu;
    if (s != cur_sys) {
      d = a.sys[s];
      cur_sys = s;
    }
    if (direct && !d.live) continue;
    // The M2L launch leaves each system's finest level to the finish launch;
    // there a target's M2L sum is evaluated at the finest level only.
    if (kKind == kTreeM2l && level >= d.depth) continue;
    const bool full = kKind != kTreeFinish || level == d.depth;
    const uint32_t dim = 1u << level;
    const uint32_t pdim = dim >> 1u;
    const uint32_t off = level_off(level);
    uint32_t px, py, pz;
    unlin(pidx, level - 1u, px, py, pz);

    // The parent's field: from memory when its level is below lo, else
    // recomputed along the chain of ancestors from level lo - 1 (fused), or
    // from the ancestors' M2L sums (finish; pf is read after the barrier).
    const bool recompute = kFused && level - 1u >= lo;
    if (kFused && recompute) ancestor_fields<kComp>(d, lo, level, px, py, pz, reinterpret_cast<unsigned char*>(dyn), pf, tid);
    PathSums path{};
    if (kKind == kTreeFinish && wave_s == 0u) path = path_loads(d, level, px, py, pz, lane);
    // Above the finest level a finish chain starts from its target's M2L
    // sum (an empty target's slot is stale and its field is not stored).
    float m2l_sum = 0.0f;
    if (kKind == kTreeFinish && !full && chain) {
      const uint32_t x = 2u * px + (ct & 1u), y = 2u * py + ((ct >> 1) & 1u), z = 2u * pz + (ct >> 2);
      m2l_sum = reinterpret_cast<gptr<float>>(G(d.m2l) + 3u * (off + lin(x, y, z, dim)))[ck];
    }
    const uint32_t tx = 2u * px + (tt & 1u), ty = 2u * py + ((tt >> 1) & 1u), tz = 2u * pz + (tt >> 2);
    // Direct items below level r read their level before the top levels'
    // pass writes it: counts and records from the children's moments.
    const bool derived = direct && level + 3u < d.depth;
    const int tcount = derived ? children_moment(d, level, tx, ty, tz).x : cell_count(d, a.parity, level, lin(tx, ty, tz, dim));
    if (direct) {
      // A direct item exists when its parent holds more than 16 bodies:
      // the sum of its 8 targets' counts (lanes 0, 8, ..., 56).
      uint32_t parent_count = 0u;
#pragma unroll
      for (int l = 0; l < 64; l += 8) parent_count += static_cast<uint32_t>(__builtin_amdgcn_readlane(tcount, l));
      if (parent_count <= static_cast<uint32_t>(kLeafCapacity)) continue;
    }
    const bool bounded =
        derived ? stage_window_from_children(win, d, level, px, py, pz, tid) : stage_window(win, d, level, px, py, pz, tid);
    const float tcx = center1(tx, dim), tcy = center1(ty, dim), tcz = center1(tz, dim);
    const bool leaf = tcount != 0 && (tcount <= kLeafCapacity || level == d.depth);
    // 32 chunks of 8 sources: the 27 source-parent blocks in pz, py, px
    // order (M2L; parents outside the level were staged as empty), then 4
    // blocks of 8 near offsets (leaf targets only), then one empty chunk.
    // Zero terms leave the sums unchanged (a running sum that starts at +0
    // never becomes -0, so adding +0 or -0 is the identity), so every item
    // walks all 32.  The waves' votes (any leaf target, any unbounded
    // source) are combined after the barrier.
    if (kKind == kTreeFinish && wave_s == 0u) {
      const float f = path_field(path, level, px, py, pz, lane);
      if (lane < 10u) pf[lane] = f;
    }
    const uint32_t vote = (__any(leaf) ? 1u : 0u) | (__any(!bounded) ? 2u : 0u);
    if (lane == 0u) votes[wave] = vote;
    __syncthreads();
    uint32_t all_votes = 0u;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) all_votes |= votes[w];
    const bool any_leaf = (all_votes & 1u) != 0u;
    // cell_epsilon, src/rainbow_fmm.wgsl:507 (0.06 * WORLD_HALF_EXTENT / dim).
    const float cell_epsilon = (0.06f * kWorldHalfExtent) * inv_pow2(dim);
    const float eps2_m2l = cell_epsilon * cell_epsilon;
    const float near_epsilon =
        fmaxf(d.softening, (0.72f * kWorldHalfExtent) * inv_pow2(dim));
    const float eps2_near = near_epsilon * near_epsilon;
    // With every source coordinate bounded by 2^60 and eps2 in [2^-94,
    // 2^100], each pair's r2 lies in [2^-94, FLT_MAX]: the pair fields use
    // the unguarded inverse square root, and lanes without a pair evaluate
    // a zero-mass source (terms of +-0) instead of branching.
    const bool fast = (all_votes & 2u) == 0u && eps2_m2l >= 0x1p-94f && eps2_near >= 0x1p-94f &&
                      eps2_m2l <= 0x1p100f && eps2_near <= 0x1p100f;

    Chain<kComp> ch;
    float inh = 0.0f;
    if (kKind != kTreeM2l && chain) {
      // L2L: this chain's component of the parent's field shifted to its
      // target's centre.
      float4 pa, pj0, pj1;
      if ((kFused && recompute) || kKind == kTreeFinish) {
        pa = make_float4(pf[0], pf[1], pf[2], pf[3]);
        pj0 = make_float4(pf[4], pf[5], pf[6], pf[7]);
        pj1 = make_float4(pf[8], pf[9], 0.0f, 0.0f);
      } else {
        load_field(G(d.fields), level_off(level - 1u) + pidx, pa, pj0, pj1);
      }
      const float dx = center1(2u * px + (ct & 1u), dim) - center1(px, pdim);
      const float dy = center1(2u * py + ((ct >> 1) & 1u), dim) - center1(py, pdim);
      const float dz = center1(2u * pz + (ct >> 2), dim) - center1(pz, pdim);
      inh = l2l_term(ck, pa, pj0, pj1, dx, dy, dz);
    }
    if (kKind == kTreeFinish && !full && chain) ch.reset(m2l_sum);
    __syncthreads();

    // Chunks that can hold a pair: the M2L blocks but the parent's own (13,
    // every pair adjacent) where this item evaluates M2L, and the near
    // chunks 27-30 when a target is a leaf (chunk 31 is padding).  The
    // other chunks' terms are zero.
    auto chunk_has_terms = [&](uint32_t q) -> bool {
      return q < 27u ? (full && q != 13u) : (kKind != kTreeM2l && any_leaf && q < 31u);
    };
    // Items without M2L chunks start at the round of chunk 24 (whose chunks
    // 24-26 they skip).
#pragma unroll 1
    for (uint32_t r0 = full ? 0u : 24u; r0 < 32u; r0 += kChunksPerRound) {
      // Each wave evaluates two chunks' 8 x 8 pair fields.
#pragma unroll
      for (uint32_t j = 0; j < kChunksPerRound / kWaves; ++j) {
        const uint32_t slot = wave_s * (kChunksPerRound / kWaves) + j;
        const uint32_t q = r0 + slot;  // uniform
        // The M2L form neither stores nor adds a chunk without terms (they
        // would all be zero); the other forms store zeros and keep their
        // chains' additions straight-line, which measured faster.
        if (kKind == kTreeM2l && !chunk_has_terms(q)) continue;
        const bool eval = chunk_has_terms(q);
        // This lane's source in the window and whether its pair counts.
        uint32_t widx = 0u;
        bool active = false;
        float eps2 = eps2_m2l;
        if (q < 27u) {
          // Source parent q's block starts at (2 qx, 2 qy, 2 qz).
          widx = static_cast<uint32_t>((block_off >> (8u * (r0 / kChunksPerRound * kPerWave + j))) & 0xFFu) +
                 lane_src;
          active = tcount != 0 && ((near_q >> q) & 1u) == 0u;
        } else {
          // Near field of leaf targets (near_all, src/rainbow_fmm.wgsl:
          // 533-571): offsets dz, dy, dx; the centre is skipped; neighbours
          // outside the level were staged as empty.
          const uint32_t idx = (q - 27u) * 8u + cc;
          if (idx < 27u) {
            widx = lane_tgt + (idx % 3u + 6u * ((idx / 3u) % 3u) + 36u * (idx / 9u)) - 43u;
            active = leaf && idx != 13u;
          }
          eps2 = eps2_near;
        }
        float t[10];
        if (eval && fast) {
          // Every lane evaluates; lanes without a pair see a zero mass.
          float4 sr = win[widx];
          if (!active) sr.w = 0.0f;
          pair_terms<true>(sr, tcx, tcy, tcz, eps2, t);
        } else {
#pragma unroll
          for (int k = 0; k < 10; ++k) t[k] = 0.0f;
          if (eval) {
            const float4 sr = win[widx];
            if (active && sr.w != 0.0f) pair_terms<false>(sr, tcx, tcy, tcz, eps2, t);
          }
        }
        // Terms of chunk `slot` stored per chain (target tt, component k)
        // as 8 consecutive sources: the chains then read contiguous runs.
#pragma unroll
        for (uint32_t k = 0; k < 10u; ++k) terms[term_index(slot, k, tt, cc)] = t[k];
      }
      __syncthreads();
      // The chains add the round's chunks in order; the L2L term goes in
      // between the last M2L chunk (26) and the first near chunk (27).
      if (chain) {
#pragma unroll
        for (uint32_t u = 0; u < kChunksPerRound; ++u) {
          if (kKind != kTreeM2l && r0 + u == 27u) ch.reset(ch.value() + inh);
          if (kKind == kTreeM2l && !chunk_has_terms(r0 + u)) continue;
          const float4 v0 = *reinterpret_cast<const float4*>(&terms[term_index(u, ck, ct, 0u)]);
          const float4 v1 = *reinterpret_cast<const float4*>(&terms[term_index(u, ck, ct, 4u)]);
          ch.add(v0.x);
          ch.add(v0.y);
          ch.add(v0.z);
          ch.add(v0.w);
          ch.add(v1.x);
          ch.add(v1.y);
          ch.add(v1.z);
          ch.add(v1.w);
        }
      }
      __syncthreads();
    }

    // Store each chain's component of its (non-empty) target's field; the
    // target's count is the one lane 8 ct of the wave loaded for the pairs.
    const int ccount = __shfl(tcount, static_cast<int>(ct << 3));
    if (chain) {
      const uint32_t x = 2u * px + (ct & 1u), y = 2u * py + ((ct >> 1) & 1u), z = 2u * pz + (ct >> 2);
      const uint32_t g = off + lin(x, y, z, dim);
      if (ccount != 0) {
        const gptr<float> f = reinterpret_cast<gptr<float>>(G(kKind == kTreeM2l ? d.m2l : d.fields) + 3u * g);
        f[ck] = ch.value();
        if (kKind != kTreeM2l && ck == 9u) {
          f[10] = 0.0f;
          f[11] = 0.0f;
        }
      }
    }
  }
}

// ------------------------------------------------------------------------
// tree pass with wave items: one item per wave
// ------------------------------------------------------------------------

// The unfused forms of k_tree (kTreeSingle, kTreeM2l, kTreeFinish) with an
// item per wave instead of per block: each wave of a block walks items of
// its own, stages the item's 6^3 source window in its own LDS, evaluates
// one chunk of 8 x 8 pair terms at a time into its own term tile (lane:
// target lane / 8, source lane mod 8) and then adds that chunk to the
// item's 80 chains (lane l: target l mod 8, component l / 8; lanes 0..15
// also component 8 + l / 8), in the reference order.  After the work table
// no block barrier is taken, so no wave waits for another's item.  A wave's
// LDS accesses execute in order, so a tile is rewritten only after the
// previous chunk's reads.
struct WaveTreeLds {
  float4 win[216];
  float terms[640];
};

constexpr uint32_t kTreeWaves = kTreeThreads / 64u;
constexpr uint32_t kWaveTreeOccupancy = 6u;  // blocks per CU (LDS: 4 tiles + the work table)

template <bool kComp, uint32_t kKind>
__global__ __launch_bounds__(kTreeThreads, kWaveTreeOccupancy) void k_tree_w(StepArgs a, uint32_t lo, uint32_t hi,
                                                                              uint32_t lead) {
  static_assert(kKind == kTreeSingle || kKind == kTreeM2l || kKind == kTreeFinish, "the unfused forms");
  __shared__ __attribute__((aligned(16))) WaveTreeLds wl[kTreeWaves];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);

  // The M2L form of a one-system step can carry the top levels (lead > 0):
  // block 0 runs the top levels' pass; the first lead - 1 waves of the
  // blocks after it take the items of levels 2..r (r = depth - 3; parents
  // at levels 1..r-1) directly, one possible parent each, and the other
  // waves the listed items of levels lo..hi.
  if constexpr (kKind == kTreeM2l) {
    static_assert(sizeof(TopLds) <= sizeof(wl), "the top levels' LDS fits the wave tiles");
    if (lead != 0u && blockIdx.x == 0u) {
      top_levels<kComp, kTreeThreads>(a, 0u, *reinterpret_cast<TopLds*>(wl));
      return;
    }
  }
  const uint32_t lead_blocks = kKind == kTreeM2l && lead != 0u ? 1u : 0u;
  const uint32_t ndirect = kKind == kTreeM2l && lead != 0u ? lead - 1u : 0u;
  const uint32_t gw = (blockIdx.x - lead_blocks) * kTreeWaves + wave;
  const uint32_t nwaves = (gridDim.x - lead_blocks) * kTreeWaves;
  const bool direct = gw < ndirect;
  const uint32_t nlisted = nwaves > ndirect ? nwaves - ndirect : 0u;
  // System 0's descriptor is loaded at entry (as in k_tree).
  SysDesc d = a.sys[0];
  uint32_t cur_sys = a.nsys == 1u ? 0u : 0xFFFFFFFFu;

  // The work table, as in k_tree: per level and stripe the inclusive prefix
  // of the stripe counts and the stripe's list offset.
  __shared__ uint32_t tab_incl[kLevels][kStripes];
  __shared__ uint32_t tab_off[kLevels][kStripes];
  if (kKind == kTreeSingle) lo = hi;
  if (tid < (hi - lo + 1u) * kStripes) {
    const uint32_t level = lo + tid / kStripes, j = tid % kStripes, e = (level - 1u) * kStripes + j;
    uint32_t incl = min(G(a.list_count)[e], G(a.stripe_cap)[e]);
    const uint32_t off = G(a.stripe_off)[e];
#pragma unroll
    for (uint32_t o = 1u; o < kStripes; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, kStripes);
      if (j >= o) incl += v;
    }
    tab_incl[level][j] = incl;
    tab_off[level][j] = off;
  }
  __syncthreads();
  uint32_t total = 0u;
  for (uint32_t level = lo; level <= hi; ++level) total += tab_incl[level][kStripes - 1u];
  auto locate = [&](uint32_t it, uint32_t& level) -> uint32_t {
    uint32_t rem = it;
    for (level = hi;; --level) {
      const uint32_t tl = tab_incl[level][kStripes - 1u];
      if (rem < tl || level == lo) break;
      rem -= tl;
    }
    const bool below = lane < kStripes && tab_incl[level][lane & (kStripes - 1u)] <= rem;
    const uint32_t j = static_cast<uint32_t>(__popcll(__ballot(below)));
    const uint32_t base = j == 0u ? 0u : tab_incl[level][j - 1u];
    return tab_off[level][j] + (rem - base);
  };

  WaveTreeLds& W = wl[wave];
  // Pair lane: target tt, source cc of the chunk (as in k_tree).
  const uint32_t tt = lane >> 3, cc = lane & 7u;
  const uint32_t wtx = 2u + (tt & 1u), wty = 2u + ((tt >> 1) & 1u), wtz = 2u + (tt >> 2);
  const uint32_t lane_src = (cc & 1u) + 6u * ((cc >> 1) & 1u) + 36u * (cc >> 2);
  const uint32_t lane_tgt = wtx + 6u * (wty + 6u * wtz);
  uint32_t near_q = 0u;
  {
    auto adjacent = [](uint32_t c, uint32_t w) -> uint32_t {
      return 2u | (c == 1u && w == 2u ? 1u : 0u) | (c == 0u && w == 3u ? 4u : 0u);
    };
    const uint32_t mx = adjacent(cc & 1u, wtx), my = adjacent((cc >> 1) & 1u, wty), mz = adjacent(cc >> 2, wtz);
    const uint32_t ys = (my & 1u) | ((my & 2u) << 2) | ((my & 4u) << 4);
    const uint32_t zs = (mz & 1u) | ((mz & 2u) << 8) | ((mz & 4u) << 16);
    near_q = (mx * ys) * zs;
  }
  // Chain lanes: chain A (target ct, component ka), chain B for lanes 0..15
  // (component kb = 8 + lane / 8).
  const uint32_t ct = lane & 7u, ka = lane >> 3;
  const bool has_b = lane < 16u;
  const uint32_t kb = has_b ? 8u + (lane >> 3) : 9u;

  uint32_t next_level = lo, next_item = 0u;
  uint32_t it0 = gw - ndirect, stride = nlisted, items = total;
  if (direct) {
    // Wave k: parent k of level 1, then k - 8 of level 2, ...
    uint32_t k = gw, pl = 1u;
    while (k >= (1u << (3u * pl))) {
      k -= 1u << (3u * pl);
      ++pl;
    }
    next_level = pl + 1u;
    next_item = pack_item(0u, k);
    it0 = 0u;
    stride = 1u;
    items = 1u;
  } else if (it0 < total) {
    next_item = G(a.lists)[locate(it0, next_level)];
  }
  for (uint32_t it = it0; it < items; it += stride) {
    const uint32_t level = __builtin_amdgcn_readfirstlane(next_level);
    const uint32_t item = __builtin_amdgcn_readfirstlane(next_item);
    if (!direct && it + stride < items) next_item = G(a.lists)[locate(it + stride, next_level)];
    const uint32_t s = item >> 16;
    if (s >= a.nsys) continue;
    const uint32_t pidx = item & 0xFFFFu;
    if (s != cur_sys) {
      d = a.sys[s];
      cur_sys = s;
    }
    if (direct && !d.live) continue;
    if (kKind == kTreeM2l && level >= d.depth) continue;
    const bool full = kKind != kTreeFinish || level == d.depth;
    const uint32_t dim = 1u << level;
    const uint32_t pdim = dim >> 1u;
    const uint32_t off = level_off(level);
    uint32_t px, py, pz;
    unlin(pidx, level - 1u, px, py, pz);

    PathSums path{};
    if (kKind == kTreeFinish) path = path_loads(d, level, px, py, pz, lane);
    // Chain lanes' target, and its M2L sum above the finest level.
    const uint32_t cx = 2u * px + (ct & 1u), cy = 2u * py + ((ct >> 1) & 1u), cz = 2u * pz + (ct >> 2);
    const uint32_t cg = off + lin(cx, cy, cz, dim);
    float m2l_a = 0.0f, m2l_b = 0.0f;
    if (kKind == kTreeFinish && !full) {
      const gptr<float> m = reinterpret_cast<gptr<float>>(G(d.m2l) + 3u * cg);
      m2l_a = m[ka];
      m2l_b = m[kb];
    }
    const uint32_t tx = 2u * px + (tt & 1u), ty = 2u * py + ((tt >> 1) & 1u), tz = 2u * pz + (tt >> 2);
    const bool derived = direct && level + 3u < d.depth;
    const int tcount = derived ? children_moment(d, level, tx, ty, tz).x : cell_count(d, a.parity, level, lin(tx, ty, tz, dim));
    if (direct) {
      uint32_t parent_count = 0u;
#pragma unroll
      for (int l = 0; l < 64; l += 8) parent_count += static_cast<uint32_t>(__builtin_amdgcn_readlane(tcount, l));
      if (parent_count <= static_cast<uint32_t>(kLeafCapacity)) continue;
    }
    wave_sync();  // the last item's window reads come first
    const bool bounded = derived ? stage_window_from_children<64u>(W.win, d, level, px, py, pz, lane)
                                 : stage_window<64u>(W.win, d, level, px, py, pz, lane);
    wave_sync();
    const float tcx = center1(tx, dim), tcy = center1(ty, dim), tcz = center1(tz, dim);
    const bool leaf = tcount != 0 && (tcount <= kLeafCapacity || level == d.depth);
    const bool any_leaf = __any(leaf);
    const float cell_epsilon = (0.06f * kWorldHalfExtent) * inv_pow2(dim);
    const float eps2_m2l = cell_epsilon * cell_epsilon;
    const float near_epsilon = fmaxf(d.softening, (0.72f * kWorldHalfExtent) * inv_pow2(dim));
    const float eps2_near = near_epsilon * near_epsilon;
    const bool fast = __all(bounded) && eps2_m2l >= 0x1p-94f && eps2_near >= 0x1p-94f && eps2_m2l <= 0x1p100f &&
                      eps2_near <= 0x1p100f;

    // L2L: each chain's component of the parent's field shifted to its
    // target's centre (the parent's field from the ancestors' M2L sums in
    // the finish form, from memory in the single-level form).
    float inh_a = 0.0f, inh_b = 0.0f;
    if (kKind != kTreeM2l) {
      float4 pa, pj0, pj1;
      if (kKind == kTreeFinish) {
        const float f = path_field(path, level, px, py, pz, lane);
        auto at = [&](int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(f), l)); };
        pa = make_float4(at(0), at(1), at(2), at(3));
        pj0 = make_float4(at(4), at(5), at(6), at(7));
        pj1 = make_float4(at(8), at(9), 0.0f, 0.0f);
      } else {
        load_field(G(d.fields), level_off(level - 1u) + pidx, pa, pj0, pj1);
      }
      const float dx = center1(cx, dim) - center1(px, pdim);
      const float dy = center1(cy, dim) - center1(py, pdim);
      const float dz = center1(cz, dim) - center1(pz, pdim);
      inh_a = l2l_term(ka, pa, pj0, pj1, dx, dy, dz);
      inh_b = l2l_term(kb, pa, pj0, pj1, dx, dy, dz);
    }
    Chain<kComp> cha, chb;
    if (kKind == kTreeFinish && !full) {
      cha.reset(m2l_a);
      chb.reset(m2l_b);
    }
    bool inh_added = kKind == kTreeM2l;
    // Chunks 0..26: source-parent blocks (pz, py, px order), the parent's
    // own block (13) without a pair; 27..30: the near offsets of leaf
    // targets.  Chunks without a pair are skipped (their terms are zero).
#pragma unroll 1
    for (uint32_t q = full ? 0u : 27u; q < 31u; ++q) {
      if (q < 27u ? q == 13u : (kKind == kTreeM2l || !any_leaf)) continue;
      if (!inh_added && q >= 27u) {
        cha.reset(cha.value() + inh_a);
        chb.reset(chb.value() + inh_b);
        inh_added = true;
      }
      uint32_t widx = 0u;
      bool active = false;
      float eps2 = eps2_m2l;
      if (q < 27u) {
        widx = 2u * (q % 3u) + 12u * ((q / 3u) % 3u) + 72u * (q / 9u) + lane_src;
        active = tcount != 0 && ((near_q >> q) & 1u) == 0u;
      } else {
        const uint32_t idx = (q - 27u) * 8u + cc;
        if (idx < 27u) {
          widx = lane_tgt + (idx % 3u + 6u * ((idx / 3u) % 3u) + 36u * (idx / 9u)) - 43u;
          active = leaf && idx != 13u;
        }
        eps2 = eps2_near;
      }
      float t[10];
      if (fast) {
        float4 sr = W.win[widx];
        if (!active) sr.w = 0.0f;
        pair_terms<true>(sr, tcx, tcy, tcz, eps2, t);
      } else {
#pragma unroll
        for (int k = 0; k < 10; ++k) t[k] = 0.0f;
        const float4 sr = W.win[widx];
        if (active && sr.w != 0.0f) pair_terms<false>(sr, tcx, tcy, tcz, eps2, t);
      }
      wave_sync();  // the last chunk's term reads come first
#pragma unroll
      for (uint32_t k = 0; k < 10u; ++k) W.terms[term_index(0u, k, tt, cc)] = t[k];
      wave_sync();
      {
        const float4 v0 = *reinterpret_cast<const float4*>(&W.terms[term_index(0u, ka, ct, 0u)]);
        const float4 v1 = *reinterpret_cast<const float4*>(&W.terms[term_index(0u, ka, ct, 4u)]);
        cha.add(v0.x);
        cha.add(v0.y);
        cha.add(v0.z);
        cha.add(v0.w);
        cha.add(v1.x);
        cha.add(v1.y);
        cha.add(v1.z);
        cha.add(v1.w);
      }
      if (has_b) {
        const float4 v0 = *reinterpret_cast<const float4*>(&W.terms[term_index(0u, kb, ct, 0u)]);
        const float4 v1 = *reinterpret_cast<const float4*>(&W.terms[term_index(0u, kb, ct, 4u)]);
        chb.add(v0.x);
        chb.add(v0.y);
        chb.add(v0.z);
        chb.add(v0.w);
        chb.add(v1.x);
        chb.add(v1.y);
        chb.add(v1.z);
        chb.add(v1.w);
      }
    }
    if (!inh_added) {
      cha.reset(cha.value() + inh_a);
      chb.reset(chb.value() + inh_b);
    }
    // Store each chain's component of its (non-empty) target's field; the
    // target's count is the one lane 8 ct loaded for the pairs.
    const int ccount = __shfl(tcount, static_cast<int>(ct << 3));
    if (ccount != 0) {
      const gptr<float> f = reinterpret_cast<gptr<float>>(G(kKind == kTreeM2l ? d.m2l : d.fields) + 3u * cg);
      f[ka] = cha.value();
      if (has_b) {
        f[kb] = chb.value();
        if (kKind != kTreeM2l && kb == 9u) {
          f[10] = 0.0f;
          f[11] = 0.0f;
        }
      }
    }
  }
}

// ------------------------------------------------------------------------
// integrate, src/rainbow_fmm.wgsl:613-771
// ------------------------------------------------------------------------

struct Leaf {
  uint32_t cell;   // global cell index of the chosen leaf
  uint32_t dim;    // its level's side
  uint32_t count;  // bodies in it
  float ax, ay, az;
  float phi;
};

// Adaptive descent (src/rainbow_fmm.wgsl:622-639), local expansion
// (:640-642) and own-leaf monopole (:644-657).  phi is the matching
// potential phi_c - a_c.d - d.(J d)/2 - G (count-1)/N / r_soft, which the
// reference does not form (energy diagnostic only).
// kDepth: the system's depth when known at compile time (0: d.depth).
template <bool kComp, bool kPhi, uint32_t kDepth = 0u>
__device__ __forceinline__ Leaf leaf_field(const SysDesc& d, uint32_t parity, float px, float py, float pz) {
  const uint32_t depth = kDepth != 0u ? kDepth : d.depth;
  const float ux = unit_coord(px), uy = unit_coord(py), uz = unit_coord(pz);
  Leaf out;
  // The cell of every level depends only on the position, so all counts are
  // fetched at once; the leaf is then the first level with count <= 16, or
  // the finest level.
  const uint32_t fx = cell1(ux, 1u << depth), fy = cell1(uy, 1u << depth), fz = cell1(uz, 1u << depth);
  const gptr<int4> finest = finest_of(d, parity);
  int lc[kLevels];
#pragma unroll
  for (uint32_t level = 1u; level < kLevels; ++level) {
    if (level <= depth) {
      const uint32_t sh = depth - level;
      const uint32_t idx = lin(fx >> sh, fy >> sh, fz >> sh, 1u << level);
      lc[level] = level == depth ? finest[idx].x : G(d.moments)[level_off(level) + idx].x;
    }
  }
  // The finest cell is the usual leaf: its field and record are loaded
  // with the counts, and used when the search ends there.
  float4 sa, sj0, sj1;
  const uint32_t fcell = level_off(depth) + lin(fx, fy, fz, 1u << depth);
  load_field(G(d.fields), fcell, sa, sj0, sj1);
  const float4 scom = G(d.source)[fcell];
  uint32_t chosen = depth;
#pragma unroll
  for (uint32_t level = kLevels - 1u; level >= 1u; --level)
    if (level < depth && lc[level] <= kLeafCapacity) chosen = level;
  const uint32_t sh = depth - chosen;
  const uint32_t c[3] = {fx >> sh, fy >> sh, fz >> sh};
  out.dim = 1u << chosen;
  out.cell = level_off(chosen) + lin(c[0], c[1], c[2], out.dim);
  const uint32_t cell = out.cell;
  float4 fa = sa, fj0 = sj0, fj1 = sj1;
  if (chosen != depth) load_field(G(d.fields), cell, fa, fj0, fj1);
  const float dx = px - center1(c[0], out.dim);
  const float dy = py - center1(c[1], out.dim);
  const float dz = pz - center1(c[2], out.dim);
  int chosen_count = lc[1];
#pragma unroll
  for (uint32_t level = 2u; level < kLevels; ++level)
    if (level == chosen) chosen_count = lc[level];
  const uint32_t count = static_cast<uint32_t>(chosen_count);
  out.count = count;
  float self[3] = {0.0f, 0.0f, 0.0f};
  float self_phi = 0.0f;
  const bool has_self = count > 1u;
  if (has_self) {
    const float4 com = chosen == depth ? scom : G(d.source)[cell];
    const float ex = com.x - px, ey = com.y - py, ez = com.z - pz;
    const float epsilon =
        fmaxf(d.softening, (0.62f * kWorldHalfExtent) * inv_pow2(out.dim));
    const float inv_r = rsqrt_rn(((ex * ex + ey * ey) + ez * ez) + epsilon * epsilon);
    const float self_mass = static_cast<float>(count - 1u) / static_cast<float>(d.n);
    self[0] = d.gravity * self_mass * ex * inv_r * inv_r * inv_r;
    self[1] = d.gravity * self_mass * ey * inv_r * inv_r * inv_r;
    self[2] = d.gravity * self_mass * ez * inv_r * inv_r * inv_r;
    if (kPhi) self_phi = -(d.gravity * self_mass) * inv_r;
  }
  const float jx = (fj0.x * dx + fj0.y * dy) + fj0.z * dz;
  const float jy = (fj0.y * dx + fj0.w * dy) + fj1.x * dz;
  const float jz = (fj0.z * dx + fj1.x * dy) + fj1.y * dz;
  if (!kComp) {
    out.ax = fa.x + jx;
    out.ay = fa.y + jy;
    out.az = fa.z + jz;
    if (has_self) {
      out.ax = out.ax + self[0];
      out.ay = out.ay + self[1];
      out.az = out.az + self[2];
    }
  } else {
    const float row[3][3] = {{fj0.x, fj0.y, fj0.z}, {fj0.y, fj0.w, fj1.x}, {fj0.z, fj1.x, fj1.y}};
    const float base[3] = {fa.x, fa.y, fa.z};
    const float dd[3] = {dx, dy, dz};
    float res[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Chain<true> s;
      s.reset(base[k]);
      s.add(row[k][0] * dd[0]);
      s.add(row[k][1] * dd[1]);
      s.add(row[k][2] * dd[2]);
      if (has_self) s.add(self[k]);
      res[k] = s.value();
    }
    out.ax = res[0];
    out.ay = res[1];
    out.az = res[2];
  }
  out.phi = 0.0f;
  if (kPhi)
    out.phi = ((fa.w - ((fa.x * dx + fa.y * dy) + fa.z * dz)) - 0.5f * ((dx * jx + dy * jy) + dz * jz)) +
              self_phi;
  return out;
}

__device__ __forceinline__ float length3(float x, float y, float z) {
  return sqrt_rn((x * x + y * y) + z * z);
}

// One body's integrate (src/rainbow_fmm.wgsl:613-771): the new state
// replaces p4 (pos, hue) and v4 (vel, excitation).
template <bool kComp, uint32_t kDepth = 0u>
__device__ __forceinline__ void integrate_body(const SysDesc& d, uint32_t parity, float4& p4, float4& v4) {
  float pos[3] = {p4.x, p4.y, p4.z};
  float velocity[3] = {v4.x, v4.y, v4.z};
  const Leaf leaf = leaf_field<kComp, false, kDepth>(d, parity, p4.x, p4.y, p4.z);
  float acceleration[3] = {leaf.ax, leaf.ay, leaf.az};
  const uint32_t count = leaf.count;

  // Shepherd control field, src/rainbow_fmm.wgsl:659-689 (off unless
  // control_point.w > 0.5).
  float control_influence = 0.0f;
  if (d.control_point[3] > 0.5f) {
    const float cdx = d.control_point[0] - pos[0];
    const float cdy = d.control_point[1] - pos[1];
    const float cdz = d.control_point[2] - pos[2];
    const float distance2 = (cdx * cdx + cdy * cdy) + cdz * cdz;
    const float control_radius = fmaxf(d.control_radius, 0.02f);
    control_influence = expf(-2.0f * distance2 / (control_radius * control_radius));
    const float control_strength = d.control_strength;
    const float isq = rsqrt_rn(distance2 + control_radius * control_radius * 0.018f);
    const float dir[3] = {cdx * isq, cdy * isq, cdz * isq};
    const float mode = d.control_axis_mode[3];
    if (mode < 1.5f) {
      for (int k = 0; k < 3; ++k)
        acceleration[k] = acceleration[k] + dir[k] * control_strength * control_influence;
    } else if (mode < 2.5f) {
      for (int k = 0; k < 3; ++k)
        acceleration[k] = acceleration[k] - dir[k] * control_strength * control_influence;
    } else {
      const float al =
          length3(d.control_axis_mode[0], d.control_axis_mode[1], d.control_axis_mode[2]);
      const float ax = d.control_axis_mode[0] / al;
      const float ay = d.control_axis_mode[1] / al;
      const float az = d.control_axis_mode[2] / al;
      const float tg[3] = {ay * cdz - az * cdy, az * cdx - ax * cdz, ax * cdy - ay * cdx};
      const float tl = length3(tg[0], tg[1], tg[2]);
      if (tl > 0.00001f)
        for (int k = 0; k < 3; ++k)
          acceleration[k] = acceleration[k] + tg[k] / tl * control_strength * control_influence;
    }
  }

  // Symplectic Euler with the excitation/hue state (src/rainbow_fmm.wgsl:691-770).
  const float dt = d.dt * d.speed;
  const float radius = fmaxf(length3(pos[0], pos[1], pos[2]), 0.0001f);
  const float velocity_length = fmaxf(length3(velocity[0], velocity[1], velocity[2]), 0.0001f);
  const float signed_radial_motion = clampf(
      ((pos[0] * velocity[0] + pos[1] * velocity[1]) + pos[2] * velocity[2]) / (radius * velocity_length),
      -1.0f, 1.0f);
  const float radial_fraction = fabsf(signed_radial_motion);
  const float radial_shock = radial_fraction * radial_fraction * radial_fraction;
  // clamp((count - 16) / 72, 0, 1); the dividend clamped to [-1, 73] leaves
  // the ramp unchanged and keeps the corrected product exact (every uint32
  // count is checked on the host by the tests).
  const float density_pulse = clampf(
      div_corrected(clampf(static_cast<float>(count) - static_cast<float>(kLeafCapacity), -1.0f, 73.0f), 72.0f,
                    1.0f / 72.0f),
      0.0f, 1.0f);
  const float field_strength = clampf(
      length3(acceleration[0], acceleration[1], acceleration[2]) * (radius + 0.12f) * 0.58f, 0.0f, 1.0f);
  const float target_excitation =
      clampf(0.025f + radial_shock * 0.56f + density_pulse * 0.26f + field_strength * 0.13f +
                 control_influence * 0.42f,
             0.0f, 1.0f);
  float excitation = v4.w;
  const float response = target_excitation > excitation ? 10.0f : 1.05f;
  const float w = clampf(dt * response, 0.0f, 1.0f);
  excitation = excitation * (1.0f - w) + target_excitation * w;

#pragma unroll
  for (int k = 0; k < 3; ++k) velocity[k] = velocity[k] + acceleration[k] * dt;
#pragma unroll
  for (int k = 0; k < 3; ++k) pos[k] = pos[k] + velocity[k] * dt;

  const float world_edge = kWorldHalfExtent * 0.985f;
  const float edge_band = 0.12f * kWorldHalfExtent;
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    const float over = fabsf(pos[axis]) - (world_edge - edge_band);
    if (over > 0.0f) {
      const float absorb = clampf(over / edge_band, 0.0f, 1.0f);
      velocity[axis] = velocity[axis] * (1.0f - absorb * 0.16f);
    }
    if (fabsf(pos[axis]) > world_edge) {
      pos[axis] = clampf(pos[axis], -world_edge, world_edge);
      velocity[axis] = velocity[axis] * -0.55f;
    }
  }

  // clamp((e - 0.30) / (0.70 - 0.30), 0, 1); the dividend clamped to [-1, 1]
  // leaves the ramp unchanged and keeps the corrected product exact (every
  // float e is checked on the host by the tests).
  constexpr float kBand = 0.70f - 0.30f;
  const float st = clampf(div_corrected(clampf(excitation - 0.30f, -1.0f, 1.0f), kBand, 1.0f / kBand), 0.0f, 1.0f);
  const float hue_drift = 1.0f - st * st * (3.0f - 2.0f * st);
  const float h = p4.w + dt * hue_drift *
                             (0.012f + signed_radial_motion * 0.030f + excitation * 0.036f +
                              control_influence * 0.080f);
  const float dynamical_hue = h - floorf(h);
  p4 = make_float4(pos[0], pos[1], pos[2], dynamical_hue);
  v4 = make_float4(velocity[0], velocity[1], velocity[2], excitation);
}

// Integrate fused with the next step's deposit: each block integrates the
// 1024 bodies of its deposit range and deposits their new positions into
// the other finest-level buffer (zeroed by this step's upward), with the
// next step's ranks when that step sorts.
//
// On a sorting step the bodies are read from the staging arrays the scatter
// filled in Morton order (and written back to the body arrays, with their
// caller indices), and each block takes the Morton range it owns from then
// on: [cell of its first body, cell of the next block's first body); a
// system's first block starts at 0 and its last ends at 2^18.
// kOne: the step holds one system, whose descriptor is then loaded with the
// block's entry instead of after it (one dependent load less per block).
// kThreads: the block's width over its 1024-body range (1024: one body per
// thread; 512: two, the second 512 after the first).
template <bool kComp, bool kOne, uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void k_integrate_deposit(StepArgs a) {
  constexpr uint32_t kPer = kDepositBodies / kThreads;
  __shared__ DepositLds L;
  __shared__ uint2 own_new;
  // The block's entry and descriptor are loaded before block 0 clears the
  // list counts (loads issued after a store could read it, so they would
  // become vector loads), and held before the live test: one scalar round
  // trip.
  const BodyBlock blk = a.deposit_blocks[xcd_tile(blockIdx.x, gridDim.x)];
  const SysDesc d = a.sys[kOne ? 0u : blk.sys];
  in_sgpr(blk.first);
  in_sgpr(d.live);
  in_sgpr(d.n);
  in_sgpr(d.depth);
  in_sgpr(d.sort_now);
  in_sgpr(d.pos);
  in_sgpr(d.vel);
  in_sgpr(d.moments);
  in_sgpr(d.fields);
  in_sgpr(d.source);
  in_sgpr(d.finest_alt);
  reset_lists(a);
  if (!d.live) return;
  const bool staged = sorts(a, d);
  const uint32_t dim = 1u << d.depth;
  float4 p[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t i = blk.first + k * kThreads + threadIdx.x;
    p[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 p4 = p[k], v4 = p[k];
    if (i < d.n) {
      p4 = staged ? G(d.tpos)[i] : G(d.pos)[i];
      v4 = staged ? G(d.tvel)[i] : G(d.vel)[i];
    }
    // One system: the deposit table is emptied while the first bodies'
    // loads are in flight (1M bodies: 0.4 us less; a batch, whose blocks
    // overlap more, measured slower and empties it in the deposit).
    if (kOne && k == 0u) deposit_clear<kThreads>(L);
    if (i < d.n) {
      if (staged) {
        G(d.perm)[i] = G(d.tperm)[i];
        if (k == 0u && threadIdx.x == 0u) {
          const uint32_t next = blk.first + kDepositBodies;
          const uint32_t lo = blk.first == 0u ? 0u : morton_of(p4, dim);
          const uint32_t hi = next >= d.n ? (1u << 18) : morton_of(G(d.tpos)[next], dim);
          own_new = make_uint2(lo, hi);
          G(d.own)[blk.first / kDepositBodies] = own_new;
        }
      }
      // Depth as a constant: the level loops of the leaf search unroll to
      // fixed shifts and offsets.
      switch (d.depth) {
        case 4u: integrate_body<kComp, 4u>(d, a.parity, p4, v4); break;
        case 5u: integrate_body<kComp, 5u>(d, a.parity, p4, v4); break;
        case 6u: integrate_body<kComp, 6u>(d, a.parity, p4, v4); break;
        default: integrate_body<kComp>(d, a.parity, p4, v4); break;
      }
      G(d.pos)[i] = p4;
      G(d.vel)[i] = v4;
      p[k] = p4;
    }
  }
  deposit_block<kThreads, kOne>(a, d, blk.first, p, a.parity ^ 1u, a.rank_next != 0u, L, staged ? &own_new : nullptr);
}

// ------------------------------------------------------------------------
// diagnostics: per-body field, FP64 energy and momentum reductions
// ------------------------------------------------------------------------

template <bool kComp>
__global__ __launch_bounds__(kIntegrateThreads) void k_accel(StepArgs a, uint32_t sys, float4* out) {
  const BodyBlock blk = a.integrate_blocks[blockIdx.x];
  if (blk.sys != sys) return;
  const SysDesc d = a.sys[blk.sys];
  const uint32_t i = blk.first + threadIdx.x;
  if (i >= d.n) return;
  const float4 p4 = G(d.pos)[i];
  const Leaf leaf = leaf_field<kComp, true>(d, a.parity, p4.x, p4.y, p4.z);
  out[G(d.perm)[i]] = make_float4(leaf.ax, leaf.ay, leaf.az, leaf.phi);
}

// Block partials of {KE, PE, Px, Py, Pz, Lx, Ly, Lz} in FP64 (equal masses
// 1/N), reduced in a fixed order so repeated runs agree bit for bit.
template <bool kComp>
__global__ __launch_bounds__(kIntegrateThreads) void k_diag_partial(StepArgs a, uint32_t sys,
                                                                     double* partial) {
  __shared__ double red[8][kIntegrateThreads];
  const BodyBlock blk = a.integrate_blocks[blockIdx.x];
  const SysDesc d = a.sys[blk.sys];
  const uint32_t i = blk.first + threadIdx.x;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (blk.sys == sys && i < d.n) {
    const float4 p4 = G(d.pos)[i];
    const float4 v4 = G(d.vel)[i];
    const Leaf leaf = leaf_field<kComp, true>(d, a.parity, p4.x, p4.y, p4.z);
    const double m = 1.0 / static_cast<double>(d.n);
    const double x = p4.x, y = p4.y, z = p4.z, vx = v4.x, vy = v4.y, vz = v4.z;
    v[0] = 0.5 * m * (vx * vx + vy * vy + vz * vz);
    v[1] = 0.5 * m * static_cast<double>(leaf.phi);
    v[2] = m * vx;
    v[3] = m * vy;
    v[4] = m * vz;
    v[5] = m * (y * vz - z * vy);
    v[6] = m * (z * vx - x * vz);
    v[7] = m * (x * vy - y * vx);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (uint32_t s = kIntegrateThreads / 2u; s > 0u; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 8; ++k) red[k][threadIdx.x] = red[k][threadIdx.x] + red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 8u) partial[8u * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The block partials summed by one block: each thread sums a strided share
// of the blocks for all eight quantities, then the threads' sums are added
// pairwise in LDS.
__global__ __launch_bounds__(256) void k_diag_final(const double* partial, uint32_t nblocks,
                                                    double* out) {
  __shared__ double red[8][256];
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < nblocks; b += 256u)
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = s[k] + partial[8u * b + k];
#pragma unroll
  for (int k = 0; k < 8; ++k) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (uint32_t h = 128u; h > 0u; h >>= 1) {
    if (threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < 8; ++k) red[k][threadIdx.x] = red[k][threadIdx.x] + red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 8u) out[threadIdx.x] = red[threadIdx.x][0];
}

// ------------------------------------------------------------------------
// FP64 direct sum (precision form): the softened Newtonian field of every
// body (mass 1/n) at target bodies, as the oracle's direct sum defines it.
// Blocks tile (256 targets) x (source chunk); partial sums per chunk are
// added in chunk order, so the result does not depend on scheduling.
// ------------------------------------------------------------------------

// The slot holding each target body (caller index targets[t]): every slot
// compares its caller index with the targets staged in LDS, so a target
// listed twice gets its slot twice.  slot_of is filled with ~0 beforehand; a
// slot that is never found stays ~0 and is clamped by the sum.
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ __launch_bounds__(256) void k_find_slots(SysDesc d, const uint32_t* targets, uint32_t nt,
                                                    uint32_t* slot_of) {
  __shared__ uint32_t tl[kDirectMaxTargetsPerPass];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const uint32_t c = i < d.n ? G(d.perm)[i] : 0xFFFFFFFFu;
  for (uint32_t base = 0; base < nt; base += kDirectMaxTargetsPerPass) {
    const uint32_t m = min(kDirectMaxTargetsPerPass, nt - base);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < m; k += 256u) tl[k] = targets[base + k];
    __syncthreads();
    if (i < d.n)
      for (uint32_t k = 0; k < m; ++k)
        if (tl[k] == c) slot_of[base + k] = i;
  }
}
__global__ __launch_bounds__(256) void k_direct_f64(SysDesc d, const uint32_t* slot_of, uint32_t nt,
                                                    uint32_t chunk, double* partial) {
  __shared__ double sx[256], sy[256], sz[256];
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const bool valid = t < nt;
  const uint32_t self = valid ? min(slot_of[t], d.n - 1u) : 0xFFFFFFFFu;
  double xi = 0.0, yi = 0.0, zi = 0.0;
  if (valid) {
    const float4 p = G(d.pos)[self];
    xi = p.x;
    yi = p.y;
    zi = p.z;
  }
  const double e2 = static_cast<double>(d.softening) * static_cast<double>(d.softening);
  const uint32_t j0 = blockIdx.y * chunk;
  const uint32_t j1 = min(d.n, j0 + chunk);
  double ax = 0.0, ay = 0.0, az = 0.0, phi = 0.0;
  for (uint32_t jt = j0; jt < j1; jt += 256u) {
    const uint32_t j = jt + threadIdx.x;
    if (j < j1) {
      const float4 q = G(d.pos)[j];
      sx[threadIdx.x] = q.x;
      sy[threadIdx.x] = q.y;
      sz[threadIdx.x] = q.z;
    }
    __syncthreads();
    const uint32_t m = min(256u, j1 - jt);
    for (uint32_t k = 0; k < m; ++k) {
      if (jt + k == self) continue;
      const double dx = sx[k] - xi, dy = sy[k] - yi, dz = sz[k] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + e2;
      const double inv_r = 1.0 / sqrt(r2);
      const double inv3 = inv_r * inv_r * inv_r;
      ax += dx * inv3;
      ay += dy * inv3;
      az += dz * inv3;
      phi -= inv_r;
    }
    __syncthreads();
  }
  if (valid) {
    double* o = partial + 4u * (static_cast<size_t>(blockIdx.y) * nt + t);
    o[0] = ax;
    o[1] = ay;
    o[2] = az;
    o[3] = phi;
  }
}

__global__ __launch_bounds__(256) void k_direct_final(SysDesc d, uint32_t nt, uint32_t nchunks,
                                                      const double* partial, double* out) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= nt) return;
  const double gm = static_cast<double>(d.gravity) / static_cast<double>(d.n);
  for (uint32_t c = 0; c < 4u; ++c) {
    double s = 0.0;
    for (uint32_t k = 0; k < nchunks; ++k) s += partial[4u * (static_cast<size_t>(k) * nt + t) + c];
    out[4u * t + c] = gm * s;
  }
}

// ------------------------------------------------------------------------
// particle transfer: the reference's interleaved Particle {pos_hue, vel_hue}
// (32 B) to and from the pos/vel streams, in caller order
// ------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_split_particles(SysDesc d, const float4* in) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= d.n) return;
  G(d.pos)[i] = in[2u * i];
  G(d.vel)[i] = in[2u * i + 1u];
  G(d.perm)[i] = i;
}

// ------------------------------------------------------------------------
// galaxy seeding: initialize, src/rainbow_fmm.wgsl:205-374
// ------------------------------------------------------------------------

// hash32 and random01, src/rainbow_fmm.wgsl:205-214.
__device__ __forceinline__ uint32_t galaxy_hash(uint32_t x) {
  x = ((x >> 16) ^ x) * 0x45d9f3bu;
  x = ((x >> 16) ^ x) * 0x45d9f3bu;
  return (x >> 16) ^ x;
}

__device__ __forceinline__ float galaxy_random(uint32_t seed) {
  return static_cast<float>(galaxy_hash(seed)) * (1.0f / 4294967296.0f);
}

// seeded_enclosed_mass, src/rainbow_fmm.wgsl:216-230: mass fraction inside
// `radius` of the 5% bulge, 70% halo, 18% disc and 7% polar populations.
__device__ __forceinline__ float galaxy_enclosed_mass(float radius) {
  const float bulge = 0.05f * powf(clampf(radius / 0.11f, 0.0f, 1.0f), 3.0f);
  const float halo = 0.70f * powf(clampf(radius / 0.70f, 0.0f, 1.0f), 3.0f);
  const float disc = 0.18f * powf(clampf((radius - 0.055f) / 0.60f, 0.0f, 1.0f), 1.0f / 0.72f);
  const float polar = 0.07f * powf(clampf((radius - 0.22f) / 0.31f, 0.0f, 1.0f), 1.0f / 0.86f);
  return fmaxf(((bulge + halo) + disc) + polar, 0.0001f);
}

__device__ __forceinline__ float3 galaxy_cross(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ float galaxy_length(float3 v) { return sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z); }

__device__ __forceinline__ float3 galaxy_normalize(float3 v) {
  const float l = galaxy_length(v);
  return make_float3(v.x / l, v.y / l, v.z / l);
}

// Body i of the reference's default scene: a live halo (with a compact
// bulge), a four-arm disc and four retrograde polar lanes.  The
// transcendental functions are the device library's (the WGSL leaves their
// accuracy to the implementation), so the seed matches the reference within
// a tolerance, not bit for bit.
__global__ __launch_bounds__(256) void k_seed_galaxy(SysDesc d) {
  constexpr float kTau = 6.28318530718f;
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= d.n) return;
  const float u0 = galaxy_random(i * 11u + 1u);
  const float u1 = galaxy_random(i * 11u + 2u);
  const float u2 = galaxy_random(i * 11u + 3u);
  const float u3 = galaxy_random(i * 11u + 4u);
  const float u4 = galaxy_random(i * 11u + 5u);
  const float u5 = galaxy_random(i * 11u + 6u);
  float3 pos, vel;
  float excitation;
  if (u5 < 0.75f) {
    // Halo (bulge when u5 < 0.05): uniform ball, circular speed about a
    // random axis.
    const float z = u0 * 2.0f - 1.0f;
    const float phi = kTau * u1;
    const float outer_radius = u5 < 0.05f ? 0.11f : 0.70f;
    const float radius = outer_radius * powf(u2, 0.3333333f);
    const float ring = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float3 radial_direction = make_float3(ring * cosf(phi), z, ring * sinf(phi));
    pos = make_float3(radius * radial_direction.x, radius * radial_direction.y, radius * radial_direction.z);
    const float tangent_z = u3 * 2.0f - 1.0f;
    const float tangent_phi = kTau * u4;
    const float tangent_ring = sqrtf(fmaxf(0.0f, 1.0f - tangent_z * tangent_z));
    const float3 random_direction =
        make_float3(tangent_ring * cosf(tangent_phi), tangent_z, tangent_ring * sinf(tangent_phi));
    float3 tangent = galaxy_cross(radial_direction, random_direction);
    if (galaxy_length(tangent) < 0.001f) tangent = galaxy_cross(radial_direction, make_float3(0.0f, 1.0f, 0.0f));
    if (galaxy_length(tangent) < 0.001f) tangent = galaxy_cross(radial_direction, make_float3(1.0f, 0.0f, 0.0f));
    const float orbital = 0.985f * sqrtf(d.gravity * galaxy_enclosed_mass(radius) / (radius + 0.018f));
    const float3 t = galaxy_normalize(tangent);
    vel = make_float3(orbital * t.x, orbital * t.y, orbital * t.z);
    excitation = u5 < 0.05f ? 0.09f : 0.025f + u4 * 0.025f;
  } else if (u5 < 0.93f) {
    // Four-arm disc, tilted by 0.2 about x.
    const float radius = 0.055f + 0.60f * powf(u0, 0.72f);
    const float arm = static_cast<float>(i & 3u) * (kTau / 4.0f);
    const float angle = (arm + radius * 7.8f) + (u1 - 0.5f) * (0.045f + radius * 0.055f);
    const float thickness = (((u2 + u3) + u4) - 1.5f) * (0.0035f + radius * 0.010f);
    const float3 flat = make_float3(radius * cosf(angle), radius * sinf(angle) * 0.72f, thickness);
    const float tilt = 0.20f;
    pos = make_float3(flat.x, flat.y * cosf(tilt) - flat.z * sinf(tilt), flat.y * sinf(tilt) + flat.z * cosf(tilt));
    const float orbital = 0.99f * sqrtf(d.gravity * galaxy_enclosed_mass(radius) / (radius + 0.018f));
    const float3 tangent = make_float3(-sinf(angle), cosf(angle) * 0.72f * cosf(tilt), cosf(angle) * 0.72f * sinf(tilt));
    const float3 pn = galaxy_normalize(pos);
    const float wave = sinf(angle * 3.0f - radius * 13.0f);
    const float3 radial = make_float3(((pn.x * orbital) * 0.035f) * wave, ((pn.y * orbital) * 0.035f) * wave,
                                      ((pn.z * orbital) * 0.035f) * wave);
    const float3 tn = galaxy_normalize(tangent);
    vel = make_float3(orbital * tn.x + radial.x, orbital * tn.y + radial.y, orbital * tn.z + radial.z);
    excitation = 0.08f + u3 * 0.05f;
  } else {
    // Four retrograde polar lanes, each in its own tilted plane.
    const float lineage = static_cast<float>(i & 3u);
    const float radius = (0.22f + 0.31f * powf(u0, 0.86f)) + (lineage - 1.5f) * 0.007f;
    const float angle = kTau * u1 + radius * 3.5f;
    const float thickness = (((u2 + u3) + u4) - 1.5f) * (0.004f + radius * 0.009f);
    const float3 flat = make_float3(radius * cosf(angle), thickness, radius * sinf(angle));
    const float tilt = -0.54f + lineage * 0.36f;
    pos = make_float3(flat.x * cosf(tilt) - flat.y * sinf(tilt), flat.x * sinf(tilt) + flat.y * cosf(tilt), flat.z);
    const float3 tangent = make_float3(-sinf(angle) * cosf(tilt), -sinf(angle) * sinf(tilt), cosf(angle));
    const float orbital = 0.99f * sqrtf(d.gravity * galaxy_enclosed_mass(radius) / (radius + 0.018f));
    const float3 tn = galaxy_normalize(tangent);
    vel = make_float3(-orbital * tn.x, -orbital * tn.y, -orbital * tn.z);
    excitation = 0.20f + u4 * 0.08f;
  }
  float hue = 0.0f;
  if (u5 >= 0.75f) {
    const float speed = fmaxf(galaxy_length(vel), 0.0001f);
    const float radius = fmaxf(galaxy_length(pos), 0.0001f);
    const float radial_motion =
        clampf(((pos.x * vel.x + pos.y * vel.y) + pos.z * vel.z) / (radius * speed), -1.0f, 1.0f);
    const float3 angular_momentum = galaxy_cross(pos, vel);
    const float spin_signature = angular_momentum.y / fmaxf(galaxy_length(angular_momentum), 0.0001f);
    const float h = ((((atan2f(vel.y, vel.x) / kTau + (vel.z / speed) * 0.14f) + spin_signature * 0.075f) +
                      radial_motion * 0.065f) +
                     excitation * 0.085f) +
                    static_cast<float>(i & 3u) * 0.018f;
    hue = h - floorf(h);
  }
  G(d.pos)[i] = make_float4(pos.x, pos.y, pos.z, hue);
  G(d.vel)[i] = make_float4(vel.x, vel.y, vel.z, excitation);
  G(d.perm)[i] = i;
}

__global__ __launch_bounds__(256) void k_join_particles(SysDesc d, float4* out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= d.n) return;
  const uint32_t j = G(d.perm)[i] < d.n ? G(d.perm)[i] : i;
  out[2u * j] = G(d.pos)[i];
  out[2u * j + 1u] = G(d.vel)[i];
}

}  // namespace

// ------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------

hipError_t launch_direct_f64(const SysDesc& d, uint32_t nt, void* scratch, hipStream_t stream) {
  if (nt == 0u || d.n == 0u) return hipSuccess;
  const DirectLayout L = direct_layout(nt);
  char* base = static_cast<char*>(scratch);
  const uint32_t* targets = reinterpret_cast<const uint32_t*>(base + L.targets);
  uint32_t* slot_of = reinterpret_cast<uint32_t*>(base + L.slots);
  double* partial = reinterpret_cast<double*>(base + L.partial);
  double* out = reinterpret_cast<double*>(base + L.out);
  hipLaunchKernelGGL(k_fill_u32, dim3((nt + 255u) / 256u), dim3(256), 0, stream, slot_of, nt, 0xFFFFFFFFu);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_find_slots, dim3((d.n + 255u) / 256u), dim3(256), 0, stream, d, targets, nt, slot_of);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const uint32_t nchunks = L.nchunks;
  const uint32_t chunk = ((d.n + nchunks - 1u) / nchunks + 255u) / 256u * 256u;
  hipLaunchKernelGGL(k_direct_f64, dim3((nt + 255u) / 256u, nchunks), dim3(256), 0, stream, d, slot_of, nt,
                     chunk, partial);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_direct_final, dim3((nt + 255u) / 256u), dim3(256), 0, stream, d, nt, nchunks, partial, out);
  return hipGetLastError();
}

hipError_t launch_seed_galaxy(const SysDesc& d, hipStream_t stream) {
  if (d.n == 0u) return hipSuccess;
  hipLaunchKernelGGL(k_seed_galaxy, dim3((d.n + 255u) / 256u), dim3(256), 0, stream, d);
  return hipGetLastError();
}

hipError_t launch_split_particles(const SysDesc& d, const float4* in, hipStream_t stream) {
  if (d.n == 0u) return hipSuccess;
  hipLaunchKernelGGL(k_split_particles, dim3((d.n + 255u) / 256u), dim3(256), 0, stream, d, in);
  return hipGetLastError();
}

hipError_t launch_join_particles(const SysDesc& d, float4* out, hipStream_t stream) {
  if (d.n == 0u) return hipSuccess;
  hipLaunchKernelGGL(k_join_particles, dim3((d.n + 255u) / 256u), dim3(256), 0, stream, d, out);
  return hipGetLastError();
}

// Launches through hipExtLaunchKernelGGL when an event pair is given, so
// that the events carry the dispatch's own start and end times.
#define RB_LAUNCH(t, kernel, grid, block, stream, ...)                                                  \
  do {                                                                                                \
    if ((t) != nullptr)                                                                               \
      hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, (t)->start, (t)->stop, 0, __VA_ARGS__);   \
    else                                                                                              \
      hipLaunchKernelGGL(kernel, grid, block, 0, stream, __VA_ARGS__);                                \
  } while (0)

#define RB_LAUNCH_LDS(t, kernel, grid, block, lds, stream, ...)                                         \
  do {                                                                                                \
    if ((t) != nullptr)                                                                               \
      hipExtLaunchKernelGGL(kernel, grid, block, lds, stream, (t)->start, (t)->stop, 0, __VA_ARGS__); \
    else                                                                                              \
      hipLaunchKernelGGL(kernel, grid, block, lds, stream, __VA_ARGS__);                              \
  } while (0)

// A launcher with nothing to launch still records its event pair.
static hipError_t no_launch(hipStream_t stream, const Timing* t) {
  if (t == nullptr) return hipSuccess;
  hipError_t e = hipEventRecord(t->start, stream);
  return e != hipSuccess ? e : hipEventRecord(t->stop, stream);
}

hipError_t launch_deposit(const StepArgs& a, uint32_t nblocks, hipStream_t stream, const Timing* t) {
  if (nblocks == 0u) return no_launch(stream, t);
  RB_LAUNCH(t, k_deposit, dim3(nblocks), dim3(kDepositThreads), stream, a);
  return hipGetLastError();
}

hipError_t launch_upward(const StepArgs& a, uint32_t nblocks, hipStream_t stream, const Timing* t) {
  if (nblocks == 0u) return no_launch(stream, t);
  if (a.nsys == 1u)
    RB_LAUNCH(t, k_upward<true>, dim3(nblocks), dim3(kUpwardThreads), stream, a);
  else
    RB_LAUNCH(t, k_upward<false>, dim3(nblocks), dim3(kUpwardThreads), stream, a);
  return hipGetLastError();
}

hipError_t launch_top(const StepArgs& a, hipStream_t stream, const Timing* t) {
  if (a.nsys == 0u) return no_launch(stream, t);
  if (a.compensated)
    RB_LAUNCH(t, k_top<true>, dim3(a.nsys), dim3(kTopThreads), stream, a);
  else
    RB_LAUNCH(t, k_top<false>, dim3(a.nsys), dim3(kTopThreads), stream, a);
  return hipGetLastError();
}

// Blocks of a tree launch that fit on the device at once (occupancy x CUs),
// cached per kernel form, pair tile and level span (which sets the dynamic
// LDS size).  The tree kernel loops over its items, so a grid of this size
// keeps every block busy and none is launched only to find no work.
template <bool kComp, uint32_t kKind, uint32_t kCpr>
static uint32_t tree_resident_blocks(uint32_t span, uint32_t lds) {
  // Handles may be driven from different host threads: the cache entries
  // are atomics (every thread computes the same value).
  static std::atomic<uint32_t> cache[kLevels] = {};
  const uint32_t key = span % kLevels;
  const uint32_t known = cache[key].load(std::memory_order_relaxed);
  if (known != 0u) return known;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tree<kComp, kKind, kCpr>, kTreeThreads, lds) !=
          hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return kTreeMaxBlocks;
  const uint32_t blocks = static_cast<uint32_t>(cus * per_cu);
  cache[key].store(blocks, std::memory_order_relaxed);
  return blocks;
}

// Resident blocks of the wave-item tree kernel k_tree_w<kComp, kKind>.
template <bool kComp, uint32_t kKind>
static uint32_t tree_w_resident_blocks() {
  static std::atomic<uint32_t> cache{0u};
  const uint32_t known = cache.load(std::memory_order_relaxed);
  if (known != 0u) return known;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tree_w<kComp, kKind>, kTreeThreads, 0) != hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return kTreeMaxBlocks;
  const uint32_t blocks = static_cast<uint32_t>(cus * per_cu);
  cache.store(blocks, std::memory_order_relaxed);
  return blocks;
}

uint32_t tree_single_level_capacity(bool compensated, uint32_t chunks) {
  // Wave items: a block's four waves each take an item.
  if (chunks == 1u)
    return kTreeWaves * (compensated ? tree_w_resident_blocks<true, kTreeSingle>()
                                     : tree_w_resident_blocks<false, kTreeSingle>());
  if (chunks == 4u)
    return compensated ? tree_resident_blocks<true, kTreeSingle, 4u>(0u, 0u)
                       : tree_resident_blocks<false, kTreeSingle, 4u>(0u, 0u);
  return compensated ? tree_resident_blocks<true, kTreeSingle, 8u>(0u, 0u)
                     : tree_resident_blocks<false, kTreeSingle, 8u>(0u, 0u);
}

template <bool kComp, uint32_t kKind, uint32_t kCpr>
static hipError_t launch_tree_form(const StepArgs& a, uint32_t lo, uint32_t hi, uint32_t nblocks, uint32_t lead,
                                   uint32_t lds, hipStream_t stream, const Timing* t) {
  nblocks = std::min(nblocks, tree_resident_blocks<kComp, kKind, kCpr>(hi - lo, lds)) + lead;
  if (lds != 0u)
    RB_LAUNCH_LDS(t, (k_tree<kComp, kKind, kCpr>), dim3(nblocks), dim3(kTreeThreads), lds, stream, a, lo, hi, lead);
  else
    RB_LAUNCH(t, (k_tree<kComp, kKind, kCpr>), dim3(nblocks), dim3(kTreeThreads), stream, a, lo, hi, lead);
  return hipGetLastError();
}

// Wave items: `items` items (an upper bound on the launch's parents) on at
// most the resident waves, after the lead block and direct waves.
template <bool kComp, uint32_t kKind>
static hipError_t launch_tree_wave_form(const StepArgs& a, uint32_t lo, uint32_t hi, uint32_t items, uint32_t lead,
                                        hipStream_t stream, const Timing* t) {
  const uint32_t resident_waves = kTreeWaves * tree_w_resident_blocks<kComp, kKind>();
  const uint32_t ndirect = lead != 0u ? lead - 1u : 0u;
  const uint32_t listed = std::max(1u, std::min(items, resident_waves));
  const uint32_t nblocks = (lead != 0u ? 1u : 0u) + (ndirect + listed + kTreeWaves - 1u) / kTreeWaves;
  RB_LAUNCH(t, (k_tree_w<kComp, kKind>), dim3(nblocks), dim3(kTreeThreads), stream, a, lo, hi, lead);
  return hipGetLastError();
}

template <bool kComp>
static hipError_t launch_tree_wave(const StepArgs& a, uint32_t kind, uint32_t lo, uint32_t hi, uint32_t items,
                                   uint32_t lead, hipStream_t stream, const Timing* t) {
  switch (kind) {
    case kTreePassM2l: return launch_tree_wave_form<kComp, kTreeM2l>(a, lo, hi, items, lead, stream, t);
    case kTreePassFinish: return launch_tree_wave_form<kComp, kTreeFinish>(a, lo, hi, items, 0u, stream, t);
    default: return launch_tree_wave_form<kComp, kTreeSingle>(a, lo, hi, items, 0u, stream, t);
  }
}

template <bool kComp, uint32_t kCpr>
static hipError_t launch_tree_comp(const StepArgs& a, uint32_t kind, uint32_t lo, uint32_t hi, uint32_t nblocks,
                                   uint32_t lead, hipStream_t stream, const Timing* t) {
  switch (kind) {
    case kTreePassM2l: return launch_tree_form<kComp, kTreeM2l, kCpr>(a, lo, hi, nblocks, lead, 0u, stream, t);
    case kTreePassFinish: return launch_tree_form<kComp, kTreeFinish, kCpr>(a, lo, hi, nblocks, 0u, 0u, stream, t);
    default:
      if (lo < hi)
        return launch_tree_form<kComp, kTreeFused, kCpr>(
            a, lo, hi, nblocks, 0u, std::max(item_bytes(kCpr), (hi - lo) * kAncestorBytes), stream, t);
      return launch_tree_form<kComp, kTreeSingle, kCpr>(a, lo, hi, nblocks, 0u, 0u, stream, t);
  }
}

hipError_t launch_tree(const StepArgs& a, uint32_t kind, uint32_t lo, uint32_t hi, uint32_t nblocks, uint32_t lead,
                       uint32_t chunks, hipStream_t stream, const Timing* t) {
  if ((nblocks == 0u && lead == 0u) || lo < 2u || hi < lo || hi >= kLevels) return no_launch(stream, t);
  if (lead != 0u && (kind != kTreePassM2l || a.nsys != 1u)) return hipErrorInvalidValue;
  if (chunks != 1u && chunks != 4u && chunks != 8u) return hipErrorInvalidValue;
  // Wave items serve the unfused forms; a fused run of levels (which
  // recomputes the ancestor path with the whole block) takes the block form.
  if (chunks == 1u && (kind != kTreePassRun || lo == hi))
    return a.compensated ? launch_tree_wave<true>(a, kind, lo, hi, nblocks, lead, stream, t)
                         : launch_tree_wave<false>(a, kind, lo, hi, nblocks, lead, stream, t);
  if (chunks == 1u) chunks = 8u;
  if (chunks == 4u)
    return a.compensated ? launch_tree_comp<true, 4u>(a, kind, lo, hi, nblocks, lead, stream, t)
                         : launch_tree_comp<false, 4u>(a, kind, lo, hi, nblocks, lead, stream, t);
  return a.compensated ? launch_tree_comp<true, 8u>(a, kind, lo, hi, nblocks, lead, stream, t)
                       : launch_tree_comp<false, 8u>(a, kind, lo, hi, nblocks, lead, stream, t);
}

hipError_t launch_sort(const StepArgs& a, uint32_t nintegrate, hipStream_t stream) {
  if (nintegrate == 0u) return hipSuccess;
  hipLaunchKernelGGL(k_scatter<kIntegrateThreads>, dim3(nintegrate), dim3(kIntegrateThreads), 0, stream, a,
                     a.integrate_blocks);
  return hipGetLastError();
}

hipError_t launch_sort_ranges(const StepArgs& a, const BodyBlock* deposit_ranges, uint32_t nranges,
                              hipStream_t stream) {
  if (nranges == 0u) return hipSuccess;
  hipLaunchKernelGGL(k_scatter<kDepositBodies>, dim3(nranges), dim3(kDepositBodies), 0, stream, a, deposit_ranges);
  return hipGetLastError();
}

template <uint32_t kThreads>
static hipError_t launch_integrate_width(const StepArgs& a, uint32_t ndeposit, hipStream_t stream, const Timing* t) {
  const dim3 grid(ndeposit), block(kThreads);
  if (a.compensated) {
    if (a.nsys == 1u)
      RB_LAUNCH(t, (k_integrate_deposit<true, true, kThreads>), grid, block, stream, a);
    else
      RB_LAUNCH(t, (k_integrate_deposit<true, false, kThreads>), grid, block, stream, a);
  } else if (a.nsys == 1u) {
    RB_LAUNCH(t, (k_integrate_deposit<false, true, kThreads>), grid, block, stream, a);
  } else {
    RB_LAUNCH(t, (k_integrate_deposit<false, false, kThreads>), grid, block, stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_integrate_deposit(const StepArgs& a, uint32_t ndeposit, uint32_t threads, hipStream_t stream,
                                    const Timing* t) {
  if (ndeposit == 0u) return no_launch(stream, t);
  if (threads == 512u) return launch_integrate_width<512u>(a, ndeposit, stream, t);
  if (threads == 1024u) return launch_integrate_width<1024u>(a, ndeposit, stream, t);
  return hipErrorInvalidValue;
}

hipError_t launch_accel(const StepArgs& a, uint32_t nblocks, uint32_t sys, float4* out,
                        hipStream_t stream) {
  if (nblocks == 0u) return hipSuccess;
  if (a.compensated)
    hipLaunchKernelGGL(k_accel<true>, dim3(nblocks), dim3(kIntegrateThreads), 0, stream, a, sys, out);
  else
    hipLaunchKernelGGL(k_accel<false>, dim3(nblocks), dim3(kIntegrateThreads), 0, stream, a, sys, out);
  return hipGetLastError();
}

hipError_t launch_diag(const StepArgs& a, uint32_t nblocks, uint32_t sys, double* partial,
                       double* out, hipStream_t stream) {
  if (nblocks == 0u) return hipSuccess;
  if (a.compensated)
    hipLaunchKernelGGL(k_diag_partial<true>, dim3(nblocks), dim3(kIntegrateThreads), 0, stream, a,
                       sys, partial);
  else
    hipLaunchKernelGGL(k_diag_partial<false>, dim3(nblocks), dim3(kIntegrateThreads), 0, stream, a,
                       sys, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, stream, partial, nblocks, out);
  return hipGetLastError();
}

}  // namespace rainbow
