// Device-side data layout shared by the gfx950 kernels and the host runtime.
//
// HBM layout of one system of n bodies and C = total_cells(depth) cells:
//   pos     float4[n]   {x, y, z, hue}            \ the reference Particle,
//   vel     float4[n]   {vx, vy, vz, excitation}  / split into two streams
//   perm    uint32[n]   caller index of the body in each slot: bodies are
//                       kept in Morton order of their finest cell
//   moments int4[C]     {count, sum_x, sum_y, sum_z}   reference Moment
//   fields  float4[3C]  {a, phi}{Jxx,Jxy,Jxz,Jyy}{Jyz,Jzz,0,0}  reference Field
//   source  float4[C]   {com.xyz, G m} per cell, G m = 0 for an empty cell
//   active  uint8[C]    bit 0: cell was an M2L target in the last solve;
//                       finest cells, bit 1: non-empty in the last solve
//                       (only then are that solve's finest buffer cell and
//                       the cell's source record non-zero); the finest
//                       level's flags are stored subtree-major (the 8^3
//                       cells under each level-(D-3) cell together)
//   extra   int4[8^D]   finest-level moments of bodies deposited by a block
//                       that does not own their cell (merged and zeroed by
//                       upward)
//   finest_alt int4[8^D] the finest level's second buffer: steps alternate
//                       between the moments array's finest level (parity 0)
//                       and this one (parity 1); a step's upward zeroes the
//                       buffer the next step deposits into
//   xflag   uint8[8^D]  the deposit added stragglers of the cell into `extra`
//                       (subtree-major, as the finest flags of `active`)
//   m2l     float4[3 level_offset(D)]  M2L sums of the cells above the
//                       finest level, written by the tree pass's first
//                       launch and read by its second
// and the system's counting-sort state (bodies are re-sorted into Morton
// order of their finest cell every few steps):
//   rank    uint32[n]   each body's index within its cell's bodies, from
//                       the deposit (bit 31: after the cell's bodies stored
//                       by its owning block)
//   tpos, tvel float4[n] each, contiguous: the sorted bodies, written by the
//                       scatter and read by that step's integrate; between
//                       steps the 32 n bytes stage particles in the
//                       reference's interleaved layout for uploads and
//                       downloads
//   tperm   uint32[n]   caller index of each sorted body
//   cstart, xstart uint32[8^D]  offset of each finest cell (Morton order)
//                       within its 8^3 subtree, without and with the cell's
//                       owned bodies
//   sstart  uint32[8^(D-3)]  offset of each subtree within the system
//   own     uint2[ceil(n / 1024)]  Morton range of finest cells each
//                       deposit range of 1024 bodies owns (stores without
//                       atomics); (0, 0) owns nothing
// with the reference's cell numbering level_offset(l) + x + 2^l (y + 2^l z).
// A step works on a set of systems (a plain solve is a set of one); each
// system's descriptor carries its own buffer pointers and nothing of a
// system is indexed by its place in the set, so systems join and leave the
// set between steps without moving memory or losing state.
#pragma once

#include <cstdint>

namespace rainbow {

// One system's buffers, plan and physics (the per-system part of the
// reference's Params uniform, src/rainbow_fmm.wgsl:18-50).
struct alignas(16) SysDesc {
  float4* pos;
  float4* vel;
  uint32_t* perm;  // caller index of the body stored in each slot
  int4* moments;
  float4* fields;
  float4* source;
  uint8_t* active;
  int4* extra;
  int4* finest_alt;
  float4* m2l;         // M2L sums of the levels below the finest, fields layout
  uint8_t* xflag;      // per finest cell: stragglers were added into `extra`
  uint32_t* rank;      // counting sort: see the layout above
  float4* tpos;
  float4* tvel;
  uint32_t* tperm;
  uint32_t* cstart;
  uint32_t* xstart;
  uint32_t* sstart;
  uint2* own;
  uint32_t n;          // bodies (particle_count)
  uint32_t depth;      // octree depth
  float scale;        // fixed-point scale (control_settings.z)
  float dt;
  float gravity;
  float softening;
  float speed;
  float control_strength;  // control_settings.x
  float control_radius;    // control_settings.y
  uint32_t live;           // 0: every kernel skips the system
  uint32_t sort_now;       // this step sorts the system even when the step
                           // as a whole does not (its first step after it
                           // joined a running set)
  float control_point[4];
  float control_axis_mode[4];
};
static_assert(sizeof(SysDesc) == 240, "SysDesc layout");

// Work list of the particle passes: a block covers bodies [first, first +
// block size) of one system.
struct alignas(8) BodyBlock {
  uint32_t sys;
  uint32_t first;
};

// Work list of the upward pass: a block owns the 8x8x8 finest cells under
// one cell of level depth-3 of one system.
struct alignas(8) SubtreeBlock {
  uint32_t sys;
  uint32_t root;  // linear index of the subtree root within level depth-3
};

constexpr uint32_t kLevels = 7;  // levels 0..6

// Internal cells are listed per level in kStripes stripes; a block appends
// to stripe (block index mod kStripes) with one atomic per level.
constexpr uint32_t kStripes = 32;

// Everything a kernel of the step needs, passed by value.
struct StepArgs {
  const SysDesc* sys;
  const BodyBlock* deposit_blocks;
  const BodyBlock* integrate_blocks;
  const SubtreeBlock* subtree_blocks;
  // Counting sort of the bodies by the Morton code of their finest cell,
  // run within a step when `sorting` is set (each system's rank, offsets,
  // staging and owned ranges are in its descriptor): the deposit writes
  // each body's index within its cell to rank (bit 31 set: after the cell's
  // owned bodies), upward writes each cell's offset within its subtree to
  // cstart/xstart and top each subtree's offset to sstart (both in Morton
  // order), the scatter moves the bodies to tpos/tvel/tperm, and the step's
  // integrate reads them from there and sets the owned ranges.  A deposit
  // block owns the Morton range [lo, hi) of finest cells (stores without
  // atomics); bodies of other cells go to `extra`.
  uint32_t* lists;             // internal cells (count > 16)
  uint32_t* list_count;        // [kLevels][kStripes] items per stripe
  const uint32_t* stripe_off;  // [kLevels][kStripes] first slot of a stripe
  const uint32_t* stripe_cap;  // [kLevels][kStripes] slots of a stripe
  uint32_t nsys;
  uint32_t compensated;
  uint32_t sorting;
  uint32_t parity;     // finest-level buffer of this step (see finest_alt)
  uint32_t rank_next;  // the fused deposit writes ranks: the next step sorts
};

// Deposit tile: the bodies of one deposit/integrate block (its owned Morton
// range of finest cells is kept per tile).  The standalone deposit runs
// kDepositThreads threads, one body each; the fused integrate+deposit runs
// 1024 or 512 threads over the same tile (Shape::integrate_threads).
constexpr uint32_t kDepositBodies = 1024;
constexpr uint32_t kDepositThreads = 1024;
constexpr uint32_t kDepositSlots = 2048;  // LDS hash slots, 2x bodies per block
constexpr uint32_t kDepositSlotBits = 31u - static_cast<uint32_t>(__builtin_clz(kDepositSlots));
static_assert((1u << kDepositSlotBits) == kDepositSlots, "hash table size is a power of two");
constexpr uint32_t kIntegrateThreads = 256;
constexpr uint32_t kUpwardThreads = 512;  // one thread per finest cell of a subtree
constexpr uint32_t kTopThreads = 1024;  // one block per system for the levels above the subtrees
constexpr uint32_t kTreeThreads = 256;  // four waves per internal parent
constexpr uint32_t kTreeMaxBlocks = 16384;

// List item: system index in the high 16 bits, cell index within its level
// in the low 16 bits (levels <= 5 hold at most 32768 cells).
inline constexpr uint32_t pack_item(uint32_t sys, uint32_t cell) { return (sys << 16) | cell; }

}  // namespace rainbow
