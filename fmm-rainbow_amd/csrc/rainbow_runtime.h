// Host runtime objects shared by the solver, batch and server C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "rainbow_device.h"
#include "rainbow_plan.h"

struct rainbow_params;

namespace rainbow {

// The physics part of the reference's Params (src/rainbow_fmm.wgsl:28-49).
struct Physics {
  float dt = 0.0009f;
  float gravity = 0.34f;
  float softening = 0.012f;
  float speed = 1.0f;
  float control_strength = 0.7f;
  float control_radius = 0.22f;
  float scale = 0.0f;  // 0: planner scale
  float control_point[4] = {0, 0, 0, 0};
  float control_axis_mode[4] = {0, 0, 0, 0};
};

Physics physics_from(const rainbow_params* p);
void apply_physics(SysDesc& s, const Physics& ph);

int fail(int code, const char* fmt, ...);
void clear_error();
const char* last_error();

// Steps captured in the multi-step graph.
constexpr uint32_t kGraphSteps = 16;
// Bodies are re-sorted into Morton order in every this many-th step by
// default (tunable per engine; 0 = never).
constexpr uint32_t kDefaultSortInterval = 8;
// Tree launches: bit l of the split mask (2 <= l <= 6) starts a launch at
// level l; a launch runs its levels up to the next start.  Default: levels
// 2..5 together and 6 alone (tunable per engine).
constexpr uint32_t kDefaultTreeSplit = (1u << 2) | (1u << 6);

// One launch of the tree pass: form (kTreePass*) and levels lo..hi; lead
// > 0: the launch also runs the top levels (see launch_tree).
// chunks: the launch's pair tile (Shape::tree_chunks, resolved per launch
// when that is 0).
struct TreeLaunch {
  uint32_t kind, lo, hi, blocks, lead, chunks;
};

// One device allocation holding a system's bodies and cells, and its
// descriptor pointing into it.
struct SystemStorage {
  void* base = nullptr;
  size_t bytes = 0;
  SysDesc desc{};
};

// Bytes of the allocation of a system of n bodies.
size_t system_storage_bytes(uint32_t n);
// Offset in that storage from which it must start zeroed (the buffers
// before it are written before they are read).
size_t system_zeroed_from(uint32_t n);
// Allocates and zeroes a system's storage; scale 0 = planner scale.
int alloc_system(uint32_t n, const Physics& ph, float scale, hipStream_t stream, SystemStorage* out);
void free_system(SystemStorage* s);
// The descriptor of a system of n bodies whose storage starts at `base`
// (system_storage_bytes(n) bytes); scale 0 = planner scale.
SysDesc describe_system(void* base, uint32_t n, const Physics& ph, float scale);
// Copies reference-layout particles in / a buffer out (synchronous on stream).
// Particles cross the bus in the reference's interleaved layout through the
// system's own staging area (tpos/tvel, 32 n bytes); kernels split them into
// the pos/vel streams and restore caller order on the way out.
int upload_system(const SysDesc& d, const void* data, size_t bytes, hipStream_t stream);
// finest_parity: the finest-level buffer holding the last solve's moments.
int download_system(const SysDesc& d, int which, void* dst, size_t bytes, hipStream_t stream,
                    uint32_t finest_parity = 0u);

// A device buffer that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes);
  void release();
};

// Work tables, lists, stream and graphs of a step over a set of systems.
class Engine {
 public:
  Engine() = default;
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;
  ~Engine();

  int open(uint32_t options);
  // Installs the system set (descriptors with device pointers).  Rebuilds the
  // work tables and drops captured graphs.  Waits for the stream first; the
  // next step deposits every system again.
  int configure(const std::vector<SysDesc>& systems);
  // The same between steps without waiting: systems keep their state, and
  // `fresh` (indices into `systems`, storage zeroed, bodies in place) are
  // deposited by the next step, as are the systems an earlier install
  // listed that no step has deposited yet.
  int configure_async(const std::vector<SysDesc>& systems, const std::vector<uint32_t>& fresh);
  // Table capacity for later installs of up to these totals.
  int reserve(uint64_t bodies, uint32_t systems, uint64_t finest);
  // Systems join only through configure_async / append_systems from now on
  // (the server).
  void start_incremental();
  // Appends fresh systems to the set, keeping everyone's place (only the
  // tables' new tails are copied).
  int append_systems(const std::vector<SysDesc>& add);
  // Takes a system out of the set by marking it dead (kernels skip it);
  // its table entries stay until the next full install.
  int retire_system(const float4* pos);
  uint64_t dead_blocks() const;
  int push_descriptors();
  int set_physics(const Physics& ph);
  int enqueue(bool integrate, bool sorting, uint32_t step_parity, bool deposit_first, bool rank_next,
              bool record_counts = false);
  hipError_t zero_finest();
  void plan_next(bool* sorting, bool* rank_next, uint32_t* since_after) const;
  int solve_fields();
  int step(uint32_t steps);
  int step_eager(uint32_t steps);
  int sync();
  int time_steps(uint32_t steps, float* ms);
  int phase_ms(uint32_t steps, float* out, uint32_t n);
  int upload(uint32_t sys, const void* data, size_t bytes);
  int seed_galaxy(uint32_t sys);
  int download(uint32_t sys, int which, void* dst, size_t bytes);
  int accelerations(uint32_t sys, void* dst, size_t bytes);
  int diagnostics(uint32_t sys, double out[8]);
  int direct_sum_f64(uint32_t sys, const uint32_t* targets, uint32_t nt, double* out);
  int check_direct_args(uint32_t sys, const uint32_t* targets, uint32_t nt, const double* out) const;
  int direct_sum_f64_in(uint32_t sys, const uint32_t* targets, uint32_t nt, double* out, void* scratch);
  int set_sort_interval(uint32_t interval);
  int set_tree_fusion(uint32_t hi);
  int set_tree_split(uint32_t starts);  // fixes the split (no measured choice)
  // Block shapes of the integrate and tree launches (rainbow_launch.h).
  int set_shape(const Shape& s);
  // Times the block shapes, tree splits and sort intervals; keeps the fastest.
  int autotune(uint32_t steps, uint32_t* chosen);
  int snapshot(bool restore);

  hipStream_t stream = nullptr;
  int device = 0;  // current device at open(); the stream belongs to it
  uint32_t options = 0;
  uint32_t dmax = 0;
  uint64_t total_bodies = 0;
  std::vector<SysDesc> sys;
  StepArgs args{};
  uint32_t n_deposit = 0;
  uint32_t n_integrate = 0;
  uint32_t n_subtree = 0;
  uint32_t n_fresh = 0;  // deposit blocks of systems that joined and await their first deposit
  uint32_t tree_blocks[kLevels + 1] = {};
  // Items (internal parents) per level in the last measured step, 0 before
  // any measurement; they choose each launch's tile under tree_chunks 0.
  uint32_t level_items[kLevels + 1] = {};
  std::vector<TreeLaunch> tree_launches;
  uint32_t tree_split = kDefaultTreeSplit;
  Shape shape;
  // Until a split is set explicitly, the first eager step after a
  // configuration records its per-level internal-cell counts and the split
  // is chosen from them (adapt_split).
  bool split_auto = true;
  // Finest-level buffer the next solve deposits into; the last solve used
  // parity ^ 1.
  uint32_t parity = 0;

 private:
  int capture(uint32_t steps, uint32_t sort_every, bool last_rank, uint32_t start_parity, hipGraphExec_t* out);
  int prepare_graphs(uint32_t steps, uint32_t start_parity);
  uint32_t chunk_steps() const;
  void plan_tree();
  int apply_tree_split(uint32_t starts);
  int adapt_split(bool wait);
  void drop_graphs();

  int install(const std::vector<SysDesc>& systems, const std::vector<uint32_t>& fresh);
  void add_tables(uint32_t s);
  void collect_fresh(const std::vector<const float4*>& listed);
  int upload_tables(uint32_t s0, uint32_t d0, uint32_t i0, uint32_t b0);
  uint64_t dead_blocks_ = 0;
  bool descs_dirty_ = false;  // host descriptors changed since the last copy
  int flush_descriptors();

  DevBuf sys_buf, deposit_buf, fresh_buf, integrate_buf, subtree_buf, list_buf, count_buf, stripe_buf, diag_buf,
      snap_buf;
  // Host copies of the work tables (kept to reuse their storage).
  std::vector<BodyBlock> dep_, integ_, fresh_;
  // Bodies (pos pointers) of the systems the next step deposits first.
  std::vector<const float4*> fresh_pending_;
  std::vector<SubtreeBlock> sub_;
  // Pinned staging of the table copies, a ring: a buffer is refilled once
  // the copies that read it have run (its event).
  struct TableStage {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
  };
  static constexpr uint32_t kTableStages = 32;
  TableStage stages_[kTableStages];
  // Staging buffers outgrown while their copies may still be queued; freed
  // with the engine (a pinned free waits for the device).
  std::vector<void*> retired_stages_;
  uint32_t stage_next_ = 0;
  int next_stage(size_t bytes, TableStage** out);

 public:
  DevBuf io_buf;  // staging of the direct sum

 private:
  uint32_t sort_interval = kDefaultSortInterval;
  uint32_t since_sort = kDefaultSortInterval;
  // Chunk graphs per starting parity of the finest-level buffers.
  hipGraphExec_t graph_chunk[2] = {nullptr, nullptr};
  // The current finest buffer has not been deposited into for the current
  // bodies (after an upload, a field solve, a schedule change or a restore):
  // the next step runs a standalone deposit.
  bool pending_deposit = true;
  hipEvent_t ev_a = nullptr;
  hipEvent_t ev_b = nullptr;
  bool split_wanted = false;   // a measured split choice is due
  bool counts_pending = false;  // counts_host is being filled (ev_counts)
  // Changes of the set (or of the block shape) since the engine opened, and
  // the one the pending counts were recorded at: a measurement in flight
  // still applies when the set changes under it (under steady churn one
  // would never finish otherwise), and the split is measured again after.
  uint64_t set_gen_ = 0;
  uint64_t counts_gen_ = 0;
  uint32_t* counts_host = nullptr;  // pinned copy of the list counts
  hipEvent_t ev_counts = nullptr;
  std::vector<uint32_t> caps_host;  // list capacity per level and stripe
};

// A fixed set of systems (the solver is a batch of one).
class Batch {
 public:
  ~Batch();
  int init(const std::vector<uint32_t>& counts, const Physics& ph, uint32_t options,
           const std::vector<float>* scales);
  Engine engine;
  std::vector<SystemStorage> storage;
};

}  // namespace rainbow

// The C ABI handles of include/rainbow_fmm.h: a solver is a batch of one.
struct rainbow_hip {
  rainbow::Batch batch;
};

struct rainbow_batch {
  rainbow::Batch batch;
};
