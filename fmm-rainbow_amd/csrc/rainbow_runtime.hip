// Host runtime: per-system device storage, work tables, the step schedule on
// one HIP stream, HIP-graph capture, and the solver/batch C ABI.
//
// The schedule per step is clear, deposit, upward, tree(1..depth_max),
// integrate -- the physics of the reference's encode_frame
// (native/main.rs:411-421) in depth+4 launches instead of 2 depth+5.  A plain
// solver is a batch of one system.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "rainbow_device.h"
#include "rainbow_fmm.h"
#include "rainbow_launch.h"
#include "rainbow_plan.h"
#include "rainbow_runtime.h"

namespace rainbow {

namespace {
thread_local std::string g_error;

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1u) & ~(kAlign - 1u); }
}  // namespace

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

void clear_error() { g_error.clear(); }

const char* last_error() { return g_error.c_str(); }

#define RB_HIP_RC(expr) \
  do {                    \
    const int rc_ = (expr); \
    if (rc_ != RAINBOW_OK) return rc_; \
  } while (0)

#define RB_HIP(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(RAINBOW_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),       \
                  __FILE__, __LINE__);                                                        \
  } while (0)

Physics physics_from(const rainbow_params* p) {
  if (p == nullptr) {
    rainbow_params_t def;
    rainbow_write_params(0, 0, 0, 0, 1, 0, 1, 0, 0.0f, &def);
    return physics_from(&def);
  }
  Physics ph;
  ph.dt = p->dt;
  ph.gravity = p->gravity;
  ph.softening = p->softening;
  ph.speed = p->speed;
  ph.control_strength = p->control_settings[0];
  ph.control_radius = p->control_settings[1];
  ph.scale = p->control_settings[2];
  for (int k = 0; k < 4; ++k) {
    ph.control_point[k] = p->control_point[k];
    ph.control_axis_mode[k] = p->control_axis_mode[k];
  }
  return ph;
}

void apply_physics(SysDesc& s, const Physics& ph) {
  s.dt = ph.dt;
  s.gravity = ph.gravity;
  s.softening = ph.softening;
  s.speed = ph.speed;
  s.control_strength = ph.control_strength;
  s.control_radius = ph.control_radius;
  for (int k = 0; k < 4; ++k) {
    s.control_point[k] = ph.control_point[k];
    s.control_axis_mode[k] = ph.control_axis_mode[k];
  }
}

// ---------------------------------------------------------------------------
// Per-system storage: pos, vel, moments, fields, source, active in one block.
// ---------------------------------------------------------------------------

namespace {
// Byte offsets of a system's buffers in its allocation (each 256-aligned).
// The buffers every step writes before reading come first; those that must
// start zeroed (the tree, the flags, the owned ranges) follow from `zero`.
struct StorageLayout {
  size_t pos, vel, perm, moments, fields, source, active, extra, finest_alt, m2l, xflag;
  size_t rank, staging, tperm, cstart, xstart, sstart, own, zero, bytes;
};

StorageLayout storage_layout(uint32_t n) {
  const uint32_t depth = depth_for_particles(n);
  const size_t cells = total_cells(depth);
  const size_t finest = level_cells(depth);
  const size_t above = level_offset(depth);
  const size_t nb = static_cast<size_t>(n);
  const size_t blocks = (nb + kDepositBodies - 1u) / kDepositBodies;
  StorageLayout L{};
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t here = at;
    at += align_up(bytes);
    return here;
  };
  // Bodies (written by the upload's split), the sort's ranks, staging and
  // offsets (written by each sorting step's deposit, upward and top before
  // its scatter reads them).
  L.pos = take(16u * nb);
  L.vel = take(16u * nb);
  L.perm = take(4u * nb);
  L.rank = take(4u * nb);
  L.staging = take(32u * nb);  // tpos then tvel, contiguous
  L.tperm = take(4u * nb);
  L.cstart = take(4u * finest);
  L.xstart = take(4u * finest);
  L.sstart = take(4u * (finest / 512u));
  // Zeroed at admission: moments (the deposit adds into a zeroed finest
  // buffer), fields and source records (zero outside the target set),
  // flags, the straggler sums, the owned ranges ((0, 0): none).
  L.zero = at;
  L.moments = take(16u * cells);
  L.fields = take(48u * cells);
  L.source = take(16u * cells);
  L.active = take(cells);
  L.extra = take(16u * finest);
  L.finest_alt = take(16u * finest);
  L.m2l = take(48u * above);
  L.xflag = take(finest);
  L.own = take(8u * blocks);
  L.bytes = at;
  return L;
}
}  // namespace

size_t system_storage_bytes(uint32_t n) { return storage_layout(n).bytes; }
size_t system_zeroed_from(uint32_t n) { return storage_layout(n).zero; }

SysDesc describe_system(void* base, uint32_t n, const Physics& ph, float scale) {
  const Plan plan = make_plan(n);
  const StorageLayout L = storage_layout(n);
  char* p = static_cast<char*>(base);
  SysDesc d{};
  d.pos = reinterpret_cast<float4*>(p + L.pos);
  d.vel = reinterpret_cast<float4*>(p + L.vel);
  d.perm = reinterpret_cast<uint32_t*>(p + L.perm);
  d.moments = reinterpret_cast<int4*>(p + L.moments);
  d.fields = reinterpret_cast<float4*>(p + L.fields);
  d.source = reinterpret_cast<float4*>(p + L.source);
  d.active = reinterpret_cast<uint8_t*>(p + L.active);
  d.extra = reinterpret_cast<int4*>(p + L.extra);
  d.finest_alt = reinterpret_cast<int4*>(p + L.finest_alt);
  d.m2l = reinterpret_cast<float4*>(p + L.m2l);
  d.xflag = reinterpret_cast<uint8_t*>(p + L.xflag);
  d.rank = reinterpret_cast<uint32_t*>(p + L.rank);
  d.tpos = reinterpret_cast<float4*>(p + L.staging);
  d.tvel = d.tpos + n;
  d.tperm = reinterpret_cast<uint32_t*>(p + L.tperm);
  d.cstart = reinterpret_cast<uint32_t*>(p + L.cstart);
  d.xstart = reinterpret_cast<uint32_t*>(p + L.xstart);
  d.sstart = reinterpret_cast<uint32_t*>(p + L.sstart);
  d.own = reinterpret_cast<uint2*>(p + L.own);
  d.n = n;
  d.depth = plan.depth;
  d.scale = scale != 0.0f ? scale : plan.scale;
  d.live = 1u;
  apply_physics(d, ph);
  return d;
}

int alloc_system(uint32_t n, const Physics& ph, float scale, hipStream_t stream, SystemStorage* out) {
  if (n == 0u) return fail(RAINBOW_E_INVALID, "a system needs at least one body");
  const size_t bytes = system_storage_bytes(n);
  void* base = nullptr;
  hipError_t e = hipMalloc(&base, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(RAINBOW_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  e = hipMemsetAsync(base, 0, bytes, stream);
  if (e != hipSuccess) {
    (void)hipFree(base);
    return fail(RAINBOW_E_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
  }
  out->base = base;
  out->bytes = bytes;
  out->desc = describe_system(base, n, ph, scale);
  return RAINBOW_OK;
}

void free_system(SystemStorage* s) {
  if (s->base) (void)hipFree(s->base);
  s->base = nullptr;
  s->bytes = 0;
}

int DevBuf::ensure(size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (bytes <= cap) return RAINBOW_OK;
  release();
  const size_t want = std::max<size_t>(bytes + bytes / 4u, 4096u);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return fail(RAINBOW_E_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
  }
  cap = want;
  return RAINBOW_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------

Engine::~Engine() {
  if (stream) (void)hipStreamSynchronize(stream);
  drop_graphs();
  for (DevBuf* b : {&sys_buf, &deposit_buf, &fresh_buf, &integrate_buf, &subtree_buf, &list_buf, &count_buf,
                    &stripe_buf, &diag_buf, &snap_buf, &io_buf})
    b->release();
  for (TableStage& st : stages_) {
    if (st.done) (void)hipEventDestroy(st.done);
    if (st.host) (void)hipHostFree(st.host);
  }
  for (void* p : retired_stages_) (void)hipHostFree(p);
  for (hipEvent_t e : {ev_a, ev_b, ev_counts})
    if (e) (void)hipEventDestroy(e);
  if (counts_host) (void)hipHostFree(counts_host);
  if (stream) (void)hipStreamDestroy(stream);
}

void Engine::drop_graphs() {
  for (int k = 0; k < 2; ++k) {
    if (graph_chunk[k]) (void)hipGraphExecDestroy(graph_chunk[k]);
    graph_chunk[k] = nullptr;
  }
}

int Engine::open(uint32_t opts) {
  options = opts;
  RB_HIP(hipGetDevice(&device));
  RB_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  RB_HIP(hipEventCreate(&ev_a));
  RB_HIP(hipEventCreate(&ev_b));
  RB_HIP(hipEventCreateWithFlags(&ev_counts, hipEventDisableTiming));
  RB_HIP(hipHostMalloc(reinterpret_cast<void**>(&counts_host), sizeof(uint32_t) * kLevels * kStripes,
                       hipHostMallocDefault));
  return RAINBOW_OK;
}

// Installs the system set and waits for it: the next step deposits every
// system again into zeroed finest buffers (bodies may have changed).
int Engine::configure(const std::vector<SysDesc>& systems) {
  RB_HIP(hipStreamSynchronize(stream));
  drop_graphs();
  int rc = install(systems, std::vector<uint32_t>());
  if (rc != RAINBOW_OK) return rc;
  RB_HIP(hipMemsetAsync(count_buf.p, 0, sizeof(uint32_t) * kLevels * kStripes, stream));
  RB_HIP(hipStreamSynchronize(stream));
  since_sort = sort_interval;
  pending_deposit = true;
  return RAINBOW_OK;
}

// Installs a new system set between steps without waiting for the device:
// the steps already enqueued keep the old tables (the table copies run after
// them in stream order), and every system's own state (finest buffers,
// ranks, owned ranges, fields) carries over, since nothing of a system is
// indexed by its place in the set.  `fresh` lists the systems (indices into
// `systems`) whose storage is zeroed and whose bodies are not deposited yet:
// the next step deposits them, alone, before its upward.  Captured graphs
// are dropped (the only wait, and only when a graph exists).
int Engine::configure_async(const std::vector<SysDesc>& systems, const std::vector<uint32_t>& fresh) {
  if (graph_chunk[0] != nullptr || graph_chunk[1] != nullptr) {
    RB_HIP(hipStreamSynchronize(stream));
    drop_graphs();
  }
  return install(systems, fresh);
}

// Device table capacity for a set of up to `bodies` bodies, `systems`
// systems and `finest` finest cells in all, so that later installs within
// those bounds never reallocate (which would wait for the device).
int Engine::reserve(uint64_t bodies, uint32_t systems, uint64_t finest) {
  const uint64_t dep = bodies / kDepositBodies + systems;
  const uint64_t integ = bodies / kIntegrateThreads + systems;
  const uint64_t sub = finest / 512u + systems;
  // Lists: a system lists at most level_cells(k) internal cells at level
  // k < depth, (8^depth - 1) / 7 in all, fewer than finest / 7 + 1.
  const uint64_t lists = finest / 7u + systems + kLevels * kStripes;
  int rc;
  if ((rc = sys_buf.ensure(sizeof(SysDesc) * systems)) || (rc = deposit_buf.ensure(sizeof(BodyBlock) * dep)) ||
      (rc = fresh_buf.ensure(sizeof(BodyBlock) * dep)) || (rc = integrate_buf.ensure(sizeof(BodyBlock) * integ)) ||
      (rc = subtree_buf.ensure(sizeof(SubtreeBlock) * sub)) || (rc = list_buf.ensure(sizeof(uint32_t) * lists)))
    return rc;
  return RAINBOW_OK;
}

// Appends system s's work: deposit and integrate ranges, subtree blocks,
// and its share of the list capacity of every (level, stripe): a subtree
// block appends at most 64 / 8 / 1 internal cells of levels depth-1 / -2 /
// -3 to stripe (block index mod kStripes), the top block of system s
// appends the levels above to stripe (s mod kStripes); a level never holds
// more than n / 17 internal cells.
void Engine::add_tables(uint32_t s) {
  const SysDesc& d = sys[s];
  const uint32_t ib = d.n / (kAdaptiveLeafCapacity + 1u);
  for (uint32_t f = 0; f < d.n; f += kDepositBodies) dep_.push_back(BodyBlock{s, f});
  for (uint32_t f = 0; f < d.n; f += kIntegrateThreads) integ_.push_back(BodyBlock{s, f});
  // Subtrees in Morton order of their roots, as the bodies are: the upward
  // blocks an XCD runs (a contiguous eighth of the table, see xcd_tile)
  // then read the finest cells that XCD's integrate blocks deposited.
  const uint32_t rd = d.depth - 3u;
  for (uint32_t m = 0; m < level_cells(rd); ++m) {
    uint32_t r = 0u;
    for (uint32_t b = 0; b < rd; ++b)
      r |= (((m >> (3u * b)) & 1u) << b) | (((m >> (3u * b + 1u)) & 1u) << (rd + b)) |
           (((m >> (3u * b + 2u)) & 1u) << (2u * rd + b));
    const uint32_t stripe = static_cast<uint32_t>(sub_.size()) % kStripes;
    caps_host[(d.depth - 1u) * kStripes + stripe] += std::min(64u, ib);
    caps_host[(d.depth - 2u) * kStripes + stripe] += std::min(8u, ib);
    caps_host[(d.depth - 3u) * kStripes + stripe] += std::min(1u, ib);
    sub_.push_back(SubtreeBlock{s, r});
  }
  for (uint32_t k = 0; k + 3u < d.depth; ++k) caps_host[k * kStripes + s % kStripes] += std::min(level_cells(k), ib);
  dmax = std::max(dmax, d.depth);
  total_bodies += d.n;
}

// The deposit ranges of the systems awaiting their first deposit: those
// listed in `listed` (bodies pointers) and those an earlier install listed
// that no step has deposited yet (installs can follow each other between
// steps).
void Engine::collect_fresh(const std::vector<const float4*>& listed) {
  std::vector<const float4*> waiting = listed;
  for (const float4* p : fresh_pending_)
    if (std::find(waiting.begin(), waiting.end(), p) == waiting.end()) waiting.push_back(p);
  fresh_.clear();
  std::vector<const float4*> kept;
  for (uint32_t s = 0; s < sys.size(); ++s) {
    if (!sys[s].live || std::find(waiting.begin(), waiting.end(), sys[s].pos) == waiting.end()) continue;
    kept.push_back(sys[s].pos);
    for (uint32_t f = 0; f < sys[s].n; f += kDepositBodies) fresh_.push_back(BodyBlock{s, f});
  }
  fresh_pending_ = kept;
  n_fresh = static_cast<uint32_t>(fresh_.size());
}

// Builds the work tables of `systems` on the host and copies them (and the
// descriptors) to the device in stream order, through a ring of pinned
// staging buffers; nothing here waits for the device unless a buffer must
// grow or every staging buffer is still being read.
int Engine::install(const std::vector<SysDesc>& systems, const std::vector<uint32_t>& fresh) {
  if (systems.size() > 65535u) return fail(RAINBOW_E_INVALID, "at most 65535 systems");
  std::vector<const float4*> listed;
  for (uint32_t s : fresh) {
    if (s >= systems.size()) return fail(RAINBOW_E_INVALID, "fresh system %u out of range", s);
    listed.push_back(systems[s].pos);
  }
  sys = systems;
  dep_.clear();
  integ_.clear();
  sub_.clear();
  caps_host.assign(kLevels * kStripes, 0u);
  dmax = 0;
  total_bodies = 0;
  dead_blocks_ = 0;
  for (uint32_t s = 0; s < sys.size(); ++s) add_tables(s);
  collect_fresh(listed);
  return upload_tables(0u, 0u, 0u, 0u);
}

// Appends `add` to the set (each fresh: storage zeroed, bodies in place);
// the systems already in the set keep their places, so only the tables'
// new tails, the stripe layout and the fresh list are copied.  Falls back
// to a full install (without the dead systems) when the tables would
// outgrow their buffers.
int Engine::append_systems(const std::vector<SysDesc>& add) {
  if (add.empty()) return RAINBOW_OK;
  std::vector<const float4*> listed;
  for (const SysDesc& d : add) listed.push_back(d.pos);
  uint64_t bodies = 0;
  for (const SysDesc& d : add) bodies += d.n;
  const uint64_t dep = dep_.size() + bodies / kDepositBodies + add.size();
  const uint64_t integ = integ_.size() + bodies / kIntegrateThreads + add.size();
  uint64_t sub = sub_.size();
  for (const SysDesc& d : add) sub += level_cells(d.depth - 3u);
  const bool fits = sys.size() + add.size() <= 65535u && sizeof(SysDesc) * (sys.size() + add.size()) <= sys_buf.cap &&
                    sizeof(BodyBlock) * dep <= deposit_buf.cap && sizeof(BodyBlock) * integ <= integrate_buf.cap &&
                    sizeof(SubtreeBlock) * sub <= subtree_buf.cap && sizeof(BodyBlock) * dep <= fresh_buf.cap &&
                    total_bodies + bodies <= 0xFFFFFFFFull;
  if (!fits) {
    std::vector<SysDesc> live;
    std::vector<uint32_t> fr;
    for (const SysDesc& d : sys)
      if (d.live) live.push_back(d);
    for (const SysDesc& d : add) {
      fr.push_back(static_cast<uint32_t>(live.size()));
      live.push_back(d);
    }
    return configure_async(live, fr);
  }
  const uint32_t s0 = static_cast<uint32_t>(sys.size());
  const uint32_t d0 = static_cast<uint32_t>(dep_.size()), i0 = static_cast<uint32_t>(integ_.size()),
                 b0 = static_cast<uint32_t>(sub_.size());
  for (const SysDesc& d : add) {
    sys.push_back(d);
    add_tables(static_cast<uint32_t>(sys.size() - 1u));
  }
  collect_fresh(listed);
  return upload_tables(s0, d0, i0, b0);
}

// Takes the system with bodies `pos` out of the set: its descriptor's live
// flag is cleared on the device (in stream order), every kernel of a later
// step skips it, and its tables stay as tombstones until the next full
// install.  Once the steps enqueued before this call have run, its storage
// is no longer read.
int Engine::retire_system(const float4* pos) {
  uint32_t s = 0;
  while (s < sys.size() && !(sys[s].pos == pos && sys[s].live)) ++s;
  if (s == sys.size()) return fail(RAINBOW_E_INVALID, "system not in the set");
  sys[s].live = 0u;
  dead_blocks_ += (sys[s].n + kIntegrateThreads - 1u) / kIntegrateThreads;
  descs_dirty_ = true;  // copied before the next step's kernels
  return RAINBOW_OK;
}

// The descriptor table, after systems were marked dead: one copy before the
// next step (stream order keeps it after the steps already enqueued).
int Engine::flush_descriptors() {
  if (!descs_dirty_ || sys.empty()) return RAINBOW_OK;
  TableStage* st = nullptr;
  const size_t bytes = sizeof(SysDesc) * sys.size();
  int rc = next_stage(bytes, &st);
  if (rc) return rc;
  std::memcpy(st->host, sys.data(), bytes);
  RB_HIP(hipMemcpyAsync(sys_buf.p, st->host, bytes, hipMemcpyHostToDevice, stream));
  RB_HIP(hipEventRecord(st->done, stream));
  descs_dirty_ = false;
  return RAINBOW_OK;
}

// Integrate ranges of systems taken out of the set, still in the tables.
uint64_t Engine::dead_blocks() const { return dead_blocks_; }

// The next pinned staging buffer of the ring, with room for `bytes`, once
// the copies that last read it have run.  The ring is deep (a step that
// changes the set uses up to three), so the wait is rare; an outgrown
// buffer is kept until the engine goes (freeing pinned memory would wait
// for the device) and a larger one takes its place.
int Engine::next_stage(size_t bytes, TableStage** out) {
  TableStage& st = stages_[stage_next_];
  stage_next_ = (stage_next_ + 1u) % kTableStages;
  if (st.done != nullptr) RB_HIP(hipEventSynchronize(st.done));
  if (st.cap < bytes) {
    if (st.host != nullptr) retired_stages_.push_back(st.host);
    st.host = nullptr;
    st.cap = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 4u, 1u << 16);
    RB_HIP(hipHostMalloc(&st.host, want, hipHostMallocDefault));
    st.cap = want;
  }
  if (st.done == nullptr) RB_HIP(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
  *out = &st;
  return RAINBOW_OK;
}

// Copies the tables from the given first entries on (0: all), the stripe
// layout and the fresh list to the device; sets the step arguments and the
// tree launches.
int Engine::upload_tables(uint32_t s0, uint32_t d0, uint32_t i0, uint32_t b0) {
  if (descs_dirty_) s0 = 0u;  // dead flags not yet copied: copy every descriptor
  descs_dirty_ = false;
  const uint32_t nsys = static_cast<uint32_t>(sys.size());
  n_deposit = static_cast<uint32_t>(dep_.size());
  n_integrate = static_cast<uint32_t>(integ_.size());
  n_subtree = static_cast<uint32_t>(sub_.size());
  std::vector<uint32_t> stripes(2u * kLevels * kStripes);
  uint32_t list_total = 0;
  for (uint32_t i = 0; i < kLevels * kStripes; ++i) {
    stripes[i] = list_total;
    stripes[kLevels * kStripes + i] = caps_host[i];
    list_total += caps_host[i];
  }
  ++set_gen_;
  split_wanted = split_auto;
  int rc;
  if ((rc = sys_buf.ensure(sizeof(SysDesc) * nsys)) || (rc = deposit_buf.ensure(sizeof(BodyBlock) * n_deposit)) ||
      (rc = fresh_buf.ensure(sizeof(BodyBlock) * n_fresh)) ||
      (rc = integrate_buf.ensure(sizeof(BodyBlock) * n_integrate)) ||
      (rc = subtree_buf.ensure(sizeof(SubtreeBlock) * n_subtree)) ||
      (rc = list_buf.ensure(sizeof(uint32_t) * list_total)) ||
      (rc = count_buf.ensure(sizeof(uint32_t) * kLevels * kStripes)) ||
      (rc = stripe_buf.ensure(sizeof(uint32_t) * stripes.size())))
    return rc;
  args.sys = static_cast<const SysDesc*>(sys_buf.p);
  args.sorting = 0u;
  args.deposit_blocks = static_cast<const BodyBlock*>(deposit_buf.p);
  args.integrate_blocks = static_cast<const BodyBlock*>(integrate_buf.p);
  args.subtree_blocks = static_cast<const SubtreeBlock*>(subtree_buf.p);
  args.lists = static_cast<uint32_t*>(list_buf.p);
  args.list_count = static_cast<uint32_t*>(count_buf.p);
  args.stripe_off = static_cast<const uint32_t*>(stripe_buf.p);
  args.stripe_cap = static_cast<const uint32_t*>(stripe_buf.p) + kLevels * kStripes;
  args.nsys = nsys;
  args.compensated = (options & RAINBOW_OPT_COMPENSATED) ? 1u : 0u;
  for (uint32_t l = 1; l <= kLevels - 1u; ++l) {
    uint64_t items = 0;
    for (uint32_t j = 0; j < kStripes; ++j) items += caps_host[(l - 1u) * kStripes + j];
    tree_blocks[l] = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(items, 1u), kTreeMaxBlocks));
  }
  plan_tree();
  // Stage the new parts of every table in one pinned buffer and copy them.
  const size_t b_sys = sizeof(SysDesc) * (nsys - s0), b_dep = sizeof(BodyBlock) * (n_deposit - d0),
               b_int = sizeof(BodyBlock) * (n_integrate - i0), b_sub = sizeof(SubtreeBlock) * (n_subtree - b0),
               b_str = sizeof(uint32_t) * stripes.size(), b_fr = sizeof(BodyBlock) * n_fresh;
  TableStage* st = nullptr;
  if ((rc = next_stage(b_sys + b_dep + b_int + b_sub + b_str + b_fr, &st)) != RAINBOW_OK) return rc;
  char* h = static_cast<char*>(st->host);
  auto put = [&](void* dev, const void* src, size_t bytes) -> hipError_t {
    if (bytes == 0u) return hipSuccess;
    std::memcpy(h, src, bytes);
    const hipError_t e = hipMemcpyAsync(dev, h, bytes, hipMemcpyHostToDevice, stream);
    h += bytes;
    return e;
  };
  RB_HIP(put(static_cast<SysDesc*>(sys_buf.p) + s0, sys.data() + s0, b_sys));
  RB_HIP(put(static_cast<BodyBlock*>(deposit_buf.p) + d0, dep_.data() + d0, b_dep));
  RB_HIP(put(static_cast<BodyBlock*>(integrate_buf.p) + i0, integ_.data() + i0, b_int));
  RB_HIP(put(static_cast<SubtreeBlock*>(subtree_buf.p) + b0, sub_.data() + b0, b_sub));
  RB_HIP(put(stripe_buf.p, stripes.data(), b_str));
  RB_HIP(put(fresh_buf.p, fresh_.data(), b_fr));
  RB_HIP(hipEventRecord(st->done, stream));
  return RAINBOW_OK;
}

// New physics for every system.  A nonzero scale replaces the systems'
// fixed-point scale; the finest moments the last fused deposit wrote are at
// the old scale, so the next step deposits again (as after an upload; the
// reference re-deposits every frame).
int Engine::set_physics(const Physics& ph) {
  bool rescaled = false;
  for (SysDesc& d : sys) {
    apply_physics(d, ph);
    if (ph.scale != 0.0f && ph.scale != d.scale) {
      d.scale = ph.scale;
      rescaled = true;
    }
  }
  const int rc = push_descriptors();
  if (rc == RAINBOW_OK && rescaled) {
    since_sort = sort_interval;
    pending_deposit = true;
  }
  return rc;
}

int Engine::push_descriptors() {
  RB_HIP(hipStreamSynchronize(stream));
  if (!sys.empty())
    RB_HIP(hipMemcpyAsync(sys_buf.p, sys.data(), sizeof(SysDesc) * sys.size(), hipMemcpyHostToDevice, stream));
  RB_HIP(hipStreamSynchronize(stream));
  return RAINBOW_OK;
}

// Tree launches from the split mask: a launch starts at level 2 and at
// every level whose bit is set, and runs up to the next start (or dmax).
// Bit 0 instead selects the two-launch form: the M2L sums of levels
// 2..dmax-1, then every level's fields from them.
void Engine::plan_tree() {
  tree_launches.clear();
  auto blocks_of = [&](uint32_t lo, uint32_t hi) {
    uint64_t blocks = 0;
    for (uint32_t k = lo; k <= hi; ++k) blocks += tree_blocks[k];
    return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(blocks, 1u), kTreeMaxBlocks));
  };
  // The pair tile of each launch: the shape's, or under tree_chunks 0 wave
  // items for a launch whose levels held at least kWaveItemsMin items in
  // the last measured step (block items of 8 chunks otherwise).
  auto tile = [&](uint32_t kind, uint32_t lo, uint32_t hi) -> uint32_t {
    if (shape.tree_chunks != 0u) return shape.tree_chunks;
    if (kind == kTreePassRun && lo < hi) return 8u;
    uint64_t items = 0;
    for (uint32_t k = (kind == kTreePassFinish ? 2u : lo); k <= hi; ++k) items += level_items[k];
    return items >= kWaveItemsMin ? 1u : 8u;
  };
  auto push = [&](uint32_t kind, uint32_t lo, uint32_t hi, uint32_t lead) {
    tree_launches.push_back(TreeLaunch{kind, lo, hi, blocks_of(lo, hi), lead, tile(kind, lo, hi)});
  };
  if (tree_split & 1u) {
    if (sys.size() == 1u && dmax >= 4u) {
      // One system: the M2L launch also runs the top levels and the items
      // of levels 2..dmax-3 (one block per possible parent at levels
      // 1..dmax-4), so the step has no launch of its own for them.
      uint32_t lead = 1u;
      for (uint32_t l = 1u; l + 4u <= dmax; ++l) lead += 1u << (3u * l);
      push(kTreePassM2l, dmax - 2u, dmax - 1u, lead);
    } else if (dmax >= 3u) {
      push(kTreePassM2l, 2u, dmax - 1u, 0u);
    }
    if (dmax >= 2u) push(kTreePassFinish, 2u, dmax, 0u);
    return;
  }
  uint32_t l = 2u;
  while (l <= dmax) {
    uint32_t hi = l;
    while (hi + 1u <= dmax && !((tree_split >> (hi + 1u)) & 1u)) ++hi;
    push(kTreePassRun, l, hi, 0u);
    l = hi + 1u;
  }
}

int Engine::set_tree_split(uint32_t starts) {
  if (starts & ~0x7Du) return fail(RAINBOW_E_INVALID, "tree split bits must be 0 or within levels 2..6");
  split_auto = false;
  split_wanted = false;
  return apply_tree_split(starts);
}

// Launches already enqueued keep their own split (they were launched with
// it); only captured graphs must wait before they are dropped.
int Engine::apply_tree_split(uint32_t starts) {
  starts |= 1u << 2;
  if (starts != tree_split && (graph_chunk[0] != nullptr || graph_chunk[1] != nullptr)) {
    RB_HIP(hipStreamSynchronize(stream));
    drop_graphs();
  }
  tree_split = starts;
  plan_tree();
  return RAINBOW_OK;
}

// From now on systems join only through configure_async (fresh, zeroed
// storage): nothing is pending, and the next step sorts.
void Engine::start_incremental() {
  pending_deposit = false;
  since_sort = sort_interval;
}

// The measured split: a level starts a launch of its own when its internal
// parents fill two thirds or more of the resident blocks of a single-level
// launch (then a fused launch would recompute ancestors for many parents);
// the levels with fewer parents join the launch below them, where one
// launch's latency serves several levels.  When that fuses every level but
// the finest, the two-launch form replaces the runs.  (Measured on the 1M-body Plummer
// step: 1343 level-6 parents for 1536 resident blocks run best alone, 603
// level-5 parents best fused.)  Without `wait`, returns when the counts of
// the recording step are not in yet.
int Engine::adapt_split(bool wait) {
  if (!counts_pending) return RAINBOW_OK;
  if (!wait) {
    const hipError_t q = hipEventQuery(ev_counts);
    if (q == hipErrorNotReady) return RAINBOW_OK;
    if (q != hipSuccess) return fail(RAINBOW_E_HIP, "split counts: %s", hipGetErrorString(q));
  }
  RB_HIP(hipEventSynchronize(ev_counts));
  counts_pending = false;
  if (!split_auto) return RAINBOW_OK;  // a split was fixed since the counts were recorded
  // Counts recorded before the latest set change still choose a split (the
  // set changes little per step); a fresh measurement follows.
  split_wanted = counts_gen_ != set_gen_;
  const uint32_t cap =
      2u * tree_single_level_capacity((options & RAINBOW_OPT_COMPENSATED) != 0u, shape.tree_chunks) / 3u;
  uint32_t starts = 1u << 2;
  for (uint32_t level = 2u; level <= dmax && level < kLevels; ++level) {
    uint64_t items = 0;  // internal parents (level - 1) of this level's targets
    for (uint32_t j = 0; j < kStripes; ++j) {
      const uint32_t e = (level - 1u) * kStripes + j;
      items += std::min(counts_host[e], caps_host.empty() ? 0u : caps_host[e]);
    }
    level_items[level] = static_cast<uint32_t>(std::min<uint64_t>(items, 0xFFFFFFFFu));
    if (level >= 3u && items >= cap) starts |= 1u << level;
  }
  // Levels 2..dmax-1 fused below a finest level alone: the two-launch form
  // instead (the M2L sums of the upper levels without the ancestor path,
  // then all fields; 1M-body Plummer: 25.6 us against 28.2).
  if (dmax >= 4u && starts == ((1u << 2) | (1u << dmax))) starts = 1u;
  // Wave items (one item per wave) run only unfused launches, so with them
  // each level is a launch of its own (saturated serving: 2.15e10
  // particle-steps/s against 2.13 with the block items' rule and 2.0 with
  // the two-launch form).
  if (shape.tree_chunks == 1u) starts = ((1u << (dmax + 1u)) - 1u) & 0x7Cu;
  return apply_tree_split(starts);
}

// A new shape drops the captured graphs; a measured split is measured
// again (it depends on the tree launches' resident blocks).
int Engine::set_shape(const Shape& s) {
  if (!shape_valid(s))
    return fail(RAINBOW_E_INVALID, "block shape: integrate threads must be 1024 or 512 and tree chunks 8, 4, 1 or 0 (got %u, %u)",
                s.integrate_threads, s.tree_chunks);
  if (s.integrate_threads == shape.integrate_threads && s.tree_chunks == shape.tree_chunks) return RAINBOW_OK;
  RB_HIP(hipStreamSynchronize(stream));
  drop_graphs();
  shape = s;
  ++set_gen_;
  if (split_auto) split_wanted = true;
  plan_tree();  // the launches' tiles follow the shape
  return RAINBOW_OK;
}

// Levels 2..hi in one launch, every deeper level alone.
int Engine::set_tree_fusion(uint32_t hi) {
  if (hi < 2u || hi > kLevels - 1u) return fail(RAINBOW_E_INVALID, "tree fusion level must be 2..%u", kLevels - 1u);
  uint32_t starts = 1u << 2;
  for (uint32_t l = hi + 1u; l < kLevels; ++l) starts |= 1u << l;
  return set_tree_split(starts);
}

// One step (or, without integrate, one field solve) at finest-level parity
// `step_parity`.  A step's deposit normally happened at the end of the
// previous step's integrate (k_integrate_deposit); `deposit_first` runs it
// here instead, into zeroed buffers.  A sorting step reorders the bodies by
// the Morton code of their finest cell: its deposit wrote the ranks, its
// upward and top write the offsets, and the scatter runs after top, in the
// same stream (a second stream's fork and join cost more here than the
// overlap with the tree pass gains); the step's integrate reads the sorted
// bodies from the staging arrays and sets the blocks' owned ranges.
// `rank_next`: the fused deposit writes ranks for a sorting next step.
int Engine::enqueue(bool integrate, bool sorting, uint32_t step_parity, bool deposit_first, bool rank_next,
                    bool record_counts) {
  if (sys.empty()) return RAINBOW_OK;
  // The sorted bodies reach the body arrays only through integrate.
  sorting = sorting && integrate;
  // Systems that joined a running set since the last step arrive in caller
  // order: when this step does not sort everyone, it sorts them alone
  // (their descriptors carry the flag for this step only).
  std::vector<uint32_t> joined;
  if (integrate && !sorting && sort_interval != 0u && !deposit_first && n_fresh > 0u) {
    for (uint32_t s = 0; s < sys.size(); ++s)
      if (sys[s].live && std::find(fresh_pending_.begin(), fresh_pending_.end(), sys[s].pos) != fresh_pending_.end()) {
        sys[s].sort_now = 1u;
        joined.push_back(s);
      }
    if (!joined.empty()) descs_dirty_ = true;
  }
  const bool any_sort = sorting || !joined.empty();
  const int flushed = flush_descriptors();
  // The copy just staged carries the flags for this step; the host copy
  // drops them at once, so the next step's descriptors come without them
  // whatever happens below.
  for (uint32_t s : joined) sys[s].sort_now = 0u;
  if (!joined.empty()) descs_dirty_ = true;
  if (flushed != RAINBOW_OK) return flushed;
  StepArgs a = args;
  a.sorting = sorting ? 1u : 0u;
  a.parity = step_parity;
  a.rank_next = rank_next ? 1u : 0u;
  if (deposit_first) {
    RB_HIP(zero_finest());
    RB_HIP(launch_deposit(a, n_deposit, stream));
  } else if (n_fresh > 0u) {
    // Systems that joined since the last step: their own first deposit
    // (ranks too when this step sorts them), into this step's finest
    // buffer.
    StepArgs f = a;
    f.deposit_blocks = static_cast<const BodyBlock*>(fresh_buf.p);
    RB_HIP(launch_deposit(f, n_fresh, stream));
  }
  // The scatter: every system's bodies on a sorting step, else those of the
  // systems that joined (their deposit ranges).
  const uint32_t fresh_ranges = n_fresh;
  auto scatter = [&]() -> hipError_t {
    return sorting ? launch_sort(a, n_integrate, stream)
                   : launch_sort_ranges(a, static_cast<const BodyBlock*>(fresh_buf.p), fresh_ranges, stream);
  };
  n_fresh = 0u;
  fresh_pending_.clear();
  RB_HIP(launch_upward(a, n_subtree, stream));
  // The top levels run in their own launch, or inside the first tree
  // launch; the sort's scatter follows whichever runs them.
  const bool top_in_tree = !tree_launches.empty() && tree_launches[0].lead != 0u;
  if (!top_in_tree) {
    RB_HIP(launch_top(a, stream));
    if (any_sort) RB_HIP(scatter());
  }
  for (size_t i = 0; i < tree_launches.size(); ++i) {
    const TreeLaunch& t = tree_launches[i];
    RB_HIP(launch_tree(a, t.kind, t.lo, t.hi, t.blocks, t.lead, t.chunks, stream));
    if (top_in_tree && i == 0u && any_sort) RB_HIP(scatter());
  }
  if (record_counts) {
    // The step's internal-cell counts, for the measured tree split (the
    // integrate kernel clears them for the next step).
    RB_HIP(hipMemcpyAsync(counts_host, count_buf.p, sizeof(uint32_t) * kLevels * kStripes, hipMemcpyDeviceToHost,
                          stream));
    RB_HIP(hipEventRecord(ev_counts, stream));
  }
  if (integrate) RB_HIP(launch_integrate_deposit(a, n_deposit, shape.integrate_threads, stream));
  return RAINBOW_OK;
}

// Zeroes every system's finest-level buffer `p` and straggler accumulator
// (before a standalone deposit; not captured in graphs).
// Both finest buffers are cleared: upward then zeroes the other buffer's
// cells only where the last solve's count was non-zero, which holds for a
// cleared buffer whatever the flags say.
hipError_t Engine::zero_finest() {
  for (const SysDesc& d : sys) {
    const size_t fb = 16ull * level_cells(d.depth);
    hipError_t e = hipMemsetAsync(d.finest_alt, 0, fb, stream);
    if (e == hipSuccess) e = hipMemsetAsync(d.moments + level_offset(d.depth), 0, fb, stream);
    if (e == hipSuccess) e = hipMemsetAsync(d.extra, 0, fb, stream);
    if (e == hipSuccess) e = hipMemsetAsync(d.xflag, 0, level_cells(d.depth), stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Captures `steps` steps from parity start_parity, with no deposit pending.
// Steps sort at multiples of sort_every (0: none); a step's fused deposit
// writes ranks when the following step sorts, and for the last step when
// last_rank is set.
int Engine::capture(uint32_t steps, uint32_t sort_every, bool last_rank, uint32_t start_parity,
                    hipGraphExec_t* out) {
  // Descriptor changes go out before the capture: a staged copy recorded
  // into the graph would replay whatever its staging buffer holds later.
  int frc = flush_descriptors();
  if (frc != RAINBOW_OK) return frc;
  hipGraph_t graph = nullptr;
  RB_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  int rc = RAINBOW_OK;
  for (uint32_t s = 0; s < steps && rc == RAINBOW_OK; ++s) {
    const bool sorting = sort_every != 0u && s % sort_every == 0u;
    const bool rank_next = s + 1u < steps ? (sort_every != 0u && (s + 1u) % sort_every == 0u) : last_rank;
    rc = enqueue(true, sorting, start_parity ^ (s & 1u), false, rank_next);
  }
  const hipError_t e = hipStreamEndCapture(stream, &graph);
  if (rc != RAINBOW_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(RAINBOW_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(RAINBOW_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
  return RAINBOW_OK;
}

// Steps per captured multi-step graph: kGraphSteps, or for a sort interval
// of at most kGraphSteps the largest multiple of it up to kGraphSteps.
uint32_t Engine::chunk_steps() const {
  if (sort_interval == 0u || sort_interval > kGraphSteps) return kGraphSteps;
  return sort_interval * (kGraphSteps / sort_interval);
}

// The chunk graph of the current sort interval, for parity p: interval 0,
// 16 steps without sorts; interval <= kGraphSteps, chunk_steps() steps that
// start with a sort and sort every interval steps (the next step sorts
// again); a longer interval, 16 steps without sorts, run only when the
// next sort lies beyond them.
int Engine::prepare_graphs(uint32_t steps, uint32_t p) {
  const uint32_t chunk = chunk_steps();
  if (chunk < 2u || steps < chunk || graph_chunk[p] != nullptr) return RAINBOW_OK;
  const bool periodic = sort_interval != 0u && sort_interval <= kGraphSteps;
  return capture(chunk, periodic ? sort_interval : 0u, periodic, p, &graph_chunk[p]);
}

// The next step's schedule: whether it sorts, and whether its fused
// deposit must write ranks (the step after it sorts).
void Engine::plan_next(bool* sorting, bool* rank_next, uint32_t* since_after) const {
  *sorting = sort_interval != 0u && since_sort >= sort_interval;
  *since_after = *sorting ? 1u : since_sort + 1u;
  *rank_next = sort_interval != 0u && *since_after >= sort_interval;
}

// since_sort counts the steps since the last sorting step (that step
// included); a step sorts when since_sort has reached the interval.
int Engine::step(uint32_t steps) {
  if (sys.empty() || steps == 0u) return RAINBOW_OK;
  int rc;
  const uint32_t chunk = chunk_steps();
  uint32_t done = 0;
  while (done < steps) {
    const uint32_t p = parity;
    bool use_chunk = !pending_deposit && n_fresh == 0u && chunk > 1u && steps - done >= chunk;
    if (use_chunk) {
      if (sort_interval != 0u && sort_interval <= kGraphSteps)
        use_chunk = since_sort >= sort_interval;  // starts with a sort
      else if (sort_interval != 0u)
        use_chunk = since_sort + chunk < sort_interval;  // no sort within or right after
    }
    if (use_chunk) {
      if ((rc = adapt_split(true)) != 0) return rc;
      if ((rc = prepare_graphs(steps - done, p)) != 0) return rc;
      RB_HIP(hipGraphLaunch(graph_chunk[p], stream));
      done += chunk;
      parity ^= chunk & 1u;
      if (sort_interval != 0u && sort_interval <= kGraphSteps)
        since_sort = sort_interval;
      else
        since_sort += chunk;
    } else {
      if ((rc = step_eager(1u)) != 0) return rc;
      done += 1u;
    }
  }
  return RAINBOW_OK;
}

// Steps launched directly, without graphs (the first step after a change
// of bodies or schedule, steps between chunks, and a set of systems that
// changes too often to amortise a capture).
int Engine::step_eager(uint32_t steps) {
  int rc = adapt_split(false);
  if (rc) return rc;
  for (uint32_t k = 0; k < steps; ++k) {
    bool sorting, rank_next;
    uint32_t after;
    plan_next(&sorting, &rank_next, &after);
    const bool record = split_wanted && !counts_pending;
    rc = enqueue(true, sorting, parity, pending_deposit, rank_next, record);
    if (rc) return rc;
    if (record) {
      counts_pending = true;
      counts_gen_ = set_gen_;
    }
    parity ^= 1u;
    since_sort = after;
    pending_deposit = false;
  }
  return RAINBOW_OK;
}

int Engine::set_sort_interval(uint32_t interval) {
  if (interval > 1024u) return fail(RAINBOW_E_INVALID, "sort interval must be <= 1024");
  RB_HIP(hipStreamSynchronize(stream));
  if (interval != sort_interval) drop_graphs();
  sort_interval = interval;
  since_sort = interval;  // sort before the next step
  pending_deposit = true;  // with ranks for that sort
  return RAINBOW_OK;
}

// Copies every system's bodies (pos, vel, perm) to a device snapshot, or
// back from it.
int Engine::snapshot(bool restore) {
  if ((total_bodies * 36ull) > snap_buf.cap && !restore) {
    const int rc = snap_buf.ensure(36ull * total_bodies);
    if (rc) return rc;
  }
  char* sp = static_cast<char*>(snap_buf.p);
  for (const SysDesc& d : sys) {
    const size_t b16 = 16ull * d.n, b4 = 4ull * d.n;
    if (restore) {
      RB_HIP(hipMemcpyAsync(d.pos, sp, b16, hipMemcpyDeviceToDevice, stream));
      RB_HIP(hipMemcpyAsync(d.vel, sp + b16, b16, hipMemcpyDeviceToDevice, stream));
      RB_HIP(hipMemcpyAsync(d.perm, sp + 2 * b16, b4, hipMemcpyDeviceToDevice, stream));
    } else {
      RB_HIP(hipMemcpyAsync(sp, d.pos, b16, hipMemcpyDeviceToDevice, stream));
      RB_HIP(hipMemcpyAsync(sp + b16, d.vel, b16, hipMemcpyDeviceToDevice, stream));
      RB_HIP(hipMemcpyAsync(sp + 2 * b16, d.perm, b4, hipMemcpyDeviceToDevice, stream));
    }
    sp += 36ull * d.n;
  }
  RB_HIP(hipStreamSynchronize(stream));
  if (restore) pending_deposit = true;  // the last fused deposit was of other positions
  return RAINBOW_OK;
}

// Times `steps` steps from the same starting state for each candidate,
// in turn: the block shapes (each at two tree splits), every tree split,
// every sort interval; keeps the fastest of each.  The bodies are restored
// afterwards.
int Engine::autotune(uint32_t steps, uint32_t* chosen) {
  static const uint32_t kCandidates[] = {1u, 2u, 4u, 8u, 16u, 32u, 0u};
  static const Shape kShapes[] = {{1024u, 8u}, {512u, 8u}, {1024u, 4u}, {512u, 4u},
                                  {1024u, 1u}, {512u, 1u}, {1024u, 0u}, {512u, 0u}};
  if (steps == 0u) return fail(RAINBOW_E_INVALID, "steps must be >= 1");
  int rc = snapshot(false);
  if (rc) return rc;
  // Block shapes, each at the two-launch form and at one launch per level
  // (the split that suits a shape depends on its resident blocks; the
  // search below refines it for the chosen shape).
  {
    const uint32_t per_level = ((1u << (dmax + 1u)) - 1u) & 0x7Cu;
    const uint32_t splits[2] = {1u, per_level};
    float best = 0.0f;
    Shape best_shape = shape;
    for (const Shape& c : kShapes) {
      if ((rc = set_shape(c)) != 0) return rc;
      for (uint32_t sp : splits) {
        if ((rc = set_tree_split(sp)) != 0) return rc;
        float ms = 0.0f;
        if ((rc = time_steps(std::min<uint32_t>(steps, 4u), &ms)) != 0) return rc;  // warm graphs
        if ((rc = snapshot(true)) != 0) return rc;
        if ((rc = time_steps(steps, &ms)) != 0) return rc;
        if ((rc = snapshot(true)) != 0) return rc;
        if (best == 0.0f || ms < best) {
          best = ms;
          best_shape = c;
        }
      }
    }
    if ((rc = set_shape(best_shape)) != 0) return rc;
  }
  // Tree launch split: every partition of levels 2..dmax into runs, and the
  // two-launch form (mask 1, tried last).
  {
    float best = 0.0f;
    uint32_t best_split = tree_split;
    const uint32_t free_bits = dmax > 2u ? dmax - 2u : 0u;  // levels 3..dmax
    for (uint32_t m = 0; m <= (1u << free_bits); ++m) {
      if ((rc = set_tree_split(m < (1u << free_bits) ? (1u << 2) | (m << 3) : 1u)) != 0) return rc;
      float ms = 0.0f;
      if ((rc = time_steps(std::min<uint32_t>(steps, 4u), &ms)) != 0) return rc;  // warm graphs
      if ((rc = snapshot(true)) != 0) return rc;
      if ((rc = time_steps(steps, &ms)) != 0) return rc;
      if ((rc = snapshot(true)) != 0) return rc;
      if (best == 0.0f || ms < best) {
        best = ms;
        best_split = tree_split;
      }
    }
    if ((rc = set_tree_split(best_split)) != 0) return rc;
  }
  float best = 0.0f;
  uint32_t best_interval = sort_interval;
  for (uint32_t c : kCandidates) {
    if ((rc = set_sort_interval(c)) != 0) return rc;
    float ms = 0.0f;
    if ((rc = time_steps(std::min<uint32_t>(steps, 4u), &ms)) != 0) return rc;  // warm graphs
    if ((rc = snapshot(true)) != 0) return rc;
    if ((rc = time_steps(steps, &ms)) != 0) return rc;
    if ((rc = snapshot(true)) != 0) return rc;
    if (best == 0.0f || ms < best) {
      best = ms;
      best_interval = c;
    }
  }
  if ((rc = set_sort_interval(best_interval)) != 0) return rc;
  if (chosen) *chosen = best_interval;
  return RAINBOW_OK;
}

int Engine::solve_fields() {
  const int rc = enqueue(false, false, parity, pending_deposit, false);
  if (rc != RAINBOW_OK) return rc;
  parity ^= 1u;
  pending_deposit = true;  // the next step deposits its own bodies
  return RAINBOW_OK;
}

int Engine::sync() {
  RB_HIP(hipStreamSynchronize(stream));
  return RAINBOW_OK;
}

int Engine::time_steps(uint32_t steps, float* ms) {
  int rc;
  // The split is settled and graphs of both parities are captured before
  // the timed region.
  if ((rc = adapt_split(true)) != 0) return rc;
  if ((rc = prepare_graphs(steps, 0u)) != 0 || (rc = prepare_graphs(steps, 1u)) != 0) return rc;
  RB_HIP(hipStreamSynchronize(stream));
  RB_HIP(hipEventRecord(ev_a, stream));
  if ((rc = step(steps)) != 0) return rc;
  RB_HIP(hipEventRecord(ev_b, stream));
  RB_HIP(hipEventSynchronize(ev_b));
  RB_HIP(hipEventElapsedTime(ms, ev_a, ev_b));
  return RAINBOW_OK;
}

int Engine::phase_ms(uint32_t steps, float* out, uint32_t n) {
  // deposit (standalone: the first step after a change of bodies), upward,
  // top, the tree launches, integrate with the next step's deposit, then
  // the body sort (its mean per step: the steps sort when due, as in
  // step()).
  const uint32_t nkern = 4u + static_cast<uint32_t>(tree_launches.size());
  const uint32_t nphase = nkern + 1u;
  if (out == nullptr || n < nphase) return fail(RAINBOW_E_INVALID, "phase_ms needs %u entries", nphase);
  if (steps == 0u) return fail(RAINBOW_E_INVALID, "steps must be >= 1");
  std::vector<hipEvent_t> ev(2u * nkern, nullptr);
  for (auto& e : ev) RB_HIP(hipEventCreate(&e));
  std::vector<Timing> tm(nkern);
  for (uint32_t i = 0; i < nkern; ++i) tm[i] = Timing{ev[2u * i], ev[2u * i + 1u]};
  hipEvent_t sort_ev[2] = {nullptr, nullptr};
  for (auto& e : sort_ev) RB_HIP(hipEventCreate(&e));
  std::vector<double> acc(nphase, 0.0);
  hipError_t err = hipSuccess;
  for (uint32_t s = 0; s < steps && err == hipSuccess; ++s) {
    bool sorting, rank_next;
    uint32_t after;
    plan_next(&sorting, &rank_next, &after);
    StepArgs a = args;
    a.sorting = sorting ? 1u : 0u;
    a.parity = parity;
    a.rank_next = rank_next ? 1u : 0u;
    const bool deposit_first = pending_deposit;
    bool skipped_deposit = false;
    bool top_slot_empty = false;
    parity ^= 1u;
    since_sort = after;
    pending_deposit = false;
    uint32_t k = 0;
    if (deposit_first) {
      if (err == hipSuccess) err = zero_finest();
      if (err == hipSuccess) err = launch_deposit(a, n_deposit, stream, &tm[k++]);
    } else {
      skipped_deposit = true;
      ++k;
    }
    if (err == hipSuccess) err = launch_upward(a, n_subtree, stream, &tm[k++]);
    const bool top_in_tree = !tree_launches.empty() && tree_launches[0].lead != 0u;
    auto sort_timed = [&]() {
      if (err == hipSuccess) err = hipEventRecord(sort_ev[0], stream);
      if (err == hipSuccess) err = launch_sort(a, n_integrate, stream);
      if (err == hipSuccess) err = hipEventRecord(sort_ev[1], stream);
    };
    if (top_in_tree) {
      // The top levels run inside the first tree launch: their entry stays
      // zero (no launch, no events).
      top_slot_empty = true;
      ++k;
    } else {
      if (err == hipSuccess) err = launch_top(a, stream, &tm[k++]);
      if (sorting) sort_timed();
    }
    for (size_t i = 0; i < tree_launches.size(); ++i) {
      const TreeLaunch& t = tree_launches[i];
      if (err == hipSuccess)
        err = launch_tree(a, t.kind, t.lo, t.hi, t.blocks, t.lead, t.chunks, stream, &tm[k++]);
      if (top_in_tree && i == 0u && sorting) sort_timed();
    }
    if (err == hipSuccess) err = launch_integrate_deposit(a, n_deposit, shape.integrate_threads, stream, &tm[k++]);
    if (sorting && err == hipSuccess) {
      err = hipEventSynchronize(sort_ev[1]);
      float ms = 0.0f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, sort_ev[0], sort_ev[1]);
      acc[nkern] += ms;
    }
    if (err == hipSuccess) err = hipStreamSynchronize(stream);
    for (uint32_t i = skipped_deposit ? 1u : 0u; i < nkern && err == hipSuccess; ++i) {
      if (i == 2u && top_slot_empty) continue;
      float ms = 0.0f;
      err = hipEventElapsedTime(&ms, tm[i].start, tm[i].stop);
      acc[i] += ms;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (auto& e : sort_ev) (void)hipEventDestroy(e);
  if (err != hipSuccess) return fail(RAINBOW_E_HIP, "phase timing: %s", hipGetErrorString(err));
  for (uint32_t i = 0; i < nphase; ++i) out[i] = static_cast<float>(acc[i] / steps);
  return static_cast<int>(nphase);
}

// The next deposit expects empty finest-level buffers and stragglers.
static int clear_finest(const SysDesc& d, hipStream_t stream) {
  const size_t fb = 16ull * level_cells(d.depth);
  RB_HIP(hipMemsetAsync(d.moments + level_offset(d.depth), 0, fb, stream));
  RB_HIP(hipMemsetAsync(d.finest_alt, 0, fb, stream));
  RB_HIP(hipMemsetAsync(d.extra, 0, fb, stream));
  RB_HIP(hipMemsetAsync(d.xflag, 0, level_cells(d.depth), stream));
  RB_HIP(hipStreamSynchronize(stream));
  return RAINBOW_OK;
}

int upload_system(const SysDesc& d, const void* data, size_t bytes, hipStream_t stream) {
  if (data == nullptr || bytes != 32ull * d.n)
    return fail(RAINBOW_E_INVALID, "upload expects %llu bytes, got %zu",
                static_cast<unsigned long long>(32ull * d.n), bytes);
  RB_HIP(hipMemcpyAsync(d.tpos, data, bytes, hipMemcpyHostToDevice, stream));
  RB_HIP(launch_split_particles(d, d.tpos, stream));
  return clear_finest(d, stream);
}

int download_system(const SysDesc& d, int which, void* dst, size_t bytes, hipStream_t stream,
                    uint32_t finest_parity) {
  const uint64_t cells = total_cells(d.depth);
  if (dst == nullptr) return fail(RAINBOW_E_INVALID, "null destination");
  if (which == RAINBOW_PARTICLES) {
    if (bytes != 32ull * d.n)
      return fail(RAINBOW_E_INVALID, "particles need %llu bytes", static_cast<unsigned long long>(32ull * d.n));
    RB_HIP(launch_join_particles(d, d.tpos, stream));
    RB_HIP(hipMemcpyAsync(dst, d.tpos, bytes, hipMemcpyDeviceToHost, stream));
    RB_HIP(hipStreamSynchronize(stream));
    return RAINBOW_OK;
  }
  if (which == RAINBOW_MOMENTS) {
    if (bytes != 16ull * cells)
      return fail(RAINBOW_E_INVALID, "moments need %llu bytes", static_cast<unsigned long long>(16ull * cells));
    // The finest level of the last solve may sit in the second buffer.
    const size_t coarse = 16ull * level_offset(d.depth);
    const void* finest = finest_parity ? static_cast<const void*>(d.finest_alt)
                                       : static_cast<const void*>(d.moments + level_offset(d.depth));
    RB_HIP(hipMemcpyAsync(dst, d.moments, coarse, hipMemcpyDeviceToHost, stream));
    RB_HIP(hipMemcpyAsync(static_cast<char*>(dst) + coarse, finest, bytes - coarse, hipMemcpyDeviceToHost, stream));
    RB_HIP(hipStreamSynchronize(stream));
    return RAINBOW_OK;
  }
  if (which == RAINBOW_FIELDS) {
    if (bytes != 48ull * cells)
      return fail(RAINBOW_E_INVALID, "fields need %llu bytes", static_cast<unsigned long long>(48ull * cells));
    RB_HIP(hipMemcpyAsync(dst, d.fields, bytes, hipMemcpyDeviceToHost, stream));
    RB_HIP(hipStreamSynchronize(stream));
    return RAINBOW_OK;
  }
  return fail(RAINBOW_E_INVALID, "unknown buffer %d", which);
}

int Engine::upload(uint32_t s, const void* data, size_t bytes) {
  if (s >= sys.size()) return fail(RAINBOW_E_INVALID, "system %u out of range", s);
  RB_HIP(hipStreamSynchronize(stream));
  since_sort = sort_interval;
  pending_deposit = true;
  return upload_system(sys[s], data, bytes, stream);
}

// The reference's default scene into system s (initialize,
// native/main.rs:423-434), in place of an upload.
int Engine::seed_galaxy(uint32_t s) {
  if (s >= sys.size()) return fail(RAINBOW_E_INVALID, "system %u out of range", s);
  RB_HIP(hipStreamSynchronize(stream));
  since_sort = sort_interval;
  pending_deposit = true;
  RB_HIP(launch_seed_galaxy(sys[s], stream));
  return clear_finest(sys[s], stream);
}

int Engine::download(uint32_t s, int which, void* dst, size_t bytes) {
  if (s >= sys.size()) return fail(RAINBOW_E_INVALID, "system %u out of range", s);
  return download_system(sys[s], which, dst, bytes, stream, parity ^ 1u);
}

int Engine::accelerations(uint32_t s, void* dst, size_t bytes) {
  if (s >= sys.size()) return fail(RAINBOW_E_INVALID, "system %u out of range", s);
  const SysDesc& d = sys[s];
  if (dst == nullptr || bytes != 16ull * d.n)
    return fail(RAINBOW_E_INVALID, "accelerations need %llu bytes", static_cast<unsigned long long>(16ull * d.n));
  float4* out = nullptr;
  RB_HIP(hipMalloc(reinterpret_cast<void**>(&out), 16ull * d.n));
  StepArgs a = args;
  a.parity = parity ^ 1u;  // the last solve's finest level
  hipError_t e = launch_accel(a, n_integrate, s, out, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dst, out, bytes, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(out);
  if (e != hipSuccess) return fail(RAINBOW_E_HIP, "accelerations: %s", hipGetErrorString(e));
  return RAINBOW_OK;
}

int Engine::direct_sum_f64(uint32_t s, const uint32_t* targets, uint32_t nt, double* out) {
  int rc = check_direct_args(s, targets, nt, out);
  if (rc != RAINBOW_OK || nt == 0u) return rc;
  if ((rc = io_buf.ensure(direct_scratch_bytes(nt))) != RAINBOW_OK) return rc;
  return direct_sum_f64_in(s, targets, nt, out, io_buf.p);
}

int Engine::check_direct_args(uint32_t s, const uint32_t* targets, uint32_t nt, const double* out) const {
  if (s >= sys.size()) return fail(RAINBOW_E_INVALID, "system %u out of range", s);
  if (nt == 0u) return RAINBOW_OK;
  if (targets == nullptr || out == nullptr) return fail(RAINBOW_E_INVALID, "null argument");
  const SysDesc& d = sys[s];
  for (uint32_t t = 0; t < nt; ++t)
    if (targets[t] >= d.n) return fail(RAINBOW_E_INVALID, "target %u out of range", targets[t]);
  return RAINBOW_OK;
}

// The direct sum in a scratch buffer of direct_scratch_bytes(nt) bytes
// (arguments already checked): targets in, kernels, result out, then a
// wait for the stream.
int Engine::direct_sum_f64_in(uint32_t s, const uint32_t* targets, uint32_t nt, double* out, void* scratch) {
  const DirectLayout L = direct_layout(nt);
  char* base = static_cast<char*>(scratch);
  RB_HIP(hipMemcpyAsync(base + L.targets, targets, 4ull * nt, hipMemcpyHostToDevice, stream));
  RB_HIP(launch_direct_f64(sys[s], nt, scratch, stream));
  RB_HIP(hipMemcpyAsync(out, base + L.out, 32ull * nt, hipMemcpyDeviceToHost, stream));
  RB_HIP(hipStreamSynchronize(stream));
  return RAINBOW_OK;
}

int Engine::diagnostics(uint32_t s, double out[8]) {
  if (s >= sys.size()) return fail(RAINBOW_E_INVALID, "system %u out of range", s);
  const int rc = diag_buf.ensure(sizeof(double) * (8u * static_cast<size_t>(n_integrate) + 8u));
  if (rc) return rc;
  double* partial = static_cast<double*>(diag_buf.p);
  double* result = partial + 8u * static_cast<size_t>(n_integrate);
  StepArgs a = args;
  a.parity = parity ^ 1u;  // the last solve's finest level
  RB_HIP(launch_diag(a, n_integrate, s, partial, result, stream));
  RB_HIP(hipMemcpyAsync(out, result, 8u * sizeof(double), hipMemcpyDeviceToHost, stream));
  RB_HIP(hipStreamSynchronize(stream));
  return RAINBOW_OK;
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------

Batch::~Batch() {
  if (engine.stream) (void)hipStreamSynchronize(engine.stream);
  for (auto& s : storage) free_system(&s);
}

int Batch::init(const std::vector<uint32_t>& counts, const Physics& ph, uint32_t options,
                const std::vector<float>* scales) {
  if (counts.empty() || counts.size() > 65535u) return fail(RAINBOW_E_INVALID, "systems must be in [1, 65535]");
  int rc = engine.open(options);
  if (rc) return rc;
  storage.resize(counts.size());
  std::vector<SysDesc> descs(counts.size());
  for (size_t s = 0; s < counts.size(); ++s) {
    const float scale = scales != nullptr ? (*scales)[s] : 0.0f;
    if ((rc = alloc_system(counts[s], ph, scale, engine.stream, &storage[s])) != 0) return rc;
    descs[s] = storage[s].desc;
  }
  return engine.configure(descs);
}

}  // namespace rainbow

// ---------------------------------------------------------------------------
// C ABI: solver (one system) and batch.
// ---------------------------------------------------------------------------

using rainbow::Batch;
using rainbow::fail;


#define RB_CHECK_HANDLE(h) \
  if ((h) == nullptr) return fail(RAINBOW_E_INVALID, "null handle")

extern "C" {

const char* rainbow_last_error(void) { return rainbow::last_error(); }

int rainbow_hip_create_ex(uint32_t particles, const rainbow_params_t* params, uint32_t options,
                          rainbow_hip_t** out) {
  rainbow::clear_error();
  if (out == nullptr) return fail(RAINBOW_E_INVALID, "null output handle");
  *out = nullptr;
  if (particles == 0u) return fail(RAINBOW_E_INVALID, "particles must be >= 1");
  if (options & ~static_cast<uint32_t>(RAINBOW_OPT_COMPENSATED))
    return fail(RAINBOW_E_INVALID, "unknown option bits 0x%x", options);
  rainbow_hip* h = new (std::nothrow) rainbow_hip;
  if (h == nullptr) return fail(RAINBOW_E_NOMEM, "out of host memory");
  const rainbow::Physics ph = rainbow::physics_from(params);
  const std::vector<float> scales(1, ph.scale);
  const int rc = h->batch.init(std::vector<uint32_t>(1, particles), ph, options, &scales);
  if (rc != RAINBOW_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return RAINBOW_OK;
}

int rainbow_hip_create(uint32_t particles, const rainbow_params_t* params, rainbow_hip_t** out) {
  return rainbow_hip_create_ex(particles, params, 0u, out);
}

void rainbow_hip_destroy(rainbow_hip_t* h) { delete h; }

int rainbow_hip_set_params(rainbow_hip_t* h, const rainbow_params_t* params) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.set_physics(rainbow::physics_from(params));
}

int rainbow_hip_plan(rainbow_hip_t* h, uint32_t* depth, uint32_t* cells, float* scale) {
  RB_CHECK_HANDLE(h);
  const rainbow::SysDesc& d = h->batch.engine.sys[0];
  if (depth) *depth = d.depth;
  if (cells) *cells = rainbow::total_cells(d.depth);
  if (scale) *scale = d.scale;
  return RAINBOW_OK;
}

int rainbow_hip_upload_particles(rainbow_hip_t* h, const void* particles, size_t bytes) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.upload(0u, particles, bytes);
}

int rainbow_hip_seed_galaxy(rainbow_hip_t* h) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.seed_galaxy(0u);
}

int rainbow_hip_solve_fields(rainbow_hip_t* h) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.solve_fields();
}

int rainbow_hip_step(rainbow_hip_t* h, uint32_t steps) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.step(steps);
}

int rainbow_hip_step_frame(rainbow_hip_t* h, double frame_interval_ms) {
  RB_CHECK_HANDLE(h);
  const float dt = rainbow_frame_dt(frame_interval_ms);
  for (auto& s : h->batch.engine.sys) s.dt = dt;
  const int rc = h->batch.engine.push_descriptors();
  return rc != RAINBOW_OK ? rc : h->batch.engine.step(1u);
}

int rainbow_hip_synchronize(rainbow_hip_t* h) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.sync();
}

int rainbow_hip_download(rainbow_hip_t* h, int which, void* dst, size_t bytes) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.download(0u, which, dst, bytes);
}

int rainbow_hip_digest(rainbow_hip_t* h, int which, uint32_t* out) {
  RB_CHECK_HANDLE(h);
  if (out == nullptr) return fail(RAINBOW_E_INVALID, "null output");
  const rainbow::SysDesc& d = h->batch.engine.sys[0];
  const size_t cells = rainbow::total_cells(d.depth);
  size_t bytes = 0;
  if (which == RAINBOW_PARTICLES)
    bytes = 32u * static_cast<size_t>(d.n);
  else if (which == RAINBOW_MOMENTS)
    bytes = 16u * cells;
  else if (which == RAINBOW_FIELDS)
    bytes = 48u * cells;
  else
    return fail(RAINBOW_E_INVALID, "unknown buffer %d", which);
  std::vector<uint8_t> buf(bytes);
  const int rc = h->batch.engine.download(0u, which, buf.data(), bytes);
  if (rc) return rc;
  *out = rainbow_fnv1a(buf.data(), bytes);
  return RAINBOW_OK;
}

int rainbow_hip_time_steps(rainbow_hip_t* h, uint32_t steps, float* ms) {
  RB_CHECK_HANDLE(h);
  if (ms == nullptr) return fail(RAINBOW_E_INVALID, "null output");
  return h->batch.engine.time_steps(steps, ms);
}

int rainbow_hip_phase_ms(rainbow_hip_t* h, uint32_t steps, float* out, uint32_t n) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.phase_ms(steps, out, n);
}

int rainbow_hip_set_sort_interval(rainbow_hip_t* h, uint32_t interval) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.set_sort_interval(interval);
}

int rainbow_hip_set_tree_fusion(rainbow_hip_t* h, uint32_t hi) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.set_tree_fusion(hi);
}

int rainbow_hip_set_tree_split(rainbow_hip_t* h, uint32_t starts) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.set_tree_split(starts);
}

int rainbow_hip_tree_launches(rainbow_hip_t* h, uint32_t* lo_hi, uint32_t n) {
  RB_CHECK_HANDLE(h);
  const auto& tl = h->batch.engine.tree_launches;
  if (lo_hi == nullptr || n < 2u * tl.size()) return fail(RAINBOW_E_INVALID, "tree_launches needs %zu entries", 2u * tl.size());
  for (size_t i = 0; i < tl.size(); ++i) {
    const uint32_t form = tl[i].lead != 0u ? 3u : tl[i].kind;
    lo_hi[2u * i] = tl[i].lo | (form << 8);
    lo_hi[2u * i + 1u] = tl[i].hi;
  }
  return static_cast<int>(tl.size());
}

int rainbow_hip_set_shape(rainbow_hip_t* h, uint32_t integrate_threads, uint32_t tree_chunks) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.set_shape(rainbow::Shape{integrate_threads, tree_chunks});
}

int rainbow_hip_get_shape(rainbow_hip_t* h, uint32_t* integrate_threads, uint32_t* tree_chunks) {
  RB_CHECK_HANDLE(h);
  if (integrate_threads == nullptr || tree_chunks == nullptr) return fail(RAINBOW_E_INVALID, "null output");
  *integrate_threads = h->batch.engine.shape.integrate_threads;
  *tree_chunks = h->batch.engine.shape.tree_chunks;
  return RAINBOW_OK;
}

int rainbow_hip_autotune(rainbow_hip_t* h, uint32_t steps, uint32_t* interval) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.autotune(steps, interval);
}

int rainbow_hip_tree_stats(rainbow_hip_t* h, uint32_t out[8]) {
  RB_CHECK_HANDLE(h);
  if (out == nullptr) return fail(RAINBOW_E_INVALID, "null output");
  const rainbow::SysDesc& d = h->batch.engine.sys[0];
  const uint32_t cells = rainbow::total_cells(d.depth);
  std::vector<int32_t> m(4u * static_cast<size_t>(cells));
  const int rc = h->batch.engine.download(0u, RAINBOW_MOMENTS, m.data(), 16u * static_cast<size_t>(cells));
  if (rc) return rc;
  for (int k = 0; k < 8; ++k) out[k] = 0u;
  for (uint32_t l = 1; l <= d.depth; ++l) {
    const uint32_t dim = 1u << l, pdim = dim >> 1;
    for (uint32_t i = 0; i < rainbow::level_cells(l); ++i) {
      const uint32_t z = i / (dim * dim), y = (i / dim) % dim, x = i % dim;
      const uint32_t pi = rainbow::level_offset(l - 1u) + (x / 2u) + pdim * ((y / 2u) + pdim * (z / 2u));
      const int32_t c = m[4u * (rainbow::level_offset(l) + i)];
      if (c != 0 && m[4u * pi] > 16) {
        out[l] += 1u;
        if (c <= 16 || l == d.depth) out[7] += 1u;
      }
    }
  }
  return RAINBOW_OK;
}

int rainbow_hip_pair_counts(rainbow_hip_t* h, uint64_t out[16]) {
  RB_CHECK_HANDLE(h);
  if (out == nullptr) return fail(RAINBOW_E_INVALID, "null output");
  const rainbow::SysDesc& d = h->batch.engine.sys[0];
  const uint32_t cells = rainbow::total_cells(d.depth);
  std::vector<int32_t> m(4u * static_cast<size_t>(cells));
  const int rc = h->batch.engine.download(0u, RAINBOW_MOMENTS, m.data(), 16u * static_cast<size_t>(cells));
  if (rc) return rc;
  for (int k = 0; k < 16; ++k) out[k] = 0u;
  auto count = [&](uint32_t l, int x, int y, int z) -> int32_t {
    const int dim = 1 << l;
    if (x < 0 || y < 0 || z < 0 || x >= dim || y >= dim || z >= dim) return 0;
    return m[4u * (rainbow::level_offset(l) + static_cast<uint32_t>(x + dim * (y + dim * z)))];
  };
  for (uint32_t l = 1; l <= d.depth; ++l) {
    const int dim = 1 << l;
    for (int z = 0; z < dim; ++z)
      for (int y = 0; y < dim; ++y)
        for (int x = 0; x < dim; ++x) {
          const int32_t c = count(l, x, y, z);
          if (c == 0 || count(l - 1u, x / 2, y / 2, z / 2) <= 16) continue;
          // M2L: children of the parent's neighbours that are not adjacent.
          for (int pz = z / 2 - 1; pz <= z / 2 + 1; ++pz)
            for (int py = y / 2 - 1; py <= y / 2 + 1; ++py)
              for (int px = x / 2 - 1; px <= x / 2 + 1; ++px)
                for (int k = 0; k < 8; ++k) {
                  const int sx = 2 * px + (k & 1), sy = 2 * py + ((k >> 1) & 1), sz = 2 * pz + (k >> 2);
                  if (std::abs(sx - x) <= 1 && std::abs(sy - y) <= 1 && std::abs(sz - z) <= 1) continue;
                  if (count(l, sx, sy, sz) != 0) out[l] += 1u;
                }
          if (c <= 16 || l == d.depth)
            for (int dz = -1; dz <= 1; ++dz)
              for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                  if ((dx | dy | dz) != 0 && count(l, x + dx, y + dy, z + dz) != 0) out[8 + l] += 1u;
        }
  }
  return RAINBOW_OK;
}

int rainbow_hip_diagnostics(rainbow_hip_t* h, double out[8]) {
  RB_CHECK_HANDLE(h);
  if (out == nullptr) return fail(RAINBOW_E_INVALID, "null output");
  return h->batch.engine.diagnostics(0u, out);
}

int rainbow_hip_direct_sum_f64(rainbow_hip_t* h, const uint32_t* targets, uint32_t ntargets, double* out) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.direct_sum_f64(0u, targets, ntargets, out);
}

int rainbow_hip_accelerations(rainbow_hip_t* h, void* dst, size_t bytes) {
  RB_CHECK_HANDLE(h);
  return h->batch.engine.accelerations(0u, dst, bytes);
}

int rainbow_batch_create(uint32_t systems, const uint32_t* counts, const rainbow_params_t* params,
                         uint32_t options, rainbow_batch_t** out) {
  rainbow::clear_error();
  if (out == nullptr || counts == nullptr) return fail(RAINBOW_E_INVALID, "null argument");
  *out = nullptr;
  if (options & ~static_cast<uint32_t>(RAINBOW_OPT_COMPENSATED))
    return fail(RAINBOW_E_INVALID, "unknown option bits 0x%x", options);
  for (uint32_t s = 0; s < systems; ++s)
    if (counts[s] == 0u) return fail(RAINBOW_E_INVALID, "system %u has no bodies", s);
  rainbow_batch* b = new (std::nothrow) rainbow_batch;
  if (b == nullptr) return fail(RAINBOW_E_NOMEM, "out of host memory");
  rainbow::Physics ph = rainbow::physics_from(params);
  ph.scale = 0.0f;  // every system keeps its own planner scale
  const int rc = b->batch.init(std::vector<uint32_t>(counts, counts + systems), ph, options, nullptr);
  if (rc != RAINBOW_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return RAINBOW_OK;
}

void rainbow_batch_destroy(rainbow_batch_t* b) { delete b; }

int rainbow_batch_upload(rainbow_batch_t* b, uint32_t system, const void* particles, size_t bytes) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.upload(system, particles, bytes);
}

int rainbow_batch_seed_galaxy(rainbow_batch_t* b, uint32_t system) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.seed_galaxy(system);
}

int rainbow_batch_download(rainbow_batch_t* b, uint32_t system, int which, void* dst, size_t bytes) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.download(system, which, dst, bytes);
}

int rainbow_batch_solve_fields(rainbow_batch_t* b) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.solve_fields();
}

int rainbow_batch_step(rainbow_batch_t* b, uint32_t steps) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.step(steps);
}

int rainbow_batch_time_steps(rainbow_batch_t* b, uint32_t steps, float* ms) {
  RB_CHECK_HANDLE(b);
  if (ms == nullptr) return fail(RAINBOW_E_INVALID, "null output");
  return b->batch.engine.time_steps(steps, ms);
}

int rainbow_batch_synchronize(rainbow_batch_t* b) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.sync();
}

uint64_t rainbow_batch_particles(rainbow_batch_t* b) { return b ? b->batch.engine.total_bodies : 0u; }

int rainbow_batch_set_sort_interval(rainbow_batch_t* b, uint32_t interval) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.set_sort_interval(interval);
}

int rainbow_batch_set_tree_fusion(rainbow_batch_t* b, uint32_t hi) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.set_tree_fusion(hi);
}

int rainbow_batch_set_tree_split(rainbow_batch_t* b, uint32_t starts) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.set_tree_split(starts);
}

int rainbow_batch_set_shape(rainbow_batch_t* b, uint32_t integrate_threads, uint32_t tree_chunks) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.set_shape(rainbow::Shape{integrate_threads, tree_chunks});
}

int rainbow_batch_get_shape(rainbow_batch_t* b, uint32_t* integrate_threads, uint32_t* tree_chunks) {
  RB_CHECK_HANDLE(b);
  if (integrate_threads == nullptr || tree_chunks == nullptr) return fail(RAINBOW_E_INVALID, "null output");
  *integrate_threads = b->batch.engine.shape.integrate_threads;
  *tree_chunks = b->batch.engine.shape.tree_chunks;
  return RAINBOW_OK;
}

int rainbow_batch_autotune(rainbow_batch_t* b, uint32_t steps, uint32_t* interval) {
  RB_CHECK_HANDLE(b);
  return b->batch.engine.autotune(steps, interval);
}

}  // extern "C"
