// Continuous-batching server: N-body solve requests of varying size share one
// batched step.
//
// A request is admitted when a slot and device memory (counted against a
// byte budget) are free: its storage is carved from an arena of large
// device chunks (no allocation per request once the arena has grown to the
// working set), zeroed and filled by a copy stream, and it joins the
// resident set at the next step boundary, where the step deposits it
// alone.  Every step advances all resident systems; a request leaves the set
// right after its own last step (each runs exactly its step count), its
// bodies are gathered into caller order behind that step and copied out on
// the copy stream, and its memory returns to the arena once that copy has
// finished.  The host enqueues up to kWindow steps ahead of the device and
// waits only when that window is full, so uploads, downloads, admissions
// and departures overlap the running steps: the set changes between steps
// without waiting for the device (Engine::configure_async).  Each system's
// arithmetic depends only on its own data, so a request's result is the
// same alone and in any mix.
//
// The reference has no serving loop; it advances one system per frame
// (app.js:712-741).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "rainbow_fmm.h"
#include "rainbow_launch.h"
#include "rainbow_runtime.h"

namespace rainbow {
namespace {

constexpr uint32_t kWindow = 3;          // steps enqueued ahead of the device
constexpr uint32_t kStepsPerPoll = 16;   // steps one poll may enqueue
constexpr size_t kChunkBytes = 4ull << 30;
constexpr size_t kArenaAlign = 256;

double now_s() {
  using namespace std::chrono;
  return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

#define RB_HIP(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(RAINBOW_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),       \
                  __FILE__, __LINE__);                                                        \
  } while (0)

// Device memory of the resident requests: chunks of at least kChunkBytes
// (less when the budget is smaller), allocated as the working set grows and
// kept until the server is destroyed, sub-allocated first-fit with
// neighbouring free ranges merged.
class Arena {
 public:
  explicit Arena(size_t limit) : limit_(limit) {}
  ~Arena() {
    for (Chunk& c : chunks_)
      if (c.base != nullptr) (void)hipFree(c.base);
  }
  // nullptr when the request does not fit (the caller waits for frees).
  // *grew is set when a chunk had to be allocated.
  void* alloc(size_t bytes, bool* grew) {
    bytes = (bytes + kArenaAlign - 1u) & ~(kArenaAlign - 1u);
    *grew = false;
    for (size_t ci = 0; ci < chunks_.size(); ++ci)
      if (void* p = take(ci, bytes)) return p;
    const size_t want = std::max(bytes, std::min(kChunkBytes, limit_ - std::min(limit_, reserved_)));
    if (reserved_ + want > limit_ || want < bytes) return nullptr;
    void* base = nullptr;
    if (hipMalloc(&base, want) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    *grew = true;
    reserved_ += want;
    // Reuse a trimmed entry: live blocks hold chunk indices.
    size_t ci = chunks_.size();
    for (size_t k = 0; k < chunks_.size(); ++k)
      if (chunks_[k].base == nullptr) {
        ci = k;
        break;
      }
    if (ci == chunks_.size()) chunks_.push_back(Chunk{});
    chunks_[ci] = Chunk{static_cast<char*>(base), want, {{0u, want}}};
    return take(ci, bytes);
  }
  void release(void* p) {
    auto it = live_.find(p);
    if (it == live_.end()) return;
    Chunk& c = chunks_[it->second.first];
    size_t off = static_cast<size_t>(static_cast<char*>(p) - c.base), len = it->second.second;
    live_.erase(it);
    auto next = c.free_.lower_bound(off);
    if (next != c.free_.end() && off + len == next->first) {
      len += next->second;
      next = c.free_.erase(next);
    }
    if (next != c.free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += len;
        return;
      }
    }
    c.free_.emplace(off, len);
  }
  size_t reserved() const { return reserved_; }
  // Returns the chunks that hold no block to the device (a request larger
  // than every chunk's free space can then get a chunk of its own); true if
  // any was freed.  hipFree waits for the device.
  bool trim() {
    bool any = false;
    for (Chunk& c : chunks_) {
      if (c.base == nullptr || c.free_.size() != 1u || c.free_.begin()->second != c.bytes) continue;
      (void)hipFree(c.base);
      reserved_ -= c.bytes;
      c.base = nullptr;
      c.free_.clear();
      any = true;
    }
    return any;
  }

 private:
  struct Chunk {
    char* base;
    size_t bytes;
    std::map<size_t, size_t> free_;  // offset -> length
  };
  void* take(size_t ci, size_t bytes) {
    Chunk& c = chunks_[ci];
    for (auto it = c.free_.begin(); it != c.free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      c.free_.erase(it);
      if (rest > 0u) c.free_.emplace(off + bytes, rest);
      void* p = c.base + off;
      live_[p] = {ci, bytes};
      return p;
    }
    return nullptr;
  }
  size_t limit_;
  size_t reserved_ = 0;
  std::vector<Chunk> chunks_;
  std::unordered_map<void*, std::pair<size_t, size_t>> live_;  // block -> (chunk, length)
};

// Page-locked host buffers, reused: requests are copied into one at submit
// and results land in one, so transfers run as DMA at full bus speed.  At
// most kPinnedKeepBytes of unused buffers are kept (the largest go first).
constexpr size_t kPinnedKeepBytes = 16ull << 30;
class PinnedPool {
 public:
  ~PinnedPool() {
    for (auto& kv : free_) (void)hipHostFree(kv.second);
    for (auto& kv : size_) (void)hipHostFree(kv.first);
  }
  void* get(size_t bytes) {
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2u * bytes) {
      void* p = it->second;
      size_[p] = it->first;
      free_bytes_ -= it->first;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    size_[p] = bytes;
    return p;
  }
  void put(void* p) {
    if (p == nullptr) return;
    auto it = size_.find(p);
    if (it == size_.end()) return;
    free_.emplace(it->second, p);
    free_bytes_ += it->second;
    size_.erase(it);
    while (free_bytes_ > kPinnedKeepBytes && !free_.empty()) {
      auto last = std::prev(free_.end());
      free_bytes_ -= last->first;
      (void)hipHostFree(last->second);
      free_.erase(last);
    }
  }

 private:
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> size_;  // buffers in use
  size_t free_bytes_ = 0;
};

// Events, reused.
class EventPool {
 public:
  ~EventPool() {
    for (hipEvent_t e : free_) (void)hipEventDestroy(e);
  }
  hipError_t get(hipEvent_t* out, bool timing = false) {
    std::vector<hipEvent_t>& pool = timing ? timed_ : free_;
    if (!pool.empty()) {
      *out = pool.back();
      pool.pop_back();
      return hipSuccess;
    }
    return hipEventCreateWithFlags(out, timing ? hipEventDefault : hipEventDisableTiming);
  }
  void put(hipEvent_t e, bool timing = false) {
    if (e != nullptr) (timing ? timed_ : free_).push_back(e);
  }
  void destroy_timed() {
    for (hipEvent_t e : timed_) (void)hipEventDestroy(e);
    timed_.clear();
  }

 private:
  std::vector<hipEvent_t> free_, timed_;
};

struct Pending {
  uint64_t id;
  uint32_t n;
  uint32_t steps;
  double submit_s;
  double deadline;  // 0: none
  void* data;       // 32 n bytes: pinned from the pool, or the caller's (borrowed)
  bool borrowed;
};

enum class SlotState { kLoading, kResident, kDraining };

struct Slot {
  uint64_t id;
  uint32_t n;
  uint32_t steps;
  uint32_t run = 0;  // steps enqueued
  double submit_s;
  double admit_s;
  double deadline;
  uint64_t bytes;    // counted against the budget
  void* block;       // arena storage
  void* upload;      // pinned request data from the pool, returned once `loaded` has passed
  void* result = nullptr;  // pinned, 32 n bytes (a completed request)
  hipEvent_t loaded = nullptr;  // copy stream: storage zeroed and bodies copied in
  hipEvent_t stepped = nullptr; // compute stream: the last step and the gather into caller order
  hipEvent_t done = nullptr;    // the last use of the storage (download, or last step)
  SlotState state = SlotState::kLoading;
  int32_t status = RAINBOW_OK;
  SysDesc desc;
};

struct InFlight {
  hipEvent_t start, stop;
};

struct Result {
  void* data;  // pinned, 32 n bytes
  size_t bytes;
};

}  // namespace

// Device bytes a resident request holds: its storage (bodies, cells and
// sort state).  The work tables of the batched step are reserved once for
// the whole budget when the server is created.
uint64_t request_bytes(uint32_t n) { return system_storage_bytes(n); }

}  // namespace rainbow

using rainbow::fail;

struct rainbow_server {
  rainbow::Engine engine;
  rainbow::Physics physics;
  hipStream_t copy = nullptr;
  uint64_t budget = 0;
  uint64_t used = 0;
  uint32_t max_slots = 0;
  rainbow::Arena* arena = nullptr;
  std::deque<rainbow::Pending> queue;
  std::vector<rainbow::Slot> slots;  // resident and draining
  std::deque<rainbow_completion_t> done;
  rainbow::PinnedPool pinned;
  rainbow::EventPool events;
  std::deque<rainbow::InFlight> inflight;
  std::unordered_map<uint64_t, rainbow::Result> results;
  std::unordered_set<uint64_t> ids;  // queued, resident or with a result not yet taken
  std::unordered_set<uint64_t> cancel_wanted;
  // Host seconds admitting, reconfiguring, enqueuing steps, retiring;
  // steps; polls; device seconds of the steps (events); host seconds
  // blocked on the device.
  double stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};

  ~rainbow_server() {
    if (engine.stream) (void)hipStreamSynchronize(engine.stream);
    if (copy) (void)hipStreamSynchronize(copy);
    for (auto& s : slots) {
      pinned.put(s.upload);
      pinned.put(s.result);
      events.put(s.loaded);
      events.put(s.stepped);
      events.put(s.done);
    }
    for (auto& f : inflight) {
      events.put(f.start, true);
      events.put(f.stop, true);
    }
    events.destroy_timed();
    for (auto& p : queue)
      if (!p.borrowed) pinned.put(p.data);
    for (auto& r : results) pinned.put(r.second.data);
    delete arena;
    if (copy) (void)hipStreamDestroy(copy);
  }

  void complete(const rainbow::Slot& s, double t) {
    rainbow_completion_t c{};
    c.id = s.id;
    c.particles = s.n;
    c.steps = s.steps;
    c.submit_s = s.submit_s;
    c.admit_s = s.admit_s;
    c.finish_s = t;
    c.status = s.status;
    c.steps_run = s.run;
    done.push_back(c);
  }

  // Completed step events: their device time, and the window they free.
  int reap_steps(bool wait_one) {
    while (!inflight.empty()) {
      const rainbow::InFlight f = inflight.front();
      hipError_t q = hipEventQuery(f.stop);
      if (q == hipErrorNotReady && wait_one) {
        const double t0 = rainbow::now_s();
        q = hipEventSynchronize(f.stop);
        stats[7] += rainbow::now_s() - t0;
        wait_one = false;
      }
      if (q == hipErrorNotReady) return RAINBOW_OK;
      if (q != hipSuccess) return fail(RAINBOW_E_HIP, "step event: %s", hipGetErrorString(q));
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, f.start, f.stop) == hipSuccess) stats[6] += 1e-3 * ms;
      events.put(f.start, true);
      events.put(f.stop, true);
      inflight.pop_front();
    }
    return RAINBOW_OK;
  }

  // Slots whose storage is no longer used: completion, result, memory back.
  int retire(double t) {
    for (size_t i = slots.size(); i-- > 0;) {
      rainbow::Slot& s = slots[i];
      if (s.upload != nullptr && s.loaded != nullptr) {
        const hipError_t q = hipEventQuery(s.loaded);
        if (q == hipSuccess) {
          pinned.put(s.upload);
          s.upload = nullptr;
        } else if (q != hipErrorNotReady) {
          return fail(RAINBOW_E_HIP, "upload event: %s", hipGetErrorString(q));
        }
      }
      if (s.state != rainbow::SlotState::kDraining) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) return fail(RAINBOW_E_HIP, "retire event: %s", hipGetErrorString(q));
      complete(s, t);
      if (s.status == RAINBOW_OK) {
        auto old = results.find(s.id);
        if (old != results.end()) pinned.put(old->second.data);
        results[s.id] = rainbow::Result{s.result, 32ull * s.n};
      } else {
        pinned.put(s.result);
        ids.erase(s.id);
      }
      pinned.put(s.upload);
      events.put(s.loaded);
      events.put(s.stepped);
      events.put(s.done);
      arena->release(s.block);
      used -= s.bytes;
      slots.erase(slots.begin() + static_cast<std::ptrdiff_t>(i));
    }
    return RAINBOW_OK;
  }

  // A resident slot stops (cancelled or timed out) after the steps already
  // enqueued: its storage is free once they have run.
  int stop(rainbow::Slot& s, int32_t status) {
    const bool loading = s.state == rainbow::SlotState::kLoading;
    s.status = status;
    s.state = rainbow::SlotState::kDraining;
    RB_HIP(events.get(&s.done));
    // A loading system never ran: its storage is free after its upload.
    RB_HIP(hipEventRecord(s.done, loading ? copy : engine.stream));
    if (!loading) return engine.retire_system(s.desc.pos);
    return RAINBOW_OK;
  }

  // Queued requests past their deadline or cancelled leave without running;
  // resident ones stop at the next step boundary.
  int expire(double t) {
    for (auto it = queue.begin(); it != queue.end();) {
      const bool cancelled = cancel_wanted.erase(it->id) > 0u;
      if (!cancelled && !(it->deadline > 0.0 && t >= it->deadline)) {
        ++it;
        continue;
      }
      rainbow_completion_t c{};
      c.id = it->id;
      c.particles = it->n;
      c.steps = it->steps;
      c.submit_s = it->submit_s;
      c.admit_s = 0.0;
      c.finish_s = t;
      c.status = cancelled ? RAINBOW_E_CANCELLED : RAINBOW_E_TIMEOUT;
      c.steps_run = 0u;
      done.push_back(c);
      if (!it->borrowed) pinned.put(it->data);
      ids.erase(it->id);
      it = queue.erase(it);
    }
    for (rainbow::Slot& s : slots) {
      if (s.state == rainbow::SlotState::kDraining) continue;
      const bool cancelled = cancel_wanted.erase(s.id) > 0u;
      const bool late = s.deadline > 0.0 && t >= s.deadline;
      if (cancelled || late) {
        const int rc = stop(s, cancelled ? RAINBOW_E_CANCELLED : RAINBOW_E_TIMEOUT);
        if (rc) return rc;
      }
    }
    cancel_wanted.clear();
    return RAINBOW_OK;
  }

  // Admission: memory from the arena, cleared and filled on the copy stream;
  // the compute stream waits for that copy and splits the bodies into the
  // system's streams before the step that deposits it.
  int admit() {
    uint32_t occupied = 0;
    for (const auto& s : slots)
      if (s.state != rainbow::SlotState::kDraining) ++occupied;
    while (!queue.empty() && occupied < max_slots) {
      rainbow::Pending& p = queue.front();
      const uint64_t need = rainbow::request_bytes(p.n);
      if (used + need > budget) break;
      bool grew = false;
      void* block = arena->alloc(need, &grew);
      if (block == nullptr && arena->trim()) block = arena->alloc(need, &grew);
      if (block == nullptr) break;  // fragmented: wait for frees
      rainbow::Slot s;
      s.id = p.id;
      s.n = p.n;
      s.steps = p.steps;
      s.submit_s = p.submit_s;
      s.deadline = p.deadline;
      s.bytes = need;
      s.block = block;
      s.upload = p.borrowed ? nullptr : p.data;
      s.desc = rainbow::describe_system(block, p.n, physics, 0.0f);
      // Only the part that must start zeroed is cleared: the bodies and the
      // sort's buffers are written before they are read.
      const size_t zero = rainbow::system_zeroed_from(p.n);
      hipError_t e = events.get(&s.loaded);
      if (e == hipSuccess) e = hipMemsetAsync(static_cast<char*>(block) + zero, 0, need - zero, copy);
      if (e == hipSuccess) e = hipMemcpyAsync(s.desc.tpos, p.data, 32ull * p.n, hipMemcpyHostToDevice, copy);
      if (e == hipSuccess) e = hipEventRecord(s.loaded, copy);
      if (e != hipSuccess) {
        arena->release(block);
        events.put(s.loaded);
        return fail(RAINBOW_E_HIP, "admission: %s", hipGetErrorString(e));
      }
      s.admit_s = rainbow::now_s();
      used += need;
      queue.pop_front();
      slots.push_back(s);
      ++occupied;
    }
    return RAINBOW_OK;
  }

  // Loaded systems join the set at this step boundary: the compute stream
  // splits their bodies into the system's streams (the copy has already
  // finished, so no step waits for an upload still in flight), and the next
  // step deposits them.
  int join() {
    std::vector<rainbow::SysDesc> add;
    for (rainbow::Slot& s : slots) {
      if (s.state != rainbow::SlotState::kLoading) continue;
      const hipError_t q = hipEventQuery(s.loaded);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) return fail(RAINBOW_E_HIP, "upload event: %s", hipGetErrorString(q));
      RB_HIP(rainbow::launch_split_particles(s.desc, s.desc.tpos, engine.stream));
      // Requests of zero steps are complete once their bodies are in.
      if (s.steps == 0u) {
        const int rc = finish(s, false);
        if (rc) return rc;
        continue;
      }
      s.state = rainbow::SlotState::kResident;
      add.push_back(s.desc);
    }
    return add.empty() ? RAINBOW_OK : engine.append_systems(add);
  }

  // Systems that left stay in the step's tables as dead entries (skipped by
  // every kernel); when they cover more than half of the integrate ranges,
  // the tables are rebuilt from the resident systems alone (the engine
  // carries over the ones awaiting their first deposit).
  int compact() {
    if (engine.dead_blocks() * 2u <= engine.n_integrate) return RAINBOW_OK;
    std::vector<rainbow::SysDesc> descs;
    for (const auto& s : slots)
      if (s.state == rainbow::SlotState::kResident) descs.push_back(s.desc);
    return engine.configure_async(descs, std::vector<uint32_t>());
  }

  // After a slot's last step: gather its bodies into caller order behind
  // that step, copy them out on the copy stream.
  int finish(rainbow::Slot& s, bool in_set = true) {
    s.result = pinned.get(32ull * s.n);
    if (s.result == nullptr) return fail(RAINBOW_E_NOMEM, "no pinned host memory for a %u-body result", s.n);
    RB_HIP(events.get(&s.stepped));
    hipError_t e = rainbow::launch_join_particles(s.desc, s.desc.tpos, engine.stream);
    if (e == hipSuccess) e = hipEventRecord(s.stepped, engine.stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(copy, s.stepped, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(s.result, s.desc.tpos, 32ull * s.n, hipMemcpyDeviceToHost, copy);
    if (e == hipSuccess) e = events.get(&s.done);
    if (e == hipSuccess) e = hipEventRecord(s.done, copy);
    if (e != hipSuccess) return fail(RAINBOW_E_HIP, "result copy: %s", hipGetErrorString(e));
    s.state = rainbow::SlotState::kDraining;
    return in_set ? engine.retire_system(s.desc.pos) : RAINBOW_OK;
  }

  // One step of every resident system, timed by an event pair.
  int step_once() {
    rainbow::InFlight f{nullptr, nullptr};
    RB_HIP(events.get(&f.start, true));
    RB_HIP(events.get(&f.stop, true));
    RB_HIP(hipEventRecord(f.start, engine.stream));
    int rc = engine.step_eager(1u);
    if (rc) return rc;
    RB_HIP(hipEventRecord(f.stop, engine.stream));
    inflight.push_back(f);
    stats[4] += 1.0;
    for (rainbow::Slot& s : slots) {
      if (s.state != rainbow::SlotState::kResident) continue;
      s.run += 1u;
      if (s.run >= s.steps && (rc = finish(s)) != 0) return rc;
    }
    return RAINBOW_OK;
  }
};

extern "C" {

uint64_t rainbow_server_request_bytes(uint32_t particles) {
  return particles == 0u ? 0u : rainbow::request_bytes(particles);
}

int rainbow_server_create(uint64_t budget_bytes, uint32_t max_slots, const rainbow_params_t* params,
                          uint32_t options, rainbow_server_t** out) {
  rainbow::clear_error();
  if (out == nullptr) return fail(RAINBOW_E_INVALID, "null output handle");
  *out = nullptr;
  if (max_slots == 0u || max_slots > 65535u) return fail(RAINBOW_E_INVALID, "max_slots must be in [1, 65535]");
  if (options & ~static_cast<uint32_t>(RAINBOW_OPT_COMPENSATED))
    return fail(RAINBOW_E_INVALID, "unknown option bits 0x%x", options);
  rainbow_server* s = new (std::nothrow) rainbow_server;
  if (s == nullptr) return fail(RAINBOW_E_NOMEM, "out of host memory");
  s->physics = rainbow::physics_from(params);
  s->physics.scale = 0.0f;  // every request keeps its own planner scale
  s->budget = budget_bytes;
  s->max_slots = max_slots;
  int rc = s->engine.open(options);
  // Tree tiles chosen per launch from the measured items: a full served
  // set is throughput-bound (many systems' items per launch), where wave
  // items run faster than block items; a few resident systems are not.
  if (rc == RAINBOW_OK) rc = s->engine.set_shape(rainbow::Shape{1024u, 0u});
  if (rc == RAINBOW_OK && hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess)
    rc = fail(RAINBOW_E_HIP, "copy stream creation failed");
  if (rc == RAINBOW_OK) rc = s->engine.configure(std::vector<rainbow::SysDesc>());
  if (rc == RAINBOW_OK) {
    // Table capacity for the most the budget can hold: at least 76 bytes
    // of storage per body and 128 per finest cell, at most 8^6 finest cells
    // per system.
    const uint64_t bodies = std::min<uint64_t>(budget_bytes / 76u, 0xFFFFFFFFull);
    const uint64_t finest = std::min<uint64_t>(budget_bytes / 128u, 262144ull * max_slots);
    rc = s->engine.reserve(bodies, max_slots, finest);
  }
  if (rc == RAINBOW_OK) {
    s->engine.start_incremental();
    s->arena = new (std::nothrow) rainbow::Arena(budget_bytes);
    if (s->arena == nullptr) rc = fail(RAINBOW_E_NOMEM, "out of host memory");
  }
  if (rc != RAINBOW_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return RAINBOW_OK;
}

void rainbow_server_destroy(rainbow_server_t* s) { delete s; }

int rainbow_server_submit(rainbow_server_t* s, const rainbow_request_t* r) {
  if (s == nullptr || r == nullptr) return fail(RAINBOW_E_INVALID, "null argument");
  if (r->particles == 0u || r->data == nullptr) return fail(RAINBOW_E_INVALID, "empty request");
  if (!(r->timeout_s >= 0.0)) return fail(RAINBOW_E_INVALID, "timeout must be >= 0");
  if (r->flags & ~static_cast<uint32_t>(RAINBOW_REQUEST_BORROW))
    return fail(RAINBOW_E_INVALID, "unknown request flags 0x%x", r->flags);
  const uint64_t need = rainbow::request_bytes(r->particles);
  if (need > s->budget)
    return fail(RAINBOW_E_BUDGET, "request of %u bodies needs %llu bytes, budget is %llu", r->particles,
                static_cast<unsigned long long>(need), static_cast<unsigned long long>(s->budget));
  if (s->ids.count(r->id) != 0u)
    return fail(RAINBOW_E_INVALID, "request id %llu is already in use", static_cast<unsigned long long>(r->id));
  rainbow::Pending p;
  p.id = r->id;
  p.n = r->particles;
  p.steps = r->steps;
  p.submit_s = rainbow::now_s();
  p.deadline = r->timeout_s > 0.0 ? p.submit_s + r->timeout_s : 0.0;
  p.borrowed = (r->flags & RAINBOW_REQUEST_BORROW) != 0u;
  if (p.borrowed) {
    p.data = const_cast<void*>(r->data);
  } else {
    p.data = s->pinned.get(32ull * r->particles);
    if (p.data == nullptr) return fail(RAINBOW_E_NOMEM, "no pinned host memory for the request");
    std::memcpy(p.data, r->data, 32ull * r->particles);
  }
  s->queue.push_back(p);
  s->ids.insert(r->id);
  return RAINBOW_OK;
}

int rainbow_server_cancel(rainbow_server_t* s, uint64_t id) {
  if (s == nullptr) return fail(RAINBOW_E_INVALID, "null handle");
  bool live = false;
  for (const auto& p : s->queue) live = live || p.id == id;
  for (const auto& sl : s->slots) live = live || (sl.id == id && sl.state != rainbow::SlotState::kDraining);
  if (!live) return fail(RAINBOW_E_STATE, "request %llu is not queued or running", static_cast<unsigned long long>(id));
  s->cancel_wanted.insert(id);
  return RAINBOW_OK;
}

int rainbow_server_poll(rainbow_server_t* s, uint32_t* stepped) {
  if (s == nullptr) return fail(RAINBOW_E_INVALID, "null handle");
  if (stepped) *stepped = 0u;
  s->stats[5] += 1.0;
  uint32_t steps = 0;
  int rc;
  for (;;) {
    double t0 = rainbow::now_s();
    if ((rc = s->reap_steps(false)) != 0 || (rc = s->retire(t0)) != 0 || (rc = s->expire(t0)) != 0) return rc;
    double t1 = rainbow::now_s();
    s->stats[3] += t1 - t0;
    if ((rc = s->admit()) != 0) return rc;
    t0 = rainbow::now_s();
    s->stats[0] += t0 - t1;
    if ((rc = s->join()) != 0 || (rc = s->compact()) != 0) return rc;
    t1 = rainbow::now_s();
    s->stats[1] += t1 - t0;
    bool any = false;
    for (const auto& sl : s->slots) any = any || sl.state == rainbow::SlotState::kResident;
    if (!any || steps >= rainbow::kStepsPerPoll) break;
    if (s->inflight.size() >= rainbow::kWindow) {
      // The device is busy with the window: wait for its oldest step (new
      // arrivals are admitted by the next poll).
      if ((rc = s->reap_steps(true)) != 0) return rc;
      s->stats[2] += rainbow::now_s() - t1;
      if (steps > 0u) break;
      continue;
    }
    if ((rc = s->step_once()) != 0) return rc;
    ++steps;
    s->stats[2] += rainbow::now_s() - t1;
  }
  if (stepped) *stepped = steps;
  return RAINBOW_OK;
}

int rainbow_server_drain(rainbow_server_t* s) {
  if (s == nullptr) return fail(RAINBOW_E_INVALID, "null handle");
  for (;;) {
    int rc = rainbow_server_poll(s, nullptr);
    if (rc) return rc;
    if (s->queue.empty() && s->slots.empty()) return RAINBOW_OK;
    bool resident = false;
    for (const auto& sl : s->slots) resident = resident || sl.state == rainbow::SlotState::kResident;
    if (!resident) {
      // Only copies and stops left: wait for them.
      if (hipStreamSynchronize(s->engine.stream) != hipSuccess || hipStreamSynchronize(s->copy) != hipSuccess)
        return fail(RAINBOW_E_HIP, "drain: stream synchronize failed");
      if (s->queue.empty()) {
        if ((rc = s->retire(rainbow::now_s())) != 0) return rc;
        if (s->slots.empty()) return RAINBOW_OK;
      }
    }
  }
}

int rainbow_server_stats(rainbow_server_t* s, double out[8]) {
  if (s == nullptr || out == nullptr) return fail(RAINBOW_E_INVALID, "null argument");
  for (int k = 0; k < 8; ++k) out[k] = s->stats[k];
  return RAINBOW_OK;
}

int rainbow_server_completions(rainbow_server_t* s, rainbow_completion_t* out, uint32_t cap, uint32_t* n) {
  if (s == nullptr || n == nullptr || (out == nullptr && cap > 0u)) return fail(RAINBOW_E_INVALID, "null argument");
  uint32_t k = 0;
  while (k < cap && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n = k;
  return RAINBOW_OK;
}

int rainbow_server_result(rainbow_server_t* s, uint64_t id, void* dst, size_t bytes) {
  if (s == nullptr || dst == nullptr) return fail(RAINBOW_E_INVALID, "null argument");
  auto it = s->results.find(id);
  if (it == s->results.end()) return fail(RAINBOW_E_STATE, "no result for request %llu", static_cast<unsigned long long>(id));
  if (bytes != it->second.bytes)
    return fail(RAINBOW_E_INVALID, "result of request %llu is %zu bytes", static_cast<unsigned long long>(id),
                it->second.bytes);
  std::memcpy(dst, it->second.data, bytes);
  s->pinned.put(it->second.data);
  s->results.erase(it);
  s->ids.erase(id);
  return RAINBOW_OK;
}

int rainbow_server_result_view(rainbow_server_t* s, uint64_t id, const void** data, size_t* bytes) {
  if (s == nullptr || data == nullptr || bytes == nullptr) return fail(RAINBOW_E_INVALID, "null argument");
  auto it = s->results.find(id);
  if (it == s->results.end()) return fail(RAINBOW_E_STATE, "no result for request %llu", static_cast<unsigned long long>(id));
  *data = it->second.data;
  *bytes = it->second.bytes;
  return RAINBOW_OK;
}

int rainbow_server_release(rainbow_server_t* s, uint64_t id) {
  if (s == nullptr) return fail(RAINBOW_E_INVALID, "null handle");
  auto it = s->results.find(id);
  if (it == s->results.end()) return fail(RAINBOW_E_STATE, "no result for request %llu", static_cast<unsigned long long>(id));
  s->pinned.put(it->second.data);
  s->results.erase(it);
  s->ids.erase(id);
  return RAINBOW_OK;
}

void* rainbow_host_alloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0u || hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}

void rainbow_host_free(void* p) {
  if (p != nullptr) (void)hipHostFree(p);
}

int rainbow_server_state(rainbow_server_t* s, uint32_t* queued, uint32_t* resident, uint64_t* bytes) {
  if (s == nullptr) return fail(RAINBOW_E_INVALID, "null handle");
  if (queued) *queued = static_cast<uint32_t>(s->queue.size());
  if (resident) *resident = static_cast<uint32_t>(s->slots.size());
  if (bytes) *bytes = s->used;
  return RAINBOW_OK;
}

}  // extern "C"
