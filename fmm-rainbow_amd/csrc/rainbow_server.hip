// Continuous-batching server: N-body solve requests of varying size share one
// batched step.  A request is admitted into a resident slot when a slot and
// device memory (counted against a byte budget) are free; its trees and
// particles stay on the device until its last step; every scheduling
// iteration runs batched steps over all resident systems (up to
// kServeSteps, and no further than the next request's last step), so
// requests that arrive while others run join at the next iteration and
// their work overlaps on the device.  Each system's arithmetic depends only
// on its own data, so a request's result is the same alone and in any mix.
// The set of systems changes at most iterations, so steps are launched
// directly rather than captured.
//
// The reference has no serving loop; it advances one system per frame
// (app.js:712-741).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "rainbow_fmm.h"
#include "rainbow_runtime.h"

namespace rainbow {
namespace {

constexpr uint32_t kServeSteps = 8;

double now_s() {
  using namespace std::chrono;
  return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

// Page-locked host buffers, reused: requests are copied into one at submit
// and results land in one, so transfers run as DMA at full bus speed.
class PinnedPool {
 public:
  ~PinnedPool() {
    for (auto& kv : free_) (void)hipHostFree(kv.second);
  }
  void* get(size_t bytes) {
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2u * bytes) {
      void* p = it->second;
      size_[p] = it->first;
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
    size_.erase(it);
  }

 private:
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> size_;  // buffers in use
};

struct Pending {
  uint64_t id;
  uint32_t n;
  uint32_t steps;
  double submit_s;
  void* data;  // pinned, 32 n bytes
};

struct Result {
  void* data;  // pinned, 32 n bytes
  size_t bytes;
};

struct Slot {
  uint64_t id;
  uint32_t n;
  uint32_t steps;
  uint32_t left;
  double submit_s;
  double admit_s;
  uint64_t bytes;
  SystemStorage storage;
};

}  // namespace

// Device bytes a resident request holds: its storage plus its share of the
// engine's work tables, lists and sort buffers (ranks and staging: 40 bytes
// per body; cell offsets: 8 per finest cell; owned ranges: 8 per deposit
// block).
uint64_t request_bytes(uint32_t n) {
  const Plan p = make_plan(n);
  uint64_t lists = 0;
  for (uint32_t k = 0; k < p.depth; ++k) lists += std::min<uint64_t>(level_cells(k), n / 17u);
  const uint64_t deposit_blocks = (n + kDepositBodies - 1u) / kDepositBodies;
  const uint64_t blocks = deposit_blocks + (n + kIntegrateThreads - 1u) / kIntegrateThreads + level_cells(p.depth - 3u);
  return system_storage_bytes(n) + 4u * lists + 8u * blocks +
         64u * ((n + kIntegrateThreads - 1u) / kIntegrateThreads) + sizeof(SysDesc) + 4u + 40ull * n +
         8ull * level_cells(p.depth) + 4ull * level_cells(p.depth - 3u) + 8ull * deposit_blocks;
}

}  // namespace rainbow

using rainbow::fail;

struct rainbow_server {
  rainbow::Engine engine;
  rainbow::Physics physics;
  uint64_t budget = 0;
  uint64_t used = 0;
  uint32_t max_slots = 0;
  bool dirty = false;
  std::deque<rainbow::Pending> queue;
  std::vector<rainbow::Slot> slots;
  std::deque<rainbow_completion_t> done;
  rainbow::PinnedPool pinned;
  std::unordered_map<uint64_t, rainbow::Result> results;
  // Host seconds spent admitting (allocation + upload), reconfiguring,
  // stepping (launch + wait) and retiring (download + free); steps; polls.
  double stats[6] = {0, 0, 0, 0, 0, 0};

  ~rainbow_server() {
    if (engine.stream) (void)hipStreamSynchronize(engine.stream);
    for (auto& s : slots) rainbow::free_system(&s.storage);
    for (auto& p : queue) pinned.put(p.data);
    for (auto& r : results) pinned.put(r.second.data);
  }

  int reconfigure() {
    std::vector<rainbow::SysDesc> descs;
    descs.reserve(slots.size());
    for (const auto& s : slots) descs.push_back(s.storage.desc);
    dirty = false;
    return engine.configure(descs);
  }

  int retire(size_t i, double t) {
    rainbow::Slot& s = slots[i];
    const size_t bytes = 32ull * s.n;
    void* out = pinned.get(bytes);
    if (out == nullptr) return fail(RAINBOW_E_NOMEM, "no pinned host memory for a %zu-byte result", bytes);
    const int rc = rainbow::download_system(s.storage.desc, RAINBOW_PARTICLES, out, bytes, engine.stream,
                                            &engine.io_buf);
    if (rc) {
      pinned.put(out);
      return rc;
    }
    rainbow_completion_t c;
    c.id = s.id;
    c.particles = s.n;
    c.steps = s.steps;
    c.submit_s = s.submit_s;
    c.admit_s = s.admit_s;
    c.finish_s = t;
    done.push_back(c);
    auto old = results.find(s.id);
    if (old != results.end()) pinned.put(old->second.data);
    results[s.id] = rainbow::Result{out, bytes};
    rainbow::free_system(&s.storage);
    used -= s.bytes;
    slots.erase(slots.begin() + static_cast<std::ptrdiff_t>(i));
    dirty = true;
    return RAINBOW_OK;
  }

  int admit() {
    while (!queue.empty() && slots.size() < max_slots) {
      rainbow::Pending& p = queue.front();
      const uint64_t need = rainbow::request_bytes(p.n);
      if (used + need > budget) break;
      rainbow::Slot s;
      s.id = p.id;
      s.n = p.n;
      s.steps = p.steps;
      s.left = p.steps;
      s.submit_s = p.submit_s;
      s.bytes = need;
      int rc = rainbow::alloc_system(p.n, physics, 0.0f, engine.stream, &s.storage);
      if (rc) return rc;
      rc = rainbow::upload_system(s.storage.desc, p.data, 32ull * p.n, engine.stream, &engine.io_buf);
      if (rc) {
        rainbow::free_system(&s.storage);
        return rc;
      }
      pinned.put(p.data);
      s.admit_s = rainbow::now_s();
      used += need;
      queue.pop_front();
      slots.push_back(std::move(s));
      dirty = true;
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
  if (rc == RAINBOW_OK) rc = s->reconfigure();
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
  const uint64_t need = rainbow::request_bytes(r->particles);
  if (need > s->budget)
    return fail(RAINBOW_E_BUDGET, "request of %u bodies needs %llu bytes, budget is %llu", r->particles,
                static_cast<unsigned long long>(need), static_cast<unsigned long long>(s->budget));
  rainbow::Pending p;
  p.id = r->id;
  p.n = r->particles;
  p.steps = r->steps;
  p.submit_s = rainbow::now_s();
  p.data = s->pinned.get(32ull * r->particles);
  if (p.data == nullptr) return fail(RAINBOW_E_NOMEM, "no pinned host memory for the request");
  std::memcpy(p.data, r->data, 32ull * r->particles);
  s->queue.push_back(p);
  return RAINBOW_OK;
}

int rainbow_server_poll(rainbow_server_t* s, uint32_t* stepped) {
  if (s == nullptr) return fail(RAINBOW_E_INVALID, "null handle");
  if (stepped) *stepped = 0u;
  s->stats[5] += 1.0;
  double t0 = rainbow::now_s();
  int rc = s->admit();
  if (rc) return rc;
  double t1 = rainbow::now_s();
  s->stats[0] += t1 - t0;
  // Requests of zero steps complete as soon as they are resident.
  for (size_t i = s->slots.size(); i-- > 0;)
    if (s->slots[i].left == 0u && (rc = s->retire(i, t1)) != 0) return rc;
  t0 = rainbow::now_s();
  s->stats[3] += t0 - t1;
  if (s->dirty && (rc = s->reconfigure()) != 0) return rc;
  t1 = rainbow::now_s();
  s->stats[1] += t1 - t0;
  if (s->slots.empty()) return RAINBOW_OK;
  uint32_t k = rainbow::kServeSteps;
  for (const auto& slot : s->slots) k = std::min(k, slot.left);
  if ((rc = s->engine.step_eager(k)) != 0) return rc;
  if ((rc = s->engine.sync()) != 0) return rc;
  const double t = rainbow::now_s();
  s->stats[2] += t - t1;
  s->stats[4] += k;
  if (stepped) *stepped = static_cast<uint32_t>(s->slots.size());
  for (auto& slot : s->slots) slot.left -= k;
  for (size_t i = s->slots.size(); i-- > 0;)
    if (s->slots[i].left == 0u && (rc = s->retire(i, t)) != 0) return rc;
  s->stats[3] += rainbow::now_s() - t;
  return RAINBOW_OK;
}

int rainbow_server_stats(rainbow_server_t* s, double out[6]) {
  if (s == nullptr || out == nullptr) return fail(RAINBOW_E_INVALID, "null argument");
  for (int k = 0; k < 6; ++k) out[k] = s->stats[k];
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
  return RAINBOW_OK;
}

int rainbow_server_state(rainbow_server_t* s, uint32_t* queued, uint32_t* resident, uint64_t* bytes) {
  if (s == nullptr) return fail(RAINBOW_E_INVALID, "null handle");
  if (queued) *queued = static_cast<uint32_t>(s->queue.size());
  if (resident) *resident = static_cast<uint32_t>(s->slots.size());
  if (bytes) *bytes = s->used;
  return RAINBOW_OK;
}

}  // extern "C"
