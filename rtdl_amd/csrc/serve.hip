// Request loop with continuous batching and memory-budget admission.
//
// Requests (a query set of one predicate against one resident scene) arrive
// at their replay times.  The loop keeps up to `slots` batches in flight,
// each on its own stream with its own BatchEngine.  Row buffers of every
// in-flight batch come from one device arena of min(budget, slots x the
// largest admissible batch) bytes: a batch occupies one contiguous range of
// it, placed first-fit in the free space left by the batches still in
// flight, so the rows in flight never exceed the budget.  Whenever a slot is
// free, the requests that have arrived and are not yet dispatched are
// admitted in arrival order while their row buffers fit the largest free
// range and the batch stays under `max_batch_queries`; the first request
// that does not fit waits (head of line) until completions free space.  A
// request whose rows can exceed the whole arena is rejected.  Batches on
// different slots overlap on the device, so requests that arrive while
// others run join the next batch as soon as a slot and memory free up.  A
// request's latency runs from its arrival time to the moment the loop
// observes its batch's completion event.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "runtime.hpp"
#include "serve.hpp"

namespace rtdl {
namespace {

constexpr uint32_t kStatusRejected = 4u;
constexpr uint32_t kStatusFailed = 16u;
constexpr uint32_t kStatusTimeout = 32u;
constexpr uint32_t kStatusInvalid = 64u;

struct Slot {
  hipStream_t stream = nullptr;
  std::unique_ptr<BatchEngine> engine;
  hipEvent_t done = nullptr;
  bool busy = false;
  bool late = false;                       // the batch passed its deadline
  double dispatched = 0.0;
  uint64_t range_at = 0, range_bytes = 0;  // the batch's rows in the arena
  std::vector<size_t> requests;           // request indices of the batch in flight
};

uint32_t stride_of(uint32_t kind) {
  switch (kind) {
    case kJobFixedRadius:
    case kJobKnn:
    case kJobBoundedKnn: return 3;
    case kJobPipPositive: return 2;
    default: return 7;
  }
}

// Largest free range of the arena given the ranges of the busy slots.
void largest_gap(const std::vector<Slot>& slots, uint64_t arena, uint64_t* at, uint64_t* bytes) {
  std::pair<uint64_t, uint64_t> used[16];
  size_t n = 0;
  for (const Slot& s : slots)
    if (s.busy && s.range_bytes) used[n++] = {s.range_at, s.range_bytes};
  std::sort(used, used + n);
  uint64_t cursor = 0;
  *at = 0;
  *bytes = 0;
  for (size_t i = 0; i <= n; ++i) {
    const uint64_t end = i < n ? used[i].first : arena;
    if (end - cursor > *bytes) {
      *at = cursor;
      *bytes = end - cursor;
    }
    if (i < n) cursor = used[i].first + used[i].second;
  }
}

}  // namespace

void serve_replay(void* const* scenes, size_t n_scenes, const RtdlServeRequest* requests,
                  size_t n_requests, const RtdlServeConfig& cfg, RtdlServeResult* results,
                  RtdlServeStats* stats) {
  require(cfg.slots >= 1 && cfg.slots <= 16, "slots must be in [1, 16]");
  require(cfg.memory_budget_bytes >= 4096ull, "memory budget too small");
  require(cfg.max_batch_queries >= 1, "max_batch_queries must be positive");
  std::vector<SceneView> views(n_scenes);
  std::vector<SceneType> types(n_scenes);
  for (size_t i = 0; i < n_scenes; ++i) {
    Prepared* p = prepared_from(scenes[i]);
    views[i] = p->scene.view(0);
    types[i] = p->type;
  }
  // a request that cannot be served is marked and skipped; the others run
  std::vector<uint8_t> valid(n_requests, 1);
  for (size_t r = 0; r < n_requests; ++r) {
    const RtdlServeRequest& q = requests[r];
    if (r) require(q.arrival_s >= requests[r - 1].arrival_s, "arrival times must be non-decreasing");
    std::memset(&results[r], 0, sizeof(RtdlServeResult));
    results[r].arrival_s = q.arrival_s;
    bool ok = q.scene < n_scenes && q.kind <= kJobPipPositive;
    if (ok) {
      const SceneType t = types[q.scene];
      const bool point_kind = q.kind <= kJobBoundedKnn;
      const bool ray_kind = q.kind >= kJobRayAnyHit && q.kind <= kJobRayClosest;
      ok = (!point_kind || t == SceneType::Points3D) && (!ray_kind || t == SceneType::Triangles3D) &&
           (q.kind != kJobPipPositive || t == SceneType::Polygons);
      const float* arena = point_kind ? cfg.point_queries : ray_kind ? cfg.ray_queries : cfg.pip_queries;
      ok = ok && (q.query_count == 0 || arena != nullptr);
    }
    if (!ok) {
      valid[r] = 0;
      results[r].status = kStatusInvalid;
    }
  }
  require(cfg.k >= 1 && cfg.k <= 64 && cfg.k_max >= 1 && cfg.k_max <= 64,
          "served k and k_max must be in [1, 64]");

  auto row_capacity = [&](const RtdlServeRequest& q) -> uint64_t {
    return job_row_capacity(q.kind, q.query_count, q.kind == kJobKnn ? cfg.k : cfg.k_max, kPipHitCap);
  };
  auto row_need = [&](const RtdlServeRequest& q) -> uint64_t {
    return (row_capacity(q) * job_row_bytes(q.kind) + 255) & ~uint64_t(255);
  };

  const size_t n_slots = cfg.slots;
  // The arena never exceeds the budget; physically it only needs what
  // `slots` batches of the largest admissible size can emit.
  const uint64_t per_query = std::max<uint64_t>(
      {(uint64_t)cfg.k_max * sizeof(FrnRowRec), (uint64_t)cfg.k * sizeof(KnnRowRec),
       (uint64_t)kPipHitCap * sizeof(PipRowRec), sizeof(ClosestRowRec)});
  const uint64_t batch_cap = (cfg.max_batch_queries + 65536) * per_query + (uint64_t(64) << 20);
  const uint64_t arena_bytes =
      std::min<uint64_t>(cfg.memory_budget_bytes, n_slots * batch_cap) & ~uint64_t(255);
  DeviceArray<uint8_t> arena(arena_bytes);

  // Captured rows go to one pinned pool sized before the loop, so the loop
  // itself only enqueues copies; host copies are made after it.
  std::vector<uint64_t> capture_at(n_requests, 0);
  uint64_t capture_total = 0;
  for (size_t r = 0; r < n_requests; ++r) {
    if (!requests[r].capture || !valid[r]) continue;
    capture_at[r] = capture_total;
    capture_total += row_capacity(requests[r]) * job_row_bytes(requests[r].kind);
  }
  uint8_t* capture_pool = nullptr;
  if (capture_total) RTDL_HIP(hipHostMalloc((void**)&capture_pool, capture_total, hipHostMallocDefault));
  struct PinnedPool {
    uint8_t* p;
    ~PinnedPool() {
      if (p) (void)hipHostFree(p);
    }
  } pool_guard{capture_pool};

  std::vector<Slot> slots(n_slots);
  struct SlotGuard {
    std::vector<Slot>& s;
    ~SlotGuard() {
      for (Slot& x : s) {
        if (x.stream) (void)hipStreamSynchronize(x.stream);
        x.engine.reset();
        if (x.done) (void)hipEventDestroy(x.done);
        if (x.stream) (void)hipStreamDestroy(x.stream);
      }
    }
  } slot_guard{slots};
  for (Slot& s : slots) {
    RTDL_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    RTDL_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    s.engine = std::make_unique<BatchEngine>(s.stream);
    s.engine->reserve(4096, cfg.max_batch_queries + 65536);
  }
  RTDL_HIP(hipDeviceSynchronize());

  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto now = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };

  std::deque<size_t> pending;
  std::deque<size_t> solo;  // requests of a failed batch, re-run one at a time
  std::vector<uint8_t> retried(n_requests, 0);
  size_t next = 0, completed = 0, batch_id = 0;
  uint64_t inflight_bytes = 0, peak_bytes = 0, total_q = 0, total_rows = 0, batches = 0;
  uint64_t n_retried = 0, n_failed = 0;
  std::vector<HostJob> jobs;
  std::vector<uint64_t> need_of(n_requests);
  for (size_t r = 0; r < n_requests; ++r) need_of[r] = valid[r] ? row_need(requests[r]) : 0;

  while (completed < n_requests) {
    bool progress = false;
    const double t = now();
    while (next < n_requests && (!cfg.realtime || requests[next].arrival_s <= t)) {
      if (valid[next]) {
        pending.push_back(next);
      } else {
        results[next].dispatch_s = results[next].done_s = t;
        ++completed;
      }
      ++next;
    }

    // completions
    for (Slot& s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) {
        if (cfg.timeout_ms && !s.late && (now() - s.dispatched) * 1e3 > cfg.timeout_ms) s.late = true;
        continue;
      }
      RTDL_HIP(q);
      const double td = now();
      // a batch past its deadline (or a forced timeout) keeps its slot and
      // rows until the device is done with them; its requests are marked
      if (cfg.timeout_ms && (td - s.dispatched) * 1e3 > cfg.timeout_ms) s.late = true;
      if (fault_fires(FaultKind::Timeout)) s.late = true;
      const JobCounters* c = s.engine->counters();
      for (size_t j = 0; j < s.requests.size(); ++j) {
        RtdlServeResult& res = results[s.requests[j]];
        res.done_s = td;
        res.rows = c[j].rows;
        res.status |= c[j].status | (s.late ? kStatusTimeout : 0u);
        total_rows += c[j].rows;
      }
      s.late = false;
      completed += s.requests.size();
      inflight_bytes -= s.range_bytes;
      s.range_bytes = 0;
      s.requests.clear();
      s.busy = false;
      progress = true;
    }

    // dispatch: requests of a failed batch first, one per batch
    for (Slot& s : slots) {
      if (pending.empty() && solo.empty()) break;
      if (s.busy) continue;
      uint64_t gap_at = 0, gap = 0;
      largest_gap(slots, arena_bytes, &gap_at, &gap);
      jobs.clear();
      uint64_t used = 0, queries = 0;
      const bool alone = !solo.empty();
      std::deque<size_t>& from = alone ? solo : pending;
      while (!from.empty() && !(alone && !jobs.empty())) {
        const size_t r = from.front();
        const RtdlServeRequest& q = requests[r];
        const uint64_t need = need_of[r];
        if (need > arena_bytes) {  // can never be admitted
          results[r].status |= kStatusRejected;
          results[r].dispatch_s = results[r].done_s = now();
          from.pop_front();
          ++completed;
          continue;
        }
        if (used + need > gap || (!jobs.empty() && queries + q.query_count > cfg.max_batch_queries)) break;
        HostJob h;
        h.scene = views[q.scene];
        h.kind = q.kind;
        h.nq = q.query_count;
        h.qstride = stride_of(q.kind);
        const float* qa = q.kind <= kJobBoundedKnn ? cfg.point_queries
                          : q.kind == kJobPipPositive ? cfg.pip_queries
                                                      : cfg.ray_queries;
        h.queries = qa + q.query_offset * h.qstride;
        h.radius = cfg.radius;
        h.k = q.kind == kJobKnn ? cfg.k : cfg.k_max;
        h.rows = arena.get() + gap_at + used;
        h.row_capacity = row_capacity(q);
        jobs.push_back(h);
        s.requests.push_back(r);
        used += need;
        queries += q.query_count;
        from.pop_front();
      }
      if (jobs.empty()) break;  // the head of the line waits for memory
      try {
        s.engine->enqueue(jobs, cfg.morton_sort != 0);
        for (size_t j = 0; j < jobs.size(); ++j) {
          const size_t r = s.requests[j];
          if (!requests[r].capture) continue;
          const uint64_t bytes = jobs[j].row_capacity * job_row_bytes(requests[r].kind);
          if (bytes)
            RTDL_HIP(hipMemcpyAsync(capture_pool + capture_at[r], jobs[j].rows, bytes, hipMemcpyDeviceToHost,
                                    s.stream));
        }
        RTDL_HIP(hipEventRecord(s.done, s.stream));
      } catch (const Error&) {
        // whatever was enqueued finishes before the slot and rows are reused;
        // the batch's requests run again alone, so a failure stays with the
        // request that causes it
        RTDL_HIP(hipStreamSynchronize(s.stream));
        const double tf = now();
        for (size_t r : s.requests) {
          if (alone || retried[r]) {
            results[r].status |= kStatusFailed;
            results[r].dispatch_s = results[r].done_s = tf;
            ++completed;
            ++n_failed;
          } else {
            retried[r] = 1;
            ++n_retried;
          }
        }
        if (!alone)
          for (auto it = s.requests.rbegin(); it != s.requests.rend(); ++it) solo.push_front(*it);
        s.requests.clear();
        progress = true;
        continue;
      }
      s.dispatched = now();
      s.late = false;
      const double td = now();
      for (size_t r : s.requests) {
        results[r].dispatch_s = td;
        results[r].batch = (uint32_t)batch_id;
        total_q += requests[r].query_count;
      }
      ++batch_id;
      ++batches;
      s.range_at = gap_at;
      s.range_bytes = used;
      inflight_bytes += used;
      peak_bytes = std::max(peak_bytes, inflight_bytes);
      s.busy = true;
      progress = true;
    }

    if (!progress) {
      bool any_busy = false;
      for (const Slot& s : slots) any_busy = any_busy || s.busy;
      if (!any_busy && pending.empty() && solo.empty() && next < n_requests && cfg.realtime) {
        const double wait = requests[next].arrival_s - now();
        if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait * 1e6) - 100));
      } else {
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    }
  }
  const double wall = now();
  for (Slot& s : slots) RTDL_HIP(hipStreamSynchronize(s.stream));
  for (size_t r = 0; r < n_requests; ++r) {
    if (!requests[r].capture || (results[r].status & (kStatusRejected | kStatusFailed | kStatusInvalid)))
      continue;
    // an overflowing request reserved more rows than its buffer holds; only
    // the buffer was copied
    const uint64_t rows = std::min<uint64_t>(results[r].rows, row_capacity(requests[r]));
    const size_t bytes = rows * job_row_bytes(requests[r].kind);
    void* out = std::malloc(bytes ? bytes : 1);
    require(out != nullptr, "host allocation for captured rows failed");
    if (bytes) std::memcpy(out, capture_pool + capture_at[r], bytes);
    results[r].captured_rows = out;
  }
  if (stats) {
    stats->wall_s = wall;
    stats->batches = batches;
    stats->queries = total_q;
    stats->rows = total_rows;
    stats->peak_inflight_bytes = peak_bytes;
    stats->region_bytes = arena_bytes;
    stats->retried = n_retried;
    stats->failed = n_failed;
  }
}

}  // namespace rtdl
