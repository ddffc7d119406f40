// Request loop with continuous batching.
//
// Requests (a query set of one predicate against one resident scene) arrive
// at their replay times.  The loop keeps `slots` batches in flight, each on
// its own stream with its own BatchEngine and its own slice of the row
// budget.  Whenever a slot is free, every request that has arrived and not
// yet been dispatched is admitted into a new batch, in arrival order, while
// its row buffer fits the slot's slice of the memory budget and the batch
// stays under `max_batch_queries`; the batch is enqueued and the loop goes
// back to polling.  Batches on different slots overlap on the device, so
// requests that arrive while others run join the next batch as soon as a
// slot frees up.  A request's latency runs from its arrival time to the
// moment the loop observes its batch's completion event.
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

struct Slot {
  hipStream_t stream = nullptr;
  std::unique_ptr<BatchEngine> engine;
  DeviceArray<uint8_t> region;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<size_t> requests;  // request indices of the batch in flight
  std::vector<std::pair<size_t, uint8_t*>> captures;  // (request, pinned copy)
  std::vector<uint64_t> capture_rows;                  // row capacity of each copy
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

}  // namespace

void serve_replay(void* const* scenes, size_t n_scenes, const RtdlServeRequest* requests,
                  size_t n_requests, const RtdlServeConfig& cfg, RtdlServeResult* results,
                  RtdlServeStats* stats) {
  require(cfg.slots >= 1 && cfg.slots <= 16, "slots must be in [1, 16]");
  require(cfg.memory_budget_bytes >= cfg.slots * 4096ull, "memory budget too small");
  require(cfg.max_batch_queries >= 1, "max_batch_queries must be positive");
  std::vector<SceneView> views(n_scenes);
  std::vector<SceneType> types(n_scenes);
  for (size_t i = 0; i < n_scenes; ++i) {
    Prepared* p = prepared_from(scenes[i]);
    views[i] = p->scene.view(0);
    types[i] = p->type;
  }
  for (size_t r = 0; r < n_requests; ++r) {
    const RtdlServeRequest& q = requests[r];
    require(q.scene < n_scenes, "request scene index out of range");
    require(q.kind <= kJobPipPositive, "unknown request kind");
    const SceneType t = types[q.scene];
    const bool point_kind = q.kind <= kJobBoundedKnn;
    const bool ray_kind = q.kind >= kJobRayAnyHit && q.kind <= kJobRayClosest;
    require(!point_kind || t == SceneType::Points3D, "neighbour requests need a 3-D point scene");
    require(!ray_kind || t == SceneType::Triangles3D, "ray requests need a triangle scene");
    require(q.kind != kJobPipPositive || t == SceneType::Polygons, "PIP requests need a polygon scene");
    const float* arena = point_kind ? cfg.point_queries : ray_kind ? cfg.ray_queries : cfg.pip_queries;
    require(q.query_count == 0 || arena != nullptr, "query arena for a request kind is null");
    if (r) require(q.arrival_s >= requests[r - 1].arrival_s, "arrival times must be non-decreasing");
    std::memset(&results[r], 0, sizeof(RtdlServeResult));
    results[r].arrival_s = q.arrival_s;
  }
  require(cfg.k >= 1 && cfg.k <= 64 && cfg.k_max >= 1 && cfg.k_max <= 64,
          "served k and k_max must be in [1, 64]");

  const size_t n_slots = cfg.slots;
  // Each slot owns an equal share of the budget; physically it only needs
  // what its largest admissible batch can emit.
  const uint64_t per_query = std::max<uint64_t>(
      {(uint64_t)cfg.k_max * sizeof(FrnRowRec), (uint64_t)cfg.k * sizeof(KnnRowRec),
       (uint64_t)kPipHitCap * sizeof(PipRowRec), sizeof(ClosestRowRec)});
  const uint64_t batch_cap = (cfg.max_batch_queries + 65536) * per_query + (uint64_t(64) << 20);
  const uint64_t region_bytes =
      std::min<uint64_t>(cfg.memory_budget_bytes / n_slots, batch_cap) & ~uint64_t(255);
  std::vector<Slot> slots(n_slots);
  for (Slot& s : slots) {
    RTDL_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    RTDL_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    s.engine = std::make_unique<BatchEngine>(s.stream);
    s.engine->reserve(4096, cfg.max_batch_queries + 65536);
    s.region.resize(region_bytes);
  }
  RTDL_HIP(hipDeviceSynchronize());

  auto row_need = [&](const RtdlServeRequest& q) -> uint64_t {
    const uint32_t k = q.kind == kJobKnn ? cfg.k : cfg.k_max;
    const uint64_t rows = job_row_capacity(q.kind, q.query_count, k, kPipHitCap);
    return (rows * job_row_bytes(q.kind) + 255) & ~uint64_t(255);
  };

  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto now = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };

  std::deque<size_t> pending;
  size_t next = 0, completed = 0, batch_id = 0;
  uint64_t inflight_bytes = 0, peak_bytes = 0, total_q = 0, total_rows = 0, batches = 0;
  std::vector<HostJob> jobs;
  std::vector<uint64_t> need_of(n_requests);
  for (size_t r = 0; r < n_requests; ++r) need_of[r] = row_need(requests[r]);
  std::vector<uint64_t> slot_bytes(n_slots, 0);

  while (completed < n_requests) {
    bool progress = false;
    const double t = now();
    while (next < n_requests && (!cfg.realtime || requests[next].arrival_s <= t)) pending.push_back(next++);

    // completions
    for (size_t si = 0; si < n_slots; ++si) {
      Slot& s = slots[si];
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      RTDL_HIP(q);
      const double td = now();
      const JobCounters* c = s.engine->counters();
      for (size_t j = 0; j < s.requests.size(); ++j) {
        RtdlServeResult& res = results[s.requests[j]];
        res.done_s = td;
        res.rows = c[j].rows;
        res.status |= c[j].status;
        total_rows += c[j].rows;
      }
      for (size_t ci = 0; ci < s.captures.size(); ++ci) {
        const size_t r = s.captures[ci].first;
        // an overflowing request reserved more rows than its buffer holds;
        // only the buffer was copied
        const size_t bytes = std::min<uint64_t>(results[r].rows, s.capture_rows[ci]) *
                             job_row_bytes(requests[r].kind);
        void* out = std::malloc(bytes ? bytes : 1);
        require(out != nullptr, "host allocation for captured rows failed");
        std::memcpy(out, s.captures[ci].second, bytes);
        RTDL_HIP(hipHostFree(s.captures[ci].second));
        results[r].captured_rows = out;
      }
      completed += s.requests.size();
      inflight_bytes -= slot_bytes[si];
      slot_bytes[si] = 0;
      s.requests.clear();
      s.captures.clear();
      s.capture_rows.clear();
      s.busy = false;
      progress = true;
    }

    // dispatch
    for (size_t si = 0; si < n_slots && !pending.empty(); ++si) {
      Slot& s = slots[si];
      if (s.busy) continue;
      jobs.clear();
      uint64_t used = 0, queries = 0;
      while (!pending.empty()) {
        const size_t r = pending.front();
        const RtdlServeRequest& q = requests[r];
        const uint64_t need = need_of[r];
        if (need > region_bytes) {  // can never be admitted
          results[r].status |= kStatusRejected;
          results[r].dispatch_s = results[r].done_s = now();
          pending.pop_front();
          ++completed;
          continue;
        }
        if (!jobs.empty() && (used + need > region_bytes || queries + q.query_count > cfg.max_batch_queries))
          break;
        HostJob h;
        h.scene = views[q.scene];
        h.kind = q.kind;
        h.nq = q.query_count;
        h.qstride = stride_of(q.kind);
        const float* arena = q.kind <= kJobBoundedKnn ? cfg.point_queries
                             : q.kind == kJobPipPositive ? cfg.pip_queries
                                                         : cfg.ray_queries;
        h.queries = arena + q.query_offset * h.qstride;
        h.radius = cfg.radius;
        h.k = q.kind == kJobKnn ? cfg.k : cfg.k_max;
        h.rows = s.region.get() + used;
        h.row_capacity = job_row_capacity(q.kind, q.query_count, h.k, kPipHitCap);
        jobs.push_back(h);
        s.requests.push_back(r);
        used += need;
        queries += q.query_count;
        pending.pop_front();
      }
      if (jobs.empty()) continue;
      s.engine->enqueue(jobs, cfg.morton_sort != 0);
      for (size_t j = 0; j < jobs.size(); ++j) {
        const size_t r = s.requests[j];
        if (!requests[r].capture) continue;
        const uint64_t bytes = jobs[j].row_capacity * job_row_bytes(requests[r].kind);
        uint8_t* pinned = nullptr;
        RTDL_HIP(hipHostMalloc((void**)&pinned, bytes ? bytes : 1, hipHostMallocDefault));
        if (bytes)
          RTDL_HIP(hipMemcpyAsync(pinned, jobs[j].rows, bytes, hipMemcpyDeviceToHost, s.stream));
        s.captures.emplace_back(r, pinned);
        s.capture_rows.push_back(jobs[j].row_capacity);
      }
      RTDL_HIP(hipEventRecord(s.done, s.stream));
      const double td = now();
      for (size_t r : s.requests) {
        results[r].dispatch_s = td;
        results[r].batch = (uint32_t)batch_id;
        total_q += requests[r].query_count;
      }
      ++batch_id;
      ++batches;
      slot_bytes[si] = used;
      inflight_bytes += used;
      peak_bytes = std::max(peak_bytes, inflight_bytes);
      s.busy = true;
      progress = true;
    }

    if (!progress) {
      bool any_busy = false;
      for (const Slot& s : slots) any_busy = any_busy || s.busy;
      if (!any_busy && pending.empty() && next < n_requests && cfg.realtime) {
        const double wait = requests[next].arrival_s - now();
        if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait * 1e6) - 100));
      } else {
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    }
  }
  const double wall = now();
  for (Slot& s : slots) {
    (void)hipStreamSynchronize(s.stream);
    s.engine.reset();
    (void)hipEventDestroy(s.done);
    (void)hipStreamDestroy(s.stream);
  }
  if (stats) {
    stats->wall_s = wall;
    stats->batches = batches;
    stats->queries = total_q;
    stats->rows = total_rows;
    stats->peak_inflight_bytes = peak_bytes;
    stats->region_bytes = region_bytes;
  }
}

}  // namespace rtdl
