// Batched execution: many independent jobs (query sets against resident
// scenes or scene instances) in one persistent launch per predicate family
// over a device work queue, with rows compacted per job by wavefront
// prefix sums.  Everything is enqueued asynchronously on one stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

#include "common.hpp"
#include "primitives.hpp"
#include "scene.hpp"

namespace rtdl {

enum JobKind : uint32_t {
  kJobFixedRadius = 0,
  kJobKnn = 1,
  kJobBoundedKnn = 2,
  kJobRayAnyHit = 3,
  kJobRayHitCount = 4,
  kJobRayClosest = 5,
  kJobPipPositive = 6,
};

// Row records written by batch jobs (query ids are job-local indices).
struct FrnRowRec {
  uint32_t query_id;
  uint32_t neighbor_id;
  double distance;
};
struct KnnRowRec {
  uint32_t query_id;
  uint32_t neighbor_id;
  double distance;
  uint32_t neighbor_rank;
  uint32_t pad;
};
struct PairRowRec {
  uint32_t id;
  uint32_t value;
};
struct ClosestRowRec {
  uint32_t ray_id;
  uint32_t triangle_id;
  double t;
};
struct PipRowRec {
  uint32_t point_id;
  uint32_t polygon_id;
  uint32_t contains;
};

size_t job_row_bytes(uint32_t kind);
// Upper bound of rows a job can emit (PIP: `pip_cap` per point).
uint64_t job_row_capacity(uint32_t kind, uint32_t nq, uint32_t k, uint32_t pip_cap);

constexpr uint32_t kPipHitCap = 16;   // default PIP row budget per point of a served request
constexpr int kPointBuckets = 7;    // KNN list sizes 1, 2, 4, ..., 64
constexpr int kLaunchFamilies = kPointBuckets + 5;  // + fixed radius, ray any / count / closest, PIP
constexpr uint32_t kStatusRowOverflow = 1u;
constexpr uint32_t kStatusInternal = 8u;

// Per-job counters.  `rows` counts reserved rows in 64 bits, so reservations
// past a job's capacity (which set kStatusRowOverflow and write nothing)
// never wrap round to offsets inside it; capacities stay below 2^32, so every
// row actually written has a 32-bit offset.
struct JobCounters {
  unsigned long long rows;
  uint32_t status;
  uint32_t ambiguous;
};
static_assert(sizeof(JobCounters) == 16, "JobCounters is cleared as four words");
constexpr uint64_t kMaxJobRows = 0xFFFFFFFFull;

struct DeviceJob {
  SceneView scene;        // root / end delimit the job's instance subtree
  const float* queries;   // FP32 queries, `qstride` floats each
  const uint32_t* order;  // batch traversal order (job-local index per position) or null
  void* rows;
  JobCounters* counters;
  uint64_t row_capacity;
  double radius;          // +inf for KNN
  uint32_t kind;
  uint32_t nq;
  uint32_t k;
  uint32_t qstride;       // 2 (2-D points), 3 (3-D points) or 7 (rays)
  uint32_t order_base;
  uint32_t pad;
};

// Per-query plan of a neighbour job between its count, emit and sort
// launches (indexed by work item * 64 + lane).
struct NbrPlan {
  uint64_t row_at;   // first output row
  uint64_t cand_at;  // first parked candidate (queries with more than k neighbours)
  double radius;     // ball walked: the job's radius, or the KNN bound
  double radius_sq;  // membership: squared distance <= radius_sq
  uint32_t n_rows;   // rows of the query (0: nothing to emit or sort here)
  uint32_t n_cand;   // parked candidates; 0 when the rows are written directly
  uint32_t bucket;   // last distance bucket parked
  uint32_t qi;       // job-local query index
  uint32_t staged;   // slots the count walk staged (rows come from them), or 0
  uint32_t pad;
};

// A parked neighbour candidate.
struct NbrCand {
  double distance;
  uint32_t id;
  uint32_t pad;
};

struct WorkItem {
  uint32_t job;
  uint32_t first;  // first position in the job
  uint32_t count;  // <= 64
  uint32_t group;  // launch family << 24 | scene index (items sort by group, then place)
};

struct HostJob {
  SceneView scene;
  uint32_t kind = kJobFixedRadius;
  uint32_t nq = 0;
  const float* queries = nullptr;
  uint32_t qstride = 3;
  double radius = 0.0;
  uint32_t k = 1;
  void* rows = nullptr;
  uint64_t row_capacity = 0;
};

// Memory an engine works in when it must allocate nothing: every device
// array is carved out of `device`, every host staging array out of `host`
// (plain memory: copies from it are staged by the runtime), sized for at most
// `max_jobs` jobs and `max_queries` queries per batch (see fixed_bytes).
struct EngineStorage {
  void* device = nullptr;
  size_t device_bytes = 0;
  void* host = nullptr;
  size_t host_bytes = 0;
  size_t max_jobs = 0;
  size_t max_queries = 0;
};

// One engine drives one batch at a time on its stream; reuse it only after
// the previous batch has completed (done_event / stream sync).
class BatchEngine {
 public:
  // Grows its device buffers (caching pool) and pinned host staging on demand.
  explicit BatchEngine(hipStream_t stream);
  // Allocation-free engine over caller-provided storage: it creates no
  // events, never grows, and rejects batches beyond the storage's limits.
  BatchEngine(hipStream_t stream, const EngineStorage& storage);
  ~BatchEngine();
  BatchEngine(const BatchEngine&) = delete;
  BatchEngine& operator=(const BatchEngine&) = delete;

  // Bytes of device and host storage a fixed engine needs for these limits.
  static void fixed_bytes(size_t max_jobs, size_t max_queries, size_t* device_bytes, size_t* host_bytes);

  void reserve(size_t max_jobs, size_t max_queries);
  // Sizes the fixed-radius parking space and stage for a batch's jobs.
  void reserve_nbr(const HostJob* jobs, size_t n);
  // Enqueues the whole batch; returns immediately.
  void enqueue(const HostJob* jobs, size_t n, bool morton_sort);
  void enqueue(const std::vector<HostJob>& jobs, bool morton_sort) {
    enqueue(jobs.data(), jobs.size(), morton_sort);
  }
  // Records the batch into an executable hipGraph instead of running it:
  // buffers are sized first, then the stream is captured while the batch is
  // enqueued, so every launch of the graph replays the whole batch (job
  // descriptors, ordering sort, family launches, counter copy) with one
  // call.  Timing and visit counting are off in a captured batch.  The
  // engine then belongs to the graph: it must not enqueue other batches.
  // `prologue` (optional) enqueues work recorded before the batch, such as
  // scene rebuilds; it must only enqueue capturable work on the stream given.
  hipGraphExec_t capture(const std::vector<HostJob>& jobs, bool morton_sort,
                         const std::function<void(hipStream_t)>& prologue = {});
  bool fixed() const { return fixed_; }
  // Valid once the batch has completed.
  const JobCounters* counters() const { return host_counters_; }
  hipStream_t stream() const { return stream_; }
  hipEvent_t done_event() const { return done_; }
  // node / primitive visit counters of the last batch (when stats enabled)
  void enable_stats(bool on) { stats_on_ = on; }
  unsigned long long* device_stats() { return stats_.get(); }
  // per-family kernel times of the last batch, from events on the stream
  // (points incl. the exact fix-up, ray any-hit, ray hit-count, ray closest,
  // PIP); valid once the batch has completed
  void enable_timing(bool on) { timing_on_ = on; }
  void family_seconds(double out[5]) const;

 private:
  struct MeasureTag {};
  explicit BatchEngine(MeasureTag) : stream_(nullptr), fixed_(true) {}
  // Lays every buffer out in `dev` / `host` (measures when their bases are null).
  void carve(Carver& dev, Carver& host, size_t max_jobs, size_t max_queries);
  hipStream_t stream_;
  hipEvent_t done_ = nullptr;
  size_t cap_jobs_ = 0, cap_queries_ = 0, cap_items_ = 0;
  DeviceJob* host_jobs_ = nullptr;       // pinned (fixed engine: caller storage)
  WorkItem* host_items_ = nullptr;       // pinned (fixed engine: caller storage)
  uint64_t* host_starts_ = nullptr;      // pinned (fixed engine: caller storage)
  JobCounters* host_counters_ = nullptr; // pinned (fixed engine: caller storage)
  uint32_t* host_scene_order_ = nullptr; // jobs by scene key (fixed engine: caller storage)
  uint32_t* host_scene_idx_ = nullptr;   // scene index of each job
  std::vector<uint32_t> host_scene_storage_;  // the two arrays above (growing engine)
  bool fixed_ = false;
  DeviceArray<DeviceJob> jobs_;
  DeviceArray<WorkItem> items_;
  DeviceArray<WorkItem> items_sorted_;   // items_ in (family, scene, Morton) order
  DeviceArray<uint64_t> item_keys_;
  DeviceArray<uint32_t> item_vals_;
  DeviceArray<uint64_t> starts_;
  DeviceArray<JobCounters> counters_;
  DeviceArray<uint32_t> item_counters_;  // eight per launch (one per XCD)
  DeviceArray<uint64_t> keys_;
  DeviceArray<uint32_t> order_;
  DeviceArray<uint4> amb_;               // (job, query, row base, count)
  DeviceArray<uint4> pip_over_;          // (job, point, row base, hit count)
  DeviceArray<uint32_t> amb_count_;
  DeviceArray<NbrPlan> nbr_plan_;        // cap_items_ * 64
  DeviceArray<NbrCand> nbr_cand_;        // parked candidates of a batch
  DeviceArray<uint32_t> nbr_stage_;      // kNbrStageMax slots per fixed-radius query of a sub-launch
  DeviceArray<uint32_t> nbr_emit_;       // plan indices of the queries the emit walk writes
  DeviceArray<unsigned long long> nbr_cand_count_;
  DeviceArray<unsigned long long> stats_;
  SortWorkspace sort_ws_;
  bool stats_on_ = false;
  bool timing_on_ = false;
  bool capturing_ = false;
  bool timed_[kLaunchFamilies] = {};
  hipEvent_t fam_start_[kLaunchFamilies] = {}, fam_stop_[kLaunchFamilies] = {};
  int cus_ = 256;
};

}  // namespace rtdl
