// Batched execution engine: persistent kernels over a device work queue.
//
// A batch holds many independent jobs (a query set against one resident
// scene instance).  Each job's queries are cut into 64-query work items; one
// persistent launch per predicate family lets every wavefront pull items
// from an atomic counter until the queue is empty, so uneven jobs balance
// across the 256 CUs.  Optionally the batch's queries are first put in
// (job, Morton code) order with the LDS radix sort so that the lanes of a
// wavefront traverse neighbouring parts of the tree.  Rows are compacted per
// job with a wavefront exclusive prefix sum and one atomic per wavefront.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>

#include "common.hpp"
#include "engine.hpp"
#include "primitives.hpp"
#include "queries.hpp"
#include "traverse.hpp"

namespace rtdl {

size_t job_row_bytes(uint32_t kind) {
  switch (kind) {
    case kJobFixedRadius: return sizeof(FrnRowRec);
    case kJobKnn:
    case kJobBoundedKnn: return sizeof(KnnRowRec);
    case kJobRayAnyHit:
    case kJobRayHitCount: return sizeof(PairRowRec);
    case kJobRayClosest: return sizeof(ClosestRowRec);
    default: return sizeof(PipRowRec);
  }
}

uint64_t job_row_capacity(uint32_t kind, uint32_t nq, uint32_t k, uint32_t pip_cap) {
  switch (kind) {
    case kJobFixedRadius:
    case kJobKnn:
    case kJobBoundedKnn: return (uint64_t)nq * k;
    case kJobPipPositive: return (uint64_t)nq * pip_cap;
    default: return nq;
  }
}

namespace {

// The skeleton of every persistent loop below is free of lane-dependent
// branches: each wave-wide step is an atomic that all 64 lanes execute (lanes
// other than 0 add zero) followed by readfirstlane, whose SGPR result is
// uniform by construction.  A `if (lane == 0) atomicAdd(...)` at the end of
// the body and another at the head of the loop let the compiler thread the
// two branches across the latch, which split the loop so that lane 0 left to
// fetch the next item while lanes 1-63 kept re-reading the old one and never
// exited; a shuffle-broadcast item index (divergent to the compiler) allows
// the same split.  Every caller runs with all 64 lanes active, so the first
// active lane is lane 0.
__device__ inline uint32_t lane0_add(uint32_t* counter, uint32_t value) {
  const uint32_t mine = (threadIdx.x & 63) == 0 ? value : 0u;
  // counters live in device global memory; the global-address-space atomic
  // lets the compiler combine the wave's 64 adds into a single atomic
  auto* g = (__attribute__((address_space(1))) uint32_t*)counter;
  return __builtin_amdgcn_readfirstlane(
      __hip_atomic_fetch_add(g, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ inline uint32_t wave_exclusive_sum(uint32_t v, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    x += lane >= (uint32_t)d ? y : 0u;
  }
  *total = __builtin_amdgcn_readlane(x, 63);
  return x - v;
}

// Work queue of one persistent launch, split by XCD.  Workgroups are handed
// to the eight XCDs in turn, and each XCD has its own L2, so XCD x first takes
// items from the x-th eighth of the queue (consecutive items are neighbouring
// Morton ranges of a job, so its L2 keeps serving the same part of the tree)
// and then helps with the other eighths.  Each call makes one attempt:
// kItemRetry means "this eighth is exhausted, call again"; kItemDone means
// all eight are.  Every value is wave-uniform.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kItemDone = 0xFFFFFFFFu;
constexpr uint32_t kItemRetry = 0xFFFFFFFEu;

struct ItemQueue {
  uint32_t* counters;  // kXcds counters
  uint32_t n;
  uint32_t src;
  uint32_t tries = 0;
  __device__ ItemQueue(uint32_t* c, uint32_t n_items) : counters(c), n(n_items), src(blockIdx.x % kXcds) {}
  __device__ inline uint32_t next() {
    if (tries >= kXcds) return kItemDone;
    const uint32_t lo = (uint32_t)(((uint64_t)n * src) / kXcds);
    const uint32_t hi = (uint32_t)(((uint64_t)n * (src + 1)) / kXcds);
    const uint32_t it = lo + lane0_add(counters + src, 1u);
    if (it < hi) return it;
    src = (src + 1) % kXcds;
    ++tries;
    return kItemRetry;
  }
};

__device__ inline uint64_t lane0_add64(unsigned long long* counter, uint32_t value) {
  const unsigned long long mine = (threadIdx.x & 63) == 0 ? value : 0ull;
  auto* g = (__attribute__((address_space(1))) unsigned long long*)counter;
  const unsigned long long old = __hip_atomic_fetch_add(g, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)old);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(old >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Reserves `c` rows per lane in the job's row buffer; returns the lane's
// first row.
__device__ inline uint64_t wave_reserve(JobCounters* counters, uint32_t c) {
  uint32_t total;
  const uint32_t prefix = wave_exclusive_sum(c, &total);
  return lane0_add64(&counters->rows, total) + prefix;
}

__device__ inline uint32_t job_query(const DeviceJob& j, uint32_t pos) {
  return j.order ? j.order[j.order_base + pos] : pos;
}

__device__ inline void write_neighbor_row(const DeviceJob& j, uint64_t at, uint32_t qi, uint32_t id,
                                          double dist, uint32_t rank) {
  if (j.kind == kJobFixedRadius) {
    static_cast<FrnRowRec*>(j.rows)[at] = FrnRowRec{qi, id, dist};
  } else {
    static_cast<KnnRowRec*>(j.rows)[at] = KnnRowRec{qi, id, dist, rank, 0u};
  }
}

__device__ inline void read_neighbor_row(const DeviceJob& j, uint64_t at, double& dist, uint32_t& id) {
  if (j.kind == kJobFixedRadius) {
    const FrnRowRec r = static_cast<const FrnRowRec*>(j.rows)[at];
    dist = r.distance;
    id = r.neighbor_id;
  } else {
    const KnnRowRec r = static_cast<const KnnRowRec*>(j.rows)[at];
    dist = r.distance;
    id = r.neighbor_id;
  }
}

// KNN and bounded-KNN jobs: each lane walks its query near-first with a
// PackedTopK list in registers whose k-th key shrinks the walk; K is the list
// size the job's k needs.  Queries whose FP32 keys tie at the boundary go to
// batch_fix_kernel.
template <int K, bool STATS, uint32_t DEPTH>
__global__ __launch_bounds__(256) void batch_points_kernel(const DeviceJob* __restrict__ jobs,
                                                           const WorkItem* __restrict__ items,
                                                           uint32_t n_items, uint32_t* item_counter,
                                                           uint4* amb, uint32_t* amb_count,
                                                           unsigned long long* stats) {
  __shared__ uint32_t stack[DEPTH * 256];  // walk_nearest's per-lane stacks
  const uint32_t lane = threadIdx.x & 63;
  LaneStats st;
  ItemQueue queue(item_counter, n_items);
  for (;;) {
    const uint32_t it = queue.next();
    if (it == kItemDone) break;
    if (it == kItemRetry) continue;
    const WorkItem w = items[it];
    const DeviceJob& j = jobs[w.job];
    const bool active = lane < w.count;
    uint32_t qi = 0;
    double qx = 0.0, qy = 0.0, qz = 0.0;
    PackedTopK<K> list;
    list.init(1);
    bool ambiguous = false;
    uint32_t c = 0;
    const uint32_t k = j.k;
    if (active) {
      qi = job_query(j, w.first + lane);
      const float* qp = j.queries + (size_t)qi * j.qstride;
      qx = qp[0];
      qy = qp[1];
      qz = j.qstride >= 3 ? (double)qp[2] : 0.0;
      const SceneView sc = j.scene;  // wave-uniform: kept in scalar registers
      ambiguous = collect_packed<K, STATS, DEPTH>(sc, sc.root, sc.end, qx, qy, qz, j.radius, k, list, st,
                                           stack + threadIdx.x, 256);
#pragma unroll
      for (int i = 0; i < K; ++i)
        if (i >= K - (int)k && list.e[i] != ~0ull) ++c;
    }
    const uint64_t at0 = wave_reserve(j.counters, c);
    if (active && c) {
      if (at0 + c > j.row_capacity) {
        atomicOr(&j.counters->status, kStatusRowOverflow);
      } else {
        uint32_t r = 0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
          if (i >= K - (int)k && list.e[i] != ~0ull) {
            const uint32_t slot = (uint32_t)list.e[i];
            const float4 p = load_prim(j.scene.prims, slot);
            const double dist = sqrt(point_d2(j.scene, slot, p, qx, qy, qz));
            write_neighbor_row(j, at0 + r, qi, __float_as_uint(p.w), dist, r + 1);
            ++r;
          }
        }
        if (ambiguous) {
          amb[atomicAdd(amb_count, 1u)] = make_uint4(w.job, qi, (uint32_t)at0, c);
          atomicAdd(&j.counters->ambiguous, 1u);
        }
      }
    }
    if (STATS) close_item(st);
  }
  if (STATS) flush_lane_stats(stats, st);
}

// Computes the rows of the neighbour queries that the list kernels flag as
// ambiguous, or whose candidates do not fit the parking space of the
// fixed-radius pipeline, with an exact sorted (FP64 distance, id) list per
// lane, and writes them at the rows reserved for them.  The list lives in LDS (slot-major, so the 64 lanes of a
// wavefront hit consecutive words): a 64-entry FP64 list does not fit a
// lane's registers, and these queries are rare.
constexpr int kFixBlock = 64;

__global__ __launch_bounds__(kFixBlock) void batch_fix_kernel(const DeviceJob* __restrict__ jobs,
                                                             const uint4* __restrict__ amb,
                                                             const uint32_t* amb_count) {
  __shared__ double list_d[kMaxListK * kFixBlock];
  __shared__ uint32_t list_id[kMaxListK * kFixBlock];
  const uint32_t lane = threadIdx.x;
  double* ld = list_d + lane;       // entry e at ld[e * kFixBlock]
  uint32_t* li = list_id + lane;
  const uint32_t n = *amb_count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint4 a = amb[i];
    const DeviceJob& j = jobs[a.x];
    const uint32_t k = j.k;  // 1 <= k <= kMaxListK, checked on the host
    const float* qp = j.queries + (size_t)a.y * j.qstride;
    const double qx = qp[0], qy = qp[1], qz = j.qstride >= 3 ? (double)qp[2] : 0.0;
    const float qf[3] = {(float)qx, (float)qy, (float)qz};
    const float m = coord_bound(scene_m(j.scene), qx, qy, qz);
    const double radius_sq = j.radius * j.radius;  // reference.py:357
    // A KNN query arrives with the k rows of its packed walk already written:
    // k real points within their largest distance bound the k-th nearest
    // distance, so the exact walk starts pruned to that ball.  (Fixed-radius
    // queries arrive with their rows unwritten.)
    double bound = j.radius;
    if (j.kind != kJobFixedRadius && a.w == k) {
      double dmax = 0.0;
      for (uint32_t r = 0; r < a.w; ++r) {
        double d;
        uint32_t id;
        read_neighbor_row(j, (uint64_t)a.z + r, d, id);
        dmax = fmax(dmax, d);
      }
      bound = fmin(bound, dmax);
    }
    float thr2 = frn_threshold2(bound, m);
    uint32_t count = 0;  // entries [0, count) sorted by (distance, id)
    LaneStats st;
    if (j.scene.root == kEndNode) continue;
    auto visit = [&](uint32_t slot, const float4& p, float) {
      const double d2 = point_d2(j.scene, slot, p, qx, qy, qz);
      if (d2 > radius_sq) return;  // reference.py:361 inclusive radius
      const double dist = sqrt(d2);
      const uint32_t ident = __float_as_uint(p.w);
      if (count == k && !key_less(dist, ident, ld[(k - 1) * kFixBlock], li[(k - 1) * kFixBlock]))
        return;
      uint32_t pos = count < k ? count++ : k - 1;
      while (pos > 0 &&
             key_less(dist, ident, ld[(pos - 1) * kFixBlock], li[(pos - 1) * kFixBlock])) {
        ld[pos * kFixBlock] = ld[(pos - 1) * kFixBlock];
        li[pos * kFixBlock] = li[(pos - 1) * kFixBlock];
        --pos;
      }
      ld[pos * kFixBlock] = dist;
      li[pos * kFixBlock] = ident;
      if (count == k) thr2 = frn_threshold2(fmin(j.radius, ld[(k - 1) * kFixBlock]), m);
    };
    walk_neighbors<false>(j.scene, j.scene.root, j.scene.end, qf, j.radius, thr2, st, visit);
    const uint32_t rows = min(count, a.w);
    for (uint32_t r = 0; r < rows; ++r)
      write_neighbor_row(j, (uint64_t)a.z + r, a.y, li[r * kFixBlock], ld[r * kFixBlock], r + 1);
  }
}

// ----------------------------------------------------------- fixed radius --
//
// Fixed-radius jobs run without a top-k list, in three launches.
// nbr_count_kernel walks each query's ball, counting the neighbours inside it
// with a histogram of their distances over kNbrBuckets equal slices of
// [0, radius].  Once k neighbours lie in buckets 0..b the walk shrinks to
// bucket b's edge (farther points cannot be among the k nearest); when bucket
// b still holds more candidates than the parking takes, the walk counts again
// over kNbrBuckets slices of bucket b's ball.  A query with c <= k neighbours
// reserves c rows, which nbr_emit_kernel writes in walk order and
// nbr_sort_kernel sorts into the (distance, id) order (reference.py:362-366).
// A query with more keeps its k nearest: the first bucket b whose cumulative
// count reaches k holds the k-th distance, the emit walk parks only the
// neighbours of buckets 0..b (at most kNbrCandMax) and the sort writes the
// first k of them.  Every decision uses the FP64 distance in the reference's
// operation order and equal distances share a bucket, so the rows are exact
// by construction.  A query whose buckets 0..b still hold more than
// kNbrCandMax neighbours after four counts (or that finds the parking space
// full) is computed by batch_fix_kernel.
// The count walk also stages the slots it counts, up to `stage_slots` of them
// per query (a multiple of 4, at most kNbrStageMax, chosen per batch by its
// memory).  When the last count round stays within that, the staged slots
// are a superset of the rows (all of them for a query with c <= k, those in
// buckets 0..b otherwise): the emit walk skips the query and nbr_sort_kernel
// recomputes the FP64 distances of the staged slots, filters and sorts them.
constexpr uint32_t kNbrBuckets = 32;
constexpr uint32_t kNbrCandMax = 128;
constexpr uint32_t kNbrStageMax = 256;
constexpr int kSortBlock = 64;
// histogram counters are 16-bit, two per LDS word; a query that meets more
// neighbours than one counter holds goes to the exact fix-up
constexpr uint32_t kNbrCountMax = 0xFFFFu;

// Distance bucket in [0, kNbrBuckets): dist * scale is monotone in dist.
__device__ inline uint32_t nbr_bucket(double dist, double scale) {
  const double x = dist * scale;
  return x < (double)kNbrBuckets ? (x >= 0.0 ? (uint32_t)x : 0u) : kNbrBuckets - 1;
}

// Distance bound of buckets 0..b: bucket(dist) <= b means dist * scale < b + 1,
// so dist < (b + 1) / scale; the factor covers the rounding of the quotient.
__device__ inline double nbr_bucket_edge(uint32_t b, double scale) {
  return (double)(b + 1) / scale * 1.0000000000009094947;  // 1 + 2^-40
}

__device__ inline void load_query3(const DeviceJob& j, uint32_t qi, double& qx, double& qy, double& qz) {
  const float* qp = j.queries + (size_t)qi * j.qstride;
  qx = qp[0];
  qy = qp[1];
  qz = j.qstride >= 3 ? (double)qp[2] : 0.0;
}

// Result of the count walks of one query.
struct NbrCount {
  double radius;     // final ball
  double radius_sq;  // membership: squared distance <= radius_sq
  uint32_t count;    // neighbours counted in buckets 0..top
  uint32_t bucket;   // more than k: the bucket holding the k-th distance
  uint32_t n_cand;   // more than k: neighbours in buckets 0..bucket
  uint32_t n_staged; // slots in the stage (a superset of the rows) when staged
  bool heavy;        // more than k neighbours: keep the k nearest
  bool refine;       // leave the query to the exact fix-up
  bool staged;       // the last round's stage holds every row's slot
};

// A query's stage cursor once its stage could not keep every row: further
// slots are dropped and the emit walk writes the rows instead.
constexpr uint32_t kStageClosed = 0x80000000u;

// Stage entries of scenes with at most 2^27 primitives carry the slot's
// distance bucket in their top 5 bits, so compaction and the sort filter
// without fetching the point again.
constexpr uint32_t kStageSlotBits = 27;
constexpr uint32_t kStageSlotMask = (1u << kStageSlotBits) - 1u;
static_assert(kNbrBuckets <= (1u << (32 - kStageSlotBits)), "bucket must fit the stage entry");

__device__ inline bool stage_packs(const SceneView& sc) { return sc.n_prims <= (1u << kStageSlotBits); }

// The count walk with the FP64 visit taken out of the lanes' walks.  Each
// lane walks its query's ball (one node step per lane per iteration, two when
// it descends: the child that follows its parent in preorder is fetched with
// it) and only prefilters leaf points in FP32; the points that pass go into a per-wavefront
// queue in LDS as (owner lane, slot, FP32 coordinates).  Whenever 64 are queued, or no lane is
// still walking, the wavefront drains 64 of them at once with every lane
// active: FP64 squared distance in the reference's order, the inclusive
// radius test, the distance bucket, the owner's histogram (LDS atomics) and
// stage (an LDS atomic cursor gives the position).  A divergent walk thus no
// longer runs the FP64 visit a few lanes at a time.  After each drain every
// owner whose count reached its checkpoint shrinks its ball to the k-th
// bucket's edge, as the walk itself did before.  Counting rounds are
// wavefront-synchronous: lanes that must recount over finer buckets do so
// together in the next round.  Every decision still rests on the FP64
// distance of each point, and the set of points counted in buckets up to the
// final cutoff does not depend on the order they are drained in, so the rows
// are exact as before.
// Leaves reach the prefilter through a second per-wavefront queue: a lane
// that meets a leaf only queues (owner lane, leaf), and whenever 64 leaves are
// queued, or no lane is still walking, the wavefront takes 64 of them, one per
// lane, and prefilters their points against the owners' current thresholds.
// Lanes meet leaves in different steps, so prefiltering in the walk itself
// would run the leaf loop nearly every step for a few lanes.  The owners'
// thresholds only shrink, so filtering later keeps every point that can still
// be counted.
// Point drains run after each point index of the leaves being prefiltered, so
// the point queue never holds more than 63 + 64 entries.
// A drain adds at most 64 slots to one stage, so after each drain a stage
// holding more than stage_slots - 64 is compacted in place by the wavefront:
// only slots whose bucket is within the owner's current cutoff (top) stay,
// as the others can no longer be rows (top only decreases within a round).
// A stage that compaction cannot bring under that mark is closed.
constexpr uint32_t kNbrQueueCap = 128;
constexpr uint32_t kNbrLeafQueueCap = 128;

__device__ inline double shfl_f64(double x, uint32_t src) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __shfl((uint32_t)b, (int)src, 64);
  const uint32_t hi = __shfl((uint32_t)(b >> 32), (int)src, 64);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Lane q's value (q wave-uniform).
__device__ inline uint64_t readlane64(uint64_t x, uint32_t q) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, q);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), q);
  return ((uint64_t)hi << 32) | lo;
}

__device__ inline double readlane_f64(double x, uint32_t q) {
  return __builtin_bit_cast(double, readlane64(__builtin_bit_cast(uint64_t, x), q));
}

__device__ inline void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

__device__ inline uint64_t shfl_xor64(uint64_t x, uint32_t m) {
  const uint32_t lo = __shfl_xor((uint32_t)x, (int)m, 64);
  const uint32_t hi = __shfl_xor((uint32_t)(x >> 32), (int)m, 64);
  return ((uint64_t)hi << 32) | lo;
}

// (distance bits, id) order; distances are non-negative, so their bit
// patterns order like the values.
__device__ inline bool dist_id_less(uint64_t da, uint32_t ia, uint64_t db, uint32_t ib) {
  return da < db || (da == db && ia < ib);
}

// One bitonic compare-exchange of element `idx` (held by this lane) with
// element idx ^ stride (held by lane ^ stride).
__device__ inline void bitonic_lanes(uint64_t& d, uint32_t& id, uint32_t idx, uint32_t size,
                                     uint32_t stride) {
  const uint64_t od = shfl_xor64(d, stride);
  const uint32_t oi = __shfl_xor(id, (int)stride, 64);
  const bool keep_min = ((idx & stride) == 0) == ((idx & size) == 0);
  const bool take = keep_min ? dist_id_less(od, oi, d, id) : dist_id_less(d, id, od, oi);
  d = take ? od : d;
  id = take ? oi : id;
}

// Sorts n <= 128 (distance bits, id) pairs held two per lane (element lane in
// (d0, i0), element lane + 64 in (d1, i1); pads are ~0) with a bitonic
// network over the next power of two >= n.  Every lane must take part.
__device__ inline void wave_sort_dist_id(uint64_t& d0, uint32_t& i0, uint64_t& d1, uint32_t& i1,
                                         uint32_t n) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t size_max = 2;
  while (size_max < n) size_max <<= 1;
  for (uint32_t size = 2; size <= size_max; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride == 64) {  // element lane vs lane + 64, one ascending block
        const bool sw = dist_id_less(d1, i1, d0, i0);
        const uint64_t td = d0;
        const uint32_t ti = i0;
        d0 = sw ? d1 : d0;
        i0 = sw ? i1 : i0;
        d1 = sw ? td : d1;
        i1 = sw ? ti : i1;
      } else {
        bitonic_lanes(d0, i0, lane, size, stride);
        if (size_max > 64) bitonic_lanes(d1, i1, lane + 64, size, stride);
      }
    }
  }
}

// Writes rows 1..n_rows of a query from n <= kNbrCandMax (distance bits, id)
// pairs in (distance, id) order.  Element `lane` is in (d0, i0), element
// lane + 64 in (d1, i1) (pads ~0).  p_lo, p_hi and p_id are three LDS planes
// of kNbrCandMax words of this wavefront; `in_planes` says the n pairs are
// already there (element x at index x).  Every lane must take part.
__device__ inline void nbr_write_sorted(const DeviceJob& j, uint64_t row_at, uint32_t qi, uint32_t n_rows,
                                        uint32_t n, uint64_t d0, uint32_t i0, uint64_t d1, uint32_t i1,
                                        uint32_t* p_lo, uint32_t* p_hi, uint32_t* p_id, bool in_planes) {
  const uint32_t lane = threadIdx.x & 63;
  if (n <= 64) {
    // rank sort: element `lane` goes to the number of elements before it in
    // (distance, id) order (the pairs are distinct: ids are), found by reading
    // every element from LDS as a broadcast
    if (!in_planes && lane < n) {
      p_lo[lane] = (uint32_t)d0;
      p_hi[lane] = (uint32_t)(d0 >> 32);
      p_id[lane] = i0;
    }
    wave_fence();
    uint32_t rank = 0;
#pragma unroll 4
    for (uint32_t x = 0; x < n; ++x) {
      const uint64_t dx = ((uint64_t)p_hi[x] << 32) | p_lo[x];
      rank += dist_id_less(dx, p_id[x], d0, i0) ? 1u : 0u;
    }
    wave_fence();  // every lane has read the planes before they are written again
    if (lane < n && rank < n_rows)
      write_neighbor_row(j, row_at + rank, qi, i0, __builtin_bit_cast(double, d0), rank + 1);
  } else {
    wave_sort_dist_id(d0, i0, d1, i1, n);
    if (lane < n_rows)
      write_neighbor_row(j, row_at + lane, qi, i0, __builtin_bit_cast(double, d0), lane + 1);
  }
}

// Sorts one staged query's rows: recomputes the FP64 distances of the slots
// the count walk staged, keeps buckets 0..bucket (all of them for a query
// with at most k neighbours, n_cand == 0) and writes the first n_rows in
// (distance, id) order.  The buckets order the rows already (a bucket is a
// monotone function of the distance), and the owner's histogram of its last
// counting round holds how many kept rows each bucket has: each kept row is
// placed in its bucket's region of the LDS planes and then ranked against
// the rows of that region only.  hist_q is the owner's histogram column;
// p_lo, p_hi and p_id are three planes of kNbrCandMax words, p_b one of
// kNbrCandMax bytes and b_cnt kNbrBuckets words, all of this wavefront.
// Every lane must take part.
__device__ inline void nbr_sort_staged(const DeviceJob& j, const uint32_t* sq, uint32_t n_staged, uint32_t n_rows,
                                       uint32_t n_cand, uint32_t bucket, double scale, double qx, double qy,
                                       double qz, uint64_t row_at, uint32_t qi, const uint32_t* hist_q,
                                       uint32_t* p_lo, uint32_t* p_hi, uint32_t* p_id, uint8_t* p_b,
                                       uint32_t* b_cnt) {
  const uint32_t lane = threadIdx.x & 63;
  const bool packed = stage_packs(j.scene);
  const uint32_t last = n_cand ? bucket : kNbrBuckets - 1;  // the last bucket kept
  // lane c < kNbrBuckets: rows in bucket c and the region's first position
  uint32_t cnt_c = 0;
  if (lane <= last) cnt_c = (hist_q[256 * (lane >> 1)] >> (16 * (lane & 1))) & 0xFFFFu;  // last < kNbrBuckets
  uint32_t total;
  const uint32_t off_c = wave_exclusive_sum(cnt_c, &total);
  if (lane < kNbrBuckets) b_cnt[lane] = 0u;
  wave_fence();
  uint32_t kept = 0;
  for (uint32_t e0 = 0; e0 < n_staged; e0 += 64) {
    const uint32_t e = e0 + lane;
    bool keep = false;
    uint64_t d = 0;
    uint32_t id = 0, b = 0;
    if (e < n_staged) {
      const uint32_t v = sq[e];
      const uint32_t slot = packed ? v & kStageSlotMask : v;
      // a packed bucket past the last one kept is dropped without a fetch
      if (!packed || (v >> kStageSlotBits) <= last) {
        const float4 p = load_prim(j.scene.prims, slot);
        const double dist = sqrt(point_d2(j.scene, slot, p, qx, qy, qz));
        b = nbr_bucket(dist, scale);
        keep = b <= last;
        d = __builtin_bit_cast(uint64_t, dist);
        id = __float_as_uint(p.w);
      }
    }
    const uint32_t base = __shfl(off_c, (int)(keep ? b : 0u), 64);
    if (keep) {
      const uint32_t pos = base + __hip_atomic_fetch_add(&b_cnt[b], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (pos < kNbrCandMax) {
        p_lo[pos] = (uint32_t)d;
        p_hi[pos] = (uint32_t)(d >> 32);
        p_id[pos] = id;
        p_b[pos] = (uint8_t)b;
      }
    }
    kept += (uint32_t)__popcll(__ballot(keep));
  }
  wave_fence();
  // the kept rows are exactly the points counted in buckets 0..last
  if ((kept > kNbrCandMax || kept != total) && lane == 0) atomicOr(&j.counters->status, kStatusInternal);
  const uint32_t n = kept < kNbrCandMax ? kept : kNbrCandMax;
  for (uint32_t h = 0; h < n; h += 64) {
    const uint32_t x = h + lane;
    const bool has = x < n;
    uint64_t d = 0;
    uint32_t id = 0, b = 0;
    if (has) {
      d = ((uint64_t)p_hi[x] << 32) | p_lo[x];
      id = p_id[x];
      b = p_b[x];
    }
    // every lane takes part in both shuffles (a shuffle reads only active lanes)
    const uint32_t start = __shfl(off_c, (int)b, 64), len_b = __shfl(cnt_c, (int)b, 64);
    // (a region never reaches past the planes: start + len_b <= total, and
    // total <= kNbrCandMax unless the status flag above is set)
    const uint32_t len = has ? min(len_b, kNbrCandMax - min(start, kNbrCandMax)) : 0u;
    // rank within the bucket's region (the pairs are distinct: ids are)
    uint32_t rank = 0;
    for (uint32_t t = 0; __ballot(t < len); ++t) {
      if (t < len) {
        const uint32_t y = start + t;
        rank += dist_id_less(((uint64_t)p_hi[y] << 32) | p_lo[y], p_id[y], d, id) ? 1u : 0u;
      }
    }
    const uint32_t pos = start + rank;
    if (has && pos < n_rows) write_neighbor_row(j, row_at + pos, qi, id, __builtin_bit_cast(double, d), pos + 1);
  }
  wave_fence();  // the planes are free again once every lane has read them
}

template <bool STATS>
__global__ __launch_bounds__(256, 5) void nbr_count_kernel(
    const DeviceJob* __restrict__ jobs, const WorkItem* __restrict__ items, uint32_t n_items,
    uint32_t* item_counter, NbrPlan* __restrict__ plans, unsigned long long* cand_count,
    uint64_t cand_cap, uint32_t* __restrict__ stage, uint32_t stage_slots, uint32_t* __restrict__ emit_list,
    uint32_t* emit_count, uint4* amb, uint32_t* amb_count, unsigned long long* stats) {
  // buckets 2w and 2w + 1 of thread t in the low and high half of hist[w * 256 + t]
  __shared__ uint32_t hist[kNbrBuckets / 2 * 256];
  __shared__ uint32_t q_slot[4 * kNbrQueueCap];
  __shared__ __align__(16) uint8_t q_owner[4 * kNbrQueueCap];
  // the queued points' FP32 coordinates, so that the drain does not fetch the
  // point again (an FP64 scene's drain still reads its exact coordinates)
  __shared__ float q_px[4 * kNbrQueueCap], q_py[4 * kNbrQueueCap], q_pz[4 * kNbrQueueCap];
  // queued leaves: the leaf's link word (count and first slot) and its owner
  __shared__ uint32_t lq_link[4 * kNbrLeafQueueCap];
  __shared__ __align__(16) uint8_t lq_owner[4 * kNbrLeafQueueCap];
  // per thread: neighbours counted this round (high word) and stage cursor
  // (low word), updated together by one LDS atomic
  __shared__ unsigned long long l_cs[256];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = threadIdx.x & ~63u;
  uint32_t* my_hist = hist + threadIdx.x;
  uint32_t* qs = q_slot + (threadIdx.x >> 6) * kNbrQueueCap;
  uint8_t* qo = q_owner + (threadIdx.x >> 6) * kNbrQueueCap;
  float* const qpx = q_px + (threadIdx.x >> 6) * kNbrQueueCap;
  float* const qpy = q_py + (threadIdx.x >> 6) * kNbrQueueCap;
  float* const qpz = q_pz + (threadIdx.x >> 6) * kNbrQueueCap;
  uint32_t* lql = lq_link + (threadIdx.x >> 6) * kNbrLeafQueueCap;
  uint8_t* lqo = lq_owner + (threadIdx.x >> 6) * kNbrLeafQueueCap;
  auto bucket_count = [&](uint32_t b) { return (my_hist[256 * (b >> 1)] >> (16 * (b & 1))) & 0xFFFFu; };
  LaneStats st;
  ItemQueue queue(item_counter, n_items);
  for (;;) {
    const uint32_t it = queue.next();
    if (it == kItemDone) break;
    if (it == kItemRetry) continue;
    const WorkItem w = items[it];
    const DeviceJob& j = jobs[w.job];
    const SceneView sc = j.scene;  // wave-uniform: kept in scalar registers
    const bool active = lane < w.count;
    const uint32_t k = j.k;
    // a lane checks for a k-th bucket each time it has counted another
    // max(k/8, 4) neighbours: frequent checks shrink the ball early, and a
    // check reads the histogram eight buckets per LDS round trip
    const uint32_t step = k / 8 > 4u ? k / 8 : 4u;
    uint32_t qi = 0;
    double qx = 0.0, qy = 0.0, qz = 0.0;
    if (active) {
      qi = job_query(j, w.first + lane);
      load_query3(j, qi, qx, qy, qz);
    }
    // queries are FP32, so qf holds them exactly: the FP64 coordinates are
    // (double)qf and are not kept in registers
    const float qf[3] = {(float)qx, (float)qy, (float)qz};
    const float m = coord_bound(scene_m(sc), qx, qy, qz);
    const bool pack = stage_packs(sc);
    uint32_t* const stage_it = stage + (size_t)it * 64 * stage_slots;
    // the walk reads the scene's compact nodes when its frame is usable
    const bool compact = sc.qnodes != nullptr && sc.qframe[6] == 0.0f;
    float q_o[3] = {0.0f, 0.0f, 0.0f}, q_s[3] = {0.0f, 0.0f, 0.0f};
    if (compact) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        q_o[a] = sc.qframe[a];
        q_s[a] = sc.qframe[3 + a];
      }
    }
    NbrCount r{j.radius, j.radius * j.radius, 0u, kNbrBuckets - 1, 0u, false, false, false};  // reference.py:357
    bool going = active;  // counts in the current round
    double scale = 0.0;  // bucket scale of the lane's last counting round
    for (uint32_t round = 0; round < 4 && __ballot(going); ++round) {
      scale = (double)kNbrBuckets / r.radius;
      float thr2 = frn_threshold2(r.radius, m);
      uint32_t top = kNbrBuckets - 1, next_check = k;
      bool shrunk = false;
      if (going) {
        for (uint32_t w2 = 0; w2 < kNbrBuckets / 2; ++w2) my_hist[256 * w2] = 0u;
        l_cs[threadIdx.x] = 0ull;
      }
      uint32_t node = sc.root;
      bool walking = going && sc.root != kEndNode;
      uint32_t guard = sc.n_nodes + 1;  // a preorder walk visits each node at most once
      uint32_t head = 0, size = 0;      // the wavefront's point queue (uniform)
      uint32_t lhead = 0, lsize = 0;    // the wavefront's leaf queue (uniform)
      for (;;) {
        if (STATS && lane == 0) ++st.steps;
        uint32_t leaf_link = 0;  // the leaf this lane met in this step, if any
        if (walking) {
          // the node and the one after it in preorder, fetched together: a
          // descent goes there next, and its test then costs no second
          // round trip to memory
          const uint32_t after = node + 1 < sc.n_nodes ? node + 1 : node;
          QNode nd, nd1;
          if (compact) {
            nd = load_qnode(sc.qnodes, node, q_o, q_s);
            nd1 = load_qnode(sc.qnodes, after, q_o, q_s);
          } else {
            nd = qnode_of(load_node(sc.nodes, node), sc.n_nodes);
            nd1 = qnode_of(load_node(sc.nodes, after), sc.n_nodes);
          }
          if (STATS) ++st.nodes;
          const bool in = box_dist2(qf, nd.lo, nd.hi) <= thr2;
          const bool leaf = node_is_leaf(nd.word);
          if (in && leaf) leaf_link = nd.word;
          // into the left child (the next record) or past a leaf (also the
          // next record), else to the escape
          uint32_t next = (in || leaf) ? node + 1 : nd.word;
          walking = next != sc.end && next < sc.n_nodes && --guard != 0;
          if (walking && in && !leaf) {  // descended: test the child (fetched with its parent) now
            if (STATS) ++st.nodes;
            const bool in1 = box_dist2(qf, nd1.lo, nd1.hi) <= thr2;
            const bool leaf1 = node_is_leaf(nd1.word);
            if (in1 && leaf1) leaf_link = nd1.word;
            next = (in1 || leaf1) ? after + 1 : nd1.word;
            walking = next != sc.end && next < sc.n_nodes && --guard != 0;
          }
          node = next;
        }
        // One drain: the first min(64, size) queued points, one per lane, then
        // the owners' checkpoints (shrink to the k-th bucket's edge once k
        // neighbours are counted).
        auto drain = [&]() {
          const uint32_t n = size < 64 ? size : 64;
          wave_fence();
          const bool has = lane < n;
          const uint32_t e = (head + (has ? lane : 0u)) & (kNbrQueueCap - 1);
          const uint32_t slot = qs[e];
          const uint32_t o = has ? qo[e] : lane;
          // the owner's query and ball (every lane takes part in the shuffles)
          const double ox = __shfl(qf[0], (int)o, 64), oy = __shfl(qf[1], (int)o, 64);
          const double oz = __shfl(qf[2], (int)o, 64);
          const double orsq = shfl_f64(r.radius_sq, o), oscale = shfl_f64(scale, o);
          const uint32_t otop = __shfl(top, (int)o, 64);
          if (has) {
            const float4 p = make_float4(qpx[e], qpy[e], qpz[e], 0.f);
            const double d2 = point_d2(sc, slot, p, ox, oy, oz);
            if (d2 <= orsq) {  // reference.py:361 inclusive radius
              const uint32_t b = nbr_bucket(sqrt(d2), oscale);
              if (b <= otop) {  // not beyond the owner's k nearest
                __hip_atomic_fetch_add(&hist[256 * (b >> 1) + wbase + o], 1u << (16 * (b & 1)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint32_t at = (uint32_t)__hip_atomic_fetch_add(&l_cs[wbase + o], (1ull << 32) | 1ull,
                                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (at < stage_slots) stage_it[(size_t)o * stage_slots + at] = pack ? slot | (b << kStageSlotBits) : slot;
              }
            }
          }
          head = (head + n) & (kNbrQueueCap - 1);
          size -= n;
          wave_fence();
          if (going) {
            const uint32_t count = (uint32_t)(l_cs[threadIdx.x] >> 32);
            if (count >= next_check) {
              next_check = count + step;
              // the first bucket c < top whose cumulative count reaches k;
              // the histogram words are read four at a time (eight buckets),
              // so their LDS latencies overlap instead of adding up
              uint32_t cum = 0, kth = top;
              for (uint32_t w0 = 0; w0 < kNbrBuckets / 2 && 2 * w0 < top && kth == top; w0 += 4) {
                uint32_t wv[4];
#pragma unroll
                for (uint32_t i = 0; i < 4; ++i) wv[i] = my_hist[256 * (w0 + i)];
#pragma unroll
                for (uint32_t i = 0; i < 8; ++i) {
                  const uint32_t c = 2 * w0 + i;
                  if (kth == top && c < top) {
                    cum += (wv[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
                    if (cum >= k) kth = c;
                  }
                }
              }
              if (kth < top) {  // the k nearest lie in buckets 0..kth: walk no farther
                top = kth;
                shrunk = true;
                thr2 = frn_threshold2(fmin(r.radius, nbr_bucket_edge(kth, scale)), m);
              }
            }
          }
          // compaction of the stages past the mark, one owner at a time
          const uint32_t cur = (uint32_t)l_cs[threadIdx.x];
          uint64_t full = __ballot(going && cur > stage_slots - 64 && cur < kStageClosed);
          while (full) {
            const uint32_t o = (uint32_t)__builtin_ctzll(full);
            full &= full - 1;
            const double ox = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(qf[0]), o));
            const double oy = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(qf[1]), o));
            const double oz = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(qf[2]), o));
            const double oscale = readlane_f64(scale, o);
            const uint32_t otop = __builtin_amdgcn_readlane(top, o);
            const uint32_t n_st = min(__builtin_amdgcn_readlane(cur, o), stage_slots);
            uint32_t* const base = stage_it + (size_t)o * stage_slots;
            uint32_t kept = 0;
            for (uint32_t e0 = 0; e0 < n_st; e0 += 64) {
              const uint32_t e = e0 + lane;
              bool keep = false;
              uint32_t v = 0;
              if (e < n_st) {
                v = base[e];
                if (pack) {
                  keep = (v >> kStageSlotBits) <= otop;
                } else {
                  const float4 p = load_prim(sc.prims, v);
                  keep = nbr_bucket(sqrt(point_d2(sc, v, p, ox, oy, oz)), oscale) <= otop;
                }
              }
              // every lane has read its slot before any is rewritten (the
              // ballot needs them all); later chunks lie past every write
              const uint64_t km = __ballot(keep);
              if (keep) base[kept + (uint32_t)__popcll(km & ((1ull << lane) - 1ull))] = v;
              kept += (uint32_t)__popcll(km);
            }
            if (lane == o) {
              const unsigned long long v = l_cs[threadIdx.x];
              l_cs[threadIdx.x] = (v & ~0xFFFFFFFFull) | (kept > stage_slots - 64 ? kStageClosed : kept);
            }
          }
          wave_fence();
        };
        // One leaf drain: the first min(64, lsize) queued leaves, one per
        // lane; their points are prefiltered in FP32 against the owner's
        // current threshold and the survivors queued in (point index within
        // the leaf, queue position) order, with point drains between indices.
        auto leaf_drain = [&]() {
          const uint32_t n = lsize < 64 ? lsize : 64;
          wave_fence();
          const bool has = lane < n;
          const uint32_t e = (lhead + (has ? lane : 0u)) & (kNbrLeafQueueCap - 1);
          const uint32_t link = lql[e];
          const uint32_t o = has ? lqo[e] : lane;
          // the owner's query and threshold (every lane takes part in the shuffles)
          const float ox = __shfl(qf[0], (int)o, 64), oy = __shfl(qf[1], (int)o, 64);
          const float oz = __shfl(qf[2], (int)o, 64), othr2 = __shfl(thr2, (int)o, 64);
          const uint32_t first = leaf_first(link), cnt = has ? leaf_count(link) : 0u;
          lhead = (lhead + n) & (kNbrLeafQueueCap - 1);
          lsize -= n;
          for (uint32_t i = 0;; ++i) {
            const bool act = i < cnt;
            if (!__ballot(act)) break;
            bool push = false;
            const uint32_t slot = first + i;
            float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
            if (act) {
              p = load_prim(sc.prims, slot);
              if (STATS) ++st.prims;
              const float fx = p.x - ox, fy = p.y - oy, fz = p.z - oz;
              push = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx)) <= othr2;
            }
            const uint64_t pm = __ballot(push);
            if (push) {
              const uint32_t pos = (head + size + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull))) & (kNbrQueueCap - 1);
              qs[pos] = slot;
              qo[pos] = (uint8_t)o;
              qpx[pos] = p.x;
              qpy[pos] = p.y;
              qpz[pos] = p.z;
            }
            size += (uint32_t)__popcll(pm);
            while (size >= 64) drain();
          }
          wave_fence();
        };
        // the leaves met in this step join the leaf queue in lane order
        const uint64_t lm = __ballot(leaf_link != 0u);
        if (leaf_link != 0u) {
          const uint32_t pos = (lhead + lsize + (uint32_t)__popcll(lm & ((1ull << lane) - 1ull))) & (kNbrLeafQueueCap - 1);
          lql[pos] = leaf_link;
          lqo[pos] = (uint8_t)lane;
        }
        lsize += (uint32_t)__popcll(lm);
        while (lsize >= 64) leaf_drain();
        if (__ballot(walking) == 0) {  // the round's walks are over: empty both queues
          while (lsize != 0) leaf_drain();
          while (size != 0) drain();
          break;
        }
      }
      wave_fence();
      if (going) {
        const unsigned long long cs = l_cs[threadIdx.x];
        const uint32_t count = (uint32_t)(cs >> 32);
        going = false;
        r.count = count;
        r.heavy = shrunk || count > k;
        r.bucket = top;
        r.n_cand = 0;
        r.n_staged = (uint32_t)cs;
        r.staged = r.n_staged <= stage_slots;
        if (r.heavy && count > kNbrCountMax) {
          r.refine = true;
        } else if (r.heavy) {
          // the first bucket c <= top whose cumulative count reaches k (the
          // histogram read four words at a time, as at the checkpoints)
          uint32_t cum = 0;
          bool found = false;
          for (uint32_t w0 = 0; w0 < kNbrBuckets / 2 && 2 * w0 <= top && !found; w0 += 4) {
            uint32_t wv[4];
#pragma unroll
            for (uint32_t i = 0; i < 4; ++i) wv[i] = my_hist[256 * (w0 + i)];
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i) {
              const uint32_t c = 2 * w0 + i;
              if (!found && c <= top) {
                cum += (wv[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
                if (cum >= k) {
                  r.bucket = c;
                  r.n_cand = cum;
                  found = true;
                }
              }
            }
          }
          if (r.n_cand > kNbrCandMax) {
            if (round == 3) {
              r.refine = true;
            } else {
              // too coarse: count again with kNbrBuckets slices of bucket b's
              // ball, which holds the k nearest
              r.radius = fmin(r.radius, nbr_bucket_edge(r.bucket, scale));
              r.radius_sq = fmin(r.radius_sq, r.radius * r.radius);
              going = true;
            }
          }
        }
      }
    }
    uint32_t n_rows = r.heavy ? k : r.count;
    bool refine = r.refine;
    const uint64_t row_at = wave_reserve(j.counters, n_rows);
    if (n_rows && row_at + n_rows > j.row_capacity) {
      atomicOr(&j.counters->status, kStatusRowOverflow);
      n_rows = 0;
      refine = false;
    }
    const bool staged = n_rows && !refine && r.staged;
    const uint32_t want = (n_rows && r.heavy && !refine && !staged) ? r.n_cand : 0u;
    uint32_t total;
    const uint32_t prefix = wave_exclusive_sum(want, &total);
    const uint64_t cand_at = lane0_add64(cand_count, total) + prefix;
    if (want && cand_at + want > cand_cap) refine = true;
    if (n_rows && refine) {
      amb[atomicAdd(amb_count, 1u)] = make_uint4(w.job, qi, (uint32_t)row_at, n_rows);
      atomicAdd(&j.counters->ambiguous, 1u);
    }
    if (active)
      plans[(size_t)it * 64 + lane] = NbrPlan{row_at,   cand_at, r.radius, r.radius_sq, refine ? 0u : n_rows,
                                              r.heavy ? r.n_cand : 0u, r.bucket, qi, staged ? r.n_staged : 0u, 0u};
    // queries whose rows the emit walk writes (not staged) join the emit list
    const bool emit = n_rows && !refine && !staged;
    const uint64_t em = __ballot(emit);
    if (em) {
      const uint32_t base = lane0_add(emit_count, (uint32_t)__popcll(em));
      if (emit) emit_list[base + (uint32_t)__popcll(em & ((1ull << lane) - 1ull))] = it * 64 + lane;
    }
    // Staged queries are sorted here, one at a time by the whole wavefront,
    // while their stages and points are still in this XCD's L2 (the sort
    // kernel only takes the emit list).  Between items the wavefront's point
    // and leaf queues and its stage cursors are free: they hold the sort's
    // three LDS planes of kNbrCandMax words.
    uint64_t todo = __ballot(staged);
    if (todo) {
      wave_fence();
      uint32_t* const p_id = reinterpret_cast<uint32_t*>(l_cs + wbase);
      const uint32_t n_cand = r.heavy ? r.n_cand : 0u;
      while (todo) {
        const uint32_t q = (uint32_t)__builtin_ctzll(todo);
        todo &= todo - 1;
        const double ox = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(qf[0]), q));
        const double oy = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(qf[1]), q));
        const double oz = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(qf[2]), q));
        nbr_sort_staged(j, stage_it + (size_t)q * stage_slots, __builtin_amdgcn_readlane(r.n_staged, q),
                        __builtin_amdgcn_readlane(n_rows, q), __builtin_amdgcn_readlane(n_cand, q),
                        __builtin_amdgcn_readlane(r.bucket, q), readlane_f64(scale, q), ox, oy, oz,
                        readlane64(row_at, q), __builtin_amdgcn_readlane(qi, q), hist + wbase + q, qs, lql, p_id,
                        lqo, reinterpret_cast<uint32_t*>(qo));
      }
      wave_fence();
    }
  }
  if (STATS) {
    st.mark = st.nodes;  // steps were counted per iteration
    flush_lane_stats(stats, st);
  }
}

// Writes one query's rows (or parks its candidates) on a second walk of its
// ball; the walk meets the same candidates in the same FP64 arithmetic as the
// count.
__device__ inline void nbr_emit_query(const DeviceJob& j, const NbrPlan& pl, NbrCand* cands, LaneStats& st) {
  const SceneView sc = j.scene;
  double qx, qy, qz;
  load_query3(j, pl.qi, qx, qy, qz);
  const float qf[3] = {(float)qx, (float)qy, (float)qz};
  const double scale = (double)kNbrBuckets / pl.radius;
  const bool parked = pl.n_cand != 0;
  const double bound = parked ? fmin(pl.radius, nbr_bucket_edge(pl.bucket, scale)) : pl.radius;
  float thr2 = frn_threshold2(bound, coord_bound(scene_m(sc), qx, qy, qz));
  uint32_t r = 0;
  NbrCand* park = cands + pl.cand_at;
  walk_points<false>(sc, sc.root, sc.end, qf, thr2, st,
                     [&](uint32_t slot, const float4& p, float) {
                       const double d2 = point_d2(sc, slot, p, qx, qy, qz);
                       if (d2 > pl.radius_sq) return;
                       const double dist = sqrt(d2);
                       const uint32_t id = __float_as_uint(p.w);
                       if (!parked) {  // ranks of more than one row are set by the sort
                         if (r < pl.n_rows) write_neighbor_row(j, pl.row_at + r, pl.qi, id, dist, r + 1);
                         ++r;
                       } else if (nbr_bucket(dist, scale) <= pl.bucket) {
                         if (r < pl.n_cand) park[r] = NbrCand{dist, id, 0u};
                         ++r;
                       }
                     });
}

// Walks the queries of the emit list, 64 at a time (one per lane, any mix of
// jobs), so a few unstaged queries no longer hold a wavefront per work item.
__global__ __launch_bounds__(256) void nbr_emit_kernel(const DeviceJob* __restrict__ jobs,
                                                       const WorkItem* __restrict__ items,
                                                       const NbrPlan* __restrict__ plans,
                                                       NbrCand* __restrict__ cands,
                                                       const uint32_t* __restrict__ emit_list,
                                                       const uint32_t* emit_count, uint32_t* chunk_counter) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t n = __builtin_amdgcn_readfirstlane(*emit_count);
  LaneStats st;
  for (;;) {
    const uint32_t chunk = lane0_add(chunk_counter, 1u);
    if ((uint64_t)chunk * 64 >= n) break;
    const uint32_t e = chunk * 64 + lane;
    if (e < n) {
      const uint32_t p = emit_list[e];
      const NbrPlan pl = plans[p];
      nbr_emit_query(jobs[items[p >> 6].job], pl, cands, st);
    }
  }
}

// Sorts the rows of the queries of the emit list (the count walk sorts the
// staged ones itself), one query at a time per wavefront: parked candidates
// (the first n_rows of them in (distance, id) order are the rows) or rows
// the emit walk wrote in walk order.
__global__ __launch_bounds__(kSortBlock) void nbr_sort_kernel(const DeviceJob* __restrict__ jobs,
                                                             const WorkItem* __restrict__ items,
                                                             const NbrPlan* __restrict__ plans,
                                                             const NbrCand* __restrict__ cands,
                                                             const uint32_t* __restrict__ emit_list,
                                                             const uint32_t* emit_count, uint32_t* counter) {
  __shared__ uint32_t planes[3 * kNbrCandMax];
  const uint32_t lane = threadIdx.x;
  const uint32_t n_list = __builtin_amdgcn_readfirstlane(*emit_count);
  for (;;) {
    const uint32_t i = lane0_add(counter, 1u);
    if (i >= n_list) break;
    const uint32_t p = __builtin_amdgcn_readfirstlane(emit_list[i]);
    const NbrPlan pl = plans[p];
    const uint32_t n_rows = __builtin_amdgcn_readfirstlane(pl.n_rows);
    const uint32_t n_cand = __builtin_amdgcn_readfirstlane(pl.n_cand);
    if (n_rows == 0 || (n_cand == 0 && n_rows == 1)) continue;
    const DeviceJob& j = jobs[items[p >> 6].job];
    uint64_t d0 = ~0ull, d1 = ~0ull;
    uint32_t i0 = ~0u, i1 = ~0u;
    uint32_t n;
    if (n_cand == 0) {  // rows written in walk order, sorted in place
      n = n_rows;
      if (lane < n) {
        double dist;
        read_neighbor_row(j, pl.row_at + lane, dist, i0);
        d0 = __builtin_bit_cast(uint64_t, dist);
      }
    } else {  // parked candidates; the first n_rows of them are the rows
      n = n_cand;
      if (lane < n) {
        const NbrCand c = cands[pl.cand_at + lane];
        d0 = __builtin_bit_cast(uint64_t, c.distance);
        i0 = c.id;
      }
      if (lane + 64 < n) {
        const NbrCand c = cands[pl.cand_at + lane + 64];
        d1 = __builtin_bit_cast(uint64_t, c.distance);
        i1 = c.id;
      }
    }
    nbr_write_sorted(j, pl.row_at, pl.qi, n_rows, n, d0, i0, d1, i1, planes, planes + kNbrCandMax,
                     planes + 2 * kNbrCandMax, false);
  }
}

template <int MODE, bool STATS>
__global__ __launch_bounds__(256) void batch_rays_kernel(const DeviceJob* __restrict__ jobs,
                                                         const WorkItem* __restrict__ items,
                                                         uint32_t n_items, uint32_t* item_counter,
                                                         unsigned long long* stats) {
  const uint32_t lane = threadIdx.x & 63;
  LaneStats st;
  ItemQueue queue(item_counter, n_items);
  for (;;) {
    const uint32_t it = queue.next();
    if (it == kItemDone) break;
    if (it == kItemRetry) continue;
    const WorkItem w = items[it];
    const DeviceJob& j = jobs[w.job];
    const bool active = lane < w.count;
    RayResult res;
    uint32_t qi = 0;
    if (active) {
      qi = job_query(j, w.first + lane);
      const float* r = j.queries + (size_t)qi * j.qstride;
      const SceneView sc = j.scene;  // wave-uniform: kept in scalar registers
      res = trace_ray<MODE, STATS>(sc, sc.root, sc.end, r[0], r[1], r[2], r[3], r[4], r[5], r[6], st);
    }
    if (MODE != 2) {
      if (active)
        static_cast<PairRowRec*>(j.rows)[qi] = PairRowRec{qi, MODE == 0 ? (res.hits ? 1u : 0u) : res.hits};
      (void)lane0_add64(&j.counters->rows, w.count);
    } else {
      const uint32_t c = (active && res.best_t == res.best_t) ? 1u : 0u;
      const uint64_t at = wave_reserve(j.counters, c);
      if (c) {
        if (at >= j.row_capacity)
          atomicOr(&j.counters->status, kStatusRowOverflow);
        else
          static_cast<ClosestRowRec*>(j.rows)[at] = ClosestRowRec{qi, res.best_id, res.best_t};
      }
    }
    if (STATS) close_item(st);
  }
  if (STATS) flush_lane_stats(stats, st);
}

// PIP jobs with the polygon test taken out of the lanes' walks.  Each lane
// walks its point's rope preorder walk (two node records per step); the
// polygons of the leaves it reaches go into a per-wavefront queue in LDS as
// (owner lane, slot), and whenever 64 are queued, or no lane is still walking,
// the wavefront tests 64 of them at once with every lane active: the exact
// box prefilter of positive hits (reference.py:122-125), then the
// point-in-polygon test (reference.py:717-742).  Lanes reach leaves in
// different steps, so testing in the walk itself ran the long edge loop for a
// few lanes at a time.  A containing polygon goes to its owner's hit list in
// LDS (input index and id of the first kPipLdsHits, every hit counted).  At
// the end of the work item each lane writes its rows ordered by polygon input
// index.  A point inside more than kPipLdsHits polygons reserves all its rows
// and is queued for batch_pip_over_kernel, which walks it again and writes
// every hit, so the row set stays exact.
constexpr uint32_t kPipQueueCap = 128;
constexpr uint32_t kPipLdsHits = 16;

// One wavefront's share of the PIP kernel's LDS (dynamic shared memory, one
// per wavefront of the workgroup).
struct PipWaveLds {
  uint32_t q_slot[kPipQueueCap];
  uint32_t hit_count[64];
  uint32_t hit_input[kPipLdsHits * 64];  // entry i of lane l at [i * 64 + l]
  uint32_t hit_id[kPipLdsHits * 64];
  uint8_t q_owner[kPipQueueCap];
};

template <bool STATS>
__global__ __launch_bounds__(256, 6) void batch_pip_kernel(const DeviceJob* __restrict__ jobs,
                                                        const WorkItem* __restrict__ items,
                                                        uint32_t n_items, uint32_t* item_counter,
                                                        double eps_sq, uint4* over, uint32_t* over_count,
                                                        unsigned long long* stats) {
  extern __shared__ uint4 pip_lds[];
  PipWaveLds& L = reinterpret_cast<PipWaveLds*>(pip_lds)[threadIdx.x >> 6];
  const uint32_t lane = threadIdx.x & 63;
  LaneStats st;
  ItemQueue queue(item_counter, n_items);
  for (;;) {
    const uint32_t it = queue.next();
    if (it == kItemDone) break;
    if (it == kItemRetry) continue;
    const WorkItem w = items[it];
    const DeviceJob& j = jobs[w.job];
    const SceneView sc = j.scene;  // wave-uniform: kept in scalar registers
    const bool active = lane < w.count;
    uint32_t qi = 0;
    float pf0 = 0.0f, pf1 = 0.0f;  // queries are FP32: (double)pf is the point exactly
    if (active) {
      qi = job_query(j, w.first + lane);
      const float* p = j.queries + (size_t)qi * j.qstride;
      pf0 = p[0];
      pf1 = p[1];
    }
    const float m = coord_bound(scene_m(sc), pf0, pf1, 0.0);
    const float pad = m * 9.5367431640625e-07f + 1.0e-11f;
    L.hit_count[lane] = 0u;
    uint32_t node = sc.root;
    bool walking = active && sc.root != sc.end;
    uint32_t guard = sc.n_nodes + 1;  // a preorder walk visits each node at most once
    uint32_t head = 0, size = 0;      // the wavefront's queue (uniform)
    // One drain: the first min(64, size) queued (owner, polygon) pairs, one
    // per lane.
    auto drain = [&]() {
      const uint32_t n = size < 64 ? size : 64;
      wave_fence();
      const bool has = lane < n;
      const uint32_t e = (head + (has ? lane : 0u)) & (kPipQueueCap - 1);
      const uint32_t slot = L.q_slot[e];
      const uint32_t o = has ? L.q_owner[e] : lane;
      const double x = __shfl(pf0, (int)o, 64), y = __shfl(pf1, (int)o, 64);
      if (has) {
        const size_t bx = 4 * (size_t)slot;  // reference.py:122-125 exact bbox prefilter
        if (!(x < load_f64(sc.poly_box, bx) || x > load_f64(sc.poly_box, bx + 2) ||
              y < load_f64(sc.poly_box, bx + 1) || y > load_f64(sc.poly_box, bx + 3))) {
          const uint4 ref = load_u4(sc.poly, slot);
          if (point_in_polygon(x, y, sc.poly_verts + 2 * (size_t)ref.x, ref.y, eps_sq)) {
            const uint32_t h = __hip_atomic_fetch_add(&L.hit_count[o], 1u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_WORKGROUP);
            if (h < kPipLdsHits) {
              L.hit_input[h * 64 + o] = sc.slot_to_input[slot];
              L.hit_id[h * 64 + o] = ref.z;
            }
          }
        }
      }
      head = (head + n) & (kPipQueueCap - 1);
      size -= n;
      wave_fence();
    };
    for (;;) {
      uint32_t leaf_link = 0;
      if (walking) {
        BvhNode nd;
        bool in;
        rope_pair_step<STATS>(sc, node, nd, in, st, [&](const BvhNode& bb, uint32_t) {
          return (pf0 >= bb.lo[0] - pad) & (pf0 <= bb.hi[0] + pad) & (pf1 >= bb.lo[1] - pad) &
                 (pf1 <= bb.hi[1] + pad);
        });
        const bool leaf = node_is_leaf(nd.link);
        if (in && leaf) leaf_link = nd.link;
        // into an overlapped internal node's subtree, else past the node
        node = (in && !leaf) ? node + 1 : nd.escape;  // left child: the next record
        walking = node != sc.end && --guard != 0;
      }
      // the leaf's polygons join the queue in (index within the leaf, lane) order
      const uint32_t first = leaf_first(leaf_link), cnt = leaf_link ? leaf_count(leaf_link) : 0u;
      for (uint32_t e = 0;; ++e) {
        const bool act = e < cnt;
        const uint64_t pm = __ballot(act);
        if (!pm) break;
        if (act) {
          if (STATS) ++st.prims;
          const uint32_t pos = (head + size + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull))) & (kPipQueueCap - 1);
          L.q_slot[pos] = first + e;
          L.q_owner[pos] = (uint8_t)lane;
        }
        size += (uint32_t)__popcll(pm);
        while (size >= 64) drain();
      }
      if (__ballot(walking) == 0) {
        while (size != 0) drain();
        break;
      }
    }
    wave_fence();
    const uint32_t c = active ? L.hit_count[lane] : 0u;
    const uint64_t at0 = wave_reserve(j.counters, c);
    if (c) {
      if (at0 + c > j.row_capacity) {
        atomicOr(&j.counters->status, kStatusRowOverflow);
      } else if (c > kPipLdsHits) {
        over[atomicAdd(over_count, 1u)] = make_uint4(w.job, qi, (uint32_t)at0, c);
      } else {
        // rows in polygon input order: each hit goes to its rank among the
        // lane's hits (input indices are distinct)
        for (uint32_t h = 0; h < c; ++h) {
          const uint32_t in_h = L.hit_input[h * 64 + lane];
          uint32_t rank = 0;
          for (uint32_t g = 0; g < c; ++g) rank += L.hit_input[g * 64 + lane] < in_h ? 1u : 0u;
          static_cast<PipRowRec*>(j.rows)[at0 + rank] = PipRowRec{qi, L.hit_id[h * 64 + lane], 1u};
        }
      }
    }
    wave_fence();  // every lane has read its hits before the next item clears the counts
    if (STATS) close_item(st);
  }
  if (STATS) flush_lane_stats(stats, st);
}

// Writes every containing polygon of the queued points (see batch_pip_kernel).
__global__ __launch_bounds__(64) void batch_pip_over_kernel(const DeviceJob* __restrict__ jobs,
                                                            const uint4* __restrict__ over,
                                                            const uint32_t* over_count, double eps_sq) {
  // one wavefront per point: the walk is the same in every lane, and each
  // polygon met is tested with its edges spread over the lanes (points inside
  // many polygons lie where polygons are dense: long walks, many tests)
  const uint32_t n = *over_count;
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint4 o = over[i];
    const DeviceJob& j = jobs[o.x];
    const SceneView sc = j.scene;
    const float* p = j.queries + (size_t)o.y * j.qstride;
    const double x = p[0], y = p[1];
    PipRowRec* rows = static_cast<PipRowRec*>(j.rows) + o.z;
    uint32_t r = 0;
    const float pf0 = p[0], pf1 = p[1];
    const float m = coord_bound(scene_m(sc), x, y, 0.0);
    const float pad = m * 9.5367431640625e-07f + 1.0e-11f;  // as walk_pip
    LaneStats st;
    uint32_t node = sc.root;
    for (uint32_t guard = sc.n_nodes + 1; node != sc.end && guard; --guard) {
      BvhNode nd;
      bool in;
      rope_pair_step<false>(sc, node, nd, in, st, [&](const BvhNode& bb, uint32_t) {
        return (pf0 >= bb.lo[0] - pad) & (pf0 <= bb.hi[0] + pad) & (pf1 >= bb.lo[1] - pad) &
               (pf1 <= bb.hi[1] + pad);
      });
      const bool leaf = node_is_leaf(nd.link);
      if (!(in && leaf)) {
        node = (in && !leaf) ? node + 1 : nd.escape;  // left child: the next record
        continue;
      }
      const uint32_t first = leaf_first(nd.link), cnt = leaf_count(nd.link);
      for (uint32_t e = 0; e < cnt; ++e) {
        const uint32_t slot = first + e;
        const size_t bx = 4 * (size_t)slot;  // reference.py:122-125 exact bbox prefilter
        if (x < load_f64(sc.poly_box, bx) || x > load_f64(sc.poly_box, bx + 2) || y < load_f64(sc.poly_box, bx + 1) ||
            y > load_f64(sc.poly_box, bx + 3))
          continue;
        const uint4 ref = load_u4(sc.poly, slot);
        if (point_in_polygon_wave(x, y, sc.poly_verts + 2 * (size_t)ref.x, ref.y, eps_sq)) {
          if (lane == 0 && r < o.w) rows[r] = PipRowRec{o.y, ref.z, 1u};
          ++r;
        }
      }
      node = nd.escape;
    }
  }
}

__device__ inline uint32_t spread10b(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Item locality keys: family, scene index, then the Morton code of the
// item's first query (keys holds the sorted query keys, job << 32 | code).
__global__ void item_keys_kernel(const WorkItem* __restrict__ items, const uint64_t* __restrict__ starts,
                                 const uint64_t* __restrict__ keys, size_t n, int scene_bits,
                                 uint64_t* __restrict__ item_keys, uint32_t* __restrict__ item_vals) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const WorkItem w = items[i];
  const uint64_t code = keys[starts[w.job] + w.first] & 0x3FFFFFFFull;
  const uint64_t fam = w.group >> 24, scene = w.group & 0xFFFFFFu;
  item_keys[i] = (fam << (30 + scene_bits)) | (scene << 30) | code;
  item_vals[i] = (uint32_t)i;
}

__global__ void gather_items_kernel(const WorkItem* __restrict__ items, const uint32_t* __restrict__ idx,
                                    size_t n, WorkItem* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = items[idx[i]];
}

__global__ void batch_keys_kernel(const DeviceJob* __restrict__ jobs,
                                  const uint64_t* __restrict__ starts, uint32_t n_jobs,
                                  uint64_t total, uint64_t* __restrict__ keys,
                                  uint32_t* __restrict__ vals) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total) return;
  uint32_t lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (starts[mid] <= p) lo = mid; else hi = mid;
  }
  const DeviceJob& j = jobs[lo];
  const uint32_t local = (uint32_t)(p - starts[lo]);
  const float* q = j.queries + (size_t)local * j.qstride;
  const float c[3] = {q[0], q[1], j.qstride >= 3 ? q[2] : 0.0f};
  uint32_t code = 0;
  if (j.scene.root != kEndNode) {
    const BvhNode r = load_node(j.scene.nodes, j.scene.root);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float e = r.hi[a] - r.lo[a];
      float f = e > 0.0f ? (c[a] - r.lo[a]) / e * 1024.0f : 0.0f;
      f = fminf(fmaxf(f, 0.0f), 1023.0f);
      if (!(f == f)) f = 0.0f;
      code |= spread10b((uint32_t)f) << (2 - a);
    }
  }
  keys[p] = ((uint64_t)lo << 32) | code;
  vals[p] = local;
}

__global__ void zero_words_kernel(uint32_t* a, size_t na, uint32_t* b, size_t nb,
                                  unsigned long long* c, size_t nc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < na) a[i] = 0u;
  if (i < nb) b[i] = 0u;
  if (i < nc) c[i] = 0ull;
}

// One persistent launch (or chain of launches) per family.  KNN jobs are
// split by the list size their k needs (1, 2, 4, ..., 64), so a small-k job
// never pays for the registers of a larger one in the same batch.
enum Family : int {
  kFamPoints1 = 0,          // ... kFamPoints1 + 6 for K = 64
  kFamNbr = kPointBuckets,  // fixed radius: count, emit, sort
  kFamRayAny,
  kFamRayCount,
  kFamRayClosest,
  kFamPip,
};
static_assert(kFamPip + 1 == kLaunchFamilies, "family count");
// work queues: one per family, plus the fixed-radius emit and sort launches
constexpr int kItemCounterSets = kLaunchFamilies + 2;
constexpr int kQueueNbrEmit = kLaunchFamilies;
constexpr int kQueueNbrSort = kLaunchFamilies + 1;
// parking space for fixed-radius candidates, per fixed-radius query of a batch
constexpr uint64_t kNbrCandPerQuery = 80;
// A fixed-radius family of more work items than this runs as consecutive
// sub-launches (count, emit, sort) of at most this many items, in stream
// order, reusing one stage (1 KiB per query) and one parking space (1.25 KiB
// per query): the scratch is bounded by 4M queries however large the batch.
constexpr uint32_t kNbrChunkItems = 65536;

int family_of(uint32_t kind, uint32_t k) {
  switch (kind) {
    case kJobFixedRadius: return kFamNbr;
    case kJobKnn:
    case kJobBoundedKnn: {
      int b = 0;
      while ((1u << b) < k) ++b;  // list_capacity(k) == 1 << b
      return kFamPoints1 + b;
    }
    case kJobRayAnyHit: return kFamRayAny;
    case kJobRayHitCount: return kFamRayCount;
    case kJobRayClosest: return kFamRayClosest;
    default: return kFamPip;
  }
}

// report group of a launch family: neighbours, any-hit, hit-count, closest, PIP
int group_of(int family) { return family <= kFamNbr ? 0 : family - kFamNbr; }

template <int K, uint32_t DEPTH>
void launch_points_depth(bool stats, unsigned grid, const DeviceJob* jobs, const WorkItem* items,
                         uint32_t n_items, uint32_t* counter, uint4* amb, uint32_t* amb_count,
                         unsigned long long* st, hipStream_t s) {
  if (stats)
    batch_points_kernel<K, true, DEPTH><<<grid, 256, 0, s>>>(jobs, items, n_items, counter, amb, amb_count, st);
  else
    batch_points_kernel<K, false, DEPTH><<<grid, 256, 0, s>>>(jobs, items, n_items, counter, amb, amb_count, st);
  RTDL_LAUNCH_CHECK();
}

// The KNN walk's stack depth is a tuned template variant (Tuning::knn_stack).
template <int K>
void launch_points(bool stats, unsigned grid, const DeviceJob* jobs, const WorkItem* items,
                   uint32_t n_items, uint32_t* counter, uint4* amb, uint32_t* amb_count,
                   unsigned long long* st, hipStream_t s) {
  switch (current_tuning().knn_stack) {
    case 16: launch_points_depth<K, 16>(stats, grid, jobs, items, n_items, counter, amb, amb_count, st, s); break;
    case 48: launch_points_depth<K, 48>(stats, grid, jobs, items, n_items, counter, amb, amb_count, st, s); break;
    default: launch_points_depth<K, 32>(stats, grid, jobs, items, n_items, counter, amb, amb_count, st, s); break;
  }
}

// `threads` per workgroup: Tuning::walk_block (the kernels hold no LDS, so
// only the scheduling granularity changes).
template <int MODE>
void launch_rays_batch(bool stats, unsigned grid, unsigned threads, const DeviceJob* jobs,
                       const WorkItem* items, uint32_t n_items, uint32_t* counter, unsigned long long* st,
                       hipStream_t s) {
  if (stats)
    batch_rays_kernel<MODE, true><<<grid, threads, 0, s>>>(jobs, items, n_items, counter, st);
  else
    batch_rays_kernel<MODE, false><<<grid, threads, 0, s>>>(jobs, items, n_items, counter, st);
  RTDL_LAUNCH_CHECK();
}

}  // namespace

BatchEngine::BatchEngine(hipStream_t stream) : stream_(stream) {
  RTDL_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
  int dev = 0;
  RTDL_HIP(hipGetDevice(&dev));
  RTDL_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
  stats_.resize(3);
  item_counters_.resize(kItemCounterSets * kXcds);
  amb_count_.resize(2);  // [0] neighbour queries for the exact fix-up, [1] PIP points past the hit cap
  nbr_cand_count_.resize(2);  // [0] parking cursor, [1] (low word) emit list length
  for (int f = 0; f < kLaunchFamilies; ++f) {
    RTDL_HIP(hipEventCreate(&fam_start_[f]));
    RTDL_HIP(hipEventCreate(&fam_stop_[f]));
  }
}

// Fixed layout, in the order the growing engine sizes the same buffers.
void BatchEngine::carve(Carver& dev, Carver& host, size_t max_jobs, size_t max_queries) {
  max_jobs = std::max<size_t>(max_jobs, 1);
  max_queries = std::max<size_t>(max_queries, 1);
  const size_t max_items = max_queries / 64 + max_jobs;
  dev.take(stats_, 3);
  dev.take(item_counters_, kItemCounterSets * kXcds);
  dev.take(amb_count_, 2);
  dev.take(nbr_cand_count_, 2);
  dev.take(jobs_, max_jobs);
  dev.take(starts_, max_jobs + 1);
  dev.take(counters_, max_jobs);
  dev.take(items_, max_items);
  dev.take(items_sorted_, max_items);
  dev.take(item_keys_, max_items);
  dev.take(item_vals_, max_items);
  dev.take(nbr_plan_, max_items * 64);
  dev.take(keys_, max_queries);
  dev.take(order_, max_queries);
  dev.take(amb_, max_queries);
  dev.take(pip_over_, max_queries);
  sort_ws_.carve(dev, std::max(max_queries, max_items));
  dev.take(nbr_cand_, std::max<size_t>(max_queries * kNbrCandPerQuery, 1u << 16));
  dev.take(nbr_stage_, max_items * 64 * kNbrStageMax);
  dev.take(nbr_emit_, max_items * 64);
  host_jobs_ = host.take<DeviceJob>(max_jobs);
  host_starts_ = host.take<uint64_t>(max_jobs + 1);
  host_counters_ = host.take<JobCounters>(max_jobs);
  host_items_ = host.take<WorkItem>(max_items);
  host_scene_order_ = host.take<uint32_t>(max_jobs);
  host_scene_idx_ = host.take<uint32_t>(max_jobs);
  cap_jobs_ = max_jobs;
  cap_queries_ = max_queries;
  cap_items_ = max_items;
}

void BatchEngine::fixed_bytes(size_t max_jobs, size_t max_queries, size_t* device_bytes, size_t* host_bytes) {
  BatchEngine probe{MeasureTag{}};
  Carver dev, host;
  probe.carve(dev, host, max_jobs, max_queries);
  *device_bytes = dev.used;
  *host_bytes = host.used;
}

BatchEngine::BatchEngine(hipStream_t stream, const EngineStorage& st) : stream_(stream), fixed_(true) {
  require(st.device != nullptr && st.host != nullptr, "engine storage must not be null");
  int dev = 0;
  RTDL_HIP(hipGetDevice(&dev));
  RTDL_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
  Carver d{static_cast<uint8_t*>(st.device), st.device_bytes};
  Carver h{static_cast<uint8_t*>(st.host), st.host_bytes};
  carve(d, h, st.max_jobs, st.max_queries);
}

void BatchEngine::family_seconds(double out[5]) const {
  for (int g = 0; g < 5; ++g) out[g] = 0.0;
  for (int f = 0; f < kLaunchFamilies; ++f) {
    if (!timed_[f]) continue;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, fam_start_[f], fam_stop_[f]) == hipSuccess)
      out[group_of(f)] += ms * 1e-3;
    else
      (void)hipGetLastError();
  }
}

BatchEngine::~BatchEngine() {
  for (int f = 0; f < kLaunchFamilies; ++f) {
    if (fam_start_[f]) (void)hipEventDestroy(fam_start_[f]);
    if (fam_stop_[f]) (void)hipEventDestroy(fam_stop_[f]);
  }
  if (done_) (void)hipEventDestroy(done_);
  if (fixed_) return;  // host arrays belong to the caller's storage
  if (host_jobs_) (void)hipHostFree(host_jobs_);
  if (host_items_) (void)hipHostFree(host_items_);
  if (host_starts_) (void)hipHostFree(host_starts_);
  if (host_counters_) (void)hipHostFree(host_counters_);
}

void BatchEngine::reserve(size_t max_jobs, size_t max_queries) {
  max_jobs = std::max<size_t>(max_jobs, 1);
  max_queries = std::max<size_t>(max_queries, 1);
  const size_t max_items = max_queries / 64 + max_jobs;
  if (fixed_) {
    if (max_jobs > cap_jobs_ || max_queries > cap_queries_)
      throw Error("batch exceeds the fixed engine's limits of " + std::to_string(cap_jobs_) + " jobs and " +
                  std::to_string(cap_queries_) + " queries");
    return;
  }
  if (max_jobs > cap_jobs_) {
    host_scene_storage_.resize(2 * max_jobs);
    host_scene_order_ = host_scene_storage_.data();
    host_scene_idx_ = host_scene_storage_.data() + max_jobs;
    if (host_jobs_) RTDL_HIP(hipHostFree(host_jobs_));
    if (host_starts_) RTDL_HIP(hipHostFree(host_starts_));
    if (host_counters_) RTDL_HIP(hipHostFree(host_counters_));
    RTDL_HIP(hipHostMalloc((void**)&host_jobs_, max_jobs * sizeof(DeviceJob), hipHostMallocDefault));
    RTDL_HIP(hipHostMalloc((void**)&host_starts_, (max_jobs + 1) * sizeof(uint64_t), hipHostMallocDefault));
    RTDL_HIP(hipHostMalloc((void**)&host_counters_, max_jobs * sizeof(JobCounters), hipHostMallocDefault));
    jobs_.resize(max_jobs);
    starts_.resize(max_jobs + 1);
    counters_.resize(max_jobs);
    cap_jobs_ = max_jobs;
  }
  if (max_items > cap_items_) {
    if (host_items_) RTDL_HIP(hipHostFree(host_items_));
    RTDL_HIP(hipHostMalloc((void**)&host_items_, max_items * sizeof(WorkItem), hipHostMallocDefault));
    items_.resize(max_items);
    items_sorted_.resize(max_items);
    item_keys_.resize(max_items);
    item_vals_.resize(max_items);
    nbr_plan_.resize(max_items * 64);
    cap_items_ = max_items;
  }
  if (max_queries > cap_queries_) {
    keys_.resize(max_queries);
    order_.resize(max_queries);
    amb_.resize(max_queries);
    pip_over_.resize(max_queries);
    sort_ws_.reserve(max_queries);
    cap_queries_ = max_queries;
  }
}

// Parking space and stage of the fixed-radius jobs of a batch.
void BatchEngine::reserve_nbr(const HostJob* jobs, size_t n) {
  uint64_t nbr_queries = 0, nbr_items = 0;
  for (size_t i = 0; i < n; ++i)
    if (jobs[i].kind == kJobFixedRadius) {
      nbr_queries += jobs[i].nq;
      nbr_items += (jobs[i].nq + 63) / 64;
    }
  if (!nbr_queries) return;
  // one sub-launch's worth (see kNbrChunkItems)
  nbr_items = std::min<uint64_t>(nbr_items, kNbrChunkItems);
  nbr_queries = std::min<uint64_t>(nbr_queries, (uint64_t)kNbrChunkItems * 64);
  const size_t want = (size_t)std::max<uint64_t>(nbr_queries * kNbrCandPerQuery, 1u << 16);
  const size_t stage = (size_t)nbr_items * 64 * kNbrStageMax;
  if (fixed_) {  // carved for max_queries / max_items, which reserve() checked
    require(nbr_cand_.size() >= want && nbr_stage_.size() >= stage && nbr_emit_.size() >= nbr_items * 64,
            "fixed engine: fixed-radius space");
    return;
  }
  if (nbr_cand_.size() < want) nbr_cand_.resize(want);
  if (nbr_stage_.size() < stage) nbr_stage_.resize(stage);
  if (nbr_emit_.size() < nbr_items * 64) nbr_emit_.resize(nbr_items * 64);
}

void BatchEngine::enqueue(const HostJob* jobs, size_t n, bool morton_sort) {
  require(n == 0 || jobs != nullptr, "jobs must not be null");
  uint64_t total = 0;
  for (size_t i = 0; i < n; ++i) total += jobs[i].nq;
  require(total < 0xFFFFFFFFull, "batch exceeds 2^32 queries");
  reserve(n, total);
  // A scene is one instance's tree: jobs on it share its nodes.  Jobs sorted
  // by (node array, root) get one index per distinct scene (in place, so a
  // batch allocates nothing on the host).
  auto scene_less = [&](uint32_t a, uint32_t b) {
    const SceneView &x = jobs[a].scene, &y = jobs[b].scene;
    return x.nodes != y.nodes ? std::less<const void*>()(x.nodes, y.nodes) : x.root < y.root;
  };
  for (size_t i = 0; i < n; ++i) host_scene_order_[i] = (uint32_t)i;
  std::sort(host_scene_order_, host_scene_order_ + n, scene_less);
  uint32_t n_scenes = 0;
  for (size_t i = 0; i < n; ++i) {
    const uint32_t j = host_scene_order_[i];
    if (i == 0 || scene_less(host_scene_order_[i - 1], j)) ++n_scenes;
    host_scene_idx_[j] = n_scenes - 1;
  }
  // order jobs' items by predicate family; each family gets one launch
  size_t item_off[kLaunchFamilies + 1] = {0};
  uint64_t start = 0;
  for (size_t i = 0; i < n; ++i) {
    const HostJob& h = jobs[i];
    require(h.qstride == 2 || h.qstride == 3 || h.qstride == 7, "query stride must be 2, 3 or 7");
    require(h.nq == 0 || h.queries != nullptr, "job queries must not be null");
    require(h.nq == 0 || h.rows != nullptr, "job rows must not be null");
    if (h.kind <= kJobBoundedKnn) require(h.k >= 1 && h.k <= 64, "batched neighbour jobs need 1 <= k <= 64");
    require(h.row_capacity <= kMaxJobRows && job_row_capacity(h.kind, h.nq, h.k, 0) <= kMaxJobRows,
            "a batch job holds at most 2^32 - 1 rows (split it into smaller jobs)");
    const int fam = family_of(h.kind, h.k);
    DeviceJob& d = host_jobs_[i];
    std::memset(&d, 0, sizeof(d));
    d.scene = h.scene;
    d.queries = h.queries;
    d.order = morton_sort ? order_.get() : nullptr;
    d.rows = h.rows;
    d.counters = counters_.get() + i;
    d.row_capacity = h.row_capacity;
    d.radius = h.kind == kJobKnn ? __builtin_inf() : h.radius;
    d.kind = h.kind;
    d.nq = h.nq;
    d.k = h.k;
    d.qstride = h.qstride;
    d.order_base = (uint32_t)start;
    host_starts_[i] = start;
    item_off[fam + 1] += (h.nq + 63) / 64;
    start += h.nq;
  }
  host_starts_[n] = start;
  reserve_nbr(jobs, n);
  for (int f = 0; f < kLaunchFamilies; ++f) item_off[f + 1] += item_off[f];
  {
    size_t cursor[kLaunchFamilies];
    for (int f = 0; f < kLaunchFamilies; ++f) cursor[f] = item_off[f];
    for (size_t i = 0; i < n; ++i) {
      const HostJob& h = jobs[i];
      const int fam = family_of(h.kind, h.k);
      const uint32_t group = ((uint32_t)fam << 24) | host_scene_idx_[i];
      for (uint64_t f = 0; f < h.nq; f += 64)
        host_items_[cursor[fam]++] =
            WorkItem{(uint32_t)i, (uint32_t)f, (uint32_t)std::min<uint64_t>(64, h.nq - f), group};
    }
  }
  hipStream_t s = stream_;
  if (n) {
    RTDL_HIP(hipMemcpyAsync(jobs_.get(), host_jobs_, n * sizeof(DeviceJob), hipMemcpyHostToDevice, s));
    RTDL_HIP(hipMemcpyAsync(starts_.get(), host_starts_, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  }
  if (item_off[kLaunchFamilies])
    RTDL_HIP(hipMemcpyAsync(items_.get(), host_items_, item_off[kLaunchFamilies] * sizeof(WorkItem),
                            hipMemcpyHostToDevice, s));
  const size_t zero_n = std::max<size_t>(n * 4, kItemCounterSets * kXcds + 1);
  zero_words_kernel<<<blocks_for(zero_n, 256), 256, 0, s>>>(
      reinterpret_cast<uint32_t*>(counters_.get()), n * 4, item_counters_.get(), kItemCounterSets * kXcds,
      stats_.get(), 3);
  RTDL_LAUNCH_CHECK();
  fill_u32(amb_count_.get(), 2, 0u, s);
  fill_u32(reinterpret_cast<uint32_t*>(nbr_cand_count_.get()), 4, 0u, s);
  if (morton_sort && total > 1) {
    batch_keys_kernel<<<blocks_for(total, 256), 256, 0, s>>>(jobs_.get(), starts_.get(), (uint32_t)n,
                                                              total, keys_.get(), order_.get());
    RTDL_LAUNCH_CHECK();
    int job_bits = 0;
    while ((1ull << job_bits) < n) ++job_bits;
    radix_sort_pairs_async(keys_.get(), order_.get(), total, 30 + 2 + job_bits, sort_ws_, s);
  } else if (morton_sort && total == 1) {
    fill_u32(order_.get(), 1, 0u, s);
  }
  // Items of one family run in (scene, Morton code of their first query)
  // order across jobs: a queue's contiguous share (one XCD's) then covers one
  // region of a scene at a time, whose nodes and primitives stay in that
  // XCD's L2, instead of every job's whole scene.
  const size_t n_items_all = item_off[kLaunchFamilies];
  const WorkItem* items_run = items_.get();
  if (morton_sort && total > 1 && n_items_all > 1) {
    int scene_bits = 0;
    while ((1ull << scene_bits) < n_scenes) ++scene_bits;
    require(scene_bits <= 24, "batch spans more than 2^24 scenes");
    item_keys_kernel<<<blocks_for(n_items_all, 256), 256, 0, s>>>(items_.get(), starts_.get(), keys_.get(),
                                                                  n_items_all, scene_bits, item_keys_.get(),
                                                                  item_vals_.get());
    RTDL_LAUNCH_CHECK();
    radix_sort_pairs_async(item_keys_.get(), item_vals_.get(), n_items_all, 30 + scene_bits + 4, sort_ws_, s);
    gather_items_kernel<<<blocks_for(n_items_all, 256), 256, 0, s>>>(items_.get(), item_vals_.get(),
                                                                     n_items_all, items_sorted_.get());
    RTDL_LAUNCH_CHECK();
    items_run = items_sorted_.get();
  }
  const Tuning tuning = current_tuning();
  const unsigned max_grid = (unsigned)cus_ * tuning.blocks_per_cu;
  auto grid_for = [&](size_t items) {
    return (unsigned)std::max<size_t>(1, std::min<size_t>((items + 3) / 4, max_grid));
  };
  // ray and PIP walks: workgroups of walk_block threads, the same wavefront
  // budget (max_grid 256-thread workgroups) and one wavefront per item
  // the exact fix-up: its 48 KB LDS lists fit three workgroups per CU, and
  // every one of them takes a share of the flagged queries
  const unsigned fix_grid = (unsigned)cus_ * 3;
  const unsigned walk_block = tuning.walk_block, walk_waves = walk_block / 64;
  auto walk_grid_for = [&](size_t items) {
    return (unsigned)std::max<size_t>(
        1, std::min<size_t>((items + walk_waves - 1) / walk_waves, (size_t)max_grid * 4 / walk_waves));
  };
  unsigned long long* st = stats_on_ ? stats_.get() : nullptr;
  const DeviceJob* dj = jobs_.get();
  int last_nbr = -1;  // the exact fix-up runs once, after the last neighbour launch
  for (int f = 0; f <= kFamNbr; ++f)
    if (item_off[f + 1] > item_off[f]) last_nbr = f;
  for (int f = 0; f < kLaunchFamilies; ++f) timed_[f] = false;
  for (int f = 0; f < kLaunchFamilies; ++f) {
    const uint32_t n_items = (uint32_t)(item_off[f + 1] - item_off[f]);
    if (!n_items) continue;
    const WorkItem* it = items_run + item_off[f];
    uint32_t* ctr = item_counters_.get() + f * kXcds;
    const unsigned grid = grid_for(n_items);
    if (timing_on_) {
      RTDL_HIP(hipEventRecord(fam_start_[f], s));
      timed_[f] = true;
    }
    if (f < kPointBuckets) {
      switch (f) {
        case 0: launch_points<1>(stats_on_, grid, dj, it, n_items, ctr, amb_.get(), amb_count_.get(), st, s); break;
        case 1: launch_points<2>(stats_on_, grid, dj, it, n_items, ctr, amb_.get(), amb_count_.get(), st, s); break;
        case 2: launch_points<4>(stats_on_, grid, dj, it, n_items, ctr, amb_.get(), amb_count_.get(), st, s); break;
        case 3: launch_points<8>(stats_on_, grid, dj, it, n_items, ctr, amb_.get(), amb_count_.get(), st, s); break;
        case 4: launch_points<16>(stats_on_, grid, dj, it, n_items, ctr, amb_.get(), amb_count_.get(), st, s); break;
        case 5: launch_points<32>(stats_on_, grid, dj, it, n_items, ctr, amb_.get(), amb_count_.get(), st, s); break;
        default: launch_points<64>(stats_on_, grid, dj, it, n_items, ctr, amb_.get(), amb_count_.get(), st, s); break;
      }
      if (f == last_nbr) {
        batch_fix_kernel<<<fix_grid, kFixBlock, 0, s>>>(dj, amb_.get(), amb_count_.get());
        RTDL_LAUNCH_CHECK();
      }
    } else if (f == kFamNbr) {
      // sub-launches of at most kNbrChunkItems items; the first one's queue
      // counters and parking cursor were cleared above, later ones clear
      // theirs (fill kernels: graph-safe) after the previous sort
      for (uint32_t c0 = 0; c0 < n_items; c0 += kNbrChunkItems) {
        const uint32_t m = std::min<uint32_t>(kNbrChunkItems, n_items - c0);
        const WorkItem* ic = it + c0;
        const unsigned gm = grid_for(m);
        require(nbr_stage_.size() >= (size_t)m * 64 * kNbrStageMax && nbr_emit_.size() >= (size_t)m * 64,
                "fixed-radius stage not reserved");
        if (c0) {
          fill_u32(ctr, kXcds, 0u, s);
          fill_u32(item_counters_.get() + kQueueNbrEmit * kXcds, 2 * kXcds, 0u, s);
          fill_u32(reinterpret_cast<uint32_t*>(nbr_cand_count_.get()), 4, 0u, s);
        }
        uint32_t* emit_count = reinterpret_cast<uint32_t*>(nbr_cand_count_.get() + 1);
        if (stats_on_)
          nbr_count_kernel<true><<<gm, 256, 0, s>>>(dj, ic, m, ctr, nbr_plan_.get(), nbr_cand_count_.get(),
                                                    nbr_cand_.size(), nbr_stage_.get(), kNbrStageMax,
                                                    nbr_emit_.get(), emit_count, amb_.get(), amb_count_.get(), st);
        else
          nbr_count_kernel<false><<<gm, 256, 0, s>>>(dj, ic, m, ctr, nbr_plan_.get(), nbr_cand_count_.get(),
                                                     nbr_cand_.size(), nbr_stage_.get(), kNbrStageMax,
                                                     nbr_emit_.get(), emit_count, amb_.get(), amb_count_.get(), st);
        RTDL_LAUNCH_CHECK();
        nbr_emit_kernel<<<gm, 256, 0, s>>>(dj, ic, nbr_plan_.get(), nbr_cand_.get(), nbr_emit_.get(), emit_count,
                                           item_counters_.get() + kQueueNbrEmit * kXcds);
        RTDL_LAUNCH_CHECK();
        const unsigned g64 = (unsigned)std::max<size_t>(1, std::min<size_t>(m, (size_t)cus_ * 32));
        nbr_sort_kernel<<<g64, kSortBlock, 0, s>>>(dj, ic, nbr_plan_.get(), nbr_cand_.get(), nbr_emit_.get(),
                                                   emit_count, item_counters_.get() + kQueueNbrSort * kXcds);
        RTDL_LAUNCH_CHECK();
      }
      if (f == last_nbr) {
        batch_fix_kernel<<<fix_grid, kFixBlock, 0, s>>>(dj, amb_.get(), amb_count_.get());
        RTDL_LAUNCH_CHECK();
      }
    } else if (f == kFamRayAny) {
      launch_rays_batch<0>(stats_on_, walk_grid_for(n_items), walk_block, dj, it, n_items, ctr, st, s);
    } else if (f == kFamRayCount) {
      launch_rays_batch<1>(stats_on_, walk_grid_for(n_items), walk_block, dj, it, n_items, ctr, st, s);
    } else if (f == kFamRayClosest) {
      launch_rays_batch<2>(stats_on_, walk_grid_for(n_items), walk_block, dj, it, n_items, ctr, st, s);
    } else {
      const double eps_sq = std::pow(1.0e-12, 2.0);  // reference.py:757 `eps ** 2`
      uint32_t* over_count = amb_count_.get() + 1;
      const unsigned wg = walk_grid_for(n_items);
      if (stats_on_)
        batch_pip_kernel<true><<<wg, walk_block, walk_waves * sizeof(PipWaveLds), s>>>(dj, it, n_items, ctr, eps_sq, pip_over_.get(),
                                                         over_count, st);
      else
        batch_pip_kernel<false><<<wg, walk_block, walk_waves * sizeof(PipWaveLds), s>>>(dj, it, n_items, ctr, eps_sq, pip_over_.get(),
                                                          over_count, st);
      RTDL_LAUNCH_CHECK();
      batch_pip_over_kernel<<<(unsigned)cus_ * 4, 64, 0, s>>>(dj, pip_over_.get(), over_count, eps_sq);
      RTDL_LAUNCH_CHECK();
    }
    if (timing_on_) RTDL_HIP(hipEventRecord(fam_stop_[f], s));
  }
  if (n)
    RTDL_HIP(hipMemcpyAsync(host_counters_, counters_.get(), n * sizeof(JobCounters),
                            hipMemcpyDeviceToHost, s));
  if (!capturing_ && done_) RTDL_HIP(hipEventRecord(done_, s));
}

hipGraphExec_t BatchEngine::capture(const std::vector<HostJob>& jobs, bool morton_sort,
                                    const std::function<void(hipStream_t)>& prologue) {
  // every allocation enqueue would make happens here, outside the capture
  uint64_t total = 0;
  for (const HostJob& h : jobs) total += h.nq;
  require(total < 0xFFFFFFFFull, "batch exceeds 2^32 queries");
  reserve(jobs.size(), total);
  reserve_nbr(jobs.data(), jobs.size());
  enable_timing(false);
  enable_stats(false);
  RTDL_HIP(hipStreamSynchronize(stream_));
  RTDL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  capturing_ = true;
  try {
    if (prologue) prologue(stream_);
    enqueue(jobs, morton_sort);
  } catch (...) {
    capturing_ = false;
    hipGraph_t broken = nullptr;
    (void)hipStreamEndCapture(stream_, &broken);
    if (broken) (void)hipGraphDestroy(broken);
    throw;
  }
  capturing_ = false;
  hipGraph_t graph = nullptr;
  RTDL_HIP(hipStreamEndCapture(stream_, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  RTDL_HIP(e);
  return exec;
}

}  // namespace rtdl
