// Host runtime: caching device allocator, per-thread streams, phase clock,
// prepared scenes and the upload -> traverse -> emit pipelines behind the
// C-ABI.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include <unistd.h>

#include "common.hpp"
#include "io.hpp"
#include "primitives.hpp"
#include "runtime.hpp"

namespace rtdl {

// ------------------------------------------------------------- device pool --
//
// Freed blocks are cached by rounded size and handed back to later requests.
// Every pipeline synchronises its stream before its temporaries go out of
// scope, so a cached block is never reused while a kernel still reads it.

namespace {
std::mutex g_pool_mutex;
std::multimap<size_t, void*> g_pool_free;
std::map<void*, size_t> g_pool_size;
size_t g_pool_cached = 0;
constexpr size_t kPoolCacheLimit = size_t(32) << 30;

size_t pool_round(size_t bytes) {
  if (bytes <= 256) return 256;
  if (bytes <= (size_t(1) << 20)) {
    size_t r = 256;
    while (r < bytes) r <<= 1;
    return r;
  }
  const size_t gran = size_t(2) << 20;
  return (bytes + gran - 1) / gran * gran;
}
}  // namespace

// ------------------------------------------------------------ fault hook --

namespace {
std::mutex g_fault_mutex;
bool g_fault_loaded = false;
FaultKind g_fault_kind = FaultKind::None;
uint64_t g_fault_nth = 0, g_fault_seen = 0;

// Parsed with C string functions: only a bad spec builds a message, so
// arming or reading the hook allocates nothing.
void parse_fault(const char* spec) {
  const char* s = spec ? spec : "";
  const char* colon = std::strchr(s, ':');
  const size_t len = colon ? (size_t)(colon - s) : std::strlen(s);
  FaultKind kind = FaultKind::None;
  if (len == 5 && std::strncmp(s, "alloc", 5) == 0) kind = FaultKind::Alloc;
  else if (len == 6 && std::strncmp(s, "launch", 6) == 0) kind = FaultKind::Launch;
  else if (len == 7 && std::strncmp(s, "timeout", 7) == 0) kind = FaultKind::Timeout;
  else if (len != 0) throw Error(std::string("fault spec must be alloc, launch or timeout[:N], not '") + s + "'");
  uint64_t nth = 1;
  if (colon) {
    char* end = nullptr;
    const unsigned long long v = std::strtoull(colon + 1, &end, 10);
    if (end == colon + 1 || *end != '\0' || v == 0)
      throw Error(std::string("fault spec: the occurrence must be a positive integer, not '") + s + "'");
    nth = v;
  }
  g_fault_kind = kind;
  g_fault_nth = nth;
  g_fault_seen = 0;
  g_fault_loaded = true;
}
}  // namespace

void set_fault(const char* spec) {
  std::lock_guard<std::mutex> g(g_fault_mutex);
  parse_fault(spec);
}

bool fault_fires(FaultKind kind) {
  std::lock_guard<std::mutex> g(g_fault_mutex);
  if (!g_fault_loaded) {
    try {
      parse_fault(std::getenv("RTDL_MI355X_FAULT"));
    } catch (const Error&) {
      g_fault_loaded = true;  // an unusable environment value arms nothing
      g_fault_kind = FaultKind::None;
    }
  }
  if (g_fault_kind != kind) return false;
  return ++g_fault_seen == g_fault_nth;
}

void* pool_alloc(size_t bytes) {
  if (fault_fires(FaultKind::Alloc))
    throw Error("injected fault (RTDL_MI355X_FAULT=alloc): device allocation of " + std::to_string(bytes) +
                " bytes failed");
  const size_t size = pool_round(bytes);
  {
    std::lock_guard<std::mutex> g(g_pool_mutex);
    auto it = g_pool_free.find(size);
    if (it != g_pool_free.end()) {
      void* p = it->second;
      g_pool_free.erase(it);
      g_pool_cached -= size;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t err = hipMalloc(&p, size);
  if (err != hipSuccess) {
    pool_trim();
    (void)hipGetLastError();
    err = hipMalloc(&p, size);
  }
  if (err != hipSuccess) {
    throw Error("device allocation of " + std::to_string(size) + " bytes failed: " +
                hipGetErrorString(err));
  }
  std::lock_guard<std::mutex> g(g_pool_mutex);
  g_pool_size[p] = size;
  return p;
}

void pool_free(void* ptr) {
  if (!ptr) return;
  std::lock_guard<std::mutex> g(g_pool_mutex);
  auto it = g_pool_size.find(ptr);
  if (it == g_pool_size.end()) return;
  const size_t size = it->second;
  if (g_pool_cached + size > kPoolCacheLimit) {
    g_pool_size.erase(it);
    (void)hipFree(ptr);
    return;
  }
  g_pool_free.emplace(size, ptr);
  g_pool_cached += size;
}

void pool_trim() {
  std::lock_guard<std::mutex> g(g_pool_mutex);
  for (auto& kv : g_pool_free) {
    g_pool_size.erase(kv.second);
    (void)hipFree(kv.second);
  }
  g_pool_free.clear();
  g_pool_cached = 0;
}

// ------------------------------------------------------- streams and clock --

namespace {
struct ThreadStream {
  hipStream_t s = nullptr;
  ~ThreadStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};
thread_local ThreadStream t_stream;
thread_local PhaseClock t_clock;
}  // namespace

namespace {
// Events and counters behind WalkStats, per thread.
struct WalkProbe {
  WalkStats stats;
  hipEvent_t start = nullptr, stop = nullptr;
  DeviceArray<unsigned long long> counters;
  ~WalkProbe() {
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
  }
};
thread_local WalkProbe t_probe;

// Marks the start of a call's traversal kernels; returns the counters the
// kernels add to while counting is on (null otherwise).
unsigned long long* walk_begin(hipStream_t s) {
  WalkProbe& w = t_probe;
  w.stats.valid = false;
  if (!w.start) {
    RTDL_HIP(hipEventCreate(&w.start));
    RTDL_HIP(hipEventCreate(&w.stop));
  }
  unsigned long long* ctr = nullptr;
  if (w.stats.counting) {
    if (!w.counters.size()) w.counters.resize(3);
    fill_u32(reinterpret_cast<uint32_t*>(w.counters.get()), 6, 0u, s);
    ctr = w.counters.get();
  }
  RTDL_HIP(hipEventRecord(w.start, s));
  return ctr;
}

void walk_end(hipStream_t s) { RTDL_HIP(hipEventRecord(t_probe.stop, s)); }

// After the stream has been synchronised past walk_end.
void walk_collect() {
  WalkProbe& w = t_probe;
  float ms = 0.0f;
  RTDL_HIP(hipEventElapsedTime(&ms, w.start, w.stop));
  w.stats.kernel_s = ms * 1e-3;
  w.stats.nodes = w.stats.prims = w.stats.steps = 0;
  if (w.stats.counting) {
    unsigned long long v[3] = {0, 0, 0};
    RTDL_HIP(hipMemcpy(v, w.counters.get(), sizeof(v), hipMemcpyDeviceToHost));
    w.stats.nodes = v[0];
    w.stats.prims = v[1];
    w.stats.steps = v[2];
  }
  w.stats.valid = true;
}
}  // namespace

WalkStats& thread_walk_stats() { return t_probe.stats; }

hipStream_t thread_stream() {
  if (!t_stream.s) RTDL_HIP(hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking));
  return t_stream.s;
}

PhaseClock& thread_clock() { return t_clock; }

void PhaseClock::begin() {
  seconds.fill(0.0);
  start = mark = std::chrono::steady_clock::now();
}
void PhaseClock::lap(int phase) {
  const auto now = std::chrono::steady_clock::now();
  seconds[phase] += std::chrono::duration<double>(now - mark).count();
  mark = now;
}
void PhaseClock::end() {
  seconds[5] = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
}

// --------------------------------------------------------------- prepared --

Prepared::Prepared(bool own_stream) : owns_stream(own_stream) {
  if (own_stream)
    RTDL_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  else
    stream = thread_stream();
}
Prepared::~Prepared() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
    if (owns_stream) (void)hipStreamDestroy(stream);
  }
  magic = 0;
}

Prepared* prepared_from(void* handle) {
  require(handle != nullptr, "prepared handle must not be null");
  auto* p = static_cast<Prepared*>(handle);
  require(p->magic == 0x52544D58u, "invalid prepared handle");
  return p;
}

namespace {
std::mutex g_tuning_mutex;
Tuning g_tuning;
}  // namespace

Tuning current_tuning() {
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  return g_tuning;
}

void set_tuning(const Tuning& t) {
  require(t.point_leaf >= 1 && t.point_leaf <= kMaxLeafSize, "point_leaf must be in [1, 8]");
  require(t.triangle_leaf >= 1 && t.triangle_leaf <= kMaxLeafSize, "triangle_leaf must be in [1, 8]");
  require(t.polygon_leaf >= 1 && t.polygon_leaf <= kMaxLeafSize, "polygon_leaf must be in [1, 8]");
  require(t.blocks_per_cu >= 1 && t.blocks_per_cu <= 64, "blocks_per_cu must be in [1, 64]");
  require(t.knn_stack == 16 || t.knn_stack == 32 || t.knn_stack == 48, "knn_stack must be 16, 32 or 48");
  require(t.walk_block == 64 || t.walk_block == 128 || t.walk_block == 256, "walk_block must be 64, 128 or 256");
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  g_tuning = t;
}

namespace {

constexpr uint32_t kSortQueriesAbove = 2048;

// Pinned staging for large host <-> device copies.  A copy from or into
// pageable memory runs at a few tens of GB/s at best; large inputs and row
// sets go instead through two pinned 32 MiB buffers, each chunk's DMA
// overlapping the host copy of the next (or previous) chunk, with the host
// copies split over threads.  One caller at a time uses the buffers; a
// concurrent caller (or another device) copies directly.
constexpr size_t kStageChunk = 32u << 20;
constexpr size_t kStageMin = 64u << 20;        // row copies back to the host
constexpr size_t kUploadStageMin = 16u << 20;  // inputs to the device
constexpr unsigned kStageThreads = 8;

// Host copy workers of the staging path: kStageThreads - 1 threads, each
// copying its share of a chunk while the caller copies the first share.
// Used by one caller at a time (under the stage lock); started on first use
// and again in a forked child, which inherits no threads.
struct CopyPool {
  std::mutex m;
  std::condition_variable go, done;
  uint64_t generation = 0;
  unsigned pending = 0;
  uint8_t* dst = nullptr;
  const uint8_t* src = nullptr;
  size_t bytes = 0, part = 0;
  pid_t pid = getpid();

  CopyPool() {
    for (unsigned t = 1; t < kStageThreads; ++t) std::thread([this, t] { run(t); }).detach();
  }

  void run(unsigned t) {
    uint64_t seen = 0;
    for (;;) {
      uint8_t* d;
      const uint8_t* s;
      size_t a, n;
      {
        std::unique_lock<std::mutex> lk(m);
        go.wait(lk, [&] { return generation != seen; });
        seen = generation;
        a = t * part;
        n = a < bytes ? std::min(part, bytes - a) : 0;
        d = dst;
        s = src;
      }
      if (n) std::memcpy(d + a, s + a, n);
      std::lock_guard<std::mutex> lk(m);
      if (--pending == 0) done.notify_one();
    }
  }

  // Returns once every share is copied; the next call starts only after
  // each worker has finished this one, so no worker misses a generation.
  void copy(uint8_t* d, const uint8_t* s, size_t b) {
    const size_t p = (b + kStageThreads - 1) / kStageThreads;
    {
      std::lock_guard<std::mutex> lk(m);
      dst = d;
      src = s;
      bytes = b;
      part = p;
      pending = kStageThreads - 1;
      ++generation;
    }
    go.notify_all();
    std::memcpy(d, s, std::min(p, b));
    std::unique_lock<std::mutex> lk(m);
    done.wait(lk, [&] { return pending == 0; });
  }
};

struct RowStage {
  std::mutex mu;
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int device = -1;
  CopyPool* pool = nullptr;  // like the stage, never freed
};

RowStage& row_stage() {
  static RowStage* stage = new RowStage();  // lives until the process ends
  return *stage;
}

// Allocates the buffers on first use; false when they belong to another device.
bool stage_ready(RowStage& st) {
  int dev = 0;
  RTDL_HIP(hipGetDevice(&dev));
  if (st.device < 0) {
    for (int i = 0; i < 2; ++i) {
      RTDL_HIP(hipHostMalloc(&st.buf[i], kStageChunk, hipHostMallocDefault));
      RTDL_HIP(hipEventCreateWithFlags(&st.ev[i], hipEventDisableTiming));
    }
    st.device = dev;
  }
  return st.device == dev;
}

// Caller holds st.mu.
void parallel_memcpy(RowStage& st, uint8_t* dst, const uint8_t* src, size_t bytes) {
  if (!st.pool || st.pool->pid != getpid()) st.pool = new CopyPool();
  st.pool->copy(dst, src, bytes);
}

// Returns false (nothing copied) when the staging buffers are busy or belong
// to another device.
bool staged_copy(uint8_t* dst, const uint8_t* src, size_t bytes, hipStream_t stream) {
  RowStage& st = row_stage();
  std::unique_lock<std::mutex> lock(st.mu, std::try_to_lock);
  if (!lock.owns_lock() || !stage_ready(st)) return false;
  const size_t n_chunks = (bytes + kStageChunk - 1) / kStageChunk;
  auto issue = [&](size_t c) {
    const size_t a = c * kStageChunk, n = std::min(kStageChunk, bytes - a);
    RTDL_HIP(hipMemcpyAsync(st.buf[c & 1], src + a, n, hipMemcpyDeviceToHost, stream));
    RTDL_HIP(hipEventRecord(st.ev[c & 1], stream));
  };
  issue(0);
  for (size_t c = 0; c < n_chunks; ++c) {
    // the buffer of chunk c + 1 held chunk c - 1, already copied out below
    if (c + 1 < n_chunks) issue(c + 1);
    RTDL_HIP(hipEventSynchronize(st.ev[c & 1]));
    const size_t a = c * kStageChunk, n = std::min(kStageChunk, bytes - a);
    parallel_memcpy(st, dst + a, static_cast<const uint8_t*>(st.buf[c & 1]), n);
  }
  return true;
}

// Host-to-device counterpart of staged_copy; returns once every chunk has
// reached the device (the buffers are free for the next caller).
bool staged_upload(uint8_t* dst, const uint8_t* src, size_t bytes, hipStream_t stream) {
  RowStage& st = row_stage();
  std::unique_lock<std::mutex> lock(st.mu, std::try_to_lock);
  if (!lock.owns_lock() || !stage_ready(st)) return false;
  const size_t n_chunks = (bytes + kStageChunk - 1) / kStageChunk;
  for (size_t c = 0; c < n_chunks; ++c) {
    const size_t a = c * kStageChunk, n = std::min(kStageChunk, bytes - a);
    // the buffer's previous chunk (c - 2) must have left it
    if (c >= 2) RTDL_HIP(hipEventSynchronize(st.ev[c & 1]));
    parallel_memcpy(st, static_cast<uint8_t*>(st.buf[c & 1]), src + a, n);
    RTDL_HIP(hipMemcpyAsync(dst + a, st.buf[c & 1], n, hipMemcpyHostToDevice, stream));
    RTDL_HIP(hipEventRecord(st.ev[c & 1], stream));
  }
  for (size_t c = n_chunks > 2 ? n_chunks - 2 : 0; c < n_chunks; ++c)
    RTDL_HIP(hipEventSynchronize(st.ev[c & 1]));
  return true;
}

template <class T>
DeviceArray<T> upload(const T* host, size_t n, hipStream_t stream) {
  DeviceArray<T> d(n);
  const size_t bytes = n * sizeof(T);
  if (n && (bytes < kUploadStageMin || !staged_upload(reinterpret_cast<uint8_t*>(d.get()),
                                                reinterpret_cast<const uint8_t*>(host), bytes, stream)))
    RTDL_HIP(hipMemcpyAsync(d.get(), host, bytes, hipMemcpyHostToDevice, stream));
  return d;
}

DeviceArray<uint8_t> upload_bytes(const void* host, size_t bytes, hipStream_t stream) {
  return upload(static_cast<const uint8_t*>(host), bytes, stream);
}

}  // namespace

namespace {

// Pinned row buffers for large per-call row sets.  The rows are copied
// straight from device memory into the buffer the caller receives: one DMA at
// the link's rate, with no host copy and no first-touch page faults of a
// fresh allocation.  A buffer the caller releases (rtdl_mi355x_free_rows) is
// kept for a later call of about its size, up to kPinnedCacheBytes of
// released buffers; beyond that it is freed.
constexpr size_t kPinnedRowsMin = 1u << 20;
constexpr size_t kPinnedCacheBytes = size_t(4) << 30;
constexpr size_t kPinnedGranule = 2u << 20;

struct PinnedRows {
  std::mutex mu;
  std::map<void*, size_t> live;         // buffers held by callers: capacity
  std::multimap<size_t, void*> cached;  // released buffers by capacity
  size_t cached_bytes = 0;
};

PinnedRows& pinned_rows() {
  static PinnedRows* rows = new PinnedRows();  // lives until the process ends
  return *rows;
}

}  // namespace

void free_host_rows(void* rows) {
  if (!rows) return;
  PinnedRows& pr = pinned_rows();
  size_t cap = 0;
  std::vector<void*> evict;
  {
    std::lock_guard<std::mutex> g(pr.mu);
    auto it = pr.live.find(rows);
    if (it == pr.live.end()) {
      std::free(rows);  // pageable rows
      return;
    }
    cap = it->second;
    pr.live.erase(it);
    if (cap <= kPinnedCacheBytes) {
      // smaller released buffers make room first: a large one costs the
      // most to pin again
      while (pr.cached_bytes + cap > kPinnedCacheBytes && !pr.cached.empty() && pr.cached.begin()->first <= cap) {
        auto small = pr.cached.begin();
        evict.push_back(small->second);
        pr.cached_bytes -= small->first;
        pr.cached.erase(small);
      }
      if (pr.cached_bytes + cap <= kPinnedCacheBytes) {
        pr.cached.emplace(cap, rows);
        pr.cached_bytes += cap;
        rows = nullptr;
      }
    }
  }
  for (void* e : evict) (void)hipHostFree(e);
  if (rows) (void)hipHostFree(rows);
}

void read_root_box(Prepared& p) {
  const SceneView v = p.scene.view(0);
  if (v.root == kEndNode) return;
  BvhNode root;
  RTDL_HIP(hipMemcpyAsync(&root, p.scene.nodes.get() + v.root, sizeof(root), hipMemcpyDeviceToHost,
                          p.stream));
  RTDL_HIP(hipStreamSynchronize(p.stream));
  for (int a = 0; a < 3; ++a) {
    p.root_lo[a] = root.lo[a];
    p.root_hi[a] = root.hi[a];
  }
}

namespace {

// A pinned buffer of at least `bytes`, or null (the caller then uses pageable
// memory): a released one that fits without wasting more than half, else a
// new allocation.
void* take_pinned_rows(size_t bytes) {
  PinnedRows& pr = pinned_rows();
  {
    std::lock_guard<std::mutex> g(pr.mu);
    auto it = pr.cached.lower_bound(bytes);
    if (it != pr.cached.end() && it->first / 2 <= bytes) {
      void* p = it->second;
      const size_t cap = it->first;
      pr.cached.erase(it);
      pr.cached_bytes -= cap;
      pr.live[p] = cap;
      return p;
    }
  }
  const size_t cap = (bytes + kPinnedGranule - 1) / kPinnedGranule * kPinnedGranule;
  void* p = nullptr;
  if (hipHostMalloc(&p, cap, hipHostMallocPortable) != hipSuccess) {
    (void)hipGetLastError();
    // pinned memory may be exhausted by released buffers: give them back and
    // try once more
    std::vector<void*> evict;
    {
      std::lock_guard<std::mutex> g(pr.mu);
      for (auto& kv : pr.cached) evict.push_back(kv.second);
      pr.cached.clear();
      pr.cached_bytes = 0;
    }
    for (void* e : evict) (void)hipHostFree(e);
    p = nullptr;
    if (evict.empty() || hipHostMalloc(&p, cap, hipHostMallocPortable) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  std::lock_guard<std::mutex> g(pr.mu);
  pr.live[p] = cap;
  return p;
}

void* host_rows(size_t count, size_t row_bytes) {
  void* rows = std::malloc(count ? count * row_bytes : 1);
  require(rows != nullptr, "host row allocation failed");
  return rows;
}

void copy_rows_back(const void* device_rows, size_t count, size_t row_bytes, void** rows_out,
                    size_t* n_out, hipStream_t stream) {
  const size_t bytes = count * row_bytes;
  void* rows = bytes >= kPinnedRowsMin ? take_pinned_rows(bytes) : nullptr;
  const bool pinned = rows != nullptr;
  if (!pinned) rows = host_rows(count, row_bytes);
  if (count) {
    try {
      if (pinned || bytes < kStageMin ||
          !staged_copy(static_cast<uint8_t*>(rows), static_cast<const uint8_t*>(device_rows), bytes, stream)) {
        RTDL_HIP(hipMemcpyAsync(rows, device_rows, bytes, hipMemcpyDeviceToHost, stream));
        RTDL_HIP(hipStreamSynchronize(stream));
      }
    } catch (const Error& e) {
      free_host_rows(rows);
      throw Error(std::string("row copy-back failed: ") + e.what());
    }
  }
  *rows_out = rows;
  *n_out = count;
}

uint64_t read_u64(const uint64_t* device, hipStream_t stream) {
  uint64_t v = 0;
  RTDL_HIP(hipMemcpyAsync(&v, device, sizeof(v), hipMemcpyDeviceToHost, stream));
  RTDL_HIP(hipStreamSynchronize(stream));
  return v;
}

// Emit order of neighbour rows: the reference sorts rows stably by query id
// (reference.py:375), so queries are emitted in (id, input index) order.
DeviceArray<uint32_t> emit_order(const uint32_t* qids, uint32_t nq, hipStream_t stream) {
  DeviceArray<uint32_t> order;
  if (nq <= 1 || is_nondecreasing(qids, nq, stream)) return order;
  order.resize(nq);
  DeviceArray<uint64_t> keys(nq);
  keys_from_u32(qids, nullptr, nq, keys.get(), stream);
  iota(order.get(), nq, stream);
  radix_sort_pairs(keys.get(), order.get(), nq, 32, stream);
  return order;
}

DeviceArray<uint32_t> traversal_order(const double* q, int stride, uint32_t n, const Prepared& p,
                                      hipStream_t stream) {
  DeviceArray<uint32_t> order;
  if (n < kSortQueriesAbove) return order;
  order.resize(n);
  query_morton_order(q, stride, n, p.root_lo, p.root_hi, order.get(), stream);
  return order;
}

}  // namespace

void prepare_points(Prepared& p, const void* raw, size_t n, int dim) {
  require(n == 0 || raw != nullptr, "search point pointer must not be null when count is nonzero");
  require(n <= kMaxLeafStart, "too many search points");
  hipStream_t s = p.stream;
  p.type = dim == 3 ? SceneType::Points3D : SceneType::Points2D;
  auto bytes = upload_bytes(raw, n * (dim == 3 ? 32 : 24), s);
  DeviceArray<double> coords(3 * n);
  DeviceArray<uint32_t> ids(n);
  unpack_points(bytes.get(), n, dim, coords.get(), ids.get(), s);
  build_scene(p.scene, PrimKind::Points, dim, n, coords.get(), ids.get(), nullptr, nullptr, 0,
              {0, (uint64_t)n}, current_tuning().point_leaf, s);
  read_root_box(p);
}

void prepare_triangles(Prepared& p, const void* raw, size_t n) {
  require(n == 0 || raw != nullptr, "triangle pointer must not be null when count is nonzero");
  require(n <= kMaxLeafStart, "too many triangles");
  hipStream_t s = p.stream;
  p.type = SceneType::Triangles3D;
  auto bytes = upload_bytes(raw, n * 76, s);
  DeviceArray<double> coords(9 * n);
  DeviceArray<uint32_t> ids(n);
  unpack_triangles(bytes.get(), n, coords.get(), ids.get(), s);
  build_scene(p.scene, PrimKind::Triangles, 3, n, coords.get(), ids.get(), nullptr, nullptr, 0,
              {0, (uint64_t)n}, current_tuning().triangle_leaf, s);
  read_root_box(p);
}

void prepare_triangles_2d(Prepared& p, const void* raw, size_t n) {
  require(n == 0 || raw != nullptr, "triangle pointer must not be null when count is nonzero");
  require(n <= kMaxLeafStart, "too many triangles");
  hipStream_t s = p.stream;
  p.type = SceneType::Triangles2D;
  auto bytes = upload_bytes(raw, n * 56, s);
  DeviceArray<double> coords(9 * n);
  DeviceArray<uint32_t> ids(n);
  unpack_triangles_2d(bytes.get(), n, coords.get(), ids.get(), s);
  build_scene(p.scene, PrimKind::Triangles, 3, n, coords.get(), ids.get(), nullptr, nullptr, 0,
              {0, (uint64_t)n}, current_tuning().triangle_leaf, s);
  read_root_box(p);
}

void prepare_polygons(Prepared& p, const void* refs, size_t n, const double* verts_xy,
                      size_t n_doubles) {
  require(n == 0 || refs != nullptr, "polygon pointer must not be null when count is nonzero");
  require(n_doubles == 0 || verts_xy != nullptr,
          "vertex pointer must not be null when vertex count is nonzero");
  require(n_doubles % 2 == 0, "vertex_xy_count must be even (x, y pairs)");
  require(n <= kMaxLeafStart, "too many polygons");
  const size_t n_verts = n_doubles / 2;
  // validate the references on the host before anything reaches the device
  const uint8_t* r = static_cast<const uint8_t*>(refs);
  for (size_t i = 0; i < n; ++i) {
    uint32_t off, cnt;
    std::memcpy(&off, r + 12 * i + 4, 4);
    std::memcpy(&cnt, r + 12 * i + 8, 4);
    require(cnt >= 3, "polygon " + std::to_string(i) + " needs at least 3 vertices");
    require((uint64_t)off + cnt <= n_verts,
            "polygon " + std::to_string(i) + " vertex range exceeds vertex_xy_count");
  }
  hipStream_t s = p.stream;
  p.type = SceneType::Polygons;
  p.n_polygons = n;
  auto bytes = upload_bytes(refs, n * 12, s);
  p.refs_input.resize(n ? n : 1);
  unpack_polygons(bytes.get(), n, p.refs_input.get(), s);
  auto verts = upload(verts_xy, n_doubles, s);
  build_scene(p.scene, PrimKind::Polygons, 2, n, nullptr, nullptr, p.refs_input.get(), verts.get(),
              n_verts, {0, (uint64_t)n}, current_tuning().polygon_leaf, s);
  read_root_box(p);
}

void prepare_point_instances(Prepared& p, const float* coords, size_t n, uint32_t leaf_size,
                             const std::vector<uint64_t>& offsets) {
  require(n == 0 || coords != nullptr, "coords must not be null when count is nonzero");
  require(leaf_size <= kMaxLeafSize, "leaf_size must be in [1, 8], or 0 for the tuned default");
  require(offsets.size() >= 2 && offsets.front() == 0 && offsets.back() == n,
          "instance offsets must start at 0 and end at the point count");
  hipStream_t s = p.stream;
  p.type = SceneType::Points3D;
  auto d_offs = upload(offsets.data(), offsets.size(), s);
  DeviceArray<double> c64(3 * n);
  DeviceArray<uint32_t> ids(n);
  widen_points(coords, n, d_offs.get(), (uint32_t)(offsets.size() - 1), c64.get(), ids.get(), s);
  build_scene(p.scene, PrimKind::Points, 3, n, c64.get(), ids.get(), nullptr, nullptr, 0, offsets,
              leaf_size ? leaf_size : current_tuning().point_leaf, s);
  read_root_box(p);
  p.instances = true;
}

void prepare_triangle_instances(Prepared& p, const float* coords, size_t n,
                                const std::vector<uint64_t>& offsets) {
  require(n == 0 || coords != nullptr, "coords must not be null when count is nonzero");
  require(offsets.size() >= 2 && offsets.front() == 0 && offsets.back() == n,
          "instance offsets must start at 0 and end at the triangle count");
  hipStream_t s = p.stream;
  p.type = SceneType::Triangles3D;
  auto d_offs = upload(offsets.data(), offsets.size(), s);
  DeviceArray<double> c64(9 * n);
  DeviceArray<uint32_t> ids(n);
  widen_f32(coords, 9 * n, c64.get(), s);
  instance_ids(n, d_offs.get(), (uint32_t)(offsets.size() - 1), ids.get(), s);
  build_scene(p.scene, PrimKind::Triangles, 3, n, c64.get(), ids.get(), nullptr, nullptr, 0, offsets,
              current_tuning().triangle_leaf, s);
  read_root_box(p);
  p.instances = true;
}

void prepare_polygon_instances(Prepared& p, const uint32_t* voff, const uint32_t* vcnt, size_t n,
                               const float* verts_xy, size_t n_verts,
                               const std::vector<uint64_t>& offsets) {
  require(n == 0 || (voff != nullptr && vcnt != nullptr),
          "polygon arrays must not be null when count is nonzero");
  require(n_verts == 0 || verts_xy != nullptr, "vertex array must not be null");
  require(offsets.size() >= 2 && offsets.front() == 0 && offsets.back() == n,
          "instance offsets must start at 0 and end at the polygon count");
  hipStream_t s = p.stream;
  std::vector<uint32_t> h_off(n), h_cnt(n);
  if (n) {
    RTDL_HIP(hipMemcpyAsync(h_off.data(), voff, n * 4, hipMemcpyDeviceToHost, s));
    RTDL_HIP(hipMemcpyAsync(h_cnt.data(), vcnt, n * 4, hipMemcpyDeviceToHost, s));
    RTDL_HIP(hipStreamSynchronize(s));
  }
  for (size_t i = 0; i < n; ++i) {
    require(h_cnt[i] >= 3, "polygon " + std::to_string(i) + " needs at least 3 vertices");
    require((uint64_t)h_off[i] + h_cnt[i] <= n_verts,
            "polygon " + std::to_string(i) + " vertex range exceeds the vertex count");
  }
  p.type = SceneType::Polygons;
  p.n_polygons = n;
  auto d_offs = upload(offsets.data(), offsets.size(), s);
  p.refs_input.resize(n ? n : 1);
  widen_polygons(voff, vcnt, n, d_offs.get(), (uint32_t)(offsets.size() - 1), p.refs_input.get(), s);
  DeviceArray<double> v64(2 * n_verts);
  widen_f32(verts_xy, 2 * n_verts, v64.get(), s);
  build_scene(p.scene, PrimKind::Polygons, 2, n, nullptr, nullptr, p.refs_input.get(), v64.get(),
              n_verts, offsets, current_tuning().polygon_leaf, s);
  read_root_box(p);
}

// -------------------------------------------------------------- neighbours --

void run_neighbors(Prepared& p, const void* raw_queries, size_t nq_in, NeighborMode mode,
                   double radius, uint32_t k, void** rows_out, size_t* n_out) {
  require(p.type == SceneType::Points2D || p.type == SceneType::Points3D,
          "prepared handle is not a point scene");
  require(nq_in == 0 || raw_queries != nullptr,
          "query pointer must not be null when query count is nonzero");
  require(nq_in < 0xFFFFFFFFull, "too many queries");
  require(k >= 1, "k must be positive");
  require(!(radius < 0.0), "radius must be non-negative");
  // Without a finite radius every scene point is a candidate of every query
  // beyond the 64-entry list, so an unbounded search keeps k <= 64 (as the
  // reference's HIPRT backend does, src/rtdsl/hiprt_runtime.py:1307-1308).
  const bool unbounded = mode == NeighborMode::Knn || !(radius < __builtin_inf());
  require(!unbounded || k <= kMaxListK,
          mode == NeighborMode::FixedRadius
              ? "fixed_radius_neighbors with an infinite radius supports k_max <= 64"
              : "knn_rows supports k <= 64 (bounded_knn_rows with a finite radius takes larger k_max)");
  const bool rank_rows = mode != NeighborMode::FixedRadius;
  const size_t row_bytes = rank_rows ? 24 : 16;
  PhaseClock& clock = thread_clock();
  const uint32_t nq = (uint32_t)nq_in;
  hipStream_t s = p.stream;
  if (nq == 0 || p.scene.n_prims == 0) {
    *rows_out = host_rows(0, row_bytes);
    *n_out = 0;
    return;
  }
  const int dim = p.type == SceneType::Points3D ? 3 : 2;
  auto bytes = upload_bytes(raw_queries, (size_t)nq * (dim == 3 ? 32 : 24), s);
  DeviceArray<double> q(3 * (size_t)nq);
  DeviceArray<uint32_t> qid(nq);
  unpack_points(bytes.get(), nq, dim, q.get(), qid.get(), s);
  auto order = traversal_order(q.get(), 3, nq, p, s);
  RTDL_HIP(hipStreamSynchronize(s));
  clock.lap(0);

  const double search_radius = mode == NeighborMode::Knn ? __builtin_inf() : radius;
  const SceneView view = p.scene.view(0);
  DeviceArray<uint32_t> count(nq);
  DeviceArray<double> cand_d;
  DeviceArray<uint32_t> cand_id;
  DeviceArray<uint64_t> src_start;  // only for the count-then-collect path
  // Temporaries that kernels or copies still read after the branch that
  // fills them: they live until the stream is synchronised below, since a
  // block returned to the pool can be handed to another thread at once.
  DeviceArray<uint32_t> amb_list, amb_count;
  DeviceArray<double> list_d;
  DeviceArray<uint32_t> list_id;
  std::vector<uint64_t> host_start;
  std::vector<uint32_t> host_count;
  uint32_t stride = k;
  const bool list_path = k <= kMaxListK;
  if (list_path) {
    cand_d.resize((size_t)nq * k);
    cand_id.resize((size_t)nq * k);
    amb_list.resize(nq);
    amb_count.resize(1);
    RTDL_HIP(hipMemsetAsync(amb_count.get(), 0, sizeof(uint32_t), s));
    unsigned long long* ctr = walk_begin(s);
    launch_collect_packed(view, q.get(), order.get(), nq, search_radius, k, cand_d.get(),
                          cand_id.get(), count.get(), amb_list.get(), amb_count.get(), ctr,
                          256, s);
    uint32_t n_amb = 0;
    RTDL_HIP(hipMemcpyAsync(&n_amb, amb_count.get(), sizeof(n_amb), hipMemcpyDeviceToHost, s));
    RTDL_HIP(hipStreamSynchronize(s));
    p.last_ambiguous = n_amb;
    // queries whose FP32 selection keys tie at the boundary are recomputed
    // with the exact FP64 list (same slots, same counts)
    launch_collect_exact(view, q.get(), amb_list.get(), n_amb, search_radius, k, false,
                         cand_d.get(), cand_id.get(), count.get(), nullptr, /*seeded=*/true, s);
    walk_end(s);
    RTDL_HIP(hipStreamSynchronize(s));
    walk_collect();
  } else {
    // Keep the 64 nearest per query and count everything inside the radius;
    // queries with more than 64 candidates are collected in full and sorted.
    stride = kMaxListK;
    DeviceArray<uint32_t> total(nq);  // read back (synchronised) before the branch ends
    list_d.resize((size_t)nq * kMaxListK);
    list_id.resize((size_t)nq * kMaxListK);
    launch_collect_exact(view, q.get(), order.get(), nq, search_radius, kMaxListK, true,
                         list_d.get(), list_id.get(), count.get(), total.get(), /*seeded=*/false, s);
    std::vector<uint32_t> host_total(nq);
    RTDL_HIP(hipMemcpyAsync(host_total.data(), total.get(), nq * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s));
    RTDL_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> heavy;
    std::vector<uint64_t> heavy_off(1, 0);
    for (uint32_t i = 0; i < nq; ++i) {
      if (host_total[i] > kMaxListK) {
        heavy.push_back(i);
        heavy_off.push_back(heavy_off.back() + host_total[i]);
      }
    }
    const uint64_t heavy_n = heavy_off.back();
    require(heavy_n < 0xFFFFFFFFull, "too many fixed-radius candidates for one call");
    const size_t light_n = (size_t)nq * kMaxListK;
    cand_d.resize(light_n + heavy_n);
    cand_id.resize(light_n + heavy_n);
    RTDL_HIP(hipMemcpyAsync(cand_d.get(), list_d.get(), light_n * sizeof(double),
                            hipMemcpyDeviceToDevice, s));
    RTDL_HIP(hipMemcpyAsync(cand_id.get(), list_id.get(), light_n * sizeof(uint32_t),
                            hipMemcpyDeviceToDevice, s));
    host_start.resize(nq);
    host_count.resize(nq);
    for (uint32_t i = 0; i < nq; ++i) {
      host_start[i] = (uint64_t)i * kMaxListK;
      host_count[i] = std::min(host_total[i], k);
    }
    if (!heavy.empty()) {
      auto d_heavy = upload(heavy.data(), heavy.size(), s);
      auto d_off = upload(heavy_off.data(), heavy_off.size(), s);
      DeviceArray<double> hd(heavy_n), hd_sorted(heavy_n);
      DeviceArray<uint32_t> hid(heavy_n), hid_sorted(heavy_n);
      launch_collect_all(view, q.get(), d_heavy.get(), (uint32_t)heavy.size(), search_radius,
                         d_off.get(), hd.get(), hid.get(), s);
      // stable LSD sort of each segment by (distance, id): sort by id, then
      // by distance, then by segment
      DeviceArray<uint64_t> keys(heavy_n);
      DeviceArray<uint32_t> perm(heavy_n), seg(heavy_n);
      iota(perm.get(), heavy_n, s);
      keys_from_u32(hid.get(), nullptr, heavy_n, keys.get(), s);
      radix_sort_pairs(keys.get(), perm.get(), heavy_n, 32, s);
      keys_from_f64(hd.get(), perm.get(), heavy_n, keys.get(), s);
      radix_sort_pairs(keys.get(), perm.get(), heavy_n, 64, s);
      segment_ids(d_off.get(), (uint32_t)heavy.size(), heavy_n, seg.get(), s);
      keys_from_u32(seg.get(), perm.get(), heavy_n, keys.get(), s);
      radix_sort_pairs(keys.get(), perm.get(), heavy_n, 32, s);
      permute_f64(hd.get(), perm.get(), heavy_n, cand_d.get() + light_n, s);
      permute_u32(hid.get(), perm.get(), heavy_n, cand_id.get() + light_n, s);
      for (size_t h = 0; h < heavy.size(); ++h) host_start[heavy[h]] = light_n + heavy_off[h];
      RTDL_HIP(hipStreamSynchronize(s));
    }
    src_start = upload(host_start.data(), nq, s);
    RTDL_HIP(hipMemcpyAsync(count.get(), host_count.data(), nq * sizeof(uint32_t),
                            hipMemcpyHostToDevice, s));
  }
  RTDL_HIP(hipStreamSynchronize(s));
  clock.lap(2);

  auto eorder = emit_order(qid.get(), nq, s);
  DeviceArray<uint32_t> counts_pos(nq);
  DeviceArray<uint64_t> offsets((size_t)nq + 1);
  gather_counts(count.get(), eorder.get(), nq, 0, counts_pos.get(), s);
  exclusive_scan(counts_pos.get(), offsets.get(), nq, s);
  const uint64_t total_rows = read_u64(offsets.get() + nq, s);
  DeviceArray<uint8_t> rows(total_rows * row_bytes);
  emit_neighbor_rows(total_rows, offsets.get(), nq, eorder.get(), src_start.get(), stride,
                     cand_d.get(), cand_id.get(), qid.get(), rank_rows, rows.get(), s);
  RTDL_HIP(hipStreamSynchronize(s));
  clock.lap(3);
  copy_rows_back(rows.get(), total_rows, row_bytes, rows_out, n_out, s);
  clock.lap(4);
}

// -------------------------------------------------------------------- rays --

// dim 3: 60-byte RtdlRay3D records against a Triangles3D scene; dim 2:
// 44-byte RtdlRay2D records against a Triangles2D scene.
static void run_rays_dim(Prepared& p, const void* raw_rays, size_t nr_in, int mode, int dim,
                         void** rows_out, size_t* n_out) {
  require(p.type == (dim == 2 ? SceneType::Triangles2D : SceneType::Triangles3D),
          dim == 2 ? "prepared handle is not a 2-D triangle scene" : "prepared handle is not a triangle scene");
  require(mode >= 0 && mode <= 2, "unknown ray mode");
  require(nr_in == 0 || raw_rays != nullptr, "ray pointer must not be null when ray count is nonzero");
  require(nr_in < 0xFFFFFFFFull, "too many rays");
  const size_t row_bytes = mode == 2 ? 16 : 8;
  PhaseClock& clock = thread_clock();
  const uint32_t nr = (uint32_t)nr_in;
  hipStream_t s = p.stream;
  if (nr == 0) {
    *rows_out = host_rows(0, row_bytes);
    *n_out = 0;
    return;
  }
  auto bytes = upload_bytes(raw_rays, (size_t)nr * (dim == 2 ? 44 : 60), s);
  DeviceArray<double> rays(7 * (size_t)nr);
  DeviceArray<uint32_t> rid(nr);
  if (dim == 2)
    unpack_rays_2d(bytes.get(), nr, rays.get(), rid.get(), s);
  else
    unpack_rays(bytes.get(), nr, rays.get(), rid.get(), s);
  auto order = traversal_order(rays.get(), 7, nr, p, s);
  RTDL_HIP(hipStreamSynchronize(s));
  clock.lap(0);
  DeviceArray<uint32_t> out_u(nr);
  DeviceArray<double> out_t(mode == 2 ? nr : 0);
  unsigned long long* ctr = walk_begin(s);
  launch_rays(p.scene.view(0), rays.get(), order.get(), nr, dim == 2 ? mode + 4 : mode, out_u.get(),
              out_t.get(), ctr, 256, s);
  walk_end(s);
  RTDL_HIP(hipStreamSynchronize(s));
  walk_collect();
  clock.lap(2);
  if (mode != 2) {
    DeviceArray<uint8_t> rows((size_t)nr * 8);
    emit_pair_rows(nr, rid.get(), out_u.get(), rows.get(), s);
    RTDL_HIP(hipStreamSynchronize(s));
    clock.lap(3);
    copy_rows_back(rows.get(), nr, 8, rows_out, n_out, s);
  } else {
    DeviceArray<uint32_t> flags(nr);
    DeviceArray<uint64_t> offsets((size_t)nr + 1);
    hit_flags(nr, out_t.get(), flags.get(), s);
    exclusive_scan(flags.get(), offsets.get(), nr, s);
    const uint64_t total = read_u64(offsets.get() + nr, s);
    DeviceArray<uint8_t> rows(total * 16);
    emit_closest_rows(nr, offsets.get(), rid.get(), out_u.get(), out_t.get(), rows.get(), s);
    RTDL_HIP(hipStreamSynchronize(s));
    clock.lap(3);
    copy_rows_back(rows.get(), total, 16, rows_out, n_out, s);
  }
  clock.lap(4);
}

void run_rays_2d(Prepared& p, const void* raw_rays, size_t nr, int mode, void** rows_out, size_t* n_out) {
  run_rays_dim(p, raw_rays, nr, mode, 2, rows_out, n_out);
}

void run_rays(Prepared& p, const void* raw_rays, size_t nr, int mode, void** rows_out, size_t* n_out) {
  run_rays_dim(p, raw_rays, nr, mode, 3, rows_out, n_out);
}

// --------------------------------------------------------------------- PIP --

void run_pip(Prepared& p, const void* raw_points, size_t np_in, bool positive_only,
             void** rows_out, size_t* n_out) {
  require(p.type == SceneType::Polygons, "prepared handle is not a polygon scene");
  require(np_in == 0 || raw_points != nullptr,
          "point pointer must not be null when point count is nonzero");
  require(np_in < 0xFFFFFFFFull, "too many points");
  PhaseClock& clock = thread_clock();
  const uint32_t np = (uint32_t)np_in;
  const uint32_t ng = (uint32_t)p.n_polygons;
  hipStream_t s = p.stream;
  if (np == 0 || ng == 0) {
    *rows_out = host_rows(0, 12);
    *n_out = 0;
    return;
  }
  require(positive_only || (uint64_t)np * ng < (uint64_t(1) << 36),
          "full_matrix PIP output exceeds 2^36 rows");
  auto bytes = upload_bytes(raw_points, (size_t)np * 24, s);
  DeviceArray<double> xy(2 * (size_t)np);
  DeviceArray<uint32_t> pid(np);
  unpack_points_xy(bytes.get(), np, xy.get(), pid.get(), s);
  auto order = traversal_order(xy.get(), 2, np, p, s);
  RTDL_HIP(hipStreamSynchronize(s));
  clock.lap(0);
  const double eps_sq = std::pow(1.0e-12, 2.0);  // reference.py:757 `eps ** 2`
  DeviceArray<uint32_t> count(np);
  DeviceArray<uint64_t> offsets((size_t)np + 1);
  const SceneView view = p.scene.view(0);
  // the measured walk is the counting pass; the writing pass repeats it
  unsigned long long* ctr = walk_begin(s);
  launch_pip(view, xy.get(), order.get(), np, positive_only, eps_sq, count.get(), nullptr, nullptr,
             ctr, 256, s);
  walk_end(s);
  exclusive_scan(count.get(), offsets.get(), np, s);
  const uint64_t total = read_u64(offsets.get() + np, s);
  walk_collect();
  DeviceArray<uint32_t> poly_idx(total ? total : 1);
  launch_pip(view, xy.get(), order.get(), np, positive_only, eps_sq, count.get(), offsets.get(),
             poly_idx.get(), nullptr, 256, s);
  RTDL_HIP(hipStreamSynchronize(s));
  clock.lap(2);
  if (positive_only) {
    DeviceArray<uint8_t> rows(total * 12);
    emit_pip_positive(total, offsets.get(), np, pid.get(), poly_idx.get(), p.refs_input.get(),
                      rows.get(), s);
    RTDL_HIP(hipStreamSynchronize(s));
    clock.lap(3);
    copy_rows_back(rows.get(), total, 12, rows_out, n_out, s);
  } else {
    const uint64_t cells = (uint64_t)np * ng;
    DeviceArray<uint8_t> rows(cells * 12);
    emit_pip_full(np, ng, offsets.get(), pid.get(), poly_idx.get(), p.refs_input.get(), rows.get(),
                  s);
    RTDL_HIP(hipStreamSynchronize(s));
    clock.lap(3);
    copy_rows_back(rows.get(), cells, 12, rows_out, n_out, s);
  }
  clock.lap(4);
}

}  // namespace rtdl
