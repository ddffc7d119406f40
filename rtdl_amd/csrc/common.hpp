// Shared host/device infrastructure of librtdl_mi355x: error handling,
// device buffers, the BVH node format and the conservative FP32 bounds.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>

namespace rtdl {

// ---------------------------------------------------------------- errors --

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define RTDL_HIP(call)                                                              \
  do {                                                                              \
    hipError_t rtdl_err_ = (call);                                                  \
    if (rtdl_err_ != hipSuccess) {                                                  \
      throw ::rtdl::Error(std::string(#call) + " failed: " + hipGetErrorString(rtdl_err_) + \
                          " (" __FILE__ ":" + std::to_string(__LINE__) + ")");      \
    }                                                                               \
  } while (0)

// Host-forced faults for testing failure handling (RTDL_MI355X_FAULT, or
// rtdl_mi355x_set_fault): "alloc", "launch" or "timeout", optionally ":N" for
// the N-th occurrence (default 1) counted from the moment the fault is armed.
// fault_fires(kind) is true exactly once, at that occurrence; the failure
// is raised on the host, so no kernel ever faults.
enum class FaultKind : int { None = 0, Alloc = 1, Launch = 2, Timeout = 3 };
bool fault_fires(FaultKind kind);
void set_fault(const char* spec);  // "" disarms; throws on an unknown spec

#define RTDL_LAUNCH_CHECK()                                                               \
  do {                                                                                    \
    if (::rtdl::fault_fires(::rtdl::FaultKind::Launch))                                  \
      throw ::rtdl::Error("injected fault (RTDL_MI355X_FAULT=launch): kernel launch failed"); \
    RTDL_HIP(hipGetLastError());                                                          \
  } while (0)

// A literal message is only turned into a string when the check fails.
inline void require(bool ok, const char* message) {
  if (!ok) throw Error(message);
}
inline void require(bool ok, const std::string& message) {
  if (!ok) throw Error(message);
}

// ------------------------------------------------------- device buffers --

// Device allocations go through a process-wide caching pool so that the
// per-call path does not pay hipMalloc/hipFree for every temporary.
void* pool_alloc(size_t bytes);
void pool_free(void* ptr);
void pool_trim();

template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  ~DeviceArray() { reset(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : ptr_(o.ptr_), n_(o.n_), owned_(o.owned_) {
    o.ptr_ = nullptr;
    o.n_ = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      reset();
      ptr_ = o.ptr_;
      n_ = o.n_;
      owned_ = o.owned_;
      o.ptr_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n) {
    require(owned_ || !ptr_, "a borrowed device array cannot be resized");
    reset();
    if (n) ptr_ = static_cast<T*>(pool_alloc(n * sizeof(T)));
    n_ = n;
  }
  // Views n elements of memory the array does not own (never freed here).
  void borrow(T* p, size_t n) {
    reset();
    ptr_ = p;
    n_ = n;
    owned_ = false;
  }
  void reset() {
    if (ptr_ && owned_) pool_free(ptr_);
    ptr_ = nullptr;
    n_ = 0;
    owned_ = true;
  }
  T* get() const { return ptr_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* ptr_ = nullptr;
  size_t n_ = 0;
  bool owned_ = true;
};

// Carves 256-byte-aligned arrays out of one block in order; with a null base
// it only measures (the bytes a set of arrays needs).
struct Carver {
  uint8_t* base = nullptr;
  size_t capacity = 0;
  size_t used = 0;
  template <class T>
  T* take(size_t n) {
    used = (used + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
    used += n * sizeof(T);
    require(!base || used <= capacity, "carved arrays exceed their block");
    return p;
  }
  template <class T>
  void take(DeviceArray<T>& a, size_t n) {
    T* p = take<T>(n);
    if (base) a.borrow(p, n);
  }
};

// ------------------------------------------------------- BVH node format --

// One BVH node is 32 bytes: an FP32 box and two links.  Nodes are visited in
// depth-first preorder without a stack: on a box hit an internal node
// continues at its left child (`link`), a leaf tests its primitives; on a
// miss, or after a leaf, traversal continues at `escape` (the next node in
// preorder once this subtree is skipped).
//   internal node: link = node index of the left child
//   leaf node:     link = kLeafFlag | (count - 1) << 28 | first primitive
constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr uint32_t kEndNode = 0xFFFFFFFFu;
constexpr uint32_t kMaxLeafSize = 8;
constexpr uint32_t kMaxLeafStart = (1u << 28) - 1;

// Process-wide launch and build choices (set by the autotuner or the caller;
// every value gives the same rows).
struct Tuning {
  uint32_t point_leaf = 2;      // primitives per LBVH leaf, point scenes
  uint32_t triangle_leaf = 1;   // triangle scenes
  uint32_t polygon_leaf = 1;    // polygon scenes
  uint32_t blocks_per_cu = 8;   // persistent batch launches: 256-thread workgroups per CU
  uint32_t knn_stack = 32;      // batched KNN walk: per-lane node stack entries (16, 32 or 48)
  uint32_t walk_block = 256;    // batched ray and PIP walks: threads per workgroup (64, 128 or 256)
};
Tuning current_tuning();
void set_tuning(const Tuning& t);  // validates; throws on out-of-range values

// A tree node as stored, in depth-first preorder: the left child of an
// internal node is the next record, `link` holds its right child (so a walk
// can fetch both children at once); a leaf's `link` holds the leaf flag, its
// point count - 1 (3 bits) and its first slot.  `escape` is the node a walk
// continues with after this node's subtree.
struct alignas(16) BvhNode {
  float lo[3];
  uint32_t link;
  float hi[3];
  uint32_t escape;
};
static_assert(sizeof(BvhNode) == 32, "BvhNode must be 32 bytes");

__host__ __device__ inline bool node_is_leaf(uint32_t link) { return (link & kLeafFlag) != 0; }
__host__ __device__ inline uint32_t leaf_first(uint32_t link) { return link & kMaxLeafStart; }
__host__ __device__ inline uint32_t leaf_count(uint32_t link) { return ((link >> 28) & 7u) + 1u; }

// Compact copy of a point scene's nodes for the fixed-radius count walk, the
// hottest rope walk: 16 bytes per node (one load instead of two).  The box
// is quantised to 16-bit steps of a per-scene frame (origin o and power-of-two
// step s per axis, chosen so that o + q * s is exact in FP32 for every q of
// the box), rounded outward, so the dequantised box contains the node's FP32
// box.  The word is a leaf's link, or an internal node's escape (the node
// count stands for the end of the walk).  A leaf needs no escape: in preorder
// the node after a leaf is the next record.  The frame holds o[3], s[3] and,
// in frame[6], 1 when a scene box is not finite (no compact walk then).
constexpr uint32_t kQFrameFloats = 8;
__host__ __device__ inline uint32_t qnode_word_of(const BvhNode& nd, uint32_t n_nodes) {
  return node_is_leaf(nd.link) ? nd.link : (nd.escape == kEndNode ? n_nodes : nd.escape);
}

// ---------------------------------------------- conservative FP32 bounds --
//
// Traversal and primitive prefilters run in FP32; the final predicate is
// always decided in FP64.  For a query q (FP64) and a search distance rho,
// with M >= every |coordinate| of the scene and the query, the FP32 squared
// distance of any point or box within true distance rho of q is at most
//   thr2(rho) = ((rho + 2^-21 * M) * (1 + 2^-18))^2,
// since each FP32 coordinate difference is within 4 * 2^-24 * M of the exact
// one (two roundings of the inputs and one of the difference), the Euclidean
// error is at most sqrt(3) * 4 * 2^-24 * M < 2^-21 * M, and the three
// squarings and two additions add at most 4 * 2^-24 relative error.  The
// (1 + 2^-18) factor also covers the FP64 evaluation of the bound itself.
__device__ inline float frn_threshold2(double rho, float m) {
  if (!(rho < 3.0e38)) return __builtin_inff();
  const double slack = (double)m * 4.76837158203125e-07;  // 2^-21
  const double t = (rho + slack) * 1.000003814697265625;  // 1 + 2^-18
  return __double2float_ru(t * t);
}

// M for one query: the scene bound and |q|, each rounded up.
__device__ inline float coord_bound(float scene_m, double x, double y, double z) {
  float m = scene_m;
  m = fmaxf(m, __double2float_ru(fabs(x)));
  m = fmaxf(m, __double2float_ru(fabs(y)));
  m = fmaxf(m, __double2float_ru(fabs(z)));
  return m;
}

__device__ inline float box_dist2(const float q[3], const float lo[3], const float hi[3]) {
  float d2 = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float g = fmaxf(fmaxf(lo[a] - q[a], q[a] - hi[a]), 0.0f);
    d2 = __builtin_fmaf(g, g, d2);
  }
  return d2;
}

// Strict (distance, id) order of the reference's candidate sort.
__host__ __device__ inline bool key_less(double d0, uint32_t i0, double d1, uint32_t i1) {
  return d0 < d1 || (d0 == d1 && i0 < i1);
}

// ----------------------------------------------------------- misc helpers --

inline unsigned blocks_for(size_t n, unsigned block) {
  return static_cast<unsigned>((n + block - 1) / block);
}

}  // namespace rtdl
