// Device-resident scenes (LBVH over points, triangles or polygons) and the
// POD view the traversal kernels receive.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "common.hpp"

namespace rtdl {

struct RebuildWork;  // the build buffers a scene keeps for in-place rebuilds

enum class PrimKind : int { Points = 0, Triangles = 1, Polygons = 2 };

// What a traversal kernel needs to know about a scene.  Primitives are
// stored in Morton order ("slots").  For points `prims` holds one float4
// (x, y, z, id bits) per slot; for triangles three float4 per slot
// ((v0, id), (v1, 0), (v2, 0)).  `exact` holds the FP64 coordinates per
// slot (3 per point, 9 per triangle) and is null when every coordinate is
// exactly representable in FP32, in which case the FP32 values are exact.
// Polygons keep FP64 vertices; `poly` holds (vertex offset, count, id, 0)
// per slot and `poly_box` the FP64 bounding box (x0, y0, x1, y1) per slot.
struct SceneView {
  const BvhNode* nodes;
  const float4* prims;
  const double* exact;
  const uint32_t* slot_to_input;
  const uint4* poly;
  const double* poly_box;
  const double* poly_verts;
  uint32_t root;
  uint32_t end;  // traversal stops when it reaches this node (escape of root)
  uint32_t n_prims;
  uint32_t n_nodes;
  float m;             // >= every |coordinate| in the scene, rounded up
  const float* m_dev;  // when set, M is read from device memory instead
  const uint4* qnodes;  // point scenes: compact nodes (16 bytes each) for the count walk
  const float* qframe;  // their quantisation frame (kQFrameFloats)
};

// A compact node (see qnode_word_of): its dequantised box and word.
struct QNode {
  float lo[3], hi[3];
  uint32_t word;
};

// Loads from scene arrays through the global address space.  The arrays
// always live in device memory; a pointer read from a job descriptor would
// otherwise be accessed with flat loads, which also count against the
// LDS / scalar wait counter.  A node is fetched as two 16-byte loads.
typedef unsigned int GlobalU32x4 __attribute__((ext_vector_type(4)));

__device__ inline GlobalU32x4 load_raw16(const void* p) {
  return *(const __attribute__((address_space(1))) GlobalU32x4*)p;
}

__device__ inline BvhNode load_node(const BvhNode* nodes, uint32_t i) {
  const GlobalU32x4 a = load_raw16(nodes + i);
  const GlobalU32x4 b = load_raw16(reinterpret_cast<const uint8_t*>(nodes + i) + 16);
  BvhNode n;
  n.lo[0] = __uint_as_float(a[0]);
  n.lo[1] = __uint_as_float(a[1]);
  n.lo[2] = __uint_as_float(a[2]);
  n.link = a[3];
  n.hi[0] = __uint_as_float(b[0]);
  n.hi[1] = __uint_as_float(b[1]);
  n.hi[2] = __uint_as_float(b[2]);
  n.escape = b[3];
  return n;
}

// A full node in the compact form (box as stored, word as qnode_word_of).
__device__ inline QNode qnode_of(const BvhNode& nd, uint32_t n_nodes) {
  QNode n;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    n.lo[ax] = nd.lo[ax];
    n.hi[ax] = nd.hi[ax];
  }
  n.word = qnode_word_of(nd, n_nodes);
  return n;
}

// Compact node i: halves (lo x, lo y), (lo z, hi x), (hi y, hi z) of three
// words, then the word; box = o + q * s per axis (exact in FP32).
__device__ inline QNode load_qnode(const uint4* qnodes, uint32_t i, const float o[3], const float s[3]) {
  const GlobalU32x4 a = load_raw16(qnodes + i);
  const uint32_t q[6] = {a[0] & 0xFFFFu, a[0] >> 16, a[1] & 0xFFFFu, a[1] >> 16, a[2] & 0xFFFFu, a[2] >> 16};
  QNode n;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    n.lo[ax] = o[ax] + (float)q[ax] * s[ax];
    n.hi[ax] = o[ax] + (float)q[3 + ax] * s[ax];
  }
  n.word = a[3];
  return n;
}

__device__ inline float4 load_prim(const float4* prims, size_t i) {
  const GlobalU32x4 v = load_raw16(prims + i);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}

__device__ inline uint4 load_u4(const uint4* a, size_t i) {
  const GlobalU32x4 v = load_raw16(a + i);
  return make_uint4(v[0], v[1], v[2], v[3]);
}

__device__ inline double load_f64(const double* a, size_t i) {
  return *(const __attribute__((address_space(1))) double*)(a + i);
}

// A set of instances built into one LBVH: instance i owns primitives
// [offsets[i], offsets[i+1]) of the input (contiguous) and gets its own
// subtree; `roots[i]` / `ends[i]` delimit its traversal.
class Scene {
 public:
  PrimKind kind = PrimKind::Points;
  int dim = 3;
  uint32_t leaf_size = 4;
  size_t n_prims = 0;
  size_t n_leaves = 0;
  size_t n_instances = 0;
  bool fp32_exact = true;
  float m = 0.0f;
  DeviceArray<BvhNode> nodes;
  DeviceArray<float4> prims;
  DeviceArray<double> exact;
  DeviceArray<uint32_t> slot_to_input;
  DeviceArray<uint4> poly;
  DeviceArray<double> poly_box;
  DeviceArray<double> poly_verts;
  DeviceArray<uint32_t> roots;  // per instance
  DeviceArray<uint32_t> ends;   // per instance
  std::vector<uint32_t> host_roots;
  std::vector<uint32_t> host_ends;
  std::vector<uint64_t> host_offsets;  // per instance, n_instances + 1
  // set by prepare_rebuild: the buffers rebuild_scene reuses, and M in
  // device memory (a rebuild changes it without the host seeing it)
  std::shared_ptr<RebuildWork> rebuild;
  const float* m_device = nullptr;
  // point scenes: the compact nodes of the fixed-radius count walk and their
  // quantisation frame, refreshed by every build and rebuild
  DeviceArray<uint4> qnodes;
  DeviceArray<float> qframe;

  SceneView view(size_t instance = 0) const {
    SceneView v{};
    v.nodes = nodes.get();
    v.prims = prims.get();
    v.exact = fp32_exact ? nullptr : exact.get();
    v.slot_to_input = slot_to_input.get();
    v.poly = poly.get();
    v.poly_box = poly_box.get();
    v.poly_verts = poly_verts.get();
    v.root = host_roots.empty() ? kEndNode : host_roots[instance];
    v.end = host_ends.empty() ? kEndNode : host_ends[instance];
    v.n_prims = (uint32_t)n_prims;
    v.n_nodes = (uint32_t)nodes.size();
    v.m = m;
    v.m_dev = m_device;
    v.qnodes = qnodes.size() ? qnodes.get() : nullptr;
    v.qframe = qnodes.size() ? qframe.get() : nullptr;
    return v;
  }
  size_t device_bytes() const {
    return nodes.bytes() + prims.bytes() + exact.bytes() + slot_to_input.bytes() + poly.bytes() +
           poly_box.bytes() + poly_verts.bytes() + roots.bytes() + ends.bytes() + qnodes.bytes() + qframe.bytes();
  }
};

// Canonical FP64 input already on the device:
//   points:    coords[3 * n] (x, y, z; z = 0 for 2-D), ids[n]
//   triangles: coords[9 * n], ids[n]
//   polygons:  refs[n] = (vertex offset, vertex count, id, 0), verts[2 * nv]
// `instance_offsets` (host, n_instances + 1 entries) splits the input into
// independently traversable instances; pass {0, n} for a single scene.
void build_scene(Scene& scene, PrimKind kind, int dim, size_t n, const double* coords,
                 const uint32_t* ids, const uint4* poly_refs, const double* poly_verts,
                 size_t n_verts, const std::vector<uint64_t>& instance_offsets,
                 uint32_t leaf_size, hipStream_t stream);

// In-place rebuilds of a point or triangle instance scene built from FP32
// input, for primitives that move while their count and instance layout stay
// (the tree is rebuilt from scratch, not refit).  prepare_rebuild allocates
// every buffer the rebuild needs for the scene's layout and switches the
// scene's M to device memory; views taken before it keep the host M and must
// not be used with rebuilt scenes.  rebuild_scene then enqueues only kernels
// on `stream` (no allocation, host copy, memset or synchronisation), so a
// stream capture can record it.  rebuild_status reads, after the stream has
// passed the rebuild, the error bits of the last one (0 when it succeeded).
constexpr uint32_t kRebuildRefitIncomplete = 1;  // the refit rounds did not reach the root
constexpr uint32_t kRebuildRootsMoved = 2;       // an instance root left its preorder position
constexpr uint32_t kRebuildNotExact = 4;         // a coordinate is not exact in FP32
void prepare_rebuild(Scene& scene, hipStream_t stream);
void rebuild_scene(Scene& scene, const float* coords, hipStream_t stream);
uint32_t rebuild_status(const Scene& scene, hipStream_t stream);

}  // namespace rtdl
