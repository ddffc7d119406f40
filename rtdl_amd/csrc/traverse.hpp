// Per-lane traversal + refine bodies shared by the per-call kernels and the
// batched / served persistent kernels.
//
// All traversals walk the rope-linked LBVH in preorder (no stack).  Boxes and
// primitives are prefiltered in FP32 against conservative bounds
// (frn_threshold2, segment_box); every surviving candidate is decided in FP64
// with the reference's operation order (the library is compiled with
// -ffp-contract=off).
#pragma once

#include "common.hpp"
#include "scene.hpp"

namespace rtdl {

struct LaneStats {
  unsigned long long nodes = 0;
  unsigned long long prims = 0;
  unsigned long long mark = 0;   // nodes at the start of the current work item
  unsigned long long steps = 0;  // lane 0: wavefront walk steps
};

// Ends a work item of the wavefront: its walk took at least as many lockstep
// steps as its busiest lane visited nodes, so lane 0 adds that maximum.  The
// ratio nodes / (64 * steps) is the walk's lane efficiency (kernels whose
// steps may visit more than one node count their steps themselves).  Every lane must
// take part.
__device__ inline void close_item(LaneStats& st) {
  unsigned long long m = st.nodes - st.mark;
  st.mark = st.nodes;
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(m, d, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) st.steps += m;
}

// Adds the wavefront's visits to stats[0..2] (nodes, primitives, walk steps).
// Every lane must take part.
__device__ inline void flush_lane_stats(unsigned long long* stats, LaneStats& st) {
  close_item(st);
  unsigned long long nodes = st.nodes, prims = st.prims;
  for (int d = 32; d >= 1; d >>= 1) {
    nodes += __shfl_xor(nodes, d, 64);
    prims += __shfl_xor(prims, d, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&stats[0], nodes);
    atomicAdd(&stats[1], prims);
    atomicAdd(&stats[2], st.steps);
  }
}

__device__ inline float scene_m(const SceneView& s) { return s.m_dev ? *s.m_dev : s.m; }

// One step of a rope (preorder + escape) walk that fetches the node together
// with its successor in preorder: a descent into an internal node continues at
// node + 1, whose record is then already at hand and is tested in the same
// step, so a chain of descents costs half the dependent fetches.  On return
// `node` / `nd` are the node the step ends on and `in` its test;
// test(record, index) -> bool.
template <bool STATS, class Test>
__device__ inline void rope_pair_step(const SceneView& s, uint32_t& node, BvhNode& nd, bool& in, LaneStats& st,
                                      Test&& test) {
  const uint32_t after = node + 1 < s.n_nodes ? node + 1 : node;
  nd = load_node(s.nodes, node);
  const BvhNode nd1 = load_node(s.nodes, after);
  if (STATS) ++st.nodes;
  in = test(nd, node);
  if (in && !node_is_leaf(nd.link)) {  // the left child is the next record
    node = after;
    nd = nd1;
    if (STATS) ++st.nodes;
    in = test(nd, node);
  }
}

// -------------------------------------------------------- point candidates --

// FP64 squared distance of slot `slot` from (qx, qy, qz), reference.py:882-890:
// dx = search - query, then dx*dx + dy*dy + dz*dz left to right.
__device__ inline double point_d2(const SceneView& s, uint32_t slot, const float4& p, double qx,
                                  double qy, double qz) {
  double px, py, pz;
  if (s.exact) {
    px = load_f64(s.exact, 3 * (size_t)slot);
    py = load_f64(s.exact, 3 * (size_t)slot + 1);
    pz = load_f64(s.exact, 3 * (size_t)slot + 2);
  } else {
    px = p.x;
    py = p.y;
    pz = p.z;
  }
  const double dx = px - qx, dy = py - qy, dz = pz - qz;
  return dx * dx + dy * dy + dz * dz;
}

// Visits every leaf primitive whose FP32 distance passes `thr2` (which the
// visitor may shrink).  visit(slot, p, d2f) -> void.
template <bool STATS, class Visit>
__device__ inline void walk_points(const SceneView& s, uint32_t root, uint32_t end, const float qf[3],
                                   float& thr2, LaneStats& st, Visit&& visit,
                                   uint32_t skip = kEndNode) {
  uint32_t node = root;
  // a preorder walk visits each node at most once; the guard only bounds the
  // loop should the tree ever be malformed
  for (uint32_t guard = s.n_nodes + 1; node != end && guard; --guard) {
    BvhNode nd;
    bool in;
    rope_pair_step<STATS>(s, node, nd, in, st, [&](const BvhNode& b, uint32_t i) {
      const float lo[3] = {b.lo[0], b.lo[1], b.lo[2]};
      const float hi[3] = {b.hi[0], b.hi[1], b.hi[2]};
      // `skip`: a subtree the caller has already walked
      return (i != skip) & (box_dist2(qf, lo, hi) <= thr2);
    });
    const bool leaf = node_is_leaf(nd.link);
    if (in && leaf) {  // the only branch of the step: the leaf's primitives
      const uint32_t first = leaf_first(nd.link), cnt = leaf_count(nd.link);
      for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t slot = first + j;
        const float4 p = load_prim(s.prims, slot);
        if (STATS) ++st.prims;
        const float fx = p.x - qf[0], fy = p.y - qf[1], fz = p.z - qf[2];
        const float d2f = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
        if (d2f <= thr2) visit(slot, p, d2f);
      }
    }
    // into an overlapped internal node's subtree (its left child, the next
    // record), else past the node
    node = (in && !leaf) ? node + 1 : nd.escape;
  }
}

// Neighbour walk in expanding rings around the query.  A greedy descent
// (nearer child box first; both children are fetched together, the left one
// being the next record and the right one the node's link) finds the query's
// leaf and its ancestors up[1..7].  The leaf
// is walked first, then each ancestor's subtree without the part already
// walked, then the rest of the tree: candidates arrive roughly nearest first,
// so a shrinking threshold (a full k list) prunes early, and every primitive
// is still offered to `visit` exactly once.
template <bool STATS, class Visit>
__device__ inline void walk_near_first(const SceneView& s, uint32_t root, uint32_t end,
                                       const float qf[3], float& thr2, LaneStats& st,
                                       Visit&& visit) {
  if (root == kEndNode) return;
  constexpr int kUp = 7;
  uint32_t up[kUp + 1];  // up[i]: i steps above the leaf; the root repeats above it
#pragma unroll
  for (int i = 0; i <= kUp; ++i) up[i] = root;
  BvhNode nd = load_node(s.nodes, root);
  uint32_t cur = root;
  if (STATS) ++st.nodes;
  for (uint32_t guard = s.n_nodes + 1; !node_is_leaf(nd.link) && guard; --guard) {
    const uint32_t l = cur + 1, r = nd.link;  // both children, fetched together
    const BvhNode ln = load_node(s.nodes, l);
    const BvhNode rn = load_node(s.nodes, r);
    if (STATS) st.nodes += 2;
    const bool right = box_dist2(qf, rn.lo, rn.hi) < box_dist2(qf, ln.lo, ln.hi);
#pragma unroll
    for (int i = kUp; i > 0; --i) up[i] = up[i - 1];
    cur = right ? r : l;
    up[0] = cur;
    nd = right ? rn : ln;
  }
  uint32_t done = kEndNode;  // root of the subtree walked so far
#pragma unroll 1
  for (int ring = 0; ring <= kUp + 1; ++ring) {
    uint32_t from = root;
    if (ring <= kUp) {
#pragma unroll
      for (int i = 0; i <= kUp; ++i)
        if (i == ring) from = up[i];
    }
    if (from == done) continue;
    const uint32_t stop = from == root ? end : load_node(s.nodes, from).escape;
    walk_points<STATS>(s, from, stop, qf, thr2, st, visit, done);
    if (from == root) break;
    done = from;
  }
}

// The walk a neighbour search uses.  With a radius, the radius bounds the walk
// and a plain preorder walk is cheapest; a search for the k nearest without a
// radius starts from the query's leaf and expands outwards, so its list fills
// with near points and prunes early.
template <bool STATS, class Visit>
__device__ inline void walk_neighbors(const SceneView& s, uint32_t root, uint32_t end,
                                      const float qf[3], double radius, float& thr2, LaneStats& st,
                                      Visit&& visit) {
  if (radius < __builtin_inf())
    walk_points<STATS>(s, root, end, qf, thr2, st, visit);
  else
    walk_near_first<STATS>(s, root, end, qf, thr2, st, visit);
}

// Default depth of the per-lane node stack of walk_nearest (the batched KNN
// walk takes the depth as a tuned template variant).
constexpr uint32_t kStackDepth = 32;

// Nearest-child-first walk for a k-nearest search: at an internal node both
// children's boxes are tested, the walk descends into the nearer one and
// pushes the farther one onto the lane's stack (LDS, entry d at
// stack[d * stride]); a popped node is tested again against the threshold,
// which has shrunk meanwhile.  Near primitives arrive first, so the list's
// k-th key prunes most of the tree.  Returns false when the stack (DEPTH
// entries) was full at some push and a subtree was left out: the caller then
// walks the whole tree again, meeting some primitives a second time.
// The walk is one flat loop in which every lane takes one action per
// iteration: expand an internal node (fetch both children: the left one is
// the next record, the right one the node's link), visit two points of a
// leaf, or test a node popped from the stack.  Each action fetches at most two
// 32-byte records, so the lanes of a wavefront issue their fetches in the
// same load instructions whatever each one is doing; nested descend and pop
// loops would instead leave most lanes idle while a few descend or pop.  The
// sequence of nodes and points of each lane is that of the nested form.
template <uint32_t DEPTH = kStackDepth>
struct NearestWalk {
  static constexpr uint32_t kExpand = 0, kLeaf = 1, kTest = 2, kDone = 3;
  uint32_t state, cur;
  uint32_t first, cnt, j;  // kLeaf: the leaf's points, next index j
  uint32_t right;          // kExpand: the right child
  uint32_t depth;          // entries on the lane's stack
  bool complete;           // no push was lost to a full stack

  __device__ inline void init(uint32_t root, uint32_t end) {
    state = (root == kEndNode || root == end) ? kDone : kTest;  // the root is tested first
    cur = root;
    first = cnt = j = right = depth = 0;
    complete = true;
  }
  __device__ inline bool done() const { return state == kDone; }

  // One action of the lane: expand, visit up to two points, or test a popped
  // node.  Returns false once the walk is over.
  template <bool STATS, class Visit>
  __device__ inline bool step(const SceneView& s, const float qf[3], float& thr2, LaneStats& st, Visit&& visit,
                              uint32_t* stack, uint32_t stride) {
    // the action's fetch: record A (a node, or points j and j + 1) and, when
    // expanding, record B (the right child)
    const uint8_t* pa;
    if (state == kLeaf)
      pa = reinterpret_cast<const uint8_t*>(s.prims + first + j);
    else
      pa = reinterpret_cast<const uint8_t*>(s.nodes + (state == kExpand ? cur + 1 : cur));
    // a leaf's odd last point: the second half of A would lie past the leaf
    const bool a_half = state != kLeaf || j + 1 < cnt;
    // the four loads are issued together, before any of them is waited for
    // (record B repeats A for a lane that does not expand), so an expansion
    // costs one round trip to memory, not two
    const uint8_t* pb = state == kExpand ? reinterpret_cast<const uint8_t*>(s.nodes + right) : pa;
    const GlobalU32x4 a0 = load_raw16(pa);
    const GlobalU32x4 a1 = load_raw16(a_half ? pa + 16 : pa);
    const GlobalU32x4 b0 = load_raw16(pb);
    const GlobalU32x4 b1 = load_raw16(a_half ? pb + 16 : pb);
    // an empty use of all four in the common path keeps the compiler from
    // sinking the B loads into the expanding branch behind a wait for A:
    // the loads are issued back to back and waited for once
    __builtin_amdgcn_sched_barrier(0);
    __asm__ volatile("" : : "v"(a0), "v"(a1), "v"(b0), "v"(b1));
    bool pop = false;
    if (state == kLeaf) {
      const float4 p0 = make_float4(__uint_as_float(a0[0]), __uint_as_float(a0[1]), __uint_as_float(a0[2]),
                                    __uint_as_float(a0[3]));
      if (STATS) ++st.prims;
      {
        const float fx = p0.x - qf[0], fy = p0.y - qf[1], fz = p0.z - qf[2];
        const float d2f = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
        if (d2f <= thr2) visit(first + j, p0, d2f);
      }
      if (j + 1 < cnt) {
        const float4 p1 = make_float4(__uint_as_float(a1[0]), __uint_as_float(a1[1]), __uint_as_float(a1[2]),
                                      __uint_as_float(a1[3]));
        if (STATS) ++st.prims;
        const float fx = p1.x - qf[0], fy = p1.y - qf[1], fz = p1.z - qf[2];
        const float d2f = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
        if (d2f <= thr2) visit(first + j + 1, p1, d2f);
      }
      j += 2;
      pop = j >= cnt;
    } else {
      // A is a node: the left child when expanding, else the node to test
      const float alo[3] = {__uint_as_float(a0[0]), __uint_as_float(a0[1]), __uint_as_float(a0[2])};
      const float ahi[3] = {__uint_as_float(a1[0]), __uint_as_float(a1[1]), __uint_as_float(a1[2])};
      const float da = box_dist2(qf, alo, ahi);
      uint32_t next = cur, link = a0[3];
      bool go = false;
      if (state == kExpand) {
        if (STATS) st.nodes += 2;
        const float blo[3] = {__uint_as_float(b0[0]), __uint_as_float(b0[1]), __uint_as_float(b0[2])};
        const float bhi[3] = {__uint_as_float(b1[0]), __uint_as_float(b1[1]), __uint_as_float(b1[2])};
        const float db = box_dist2(qf, blo, bhi);
        const bool right_first = db < da;
        const float dn = right_first ? db : da, df = right_first ? da : db;
        if (df <= thr2) {
          if (depth < DEPTH)
            stack[stride * depth++] = right_first ? cur + 1 : right;
          else
            complete = false;
        }
        go = dn <= thr2;
        next = right_first ? right : cur + 1;
        link = right_first ? b0[3] : a0[3];
      } else {  // kTest
        if (STATS) ++st.nodes;
        go = da <= thr2;
      }
      if (go) {
        cur = next;
        if (node_is_leaf(link)) {
          state = kLeaf;
          first = leaf_first(link);
          cnt = leaf_count(link);
          j = 0;
        } else {
          state = kExpand;
          right = link;
        }
      } else {
        pop = true;
      }
    }
    if (pop) {
      if (depth) {
        cur = stack[stride * --depth];
        state = kTest;
      } else {
        state = kDone;
      }
    }
    return state != kDone;
  }
};

template <bool STATS, uint32_t DEPTH = kStackDepth, class Visit>
__device__ inline bool walk_nearest(const SceneView& s, uint32_t root, uint32_t end, const float qf[3],
                                    float& thr2, LaneStats& st, Visit&& visit, uint32_t* stack,
                                    uint32_t stride) {
  NearestWalk<DEPTH> w;
  w.init(root, end);
  // every node is pushed at most once and a leaf takes at most four visits,
  // so the guard only bounds the loop should the tree ever be malformed
  for (uint32_t guard = 4 * s.n_nodes + 4; guard && !w.done(); --guard)
    w.template step<STATS>(s, qf, thr2, st, visit, stack, stride);
  return w.complete;
}

// Exact list: K (distance, id) entries in registers, k of them live in slots
// [K - k, K); slots below hold a -1 sentinel, so the k-th smallest is always
// slot K - 1 and every index is static.
template <int K>
struct ExactTopK {
  double d[K];
  uint32_t id[K];

  __device__ inline void init(uint32_t k) {
#pragma unroll
    for (int i = 0; i < K; ++i) {
      d[i] = (i < K - (int)k) ? -1.0 : __builtin_inf();
      id[i] = 0xFFFFFFFFu;
    }
  }
  __device__ inline bool accepts(double dist, uint32_t ident) const {
    return key_less(dist, ident, d[K - 1], id[K - 1]);
  }
  __device__ inline void insert(double dist, uint32_t ident) {
    d[K - 1] = dist;
    id[K - 1] = ident;
#pragma unroll
    for (int i = K - 1; i > 0; --i) {
      const bool sw = key_less(d[i], id[i], d[i - 1], id[i - 1]);
      const double td = d[i];
      const uint32_t ti = id[i];
      d[i] = sw ? d[i - 1] : td;
      id[i] = sw ? id[i - 1] : ti;
      d[i - 1] = sw ? td : d[i - 1];
      id[i - 1] = sw ? ti : id[i - 1];
    }
  }
  __device__ inline bool full() const { return d[K - 1] < __builtin_inf(); }
};

// Exact k nearest (distance, id) within `radius`; with COUNT_ALL the radius
// never shrinks and `inside` counts every candidate within it.
template <int K, bool COUNT_ALL, bool STATS>
__device__ inline void collect_exact(const SceneView& s, uint32_t root, uint32_t end, double qx,
                                     double qy, double qz, double radius, uint32_t k,
                                     ExactTopK<K>& list, uint32_t& inside, LaneStats& st) {
  const float qf[3] = {(float)qx, (float)qy, (float)qz};
  const float m = coord_bound(scene_m(s), qx, qy, qz);
  const double radius_sq = radius * radius;  // reference.py:357
  float thr2 = frn_threshold2(radius, m);
  list.init(k);
  inside = 0;
  if (root == kEndNode) return;
  auto visit = [&](uint32_t slot, const float4& p, float) {
    const double d2 = point_d2(s, slot, p, qx, qy, qz);
    if (d2 > radius_sq) return;  // reference.py:361 inclusive radius
    ++inside;
    const double dist = sqrt(d2);
    const uint32_t ident = __float_as_uint(p.w);
    if (list.accepts(dist, ident)) {
      list.insert(dist, ident);
      if (!COUNT_ALL && list.full()) thr2 = frn_threshold2(fmin(radius, list.d[K - 1]), m);
    }
  };
  walk_neighbors<STATS>(s, root, end, qf, radius, thr2, st, visit);
}

// Packed list: each entry is (FP32 bits of fl32(d2) << 32) | slot, so the
// whole (key, slot) order is one 64-bit compare and an entry costs two
// registers.  fl32 is monotone, so a smaller key never means a larger FP64
// distance, but keys do not decide ties: two squared distances on either side
// of an FP32 rounding point (adjacent doubles, say) get keys one ulp apart
// and can still have the same FP64 square root, which the reference orders by
// id.  Keys two or more ulps apart are a full FP32 ulp (relative 2^-24) apart
// in d2, so their square roots differ by far more than an FP64 ulp.  The list
// therefore reports `ambiguous` when a dropped key lies within one ulp of the
// final k-th key or when two kept keys lie within one ulp of each other; such
// queries are recomputed with the exact (distance, id) list.
template <int K>
struct PackedTopK {
  uint64_t e[K];

  __device__ inline void init(uint32_t k) {
#pragma unroll
    for (int i = 0; i < K; ++i) e[i] = (i < K - (int)k) ? 0ull : ~0ull;
  }
  __device__ inline uint64_t insert(uint64_t v) {  // returns the evicted entry
    const uint64_t out = e[K - 1];
    e[K - 1] = v;
#pragma unroll
    for (int i = K - 1; i > 0; --i) {
      const uint64_t a = e[i - 1], b = e[i];
      e[i - 1] = a < b ? a : b;
      e[i] = a < b ? b : a;
    }
    return out;
  }
  __device__ inline bool full() const { return e[K - 1] != ~0ull; }
};

__device__ inline uint64_t pack_key(double d2, uint32_t slot) {
  return ((uint64_t)__float_as_uint(__double2float_rn(d2)) << 32) | slot;
}
__device__ inline uint32_t key_bits(uint64_t e) { return (uint32_t)(e >> 32); }

// One candidate of a k-nearest search offered to its packed list: FP64
// distance in the reference's order, the radius test, the (fl32(d2), slot)
// key; once the list is full its k-th key shrinks the walk's threshold.
// drop_min keeps the smallest key bits of any candidate dropped.
template <int K>
__device__ inline void knn_offer(const SceneView& s, uint32_t slot, const float4& p, double qx, double qy,
                                 double qz, double radius, double radius_sq, float m, PackedTopK<K>& list,
                                 float& thr2, uint32_t& drop_min) {
  const double d2 = point_d2(s, slot, p, qx, qy, qz);
  if (d2 > radius_sq) return;
  const uint64_t v = pack_key(d2, slot);
  if (v < list.e[K - 1]) {
    const uint64_t out = list.insert(v);
    if (out != ~0ull) drop_min = min(drop_min, key_bits(out));
    if (list.full()) {
      // every candidate whose key lies within one ulp of the k-th key
      // stays inside (the factor covers 2^-23 of d2 and its rounding)
      const double kth = (double)__uint_as_float(key_bits(list.e[K - 1]));
      thr2 = frn_threshold2(fmin(radius, sqrt(kth) * 1.0000002384185791015625), m);
    }
  } else {
    drop_min = min(drop_min, key_bits(v));
  }
}

// Whether a finished k-nearest list may differ from the exact one: dropped
// keys are never below the final k-th key (it only shrinks), and kept keys
// ascend, so each test below is a difference of at most one ulp.
template <int K>
__device__ inline bool knn_ambiguous(const PackedTopK<K>& list, uint32_t k, uint32_t drop_min) {
  bool ambiguous = list.full() && drop_min - key_bits(list.e[K - 1]) <= 1u;
#pragma unroll
  for (int i = 1; i < K; ++i) {
    if (i > K - (int)k && list.e[i] != ~0ull && key_bits(list.e[i]) - key_bits(list.e[i - 1]) <= 1u)
      ambiguous = true;
  }
  return ambiguous;
}

// After a nearest-first walk whose stack overflowed: every primitive is
// offered again, skipping the entries already in the list.
template <int K, bool STATS>
__device__ inline void knn_rewalk(const SceneView& s, uint32_t root, uint32_t end, const float qf[3], double qx,
                                  double qy, double qz, double radius, double radius_sq, float m,
                                  PackedTopK<K>& list, float& thr2, uint32_t& drop_min, LaneStats& st) {
  walk_points<STATS>(s, root, end, qf, thr2, st, [&](uint32_t slot, const float4& p, float) {
    const uint64_t v = pack_key(point_d2(s, slot, p, qx, qy, qz), slot);
    bool dup = false;
#pragma unroll
    for (int i = 0; i < K; ++i) dup = dup || list.e[i] == v;
    if (!dup) knn_offer<K>(s, slot, p, qx, qy, qz, radius, radius_sq, m, list, thr2, drop_min);
  });
}

// `stack` (LDS, DEPTH entries at stride `stride`) selects the
// nearest-child-first walk for a search without a radius; without one the
// ring walk is used.
template <int K, bool STATS, uint32_t DEPTH = kStackDepth>
__device__ inline bool collect_packed(const SceneView& s, uint32_t root, uint32_t end, double qx,
                                      double qy, double qz, double radius, uint32_t k,
                                      PackedTopK<K>& list, LaneStats& st, uint32_t* stack = nullptr,
                                      uint32_t stride = 0) {
  const float qf[3] = {(float)qx, (float)qy, (float)qz};
  const float m = coord_bound(scene_m(s), qx, qy, qz);
  const double radius_sq = radius * radius;
  float thr2 = frn_threshold2(radius, m);
  list.init(k);
  uint32_t drop_min = 0xFFFFFFFFu;  // smallest key bits of any dropped candidate
  if (root == kEndNode) return false;
  auto visit = [&](uint32_t slot, const float4& p, float) {
    knn_offer<K>(s, slot, p, qx, qy, qz, radius, radius_sq, m, list, thr2, drop_min);
  };
  if (stack && !(radius < __builtin_inf())) {
    if (!walk_nearest<STATS, DEPTH>(s, root, end, qf, thr2, st, visit, stack, stride))
      knn_rewalk<K, STATS>(s, root, end, qf, qx, qy, qz, radius, radius_sq, m, list, thr2, drop_min, st);
  } else {
    walk_neighbors<STATS>(s, root, end, qf, radius, thr2, st, visit);
  }
  return knn_ambiguous<K>(list, k, drop_min);
}

// ------------------------------------------------------------------- rays --

// reference.py:810-846, Moller-Trumbore in this operation order.
__device__ inline bool mt_hit(double ox, double oy, double oz, double dx, double dy, double dz,
                              double tmax, const double v[9], double* t_out) {
  const double e1x = v[3] - v[0];
  const double e1y = v[4] - v[1];
  const double e1z = v[5] - v[2];
  const double e2x = v[6] - v[0];
  const double e2y = v[7] - v[1];
  const double e2z = v[8] - v[2];
  const double px = dy * e2z - dz * e2y;
  const double py = dz * e2x - dx * e2z;
  const double pz = dx * e2y - dy * e2x;
  const double det = e1x * px + e1y * py + e1z * pz;
  if (fabs(det) <= 1.0e-8) return false;
  const double inv = 1.0 / det;
  const double tx = ox - v[0];
  const double ty = oy - v[1];
  const double tz = oz - v[2];
  const double u = (tx * px + ty * py + tz * pz) * inv;
  if (u < 0.0 || u > 1.0) return false;
  const double qx = ty * e1z - tz * e1y;
  const double qy = tz * e1x - tx * e1z;
  const double qz = tx * e1y - ty * e1x;
  const double vv = (dx * qx + dy * qy + dz * qz) * inv;
  if (vv < 0.0 || (u + vv) > 1.0) return false;
  const double t = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (t >= 0.0 && t <= tmax) {
    *t_out = t;
    return true;
  }
  return false;
}

// 2-D ray any-hit (reference.py:783-803): the ray is the segment from o to
// o + d * tmax; it hits a triangle when either end lies inside it or the
// segment crosses one of its edges.  FP64 in the reference's operation order.

// reference.py:693-714, left = p -> e, right = q0 -> q1
__device__ inline bool segments_cross_2d(double px, double py, double ex, double ey, double qx0,
                                         double qy0, double qx1, double qy1) {
  const double rx = ex - px;
  const double ry = ey - py;
  const double sx = qx1 - qx0;
  const double sy = qy1 - qy0;
  const double denom = rx * sy - ry * sx;
  if (fabs(denom) < 1.0e-7) return false;
  const double qpx = qx0 - px;
  const double qpy = qy0 - py;
  const double t = (qpx * sy - qpy * sx) / denom;
  const double u = (qpx * ry - qpy * rx) / denom;
  return 0.0 <= t && t <= 1.0 && 0.0 <= u && u <= 1.0;
}

// reference.py:893-915, vertices (ax, ay), (bx, by), (cx, cy)
__device__ inline bool point_in_triangle_2d(double x, double y, double ax, double ay, double bx,
                                            double by, double cx, double cy) {
  const double v0x = cx - ax;
  const double v0y = cy - ay;
  const double v1x = bx - ax;
  const double v1y = by - ay;
  const double v2x = x - ax;
  const double v2y = y - ay;
  const double dot00 = v0x * v0x + v0y * v0y;
  const double dot01 = v0x * v1x + v0y * v1y;
  const double dot02 = v0x * v2x + v0y * v2y;
  const double dot11 = v1x * v1x + v1y * v1y;
  const double dot12 = v1x * v2x + v1y * v2y;
  const double denom = dot00 * dot11 - dot01 * dot01;
  if (fabs(denom) < 1.0e-7) return false;
  const double inv = 1.0 / denom;
  const double u = (dot11 * dot02 - dot01 * dot12) * inv;
  const double v = (dot00 * dot12 - dot01 * dot02) * inv;
  return u >= 0.0 && v >= 0.0 && (u + v) <= 1.0;
}

// v: x0 y0 z0 x1 y1 z1 x2 y2 z2 (z unused)
__device__ inline bool hit_2d(double ox, double oy, double dx, double dy, double tmax, const double v[9]) {
  const double ex = ox + dx * tmax;
  const double ey = oy + dy * tmax;
  if (point_in_triangle_2d(ox, oy, v[0], v[1], v[3], v[4], v[6], v[7])) return true;
  if (point_in_triangle_2d(ex, ey, v[0], v[1], v[3], v[4], v[6], v[7])) return true;
  return segments_cross_2d(ox, oy, ex, ey, v[0], v[1], v[3], v[4]) ||
         segments_cross_2d(ox, oy, ex, ey, v[3], v[4], v[6], v[7]) ||
         segments_cross_2d(ox, oy, ex, ey, v[6], v[7], v[0], v[1]);
}

// Conservative FP32 test of the segment o + s * D, s in [0, slim], against
// the box grown by `pad` on every side.  `pad` (2^-18 of the coordinate
// scale) dominates every FP32 rounding of o, D, the box and the slab
// parameters; the parameters get a further (1 + 2^-20) factor.
__device__ inline bool segment_box(const float o[3], const float D[3], const float inv[3],
                                   const float lo[3], const float hi[3], float pad, float slim) {
  float tn = -9.5367431640625e-07f;
  float tf = slim * 1.00000095367431640625f + 9.5367431640625e-07f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float l = lo[a] - pad, h = hi[a] + pad;
    if (D[a] == 0.0f) {
      if (o[a] < l || o[a] > h) return false;
      continue;
    }
    const float t0 = (l - o[a]) * inv[a], t1 = (h - o[a]) * inv[a];
    const float a0 = fminf(t0, t1);
    const float a1 = fmaxf(t0, t1);
    tn = fmaxf(tn, a0 - fabsf(a0) * 9.5367431640625e-07f);
    tf = fminf(tf, a1 + fabsf(a1) * 9.5367431640625e-07f);
  }
  return tn <= tf;
}

struct RayResult {
  uint32_t hits = 0;
  double best_t = __builtin_nan("");
  uint32_t best_id = 0;
};

// MODE & 3: 0 any-hit (stops at the first hit), 1 hit count, 2 closest hit;
// MODE & 4: 2-D (z of rays and triangles is 0, the hit test of
// reference.py:783-803, hit t = 0.0 as reference.py:780 returns).
template <int MODE, bool STATS>
__device__ inline RayResult trace_ray(const SceneView& s, uint32_t root, uint32_t end, double ox,
                                      double oy, double oz, double dx, double dy, double dz,
                                      double tmax, LaneStats& st) {
  constexpr int KIND = MODE & 3;
  constexpr bool D2 = (MODE & 4) != 0;
  RayResult res;
  // 3-D: t in [0, tmax] is impossible for a negative or NaN tmax
  // (reference.py:844).  2-D takes any tmax: the segment o -> o + d * tmax
  // is tested as it is (reference.py:783-803).
  if ((!D2 && !(tmax >= 0.0)) || root == kEndNode) return res;
  // The prefilter walks the segment o -> o + d * seg_t.  A 3-D ray with an
  // infinite tmax can only hit inside the scene's box [-M, M]^3, where
  // |o + t d| <= M bounds t by (M + |o|) / |d| (max norms); the walk takes
  // four times that, so the FP64 hits of the reference all lie on it.  A
  // segment with a non-finite end (2-D, or a zero or NaN direction) is
  // tested against every primitive: the box prefilter is switched off
  // (pad = inf).
  double seg_t = tmax;
  if (!D2 && !(tmax < __builtin_inf())) {
    const double dn = fmax(fmax(fabs(dx), fabs(dy)), fabs(dz));
    const double on = fmax(fmax(fabs(ox), fabs(oy)), fabs(oz));
    seg_t = 4.0 * (((double)scene_m(s) + on) / dn);
  }
  const bool bounded = fabs(dx * seg_t) < __builtin_inf() && fabs(dy * seg_t) < __builtin_inf() &&
                       fabs(dz * seg_t) < __builtin_inf();
  const float o32[3] = {(float)ox, (float)oy, (float)oz};
  float D32[3] = {0.0f, 0.0f, 0.0f};
  float inv[3] = {0.0f, 0.0f, 0.0f};
  float pad = __builtin_inff();
  if (bounded) {
    const double D[3] = {dx * seg_t, dy * seg_t, dz * seg_t};
    const double omax = fmax(fmax(fabs(ox), fabs(oy)), fabs(oz));
    const double dmax = fmax(fmax(fabs(D[0]), fabs(D[1])), fabs(D[2]));
    pad = __double2float_ru(((double)scene_m(s) + omax + dmax) * 3.814697265625e-06);  // 2^-18
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      D32[a] = (float)D[a];
      inv[a] = D32[a] != 0.0f ? 1.0f / D32[a] : 0.0f;
    }
  }
  float slim = 1.0f;
  uint32_t node = root;
  bool done = false;
  for (uint32_t guard = s.n_nodes + 1; node != end && !done && guard; --guard) {
    BvhNode nd;
    bool in;
    rope_pair_step<STATS>(s, node, nd, in, st, [&](const BvhNode& b, uint32_t) {
      const float lo[3] = {b.lo[0], b.lo[1], b.lo[2]};
      const float hi[3] = {b.hi[0], b.hi[1], b.hi[2]};
      return segment_box(o32, D32, inv, lo, hi, pad, slim);
    });
    const bool leaf = node_is_leaf(nd.link);
    // into an overlapped internal node's subtree, else past the node
    const uint32_t next = (in && !leaf) ? node + 1 : nd.escape;  // left child: the next record
    if (!(in && leaf)) {
      node = next;
      continue;
    }
    const uint32_t first = leaf_first(nd.link), cnt = leaf_count(nd.link);
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t slot = first + j;
      if (STATS) ++st.prims;
      double v[9];
      const float4 a = load_prim(s.prims, 3 * (size_t)slot);
      const uint32_t ident = __float_as_uint(a.w);
      if (s.exact) {
#pragma unroll
        for (int e = 0; e < 9; ++e) v[e] = load_f64(s.exact, 9 * (size_t)slot + e);
      } else {
        const float4 b = load_prim(s.prims, 3 * (size_t)slot + 1);
        const float4 c = load_prim(s.prims, 3 * (size_t)slot + 2);
        v[0] = a.x; v[1] = a.y; v[2] = a.z;
        v[3] = b.x; v[4] = b.y; v[5] = b.z;
        v[6] = c.x; v[7] = c.y; v[8] = c.z;
      }
      double th = 0.0;
      if (D2) {
        if (!hit_2d(ox, oy, dx, dy, tmax, v)) continue;
      } else if (!mt_hit(ox, oy, oz, dx, dy, dz, tmax, v, &th)) {
        continue;
      }
      ++res.hits;
      if (KIND == 0) {
        done = true;
        break;
      }
      if (KIND == 2) {
        if (!(res.best_t == res.best_t) || th < res.best_t || (th == res.best_t && ident < res.best_id)) {
          res.best_t = th;
          res.best_id = ident;
          // 2-D hits all report t = 0: the segment cannot be shortened
          if (!D2 && bounded && seg_t > 0.0) slim = fminf(1.0f, __double2float_ru(th / seg_t));
        }
      }
    }
    node = nd.escape;
  }
  return res;
}

// -------------------------------------------------------- point in polygon --

// reference.py:745-768 (`x ** 2` and `x ** 0.5` are evaluated as the
// correctly rounded x * x and sqrt(x); libm pow can differ from them by one
// ulp, which can only matter for a point at the 1e-12 tolerance boundary).
__device__ inline bool point_on_segment(double px, double py, double ax, double ay, double bx,
                                        double by, double eps_sq) {
  const double eps = 1.0e-12;
  const double ex = bx - ax, ey = by - ay;
  const double length_sq = ex * ex + ey * ey;
  if (length_sq <= eps_sq) return fabs(px - ax) <= eps && fabs(py - ay) <= eps;
  const double cross = (px - ax) * (by - ay) - (py - ay) * (bx - ax);
  // cross^2 > 4 eps^2 length_sq (FP64, no underflow since length_sq >
  // eps_sq) implies |cross| > 2 eps sqrt(length_sq) (1 - 2^-50), beyond any
  // rounding of the exact test below: the usual far-from-the-edge case
  // skips the square root with the same answer
  if (cross * cross > 4.0 * eps_sq * length_sq) return false;
  const double length = sqrt(length_sq);
  if (fabs(cross) > eps * length) return false;
  const double dot = (px - ax) * (bx - ax) + (py - ay) * (by - ay);
  const double along_eps = eps * length;
  if (dot < -along_eps) return false;
  if (dot - length_sq > along_eps) return false;
  return true;
}

// reference.py:717-742
// Vertex (x, y) pairs fetched together by point_in_polygon, so their memory
// latencies overlap instead of following one another edge by edge.
constexpr uint32_t kPipVertexChunk = 4;

__device__ inline void load_vertex(const double* vertices, uint32_t k, double& x, double& y) {
  const GlobalU32x4 w = load_raw16(vertices + 2 * (size_t)k);  // 16-byte aligned (x, y)
  x = __builtin_bit_cast(double, ((uint64_t)w[1] << 32) | w[0]);
  y = __builtin_bit_cast(double, ((uint64_t)w[3] << 32) | w[2]);
}

__device__ inline bool point_in_polygon(double x, double y, const double* vertices, uint32_t n,
                                        double eps_sq) {
  // One pass over the edges a = v[k] -> b = v[k + 1]: any edge holding the
  // point answers true (the reference's first loop), otherwise the
  // even-odd parity of the crossings decides (its second loop, whose pair
  // (i, j = i - 1) is (b, a) here; the parity does not depend on the order).
  bool inside = false;
  double ax, ay;
  load_vertex(vertices, 0, ax, ay);
  for (uint32_t k0 = 0; k0 < n; k0 += kPipVertexChunk) {
    double bx[kPipVertexChunk], by[kPipVertexChunk];
#pragma unroll
    for (uint32_t t = 0; t < kPipVertexChunk; ++t) {
      const uint32_t k = k0 + t;
      bx[t] = ax;
      by[t] = ay;
      if (k < n) load_vertex(vertices, k + 1 == n ? 0u : k + 1, bx[t], by[t]);
    }
#pragma unroll
    for (uint32_t t = 0; t < kPipVertexChunk; ++t) {
      if (k0 + t >= n) break;
      if (point_on_segment(x, y, ax, ay, bx[t], by[t], eps_sq)) return true;
      if ((by[t] > y) != (ay > y)) {
        double den = ay - by[t];
        if (den == 0.0) den = 1.0e-20;
        if (x <= (ax - bx[t]) * (y - by[t]) / den + bx[t]) inside = !inside;
      }
      ax = bx[t];
      ay = by[t];
    }
  }
  return inside;
}

// point_in_polygon with the edges spread over the 64 lanes of a wavefront
// (every lane takes part; x, y and the polygon are the same in all lanes):
// the edges' on-edge and crossing tests are those of point_in_polygon, and
// "any edge holds the point, else the parity of the crossings" does not
// depend on the order the edges are taken in.
__device__ inline bool point_in_polygon_wave(double x, double y, const double* vertices, uint32_t n,
                                             double eps_sq) {
  const uint32_t lane = threadIdx.x & 63;
  bool on_any = false;
  uint32_t parity = 0;
  for (uint32_t k0 = 0; k0 < n; k0 += 64) {
    const uint32_t k = k0 + lane;
    bool on = false, cross = false;
    if (k < n) {
      double ax, ay, bx, by;
      load_vertex(vertices, k, ax, ay);
      load_vertex(vertices, k + 1 == n ? 0u : k + 1, bx, by);
      on = point_on_segment(x, y, ax, ay, bx, by, eps_sq);
      if ((by > y) != (ay > y)) {
        double den = ay - by;
        if (den == 0.0) den = 1.0e-20;
        cross = x <= (ax - bx) * (y - by) / den + bx;
      }
    }
    on_any = on_any || __ballot(on) != 0;
    parity ^= (uint32_t)__popcll(__ballot(cross)) & 1u;
  }
  return on_any || parity != 0;
}

// Calls hit(slot) for every polygon that contains (x, y) under the
// reference's rules.  Boxes are grown by more than the on-edge tolerance
// (1e-12 from the boundary) plus FP32 rounding of the point.
template <bool STATS, class Hit>
__device__ inline void walk_pip(const SceneView& s, uint32_t root, uint32_t end, double x, double y,
                                bool positive_only, double eps_sq, LaneStats& st, Hit&& hit) {
  const float pf0 = (float)x, pf1 = (float)y;
  const float m = coord_bound(scene_m(s), x, y, 0.0);
  const float pad = m * 9.5367431640625e-07f + 1.0e-11f;
  uint32_t node = root;
  for (uint32_t guard = s.n_nodes + 1; node != end && guard; --guard) {
    BvhNode nd;
    bool in;
    rope_pair_step<STATS>(s, node, nd, in, st, [&](const BvhNode& b, uint32_t) {
      return (pf0 >= b.lo[0] - pad) & (pf0 <= b.hi[0] + pad) & (pf1 >= b.lo[1] - pad) & (pf1 <= b.hi[1] + pad);
    });
    const bool leaf = node_is_leaf(nd.link);
    // into an overlapped internal node's subtree, else past the node
    const uint32_t next = (in && !leaf) ? node + 1 : nd.escape;  // left child: the next record
    if (!(in && leaf)) {
      node = next;
      continue;
    }
    const uint32_t first = leaf_first(nd.link), cnt = leaf_count(nd.link);
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t slot = first + j;
      if (STATS) ++st.prims;
      if (positive_only) {  // reference.py:122-125 exact bbox prefilter
        const size_t b = 4 * (size_t)slot;
        if (x < load_f64(s.poly_box, b) || x > load_f64(s.poly_box, b + 2) || y < load_f64(s.poly_box, b + 1) ||
            y > load_f64(s.poly_box, b + 3))
          continue;
      }
      const uint4 ref = load_u4(s.poly, slot);
      if (point_in_polygon(x, y, s.poly_verts + 2 * (size_t)ref.x, ref.y, eps_sq)) hit(slot);
    }
    node = nd.escape;
  }
}

}  // namespace rtdl
