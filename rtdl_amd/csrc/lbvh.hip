// LBVH construction for points, triangles and polygons.
//
//   1. bounds:   per-primitive FP32 boxes rounded outward from the FP64 input,
//                per-instance centroid bounds, the coordinate bound M and
//                whether every coordinate is exact in FP32;
//   2. keys:     (instance << 32) | Morton code of the centroid (30-bit 3-D,
//                32-bit 2-D), sorted with the LDS radix sort (stable, so equal
//                codes keep input order);
//   3. gather:   primitives copied into Morton order;
//   4. leaves:   runs of up to `leaf_size` consecutive primitives of one
//                instance; Karras' binary radix tree over the leaf keys
//                (ties broken by leaf index);
//   5. ropes:    the escape link of a node whose leaf range ends at leaf b is
//                the highest node whose range starts at leaf b + 1 (found by
//                an atomicMax over range lengths), so traversal needs no
//                stack;
//   6. refit:    bottom-up box unions in rounds, one launch per tree level:
//                a node is refit in the launch after both children are
//                done, so no thread waits on another and no fence is needed
//                (an agent-scope fence writes back the L2 of the thread's
//                XCD, which made a climbing-thread refit the slowest step).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "io.hpp"
#include "primitives.hpp"
#include "scene.hpp"

namespace rtdl {
namespace {

__device__ inline int ordered_from_float(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__host__ inline float float_from_ordered(int i) {
  int b = i >= 0 ? i : i ^ 0x7FFFFFFF;
  float f;
  std::memcpy(&f, &b, sizeof(f));
  return f;
}

struct BoundsTask {
  uint32_t instance;
  uint32_t begin;
  uint32_t end;
};

__device__ inline void prim_box(PrimKind kind, size_t p, const double* coords, const uint4* refs,
                                const double* verts, double lo[3], double hi[3], double c[3],
                                bool* exact, double* maxabs) {
  *exact = true;
  *maxabs = 0.0;
  auto see = [&](double v) {
    *exact = *exact && ((double)(float)v == v);
    *maxabs = fmax(*maxabs, fabs(v));
  };
  if (kind == PrimKind::Points) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double v = coords[3 * p + a];
      lo[a] = hi[a] = c[a] = v;
      see(v);
    }
  } else if (kind == PrimKind::Triangles) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double v0 = coords[9 * p + a], v1 = coords[9 * p + 3 + a], v2 = coords[9 * p + 6 + a];
      lo[a] = fmin(fmin(v0, v1), v2);
      hi[a] = fmax(fmax(v0, v1), v2);
      c[a] = (v0 + v1 + v2) / 3.0;
      see(v0);
      see(v1);
      see(v2);
    }
  } else {
    const uint4 r = refs[p];
    lo[0] = lo[1] = __builtin_inf();
    hi[0] = hi[1] = -__builtin_inf();
    for (uint32_t k = 0; k < r.y; ++k) {
      const double x = verts[2 * ((size_t)r.x + k)], y = verts[2 * ((size_t)r.x + k) + 1];
      lo[0] = fmin(lo[0], x);
      hi[0] = fmax(hi[0], x);
      lo[1] = fmin(lo[1], y);
      hi[1] = fmax(hi[1], y);
      see(x);
      see(y);
    }
    lo[2] = hi[2] = 0.0;
    c[0] = 0.5 * (lo[0] + hi[0]);
    c[1] = 0.5 * (lo[1] + hi[1]);
    c[2] = 0.0;
  }
}

__global__ void bounds_kernel(PrimKind kind, const BoundsTask* tasks, const double* coords,
                              const uint4* refs, const double* verts, float4* box_lo,
                              float4* box_hi, int* inst_bounds, uint32_t* maxabs_bits,
                              uint32_t* exact_flag) {
  const BoundsTask task = tasks[blockIdx.x];
  __shared__ int red[6];
  __shared__ float red_m;
  __shared__ uint32_t red_exact;
  if (threadIdx.x < 3) {
    red[threadIdx.x] = 0x7FFFFFFF;
    red[3 + threadIdx.x] = (int)0x80000000;
  }
  if (threadIdx.x == 0) {
    red_m = 0.0f;
    red_exact = 1u;
  }
  __syncthreads();
  float cmin[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
  float cmax[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
  float local_m = 0.0f;
  bool local_exact = true;
  for (uint32_t p = task.begin + threadIdx.x; p < task.end; p += blockDim.x) {
    double lo[3], hi[3], c[3], maxabs;
    bool exact;
    prim_box(kind, p, coords, refs, verts, lo, hi, c, &exact, &maxabs);
    box_lo[p] = make_float4(__double2float_rd(lo[0]), __double2float_rd(lo[1]),
                            __double2float_rd(lo[2]), 0.0f);
    box_hi[p] = make_float4(__double2float_ru(hi[0]), __double2float_ru(hi[1]),
                            __double2float_ru(hi[2]), 0.0f);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float cf = (float)c[a];
      cmin[a] = fminf(cmin[a], cf);
      cmax[a] = fmaxf(cmax[a], cf);
    }
    local_m = fmaxf(local_m, __double2float_ru(maxabs));
    local_exact = local_exact && exact;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (cmin[a] <= cmax[a]) {
      atomicMin(&red[a], ordered_from_float(cmin[a]));
      atomicMax(&red[3 + a], ordered_from_float(cmax[a]));
    }
  }
  atomicMax((unsigned int*)&red_m, __float_as_uint(local_m));
  if (!local_exact) red_exact = 0u;
  __syncthreads();
  if (threadIdx.x < 3) {
    atomicMin(&inst_bounds[6 * task.instance + threadIdx.x], red[threadIdx.x]);
    atomicMax(&inst_bounds[6 * task.instance + 3 + threadIdx.x], red[3 + threadIdx.x]);
  }
  if (threadIdx.x == 0) {
    atomicMax(maxabs_bits, __float_as_uint(red_m));
    if (!red_exact) atomicAnd(exact_flag, 0u);
  }
}

__device__ inline uint32_t spread10(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ inline uint32_t spread16(uint32_t v) {  // 16 bits -> every second bit
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

__device__ inline uint32_t quantize(float v, float lo, float hi, uint32_t cells) {
  const float ext = hi - lo;
  if (!(ext > 0.0f)) return 0u;
  float f = (v - lo) / ext * (float)cells;
  f = fminf(fmaxf(f, 0.0f), (float)(cells - 1));
  return (uint32_t)f;
}

__global__ void morton_kernel(size_t n, int dim, const float4* box_lo, const float4* box_hi,
                              const uint64_t* inst_offsets, uint32_t n_inst,
                              const float* inst_bounds, uint64_t* keys, uint32_t* vals) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  // instance of p: last i with inst_offsets[i] <= p
  uint32_t lo = 0, hi = n_inst;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (inst_offsets[mid] <= p) lo = mid; else hi = mid;
  }
  const float* b = inst_bounds + 6 * lo;
  const float4 l = box_lo[p], h = box_hi[p];
  const float cx = 0.5f * (l.x + h.x), cy = 0.5f * (l.y + h.y), cz = 0.5f * (l.z + h.z);
  uint32_t code;
  if (dim == 3) {
    code = (spread10(quantize(cx, b[0], b[3], 1024)) << 2) |
           (spread10(quantize(cy, b[1], b[4], 1024)) << 1) | spread10(quantize(cz, b[2], b[5], 1024));
  } else {
    code = (spread16(quantize(cx, b[0], b[3], 65536)) << 1) | spread16(quantize(cy, b[1], b[4], 65536));
  }
  keys[p] = ((uint64_t)lo << 32) | code;
  vals[p] = (uint32_t)p;
}

__global__ void gather_kernel(PrimKind kind, size_t n, const uint32_t* order, const double* coords,
                              const uint32_t* ids, const uint4* refs, const float4* box_lo,
                              const float4* box_hi, float4* prims, double* exact,
                              uint32_t* slot_to_input, uint4* poly, double* poly_box,
                              float4* sorted_lo, float4* sorted_hi) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t p = order[j];
  slot_to_input[j] = p;
  sorted_lo[j] = box_lo[p];
  sorted_hi[j] = box_hi[p];
  if (kind == PrimKind::Points) {
    const double x = coords[3 * (size_t)p], y = coords[3 * (size_t)p + 1], z = coords[3 * (size_t)p + 2];
    prims[j] = make_float4((float)x, (float)y, (float)z, __uint_as_float(ids[p]));
    if (exact) {
      exact[3 * j] = x;
      exact[3 * j + 1] = y;
      exact[3 * j + 2] = z;
    }
  } else if (kind == PrimKind::Triangles) {
    const double* c = coords + 9 * (size_t)p;
    prims[3 * j] = make_float4((float)c[0], (float)c[1], (float)c[2], __uint_as_float(ids[p]));
    prims[3 * j + 1] = make_float4((float)c[3], (float)c[4], (float)c[5], 0.0f);
    prims[3 * j + 2] = make_float4((float)c[6], (float)c[7], (float)c[8], 0.0f);
    if (exact) {
#pragma unroll
      for (int a = 0; a < 9; ++a) exact[9 * j + a] = c[a];
    }
  } else {
    const uint4 r = refs[p];
    poly[j] = r;
    // exact FP64 box of the polygon (reference pip_cpu bounds, min/max of vertices)
    (void)poly_box;
  }
}

__global__ void poly_box_kernel(size_t n, const uint4* poly, const double* verts, double* poly_box) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint4 r = poly[j];
  double x0 = verts[2 * (size_t)r.x], y0 = verts[2 * (size_t)r.x + 1], x1 = x0, y1 = y0;
  for (uint32_t k = 1; k < r.y; ++k) {
    const double x = verts[2 * ((size_t)r.x + k)], y = verts[2 * ((size_t)r.x + k) + 1];
    x0 = fmin(x0, x);
    x1 = fmax(x1, x);
    y0 = fmin(y0, y);
    y1 = fmax(y1, y);
  }
  poly_box[4 * j] = x0;
  poly_box[4 * j + 1] = y0;
  poly_box[4 * j + 2] = x1;
  poly_box[4 * j + 3] = y1;
}

// leaf l of instance i covers slots [first, first + count)
__global__ void leaves_kernel(size_t m, const uint64_t* leaf_offsets, const uint64_t* inst_offsets,
                              uint32_t n_inst, uint32_t leaf_size, const uint64_t* sorted_keys,
                              uint32_t* leaf_first, uint32_t* leaf_count, uint64_t* leaf_keys) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= m) return;
  uint32_t lo = 0, hi = n_inst;
  while (hi - lo > 1) {  // last instance with leaf_offsets[i] <= l
    const uint32_t mid = (lo + hi) / 2;
    if (leaf_offsets[mid] <= l) lo = mid; else hi = mid;
  }
  // skip empty instances that share the same leaf offset
  while (lo + 1 < n_inst && leaf_offsets[lo + 1] <= l) ++lo;
  const uint64_t first = inst_offsets[lo] + (l - leaf_offsets[lo]) * (uint64_t)leaf_size;
  const uint64_t end = inst_offsets[lo + 1];
  const uint64_t count = end - first < leaf_size ? end - first : leaf_size;
  leaf_first[l] = (uint32_t)first;
  leaf_count[l] = (uint32_t)count;
  leaf_keys[l] = sorted_keys[first];
}

__device__ inline int karras_delta(const uint64_t* keys, int m, int i, int j) {
  if (j < 0 || j >= m) return -1;
  const uint64_t a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll(a ^ b);
}

__global__ void karras_kernel(int m, const uint64_t* keys, uint32_t* left, uint32_t* right,
                              uint32_t* parent, uint32_t* range_lo, uint32_t* range_hi,
                              unsigned long long* highest) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m - 1) return;
  const int d = (karras_delta(keys, m, i, i + 1) - karras_delta(keys, m, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = karras_delta(keys, m, i, i - d);
  int lmax = 2;
  while (karras_delta(keys, m, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (karras_delta(keys, m, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = karras_delta(keys, m, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) / 2;
    if (karras_delta(keys, m, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? d : 0);
  const int lo = i < j ? i : j;
  const int hi = i < j ? j : i;
  const uint32_t leaf_base = (uint32_t)(m - 1);
  const uint32_t lc = (lo == gamma) ? leaf_base + (uint32_t)gamma : (uint32_t)gamma;
  const uint32_t rc = (hi == gamma + 1) ? leaf_base + (uint32_t)(gamma + 1) : (uint32_t)(gamma + 1);
  left[i] = lc;
  right[i] = rc;
  parent[lc] = (uint32_t)i;
  parent[rc] = (uint32_t)i;
  range_lo[i] = (uint32_t)lo;
  range_hi[i] = (uint32_t)hi;
  atomicMax(&highest[lo], ((unsigned long long)(hi - lo) << 32) | (unsigned long long)i);
}

__device__ inline uint32_t escape_after(uint32_t last_leaf, uint32_t m,
                                        const unsigned long long* highest) {
  if (last_leaf + 1 >= m) return kEndNode;
  const unsigned long long h = highest[last_leaf + 1];
  if (h) return (uint32_t)(h & 0xFFFFFFFFull);
  return (m - 1) + last_leaf + 1;
}

// Internal records hold the right child (the left one is the next record in
// preorder) and the rope.
__global__ void links_kernel(uint32_t m, const uint32_t* right, const uint32_t* range_hi,
                             const unsigned long long* highest, BvhNode* nodes) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m - 1) return;
  nodes[k].link = right[k];
  nodes[k].escape = escape_after(range_hi[k], m, highest);
}

// Leaf records: box of the leaf's primitives, first slot and count, rope.
__global__ void leaf_nodes_kernel(uint32_t m, const uint32_t* leaf_first, const uint32_t* leaf_count,
                                  const float4* sorted_lo, const float4* sorted_hi,
                                  const unsigned long long* highest, BvhNode* nodes) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= m) return;
  const uint32_t first = leaf_first[l], count = leaf_count[l];
  float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
  float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
  for (uint32_t k = 0; k < count; ++k) {
    const float4 a = sorted_lo[first + k], b = sorted_hi[first + k];
    lo[0] = fminf(lo[0], a.x);
    lo[1] = fminf(lo[1], a.y);
    lo[2] = fminf(lo[2], a.z);
    hi[0] = fmaxf(hi[0], b.x);
    hi[1] = fmaxf(hi[1], b.y);
    hi[2] = fmaxf(hi[2], b.z);
  }
  BvhNode out;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    out.lo[a] = lo[a];
    out.hi[a] = hi[a];
  }
  out.link = kLeafFlag | ((count - 1) << 28) | first;
  out.escape = escape_after(l, m, highest);
  nodes[(m - 1) + l] = out;
}

constexpr uint32_t kNotRefit = 0xFFFFFFFFu;

// One round of the bottom-up refit: internal node k takes the union of its
// children's boxes once both were finished in earlier launches (leaves from
// the start, internal children with done[] < round), and records the round.
// A node finishes in the round equal to its height, so the rounds need no
// atomics or fences: every box a round reads was written by an earlier
// launch on the same stream.
__global__ void refit_round_kernel(uint32_t n_internal, uint32_t round, const uint32_t* left,
                                   const uint32_t* right, uint32_t* done, BvhNode* nodes) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  // once the root (internal node 0) is done, so is every node: later rounds
  // of a fixed-length sequence return at once
  if (k >= n_internal || done[0] != kNotRefit || done[k] != kNotRefit) return;
  const uint32_t a = left[k], b = right[k];
  const bool ready_a = a >= n_internal || done[a] < round;
  const bool ready_b = b >= n_internal || done[b] < round;
  if (!(ready_a && ready_b)) return;
  const BvhNode na = nodes[a], nb = nodes[b];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    nodes[k].lo[ax] = fminf(na.lo[ax], nb.lo[ax]);
    nodes[k].hi[ax] = fmaxf(na.hi[ax], nb.hi[ax]);
  }
  done[k] = round;
}

// Root of each instance: the node whose leaf range is exactly the instance's.
__global__ void roots_kernel(uint32_t n_inst, uint32_t m, const uint64_t* leaf_offsets,
                             const uint32_t* parent, const uint32_t* range_lo,
                             const uint32_t* range_hi, const BvhNode* nodes, uint32_t* roots,
                             uint32_t* ends) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  const uint64_t a = leaf_offsets[i], b = leaf_offsets[i + 1];
  if (a == b) {
    roots[i] = kEndNode;
    ends[i] = kEndNode;
    return;
  }
  uint32_t node = (m - 1) + (uint32_t)a;
  if (b - a > 1) {
    node = parent[node];
    while (node != kEndNode && !(range_lo[node] == a && range_hi[node] == b - 1)) node = parent[node];
  }
  roots[i] = node;
  ends[i] = node == kEndNode ? kEndNode : nodes[node].escape;
}

// Depth-first preorder position of every node.  The nodes before n in
// preorder are its ancestors and the subtrees left of its path; those subtrees
// cover the leaves [0, lo(n)) with 2*lo(n) - t nodes for t right turns, so
// pre(n) = 2*lo(n) + (number of left turns on the path from the root).
__global__ void preorder_kernel(uint32_t m, const uint32_t* parent, const uint32_t* left,
                                const uint32_t* range_lo, uint32_t* pre) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= 2 * m - 1) return;
  const uint32_t lo = n < m - 1 ? range_lo[n] : n - (m - 1);
  uint32_t turns = 0;
  uint32_t c = n;
  for (uint32_t guard = 2 * m; c != 0 && guard; --guard) {  // the root is node 0
    const uint32_t p = parent[c];
    if (left[p] == c) ++turns;
    c = p;
  }
  pre[n] = 2 * lo + turns;
}

// Stores the nodes in preorder, so a walk that descends reads consecutive
// records (the left child of an internal node is the next record; `link`
// keeps the right child's new position).
__global__ void permute_nodes_kernel(uint32_t n_nodes, const BvhNode* in, const uint32_t* pre,
                                     BvhNode* out) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  BvhNode nd = in[n];
  if (!node_is_leaf(nd.link)) nd.link = pre[nd.link];
  if (nd.escape != kEndNode) nd.escape = pre[nd.escape];
  out[pre[n]] = nd;
}

__global__ void remap_roots_kernel(uint32_t n_inst, const uint32_t* pre, uint32_t* roots,
                                   uint32_t* ends) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  if (roots[i] != kEndNode) roots[i] = pre[roots[i]];
  if (ends[i] != kEndNode) ends[i] = pre[ends[i]];
}

// The quantisation frame of a point scene's compact nodes, from the union of
// its instance roots' boxes (one workgroup of 256 threads).  Per axis the
// step s is the smallest power of two with s >= extent / 2^15 and
// s >= max|coordinate| / 2^22, and the origin o = floor(lo / s) * s, so every
// o + q * s (0 <= q < 2^16) is an integer multiple of s below 2^24 * s in
// magnitude: exact in FP32.
__global__ void qframe_kernel(uint32_t n_inst, const uint32_t* roots, const BvhNode* nodes, float* frame) {
  __shared__ float r_lo[3][256], r_hi[3][256];
  __shared__ uint32_t r_bad[256];
  const uint32_t t = threadIdx.x;
  float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
  float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
  uint32_t bad = 0;
  for (uint32_t i = t; i < n_inst; i += 256) {
    const uint32_t r = roots[i];
    if (r == kEndNode) continue;
    const BvhNode nd = nodes[r];
    for (int a = 0; a < 3; ++a) {
      if (!(fabsf(nd.lo[a]) < __builtin_inff()) || !(fabsf(nd.hi[a]) < __builtin_inff())) bad = 1;
      lo[a] = fminf(lo[a], nd.lo[a]);
      hi[a] = fmaxf(hi[a], nd.hi[a]);
    }
  }
  for (int a = 0; a < 3; ++a) {
    r_lo[a][t] = lo[a];
    r_hi[a][t] = hi[a];
  }
  r_bad[t] = bad;
  __syncthreads();
  if (t != 0) return;
  for (uint32_t u = 1; u < 256; ++u) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], r_lo[a][u]);
      hi[a] = fmaxf(hi[a], r_hi[a][u]);
    }
    bad |= r_bad[u];
  }
  for (int a = 0; a < 3; ++a) {
    double o = 0.0, s = 1.0;
    if (!bad && lo[a] <= hi[a]) {
      const double ext = (double)hi[a] - (double)lo[a];
      const double mx = fmax(fabs((double)lo[a]), fabs((double)hi[a]));
      const double need = fmax(fmax(ext * 0x1p-15, mx * 0x1p-22), 0x1p-100);
      int e = 0;
      const double f = frexp(need, &e);  // need = f * 2^e, f in [0.5, 1)
      s = ldexp(1.0, f == 0.5 ? e - 1 : e);
      o = floor((double)lo[a] / s) * s;
    }
    frame[a] = (float)o;
    frame[3 + a] = (float)s;
  }
  frame[6] = bad ? 1.0f : 0.0f;
  frame[7] = 0.0f;
}

// Compact node i: its box quantised outward in the frame (the largest q_lo
// with o + q_lo * s <= lo and the smallest q_hi with o + q_hi * s >= hi, both
// checked in the FP32 arithmetic of the walk) and its word.
__global__ void qnodes_kernel(uint32_t n_nodes, const BvhNode* nodes, const float* frame, uint4* qnodes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const BvhNode nd = nodes[i];
  uint32_t q[6] = {0, 0, 0, 0xFFFFu, 0xFFFFu, 0xFFFFu};
  if (frame[6] == 0.0f) {
    for (int a = 0; a < 3; ++a) {
      const float o = frame[a], s = frame[3 + a];
      double ql = floor(((double)nd.lo[a] - (double)o) / (double)s);
      double qh = ceil(((double)nd.hi[a] - (double)o) / (double)s);
      ql = fmin(fmax(ql, 0.0), 65535.0);
      qh = fmin(fmax(qh, 0.0), 65535.0);
      uint32_t l = (uint32_t)ql, h = (uint32_t)qh;
      while (l > 0 && o + (float)l * s > nd.lo[a]) --l;
      while (h < 0xFFFFu && o + (float)h * s < nd.hi[a]) ++h;
      q[a] = l;
      q[3 + a] = h;
    }
  }
  qnodes[i] = make_uint4(q[0] | (q[1] << 16), q[2] | (q[3] << 16), q[4] | (q[5] << 16), qnode_word_of(nd, n_nodes));
}

// Frame and compact nodes of a point scene whose nodes are in their final
// preorder places (kernels only: a captured rebuild records them too).
void build_qnodes(Scene& scene, uint32_t n_inst, hipStream_t stream) {
  const size_t n_nodes = scene.nodes.size();
  require(scene.qnodes.size() == n_nodes && scene.qframe.size() == kQFrameFloats, "compact nodes not sized");
  if (!n_nodes) return;
  qframe_kernel<<<1, 256, 0, stream>>>(n_inst, scene.roots.get(), scene.nodes.get(), scene.qframe.get());
  RTDL_LAUNCH_CHECK();
  qnodes_kernel<<<blocks_for(n_nodes, 256), 256, 0, stream>>>((uint32_t)n_nodes, scene.nodes.get(),
                                                               scene.qframe.get(), scene.qnodes.get());
  RTDL_LAUNCH_CHECK();
}

}  // namespace

void build_scene(Scene& scene, PrimKind kind, int dim, size_t n, const double* coords,
                 const uint32_t* ids, const uint4* poly_refs, const double* poly_verts,
                 size_t n_verts, const std::vector<uint64_t>& instance_offsets,
                 uint32_t leaf_size, hipStream_t stream) {
  require(instance_offsets.size() >= 2 && instance_offsets.front() == 0 &&
              instance_offsets.back() == n,
          "instance offsets must start at 0 and end at the primitive count");
  require(leaf_size >= 1 && leaf_size <= kMaxLeafSize, "leaf size must be in [1, 8]");
  require(n <= kMaxLeafStart, "scene exceeds 2^28 primitives");
  const uint32_t n_inst = (uint32_t)(instance_offsets.size() - 1);
  scene.kind = kind;
  scene.dim = dim;
  scene.leaf_size = leaf_size;
  scene.n_prims = n;
  scene.n_instances = n_inst;
  scene.host_offsets = instance_offsets;

  // leaves per instance
  std::vector<uint64_t> leaf_offsets(n_inst + 1, 0);
  for (uint32_t i = 0; i < n_inst; ++i) {
    const uint64_t ni = instance_offsets[i + 1] - instance_offsets[i];
    require(instance_offsets[i + 1] >= instance_offsets[i], "instance offsets must be non-decreasing");
    leaf_offsets[i + 1] = leaf_offsets[i] + (ni + leaf_size - 1) / leaf_size;
  }
  const size_t m = leaf_offsets[n_inst];
  scene.n_leaves = m;
  if (kind == PrimKind::Polygons && n_verts) {
    scene.poly_verts.resize(2 * n_verts);
    RTDL_HIP(hipMemcpyAsync(scene.poly_verts.get(), poly_verts, 2 * n_verts * sizeof(double),
                            hipMemcpyDeviceToDevice, stream));
  }
  scene.host_roots.assign(n_inst, kEndNode);
  scene.host_ends.assign(n_inst, kEndNode);
  scene.roots.resize(n_inst ? n_inst : 1);
  scene.ends.resize(n_inst ? n_inst : 1);
  if (n == 0) {
    scene.m = 0.0f;
    scene.fp32_exact = true;
    RTDL_HIP(hipMemcpyAsync(scene.roots.get(), scene.host_roots.data(), n_inst * sizeof(uint32_t),
                            hipMemcpyHostToDevice, stream));
    RTDL_HIP(hipMemcpyAsync(scene.ends.get(), scene.host_ends.data(), n_inst * sizeof(uint32_t),
                            hipMemcpyHostToDevice, stream));
    RTDL_HIP(hipStreamSynchronize(stream));
    return;
  }

  // 1. bounds
  std::vector<BoundsTask> tasks;
  constexpr uint32_t kChunk = 16384;
  for (uint32_t i = 0; i < n_inst; ++i)
    for (uint64_t b = instance_offsets[i]; b < instance_offsets[i + 1]; b += kChunk)
      tasks.push_back({i, (uint32_t)b, (uint32_t)std::min<uint64_t>(b + kChunk, instance_offsets[i + 1])});
  DeviceArray<BoundsTask> d_tasks(tasks.size());
  DeviceArray<float4> box_lo(n), box_hi(n);
  DeviceArray<int> inst_bounds(6 * (size_t)n_inst);
  DeviceArray<uint32_t> flags(2);  // [0] max-abs bits, [1] exact flag
  std::vector<int> init_bounds(6 * (size_t)n_inst);
  for (uint32_t i = 0; i < n_inst; ++i)
    for (int a = 0; a < 3; ++a) {
      init_bounds[6 * i + a] = 0x7FFFFFFF;
      init_bounds[6 * i + 3 + a] = (int)0x80000000;
    }
  const uint32_t init_flags[2] = {0u, 1u};
  RTDL_HIP(hipMemcpyAsync(d_tasks.get(), tasks.data(), tasks.size() * sizeof(BoundsTask),
                          hipMemcpyHostToDevice, stream));
  RTDL_HIP(hipMemcpyAsync(inst_bounds.get(), init_bounds.data(), init_bounds.size() * sizeof(int),
                          hipMemcpyHostToDevice, stream));
  RTDL_HIP(hipMemcpyAsync(flags.get(), init_flags, sizeof(init_flags), hipMemcpyHostToDevice, stream));
  bounds_kernel<<<(unsigned)tasks.size(), 256, 0, stream>>>(kind, d_tasks.get(), coords, poly_refs,
                                                             poly_verts, box_lo.get(), box_hi.get(),
                                                             inst_bounds.get(), flags.get(),
                                                             flags.get() + 1);
  RTDL_LAUNCH_CHECK();
  std::vector<int> enc_bounds(6 * (size_t)n_inst);
  uint32_t host_flags[2];
  RTDL_HIP(hipMemcpyAsync(enc_bounds.data(), inst_bounds.get(), enc_bounds.size() * sizeof(int),
                          hipMemcpyDeviceToHost, stream));
  RTDL_HIP(hipMemcpyAsync(host_flags, flags.get(), sizeof(host_flags), hipMemcpyDeviceToHost, stream));
  RTDL_HIP(hipStreamSynchronize(stream));
  std::memcpy(&scene.m, &host_flags[0], sizeof(float));
  scene.fp32_exact = host_flags[1] != 0;
  std::vector<float> fbounds(enc_bounds.size());
  for (size_t i = 0; i < enc_bounds.size(); ++i) fbounds[i] = float_from_ordered(enc_bounds[i]);

  // 2. Morton keys and sort
  DeviceArray<float> d_fbounds(fbounds.size());
  DeviceArray<uint64_t> d_inst_offsets(n_inst + 1), d_leaf_offsets(n_inst + 1);
  RTDL_HIP(hipMemcpyAsync(d_fbounds.get(), fbounds.data(), fbounds.size() * sizeof(float),
                          hipMemcpyHostToDevice, stream));
  RTDL_HIP(hipMemcpyAsync(d_inst_offsets.get(), instance_offsets.data(),
                          (n_inst + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
  RTDL_HIP(hipMemcpyAsync(d_leaf_offsets.get(), leaf_offsets.data(), (n_inst + 1) * sizeof(uint64_t),
                          hipMemcpyHostToDevice, stream));
  DeviceArray<uint64_t> keys(n);
  DeviceArray<uint32_t> order(n);
  morton_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(n, dim, box_lo.get(), box_hi.get(),
                                                         d_inst_offsets.get(), n_inst,
                                                         d_fbounds.get(), keys.get(), order.get());
  RTDL_LAUNCH_CHECK();
  int inst_bits = 0;
  while ((1u << inst_bits) < n_inst) ++inst_bits;
  radix_sort_pairs(keys.get(), order.get(), n, 32 + inst_bits, stream);

  // 3. gather into Morton order
  if (kind == PrimKind::Points) scene.prims.resize(n);
  if (kind == PrimKind::Triangles) scene.prims.resize(3 * n);
  if (kind == PrimKind::Polygons) {
    scene.prims.reset();
    scene.poly.resize(n);
    scene.poly_box.resize(4 * n);
  }
  scene.exact.reset();
  if (!scene.fp32_exact && kind != PrimKind::Polygons)
    scene.exact.resize((kind == PrimKind::Points ? 3 : 9) * n);
  scene.slot_to_input.resize(n);
  DeviceArray<float4> sorted_lo(n), sorted_hi(n);
  gather_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(
      kind, n, order.get(), coords, ids, poly_refs, box_lo.get(), box_hi.get(), scene.prims.get(),
      scene.exact.get(), scene.slot_to_input.get(), scene.poly.get(), scene.poly_box.get(),
      sorted_lo.get(), sorted_hi.get());
  RTDL_LAUNCH_CHECK();
  if (kind == PrimKind::Polygons) {
    poly_box_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(n, scene.poly.get(), scene.poly_verts.get(),
                                                             scene.poly_box.get());
    RTDL_LAUNCH_CHECK();
  }

  // 4. leaves + Karras hierarchy
  DeviceArray<uint32_t> leaf_first(m), leaf_count(m);
  DeviceArray<uint64_t> leaf_keys(m);
  leaves_kernel<<<blocks_for(m, 256), 256, 0, stream>>>(m, d_leaf_offsets.get(), d_inst_offsets.get(),
                                                         n_inst, leaf_size, keys.get(),
                                                         leaf_first.get(), leaf_count.get(),
                                                         leaf_keys.get());
  RTDL_LAUNCH_CHECK();
  const size_t n_nodes = 2 * m - 1;
  scene.nodes.resize(n_nodes);
  const size_t n_internal = m - 1;
  DeviceArray<uint32_t> left(n_internal ? n_internal : 1), right(n_internal ? n_internal : 1);
  DeviceArray<uint32_t> parent(n_nodes), range_lo(n_internal ? n_internal : 1),
      range_hi(n_internal ? n_internal : 1), done(n_internal ? n_internal : 1);
  DeviceArray<unsigned long long> highest(m);
  RTDL_HIP(hipMemsetAsync(highest.get(), 0, m * sizeof(unsigned long long), stream));
  RTDL_HIP(hipMemsetAsync(done.get(), 0xFF, done.bytes(), stream));
  RTDL_HIP(hipMemsetAsync(parent.get(), 0xFF, parent.bytes(), stream));
  if (n_internal) {
    karras_kernel<<<blocks_for(n_internal, 256), 256, 0, stream>>>(
        (int)m, leaf_keys.get(), left.get(), right.get(), parent.get(), range_lo.get(),
        range_hi.get(), highest.get());
    RTDL_LAUNCH_CHECK();
    links_kernel<<<blocks_for(n_internal, 256), 256, 0, stream>>>(
        (uint32_t)m, right.get(), range_hi.get(), highest.get(), scene.nodes.get());
    RTDL_LAUNCH_CHECK();
  }
  // 5. leaf records and ropes
  leaf_nodes_kernel<<<blocks_for(m, 256), 256, 0, stream>>>(
      (uint32_t)m, leaf_first.get(), leaf_count.get(), sorted_lo.get(), sorted_hi.get(), highest.get(),
      scene.nodes.get());
  RTDL_LAUNCH_CHECK();
  // 6. bottom-up refit, one launch per tree level, in groups of 16 launches
  // until the root (internal node 0) is done; a Karras tree over 64-bit keys
  // with index tie-breaks is at most about 100 levels deep
  if (n_internal) {
    uint32_t root_done = kNotRefit;
    uint32_t round = 1;
    while (root_done == kNotRefit) {
      require(round < 4096, "LBVH refit did not reach the root (malformed hierarchy)");
      for (int r = 0; r < 16; ++r, ++round) {
        refit_round_kernel<<<blocks_for(n_internal, 256), 256, 0, stream>>>(
            (uint32_t)n_internal, round, left.get(), right.get(), done.get(), scene.nodes.get());
        RTDL_LAUNCH_CHECK();
      }
      RTDL_HIP(hipMemcpyAsync(&root_done, done.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
      RTDL_HIP(hipStreamSynchronize(stream));
    }
  }
  roots_kernel<<<blocks_for(n_inst, 64), 64, 0, stream>>>(
      n_inst, (uint32_t)m, d_leaf_offsets.get(), parent.get(), range_lo.get(), range_hi.get(),
      scene.nodes.get(), scene.roots.get(), scene.ends.get());
  RTDL_LAUNCH_CHECK();
  // 7. store the nodes in depth-first preorder
  {
    DeviceArray<uint32_t> pre(n_nodes);
    DeviceArray<BvhNode> ordered(n_nodes);
    preorder_kernel<<<blocks_for(n_nodes, 256), 256, 0, stream>>>((uint32_t)m, parent.get(), left.get(),
                                                                   range_lo.get(), pre.get());
    RTDL_LAUNCH_CHECK();
    permute_nodes_kernel<<<blocks_for(n_nodes, 256), 256, 0, stream>>>((uint32_t)n_nodes, scene.nodes.get(),
                                                                        pre.get(), ordered.get());
    RTDL_LAUNCH_CHECK();
    remap_roots_kernel<<<blocks_for(n_inst, 64), 64, 0, stream>>>(n_inst, pre.get(), scene.roots.get(),
                                                                  scene.ends.get());
    RTDL_LAUNCH_CHECK();
    RTDL_HIP(hipStreamSynchronize(stream));  // before the old nodes return to the pool
    scene.nodes = std::move(ordered);
  }
  // 8. point scenes: the compact nodes of the fixed-radius count walk
  if (kind == PrimKind::Points) {
    scene.qnodes.resize(n_nodes);
    scene.qframe.resize(kQFrameFloats);
    build_qnodes(scene, n_inst, stream);
  }
  RTDL_HIP(hipMemcpyAsync(scene.host_roots.data(), scene.roots.get(), n_inst * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, stream));
  RTDL_HIP(hipMemcpyAsync(scene.host_ends.data(), scene.ends.get(), n_inst * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, stream));
  RTDL_HIP(hipStreamSynchronize(stream));
}


// ------------------------------------------------------------- rebuilds --

struct RebuildWork {
  PrimKind kind = PrimKind::Points;
  size_t n = 0, m = 0, n_nodes = 0, n_internal = 0, n_tasks = 0;
  uint32_t n_inst = 0;
  int inst_bits = 0;
  DeviceArray<uint64_t> inst_offsets, leaf_offsets;
  DeviceArray<double> coords;
  DeviceArray<uint32_t> ids;
  DeviceArray<BoundsTask> tasks;
  DeviceArray<float4> box_lo, box_hi, sorted_lo, sorted_hi;
  DeviceArray<int> inst_bounds;
  DeviceArray<uint32_t> flags;  // [0] M bits (read by traversals), [1] exact flag, [2] status
  DeviceArray<float> fbounds;
  DeviceArray<uint64_t> keys, leaf_keys;
  DeviceArray<uint32_t> order, leaf_first, leaf_count, left, right, parent, range_lo, range_hi, done, pre,
      roots, ends;
  DeviceArray<unsigned long long> highest;
  DeviceArray<BvhNode> work_nodes;
  SortWorkspace sort_ws;
};

namespace {

// The initial values build_scene uploads or memsets, written by a kernel.
__global__ void rebuild_init_kernel(uint32_t n_inst, int* inst_bounds, uint32_t* flags, size_t m,
                                    unsigned long long* highest, size_t n_internal, uint32_t* done,
                                    size_t n_nodes, uint32_t* parent) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 6 * (size_t)n_inst) inst_bounds[i] = (i % 6) < 3 ? 0x7FFFFFFF : (int)0x80000000;
  if (i == 0) {
    flags[0] = 0u;
    flags[1] = 1u;
    flags[2] = 0u;
  }
  if (i < m) highest[i] = 0ull;
  if (i < n_internal) done[i] = kNotRefit;
  if (i < n_nodes) parent[i] = kEndNode;
}

__device__ inline float float_from_ordered_dev(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void decode_bounds_kernel(size_t n, const int* enc, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = float_from_ordered_dev(enc[i]);
}

// Error bits of a rebuild: the refit reached the root, the instance roots kept
// their preorder positions (the jobs' scene views hold them) and every
// coordinate stayed exact in FP32 (the scene keeps no FP64 copy).
__global__ void rebuild_check_kernel(uint32_t n_inst, const uint32_t* roots, const uint32_t* ends,
                                     const uint32_t* scene_roots, const uint32_t* scene_ends,
                                     size_t n_internal, const uint32_t* done, uint32_t* flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    if (n_internal && done[0] == kNotRefit) atomicOr(&flags[2], kRebuildRefitIncomplete);
    if (flags[1] == 0u) atomicOr(&flags[2], kRebuildNotExact);
  }
  if (i < n_inst && (roots[i] != scene_roots[i] || ends[i] != scene_ends[i]))
    atomicOr(&flags[2], kRebuildRootsMoved);
}

// Refit rounds of a rebuild: a Karras tree over 64-bit keys with index
// tie-breaks is at most about 100 levels deep, and rounds after the root's
// return at once.
constexpr uint32_t kRebuildRefitRounds = 128;

}  // namespace

void prepare_rebuild(Scene& scene, hipStream_t stream) {
  require(scene.kind == PrimKind::Points || scene.kind == PrimKind::Triangles,
          "only point and triangle instance scenes can be rebuilt");
  require(scene.fp32_exact, "rebuilds need FP32-exact input (instance scenes from FP32 coordinates)");
  if (scene.rebuild) return;
  auto w = std::make_shared<RebuildWork>();
  const size_t n = scene.n_prims;
  const uint32_t n_inst = (uint32_t)scene.n_instances;
  w->kind = scene.kind;
  w->n = n;
  w->n_inst = n_inst;
  w->m = scene.n_leaves;
  w->n_nodes = n ? 2 * w->m - 1 : 0;
  w->n_internal = n ? w->m - 1 : 0;
  while ((1u << w->inst_bits) < n_inst) ++w->inst_bits;
  const std::vector<uint64_t>& offs = scene.host_offsets;
  std::vector<uint64_t> leaf_offsets(n_inst + 1, 0);
  for (uint32_t i = 0; i < n_inst; ++i)
    leaf_offsets[i + 1] = leaf_offsets[i] + (offs[i + 1] - offs[i] + scene.leaf_size - 1) / scene.leaf_size;
  std::vector<BoundsTask> tasks;
  constexpr uint32_t kChunk = 16384;  // as build_scene
  for (uint32_t i = 0; i < n_inst; ++i)
    for (uint64_t b = offs[i]; b < offs[i + 1]; b += kChunk)
      tasks.push_back({i, (uint32_t)b, (uint32_t)std::min<uint64_t>(b + kChunk, offs[i + 1])});
  w->n_tasks = tasks.size();
  const size_t per = scene.kind == PrimKind::Points ? 3 : 9;
  auto at_least1 = [](size_t x) { return x ? x : (size_t)1; };
  w->inst_offsets.resize(n_inst + 1);
  w->leaf_offsets.resize(n_inst + 1);
  w->coords.resize(at_least1(per * n));
  w->ids.resize(at_least1(n));
  w->tasks.resize(at_least1(tasks.size()));
  w->box_lo.resize(at_least1(n));
  w->box_hi.resize(at_least1(n));
  w->sorted_lo.resize(at_least1(n));
  w->sorted_hi.resize(at_least1(n));
  w->inst_bounds.resize(at_least1(6 * (size_t)n_inst));
  w->flags.resize(3);
  w->fbounds.resize(at_least1(6 * (size_t)n_inst));
  w->keys.resize(at_least1(n));
  w->order.resize(at_least1(n));
  w->leaf_keys.resize(at_least1(w->m));
  w->leaf_first.resize(at_least1(w->m));
  w->leaf_count.resize(at_least1(w->m));
  w->left.resize(at_least1(w->n_internal));
  w->right.resize(at_least1(w->n_internal));
  w->range_lo.resize(at_least1(w->n_internal));
  w->range_hi.resize(at_least1(w->n_internal));
  w->done.resize(at_least1(w->n_internal));
  w->parent.resize(at_least1(w->n_nodes));
  w->pre.resize(at_least1(w->n_nodes));
  w->highest.resize(at_least1(w->m));
  w->work_nodes.resize(at_least1(w->n_nodes));
  w->roots.resize(at_least1(n_inst));
  w->ends.resize(at_least1(n_inst));
  w->sort_ws.reserve(at_least1(n));
  RTDL_HIP(hipMemcpyAsync(w->inst_offsets.get(), offs.data(), (n_inst + 1) * sizeof(uint64_t),
                          hipMemcpyHostToDevice, stream));
  RTDL_HIP(hipMemcpyAsync(w->leaf_offsets.get(), leaf_offsets.data(), (n_inst + 1) * sizeof(uint64_t),
                          hipMemcpyHostToDevice, stream));
  if (!tasks.empty())
    RTDL_HIP(hipMemcpyAsync(w->tasks.get(), tasks.data(), tasks.size() * sizeof(BoundsTask),
                            hipMemcpyHostToDevice, stream));
  // M of the current tree, in device memory from now on
  uint32_t m_bits;
  std::memcpy(&m_bits, &scene.m, sizeof(m_bits));
  const uint32_t init[3] = {m_bits, 1u, 0u};
  RTDL_HIP(hipMemcpyAsync(w->flags.get(), init, sizeof(init), hipMemcpyHostToDevice, stream));
  RTDL_HIP(hipStreamSynchronize(stream));
  scene.m_device = reinterpret_cast<const float*>(w->flags.get());
  scene.rebuild = std::move(w);
}

void rebuild_scene(Scene& scene, const float* coords, hipStream_t stream) {
  require(scene.rebuild != nullptr, "scene was not prepared for rebuilds");
  RebuildWork& w = *scene.rebuild;
  const size_t n = w.n;
  const uint32_t n_inst = w.n_inst;
  if (!n) return;
  require(coords != nullptr, "coords must not be null");
  // 0. FP64 coordinates and instance-local ids, as the first build
  if (w.kind == PrimKind::Points) {
    widen_points(coords, n, w.inst_offsets.get(), n_inst, w.coords.get(), w.ids.get(), stream);
  } else {
    widen_f32(coords, 9 * n, w.coords.get(), stream);
    instance_ids(n, w.inst_offsets.get(), n_inst, w.ids.get(), stream);
  }
  const size_t init_n = std::max({6 * (size_t)n_inst, w.m, w.n_nodes, (size_t)1});
  rebuild_init_kernel<<<blocks_for(init_n, 256), 256, 0, stream>>>(
      n_inst, w.inst_bounds.get(), w.flags.get(), w.m, w.highest.get(), w.n_internal, w.done.get(),
      w.n_nodes, w.parent.get());
  RTDL_LAUNCH_CHECK();
  // 1. bounds, decoded on the device
  bounds_kernel<<<(unsigned)w.n_tasks, 256, 0, stream>>>(scene.kind, w.tasks.get(), w.coords.get(), nullptr,
                                                         nullptr, w.box_lo.get(), w.box_hi.get(),
                                                         w.inst_bounds.get(), w.flags.get(), w.flags.get() + 1);
  RTDL_LAUNCH_CHECK();
  decode_bounds_kernel<<<blocks_for(6 * (size_t)n_inst, 256), 256, 0, stream>>>(
      6 * (size_t)n_inst, w.inst_bounds.get(), w.fbounds.get());
  RTDL_LAUNCH_CHECK();
  // 2. Morton keys and sort
  morton_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(n, scene.dim, w.box_lo.get(), w.box_hi.get(),
                                                         w.inst_offsets.get(), n_inst, w.fbounds.get(),
                                                         w.keys.get(), w.order.get());
  RTDL_LAUNCH_CHECK();
  radix_sort_pairs_async(w.keys.get(), w.order.get(), n, 32 + w.inst_bits, w.sort_ws, stream);
  // 3. gather into Morton order, into the scene's own arrays
  gather_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(
      scene.kind, n, w.order.get(), w.coords.get(), w.ids.get(), nullptr, w.box_lo.get(), w.box_hi.get(),
      scene.prims.get(), nullptr, scene.slot_to_input.get(), nullptr, nullptr, w.sorted_lo.get(),
      w.sorted_hi.get());
  RTDL_LAUNCH_CHECK();
  // 4. leaves + Karras hierarchy, 5. leaf records and ropes (in the work nodes)
  leaves_kernel<<<blocks_for(w.m, 256), 256, 0, stream>>>(w.m, w.leaf_offsets.get(), w.inst_offsets.get(),
                                                           n_inst, scene.leaf_size, w.keys.get(),
                                                           w.leaf_first.get(), w.leaf_count.get(),
                                                           w.leaf_keys.get());
  RTDL_LAUNCH_CHECK();
  if (w.n_internal) {
    karras_kernel<<<blocks_for(w.n_internal, 256), 256, 0, stream>>>(
        (int)w.m, w.leaf_keys.get(), w.left.get(), w.right.get(), w.parent.get(), w.range_lo.get(),
        w.range_hi.get(), w.highest.get());
    RTDL_LAUNCH_CHECK();
    links_kernel<<<blocks_for(w.n_internal, 256), 256, 0, stream>>>(
        (uint32_t)w.m, w.right.get(), w.range_hi.get(), w.highest.get(), w.work_nodes.get());
    RTDL_LAUNCH_CHECK();
  }
  leaf_nodes_kernel<<<blocks_for(w.m, 256), 256, 0, stream>>>(
      (uint32_t)w.m, w.leaf_first.get(), w.leaf_count.get(), w.sorted_lo.get(), w.sorted_hi.get(),
      w.highest.get(), w.work_nodes.get());
  RTDL_LAUNCH_CHECK();
  // 6. refit: a fixed number of rounds instead of a host poll of the root
  if (w.n_internal)
    for (uint32_t round = 1; round <= kRebuildRefitRounds; ++round) {
      refit_round_kernel<<<blocks_for(w.n_internal, 256), 256, 0, stream>>>(
          (uint32_t)w.n_internal, round, w.left.get(), w.right.get(), w.done.get(), w.work_nodes.get());
      RTDL_LAUNCH_CHECK();
    }
  roots_kernel<<<blocks_for(n_inst, 64), 64, 0, stream>>>(
      n_inst, (uint32_t)w.m, w.leaf_offsets.get(), w.parent.get(), w.range_lo.get(), w.range_hi.get(),
      w.work_nodes.get(), w.roots.get(), w.ends.get());
  RTDL_LAUNCH_CHECK();
  // 7. preorder, written into the scene's node array
  preorder_kernel<<<blocks_for(w.n_nodes, 256), 256, 0, stream>>>((uint32_t)w.m, w.parent.get(), w.left.get(),
                                                                   w.range_lo.get(), w.pre.get());
  RTDL_LAUNCH_CHECK();
  permute_nodes_kernel<<<blocks_for(w.n_nodes, 256), 256, 0, stream>>>((uint32_t)w.n_nodes, w.work_nodes.get(),
                                                                        w.pre.get(), scene.nodes.get());
  RTDL_LAUNCH_CHECK();
  remap_roots_kernel<<<blocks_for(n_inst, 64), 64, 0, stream>>>(n_inst, w.pre.get(), w.roots.get(), w.ends.get());
  RTDL_LAUNCH_CHECK();
  rebuild_check_kernel<<<blocks_for(std::max<size_t>(n_inst, 1), 64), 64, 0, stream>>>(
      n_inst, w.roots.get(), w.ends.get(), scene.roots.get(), scene.ends.get(), w.n_internal, w.done.get(),
      w.flags.get());
  RTDL_LAUNCH_CHECK();
  // 8. the compact nodes (the instance roots keep their places, as checked)
  if (scene.qnodes.size()) build_qnodes(scene, n_inst, stream);
}

uint32_t rebuild_status(const Scene& scene, hipStream_t stream) {
  require(scene.rebuild != nullptr, "scene was not prepared for rebuilds");
  uint32_t st = 0;
  RTDL_HIP(hipMemcpyAsync(&st, scene.rebuild->flags.get() + 2, sizeof(st), hipMemcpyDeviceToHost, stream));
  RTDL_HIP(hipStreamSynchronize(stream));
  return st;
}

}  // namespace rtdl
