// Traversal + refine kernels for fixed-radius / KNN point queries, 3-D ray
// queries against triangles, and point-in-polygon.
//
// One lane per query.  Queries run in Morton order (via `order`) so the 64
// lanes of a wavefront walk nearly the same part of the tree; the stackless
// rope traversal keeps per-lane state in a few registers.  Box and primitive
// prefilters run in FP32 against conservative bounds (common.hpp); every
// candidate that survives is decided in FP64 with the reference's operation
// order (this file is compiled with -ffp-contract=off).
#include <cmath>

#include "common.hpp"
#include "primitives.hpp"
#include "queries.hpp"

namespace rtdl {
namespace {

__device__ inline void load_box(const BvhNode& nd, float lo[3], float hi[3]) {
  lo[0] = nd.lo[0];
  lo[1] = nd.lo[1];
  lo[2] = nd.lo[2];
  hi[0] = nd.hi[0];
  hi[1] = nd.hi[1];
  hi[2] = nd.hi[2];
}

__device__ inline void flush_stats(TraversalStats* stats, unsigned long long nodes,
                                   unsigned long long prims) {
  // one atomic per wavefront
  for (int d = 32; d >= 1; d >>= 1) {
    nodes += __shfl_xor(nodes, d, 64);
    prims += __shfl_xor(prims, d, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&stats->nodes, nodes);
    atomicAdd(&stats->prims, prims);
  }
}

// ------------------------------------------------------------ point queries --

// Sorted (distance, id) list of capacity K holding the k smallest entries in
// slots [K - k, K): slots below K - k hold a -1 sentinel that never moves, so
// the k-th smallest is always slot K - 1 and every index is static.
template <int K>
struct TopK {
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

template <int K, bool COUNT_ALL, bool STATS>
__global__ __launch_bounds__(256) void collect_kernel(SceneView s, const double* __restrict__ q,
                                                      const uint32_t* __restrict__ order,
                                                      uint32_t nq, double radius, uint32_t k,
                                                      double* __restrict__ list_d,
                                                      uint32_t* __restrict__ list_id,
                                                      uint32_t* __restrict__ count,
                                                      uint32_t* __restrict__ total,
                                                      TraversalStats* stats) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_nodes = 0, n_prims = 0;
  if (t < nq) {
    const uint32_t qi = order ? order[t] : t;
    const double qx = q[3 * (size_t)qi], qy = q[3 * (size_t)qi + 1], qz = q[3 * (size_t)qi + 2];
    const float qf[3] = {(float)qx, (float)qy, (float)qz};
    const float m = coord_bound(s.m, qx, qy, qz);
    const double radius_sq = radius * radius;  // reference.py:357 radius_sq = radius * radius
    float thr2 = frn_threshold2(radius, m);
    TopK<K> list;
    list.init(k);
    uint32_t inside = 0;
    uint32_t node = s.root;
    // a preorder walk visits each node at most once; the guard only bounds
    // the loop should the tree ever be malformed
    for (uint32_t guard = s.n_nodes + 1; node != s.end && guard; --guard) {
      const BvhNode nd = s.nodes[node];
      if (STATS) ++n_nodes;
      float lo[3], hi[3];
      load_box(nd, lo, hi);
      if (!(box_dist2(qf, lo, hi) <= thr2)) {
        node = nd.escape;
        continue;
      }
      if (!node_is_leaf(nd.link)) {
        node = nd.link;
        continue;
      }
      const uint32_t first = leaf_first(nd.link), cnt = leaf_count(nd.link);
      for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t slot = first + j;
        const float4 p = s.prims[slot];
        if (STATS) ++n_prims;
        const float fx = p.x - qf[0], fy = p.y - qf[1], fz = p.z - qf[2];
        const float d2f = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
        if (!(d2f <= thr2)) continue;
        double px, py, pz;
        if (s.exact) {
          px = s.exact[3 * (size_t)slot];
          py = s.exact[3 * (size_t)slot + 1];
          pz = s.exact[3 * (size_t)slot + 2];
        } else {
          px = p.x;
          py = p.y;
          pz = p.z;
        }
        // reference.py:882-890: dx = right - left (search - query), dx*dx + dy*dy + dz*dz
        const double dx = px - qx, dy = py - qy, dz = pz - qz;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > radius_sq) continue;  // reference.py:361 inclusive radius
        ++inside;
        const double dist = sqrt(d2);
        const uint32_t ident = __float_as_uint(p.w);
        if (list.accepts(dist, ident)) {
          list.insert(dist, ident);
          if (!COUNT_ALL && list.full()) thr2 = frn_threshold2(fmin(radius, list.d[K - 1]), m);
        }
      }
      node = nd.escape;
    }
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (i >= K - (int)k && list.d[i] < __builtin_inf()) {
        const uint32_t r = (uint32_t)(i - (K - (int)k));
        list_d[(size_t)qi * k + r] = list.d[i];
        list_id[(size_t)qi * k + r] = list.id[i];
        ++c;
      }
    }
    count[qi] = c;
    if (COUNT_ALL) total[qi] = inside;
  }
  if (STATS) flush_stats(stats, n_nodes, n_prims);
}

__global__ __launch_bounds__(256) void collect_all_kernel(SceneView s, const double* __restrict__ q,
                                                          const uint32_t* __restrict__ queries,
                                                          uint32_t n, double radius,
                                                          const uint64_t* __restrict__ offset,
                                                          double* __restrict__ out_d,
                                                          uint32_t* __restrict__ out_id) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t qi = queries[t];
  const double qx = q[3 * (size_t)qi], qy = q[3 * (size_t)qi + 1], qz = q[3 * (size_t)qi + 2];
  const float qf[3] = {(float)qx, (float)qy, (float)qz};
  const float m = coord_bound(s.m, qx, qy, qz);
  const double radius_sq = radius * radius;
  const float thr2 = frn_threshold2(radius, m);
  uint64_t w = offset[t];
  const uint64_t w_end = offset[t + 1];
  uint32_t node = s.root;
  for (uint32_t guard = s.n_nodes + 1; node != s.end && guard; --guard) {
    const BvhNode nd = s.nodes[node];
    float lo[3], hi[3];
    load_box(nd, lo, hi);
    if (!(box_dist2(qf, lo, hi) <= thr2)) {
      node = nd.escape;
      continue;
    }
    if (!node_is_leaf(nd.link)) {
      node = nd.link;
      continue;
    }
    const uint32_t first = leaf_first(nd.link), cnt = leaf_count(nd.link);
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t slot = first + j;
      const float4 p = s.prims[slot];
      const float fx = p.x - qf[0], fy = p.y - qf[1], fz = p.z - qf[2];
      const float d2f = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
      if (!(d2f <= thr2)) continue;
      double px, py, pz;
      if (s.exact) {
        px = s.exact[3 * (size_t)slot];
        py = s.exact[3 * (size_t)slot + 1];
        pz = s.exact[3 * (size_t)slot + 2];
      } else {
        px = p.x;
        py = p.y;
        pz = p.z;
      }
      const double dx = px - qx, dy = py - qy, dz = pz - qz;
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 > radius_sq) continue;
      if (w < w_end) {
        out_d[w] = sqrt(d2);
        out_id[w] = __float_as_uint(p.w);
      }
      ++w;
    }
    node = nd.escape;
  }
}

template <int K, bool COUNT_ALL, bool STATS>
void launch_collect_t(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq,
                      double radius, uint32_t k, double* list_d, uint32_t* list_id,
                      uint32_t* count, uint32_t* total, TraversalStats* stats, int block,
                      hipStream_t stream) {
  collect_kernel<K, COUNT_ALL, STATS><<<blocks_for(nq, block), block, 0, stream>>>(
      s, q, order, nq, radius, k, list_d, list_id, count, total, stats);
  RTDL_LAUNCH_CHECK();
}

template <bool STATS>
void dispatch_collect(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq,
                      double radius, uint32_t k, bool count_all, double* list_d,
                      uint32_t* list_id, uint32_t* count, uint32_t* total,
                      TraversalStats* stats, int block, hipStream_t stream) {
  if (count_all) {
    require(k == kMaxListK, "count_all collection needs k == 64");
    launch_collect_t<64, true, STATS>(s, q, order, nq, radius, k, list_d, list_id, count, total,
                                      stats, block, stream);
    return;
  }
  if (k <= 1)
    launch_collect_t<1, false, STATS>(s, q, order, nq, radius, k, list_d, list_id, count, total, stats, block, stream);
  else if (k <= 2)
    launch_collect_t<2, false, STATS>(s, q, order, nq, radius, k, list_d, list_id, count, total, stats, block, stream);
  else if (k <= 4)
    launch_collect_t<4, false, STATS>(s, q, order, nq, radius, k, list_d, list_id, count, total, stats, block, stream);
  else if (k <= 8)
    launch_collect_t<8, false, STATS>(s, q, order, nq, radius, k, list_d, list_id, count, total, stats, block, stream);
  else if (k <= 16)
    launch_collect_t<16, false, STATS>(s, q, order, nq, radius, k, list_d, list_id, count, total, stats, block, stream);
  else if (k <= 32)
    launch_collect_t<32, false, STATS>(s, q, order, nq, radius, k, list_d, list_id, count, total, stats, block, stream);
  else
    launch_collect_t<64, false, STATS>(s, q, order, nq, radius, k, list_d, list_id, count, total, stats, block, stream);
}

// -------------------------------------------------------------------- rays --

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

// Conservative FP32 test of the segment o + s * D, s in [0, slim], against
// the box grown by `pad` on every side.  `pad` (2^-18 of the coordinate
// scale) dominates every FP32 rounding of o, D, the box and the slab
// parameters, and the exit parameters get a further (1 + 2^-20) factor.
__device__ inline bool segment_box(const float o[3], const float D[3], const float inv[3],
                                   const float lo[3], const float hi[3], float pad, float slim) {
  float tn = -9.5367431640625e-07f;  // -2^-20
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

template <int MODE, bool STATS>
__global__ __launch_bounds__(256) void ray_kernel(SceneView s, const double* __restrict__ rays,
                                                  const uint32_t* __restrict__ order, uint32_t nr,
                                                  uint32_t* __restrict__ out_u,
                                                  double* __restrict__ out_t,
                                                  TraversalStats* stats) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_nodes = 0, n_prims = 0;
  if (t < nr) {
    const uint32_t ri = order ? order[t] : t;
    const double* r = rays + 7 * (size_t)ri;
    const double ox = r[0], oy = r[1], oz = r[2], dx = r[3], dy = r[4], dz = r[5], tmax = r[6];
    uint32_t hits = 0;
    double best_t = __builtin_nan("");
    uint32_t best_id = 0;
    // t in [0, tmax] is impossible for a negative or NaN tmax (reference.py:844)
    if (tmax >= 0.0 && s.root != kEndNode) {
      const bool bounded = tmax < __builtin_inf();
      float o32[3] = {(float)ox, (float)oy, (float)oz};
      float D32[3] = {0.0f, 0.0f, 0.0f};
      float inv[3] = {0.0f, 0.0f, 0.0f};
      float pad = __builtin_inff();
      if (bounded) {
        const double D[3] = {dx * tmax, dy * tmax, dz * tmax};
        double scale = (double)s.m;
        double omax = fmax(fmax(fabs(ox), fabs(oy)), fabs(oz));
        double dmax = fmax(fmax(fabs(D[0]), fabs(D[1])), fabs(D[2]));
        pad = __double2float_ru((scale + omax + dmax) * 3.814697265625e-06);  // 2^-18
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          D32[a] = (float)D[a];
          inv[a] = D32[a] != 0.0f ? 1.0f / D32[a] : 0.0f;
        }
      }
      float slim = 1.0f;
      uint32_t node = s.root;
      bool done = false;
      for (uint32_t guard = s.n_nodes + 1; node != s.end && !done && guard; --guard) {
        const BvhNode nd = s.nodes[node];
        if (STATS) ++n_nodes;
        float lo[3], hi[3];
        load_box(nd, lo, hi);
        if (!segment_box(o32, D32, inv, lo, hi, pad, slim)) {
          node = nd.escape;
          continue;
        }
        if (!node_is_leaf(nd.link)) {
          node = nd.link;
          continue;
        }
        const uint32_t first = leaf_first(nd.link), cnt = leaf_count(nd.link);
        for (uint32_t j = 0; j < cnt; ++j) {
          const uint32_t slot = first + j;
          if (STATS) ++n_prims;
          double v[9];
          uint32_t ident;
          const float4 a = s.prims[3 * (size_t)slot];
          ident = __float_as_uint(a.w);
          if (s.exact) {
#pragma unroll
            for (int e = 0; e < 9; ++e) v[e] = s.exact[9 * (size_t)slot + e];
          } else {
            const float4 b = s.prims[3 * (size_t)slot + 1];
            const float4 c = s.prims[3 * (size_t)slot + 2];
            v[0] = a.x; v[1] = a.y; v[2] = a.z;
            v[3] = b.x; v[4] = b.y; v[5] = b.z;
            v[6] = c.x; v[7] = c.y; v[8] = c.z;
          }
          double th;
          if (!mt_hit(ox, oy, oz, dx, dy, dz, tmax, v, &th)) continue;
          ++hits;
          if (MODE == 0) {
            done = true;
            break;
          }
          if (MODE == 2) {
            if (!(best_t == best_t) || th < best_t || (th == best_t && ident < best_id)) {
              best_t = th;
              best_id = ident;
              if (bounded && tmax > 0.0) slim = fminf(1.0f, __double2float_ru(th / tmax));
            }
          }
        }
        node = nd.escape;
      }
    }
    if (MODE == 0) out_u[ri] = hits ? 1u : 0u;
    if (MODE == 1) out_u[ri] = hits;
    if (MODE == 2) {
      out_u[ri] = best_id;
      out_t[ri] = best_t;
    }
  }
  if (STATS) flush_stats(stats, n_nodes, n_prims);
}

template <int MODE>
void launch_rays_mode(const SceneView& s, const double* rays, const uint32_t* order, uint32_t nr,
                      uint32_t* out_u, double* out_t, TraversalStats* stats, int block,
                      hipStream_t stream) {
  if (stats)
    ray_kernel<MODE, true><<<blocks_for(nr, block), block, 0, stream>>>(s, rays, order, nr, out_u, out_t, stats);
  else
    ray_kernel<MODE, false><<<blocks_for(nr, block), block, 0, stream>>>(s, rays, order, nr, out_u, out_t, stats);
  RTDL_LAUNCH_CHECK();
}

// --------------------------------------------------------- point in polygon --

// reference.py:745-768 (`x ** 2` and `x ** 0.5` are evaluated as the
// correctly rounded x * x and sqrt(x); libm pow differs from them by at most
// one ulp, which can only matter for a point at the 1e-12 tolerance boundary).
__device__ inline bool point_on_segment(double px, double py, double ax, double ay, double bx,
                                        double by, double eps_sq) {
  const double eps = 1.0e-12;
  const double ex = bx - ax, ey = by - ay;
  const double length_sq = ex * ex + ey * ey;
  if (length_sq <= eps_sq) return fabs(px - ax) <= eps && fabs(py - ay) <= eps;
  const double cross = (px - ax) * (by - ay) - (py - ay) * (bx - ax);
  const double length = sqrt(length_sq);
  if (fabs(cross) > eps * length) return false;
  const double dot = (px - ax) * (bx - ax) + (py - ay) * (by - ay);
  const double along_eps = eps * length;
  if (dot < -along_eps) return false;
  if (dot - length_sq > along_eps) return false;
  return true;
}

// reference.py:717-742
__device__ inline bool point_in_polygon(double x, double y, const double* v, uint32_t n,
                                        double eps_sq) {
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t j = (i + 1 == n) ? 0 : i + 1;
    if (point_on_segment(x, y, v[2 * i], v[2 * i + 1], v[2 * j], v[2 * j + 1], eps_sq)) return true;
  }
  bool inside = false;
  uint32_t j = n - 1;
  for (uint32_t i = 0; i < n; ++i) {
    const double xi = v[2 * i], yi = v[2 * i + 1], xj = v[2 * j], yj = v[2 * j + 1];
    if ((yi > y) != (yj > y)) {
      double den = yj - yi;
      if (den == 0.0) den = 1.0e-20;
      if (x <= (xj - xi) * (y - yi) / den + xi) inside = !inside;
    }
    j = i;
  }
  return inside;
}

template <bool WRITE, bool STATS>
__global__ __launch_bounds__(256) void pip_kernel(SceneView s, const double* __restrict__ pts,
                                                  const uint32_t* __restrict__ order, uint32_t np,
                                                  int positive_only, double eps_sq,
                                                  uint32_t* __restrict__ count,
                                                  const uint64_t* __restrict__ offset,
                                                  uint32_t* __restrict__ out_poly,
                                                  TraversalStats* stats) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_nodes = 0, n_prims = 0;
  if (t < np) {
    const uint32_t pi = order ? order[t] : t;
    const double x = pts[2 * (size_t)pi], y = pts[2 * (size_t)pi + 1];
    const float pf[3] = {(float)x, (float)y, 0.0f};
    // Boxes are grown by more than the on-edge tolerance (1e-12 from the
    // boundary) plus FP32 rounding of the point.
    const float m = coord_bound(s.m, x, y, 0.0);
    const float pad = m * 9.5367431640625e-07f + 1.0e-11f;
    uint32_t found = 0;
    uint64_t w = WRITE ? offset[pi] : 0;
    uint32_t node = s.root;
    for (uint32_t guard = s.n_nodes + 1; node != s.end && guard; --guard) {
      const BvhNode nd = s.nodes[node];
      if (STATS) ++n_nodes;
      const bool in = pf[0] >= nd.lo[0] - pad && pf[0] <= nd.hi[0] + pad &&
                      pf[1] >= nd.lo[1] - pad && pf[1] <= nd.hi[1] + pad;
      if (!in) {
        node = nd.escape;
        continue;
      }
      if (!node_is_leaf(nd.link)) {
        node = nd.link;
        continue;
      }
      const uint32_t first = leaf_first(nd.link), cnt = leaf_count(nd.link);
      for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t slot = first + j;
        if (STATS) ++n_prims;
        if (positive_only) {  // reference.py:122-125 exact bbox prefilter
          const double* bb = s.poly_box + 4 * (size_t)slot;
          if (x < bb[0] || x > bb[2] || y < bb[1] || y > bb[3]) continue;
        }
        const uint4 ref = s.poly[slot];
        if (!point_in_polygon(x, y, s.poly_verts + 2 * (size_t)ref.x, ref.y, eps_sq)) continue;
        if (WRITE && w < offset[pi + 1]) {
          // insertion keeps the segment ascending by polygon input index
          const uint32_t pidx = s.slot_to_input[slot];
          uint64_t i = w;
          const uint64_t base = offset[pi];
          while (i > base && out_poly[i - 1] > pidx) {
            out_poly[i] = out_poly[i - 1];
            --i;
          }
          out_poly[i] = pidx;
          ++w;
        }
        ++found;
      }
      node = nd.escape;
    }
    if (!WRITE) count[pi] = found;
  }
  if (STATS) flush_stats(stats, n_nodes, n_prims);
}

// ------------------------------------------------------------ query order --

__device__ inline uint32_t spread10q(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void query_morton_kernel(const double* q, int stride, uint32_t n, float lx, float ly,
                                    float lz, float sx, float sy, float sz, uint64_t* keys,
                                    uint32_t* vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float c[3] = {(float)q[(size_t)stride * i], (float)q[(size_t)stride * i + 1],
                      stride >= 3 ? (float)q[(size_t)stride * i + 2] : 0.0f};
  const float l[3] = {lx, ly, lz}, sc[3] = {sx, sy, sz};
  uint32_t code = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float f = (c[a] - l[a]) * sc[a];
    f = fminf(fmaxf(f, 0.0f), 1023.0f);
    if (!(f == f)) f = 0.0f;
    code |= spread10q((uint32_t)f) << (2 - a);
  }
  keys[i] = code;
  vals[i] = i;
}

}  // namespace

void launch_collect(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq,
                    double radius, uint32_t k, bool count_all, double* list_d, uint32_t* list_id,
                    uint32_t* count, uint32_t* total, TraversalStats* stats, int block,
                    hipStream_t stream) {
  if (!nq) return;
  require(k >= 1 && k <= kMaxListK, "collect k must be in [1, 64]");
  if (stats)
    dispatch_collect<true>(s, q, order, nq, radius, k, count_all, list_d, list_id, count, total,
                           stats, block, stream);
  else
    dispatch_collect<false>(s, q, order, nq, radius, k, count_all, list_d, list_id, count, total,
                            stats, block, stream);
}

void launch_collect_all(const SceneView& s, const double* q, const uint32_t* queries, uint32_t n,
                        double radius, const uint64_t* offset, double* out_d, uint32_t* out_id,
                        hipStream_t stream) {
  if (!n) return;
  collect_all_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(s, q, queries, n, radius, offset,
                                                              out_d, out_id);
  RTDL_LAUNCH_CHECK();
}

void launch_rays(const SceneView& s, const double* rays, const uint32_t* order, uint32_t nr,
                 int mode, uint32_t* out_u, double* out_t, TraversalStats* stats, int block,
                 hipStream_t stream) {
  if (!nr) return;
  if (mode == 0) launch_rays_mode<0>(s, rays, order, nr, out_u, out_t, stats, block, stream);
  else if (mode == 1) launch_rays_mode<1>(s, rays, order, nr, out_u, out_t, stats, block, stream);
  else launch_rays_mode<2>(s, rays, order, nr, out_u, out_t, stats, block, stream);
}

void launch_pip(const SceneView& s, const double* pts, const uint32_t* order, uint32_t np,
                bool positive_only, double eps_sq, uint32_t* count, const uint64_t* offset,
                uint32_t* out_poly, TraversalStats* stats, int block, hipStream_t stream) {
  if (!np) return;
  const unsigned grid = blocks_for(np, block);
  if (out_poly) {
    pip_kernel<true, false><<<grid, block, 0, stream>>>(s, pts, order, np, positive_only ? 1 : 0,
                                                         eps_sq, count, offset, out_poly, nullptr);
  } else if (stats) {
    pip_kernel<false, true><<<grid, block, 0, stream>>>(s, pts, order, np, positive_only ? 1 : 0,
                                                         eps_sq, count, offset, out_poly, stats);
  } else {
    pip_kernel<false, false><<<grid, block, 0, stream>>>(s, pts, order, np, positive_only ? 1 : 0,
                                                          eps_sq, count, offset, out_poly, nullptr);
  }
  RTDL_LAUNCH_CHECK();
}

void query_morton_order(const double* q, int stride, uint32_t n, const float lo[3],
                        const float hi[3], uint32_t* order, hipStream_t stream) {
  if (!n) return;
  float sc[3];
  for (int a = 0; a < 3; ++a) {
    const float e = hi[a] - lo[a];
    sc[a] = e > 0.0f ? 1024.0f / e : 0.0f;
  }
  DeviceArray<uint64_t> keys(n);
  query_morton_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(q, stride, n, lo[0], lo[1], lo[2],
                                                               sc[0], sc[1], sc[2], keys.get(),
                                                               order);
  RTDL_LAUNCH_CHECK();
  radix_sort_pairs(keys.get(), order, n, 30, stream);
}

}  // namespace rtdl
