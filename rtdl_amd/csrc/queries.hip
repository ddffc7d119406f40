// Per-call traversal kernels: one lane per query, queries in Morton order (via
// `order`) so that the 64 lanes of a wavefront walk nearly the same part of
// the tree.  The traversal bodies live in traverse.hpp.
#include <cmath>

#include "common.hpp"
#include "primitives.hpp"
#include "queries.hpp"
#include "traverse.hpp"

namespace rtdl {
namespace {

template <int K, bool STATS>
__global__ __launch_bounds__(256) void collect_packed_kernel(
    SceneView s, const double* __restrict__ q, const uint32_t* __restrict__ order, uint32_t nq,
    double radius, uint32_t k, double* __restrict__ list_d, uint32_t* __restrict__ list_id,
    uint32_t* __restrict__ count, uint32_t* __restrict__ amb_list, uint32_t* __restrict__ amb_count,
    unsigned long long* stats) {
  __shared__ uint32_t stack[kStackDepth * 256];  // walk_nearest's per-lane stacks
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStats st;
  if (t < nq) {
    const uint32_t qi = order ? order[t] : t;
    const double qx = q[3 * (size_t)qi], qy = q[3 * (size_t)qi + 1], qz = q[3 * (size_t)qi + 2];
    PackedTopK<K> list;
    const bool ambiguous = collect_packed<K, STATS>(s, s.root, s.end, qx, qy, qz, radius, k, list, st,
                                                    stack + threadIdx.x, blockDim.x);
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (i >= K - (int)k && list.e[i] != ~0ull) {
        const uint32_t slot = (uint32_t)list.e[i];
        const float4 p = load_prim(s.prims, slot);
        const size_t at = (size_t)qi * k + (uint32_t)(i - (K - (int)k));
        list_d[at] = sqrt(point_d2(s, slot, p, qx, qy, qz));
        list_id[at] = __float_as_uint(p.w);
        ++c;
      }
    }
    count[qi] = c;
    if (ambiguous) amb_list[atomicAdd(amb_count, 1u)] = qi;
  }
  if (STATS) flush_lane_stats(stats, st);
}

template <int K, bool COUNT_ALL>
__global__ __launch_bounds__(256) void collect_exact_kernel(
    SceneView s, const double* __restrict__ q, const uint32_t* __restrict__ order, uint32_t nq,
    double radius, uint32_t k, double* __restrict__ list_d, uint32_t* __restrict__ list_id,
    uint32_t* __restrict__ count, uint32_t* __restrict__ total, bool seeded) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  LaneStats st;
  const uint32_t qi = order ? order[t] : t;
  const double qx = q[3 * (size_t)qi], qy = q[3 * (size_t)qi + 1], qz = q[3 * (size_t)qi + 2];
  // Seeded: the row holds the FP64 distances of k distinct points found by
  // the packed walk, so the k nearest (and every point tied with the k-th)
  // lie within the largest of them; the factor 1 + 2^-40 keeps each of those
  // points inside the squared radius after rounding.
  double rq = radius;
  if (seeded && count[qi] == k) {
    double dmax = 0.0;
    bool finite = true;
    for (uint32_t i = 0; i < k; ++i) {
      const double d = list_d[(size_t)qi * k + i];
      finite = finite && d < __builtin_inf();
      dmax = d > dmax ? d : dmax;
    }
    if (finite) rq = fmin(radius, dmax * 1.0000000000009094947);
  }
  ExactTopK<K> list;
  uint32_t inside = 0;
  collect_exact<K, COUNT_ALL, false>(s, s.root, s.end, qx, qy, qz, rq, k, list, inside, st);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (i >= K - (int)k && list.d[i] < __builtin_inf()) {
      const size_t at = (size_t)qi * k + (uint32_t)(i - (K - (int)k));
      list_d[at] = list.d[i];
      list_id[at] = list.id[i];
      ++c;
    }
  }
  count[qi] = c;
  if (COUNT_ALL) total[qi] = inside;
}

__global__ __launch_bounds__(256) void collect_all_kernel(SceneView s, const double* __restrict__ q,
                                                          const uint32_t* __restrict__ queries,
                                                          uint32_t n, double radius,
                                                          const uint64_t* __restrict__ offset,
                                                          double* __restrict__ out_d,
                                                          uint32_t* __restrict__ out_id) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  LaneStats st;
  const uint32_t qi = queries[t];
  const double qx = q[3 * (size_t)qi], qy = q[3 * (size_t)qi + 1], qz = q[3 * (size_t)qi + 2];
  const float qf[3] = {(float)qx, (float)qy, (float)qz};
  const float m = coord_bound(scene_m(s), qx, qy, qz);
  const double radius_sq = radius * radius;
  float thr2 = frn_threshold2(radius, m);
  uint64_t w = offset[t];
  const uint64_t w_end = offset[t + 1];
  walk_points<false>(s, s.root, s.end, qf, thr2, st, [&](uint32_t slot, const float4& p, float) {
    const double d2 = point_d2(s, slot, p, qx, qy, qz);
    if (d2 > radius_sq) return;
    if (w < w_end) {
      out_d[w] = sqrt(d2);
      out_id[w] = __float_as_uint(p.w);
    }
    ++w;
  });
}

template <int MODE, bool STATS>
__global__ __launch_bounds__(256) void ray_kernel(SceneView s, const double* __restrict__ rays,
                                                  const uint32_t* __restrict__ order, uint32_t nr,
                                                  uint32_t* __restrict__ out_u,
                                                  double* __restrict__ out_t,
                                                  unsigned long long* stats) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStats st;
  if (t < nr) {
    const uint32_t ri = order ? order[t] : t;
    const double* r = rays + 7 * (size_t)ri;
    const RayResult res =
        trace_ray<MODE, STATS>(s, s.root, s.end, r[0], r[1], r[2], r[3], r[4], r[5], r[6], st);
    if ((MODE & 3) == 0) out_u[ri] = res.hits ? 1u : 0u;
    if ((MODE & 3) == 1) out_u[ri] = res.hits;
    if ((MODE & 3) == 2) {
      out_u[ri] = res.best_id;
      out_t[ri] = res.best_t;
    }
  }
  if (STATS) flush_lane_stats(stats, st);
}

template <bool WRITE, bool STATS>
__global__ __launch_bounds__(256) void pip_kernel(SceneView s, const double* __restrict__ pts,
                                                  const uint32_t* __restrict__ order, uint32_t np,
                                                  int positive_only, double eps_sq,
                                                  uint32_t* __restrict__ count,
                                                  const uint64_t* __restrict__ offset,
                                                  uint32_t* __restrict__ out_poly,
                                                  unsigned long long* stats) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStats st;
  if (t < np) {
    const uint32_t pi = order ? order[t] : t;
    const double x = pts[2 * (size_t)pi], y = pts[2 * (size_t)pi + 1];
    uint32_t found = 0;
    const uint64_t base = WRITE ? offset[pi] : 0;
    const uint64_t limit = WRITE ? offset[pi + 1] : 0;
    walk_pip<STATS>(s, s.root, s.end, x, y, positive_only != 0, eps_sq, st, [&](uint32_t slot) {
      if (WRITE && base + found < limit) {
        // insertion keeps the segment ascending by polygon input index
        const uint32_t pidx = s.slot_to_input[slot];
        uint64_t i = base + found;
        while (i > base && out_poly[i - 1] > pidx) {
          out_poly[i] = out_poly[i - 1];
          --i;
        }
        out_poly[i] = pidx;
      }
      ++found;
    });
    if (!WRITE) count[pi] = found;
  }
  if (STATS) flush_lane_stats(stats, st);
}

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

template <int K>
void packed_t(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq, double radius,
              uint32_t k, double* list_d, uint32_t* list_id, uint32_t* count, uint32_t* amb_list,
              uint32_t* amb_count, unsigned long long* stats, int block, hipStream_t stream) {
  if (stats)
    collect_packed_kernel<K, true><<<blocks_for(nq, block), block, 0, stream>>>(
        s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats);
  else
    collect_packed_kernel<K, false><<<blocks_for(nq, block), block, 0, stream>>>(
        s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats);
  RTDL_LAUNCH_CHECK();
}

template <int K, bool COUNT_ALL>
void exact_t(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq, double radius,
             uint32_t k, double* list_d, uint32_t* list_id, uint32_t* count, uint32_t* total,
             bool seeded, hipStream_t stream) {
  collect_exact_kernel<K, COUNT_ALL><<<blocks_for(nq, 64), 64, 0, stream>>>(
      s, q, order, nq, radius, k, list_d, list_id, count, total, seeded);
  RTDL_LAUNCH_CHECK();
}

template <int MODE>
void rays_t(const SceneView& s, const double* rays, const uint32_t* order, uint32_t nr,
            uint32_t* out_u, double* out_t, unsigned long long* stats, int block,
            hipStream_t stream) {
  if (stats)
    ray_kernel<MODE, true><<<blocks_for(nr, block), block, 0, stream>>>(s, rays, order, nr, out_u, out_t, stats);
  else
    ray_kernel<MODE, false><<<blocks_for(nr, block), block, 0, stream>>>(s, rays, order, nr, out_u, out_t, stats);
  RTDL_LAUNCH_CHECK();
}

}  // namespace

uint32_t list_capacity(uint32_t k) {
  uint32_t K = 1;
  while (K < k) K <<= 1;
  return K;
}

void launch_collect_packed(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq,
                           double radius, uint32_t k, double* list_d, uint32_t* list_id,
                           uint32_t* count, uint32_t* amb_list, uint32_t* amb_count,
                           unsigned long long* stats, int block, hipStream_t stream) {
  if (!nq) return;
  require(k >= 1 && k <= kMaxListK, "collect k must be in [1, 64]");
  switch (list_capacity(k)) {
    case 1: packed_t<1>(s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats, block, stream); break;
    case 2: packed_t<2>(s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats, block, stream); break;
    case 4: packed_t<4>(s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats, block, stream); break;
    case 8: packed_t<8>(s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats, block, stream); break;
    case 16: packed_t<16>(s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats, block, stream); break;
    case 32: packed_t<32>(s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats, block, stream); break;
    default: packed_t<64>(s, q, order, nq, radius, k, list_d, list_id, count, amb_list, amb_count, stats, block, stream); break;
  }
}

void launch_collect_exact(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq,
                          double radius, uint32_t k, bool count_all, double* list_d,
                          uint32_t* list_id, uint32_t* count, uint32_t* total, bool seeded,
                          hipStream_t stream) {
  if (!nq) return;
  require(k >= 1 && k <= kMaxListK, "collect k must be in [1, 64]");
  if (count_all) {
    require(k == kMaxListK, "count_all collection needs k == 64");
    exact_t<64, true>(s, q, order, nq, radius, k, list_d, list_id, count, total, seeded, stream);
    return;
  }
  switch (list_capacity(k)) {
    case 1: exact_t<1, false>(s, q, order, nq, radius, k, list_d, list_id, count, total, seeded, stream); break;
    case 2: exact_t<2, false>(s, q, order, nq, radius, k, list_d, list_id, count, total, seeded, stream); break;
    case 4: exact_t<4, false>(s, q, order, nq, radius, k, list_d, list_id, count, total, seeded, stream); break;
    case 8: exact_t<8, false>(s, q, order, nq, radius, k, list_d, list_id, count, total, seeded, stream); break;
    case 16: exact_t<16, false>(s, q, order, nq, radius, k, list_d, list_id, count, total, seeded, stream); break;
    case 32: exact_t<32, false>(s, q, order, nq, radius, k, list_d, list_id, count, total, seeded, stream); break;
    default: exact_t<64, false>(s, q, order, nq, radius, k, list_d, list_id, count, total, seeded, stream); break;
  }
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
                 int mode, uint32_t* out_u, double* out_t, unsigned long long* stats, int block,
                 hipStream_t stream) {
  if (!nr) return;
  switch (mode) {
    case 0: rays_t<0>(s, rays, order, nr, out_u, out_t, stats, block, stream); break;
    case 1: rays_t<1>(s, rays, order, nr, out_u, out_t, stats, block, stream); break;
    case 2: rays_t<2>(s, rays, order, nr, out_u, out_t, stats, block, stream); break;
    case 4: rays_t<4>(s, rays, order, nr, out_u, out_t, stats, block, stream); break;
    case 5: rays_t<5>(s, rays, order, nr, out_u, out_t, stats, block, stream); break;
    case 6: rays_t<6>(s, rays, order, nr, out_u, out_t, stats, block, stream); break;
    default: throw Error("unknown ray mode " + std::to_string(mode));
  }
}

void launch_pip(const SceneView& s, const double* pts, const uint32_t* order, uint32_t np,
                bool positive_only, double eps_sq, uint32_t* count, const uint64_t* offset,
                uint32_t* out_poly, unsigned long long* stats, int block, hipStream_t stream) {
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
