// Launch wrappers of the per-call traversal + refine kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "scene.hpp"

namespace rtdl {

// Largest per-query selection kept in registers; larger fixed-radius k_max
// values use the count-then-collect path.
constexpr uint32_t kMaxListK = 64;

// Power-of-two register list capacity used for a selection of k.
uint32_t list_capacity(uint32_t k);

// `stats`, when not null, points at two device counters (node records
// fetched, primitive records tested) that the kernel adds to.

// ---- point queries --------------------------------------------------------
// q: FP64 query coordinates (3 per query, z = 0 in 2-D); order: traversal
// order (position -> query index) or null.  For every query the k smallest
// (distance, id) candidates with squared FP64 distance <= radius * radius are
// written, ascending, to list_d/list_id[query * k + r], with count[query].
// The packed kernel appends queries whose FP32 selection keys tie at the
// selection boundary to amb_list (amb_count is the device counter); those
// must be recomputed with launch_collect_exact.
void launch_collect_packed(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq,
                           double radius, uint32_t k, double* list_d, uint32_t* list_id,
                           uint32_t* count, uint32_t* amb_list, uint32_t* amb_count,
                           unsigned long long* stats, int block, hipStream_t stream);
// Exact FP64 selection; with count_all (k must be 64) the radius never
// shrinks and total[query] receives the number of candidates in the radius.
// Seeded: a query whose row already holds k FP64 distances (from the packed
// kernel) searches only the ball of the largest of them.
void launch_collect_exact(const SceneView& s, const double* q, const uint32_t* order, uint32_t nq,
                          double radius, uint32_t k, bool count_all, double* list_d,
                          uint32_t* list_id, uint32_t* count, uint32_t* total, bool seeded,
                          hipStream_t stream);

// Writes every candidate inside the radius for the queries listed in
// `queries` (n entries) into [offset[i], offset[i + 1]) of out_d / out_id,
// unsorted.
void launch_collect_all(const SceneView& s, const double* q, const uint32_t* queries, uint32_t n,
                        double radius, const uint64_t* offset, double* out_d, uint32_t* out_id,
                        hipStream_t stream);

// ---- rays ---------------------------------------------------------------
// rays: 7 FP64 per ray (ox, oy, oz, dx, dy, dz, tmax).  mode 0: any-hit flag,
// 1: hit count, 2: closest hit (triangle id in out_u, t in out_t, NaN t when
// the ray hits nothing).  mode + 4: the same for 2-D rays (oz = dz = 0 and
// triangles with z = 0), whose hits all have t = 0.
void launch_rays(const SceneView& s, const double* rays, const uint32_t* order, uint32_t nr,
                 int mode, uint32_t* out_u, double* out_t, unsigned long long* stats, int block,
                 hipStream_t stream);

// ---- point in polygon -----------------------------------------------------
// pts: 2 FP64 per point.  Pass 1 (out_poly == null) writes count[point];
// pass 2 writes the input indices of containing polygons, ascending, into
// [offset[point], offset[point + 1]).
void launch_pip(const SceneView& s, const double* pts, const uint32_t* order, uint32_t np,
                bool positive_only, double eps_sq, uint32_t* count, const uint64_t* offset,
                uint32_t* out_poly, unsigned long long* stats, int block, hipStream_t stream);

// ---- helpers ----------------------------------------------------------------
// Morton order of query positions inside the box [lo, hi] (`stride` FP64 per
// query, position in the first 2 or 3); returns the permutation in `order`.
void query_morton_order(const double* q, int stride, uint32_t n, const float lo[3],
                        const float hi[3], uint32_t* order, hipStream_t stream);

}  // namespace rtdl
