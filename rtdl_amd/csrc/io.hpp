// Unpacking of the C-ABI input records into canonical device arrays, and
// emission of C-ABI row records from device results.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtdl {

// ---- unpack (raw = device copy of the caller's record array) ----
// points: RtdlPoint (24 B) when dim == 2, RtdlPoint3D (32 B) when dim == 3;
// coords receives 3 doubles per point (z = +0.0 in 2-D).
void unpack_points(const void* raw, size_t n, int dim, double* coords, uint32_t* ids,
                   hipStream_t stream);
// RtdlTriangle3D (76 B, packed): coords 9 doubles per triangle.
void unpack_triangles(const void* raw, size_t n, double* coords, uint32_t* ids, hipStream_t stream);
// RtdlRay3D (60 B, packed): 7 doubles per ray.
void unpack_rays(const void* raw, size_t n, double* rays, uint32_t* ids, hipStream_t stream);
// RtdlTriangle (56 B): 9 doubles per triangle with z = 0.
void unpack_triangles_2d(const void* raw, size_t n, double* coords, uint32_t* ids, hipStream_t stream);
// RtdlRay2D (44 B, packed): ox oy 0 dx dy 0 tmax per ray.
void unpack_rays_2d(const void* raw, size_t n, double* rays, uint32_t* ids, hipStream_t stream);
// RtdlPolygonRef (12 B): (vertex_offset, vertex_count, id, 0).
void unpack_polygons(const void* raw, size_t n, uint4* refs, hipStream_t stream);
// RtdlPoint (24 B) -> 2 doubles per point (PIP probes).
void unpack_points_xy(const void* raw, size_t n, double* xy, uint32_t* ids, hipStream_t stream);

// ---- emit ----
// emit position j holds query order[j] (identity when order is null).
// counts_pos[j] = min(count[order[j]], cap) (cap 0 = no cap).
void gather_counts(const uint32_t* count, const uint32_t* order, uint32_t n, uint32_t cap,
                   uint32_t* counts_pos, hipStream_t stream);
// Rows of neighbour queries: row i belongs to the emit position j with
// offsets[j] <= i < offsets[j + 1]; its candidate is cand[src_start[q] + r]
// (src_start null: q * stride).  rank_rows selects the 24-byte KNN row.
void emit_neighbor_rows(uint64_t total, const uint64_t* offsets, uint32_t n,
                        const uint32_t* order, const uint64_t* src_start, uint32_t stride,
                        const double* cand_d, const uint32_t* cand_id, const uint32_t* qids,
                        bool rank_rows, void* rows, hipStream_t stream);
// (id, value) rows in input order.
void emit_pair_rows(uint32_t n, const uint32_t* ids, const uint32_t* values, void* rows,
                    hipStream_t stream);
// closest-hit: flag[i] = 1 when t[i] is not NaN.
void hit_flags(uint32_t n, const double* t, uint32_t* flags, hipStream_t stream);
void emit_closest_rows(uint32_t n, const uint64_t* offsets, const uint32_t* ray_ids,
                       const uint32_t* tri_ids, const double* t, void* rows, hipStream_t stream);
// PIP positive rows: point i's containing polygons (input indices) are
// poly_idx[offsets[i] .. offsets[i+1]).
void emit_pip_positive(uint64_t total, const uint64_t* offsets, uint32_t n, const uint32_t* pids,
                       const uint32_t* poly_idx, const uint4* refs_input_order, void* rows,
                       hipStream_t stream);
void emit_pip_full(uint32_t n_points, uint32_t n_polys, const uint64_t* offsets,
                   const uint32_t* pids, const uint32_t* poly_idx, const uint4* refs_input_order,
                   void* rows, hipStream_t stream);

// Segment-sort support for the count-then-collect path: key of a candidate
// for the stable multi-pass sort.
void keys_from_u32(const uint32_t* src, const uint32_t* perm, size_t n, uint64_t* keys,
                   hipStream_t stream);
void keys_from_f64(const double* src, const uint32_t* perm, size_t n, uint64_t* keys,
                   hipStream_t stream);
// seg[i] = segment of candidate i, from segment offsets (n_seg + 1 entries).
void segment_ids(const uint64_t* offsets, uint32_t n_seg, size_t n, uint32_t* seg,
                 hipStream_t stream);
void permute_u32(const uint32_t* src, const uint32_t* perm, size_t n, uint32_t* dst,
                 hipStream_t stream);
void permute_f64(const double* src, const uint32_t* perm, size_t n, double* dst,
                 hipStream_t stream);

// FP32 device inputs widened to the canonical FP64 form; ids are
// instance-local indices (offs: n_inst + 1 device entries).
void widen_points(const float* in, size_t n, const uint64_t* offs, uint32_t n_inst, double* coords,
                  uint32_t* ids, hipStream_t stream);
void widen_polygons(const uint32_t* voff, const uint32_t* vcnt, size_t n, const uint64_t* offs,
                    uint32_t n_inst, uint4* refs, hipStream_t stream);
void widen_f32(const float* in, size_t n, double* out, hipStream_t stream);
// ids[i] = i - offs[instance of i]
void instance_ids(size_t n, const uint64_t* offs, uint32_t n_inst, uint32_t* ids, hipStream_t stream);

// FP64 self-test: out_sqrt[i] = sqrt(a[i]), out_div[i] = a[i] / b[i].
void fp64_selftest(const double* a, const double* b, size_t n, double* out_sqrt, double* out_div,
                   hipStream_t stream);

}  // namespace rtdl
