// Record unpacking and row emission kernels.  Input records follow the
// reference's ctypes layouts (some packed, so doubles are read as two 32-bit
// words); output rows follow the reference's row structs.
#include "common.hpp"
#include "io.hpp"

namespace rtdl {
namespace {

__device__ inline double f64_at(const uint8_t* base, size_t byte_offset) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(base + byte_offset);
  return __hiloint2double((int)w[1], (int)w[0]);
}
__device__ inline uint32_t u32_at(const uint8_t* base, size_t byte_offset) {
  return *reinterpret_cast<const uint32_t*>(base + byte_offset);
}

__global__ void unpack_points_kernel(const uint8_t* raw, size_t n, int dim, double* coords,
                                     uint32_t* ids) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t rec = dim == 3 ? 32 : 24;
  const size_t b = i * rec;
  ids[i] = u32_at(raw, b);
  coords[3 * i] = f64_at(raw, b + 8);
  coords[3 * i + 1] = f64_at(raw, b + 16);
  coords[3 * i + 2] = dim == 3 ? f64_at(raw, b + 24) : 0.0;
}

__global__ void unpack_points_xy_kernel(const uint8_t* raw, size_t n, double* xy, uint32_t* ids) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = i * 24;
  ids[i] = u32_at(raw, b);
  xy[2 * i] = f64_at(raw, b + 8);
  xy[2 * i + 1] = f64_at(raw, b + 16);
}

__global__ void unpack_triangles_kernel(const uint8_t* raw, size_t n, double* coords,
                                        uint32_t* ids) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = i * 76;
  ids[i] = u32_at(raw, b);
#pragma unroll
  for (int a = 0; a < 9; ++a) coords[9 * i + a] = f64_at(raw, b + 4 + 8 * a);
}

__global__ void unpack_rays_kernel(const uint8_t* raw, size_t n, double* rays, uint32_t* ids) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = i * 60;
  ids[i] = u32_at(raw, b);
#pragma unroll
  for (int a = 0; a < 7; ++a) rays[7 * i + a] = f64_at(raw, b + 4 + 8 * a);
}

// RtdlTriangle (56 B, natural alignment: id, pad, x0 y0 x1 y1 x2 y2) -> 9
// doubles with z = 0.
__global__ void unpack_triangles_2d_kernel(const uint8_t* raw, size_t n, double* coords,
                                           uint32_t* ids) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = i * 56;
  ids[i] = u32_at(raw, b);
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    coords[9 * i + 3 * v] = f64_at(raw, b + 8 + 16 * v);
    coords[9 * i + 3 * v + 1] = f64_at(raw, b + 16 + 16 * v);
    coords[9 * i + 3 * v + 2] = 0.0;
  }
}

// RtdlRay2D (44 B, packed: id, ox oy dx dy tmax) -> ox oy 0 dx dy 0 tmax.
__global__ void unpack_rays_2d_kernel(const uint8_t* raw, size_t n, double* rays, uint32_t* ids) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = i * 44;
  ids[i] = u32_at(raw, b);
  double* r = rays + 7 * i;
  r[0] = f64_at(raw, b + 4);
  r[1] = f64_at(raw, b + 12);
  r[2] = 0.0;
  r[3] = f64_at(raw, b + 20);
  r[4] = f64_at(raw, b + 28);
  r[5] = 0.0;
  r[6] = f64_at(raw, b + 36);
}

__global__ void unpack_polygons_kernel(const uint8_t* raw, size_t n, uint4* refs) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = i * 12;
  refs[i] = make_uint4(u32_at(raw, b + 4), u32_at(raw, b + 8), u32_at(raw, b), 0u);
}

__global__ void gather_counts_kernel(const uint32_t* count, const uint32_t* order, uint32_t n,
                                     uint32_t cap, uint32_t* out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t c = count[order ? order[j] : j];
  if (cap && c > cap) c = cap;
  out[j] = c;
}

__device__ inline uint32_t owner_of(const uint64_t* offsets, uint32_t n, uint64_t i) {
  // largest j with offsets[j] <= i (offsets has n + 1 entries, offsets[n] > i)
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

struct FrnRow {
  uint32_t query_id;
  uint32_t neighbor_id;
  double distance;
};
struct KnnRow {
  uint32_t query_id;
  uint32_t neighbor_id;
  double distance;
  uint32_t neighbor_rank;
  uint32_t pad;
};
struct PairRow {
  uint32_t id;
  uint32_t value;
};
struct ClosestRow {
  uint32_t ray_id;
  uint32_t triangle_id;
  double t;
};
struct PipRow {
  uint32_t point_id;
  uint32_t polygon_id;
  uint32_t contains;
};

__global__ void emit_neighbor_kernel(uint64_t total, const uint64_t* offsets, uint32_t n,
                                     const uint32_t* order, const uint64_t* src_start,
                                     uint32_t stride, const double* cand_d,
                                     const uint32_t* cand_id, const uint32_t* qids, int rank_rows,
                                     void* rows) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  // skip empty positions: owner_of returns the last position starting at <= i
  const uint32_t j = owner_of(offsets, n, i);
  const uint32_t qi = order ? order[j] : j;
  const uint64_t r = i - offsets[j];
  const uint64_t src = (src_start ? src_start[qi] : (uint64_t)qi * stride) + r;
  if (rank_rows) {
    KnnRow row;
    row.query_id = qids[qi];
    row.neighbor_id = cand_id[src];
    row.distance = cand_d[src];
    row.neighbor_rank = (uint32_t)(r + 1);
    row.pad = 0;
    static_cast<KnnRow*>(rows)[i] = row;
  } else {
    FrnRow row;
    row.query_id = qids[qi];
    row.neighbor_id = cand_id[src];
    row.distance = cand_d[src];
    static_cast<FrnRow*>(rows)[i] = row;
  }
}

__global__ void emit_pair_kernel(uint32_t n, const uint32_t* ids, const uint32_t* values,
                                 PairRow* rows) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rows[i] = PairRow{ids[i], values[i]};
}

__global__ void hit_flags_kernel(uint32_t n, const double* t, uint32_t* flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = (t[i] == t[i]) ? 1u : 0u;
}

__global__ void emit_closest_kernel(uint32_t n, const uint64_t* offsets, const uint32_t* ray_ids,
                                    const uint32_t* tri_ids, const double* t, ClosestRow* rows) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (offsets[i + 1] == offsets[i]) return;
  rows[offsets[i]] = ClosestRow{ray_ids[i], tri_ids[i], t[i]};
}

__global__ void emit_pip_positive_kernel(uint64_t total, const uint64_t* offsets, uint32_t n,
                                         const uint32_t* pids, const uint32_t* poly_idx,
                                         const uint4* refs, PipRow* rows) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint32_t p = owner_of(offsets, n, i);
  rows[i] = PipRow{pids[p], refs[poly_idx[i]].z, 1u};
}

__global__ void emit_pip_full_init_kernel(uint32_t n_points, uint32_t n_polys,
                                          const uint32_t* pids, const uint4* refs, PipRow* rows) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)n_points * n_polys) return;
  const uint32_t p = (uint32_t)(i / n_polys), g = (uint32_t)(i % n_polys);
  rows[i] = PipRow{pids[p], refs[g].z, 0u};
}

__global__ void emit_pip_full_set_kernel(uint64_t total, const uint64_t* offsets, uint32_t n,
                                         uint32_t n_polys, const uint32_t* poly_idx,
                                         PipRow* rows) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint32_t p = owner_of(offsets, n, i);
  rows[(uint64_t)p * n_polys + poly_idx[i]].contains = 1u;
}

__global__ void keys_u32_kernel(const uint32_t* src, const uint32_t* perm, size_t n,
                                uint64_t* keys) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = src[perm ? perm[i] : i];
}

__global__ void keys_f64_kernel(const double* src, const uint32_t* perm, size_t n,
                                uint64_t* keys) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // non-negative doubles order like their bit patterns
  keys[i] = (uint64_t)__double_as_longlong(src[perm ? perm[i] : i]);
}

__global__ void segment_ids_kernel(const uint64_t* offsets, uint32_t n_seg, size_t n,
                                   uint32_t* seg) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  seg[i] = owner_of(offsets, n_seg, i);
}

__global__ void permute_u32_kernel(const uint32_t* src, const uint32_t* perm, size_t n,
                                   uint32_t* dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ void permute_f64_kernel(const double* src, const uint32_t* perm, size_t n,
                                   double* dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ void widen_points_kernel(const float* in, size_t n, const uint64_t* offs,
                                    uint32_t n_inst, double* coords, uint32_t* ids) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  coords[3 * i] = in[3 * i];
  coords[3 * i + 1] = in[3 * i + 1];
  coords[3 * i + 2] = in[3 * i + 2];
  ids[i] = (uint32_t)(i - offs[owner_of(offs, n_inst, i)]);
}

__global__ void widen_polygons_kernel(const uint32_t* voff, const uint32_t* vcnt, size_t n,
                                      const uint64_t* offs, uint32_t n_inst, uint4* refs) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  refs[i] = make_uint4(voff[i], vcnt[i], (uint32_t)(i - offs[owner_of(offs, n_inst, i)]), 0u);
}

__global__ void instance_ids_kernel(size_t n, const uint64_t* offs, uint32_t n_inst, uint32_t* ids) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ids[i] = (uint32_t)(i - offs[owner_of(offs, n_inst, i)]);
}

__global__ void widen_f32_kernel(const float* in, size_t n, double* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

__global__ void fp64_selftest_kernel(const double* a, const double* b, size_t n, double* s,
                                     double* d) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  s[i] = sqrt(a[i]);
  d[i] = a[i] / b[i];
}

}  // namespace

#define RTDL_GRID(n) blocks_for((n), 256), 256, 0, stream

void unpack_points(const void* raw, size_t n, int dim, double* coords, uint32_t* ids,
                   hipStream_t stream) {
  if (!n) return;
  unpack_points_kernel<<<RTDL_GRID(n)>>>(static_cast<const uint8_t*>(raw), n, dim, coords, ids);
  RTDL_LAUNCH_CHECK();
}
void unpack_points_xy(const void* raw, size_t n, double* xy, uint32_t* ids, hipStream_t stream) {
  if (!n) return;
  unpack_points_xy_kernel<<<RTDL_GRID(n)>>>(static_cast<const uint8_t*>(raw), n, xy, ids);
  RTDL_LAUNCH_CHECK();
}
void unpack_triangles(const void* raw, size_t n, double* coords, uint32_t* ids,
                      hipStream_t stream) {
  if (!n) return;
  unpack_triangles_kernel<<<RTDL_GRID(n)>>>(static_cast<const uint8_t*>(raw), n, coords, ids);
  RTDL_LAUNCH_CHECK();
}
void unpack_rays(const void* raw, size_t n, double* rays, uint32_t* ids, hipStream_t stream) {
  if (!n) return;
  unpack_rays_kernel<<<RTDL_GRID(n)>>>(static_cast<const uint8_t*>(raw), n, rays, ids);
  RTDL_LAUNCH_CHECK();
}
void unpack_triangles_2d(const void* raw, size_t n, double* coords, uint32_t* ids,
                         hipStream_t stream) {
  if (!n) return;
  unpack_triangles_2d_kernel<<<RTDL_GRID(n)>>>(static_cast<const uint8_t*>(raw), n, coords, ids);
  RTDL_LAUNCH_CHECK();
}
void unpack_rays_2d(const void* raw, size_t n, double* rays, uint32_t* ids, hipStream_t stream) {
  if (!n) return;
  unpack_rays_2d_kernel<<<RTDL_GRID(n)>>>(static_cast<const uint8_t*>(raw), n, rays, ids);
  RTDL_LAUNCH_CHECK();
}
void unpack_polygons(const void* raw, size_t n, uint4* refs, hipStream_t stream) {
  if (!n) return;
  unpack_polygons_kernel<<<RTDL_GRID(n)>>>(static_cast<const uint8_t*>(raw), n, refs);
  RTDL_LAUNCH_CHECK();
}
void gather_counts(const uint32_t* count, const uint32_t* order, uint32_t n, uint32_t cap,
                   uint32_t* counts_pos, hipStream_t stream) {
  if (!n) return;
  gather_counts_kernel<<<RTDL_GRID(n)>>>(count, order, n, cap, counts_pos);
  RTDL_LAUNCH_CHECK();
}
void emit_neighbor_rows(uint64_t total, const uint64_t* offsets, uint32_t n,
                        const uint32_t* order, const uint64_t* src_start, uint32_t stride,
                        const double* cand_d, const uint32_t* cand_id, const uint32_t* qids,
                        bool rank_rows, void* rows, hipStream_t stream) {
  if (!total) return;
  emit_neighbor_kernel<<<RTDL_GRID(total)>>>(total, offsets, n, order, src_start, stride, cand_d,
                                             cand_id, qids, rank_rows ? 1 : 0, rows);
  RTDL_LAUNCH_CHECK();
}
void emit_pair_rows(uint32_t n, const uint32_t* ids, const uint32_t* values, void* rows,
                    hipStream_t stream) {
  if (!n) return;
  emit_pair_kernel<<<RTDL_GRID(n)>>>(n, ids, values, static_cast<PairRow*>(rows));
  RTDL_LAUNCH_CHECK();
}
void hit_flags(uint32_t n, const double* t, uint32_t* flags, hipStream_t stream) {
  if (!n) return;
  hit_flags_kernel<<<RTDL_GRID(n)>>>(n, t, flags);
  RTDL_LAUNCH_CHECK();
}
void emit_closest_rows(uint32_t n, const uint64_t* offsets, const uint32_t* ray_ids,
                       const uint32_t* tri_ids, const double* t, void* rows, hipStream_t stream) {
  if (!n) return;
  emit_closest_kernel<<<RTDL_GRID(n)>>>(n, offsets, ray_ids, tri_ids, t,
                                        static_cast<ClosestRow*>(rows));
  RTDL_LAUNCH_CHECK();
}
void emit_pip_positive(uint64_t total, const uint64_t* offsets, uint32_t n, const uint32_t* pids,
                       const uint32_t* poly_idx, const uint4* refs, void* rows,
                       hipStream_t stream) {
  if (!total) return;
  emit_pip_positive_kernel<<<RTDL_GRID(total)>>>(total, offsets, n, pids, poly_idx, refs,
                                                 static_cast<PipRow*>(rows));
  RTDL_LAUNCH_CHECK();
}
void emit_pip_full(uint32_t n_points, uint32_t n_polys, const uint64_t* offsets,
                   const uint32_t* pids, const uint32_t* poly_idx, const uint4* refs, void* rows,
                   hipStream_t stream) {
  const uint64_t cells = (uint64_t)n_points * n_polys;
  if (!cells) return;
  emit_pip_full_init_kernel<<<RTDL_GRID(cells)>>>(n_points, n_polys, pids, refs,
                                                  static_cast<PipRow*>(rows));
  RTDL_LAUNCH_CHECK();
  uint64_t total = 0;
  RTDL_HIP(hipMemcpyAsync(&total, offsets + n_points, sizeof(total), hipMemcpyDeviceToHost, stream));
  RTDL_HIP(hipStreamSynchronize(stream));
  if (total) {
    emit_pip_full_set_kernel<<<RTDL_GRID(total)>>>(total, offsets, n_points, n_polys, poly_idx,
                                                   static_cast<PipRow*>(rows));
    RTDL_LAUNCH_CHECK();
  }
}
void keys_from_u32(const uint32_t* src, const uint32_t* perm, size_t n, uint64_t* keys,
                   hipStream_t stream) {
  if (!n) return;
  keys_u32_kernel<<<RTDL_GRID(n)>>>(src, perm, n, keys);
  RTDL_LAUNCH_CHECK();
}
void keys_from_f64(const double* src, const uint32_t* perm, size_t n, uint64_t* keys,
                   hipStream_t stream) {
  if (!n) return;
  keys_f64_kernel<<<RTDL_GRID(n)>>>(src, perm, n, keys);
  RTDL_LAUNCH_CHECK();
}
void segment_ids(const uint64_t* offsets, uint32_t n_seg, size_t n, uint32_t* seg,
                 hipStream_t stream) {
  if (!n) return;
  segment_ids_kernel<<<RTDL_GRID(n)>>>(offsets, n_seg, n, seg);
  RTDL_LAUNCH_CHECK();
}
void permute_u32(const uint32_t* src, const uint32_t* perm, size_t n, uint32_t* dst,
                 hipStream_t stream) {
  if (!n) return;
  permute_u32_kernel<<<RTDL_GRID(n)>>>(src, perm, n, dst);
  RTDL_LAUNCH_CHECK();
}
void permute_f64(const double* src, const uint32_t* perm, size_t n, double* dst,
                 hipStream_t stream) {
  if (!n) return;
  permute_f64_kernel<<<RTDL_GRID(n)>>>(src, perm, n, dst);
  RTDL_LAUNCH_CHECK();
}
void widen_points(const float* in, size_t n, const uint64_t* offs, uint32_t n_inst, double* coords,
                  uint32_t* ids, hipStream_t stream) {
  if (!n) return;
  widen_points_kernel<<<RTDL_GRID(n)>>>(in, n, offs, n_inst, coords, ids);
  RTDL_LAUNCH_CHECK();
}
void widen_polygons(const uint32_t* voff, const uint32_t* vcnt, size_t n, const uint64_t* offs,
                    uint32_t n_inst, uint4* refs, hipStream_t stream) {
  if (!n) return;
  widen_polygons_kernel<<<RTDL_GRID(n)>>>(voff, vcnt, n, offs, n_inst, refs);
  RTDL_LAUNCH_CHECK();
}
void widen_f32(const float* in, size_t n, double* out, hipStream_t stream) {
  if (!n) return;
  widen_f32_kernel<<<RTDL_GRID(n)>>>(in, n, out);
  RTDL_LAUNCH_CHECK();
}
void instance_ids(size_t n, const uint64_t* offs, uint32_t n_inst, uint32_t* ids, hipStream_t stream) {
  if (!n) return;
  instance_ids_kernel<<<RTDL_GRID(n)>>>(n, offs, n_inst, ids);
  RTDL_LAUNCH_CHECK();
}
void fp64_selftest(const double* a, const double* b, size_t n, double* out_sqrt, double* out_div,
                   hipStream_t stream) {
  if (!n) return;
  fp64_selftest_kernel<<<RTDL_GRID(n)>>>(a, b, n, out_sqrt, out_div);
  RTDL_LAUNCH_CHECK();
}

}  // namespace rtdl
