/*
 * rtdl_mi355x.h — C-ABI of the MI355X-native RTDL backend (librtdl_mi355x.so).
 *
 * The plain per-call and prepared entry points replace the reference's
 * HIPRT C-ABI one for one (same record layouts, same argument meaning, same
 * status/error-buffer convention); each declaration cites the reference entry
 * point it replaces.  The device-resident batch and serving entry points at
 * the end have no counterpart in the reference: they extend the backend with
 * the batched and served execution patterns.
 *
 * Conventions (reference: src/native/hiprt/rtdl_hiprt_core.cpp:3323-3335,
 * src/native/hiprt/rtdl_hiprt_api.cpp:288-330):
 *   - return 0 on success; non-zero on failure with a NUL-terminated message
 *     truncated into `error_out` (size `error_size`);
 *   - `*rows_out` / `*row_count_out` are zeroed on entry; row buffers are
 *     allocated by the library and released with rtdl_mi355x_free_rows;
 *   - a null input pointer with a non-zero count is an error;
 *   - an empty probe set yields 0 rows; an empty build set yields all-zero
 *     any-hit / hit-count rows and no neighbour, closest-hit or PIP rows;
 *   - inputs are borrowed for the duration of the call only; prepared handles
 *     are owned by the caller and released with the matching destroy call.
 *   - all entry points are thread-safe; calls on one prepared handle are
 *     serialised by the handle.
 */
#ifndef RTDL_MI355X_H
#define RTDL_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- input records (reference: src/rtdsl/hiprt_runtime.py:108-260) ---- */

typedef struct RtdlPoint {
  uint32_t id;
  double x;
  double y;
} RtdlPoint; /* 24 bytes */

typedef struct RtdlPoint3D {
  uint32_t id;
  double x;
  double y;
  double z;
} RtdlPoint3D; /* 32 bytes */

typedef struct RtdlPolygonRef {
  uint32_t id;
  uint32_t vertex_offset; /* in vertices, into vertices_xy (2 doubles each) */
  uint32_t vertex_count;
} RtdlPolygonRef; /* 12 bytes */

#pragma pack(push, 1)
typedef struct RtdlTriangle3D {
  uint32_t id;
  double x0, y0, z0, x1, y1, z1, x2, y2, z2;
} RtdlTriangle3D; /* 76 bytes, packed */

typedef struct RtdlRay3D {
  uint32_t id;
  double ox, oy, oz, dx, dy, dz, tmax;
} RtdlRay3D; /* 60 bytes, packed */

typedef struct RtdlRay2D {
  uint32_t id;
  double ox, oy, dx, dy, tmax;
} RtdlRay2D; /* 44 bytes, packed (hiprt_runtime.py:137-148) */
#pragma pack(pop)

typedef struct RtdlTriangle {
  uint32_t id;
  double x0, y0, x1, y1, x2, y2;
} RtdlTriangle; /* 56 bytes, natural alignment (hiprt_runtime.py:125-135) */

/* ---- output rows (reference: src/rtdsl/hiprt_runtime.py:172-267,
 *      src/rtdsl/embree_runtime.py:420-479) ---- */

typedef struct RtdlFixedRadiusNeighborRow {
  uint32_t query_id;
  uint32_t neighbor_id;
  double distance;
} RtdlFixedRadiusNeighborRow; /* 16 bytes */

typedef struct RtdlKnnNeighborRow {
  uint32_t query_id;
  uint32_t neighbor_id;
  double distance;
  uint32_t neighbor_rank;
} RtdlKnnNeighborRow; /* 24 bytes */

typedef struct RtdlRayAnyHitRow {
  uint32_t ray_id;
  uint32_t any_hit;
} RtdlRayAnyHitRow;

typedef struct RtdlRayHitCountRow {
  uint32_t ray_id;
  uint32_t hit_count;
} RtdlRayHitCountRow;

typedef struct RtdlRayClosestHitRow {
  uint32_t ray_id;
  uint32_t triangle_id;
  double t;
} RtdlRayClosestHitRow; /* 16 bytes */

typedef struct RtdlPipRow {
  uint32_t point_id;
  uint32_t polygon_id;
  uint32_t contains;
} RtdlPipRow;

/* ---- library / device ---- */

/* replaces rtdl_hiprt_get_version (src/native/hiprt/rtdl_hiprt_api.cpp:1-9) */
int rtdl_mi355x_get_version(int* major, int* minor, int* patch);

/* replaces rtdl_hiprt_context_probe (src/native/hiprt/rtdl_hiprt_api.cpp:1059-1100):
 * device_type = 1 for a GPU; api_version = HIP runtime version. */
int rtdl_mi355x_context_probe(char* device_name, size_t device_name_size, int* device_type,
                              int* api_version, char* error_out, size_t error_size);

/* replaces rtdl_hiprt_free_rows (src/native/hiprt/rtdl_hiprt_api.cpp:11-13).
 * Row sets of 1 MiB or more live in pinned host memory; a released one is
 * kept (up to 4 GiB of released buffers) for a later call of about its size.
 * Only pointers returned as rows by this library may be passed. */
void rtdl_mi355x_free_rows(void* rows);

/* Per-phase wall times of the last call on this thread, in seconds:
 * [0] upload+pack, [1] build, [2] traverse, [3] emit, [4] copy-back, [5] total.
 * (cf. src/rtdsl/optix_runtime.py:4880-4902).  Returns the count written. */
int rtdl_mi355x_get_last_phase_timings(double* seconds_out, size_t capacity);

/* Traversal measurement of this thread's per-call and prepared calls (no
 * counterpart in the reference; for the roofline of the per-call kernels).
 * get_last_walk_stats gives the last call's traversal kernel time (HIP
 * events on the call's stream: the neighbour walk with its exact re-run for
 * k <= 64, the ray walk, the PIP counting walk) and, while counting is on,
 * visits[0] node records fetched, visits[1] primitive records tested and
 * visits[2] wavefront walk steps (per wavefront and work item, the node
 * visits of its busiest lane, summed; for the batched fixed-radius count
 * walk, which may test two nodes per step, its loop iterations), so that
 * visits[0] / (64 * visits[2]) is the node visits per lane and step.  visits
 * has 3 entries.  Counting kernels are slower: time and count in separate
 * calls. */
int rtdl_mi355x_set_walk_counting(uint32_t on, char* error_out, size_t error_size);

/* Arms a host-forced fault for testing failure handling (no counterpart in
 * the reference; the RTDL_MI355X_FAULT environment variable arms one at load):
 * "alloc", "launch" or "timeout", optionally ":N" for the N-th occurrence
 * (default 1) from now: the N-th device allocation fails, the N-th kernel
 * launch check fails, or the N-th served batch to complete is treated as
 * past its deadline.  Each fires once.  "" disarms.  No kernel faults: the
 * failure is raised on the host. */
int rtdl_mi355x_set_fault(const char* spec, char* error_out, size_t error_size);
int rtdl_mi355x_get_last_walk_stats(double* kernel_seconds, uint64_t* visits, char* error_out,
                                    size_t error_size);

/* ---- plain per-call entry points ---- */

/* replaces rtdl_hiprt_run_fixed_radius_neighbors_2d (src/native/hiprt/rtdl_hiprt_api.cpp:488-515) */
int rtdl_mi355x_run_fixed_radius_neighbors_2d(const RtdlPoint* queries, size_t query_count,
                                              const RtdlPoint* search_points, size_t search_count,
                                              double radius, uint32_t k_max,
                                              RtdlFixedRadiusNeighborRow** rows_out,
                                              size_t* row_count_out, char* error_out,
                                              size_t error_size);

/* replaces rtdl_hiprt_run_fixed_radius_neighbors_3d (src/native/hiprt/rtdl_hiprt_api.cpp:410-437) */
int rtdl_mi355x_run_fixed_radius_neighbors_3d(const RtdlPoint3D* queries, size_t query_count,
                                              const RtdlPoint3D* search_points, size_t search_count,
                                              double radius, uint32_t k_max,
                                              RtdlFixedRadiusNeighborRow** rows_out,
                                              size_t* row_count_out, char* error_out,
                                              size_t error_size);

/* replaces the KNN path of run_hiprt (src/rtdsl/hiprt_runtime.py:1398-1475), with the
 * row shape of rtdl_oracle_run_k_closest_hits (src/native/oracle/rtdl_oracle_abi.h:390-409) */
int rtdl_mi355x_run_knn_rows_2d(const RtdlPoint* queries, size_t query_count,
                                const RtdlPoint* search_points, size_t search_count, uint32_t k,
                                RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                char* error_out, size_t error_size);
int rtdl_mi355x_run_knn_rows_3d(const RtdlPoint3D* queries, size_t query_count,
                                const RtdlPoint3D* search_points, size_t search_count, uint32_t k,
                                RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                char* error_out, size_t error_size);

/* replaces the bounded-KNN path of run_hiprt (src/rtdsl/hiprt_runtime.py:1478-1510) */
int rtdl_mi355x_run_bounded_knn_rows_2d(const RtdlPoint* queries, size_t query_count,
                                        const RtdlPoint* search_points, size_t search_count,
                                        double radius, uint32_t k_max,
                                        RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                        char* error_out, size_t error_size);
int rtdl_mi355x_run_bounded_knn_rows_3d(const RtdlPoint3D* queries, size_t query_count,
                                        const RtdlPoint3D* search_points, size_t search_count,
                                        double radius, uint32_t k_max,
                                        RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                        char* error_out, size_t error_size);

/* replaces rtdl_hiprt_run_ray_anyhit_3d (src/native/hiprt/rtdl_hiprt_api.cpp:288-408) */
int rtdl_mi355x_run_ray_anyhit_3d(const RtdlRay3D* rays, size_t ray_count,
                                  const RtdlTriangle3D* triangles, size_t triangle_count,
                                  RtdlRayAnyHitRow** rows_out, size_t* row_count_out,
                                  char* error_out, size_t error_size);

/* replaces rtdl_hiprt_run_ray_anyhit_2d (src/native/hiprt/rtdl_hiprt_api.cpp:555-572); the ray
 * is the segment o -> o + d * tmax (reference.py:783-803) */
int rtdl_mi355x_run_ray_anyhit_2d(const RtdlRay2D* rays, size_t ray_count,
                                  const RtdlTriangle* triangles, size_t triangle_count,
                                  RtdlRayAnyHitRow** rows_out, size_t* row_count_out,
                                  char* error_out, size_t error_size);

/* replaces rtdl_hiprt_run_ray_hitcount_2d (src/native/hiprt/rtdl_hiprt_api.cpp:536-553):
 * one row per ray, the number of triangles the segment hits */
int rtdl_mi355x_run_ray_hitcount_2d(const RtdlRay2D* rays, size_t ray_count,
                                    const RtdlTriangle* triangles, size_t triangle_count,
                                    RtdlRayHitCountRow** rows_out, size_t* row_count_out,
                                    char* error_out, size_t error_size);

/* 2-D closest hit as the reference's CPU path defines it (reference.py:248-266,
 * 775-780; no HIPRT counterpart): every 2-D hit has t = 0.0, so a hit ray's row
 * carries the lowest hit triangle id and t = 0.0; rays without hits emit no row */
int rtdl_mi355x_run_ray_closest_hit_2d(const RtdlRay2D* rays, size_t ray_count,
                                       const RtdlTriangle* triangles, size_t triangle_count,
                                       RtdlRayClosestHitRow** rows_out, size_t* row_count_out,
                                       char* error_out, size_t error_size);

/* replaces rtdl_hiprt_run_ray_hitcount_3d (src/native/hiprt/rtdl_hiprt_api.cpp:183-286) */
int rtdl_mi355x_run_ray_hitcount_3d(const RtdlRay3D* rays, size_t ray_count,
                                    const RtdlTriangle3D* triangles, size_t triangle_count,
                                    RtdlRayHitCountRow** rows_out, size_t* row_count_out,
                                    char* error_out, size_t error_size);

/* replaces the closest-hit runner of the Embree backend
 * (src/native/embree/rtdl_embree_api.cpp:1195, rtdl_embree_run_ray_closest_hit_3d) */
int rtdl_mi355x_run_ray_closest_hit_3d(const RtdlRay3D* rays, size_t ray_count,
                                       const RtdlTriangle3D* triangles, size_t triangle_count,
                                       RtdlRayClosestHitRow** rows_out, size_t* row_count_out,
                                       char* error_out, size_t error_size);

/* replaces rtdl_hiprt_run_point_primitive_anyhit_packet (src/native/hiprt/rtdl_hiprt_api.cpp:574-593);
 * positive_only follows the Embree signature (src/rtdsl/embree_runtime.py:3933):
 * 0 = full_matrix rows, 1 = positive_hits rows. */
int rtdl_mi355x_run_point_in_polygon(const RtdlPoint* points, size_t point_count,
                                     const RtdlPolygonRef* polygons, size_t polygon_count,
                                     const double* vertices_xy, size_t vertex_xy_count,
                                     uint32_t positive_only, RtdlPipRow** rows_out,
                                     size_t* row_count_out, char* error_out, size_t error_size);

/* ---- prepared (scene-resident) entry points ---- */

/* replaces rtdl_hiprt_prepare_fixed_radius_neighbors_3d (src/native/hiprt/rtdl_hiprt_api.cpp:439-454).
 * The prepared point scene serves fixed-radius, KNN and bounded-KNN probes. */
int rtdl_mi355x_prepare_points_2d(const RtdlPoint* search_points, size_t search_count,
                                  void** prepared_out, char* error_out, size_t error_size);
int rtdl_mi355x_prepare_points_3d(const RtdlPoint3D* search_points, size_t search_count,
                                  void** prepared_out, char* error_out, size_t error_size);

/* replaces rtdl_hiprt_run_prepared_fixed_radius_neighbors_3d (src/native/hiprt/rtdl_hiprt_api.cpp:460-486);
 * `queries` points to RtdlPoint or RtdlPoint3D records matching the prepared dimension. */
int rtdl_mi355x_run_prepared_fixed_radius_neighbors(void* prepared, const void* queries,
                                                    size_t query_count, double radius,
                                                    uint32_t k_max,
                                                    RtdlFixedRadiusNeighborRow** rows_out,
                                                    size_t* row_count_out, char* error_out,
                                                    size_t error_size);
int rtdl_mi355x_run_prepared_knn_rows(void* prepared, const void* queries, size_t query_count,
                                      uint32_t k, RtdlKnnNeighborRow** rows_out,
                                      size_t* row_count_out, char* error_out, size_t error_size);
int rtdl_mi355x_run_prepared_bounded_knn_rows(void* prepared, const void* queries,
                                              size_t query_count, double radius, uint32_t k_max,
                                              RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                              char* error_out, size_t error_size);

/* replace rtdl_hiprt_prepare_ray_anyhit_2d / _run_prepared_ray_anyhit_2d
 * (src/native/hiprt/rtdl_hiprt_api.cpp:81-181); the handle is released with
 * rtdl_mi355x_destroy_prepared. */
int rtdl_mi355x_prepare_triangles_2d(const RtdlTriangle* triangles, size_t triangle_count,
                                     void** prepared_out, char* error_out, size_t error_size);
int rtdl_mi355x_run_prepared_ray_anyhit_2d(void* prepared, const RtdlRay2D* rays, size_t ray_count,
                                           RtdlRayAnyHitRow** rows_out, size_t* row_count_out,
                                           char* error_out, size_t error_size);
/* prepared forms of rtdl_mi355x_run_ray_hitcount_2d / _closest_hit_2d above */
int rtdl_mi355x_run_prepared_ray_hitcount_2d(void* prepared, const RtdlRay2D* rays, size_t ray_count,
                                             RtdlRayHitCountRow** rows_out, size_t* row_count_out,
                                             char* error_out, size_t error_size);
int rtdl_mi355x_run_prepared_ray_closest_hit_2d(void* prepared, const RtdlRay2D* rays, size_t ray_count,
                                                RtdlRayClosestHitRow** rows_out, size_t* row_count_out,
                                                char* error_out, size_t error_size);

/* replaces rtdl_hiprt_prepare_ray_hitcount_3d / _run_prepared_ray_hitcount_3d
 * (src/native/hiprt/rtdl_hiprt_api.cpp:15-79); the prepared triangle scene serves
 * any-hit, hit-count and closest-hit probes. */
int rtdl_mi355x_prepare_triangles_3d(const RtdlTriangle3D* triangles, size_t triangle_count,
                                     void** prepared_out, char* error_out, size_t error_size);
int rtdl_mi355x_run_prepared_ray_anyhit_3d(void* prepared, const RtdlRay3D* rays,
                                           size_t ray_count, RtdlRayAnyHitRow** rows_out,
                                           size_t* row_count_out, char* error_out,
                                           size_t error_size);
int rtdl_mi355x_run_prepared_ray_hitcount_3d(void* prepared, const RtdlRay3D* rays,
                                             size_t ray_count, RtdlRayHitCountRow** rows_out,
                                             size_t* row_count_out, char* error_out,
                                             size_t error_size);
int rtdl_mi355x_run_prepared_ray_closest_hit_3d(void* prepared, const RtdlRay3D* rays,
                                                size_t ray_count, RtdlRayClosestHitRow** rows_out,
                                                size_t* row_count_out, char* error_out,
                                                size_t error_size);

/* replaces the prepared PIP path of the Embree backend
 * (src/rtdsl/embree_runtime.py:646-768 prepared executions over polygon scenes) */
int rtdl_mi355x_prepare_polygons(const RtdlPolygonRef* polygons, size_t polygon_count,
                                 const double* vertices_xy, size_t vertex_xy_count,
                                 void** prepared_out, char* error_out, size_t error_size);
int rtdl_mi355x_run_prepared_point_in_polygon(void* prepared, const RtdlPoint* points,
                                              size_t point_count, uint32_t positive_only,
                                              RtdlPipRow** rows_out, size_t* row_count_out,
                                              char* error_out, size_t error_size);

/* replaces rtdl_hiprt_destroy_prepared_* (src/native/hiprt/rtdl_hiprt_api.cpp:48-50,456-458) */
void rtdl_mi355x_destroy_prepared(void* prepared);

/* ---- reference-shaped prepared handles ----
 * The HIPRT backend's prepared calls with their exact parameter lists, so a
 * binding written for them needs only the prefix changed.  They share the
 * prepared scenes above (a handle from one prepare call may also be passed
 * to the generic run/destroy calls of its scene type). */

/* replaces rtdl_hiprt_prepare_fixed_radius_neighbors_3d (src/native/hiprt/rtdl_hiprt_api.cpp:439-454):
 * the radius is bound at prepare time; radius < 0 (or NaN) and an empty
 * search set are errors, as in rtdl_hiprt_core.cpp:1048-1063. */
int rtdl_mi355x_prepare_fixed_radius_neighbors_3d(const RtdlPoint3D* search_points, size_t search_count,
                                                  double radius, void** prepared_out, char* error_out,
                                                  size_t error_size);
/* replaces rtdl_hiprt_run_prepared_fixed_radius_neighbors_3d (src/native/hiprt/rtdl_hiprt_api.cpp:460-486).
 * k_max = 0 is an error; unlike HIPRT (k_max <= 64) any k_max is taken: the
 * bound radius limits the candidates gathered beyond 64. */
int rtdl_mi355x_run_prepared_fixed_radius_neighbors_3d(void* prepared, const RtdlPoint3D* queries,
                                                       size_t query_count, uint32_t k_max,
                                                       RtdlFixedRadiusNeighborRow** rows_out,
                                                       size_t* row_count_out, char* error_out,
                                                       size_t error_size);
/* replaces rtdl_hiprt_destroy_prepared_fixed_radius_neighbors_3d (src/native/hiprt/rtdl_hiprt_api.cpp:456-458) */
void rtdl_mi355x_destroy_prepared_fixed_radius_neighbors_3d(void* prepared);

/* replaces rtdl_hiprt_prepare_ray_hitcount_3d (src/native/hiprt/rtdl_hiprt_api.cpp:15-46); run with
 * rtdl_mi355x_run_prepared_ray_hitcount_3d above (same parameters as :52-79) */
int rtdl_mi355x_prepare_ray_hitcount_3d(const RtdlTriangle3D* triangles, size_t triangle_count,
                                        void** prepared_out, char* error_out, size_t error_size);
/* replaces rtdl_hiprt_destroy_prepared_ray_hitcount_3d (src/native/hiprt/rtdl_hiprt_api.cpp:48-50) */
void rtdl_mi355x_destroy_prepared_ray_hitcount_3d(void* prepared);

/* replaces rtdl_hiprt_prepare_ray_anyhit_2d (src/native/hiprt/rtdl_hiprt_api.cpp:81-151); run with
 * rtdl_mi355x_run_prepared_ray_anyhit_2d above (same parameters as :157-181) */
int rtdl_mi355x_prepare_ray_anyhit_2d(const RtdlTriangle* triangles, size_t triangle_count,
                                      void** prepared_out, char* error_out, size_t error_size);
/* replaces rtdl_hiprt_destroy_prepared_ray_anyhit_2d (src/native/hiprt/rtdl_hiprt_api.cpp:153-155) */
void rtdl_mi355x_destroy_prepared_ray_anyhit_2d(void* prepared);

/* ---- batched and served execution (no counterpart in the reference) ----
 *
 * These entry points take device-resident inputs (FP32 query arrays whose
 * values are exact in FP64) and write device-resident row buffers.  Query
 * ids in the rows are indices into the job's query array.  Neighbour rows
 * use the RtdlFixedRadiusNeighborRow / RtdlKnnNeighborRow layouts, ray rows
 * RtdlRayAnyHitRow / RtdlRayHitCountRow / RtdlRayClosestHitRow, PIP rows
 * RtdlPipRow (positive hits only, at most 16 per point). */

enum {
  RTDL_JOB_FIXED_RADIUS = 0,
  RTDL_JOB_KNN = 1,
  RTDL_JOB_BOUNDED_KNN = 2,
  RTDL_JOB_RAY_ANYHIT = 3,
  RTDL_JOB_RAY_HITCOUNT = 4,
  RTDL_JOB_RAY_CLOSEST = 5,
  RTDL_JOB_PIP_POSITIVE = 6
};

/* status bits of a job / request.  A PIP job's default row budget is 16 rows
 * per point on average (row_capacity = 16 * query_count); one point may have
 * any number of containing polygons within that budget. */
#define RTDL_STATUS_ROW_OVERFLOW 1u /* the job's rows did not fit row_capacity: rows are incomplete */
#define RTDL_STATUS_REJECTED 4u     /* served request whose rows can exceed the serving row arena */
#define RTDL_STATUS_INTERNAL 8u     /* an internal invariant failed: the job's rows are not valid */
#define RTDL_STATUS_FAILED 16u      /* served request whose batch failed, and which failed again alone */
#define RTDL_STATUS_TIMEOUT 32u     /* served request not done by its deadline (its rows may be late) */
#define RTDL_STATUS_INVALID 64u     /* served request naming a missing scene, a scene of another type or a null arena */

typedef struct RtdlBatchJob {
  void* scene;           /* prepared handle */
  uint32_t instance;     /* instance of a multi-instance scene, 0 otherwise */
  uint32_t kind;         /* RTDL_JOB_* */
  const float* queries;  /* device pointer, query_stride floats per query */
  uint32_t query_count;
  uint32_t query_stride; /* 2 (2-D points), 3 (3-D points, z = 0 for 2-D) or 7 (rays) */
  double radius;         /* fixed radius / bounded KNN */
  uint32_t k;            /* k (KNN) or k_max (fixed radius, bounded KNN), 1..64 */
  uint32_t pad;
  void* rows;            /* device row buffer */
  uint64_t row_capacity; /* rows the buffer holds, < 2^32: at least query_count for any-hit and
                            hit-count jobs; other kinds may hold fewer than the most they can
                            emit (k per query, ...) and set RTDL_STATUS_ROW_OVERFLOW, with
                            incomplete rows, if they run out */
} RtdlBatchJob;

typedef struct RtdlBatchJobResult {
  uint64_t rows;      /* rows written */
  uint32_t status;    /* RTDL_STATUS_* bits */
  uint32_t ambiguous; /* selections recomputed with the exact FP64 list */
} RtdlBatchJobResult;

/* Runs all jobs as one batch: one persistent launch per predicate family
 * over a device work queue, optional (job, Morton) query ordering, rows
 * compacted per job by wavefront prefix sums.  Blocks until the batch is
 * done; seconds_out (optional) receives the device time of the batch. */
int rtdl_mi355x_run_batch(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                          RtdlBatchJobResult* results, double* seconds_out, char* error_out,
                          size_t error_size);

/* rtdl_mi355x_run_batch for callers that cannot allocate (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It has no
 * counterpart in the reference.  Same jobs, rows and results as
 * rtdl_mi355x_run_batch, but the call allocates nothing: no device, pinned
 * or host heap memory, no stream and no event.  Every scratch buffer of the
 * batch (job descriptors, work items, counters, sort keys, fixed-radius
 * plans, stage and parking space) lives in a static 64 MiB device array that
 * each GPU holds until the process ends; the caller passes the inputs and the
 * row buffers as for rtdl_mi355x_run_batch.  Limits: at most 128 jobs and
 * 16,384 queries per call (larger batches are rejected with an error).  Runs
 * on the null stream and blocks until the batch is done; seconds_out
 * (optional) receives the call's wall time.  Calls are serialised in a
 * process: the array holds one batch at a time, and each call clears the
 * counters it needs. */
int rtdl_mi355x_run_batch_noalloc(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                                  RtdlBatchJobResult* results, double* seconds_out, char* error_out,
                                  size_t error_size);

/* rtdl_mi355x_run_batch with measurement: family_seconds (5 entries:
 * neighbour, ray any-hit, ray hit-count, ray closest-hit, PIP) receives each
 * family's kernel time from events on the launch stream; visits (3 entries,
 * optional, counting kernels) receives the node records fetched, the
 * primitive records tested and the wavefront walk steps (as for
 * rtdl_mi355x_get_last_walk_stats) of all traversals of the batch. */
int rtdl_mi355x_run_batch_profiled(const RtdlBatchJob* jobs, size_t job_count,
                                   uint32_t morton_sort, RtdlBatchJobResult* results,
                                   double* family_seconds, uint64_t* visits, char* error_out,
                                   size_t error_size);

/* Host copy of a prepared point or triangle scene (no counterpart in the
 * reference; for checking the traversal visit counters against a replay of
 * the walk on the host).  Call once with null buffers to get the sizes, then
 * with buffers of node_count 32-byte nodes, prim_count * prim_float4s
 * float4 records, prim_count * exact_doubles doubles (when exact_doubles is
 * nonzero) and instance_count roots and ends. */
typedef struct RtdlSceneExport {
  uint64_t node_count;     /* BvhNode {float lo[3]; uint32 link; float hi[3]; uint32 escape} */
  uint64_t prim_count;
  uint64_t instance_count;
  uint32_t prim_float4s;   /* float4 records per primitive: 1 (point), 3 (triangle) */
  uint32_t exact_doubles;  /* FP64 coordinates per primitive kept for refine, 0 if FP32 is exact */
  float m;                 /* bound on every |coordinate| of the scene */
  uint32_t pad;
} RtdlSceneExport;

int rtdl_mi355x_scene_export(void* prepared, RtdlSceneExport* info, void* nodes, void* prims,
                             double* exact, uint32_t* roots, uint32_t* ends, char* error_out,
                             size_t error_size);

/* Host copy of a prepared point scene's compact nodes (no counterpart in the
 * reference; for checking them and the fixed-radius count walk that reads
 * them against a replay on the host): node_count 16-byte records (three
 * words of 16-bit box steps (lo x, lo y), (lo z, hi x), (hi y, hi z), then a
 * leaf's link or an internal node's escape, node_count meaning the end) and
 * the 8-float frame (origin[3], step[3], unusable flag, 0).  has_compact
 * receives 0 for scenes without them (non-point or empty scenes). */
int rtdl_mi355x_scene_export_compact(void* prepared, uint32_t* has_compact, void* qnodes, float* frame,
                                     char* error_out, size_t error_size);

/* Process-wide build and launch choices (no counterpart in the reference;
 * for the autotuner in rtdl_amd/autotune.py and callers that pin a tuned
 * configuration).  Every setting gives the same rows; scenes built and
 * graphs captured earlier keep the values they were made with. */
typedef struct RtdlTuning {
  uint32_t point_leaf;     /* primitives per LBVH leaf in point scenes, 1..8 (default 2) */
  uint32_t triangle_leaf;  /* triangle scenes, 1..8 (default 1) */
  uint32_t polygon_leaf;   /* polygon scenes, 1..8 (default 1) */
  uint32_t blocks_per_cu;  /* persistent batch launches: 256-thread workgroups per CU, 1..64 (default 8) */
  uint32_t knn_stack;      /* batched KNN walk: per-lane LDS node stack entries, 16, 32 or 48 (default 32);
                              a walk that fills it walks the tree again */
  uint32_t walk_block;     /* batched ray and PIP walks: threads per workgroup, 64, 128 or 256 (default 256) */
} RtdlTuning;

int rtdl_mi355x_get_tuning(RtdlTuning* out, char* error_out, size_t error_size);
int rtdl_mi355x_set_tuning(const RtdlTuning* tuning, char* error_out, size_t error_size);

/* A batch recorded once into an executable hipGraph and replayed with one
 * call (no counterpart in the reference; for callers that run the same
 * batch shape repeatedly, e.g. new query contents in the same device buffers
 * every frame).  Create validates the jobs like rtdl_mi355x_run_batch,
 * sizes every buffer and captures the whole batch (job descriptors,
 * ordering sort, family launches, counter copy) on a stream of its own;
 * launch replays it, blocks until it is done and fills results (job_count
 * entries) and seconds_out (optional, device time).  The jobs' scenes,
 * query and row buffers are read at every launch and must outlive the
 * graph; their contents may change between launches.  Launches of one graph
 * must not overlap. */
int rtdl_mi355x_batch_graph_create(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                                   void** graph_out, char* error_out, size_t error_size);
int rtdl_mi355x_batch_graph_launch(void* graph, RtdlBatchJobResult* results, double* seconds_out,
                                   char* error_out, size_t error_size);
int rtdl_mi355x_batch_graph_destroy(void* graph, char* error_out, size_t error_size);

/* Moving primitives (no counterpart in the reference's HIPRT API; the OptiX
 * backend's rebuild-in-place of a geometry, src/native/optix/
 * rtdl_optix_api.cpp:4323-4393 and 6249-6288, is the closest).
 * rebuild_instances rebuilds the trees of a point or triangle instance scene
 * (made by rtdl_mi355x_prepare_point_instances[_leaf] /
 * _prepare_triangle_instances) from new FP32 device coordinates of the same
 * primitives: same count, layout and instance offsets, so handles, ids and
 * jobs stay valid.  batch_graph_create_rebuild records, in front of the batch,
 * the rebuild of scenes[i] from the device buffer coords[i] (scene_count
 * pairs; each scene at most once), so every launch rebuilds those scenes from
 * the buffers' current contents and then runs the batch; launch reports an
 * error a rebuild met (a coordinate that is not exact in FP32, such as NaN).
 * The first rebuild of a scene keeps its build buffers for later ones (about
 * 100 bytes per primitive); a scene in such a graph must not be queried,
 * rebuilt or destroyed outside it while the graph is alive. */
int rtdl_mi355x_rebuild_instances(void* prepared, const float* coords, char* error_out, size_t error_size);
int rtdl_mi355x_batch_graph_create_rebuild(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                                           void* const* scenes, const float* const* coords,
                                           size_t scene_count, void** graph_out, char* error_out,
                                           size_t error_size);

/* Multi-instance scenes for batched jobs: instance i owns the primitives
 * [instance_offsets[i], instance_offsets[i + 1]) of the device input and
 * gets its own LBVH subtree; neighbour / polygon ids are instance-local
 * indices.  coords: 3 FP32 per point (z = 0 for 2-D instances).
 * Polygons: refs (vertex offset, vertex count) per polygon with offsets into
 * the device vertex array (2 FP32 per vertex). */
int rtdl_mi355x_prepare_point_instances(const float* coords, size_t point_count,
                                        const uint64_t* instance_offsets, size_t instance_count,
                                        void** prepared_out, char* error_out, size_t error_size);
/* The same with the points per leaf of this scene's trees (1-8; 0 takes the
 * process tuning): a scene queried by fixed-radius jobs walks fewer nodes
 * and fetches fewer bytes with larger leaves, while KNN walks prefer small
 * ones, so each scene can be built for the jobs it serves (no counterpart in
 * the reference). */
int rtdl_mi355x_prepare_point_instances_leaf(const float* coords, size_t point_count,
                                             const uint64_t* instance_offsets, size_t instance_count,
                                             uint32_t leaf_size, void** prepared_out, char* error_out,
                                             size_t error_size);
/* coords: 9 FP32 per triangle (three vertices); ids are instance-local. */
int rtdl_mi355x_prepare_triangle_instances(const float* coords, size_t triangle_count,
                                           const uint64_t* instance_offsets, size_t instance_count,
                                           void** prepared_out, char* error_out, size_t error_size);
int rtdl_mi355x_prepare_polygon_instances(const uint32_t* vertex_offsets,
                                          const uint32_t* vertex_counts, size_t polygon_count,
                                          const float* vertices_xy, size_t vertex_count,
                                          const uint64_t* instance_offsets, size_t instance_count,
                                          void** prepared_out, char* error_out, size_t error_size);

typedef struct RtdlServeRequest {
  uint32_t kind;         /* RTDL_JOB_* */
  uint32_t scene;        /* index into the scene handle array */
  uint64_t query_offset; /* first query of the request in the kind's arena */
  uint32_t query_count;
  uint32_t capture;      /* 1: return this request's rows to the host */
  double arrival_s;      /* arrival time from the start of the replay; non-decreasing */
} RtdlServeRequest;

typedef struct RtdlServeConfig {
  double radius;                /* fixed-radius requests */
  uint32_t k_max;               /* fixed-radius requests, 1..64 */
  uint32_t k;                   /* KNN requests, 1..64 */
  uint32_t slots;               /* batches in flight, each on its own stream */
  uint32_t morton_sort;         /* order each batch's queries by (request, Morton code) */
  uint64_t memory_budget_bytes; /* device bytes for the row buffers of all in-flight batches */
  uint64_t max_batch_queries;
  const float* point_queries;   /* device arena, 3 FP32 per query (neighbour requests) */
  const float* ray_queries;     /* device arena, 7 FP32 per ray */
  const float* pip_queries;     /* device arena, 2 FP32 per point */
  uint32_t realtime;            /* 1: honour arrival times; 0: everything arrives at 0 */
  uint32_t timeout_ms;          /* > 0: requests whose batch is not done this long after dispatch
                                   are marked RTDL_STATUS_TIMEOUT (the loop still waits for it) */
} RtdlServeConfig;

typedef struct RtdlServeResult {
  double arrival_s;
  double dispatch_s;
  double done_s;
  uint64_t rows;
  uint32_t status;     /* RTDL_STATUS_* bits */
  uint32_t batch;
  void* captured_rows; /* with capture: host rows, release with rtdl_mi355x_free_rows */
} RtdlServeResult;

typedef struct RtdlServeStats {
  double wall_s;
  uint64_t batches;
  uint64_t queries;
  uint64_t rows;
  uint64_t peak_inflight_bytes; /* largest sum of in-flight row buffers (<= region_bytes) */
  uint64_t region_bytes;        /* the row arena: min(budget, slots x the largest admissible batch) */
  uint64_t retried;             /* requests re-run alone after their batch failed */
  uint64_t failed;              /* requests that failed alone as well (RTDL_STATUS_FAILED) */
} RtdlServeStats;

/* Replays a request stream against resident scenes with continuous
 * batching: up to `slots` batches in flight, each admitted from the
 * requests that have arrived, in order, while their row buffers fit the
 * largest free range of one row arena bounded by the memory budget (the
 * first request that does not fit waits for completions to free memory; one
 * that can never fit is rejected).  Latency = done_s - arrival_s.
 * Failures stay with the requests they belong to: an invalid request is
 * marked RTDL_STATUS_INVALID and skipped; a batch whose launch fails is
 * re-run one request at a time, and only a request that fails alone too is
 * marked RTDL_STATUS_FAILED; the other requests are served normally. */
int rtdl_mi355x_serve_replay(void* const* scenes, size_t scene_count,
                             const RtdlServeRequest* requests, size_t request_count,
                             const RtdlServeConfig* config, RtdlServeResult* results,
                             RtdlServeStats* stats, char* error_out, size_t error_size);

/* ---- device buffers (no counterpart in the reference) ----
 * Plumbing for callers that keep query arenas, instance inputs and row
 * buffers resident in HBM across batches (the batched and served entry
 * points above take device pointers).  Allocation is hipMalloc on the
 * current device; copies are synchronous with respect to the host and to
 * every batch the library has completed. */
int rtdl_mi355x_device_alloc(size_t bytes, void** ptr_out, char* error_out, size_t error_size);
/* the calling thread's current HIP device (the device of its allocations) */
int rtdl_mi355x_get_device(int* device_out, char* error_out, size_t error_size);
int rtdl_mi355x_device_free(void* ptr, char* error_out, size_t error_size);
int rtdl_mi355x_copy_to_device(void* dst, const void* src, size_t bytes, char* error_out,
                               size_t error_size);
int rtdl_mi355x_copy_to_host(void* dst, const void* src, size_t bytes, char* error_out,
                             size_t error_size);
/* device-to-device copy on the calling thread's stream (synchronous) */
int rtdl_mi355x_copy_device(void* dst, const void* src, size_t bytes, char* error_out,
                            size_t error_size);

/* ---- diagnostics (no counterpart in the reference) ---- */

/* Evaluates sqrt(a[i]) and a[i] / b[i] in FP64 on the device with the same
 * code generation as the refine kernels, so tests can confirm both are
 * correctly rounded (the refine step relies on it for bit-exact rows). */
int rtdl_mi355x_selftest_fp64(const double* a, const double* b, size_t n, double* sqrt_out,
                              double* div_out, char* error_out, size_t error_size);

#ifdef __cplusplus
}
#endif

#endif /* RTDL_MI355X_H */
