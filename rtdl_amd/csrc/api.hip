// extern "C" entry points of librtdl_mi355x.so (declared in
// include/rtdl_mi355x.h).  Each wraps a runtime pipeline in the status /
// error-buffer convention of the reference's native backends
// (reference: src/native/hiprt/rtdl_hiprt_core.cpp:3323-3335).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "../../include/rtdl_mi355x.h"
#include "common.hpp"
#include "engine.hpp"
#include "io.hpp"
#include "noalloc.hpp"
#include "runtime.hpp"
#include "serve.hpp"

using namespace rtdl;

namespace {

constexpr int kVersionMajor = 0;
constexpr int kVersionMinor = 1;
constexpr int kVersionPatch = 0;

template <class Row>
void zero_outputs(Row** rows_out, size_t* n_out) {
  require(rows_out != nullptr && n_out != nullptr, "output pointers must not be null");
  *rows_out = nullptr;
  *n_out = 0;
}

// Per-call runs build a throw-away scene on the calling thread's stream.
template <class Build, class Run>
void per_call(Build&& build, Run&& run) {
  PhaseClock& clock = thread_clock();
  clock.begin();
  Prepared scene(false);
  build(scene);
  RTDL_HIP(hipStreamSynchronize(scene.stream));
  clock.lap(1);
  run(scene);
  clock.end();
}

template <class Run>
void on_prepared(void* handle, Run&& run) {
  Prepared* p = prepared_from(handle);
  std::lock_guard<std::mutex> g(p->lock);
  PhaseClock& clock = thread_clock();
  clock.begin();
  run(*p);
  clock.end();
}

template <class Build>
void make_prepared(void** prepared_out, Build&& build) {
  require(prepared_out != nullptr, "prepared_out must not be null");
  *prepared_out = nullptr;
  PhaseClock& clock = thread_clock();
  clock.begin();
  auto p = std::make_unique<Prepared>(true);
  build(*p);
  RTDL_HIP(hipStreamSynchronize(p->stream));
  clock.lap(1);
  clock.end();
  *prepared_out = p.release();
}

}  // namespace

extern "C" {

int rtdl_mi355x_get_version(int* major, int* minor, int* patch) {
  if (!major || !minor || !patch) return 1;
  *major = kVersionMajor;
  *minor = kVersionMinor;
  *patch = kVersionPatch;
  return 0;
}

int rtdl_mi355x_context_probe(char* device_name, size_t device_name_size, int* device_type,
                              int* api_version, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(device_name != nullptr && device_name_size > 0 && device_type != nullptr &&
                api_version != nullptr,
            "null output passed to rtdl_mi355x_context_probe");
    device_name[0] = '\0';
    *device_type = -1;
    int runtime = 0;
    RTDL_HIP(hipRuntimeGetVersion(&runtime));
    *api_version = runtime;
    int count = 0;
    RTDL_HIP(hipGetDeviceCount(&count));
    require(count > 0, "no HIP device is visible");
    hipDeviceProp_t props;
    RTDL_HIP(hipGetDeviceProperties(&props, 0));
    std::snprintf(device_name, device_name_size, "%s (%s)", props.name, props.gcnArchName);
    *device_type = 1;
  });
}

void rtdl_mi355x_free_rows(void* rows) { free_host_rows(rows); }

int rtdl_mi355x_set_fault(const char* spec, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] { set_fault(spec); });
}

int rtdl_mi355x_set_walk_counting(uint32_t on, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] { thread_walk_stats().counting = on != 0; });
}

int rtdl_mi355x_get_last_walk_stats(double* kernel_seconds, uint64_t* visits, char* error_out,
                                    size_t error_size) {
  return guarded(error_out, error_size, [&] {
    const WalkStats& w = thread_walk_stats();
    require(w.valid, "no traversal has been measured on this thread");
    if (kernel_seconds) *kernel_seconds = w.kernel_s;
    if (visits) {
      visits[0] = w.nodes;
      visits[1] = w.prims;
      visits[2] = w.steps;
    }
  });
}

int rtdl_mi355x_get_last_phase_timings(double* seconds_out, size_t capacity) {
  if (!seconds_out) return 0;
  const auto& s = thread_clock().seconds;
  const size_t n = capacity < s.size() ? capacity : s.size();
  for (size_t i = 0; i < n; ++i) seconds_out[i] = s[i];
  return (int)n;
}

// ---------------------------------------------------------- neighbours --

static int neighbors_per_call(const void* queries, size_t nq, const void* search, size_t ns,
                              int dim, NeighborMode mode, double radius, uint32_t k,
                              void** rows_out, size_t* n_out, char* error_out,
                              size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, n_out);
    require((nq == 0 || queries != nullptr) && (ns == 0 || search != nullptr),
            "input pointers must not be null when counts are nonzero");
    require(k >= 1, mode == NeighborMode::Knn ? "knn_rows k must be positive"
                                              : "k_max must be positive");
    require(!(radius < 0.0), "radius must be non-negative");
    per_call([&](Prepared& p) { prepare_points(p, search, ns, dim); },
             [&](Prepared& p) { run_neighbors(p, queries, nq, mode, radius, k, rows_out, n_out); });
  });
}

int rtdl_mi355x_run_fixed_radius_neighbors_2d(const RtdlPoint* queries, size_t query_count,
                                              const RtdlPoint* search_points, size_t search_count,
                                              double radius, uint32_t k_max,
                                              RtdlFixedRadiusNeighborRow** rows_out,
                                              size_t* row_count_out, char* error_out,
                                              size_t error_size) {
  return neighbors_per_call(queries, query_count, search_points, search_count, 2,
                            NeighborMode::FixedRadius, radius, k_max, (void**)rows_out,
                            row_count_out, error_out, error_size);
}

int rtdl_mi355x_run_fixed_radius_neighbors_3d(const RtdlPoint3D* queries, size_t query_count,
                                              const RtdlPoint3D* search_points, size_t search_count,
                                              double radius, uint32_t k_max,
                                              RtdlFixedRadiusNeighborRow** rows_out,
                                              size_t* row_count_out, char* error_out,
                                              size_t error_size) {
  return neighbors_per_call(queries, query_count, search_points, search_count, 3,
                            NeighborMode::FixedRadius, radius, k_max, (void**)rows_out,
                            row_count_out, error_out, error_size);
}

int rtdl_mi355x_run_knn_rows_2d(const RtdlPoint* queries, size_t query_count,
                                const RtdlPoint* search_points, size_t search_count, uint32_t k,
                                RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                char* error_out, size_t error_size) {
  return neighbors_per_call(queries, query_count, search_points, search_count, 2,
                            NeighborMode::Knn, 0.0, k, (void**)rows_out, row_count_out, error_out,
                            error_size);
}

int rtdl_mi355x_run_knn_rows_3d(const RtdlPoint3D* queries, size_t query_count,
                                const RtdlPoint3D* search_points, size_t search_count, uint32_t k,
                                RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                char* error_out, size_t error_size) {
  return neighbors_per_call(queries, query_count, search_points, search_count, 3,
                            NeighborMode::Knn, 0.0, k, (void**)rows_out, row_count_out, error_out,
                            error_size);
}

int rtdl_mi355x_run_bounded_knn_rows_2d(const RtdlPoint* queries, size_t query_count,
                                        const RtdlPoint* search_points, size_t search_count,
                                        double radius, uint32_t k_max,
                                        RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                        char* error_out, size_t error_size) {
  return neighbors_per_call(queries, query_count, search_points, search_count, 2,
                            NeighborMode::BoundedKnn, radius, k_max, (void**)rows_out,
                            row_count_out, error_out, error_size);
}

int rtdl_mi355x_run_bounded_knn_rows_3d(const RtdlPoint3D* queries, size_t query_count,
                                        const RtdlPoint3D* search_points, size_t search_count,
                                        double radius, uint32_t k_max,
                                        RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                        char* error_out, size_t error_size) {
  return neighbors_per_call(queries, query_count, search_points, search_count, 3,
                            NeighborMode::BoundedKnn, radius, k_max, (void**)rows_out,
                            row_count_out, error_out, error_size);
}

// ------------------------------------------------------------------ rays --

static int rays_per_call(const RtdlRay3D* rays, size_t nr, const RtdlTriangle3D* tris, size_t nt,
                         int mode, void** rows_out, size_t* n_out, char* error_out,
                         size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, n_out);
    require((nr == 0 || rays != nullptr) && (nt == 0 || tris != nullptr),
            "input pointers must not be null when counts are nonzero");
    per_call([&](Prepared& p) { prepare_triangles(p, tris, nt); },
             [&](Prepared& p) { run_rays(p, rays, nr, mode, rows_out, n_out); });
  });
}

int rtdl_mi355x_run_ray_anyhit_3d(const RtdlRay3D* rays, size_t ray_count,
                                  const RtdlTriangle3D* triangles, size_t triangle_count,
                                  RtdlRayAnyHitRow** rows_out, size_t* row_count_out,
                                  char* error_out, size_t error_size) {
  return rays_per_call(rays, ray_count, triangles, triangle_count, 0, (void**)rows_out,
                       row_count_out, error_out, error_size);
}

static int rays_2d_per_call(const RtdlRay2D* rays, size_t ray_count, const RtdlTriangle* triangles,
                            size_t triangle_count, int mode, void** rows_out, size_t* row_count_out,
                            char* error_out, size_t error_size) {
  static_assert(sizeof(RtdlRay2D) == 44 && sizeof(RtdlTriangle) == 56, "reference record layouts");
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, row_count_out);
    require((ray_count == 0 || rays != nullptr) && (triangle_count == 0 || triangles != nullptr),
            "input pointers must not be null when counts are nonzero");
    per_call([&](Prepared& p) { prepare_triangles_2d(p, triangles, triangle_count); },
             [&](Prepared& p) { run_rays_2d(p, rays, ray_count, mode, rows_out, row_count_out); });
  });
}

int rtdl_mi355x_run_ray_anyhit_2d(const RtdlRay2D* rays, size_t ray_count,
                                  const RtdlTriangle* triangles, size_t triangle_count,
                                  RtdlRayAnyHitRow** rows_out, size_t* row_count_out,
                                  char* error_out, size_t error_size) {
  return rays_2d_per_call(rays, ray_count, triangles, triangle_count, 0, (void**)rows_out, row_count_out,
                          error_out, error_size);
}

int rtdl_mi355x_run_ray_hitcount_2d(const RtdlRay2D* rays, size_t ray_count,
                                    const RtdlTriangle* triangles, size_t triangle_count,
                                    RtdlRayHitCountRow** rows_out, size_t* row_count_out,
                                    char* error_out, size_t error_size) {
  return rays_2d_per_call(rays, ray_count, triangles, triangle_count, 1, (void**)rows_out, row_count_out,
                          error_out, error_size);
}

int rtdl_mi355x_run_ray_closest_hit_2d(const RtdlRay2D* rays, size_t ray_count,
                                       const RtdlTriangle* triangles, size_t triangle_count,
                                       RtdlRayClosestHitRow** rows_out, size_t* row_count_out,
                                       char* error_out, size_t error_size) {
  return rays_2d_per_call(rays, ray_count, triangles, triangle_count, 2, (void**)rows_out, row_count_out,
                          error_out, error_size);
}

int rtdl_mi355x_run_ray_hitcount_3d(const RtdlRay3D* rays, size_t ray_count,
                                    const RtdlTriangle3D* triangles, size_t triangle_count,
                                    RtdlRayHitCountRow** rows_out, size_t* row_count_out,
                                    char* error_out, size_t error_size) {
  return rays_per_call(rays, ray_count, triangles, triangle_count, 1, (void**)rows_out,
                       row_count_out, error_out, error_size);
}

int rtdl_mi355x_run_ray_closest_hit_3d(const RtdlRay3D* rays, size_t ray_count,
                                       const RtdlTriangle3D* triangles, size_t triangle_count,
                                       RtdlRayClosestHitRow** rows_out, size_t* row_count_out,
                                       char* error_out, size_t error_size) {
  return rays_per_call(rays, ray_count, triangles, triangle_count, 2, (void**)rows_out,
                       row_count_out, error_out, error_size);
}

// ------------------------------------------------------------------- PIP --

int rtdl_mi355x_run_point_in_polygon(const RtdlPoint* points, size_t point_count,
                                     const RtdlPolygonRef* polygons, size_t polygon_count,
                                     const double* vertices_xy, size_t vertex_xy_count,
                                     uint32_t positive_only, RtdlPipRow** rows_out,
                                     size_t* row_count_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, row_count_out);
    require((point_count == 0 || points != nullptr) && (polygon_count == 0 || polygons != nullptr),
            "input pointers must not be null when counts are nonzero");
    per_call(
        [&](Prepared& p) { prepare_polygons(p, polygons, polygon_count, vertices_xy, vertex_xy_count); },
        [&](Prepared& p) {
          run_pip(p, points, point_count, positive_only != 0, (void**)rows_out, row_count_out);
        });
  });
}

// -------------------------------------------------------------- prepared --

int rtdl_mi355x_prepare_points_2d(const RtdlPoint* search_points, size_t search_count,
                                  void** prepared_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(search_count == 0 || search_points != nullptr,
            "search point pointer must not be null when count is nonzero");
    make_prepared(prepared_out, [&](Prepared& p) { prepare_points(p, search_points, search_count, 2); });
  });
}

int rtdl_mi355x_prepare_points_3d(const RtdlPoint3D* search_points, size_t search_count,
                                  void** prepared_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(search_count == 0 || search_points != nullptr,
            "search point pointer must not be null when count is nonzero");
    make_prepared(prepared_out, [&](Prepared& p) { prepare_points(p, search_points, search_count, 3); });
  });
}

int rtdl_mi355x_run_prepared_fixed_radius_neighbors(void* prepared, const void* queries,
                                                    size_t query_count, double radius,
                                                    uint32_t k_max,
                                                    RtdlFixedRadiusNeighborRow** rows_out,
                                                    size_t* row_count_out, char* error_out,
                                                    size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, row_count_out);
    require(k_max >= 1, "k_max must be positive");
    on_prepared(prepared, [&](Prepared& p) {
      run_neighbors(p, queries, query_count, NeighborMode::FixedRadius, radius, k_max,
                    (void**)rows_out, row_count_out);
    });
  });
}

int rtdl_mi355x_run_prepared_knn_rows(void* prepared, const void* queries, size_t query_count,
                                      uint32_t k, RtdlKnnNeighborRow** rows_out,
                                      size_t* row_count_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, row_count_out);
    require(k >= 1, "knn_rows k must be positive");
    on_prepared(prepared, [&](Prepared& p) {
      run_neighbors(p, queries, query_count, NeighborMode::Knn, 0.0, k, (void**)rows_out,
                    row_count_out);
    });
  });
}

int rtdl_mi355x_run_prepared_bounded_knn_rows(void* prepared, const void* queries,
                                              size_t query_count, double radius, uint32_t k_max,
                                              RtdlKnnNeighborRow** rows_out, size_t* row_count_out,
                                              char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, row_count_out);
    require(k_max >= 1, "k_max must be positive");
    on_prepared(prepared, [&](Prepared& p) {
      run_neighbors(p, queries, query_count, NeighborMode::BoundedKnn, radius, k_max,
                    (void**)rows_out, row_count_out);
    });
  });
}

int rtdl_mi355x_prepare_triangles_3d(const RtdlTriangle3D* triangles, size_t triangle_count,
                                     void** prepared_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(triangle_count == 0 || triangles != nullptr,
            "triangle pointer must not be null when count is nonzero");
    make_prepared(prepared_out, [&](Prepared& p) { prepare_triangles(p, triangles, triangle_count); });
  });
}

int rtdl_mi355x_prepare_triangles_2d(const RtdlTriangle* triangles, size_t triangle_count,
                                     void** prepared_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(triangle_count == 0 || triangles != nullptr,
            "triangle pointer must not be null when count is nonzero");
    make_prepared(prepared_out, [&](Prepared& p) { prepare_triangles_2d(p, triangles, triangle_count); });
  });
}

static int rays_2d_prepared(void* prepared, const RtdlRay2D* rays, size_t nr, int mode, void** rows_out,
                            size_t* n_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, n_out);
    on_prepared(prepared, [&](Prepared& p) { run_rays_2d(p, rays, nr, mode, rows_out, n_out); });
  });
}

int rtdl_mi355x_run_prepared_ray_anyhit_2d(void* prepared, const RtdlRay2D* rays, size_t ray_count,
                                           RtdlRayAnyHitRow** rows_out, size_t* row_count_out,
                                           char* error_out, size_t error_size) {
  return rays_2d_prepared(prepared, rays, ray_count, 0, (void**)rows_out, row_count_out, error_out,
                          error_size);
}

int rtdl_mi355x_run_prepared_ray_hitcount_2d(void* prepared, const RtdlRay2D* rays, size_t ray_count,
                                             RtdlRayHitCountRow** rows_out, size_t* row_count_out,
                                             char* error_out, size_t error_size) {
  return rays_2d_prepared(prepared, rays, ray_count, 1, (void**)rows_out, row_count_out, error_out,
                          error_size);
}

int rtdl_mi355x_run_prepared_ray_closest_hit_2d(void* prepared, const RtdlRay2D* rays, size_t ray_count,
                                                RtdlRayClosestHitRow** rows_out, size_t* row_count_out,
                                                char* error_out, size_t error_size) {
  return rays_2d_prepared(prepared, rays, ray_count, 2, (void**)rows_out, row_count_out, error_out,
                          error_size);
}

static int rays_prepared(void* prepared, const RtdlRay3D* rays, size_t nr, int mode,
                         void** rows_out, size_t* n_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, n_out);
    on_prepared(prepared, [&](Prepared& p) { run_rays(p, rays, nr, mode, rows_out, n_out); });
  });
}

int rtdl_mi355x_run_prepared_ray_anyhit_3d(void* prepared, const RtdlRay3D* rays,
                                           size_t ray_count, RtdlRayAnyHitRow** rows_out,
                                           size_t* row_count_out, char* error_out,
                                           size_t error_size) {
  return rays_prepared(prepared, rays, ray_count, 0, (void**)rows_out, row_count_out, error_out,
                       error_size);
}

int rtdl_mi355x_run_prepared_ray_hitcount_3d(void* prepared, const RtdlRay3D* rays,
                                             size_t ray_count, RtdlRayHitCountRow** rows_out,
                                             size_t* row_count_out, char* error_out,
                                             size_t error_size) {
  return rays_prepared(prepared, rays, ray_count, 1, (void**)rows_out, row_count_out, error_out,
                       error_size);
}

int rtdl_mi355x_run_prepared_ray_closest_hit_3d(void* prepared, const RtdlRay3D* rays,
                                                size_t ray_count, RtdlRayClosestHitRow** rows_out,
                                                size_t* row_count_out, char* error_out,
                                                size_t error_size) {
  return rays_prepared(prepared, rays, ray_count, 2, (void**)rows_out, row_count_out, error_out,
                       error_size);
}

int rtdl_mi355x_prepare_polygons(const RtdlPolygonRef* polygons, size_t polygon_count,
                                 const double* vertices_xy, size_t vertex_xy_count,
                                 void** prepared_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(polygon_count == 0 || polygons != nullptr,
            "polygon pointer must not be null when count is nonzero");
    require(vertex_xy_count == 0 || vertices_xy != nullptr,
            "vertex pointer must not be null when vertex count is nonzero");
    make_prepared(prepared_out, [&](Prepared& p) {
      prepare_polygons(p, polygons, polygon_count, vertices_xy, vertex_xy_count);
    });
  });
}

int rtdl_mi355x_run_prepared_point_in_polygon(void* prepared, const RtdlPoint* points,
                                              size_t point_count, uint32_t positive_only,
                                              RtdlPipRow** rows_out, size_t* row_count_out,
                                              char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, row_count_out);
    on_prepared(prepared, [&](Prepared& p) {
      run_pip(p, points, point_count, positive_only != 0, (void**)rows_out, row_count_out);
    });
  });
}

void rtdl_mi355x_destroy_prepared(void* prepared) {
  if (!prepared) return;
  auto* p = static_cast<Prepared*>(prepared);
  if (p->magic != 0x52544D58u) return;
  delete p;
}

// ------------------------------------- reference-shaped prepared handles --
//
// The HIPRT backend's prepared calls, with their parameter lists, over the
// same prepared scenes (src/native/hiprt/rtdl_hiprt_api.cpp:15-79, 81-181,
// 439-486).

int rtdl_mi355x_prepare_fixed_radius_neighbors_3d(const RtdlPoint3D* search_points, size_t search_count,
                                                  double radius, void** prepared_out, char* error_out,
                                                  size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(prepared_out != nullptr, "prepared_out must not be null");
    *prepared_out = nullptr;
    // rtdl_hiprt_core.cpp:1048-1063
    require(!(radius < 0.0) && radius == radius, "fixed_radius_neighbors radius must be non-negative");
    require(search_points != nullptr || search_count == 0,
            "search point pointer must not be null when count is nonzero");
    require(search_count > 0, "prepared fixed_radius_neighbors_3d requires at least one search point");
    make_prepared(prepared_out, [&](Prepared& p) {
      prepare_points(p, search_points, search_count, 3);
      p.radius = radius;
    });
  });
}

int rtdl_mi355x_run_prepared_fixed_radius_neighbors_3d(void* prepared, const RtdlPoint3D* queries,
                                                       size_t query_count, uint32_t k_max,
                                                       RtdlFixedRadiusNeighborRow** rows_out,
                                                       size_t* row_count_out, char* error_out,
                                                       size_t error_size) {
  return guarded(error_out, error_size, [&] {
    zero_outputs(rows_out, row_count_out);
    require(prepared != nullptr, "prepared fixed-radius-neighbors handle must not be null");
    require(k_max > 0, "fixed_radius_neighbors k_max must be positive");
    on_prepared(prepared, [&](Prepared& p) {
      require(p.type == SceneType::Points3D, "prepared handle is not a 3-D fixed-radius scene");
      run_neighbors(p, queries, query_count, NeighborMode::FixedRadius, p.radius, k_max, (void**)rows_out,
                    row_count_out);
    });
  });
}

void rtdl_mi355x_destroy_prepared_fixed_radius_neighbors_3d(void* prepared) {
  rtdl_mi355x_destroy_prepared(prepared);
}

int rtdl_mi355x_prepare_ray_hitcount_3d(const RtdlTriangle3D* triangles, size_t triangle_count,
                                        void** prepared_out, char* error_out, size_t error_size) {
  return rtdl_mi355x_prepare_triangles_3d(triangles, triangle_count, prepared_out, error_out, error_size);
}

void rtdl_mi355x_destroy_prepared_ray_hitcount_3d(void* prepared) { rtdl_mi355x_destroy_prepared(prepared); }

int rtdl_mi355x_prepare_ray_anyhit_2d(const RtdlTriangle* triangles, size_t triangle_count,
                                      void** prepared_out, char* error_out, size_t error_size) {
  return rtdl_mi355x_prepare_triangles_2d(triangles, triangle_count, prepared_out, error_out, error_size);
}

void rtdl_mi355x_destroy_prepared_ray_anyhit_2d(void* prepared) { rtdl_mi355x_destroy_prepared(prepared); }

// ------------------------------------------------------ batched / served --

int rtdl_mi355x_prepare_point_instances(const float* coords, size_t point_count,
                                        const uint64_t* instance_offsets, size_t instance_count,
                                        void** prepared_out, char* error_out, size_t error_size) {
  return rtdl_mi355x_prepare_point_instances_leaf(coords, point_count, instance_offsets, instance_count, 0u,
                                                  prepared_out, error_out, error_size);
}

int rtdl_mi355x_prepare_point_instances_leaf(const float* coords, size_t point_count,
                                             const uint64_t* instance_offsets, size_t instance_count,
                                             uint32_t leaf_size, void** prepared_out, char* error_out,
                                             size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(instance_offsets != nullptr && instance_count >= 1, "instance offsets are required");
    std::vector<uint64_t> offs(instance_offsets, instance_offsets + instance_count + 1);
    make_prepared(prepared_out,
                  [&](Prepared& p) { prepare_point_instances(p, coords, point_count, leaf_size, offs); });
  });
}

int rtdl_mi355x_prepare_triangle_instances(const float* coords, size_t triangle_count,
                                           const uint64_t* instance_offsets, size_t instance_count,
                                           void** prepared_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(instance_offsets != nullptr && instance_count >= 1, "instance offsets are required");
    std::vector<uint64_t> offs(instance_offsets, instance_offsets + instance_count + 1);
    make_prepared(prepared_out, [&](Prepared& p) { prepare_triangle_instances(p, coords, triangle_count, offs); });
  });
}

int rtdl_mi355x_prepare_polygon_instances(const uint32_t* vertex_offsets,
                                          const uint32_t* vertex_counts, size_t polygon_count,
                                          const float* vertices_xy, size_t vertex_count,
                                          const uint64_t* instance_offsets, size_t instance_count,
                                          void** prepared_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(instance_offsets != nullptr && instance_count >= 1, "instance offsets are required");
    std::vector<uint64_t> offs(instance_offsets, instance_offsets + instance_count + 1);
    make_prepared(prepared_out, [&](Prepared& p) {
      prepare_polygon_instances(p, vertex_offsets, vertex_counts, polygon_count, vertices_xy,
                                vertex_count, offs);
    });
  });
}

static void run_batch_impl(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                           RtdlBatchJobResult* results, double* seconds_out,
                           double* family_seconds, unsigned long long* visits);

int rtdl_mi355x_run_batch(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                          RtdlBatchJobResult* results, double* seconds_out, char* error_out,
                          size_t error_size) {
  return guarded(error_out, error_size, [&] {
    run_batch_impl(jobs, job_count, morton_sort, results, seconds_out, nullptr, nullptr);
  });
}

int rtdl_mi355x_run_batch_profiled(const RtdlBatchJob* jobs, size_t job_count,
                                   uint32_t morton_sort, RtdlBatchJobResult* results,
                                   double* family_seconds, uint64_t* visits, char* error_out,
                                   size_t error_size) {
  return guarded(error_out, error_size, [&] {
    run_batch_impl(jobs, job_count, morton_sort, results, nullptr, family_seconds,
                   reinterpret_cast<unsigned long long*>(visits));
  });
}

// A validated engine job (the handles and pointers are the caller's).
static void host_job_from(const RtdlBatchJob& b, HostJob& h) {
  {
    Prepared* p = prepared_from(b.scene);
    require(b.instance < p->scene.n_instances || (p->scene.n_instances == 0 && b.instance == 0),
            "job instance out of range");
    const bool point_kind = b.kind <= kJobBoundedKnn;
    const bool ray_kind = b.kind >= kJobRayAnyHit && b.kind <= kJobRayClosest;
    require(b.kind <= kJobPipPositive, "unknown job kind");
    require(!point_kind || p->type == SceneType::Points2D || p->type == SceneType::Points3D,
            "neighbour jobs need a point scene");
    require(!ray_kind || p->type == SceneType::Triangles3D, "ray jobs need a triangle scene");
    require(b.kind != kJobPipPositive || p->type == SceneType::Polygons, "PIP jobs need a polygon scene");
    require(!ray_kind || b.query_stride == 7, "ray jobs need query_stride 7");
    require(b.kind != kJobPipPositive || b.query_stride == 2, "PIP jobs need query_stride 2");
    require(!point_kind || b.query_stride == 2 || b.query_stride == 3,
            "neighbour jobs need query_stride 2 or 3");
    h = HostJob{};
    h.scene = p->scene.view(b.instance);
    h.kind = b.kind;
    h.nq = b.query_count;
    h.queries = b.queries;
    h.qstride = b.query_stride;
    h.radius = b.radius;
    h.k = b.k;
    h.rows = b.rows;
    h.row_capacity = b.row_capacity;
    // any-hit and hit-count rows are written at their ray's index; every
    // other kind reserves rows and reports RTDL_STATUS_ROW_OVERFLOW past the
    // capacity, so a buffer may be sized by the rows expected
    const bool indexed = b.kind == kJobRayAnyHit || b.kind == kJobRayHitCount;
    require(!indexed || h.row_capacity >= b.query_count, "any-hit and hit-count jobs need one row per ray");
  }
}

static std::vector<HostJob> host_jobs_from(const RtdlBatchJob* jobs, size_t job_count) {
  require(job_count == 0 || jobs != nullptr, "jobs must not be null when job_count is nonzero");
  std::vector<HostJob> host(job_count);
  for (size_t i = 0; i < job_count; ++i) host_job_from(jobs[i], host[i]);
  return host;
}

int rtdl_mi355x_run_batch_noalloc(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                                  RtdlBatchJobResult* results, double* seconds_out, char* error_out,
                                  size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(job_count == 0 || (jobs != nullptr && results != nullptr),
            "jobs and results must not be null when job_count is nonzero");
    if (job_count > kNoAllocMaxJobs)
      throw Error("rtdl_mi355x_run_batch_noalloc takes at most " + std::to_string(kNoAllocMaxJobs) + " jobs");
    uint64_t total = 0;
    for (size_t i = 0; i < job_count; ++i) total += jobs[i].query_count;
    if (total > kNoAllocMaxQueries)
      throw Error("rtdl_mi355x_run_batch_noalloc takes at most " + std::to_string(kNoAllocMaxQueries) +
                  " queries per batch");
    std::lock_guard<std::mutex> lock(noalloc_mutex());
    const NoAllocContext ctx = noalloc_context();
    for (size_t i = 0; i < job_count; ++i) host_job_from(jobs[i], ctx.jobs[i]);
    const auto t0 = std::chrono::steady_clock::now();
    ctx.engine->enqueue(ctx.jobs, job_count, morton_sort != 0);
    RTDL_HIP(hipStreamSynchronize(ctx.engine->stream()));
    if (seconds_out)
      *seconds_out = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const JobCounters* c = ctx.engine->counters();
    for (size_t i = 0; i < job_count; ++i) {
      results[i].rows = c[i].rows;
      results[i].status = c[i].status;
      results[i].ambiguous = c[i].ambiguous;
    }
  });
}

int rtdl_mi355x_scene_export(void* prepared, RtdlSceneExport* info, void* nodes, void* prims,
                             double* exact, uint32_t* roots, uint32_t* ends, char* error_out,
                             size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(info != nullptr, "info must not be null");
    Prepared* p = prepared_from(prepared);
    std::lock_guard<std::mutex> lock(p->lock);
    const Scene& sc = p->scene;
    require(sc.kind == PrimKind::Points || sc.kind == PrimKind::Triangles,
            "scene export covers point and triangle scenes");
    RtdlSceneExport out{};
    out.node_count = sc.nodes.size();
    out.prim_count = sc.n_prims;
    out.instance_count = sc.host_roots.size();
    out.prim_float4s = sc.kind == PrimKind::Triangles ? 3u : 1u;
    out.exact_doubles = sc.fp32_exact ? 0u : (sc.kind == PrimKind::Triangles ? 9u : 3u);
    out.m = sc.m;
    *info = out;
    if (nodes) RTDL_HIP(hipMemcpy(nodes, sc.nodes.get(), sc.nodes.bytes(), hipMemcpyDeviceToHost));
    if (prims)
      RTDL_HIP(hipMemcpy(prims, sc.prims.get(), out.prim_count * out.prim_float4s * sizeof(float4),
                         hipMemcpyDeviceToHost));
    if (exact && out.exact_doubles)
      RTDL_HIP(hipMemcpy(exact, sc.exact.get(), out.prim_count * out.exact_doubles * sizeof(double),
                         hipMemcpyDeviceToHost));
    if (roots) std::copy(sc.host_roots.begin(), sc.host_roots.end(), roots);
    if (ends) std::copy(sc.host_ends.begin(), sc.host_ends.end(), ends);
  });
}

int rtdl_mi355x_scene_export_compact(void* prepared, uint32_t* has_compact, void* qnodes, float* frame,
                                     char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(has_compact != nullptr, "has_compact must not be null");
    Prepared* p = prepared_from(prepared);
    std::lock_guard<std::mutex> lock(p->lock);
    const Scene& sc = p->scene;
    *has_compact = sc.qnodes.size() ? 1u : 0u;
    if (!sc.qnodes.size()) return;
    if (qnodes) RTDL_HIP(hipMemcpy(qnodes, sc.qnodes.get(), sc.qnodes.bytes(), hipMemcpyDeviceToHost));
    if (frame) RTDL_HIP(hipMemcpy(frame, sc.qframe.get(), sc.qframe.bytes(), hipMemcpyDeviceToHost));
  });
}

int rtdl_mi355x_get_tuning(RtdlTuning* out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(out != nullptr, "out must not be null");
    const Tuning t = current_tuning();
    *out = RtdlTuning{t.point_leaf, t.triangle_leaf, t.polygon_leaf, t.blocks_per_cu, t.knn_stack, t.walk_block};
  });
}

int rtdl_mi355x_set_tuning(const RtdlTuning* tuning, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(tuning != nullptr, "tuning must not be null");
    Tuning t;
    t.point_leaf = tuning->point_leaf;
    t.triangle_leaf = tuning->triangle_leaf;
    t.polygon_leaf = tuning->polygon_leaf;
    t.blocks_per_cu = tuning->blocks_per_cu;
    t.knn_stack = tuning->knn_stack;
    t.walk_block = tuning->walk_block;
    set_tuning(t);
  });
}

namespace {
// Raises the error a rebuild left in its status bits (read after `s` passed it).
void check_rebuild(const Scene& scene, hipStream_t s) {
  const uint32_t st = rebuild_status(scene, s);
  require(!(st & kRebuildNotExact), "rebuild: a coordinate is not exact in FP32 (NaN?)");
  require(!(st & kRebuildRefitIncomplete), "rebuild: the refit did not reach the root");
  require(!(st & kRebuildRootsMoved), "rebuild: an instance root moved");
}

Prepared* rebuildable(void* handle) {
  Prepared* p = prepared_from(handle);
  require(p->instances && (p->scene.kind == PrimKind::Points || p->scene.kind == PrimKind::Triangles),
          "only point and triangle instance scenes can be rebuilt");
  return p;
}

struct BatchGraph {
  hipStream_t stream = nullptr;
  std::unique_ptr<BatchEngine> engine;
  hipGraphExec_t exec = nullptr;
  size_t job_count = 0;
  std::vector<Prepared*> rebuilt;  // scenes the graph rebuilds before the batch
  ~BatchGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    engine.reset();
    if (stream) (void)hipStreamDestroy(stream);
  }
};
}  // namespace

int rtdl_mi355x_batch_graph_create(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                                   void** graph_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(graph_out != nullptr, "graph_out must not be null");
    *graph_out = nullptr;
    const std::vector<HostJob> host = host_jobs_from(jobs, job_count);
    auto g = std::make_unique<BatchGraph>();
    RTDL_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    g->engine = std::make_unique<BatchEngine>(g->stream);
    g->exec = g->engine->capture(host, morton_sort != 0);
    g->job_count = job_count;
    *graph_out = g.release();
  });
}

int rtdl_mi355x_rebuild_instances(void* prepared, const float* coords, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    Prepared* p = rebuildable(prepared);
    std::lock_guard<std::mutex> lock(p->lock);
    require(p->scene.n_prims == 0 || coords != nullptr, "coords must not be null");
    prepare_rebuild(p->scene, p->stream);
    rebuild_scene(p->scene, coords, p->stream);
    check_rebuild(p->scene, p->stream);
    read_root_box(*p);
  });
}

int rtdl_mi355x_batch_graph_create_rebuild(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                                           void* const* scenes, const float* const* coords,
                                           size_t scene_count, void** graph_out, char* error_out,
                                           size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(graph_out != nullptr, "graph_out must not be null");
    *graph_out = nullptr;
    require(scene_count == 0 || (scenes != nullptr && coords != nullptr), "scenes and coords must not be null");
    auto g = std::make_unique<BatchGraph>();
    for (size_t i = 0; i < scene_count; ++i) {
      Prepared* p = rebuildable(scenes[i]);
      require(p->scene.n_prims == 0 || coords[i] != nullptr, "coords must not be null");
      for (Prepared* q : g->rebuilt) require(q != p, "a scene is rebuilt twice in one graph");
      std::lock_guard<std::mutex> lock(p->lock);
      prepare_rebuild(p->scene, p->stream);  // before the job views are taken: M in device memory
      g->rebuilt.push_back(p);
    }
    const std::vector<HostJob> host = host_jobs_from(jobs, job_count);
    RTDL_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    g->engine = std::make_unique<BatchEngine>(g->stream);
    g->exec = g->engine->capture(host, morton_sort != 0, [&](hipStream_t s) {
      for (size_t i = 0; i < scene_count; ++i) rebuild_scene(g->rebuilt[i]->scene, coords[i], s);
    });
    g->job_count = job_count;
    *graph_out = g.release();
  });
}

int rtdl_mi355x_batch_graph_launch(void* graph, RtdlBatchJobResult* results, double* seconds_out,
                                   char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(graph != nullptr, "graph must not be null");
    BatchGraph* g = static_cast<BatchGraph*>(graph);
    require(g->job_count == 0 || results != nullptr, "results must not be null");
    hipEvent_t start, stop;
    RTDL_HIP(hipEventCreate(&start));
    RTDL_HIP(hipEventCreate(&stop));
    RTDL_HIP(hipEventRecord(start, g->stream));
    RTDL_HIP(hipGraphLaunch(g->exec, g->stream));
    RTDL_HIP(hipEventRecord(stop, g->stream));
    RTDL_HIP(hipStreamSynchronize(g->stream));
    float ms = 0.0f;
    RTDL_HIP(hipEventElapsedTime(&ms, start, stop));
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    if (seconds_out) *seconds_out = ms * 1e-3;
    for (Prepared* p : g->rebuilt) check_rebuild(p->scene, g->stream);
    const JobCounters* c = g->engine->counters();
    for (size_t i = 0; i < g->job_count; ++i) {
      results[i].rows = c[i].rows;
      results[i].status = c[i].status;
      results[i].ambiguous = c[i].ambiguous;
    }
  });
}

int rtdl_mi355x_batch_graph_destroy(void* graph, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    BatchGraph* g = static_cast<BatchGraph*>(graph);
    if (g) RTDL_HIP(hipStreamSynchronize(g->stream));
    delete g;
  });
}

static void run_batch_impl(const RtdlBatchJob* jobs, size_t job_count, uint32_t morton_sort,
                           RtdlBatchJobResult* results, double* seconds_out,
                           double* family_seconds, unsigned long long* visits) {
  {
    require(job_count == 0 || results != nullptr, "results must not be null when job_count is nonzero");
    const std::vector<HostJob> host = host_jobs_from(jobs, job_count);
    static thread_local std::unique_ptr<BatchEngine> engine;
    hipStream_t s = thread_stream();
    if (!engine) engine = std::make_unique<BatchEngine>(s);
    engine->enable_timing(family_seconds != nullptr);
    engine->enable_stats(visits != nullptr);
    hipEvent_t start, stop;
    RTDL_HIP(hipEventCreate(&start));
    RTDL_HIP(hipEventCreate(&stop));
    RTDL_HIP(hipEventRecord(start, s));
    engine->enqueue(host, morton_sort != 0);
    RTDL_HIP(hipEventRecord(stop, s));
    RTDL_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    RTDL_HIP(hipEventElapsedTime(&ms, start, stop));
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    if (seconds_out) *seconds_out = ms * 1e-3;
    if (family_seconds) engine->family_seconds(family_seconds);
    if (visits) {
      RTDL_HIP(hipMemcpy(visits, engine->device_stats(), 3 * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
    }
    const JobCounters* c = engine->counters();
    for (size_t i = 0; i < job_count; ++i) {
      results[i].rows = c[i].rows;
      results[i].status = c[i].status;
      results[i].ambiguous = c[i].ambiguous;
    }
  }
}

int rtdl_mi355x_serve_replay(void* const* scenes, size_t scene_count,
                             const RtdlServeRequest* requests, size_t request_count,
                             const RtdlServeConfig* config, RtdlServeResult* results,
                             RtdlServeStats* stats, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(config != nullptr, "config must not be null");
    require(request_count == 0 || (requests != nullptr && results != nullptr),
            "requests and results must not be null when request_count is nonzero");
    require(scene_count == 0 || scenes != nullptr, "scenes must not be null");
    serve_replay(scenes, scene_count, requests, request_count, *config, results, stats);
  });
}

int rtdl_mi355x_selftest_fp64(const double* a, const double* b, size_t n, double* sqrt_out,
                              double* div_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(n == 0 || (a && b && sqrt_out && div_out), "null pointer with nonzero count");
    if (!n) return;
    hipStream_t s = thread_stream();
    DeviceArray<double> da(n), db(n), ds(n), dd(n);
    RTDL_HIP(hipMemcpyAsync(da.get(), a, n * sizeof(double), hipMemcpyHostToDevice, s));
    RTDL_HIP(hipMemcpyAsync(db.get(), b, n * sizeof(double), hipMemcpyHostToDevice, s));
    fp64_selftest(da.get(), db.get(), n, ds.get(), dd.get(), s);
    RTDL_HIP(hipMemcpyAsync(sqrt_out, ds.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
    RTDL_HIP(hipMemcpyAsync(div_out, dd.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
    RTDL_HIP(hipStreamSynchronize(s));
  });
}

int rtdl_mi355x_get_device(int* device_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(device_out != nullptr, "device_out must not be null");
    RTDL_HIP(hipGetDevice(device_out));
  });
}

int rtdl_mi355x_device_alloc(size_t bytes, void** ptr_out, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    require(ptr_out != nullptr, "ptr_out must not be null");
    *ptr_out = nullptr;
    if (fault_fires(FaultKind::Alloc))
      throw Error("injected fault (RTDL_MI355X_FAULT=alloc): device allocation of " + std::to_string(bytes) +
                  " bytes failed");
    void* p = nullptr;
    RTDL_HIP(hipMalloc(&p, bytes ? bytes : 1));
    *ptr_out = p;
  });
}

int rtdl_mi355x_device_free(void* ptr, char* error_out, size_t error_size) {
  return guarded(error_out, error_size, [&] {
    if (ptr) RTDL_HIP(hipFree(ptr));
  });
}

int rtdl_mi355x_copy_to_device(void* dst, const void* src, size_t bytes, char* error_out,
                               size_t error_size) {
  return guarded(error_out, error_size, [&] {
    if (!bytes) return;
    require(dst != nullptr && src != nullptr, "null pointer with nonzero byte count");
    hipStream_t s = thread_stream();
    RTDL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    RTDL_HIP(hipStreamSynchronize(s));
  });
}

int rtdl_mi355x_copy_device(void* dst, const void* src, size_t bytes, char* error_out,
                            size_t error_size) {
  return guarded(error_out, error_size, [&] {
    if (!bytes) return;
    require(dst != nullptr && src != nullptr, "null pointer with nonzero byte count");
    hipStream_t s = thread_stream();
    RTDL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    RTDL_HIP(hipStreamSynchronize(s));
  });
}

int rtdl_mi355x_copy_to_host(void* dst, const void* src, size_t bytes, char* error_out,
                             size_t error_size) {
  return guarded(error_out, error_size, [&] {
    if (!bytes) return;
    require(dst != nullptr && src != nullptr, "null pointer with nonzero byte count");
    hipStream_t s = thread_stream();
    RTDL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    RTDL_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
