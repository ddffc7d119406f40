// Host runtime shared by the C-ABI entry points: streams, phase timing,
// prepared scenes and the query pipelines (upload -> traverse -> emit).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "queries.hpp"
#include "scene.hpp"

namespace rtdl {

// Per-thread stream for the plain per-call path.
hipStream_t thread_stream();

// Phase wall times of the last call on this thread (seconds):
// 0 upload+pack, 1 build, 2 traverse, 3 emit, 4 copy-back, 5 total.
struct PhaseClock {
  std::array<double, 6> seconds{};
  std::chrono::steady_clock::time_point start, mark;
  void begin();
  void lap(int phase);  // attributes the time since the last mark to `phase`
  void end();
};
PhaseClock& thread_clock();

// Traversal measurement of the per-call and prepared calls on this thread:
// the time of the last call's traversal kernels (HIP events on the call's
// stream) and, while counting is on, the node records fetched and primitive
// records tested (counting kernels, slower).  Neighbour calls with k <= 64,
// ray calls and PIP calls record it.
struct WalkStats {
  bool counting = false;
  bool valid = false;
  double kernel_s = 0.0;
  unsigned long long nodes = 0, prims = 0, steps = 0;
};
WalkStats& thread_walk_stats();

enum class SceneType : int { Points2D = 0, Points3D = 1, Triangles3D = 2, Polygons = 3, Triangles2D = 4 };

// A prepared (device-resident) scene.  Calls on one handle are serialised by
// `lock`; each handle owns its stream.
struct Prepared {
  uint32_t magic = 0x52544D58u;  // "RTMX"
  SceneType type = SceneType::Points3D;
  Scene scene;
  std::mutex lock;
  hipStream_t stream = nullptr;
  float root_lo[3] = {0, 0, 0};
  float root_hi[3] = {0, 0, 0};
  DeviceArray<uint4> refs_input;  // polygons: refs in input order
  size_t n_polygons = 0;
  uint32_t last_ambiguous = 0;  // queries of the last call recomputed exactly
  double radius = 0.0;          // fixed radius bound at prepare time (reference-shaped FRN handle)
  bool instances = false;       // built by prepare_*_instances from FP32 device coordinates
  bool owns_stream = true;
  // own_stream = false borrows the calling thread's stream (per-call path).
  explicit Prepared(bool own_stream = true);
  ~Prepared();
};

Prepared* prepared_from(void* handle);

// Builders (raw = caller's record array on the host).
void prepare_points(Prepared& p, const void* raw, size_t n, int dim);
void prepare_triangles(Prepared& p, const void* raw, size_t n);
// RtdlTriangle records (2-D): a triangle scene with z = 0, for 2-D rays.
void prepare_triangles_2d(Prepared& p, const void* raw, size_t n);
void prepare_polygons(Prepared& p, const void* refs, size_t n, const double* verts_xy,
                      size_t n_doubles);

// Multi-instance builders from device FP32 data (ids are instance-local).
// Reads the root box of instance 0 into p.root_lo / root_hi (synchronises).
void read_root_box(Prepared& p);
// Releases a row buffer returned by a per-call or prepared entry point
// (pageable memory, or a pinned buffer kept for later calls).
void free_host_rows(void* rows);
// leaf_size: points per leaf of the instance trees, 0 for the process tuning
void prepare_point_instances(Prepared& p, const float* coords, size_t n, uint32_t leaf_size,
                             const std::vector<uint64_t>& offsets);
void prepare_triangle_instances(Prepared& p, const float* coords, size_t n,
                                const std::vector<uint64_t>& offsets);
void prepare_polygon_instances(Prepared& p, const uint32_t* voff, const uint32_t* vcnt, size_t n,
                               const float* verts_xy, size_t n_verts,
                               const std::vector<uint64_t>& offsets);

// Query pipelines; rows are malloc'ed host arrays (freed by free_rows).
enum class NeighborMode : int { FixedRadius = 0, Knn = 1, BoundedKnn = 2 };
void run_neighbors(Prepared& p, const void* raw_queries, size_t nq, NeighborMode mode,
                   double radius, uint32_t k, void** rows_out, size_t* n_out);
// RtdlRay2D records against a 2-D triangle scene; mode 0 any-hit, 1 hit
// count, 2 closest hit (rows as run_rays).
void run_rays_2d(Prepared& p, const void* raw_rays, size_t nr, int mode, void** rows_out, size_t* n_out);
void run_rays(Prepared& p, const void* raw_rays, size_t nr, int mode, void** rows_out,
              size_t* n_out);
void run_pip(Prepared& p, const void* raw_points, size_t np, bool positive_only, void** rows_out,
             size_t* n_out);

// Runs `body` and converts exceptions into the status / message convention.
template <class F>
int guarded(char* error_out, size_t error_size, F&& body) {
  if (error_out && error_size) error_out[0] = '\0';
  try {
    // the launch checks read hipGetLastError, which also reports the status
    // of any earlier call on this thread whose result was not examined (a
    // cleanup call, another library); start every entry point from a clean
    // record so that a launch is only blamed for its own error
    (void)hipGetLastError();
    body();
    return 0;
  } catch (const std::exception& e) {
    if (error_out && error_size) {
      std::snprintf(error_out, error_size, "%s", e.what());
    }
    return 1;
  } catch (...) {
    if (error_out && error_size) std::snprintf(error_out, error_size, "unknown error");
    return 1;
  }
}

}  // namespace rtdl
