/*
 * ORACLE — test infrastructure only, never a product path.
 *
 * An independent C restatement of the reference's FP64 `cpu_python_reference`
 * predicates, used (1) to check the pure-Python restatement on random inputs,
 * (2) as the large-size parity checker for the HIP backend, and (3) as the
 * multi-threaded CPU baseline timed by bench.py.  Only tests/, smoke() and
 * bench.py's cpu_baseline leg load it.
 *
 * Arithmetic follows the reference's Python operation order exactly
 * (compiled with -ffp-contract=off; IEEE binary64; sqrt correctly rounded;
 * `x ** y` in Python is libm pow(), so pow() is called where the reference
 * uses `**`).  A uniform grid is used only as a conservative candidate
 * filter: cells are at least (1 + 1e-6) times the search distance and are
 * visited with a one-cell halo, so every pair the brute force accepts is
 * still examined; `brute` = 1 disables the grid for cross-checks.
 *
 * Reference lines followed:
 *   fixed_radius_neighbors_cpu  src/rtdsl/reference.py:349-376
 *   knn_rows_cpu                src/rtdsl/reference.py:379-403
 *   bounded_knn_rows_cpu        src/rtdsl/reference.py:406-434
 *   _point_distance_sq          src/rtdsl/reference.py:882-890
 *   _finite_ray_triangle_hit_t_3d  src/rtdsl/reference.py:810-846
 *   ray any-hit / hit-count / closest-hit  src/rtdsl/reference.py:174-200,248-266
 *   2-D any-hit _finite_ray_hits_triangle_2d  src/rtdsl/reference.py:783-803
 *     (_segment_intersection :693-714, _point_in_triangle :893-915)
 *   pip_cpu / _point_in_polygon / _point_on_segment  src/rtdsl/reference.py:103-130,717-768
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint32_t qid;
  uint32_t nid;
  double dist;
  uint32_t rank;
  uint32_t pad;
} oracle_nn_row;

typedef struct {
  double d;
  uint32_t id;
} cand_t;

static int cand_less(const cand_t* a, const cand_t* b) {
  if (a->d < b->d) return 1;
  if (a->d > b->d) return 0;
  return a->id < b->id;
}

static int cand_cmp(const void* pa, const void* pb) {
  const cand_t* a = (const cand_t*)pa;
  const cand_t* b = (const cand_t*)pb;
  if (cand_less(a, b)) return -1;
  if (cand_less(b, a)) return 1;
  return 0;
}

/* ---------------------------------------------------------------- grid -- */

typedef struct {
  int dim;
  double lo[3];
  double cell;
  int64_t n[3];
  int64_t* start; /* cells+1 */
  uint32_t* items;
} grid_t;

static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

static int64_t cell_coord(const grid_t* g, int a, double x) {
  double f = floor((x - g->lo[a]) / g->cell);
  if (!(f == f)) f = 0.0;
  if (f < 0.0) f = 0.0;
  if (f > (double)(g->n[a] - 1)) f = (double)(g->n[a] - 1);
  return (int64_t)f;
}

/* Build a grid over item boxes [bmin, bmax] (per axis arrays, stride dim). */
static void grid_build(grid_t* g, int dim, size_t count, const double* bmin, const double* bmax,
                       double want_cell) {
  memset(g, 0, sizeof(*g));
  g->dim = dim;
  double hi[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a) {
    g->lo[a] = 0.0;
    hi[a] = 0.0;
  }
  if (count > 0) {
    for (int a = 0; a < dim; ++a) {
      g->lo[a] = bmin[a];
      hi[a] = bmax[a];
    }
    for (size_t i = 0; i < count; ++i) {
      for (int a = 0; a < dim; ++a) {
        if (bmin[i * dim + a] < g->lo[a]) g->lo[a] = bmin[i * dim + a];
        if (bmax[i * dim + a] > hi[a]) hi[a] = bmax[i * dim + a];
      }
    }
  }
  double extent = 0.0;
  for (int a = 0; a < dim; ++a) {
    double e = hi[a] - g->lo[a];
    if (e > extent) extent = e;
  }
  double cell = want_cell;
  if (!(cell > 0.0) || !isfinite(cell)) cell = extent > 0.0 ? extent : 1.0;
  /* cap the cell count at ~2^24 */
  for (;;) {
    double total = 1.0;
    for (int a = 0; a < dim; ++a) total *= floor((hi[a] - g->lo[a]) / cell) + 1.0;
    if (total <= 16777216.0) break;
    cell *= 1.5;
  }
  g->cell = cell;
  int64_t ncell = 1;
  for (int a = 0; a < 3; ++a) {
    g->n[a] = a < dim ? (int64_t)floor((hi[a] - g->lo[a]) / cell) + 1 : 1;
    ncell *= g->n[a];
  }
  g->start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
  /* count */
  for (size_t i = 0; i < count; ++i) {
    int64_t c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
      c0[a] = cell_coord(g, a, bmin[i * dim + a]);
      c1[a] = cell_coord(g, a, bmax[i * dim + a]);
    }
    for (int64_t z = c0[2]; z <= c1[2]; ++z)
      for (int64_t y = c0[1]; y <= c1[1]; ++y)
        for (int64_t x = c0[0]; x <= c1[0]; ++x) g->start[(z * g->n[1] + y) * g->n[0] + x + 1]++;
  }
  for (int64_t c = 0; c < ncell; ++c) g->start[c + 1] += g->start[c];
  g->items = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(g->start[ncell] > 0 ? g->start[ncell] : 1));
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
  memcpy(fill, g->start, sizeof(int64_t) * (size_t)ncell);
  for (size_t i = 0; i < count; ++i) {
    int64_t c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
      c0[a] = cell_coord(g, a, bmin[i * dim + a]);
      c1[a] = cell_coord(g, a, bmax[i * dim + a]);
    }
    for (int64_t z = c0[2]; z <= c1[2]; ++z)
      for (int64_t y = c0[1]; y <= c1[1]; ++y)
        for (int64_t x = c0[0]; x <= c1[0]; ++x)
          g->items[fill[(z * g->n[1] + y) * g->n[0] + x]++] = (uint32_t)i;
  }
  free(fill);
}

static void grid_free(grid_t* g) {
  free(g->start);
  free(g->items);
}

/* ----------------------------------------------------------- threading -- */

typedef void (*range_fn)(void* ctx, size_t begin, size_t end, int worker);

typedef struct {
  range_fn fn;
  void* ctx;
  size_t n;
  size_t chunk;
  size_t* next;
  pthread_mutex_t* lock;
  int worker;
} worker_arg;

static void* worker_main(void* p) {
  worker_arg* w = (worker_arg*)p;
  for (;;) {
    pthread_mutex_lock(w->lock);
    size_t b = *w->next;
    *w->next = b + w->chunk;
    pthread_mutex_unlock(w->lock);
    if (b >= w->n) break;
    size_t e = b + w->chunk < w->n ? b + w->chunk : w->n;
    w->fn(w->ctx, b, e, w->worker);
  }
  return NULL;
}

static void parallel_for(size_t n, int threads, size_t chunk, range_fn fn, void* ctx) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  size_t next = 0;
  pthread_mutex_t lock;
  pthread_mutex_init(&lock, NULL);
  pthread_t tid[256];
  worker_arg args[256];
  for (int t = 0; t < threads; ++t) {
    args[t].fn = fn;
    args[t].ctx = ctx;
    args[t].n = n;
    args[t].chunk = chunk;
    args[t].next = &next;
    args[t].lock = &lock;
    args[t].worker = t;
  }
  if (threads == 1) {
    worker_main(&args[0]);
  } else {
    for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, worker_main, &args[t]);
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  }
  pthread_mutex_destroy(&lock);
}

/* ------------------------------------------------- neighbour queries -- */

typedef struct {
  int dim;
  size_t nq, ns;
  const double *qx, *qy, *qz;
  const uint32_t* qid;
  const double *sx, *sy, *sz;
  const uint32_t* sid;
  double radius; /* <0 => unbounded (knn) */
  size_t k;
  int brute;
  grid_t grid;
  /* per-query outputs */
  cand_t** per_query;
  uint32_t* per_count;
} nn_ctx;

static double dist_sq(const nn_ctx* c, size_t q, size_t s) {
  /* reference: src/rtdsl/reference.py:882-890 — dx = right - left, z = 0.0 in 2-D */
  double dx = c->sx[s] - c->qx[q];
  double dy = c->sy[s] - c->qy[q];
  double dz = (c->dim == 3 ? c->sz[s] : 0.0) - (c->dim == 3 ? c->qz[q] : 0.0);
  return dx * dx + dy * dy + dz * dz;
}

typedef struct {
  cand_t* v;
  size_t n, cap;
} cvec;

static void cvec_push(cvec* v, cand_t c) {
  if (v->n == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 64;
    v->v = (cand_t*)realloc(v->v, v->cap * sizeof(cand_t));
  }
  v->v[v->n++] = c;
}

/* keep a sorted top-k list (insertion); returns nothing */
static void topk_insert(cand_t* list, size_t* count, size_t k, cand_t c) {
  if (*count == k) {
    if (!cand_less(&c, &list[k - 1])) return;
    (*count)--;
  }
  size_t i = *count;
  while (i > 0 && cand_less(&c, &list[i - 1])) {
    list[i] = list[i - 1];
    --i;
  }
  list[i] = c;
  (*count)++;
}

static void nn_radius_range(void* p, size_t b, size_t e, int worker) {
  (void)worker;
  nn_ctx* c = (nn_ctx*)p;
  double r2 = c->radius * c->radius;
  for (size_t q = b; q < e; ++q) {
    cvec found = {0, 0, 0};
    if (c->brute || c->ns == 0) {
      for (size_t s = 0; s < c->ns; ++s) {
        double d2 = dist_sq(c, q, s);
        if (d2 > r2) continue;
        cand_t cd = {sqrt(d2), c->sid[s]};
        cvec_push(&found, cd);
      }
    } else {
      const grid_t* g = &c->grid;
      double qc[3] = {c->qx[q], c->qy[q], c->dim == 3 ? c->qz[q] : 0.0};
      int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
      for (int a = 0; a < c->dim; ++a) {
        int64_t cc = cell_coord(g, a, qc[a]);
        lo[a] = clampi(cc - 1, 0, g->n[a] - 1);
        hi[a] = clampi(cc + 1, 0, g->n[a] - 1);
      }
      for (int64_t z = lo[2]; z <= hi[2]; ++z)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
          for (int64_t x = lo[0]; x <= hi[0]; ++x) {
            int64_t cell = (z * g->n[1] + y) * g->n[0] + x;
            for (int64_t it = g->start[cell]; it < g->start[cell + 1]; ++it) {
              uint32_t s = g->items[it];
              double d2 = dist_sq(c, q, s);
              if (d2 > r2) continue;
              cand_t cd = {sqrt(d2), c->sid[s]};
              cvec_push(&found, cd);
            }
          }
    }
    qsort(found.v, found.n, sizeof(cand_t), cand_cmp);
    size_t keep = found.n < c->k ? found.n : c->k;
    c->per_query[q] = (cand_t*)malloc(sizeof(cand_t) * (keep ? keep : 1));
    if (keep) memcpy(c->per_query[q], found.v, keep * sizeof(cand_t));
    c->per_count[q] = (uint32_t)keep;
    free(found.v);
  }
}

static void nn_knn_range(void* p, size_t b, size_t e, int worker) {
  (void)worker;
  nn_ctx* c = (nn_ctx*)p;
  size_t k = c->k < c->ns ? c->k : c->ns;
  for (size_t q = b; q < e; ++q) {
    cand_t* list = (cand_t*)malloc(sizeof(cand_t) * (k ? k : 1));
    size_t count = 0;
    if (c->brute || k == 0) {
      for (size_t s = 0; s < c->ns && k; ++s) {
        cand_t cd = {sqrt(dist_sq(c, q, s)), c->sid[s]};
        topk_insert(list, &count, k, cd);
      }
    } else {
      const grid_t* g = &c->grid;
      double qc[3] = {c->qx[q], c->qy[q], c->dim == 3 ? c->qz[q] : 0.0};
      int64_t cc[3] = {0, 0, 0};
      for (int a = 0; a < c->dim; ++a) cc[a] = cell_coord(g, a, qc[a]);
      int64_t maxring = 0;
      for (int a = 0; a < c->dim; ++a) {
        int64_t m = cc[a] > g->n[a] - 1 - cc[a] ? cc[a] : g->n[a] - 1 - cc[a];
        if (m > maxring) maxring = m;
      }
      /* Ring R covers cells at Chebyshev distance R.  Every point in a cell
       * at ring >= R+1 lies at least R*cell away (the query sits inside its
       * own cell; a 1e-9 relative slack absorbs floor() rounding). */
      for (int64_t R = 0; R <= maxring; ++R) {
        int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int a = 0; a < c->dim; ++a) {
          lo[a] = cc[a] - R;
          hi[a] = cc[a] + R;
        }
        for (int64_t z = lo[2]; z <= hi[2]; ++z) {
          if (z < 0 || z >= g->n[2]) continue;
          for (int64_t y = lo[1]; y <= hi[1]; ++y) {
            if (y < 0 || y >= g->n[1]) continue;
            for (int64_t x = lo[0]; x <= hi[0]; ++x) {
              if (x < 0 || x >= g->n[0]) continue;
              int on_ring = (x == lo[0] || x == hi[0]);
              if (c->dim >= 2) on_ring |= (y == lo[1] || y == hi[1]);
              if (c->dim == 3) on_ring |= (z == lo[2] || z == hi[2]);
              if (!on_ring && R > 0) continue;
              int64_t cell = (z * g->n[1] + y) * g->n[0] + x;
              for (int64_t it = g->start[cell]; it < g->start[cell + 1]; ++it) {
                uint32_t s = g->items[it];
                cand_t cd = {sqrt(dist_sq(c, q, s)), c->sid[s]};
                topk_insert(list, &count, k, cd);
              }
            }
          }
        }
        if (count == k && list[k - 1].d < (double)R * g->cell * (1.0 - 1e-9)) break;
      }
    }
    c->per_query[q] = list;
    c->per_count[q] = (uint32_t)count;
  }
}

/* Returns a malloc'ed row array (caller frees with oracle_free).  Rows are
 * grouped per query in input order, each group ordered by (distance, id);
 * the caller applies the final stable sort by query id.
 * mode: 0 = fixed-radius (k = k_max), 1 = knn (radius ignored), 2 = bounded knn. */
int64_t oracle_neighbors(int mode, int dim, size_t nq, const double* qx, const double* qy,
                         const double* qz, const uint32_t* qid, size_t ns, const double* sx,
                         const double* sy, const double* sz, const uint32_t* sid, double radius,
                         size_t k, int threads, int brute, oracle_nn_row** rows_out) {
  nn_ctx c;
  memset(&c, 0, sizeof(c));
  c.dim = dim;
  c.nq = nq;
  c.ns = ns;
  c.qx = qx;
  c.qy = qy;
  c.qz = qz;
  c.qid = qid;
  c.sx = sx;
  c.sy = sy;
  c.sz = sz;
  c.sid = sid;
  c.radius = radius;
  c.k = k;
  c.brute = brute;
  c.per_query = (cand_t**)calloc(nq ? nq : 1, sizeof(cand_t*));
  c.per_count = (uint32_t*)calloc(nq ? nq : 1, sizeof(uint32_t));
  if (!brute && ns > 0) {
    double* pts = (double*)malloc(sizeof(double) * ns * dim);
    for (size_t s = 0; s < ns; ++s) {
      pts[s * dim + 0] = sx[s];
      pts[s * dim + 1] = sy[s];
      if (dim == 3) pts[s * dim + 2] = sz[s];
    }
    double want;
    if (mode == 1) {
      /* knn: about 4 points per cell on average */
      double lo[3], hi[3];
      for (int a = 0; a < dim; ++a) {
        lo[a] = pts[a];
        hi[a] = pts[a];
      }
      for (size_t s = 0; s < ns; ++s)
        for (int a = 0; a < dim; ++a) {
          if (pts[s * dim + a] < lo[a]) lo[a] = pts[s * dim + a];
          if (pts[s * dim + a] > hi[a]) hi[a] = pts[s * dim + a];
        }
      double vol = 1.0;
      double ext = 0.0;
      for (int a = 0; a < dim; ++a) {
        double e = hi[a] - lo[a];
        if (e > ext) ext = e;
        vol *= e > 0 ? e : 1.0;
      }
      want = pow(vol * 4.0 / (double)ns, 1.0 / (double)dim);
      if (!(want > 0.0)) want = ext > 0 ? ext : 1.0;
    } else {
      want = radius * (1.0 + 1e-6) + 1e-300;
    }
    grid_build(&c.grid, dim, ns, pts, pts, want);
    free(pts);
  }
  parallel_for(nq, threads, 64, mode == 1 ? nn_knn_range : nn_radius_range, &c);
  int64_t total = 0;
  for (size_t q = 0; q < nq; ++q) total += c.per_count[q];
  oracle_nn_row* rows = (oracle_nn_row*)malloc(sizeof(oracle_nn_row) * (size_t)(total ? total : 1));
  int64_t w = 0;
  for (size_t q = 0; q < nq; ++q) {
    for (uint32_t i = 0; i < c.per_count[q]; ++i) {
      rows[w].qid = qid[q];
      rows[w].nid = c.per_query[q][i].id;
      rows[w].dist = c.per_query[q][i].d;
      rows[w].rank = i + 1;
      rows[w].pad = 0;
      ++w;
    }
    free(c.per_query[q]);
  }
  free(c.per_query);
  free(c.per_count);
  if (!brute && ns > 0) grid_free(&c.grid);
  *rows_out = rows;
  return total;
}

void oracle_free(void* p) { free(p); }

/* ---------------------------------------------------------- rays (3-D) -- */

/* reference: src/rtdsl/reference.py:810-846; returns 1 and *t_out on a hit. */
static int hit_t_3d(const double* r, const double* tri, double* t_out) {
  double e1x = tri[3] - tri[0];
  double e1y = tri[4] - tri[1];
  double e1z = tri[5] - tri[2];
  double e2x = tri[6] - tri[0];
  double e2y = tri[7] - tri[1];
  double e2z = tri[8] - tri[2];
  double ox = r[0], oy = r[1], oz = r[2], dx = r[3], dy = r[4], dz = r[5], tmax = r[6];
  double px = dy * e2z - dz * e2y;
  double py = dz * e2x - dx * e2z;
  double pz = dx * e2y - dy * e2x;
  double det = e1x * px + e1y * py + e1z * pz;
  if (fabs(det) <= 1.0e-8) return 0;
  double inv = 1.0 / det;
  double tx = ox - tri[0];
  double ty = oy - tri[1];
  double tz = oz - tri[2];
  double u = (tx * px + ty * py + tz * pz) * inv;
  if (u < 0.0 || u > 1.0) return 0;
  double qx = ty * e1z - tz * e1y;
  double qy = tz * e1x - tx * e1z;
  double qz = tx * e1y - ty * e1x;
  double v = (dx * qx + dy * qy + dz * qz) * inv;
  if (v < 0.0 || (u + v) > 1.0) return 0;
  double t = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (t >= 0.0 && t <= tmax) {
    *t_out = t;
    return 1;
  }
  return 0;
}

/* ------------------------------------------------------- rays (2-D any) -- */

/* reference: src/rtdsl/reference.py:693-714 (left = p -> e, right = q0 -> q1) */
static int segments_cross_2d(double px, double py, double ex, double ey, double qx0, double qy0,
                             double qx1, double qy1) {
  double rx = ex - px;
  double ry = ey - py;
  double sx = qx1 - qx0;
  double sy = qy1 - qy0;
  double denom = rx * sy - ry * sx;
  if (fabs(denom) < 1.0e-7) return 0;
  double qpx = qx0 - px;
  double qpy = qy0 - py;
  double t = (qpx * sy - qpy * sx) / denom;
  double u = (qpx * ry - qpy * rx) / denom;
  return 0.0 <= t && t <= 1.0 && 0.0 <= u && u <= 1.0;
}

/* reference: src/rtdsl/reference.py:893-915 */
static int point_in_triangle_2d(double x, double y, const double* v) {
  double v0x = v[6] - v[0];
  double v0y = v[7] - v[1];
  double v1x = v[3] - v[0];
  double v1y = v[4] - v[1];
  double v2x = x - v[0];
  double v2y = y - v[1];
  double dot00 = v0x * v0x + v0y * v0y;
  double dot01 = v0x * v1x + v0y * v1y;
  double dot02 = v0x * v2x + v0y * v2y;
  double dot11 = v1x * v1x + v1y * v1y;
  double dot12 = v1x * v2x + v1y * v2y;
  double denom = dot00 * dot11 - dot01 * dot01;
  if (fabs(denom) < 1.0e-7) return 0;
  double inv = 1.0 / denom;
  double u = (dot11 * dot02 - dot01 * dot12) * inv;
  double w = (dot00 * dot12 - dot01 * dot02) * inv;
  return u >= 0.0 && w >= 0.0 && (u + w) <= 1.0;
}

/* reference: src/rtdsl/reference.py:783-803; r = ox oy - dx dy - tmax, tri = x y - (x3) */
static int hits_2d(const double* r, const double* v) {
  double ox = r[0], oy = r[1];
  double ex = ox + r[3] * r[6];
  double ey = oy + r[4] * r[6];
  if (point_in_triangle_2d(ox, oy, v)) return 1;
  if (point_in_triangle_2d(ex, ey, v)) return 1;
  return segments_cross_2d(ox, oy, ex, ey, v[0], v[1], v[3], v[4]) ||
         segments_cross_2d(ox, oy, ex, ey, v[3], v[4], v[6], v[7]) ||
         segments_cross_2d(ox, oy, ex, ey, v[6], v[7], v[0], v[1]);
}

typedef struct {
  size_t nr, nt;
  const double* rays;  /* nr x 7: ox oy oz dx dy dz tmax */
  const double* tris;  /* nt x 9 */
  const uint32_t* tid; /* nt */
  int mode;            /* 0 any, 1 count, 2 closest; 3, 4, 5: the same in 2-D */
  int brute;
  grid_t grid;
  double pad;
  uint32_t** stamp; /* per worker, nt */
  uint32_t* out_u;  /* any / count / closest triangle id */
  double* out_t;    /* closest t */
  uint8_t* out_has; /* closest: hit present */
} ray_ctx;

/* Conservative segment-vs-box test in FP64 with a padded box. */
static int seg_box(const double* r, const double* lo, const double* hi, double pad) {
  double t0 = 0.0, t1 = r[6];
  if (!(t1 >= 0.0)) return 0;
  for (int a = 0; a < 3; ++a) {
    double o = r[a], d = r[3 + a];
    double l = lo[a] - pad, h = hi[a] + pad;
    if (d == 0.0) {
      if (o < l || o > h) return 0;
      continue;
    }
    double ta = (l - o) / d, tb = (h - o) / d;
    if (ta > tb) {
      double s = ta;
      ta = tb;
      tb = s;
    }
    ta -= fabs(ta) * 1e-9 + 1e-300;
    tb += fabs(tb) * 1e-9 + 1e-300;
    if (ta > t0) t0 = ta;
    if (tb < t1) t1 = tb;
    if (t0 > t1) return 0;
  }
  return 1;
}

static void ray_range(void* p, size_t b, size_t e, int worker) {
  ray_ctx* c = (ray_ctx*)p;
  uint32_t* stamp = c->stamp ? c->stamp[worker] : NULL;
  for (size_t i = b; i < e; ++i) {
    const double* r = c->rays + i * 7;
    /* 2-D: the grid walks the segment o -> o + d * tmax as s in [0, 1] */
    double seg[7];
    int brute = c->brute;
    const double* rg = r;
    if (c->mode >= 3) {
      seg[0] = r[0];
      seg[1] = r[1];
      seg[2] = 0.0;
      seg[3] = r[3] * r[6];
      seg[4] = r[4] * r[6];
      seg[5] = 0.0;
      seg[6] = 1.0;
      rg = seg;
      if (!isfinite(seg[3]) || !isfinite(seg[4])) brute = 1;
    }
    uint32_t count = 0;
    int any = 0;
    int has = 0;
    double best_t = 0.0;
    uint32_t best_id = 0;
#define VISIT(TRI)                                                                        \
  do {                                                                                    \
    double t_ = 0.0;                                                                      \
    const double* v_ = c->tris + (size_t)(TRI) * 9;                                       \
    if (c->mode >= 3 ? hits_2d(r, v_) : hit_t_3d(r, v_, &t_)) {                           \
      any = 1;                                                                            \
      ++count;                                                                            \
      uint32_t id_ = c->tid[TRI];                                                         \
      if (!has || t_ < best_t || (t_ == best_t && id_ < best_id)) {                       \
        has = 1;                                                                          \
        best_t = t_;                                                                      \
        best_id = id_;                                                                    \
      }                                                                                   \
    }                                                                                     \
  } while (0)
    if (brute) {
      for (size_t t = 0; t < c->nt; ++t) {
        VISIT(t);
        if (any && (c->mode == 0 || c->mode == 3)) break;
      }
    } else if (c->nt > 0) {
      const grid_t* g = &c->grid;
      double s0[3], s1[3];
      for (int a = 0; a < 3; ++a) {
        double p0 = rg[a], p1 = rg[a] + rg[3 + a] * rg[6];
        s0[a] = p0 < p1 ? p0 : p1;
        s1[a] = p0 < p1 ? p1 : p0;
        if (!isfinite(s0[a])) s0[a] = -INFINITY;
        if (!isfinite(s1[a])) s1[a] = INFINITY;
      }
      int64_t lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        double l = s0[a] - c->pad, h = s1[a] + c->pad;
        lo[a] = (l == -INFINITY) ? 0 : cell_coord(g, a, l);
        hi[a] = (h == INFINITY) ? g->n[a] - 1 : cell_coord(g, a, h);
      }
      int done = 0;
      for (int64_t z = lo[2]; z <= hi[2] && !done; ++z)
        for (int64_t y = lo[1]; y <= hi[1] && !done; ++y)
          for (int64_t x = lo[0]; x <= hi[0] && !done; ++x) {
            double clo[3] = {g->lo[0] + x * g->cell, g->lo[1] + y * g->cell, g->lo[2] + z * g->cell};
            double chi[3] = {clo[0] + g->cell, clo[1] + g->cell, clo[2] + g->cell};
            if (!seg_box(rg, clo, chi, c->pad + g->cell * 1e-9)) continue;
            int64_t cell = (z * g->n[1] + y) * g->n[0] + x;
            for (int64_t it = g->start[cell]; it < g->start[cell + 1]; ++it) {
              uint32_t t = g->items[it];
              if (stamp) {
                if (stamp[t] == (uint32_t)(i + 1)) continue;
                stamp[t] = (uint32_t)(i + 1);
              }
              VISIT(t);
              if (any && (c->mode == 0 || c->mode == 3)) {
                done = 1;
                break;
              }
            }
          }
    }
#undef VISIT
    if (c->mode == 0 || c->mode == 3)
      c->out_u[i] = (uint32_t)any;
    else if (c->mode == 1 || c->mode == 4)
      c->out_u[i] = count;
    else {
      c->out_u[i] = best_id;
      c->out_t[i] = best_t;
      c->out_has[i] = (uint8_t)has;
    }
  }
}

/* mode 0: out_u = any_hit; 1: out_u = hit_count; 2: closest (out_u id, out_t, out_has);
 * 3, 4, 5: the same for 2-D (rays ox oy 0 dx dy 0 tmax, triangles with z = 0;
 * a 2-D hit has t = 0.0, reference.py:775-780). */
int oracle_rays_3d(int mode, size_t nr, const double* rays, size_t nt, const double* tris,
                   const uint32_t* tid, int threads, int brute, uint32_t* out_u, double* out_t,
                   uint8_t* out_has) {
  ray_ctx c;
  memset(&c, 0, sizeof(c));
  c.nr = nr;
  c.nt = nt;
  c.rays = rays;
  c.tris = tris;
  c.tid = tid;
  c.mode = mode;
  c.brute = brute;
  c.out_u = out_u;
  c.out_t = out_t;
  c.out_has = out_has;
  if (threads < 1) threads = 1;
  if (!brute && nt > 0) {
    double* bmin = (double*)malloc(sizeof(double) * nt * 3);
    double* bmax = (double*)malloc(sizeof(double) * nt * 3);
    double ext = 0.0;
    for (size_t t = 0; t < nt; ++t)
      for (int a = 0; a < 3; ++a) {
        double v0 = tris[t * 9 + a], v1 = tris[t * 9 + 3 + a], v2 = tris[t * 9 + 6 + a];
        double l = v0 < v1 ? v0 : v1;
        l = l < v2 ? l : v2;
        double h = v0 > v1 ? v0 : v1;
        h = h > v2 ? h : v2;
        bmin[t * 3 + a] = l;
        bmax[t * 3 + a] = h;
        if (fabs(l) > ext) ext = fabs(l);
        if (fabs(h) > ext) ext = fabs(h);
      }
    /* pad every box: the FP64 refine can accept a point marginally outside */
    c.pad = ext * 1e-9 + 1e-300;
    for (size_t t = 0; t < nt * 3; ++t) {
      bmin[t] -= c.pad;
      bmax[t] += c.pad;
    }
    double vol = 1.0, lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = bmin[a];
      hi[a] = bmax[a];
    }
    for (size_t t = 0; t < nt; ++t)
      for (int a = 0; a < 3; ++a) {
        if (bmin[t * 3 + a] < lo[a]) lo[a] = bmin[t * 3 + a];
        if (bmax[t * 3 + a] > hi[a]) hi[a] = bmax[t * 3 + a];
      }
    double span = 0.0;
    for (int a = 0; a < 3; ++a) {
      vol *= (hi[a] - lo[a]) > 0 ? (hi[a] - lo[a]) : 1.0;
      if (hi[a] - lo[a] > span) span = hi[a] - lo[a];
    }
    /* about 16 triangles per cell, and at most ~48 cells per axis: a long
     * segment enumerates every cell of its bounding box */
    double want = pow(vol * 16.0 / (double)nt, 1.0 / 3.0);
    if (want < span / 48.0) want = span / 48.0;
    grid_build(&c.grid, 3, nt, bmin, bmax, want);
    free(bmin);
    free(bmax);
    if (mode != 0 && mode != 3) {
      c.stamp = (uint32_t**)calloc((size_t)threads, sizeof(uint32_t*));
      for (int w = 0; w < threads; ++w) c.stamp[w] = (uint32_t*)calloc(nt, sizeof(uint32_t));
    }
  }
  parallel_for(nr, threads, 16, ray_range, &c);
  if (c.stamp) {
    for (int w = 0; w < threads; ++w) free(c.stamp[w]);
    free(c.stamp);
  }
  if (!brute && nt > 0) grid_free(&c.grid);
  return 0;
}

/* ------------------------------------------------------------------ PIP -- */

/* reference: src/rtdsl/reference.py:745-768 */
static int point_on_segment(double px, double py, double ax, double ay, double bx, double by) {
  const double eps = 1.0e-12;
  double length_sq = pow(bx - ax, 2.0) + pow(by - ay, 2.0);
  if (length_sq <= pow(eps, 2.0)) return fabs(px - ax) <= eps && fabs(py - ay) <= eps;
  double cross = (px - ax) * (by - ay) - (py - ay) * (bx - ax);
  double length = pow(length_sq, 0.5);
  if (fabs(cross) > eps * length) return 0;
  double dot = (px - ax) * (bx - ax) + (py - ay) * (by - ay);
  double along_eps = eps * length;
  if (dot < -along_eps) return 0;
  if (dot - length_sq > along_eps) return 0;
  return 1;
}

/* reference: src/rtdsl/reference.py:717-742 */
static int point_in_polygon(double x, double y, const double* v, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t j = (i + 1) % n;
    if (point_on_segment(x, y, v[2 * i], v[2 * i + 1], v[2 * j], v[2 * j + 1])) return 1;
  }
  int inside = 0;
  uint32_t j = n - 1;
  for (uint32_t i = 0; i < n; ++i) {
    double xi = v[2 * i], yi = v[2 * i + 1], xj = v[2 * j], yj = v[2 * j + 1];
    if ((yi > y) != (yj > y)) {
      double den = (yj - yi);
      if (den == 0.0) den = 1.0e-20;
      if (x <= (xj - xi) * (y - yi) / den + xi) inside = !inside;
    }
    j = i;
  }
  return inside;
}

typedef struct {
  size_t np, ngon;
  const double *px, *py;
  const uint32_t* off;
  const uint32_t* cnt;
  const double* verts;
  const double* bbox; /* ngon x 4 */
  int positive_only;
  int brute;
  grid_t grid;
  uint32_t** hits; /* per point: polygon indices (ascending) */
  uint32_t* nhits;
} pip_ctx;

static int u32cmp(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static void pip_range(void* p, size_t b, size_t e, int worker) {
  (void)worker;
  pip_ctx* c = (pip_ctx*)p;
  for (size_t i = b; i < e; ++i) {
    double x = c->px[i], y = c->py[i];
    uint32_t cap = 8, n = 0;
    uint32_t* out = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    const uint32_t* cand = NULL;
    int64_t ncand = 0;
    int64_t cell = 0;
    if (!c->brute && c->ngon > 0) {
      int64_t cx = cell_coord(&c->grid, 0, x), cy = cell_coord(&c->grid, 1, y);
      cell = cy * c->grid.n[0] + cx;
      cand = c->grid.items + c->grid.start[cell];
      ncand = c->grid.start[cell + 1] - c->grid.start[cell];
    } else {
      ncand = (int64_t)c->ngon;
    }
    for (int64_t k = 0; k < ncand; ++k) {
      uint32_t g = cand ? cand[k] : (uint32_t)k;
      const double* bb = c->bbox + (size_t)g * 4;
      if (c->positive_only) {
        if (x < bb[0] || x > bb[2] || y < bb[1] || y > bb[3]) continue;
      }
      if (point_in_polygon(x, y, c->verts + 2 * (size_t)c->off[g], c->cnt[g])) {
        if (n == cap) {
          cap *= 2;
          out = (uint32_t*)realloc(out, sizeof(uint32_t) * cap);
        }
        out[n++] = g;
      }
    }
    qsort(out, n, sizeof(uint32_t), u32cmp);
    c->hits[i] = out;
    c->nhits[i] = n;
  }
}

/* Returns a malloc'ed (point index, polygon index) pair list of contains=1
 * pairs in (point input order, polygon input order).  With positive_only=0
 * the bbox prefilter of positive_hits mode is not applied, matching
 * full_matrix semantics; the grid then uses boxes padded by the on-edge
 * tolerance so that no accepted pair is missed. */
int64_t oracle_pip(size_t np, const double* px, const double* py, size_t ngon, const uint32_t* off,
                   const uint32_t* cnt, const double* verts, int positive_only, int threads, int brute,
                   uint32_t** pairs_out) {
  pip_ctx c;
  memset(&c, 0, sizeof(c));
  c.np = np;
  c.ngon = ngon;
  c.px = px;
  c.py = py;
  c.off = off;
  c.cnt = cnt;
  c.verts = verts;
  c.positive_only = positive_only;
  c.brute = brute;
  double* bbox = (double*)malloc(sizeof(double) * (ngon ? ngon : 1) * 4);
  double* bmin = (double*)malloc(sizeof(double) * (ngon ? ngon : 1) * 2);
  double* bmax = (double*)malloc(sizeof(double) * (ngon ? ngon : 1) * 2);
  double area = 0.0;
  for (size_t g = 0; g < ngon; ++g) {
    const double* v = verts + 2 * (size_t)off[g];
    double x0 = v[0], y0 = v[1], x1 = v[0], y1 = v[1], maxlen = 0.0;
    for (uint32_t i = 0; i < cnt[g]; ++i) {
      double x = v[2 * i], y = v[2 * i + 1];
      if (x < x0) x0 = x;
      if (x > x1) x1 = x;
      if (y < y0) y0 = y;
      if (y > y1) y1 = y;
    }
    maxlen = (x1 - x0) + (y1 - y0);
    bbox[g * 4 + 0] = x0;
    bbox[g * 4 + 1] = y0;
    bbox[g * 4 + 2] = x1;
    bbox[g * 4 + 3] = y1;
    double pad = positive_only ? 0.0 : (1e-12 * (maxlen + 1.0) + 1e-9 * (fabs(x0) + fabs(y0) + fabs(x1) + fabs(y1)));
    bmin[g * 2 + 0] = x0 - pad;
    bmin[g * 2 + 1] = y0 - pad;
    bmax[g * 2 + 0] = x1 + pad;
    bmax[g * 2 + 1] = y1 + pad;
    area += (x1 - x0 + 1e-12) * (y1 - y0 + 1e-12);
  }
  c.bbox = bbox;
  if (!brute && ngon > 0) grid_build(&c.grid, 2, ngon, bmin, bmax, sqrt(area / (double)ngon) * 1.5);
  c.hits = (uint32_t**)calloc(np ? np : 1, sizeof(uint32_t*));
  c.nhits = (uint32_t*)calloc(np ? np : 1, sizeof(uint32_t));
  parallel_for(np, threads, 64, pip_range, &c);
  int64_t total = 0;
  for (size_t i = 0; i < np; ++i) total += c.nhits[i];
  uint32_t* pairs = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)(total ? total : 1));
  int64_t w = 0;
  for (size_t i = 0; i < np; ++i) {
    for (uint32_t k = 0; k < c.nhits[i]; ++k) {
      pairs[2 * w] = (uint32_t)i;
      pairs[2 * w + 1] = c.hits[i][k];
      ++w;
    }
    free(c.hits[i]);
  }
  free(c.hits);
  free(c.nhits);
  free(bbox);
  free(bmin);
  free(bmax);
  if (!brute && ngon > 0) grid_free(&c.grid);
  *pairs_out = pairs;
  return total;
}
