"""ORACLE — test infrastructure only, never a product path.

A clean-room, single-threaded, pure-Python FP64 restatement of the
reference's ``cpu_python_reference`` backend for the hot-path predicates.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker / CPU baseline.

Every function names the reference lines it follows.  The arithmetic is
written in the reference's operation order (Python floats are IEEE binary64,
``math.sqrt`` is correctly rounded, ``x ** y`` is libm ``pow``), so results are
bit-identical to the reference on the same inputs.

Pins (see tests/test_oracle_pins.py): the reference's authored cases
(goal198/goal204/goal548/goal552/goal632 tests), closed forms (ordering,
inclusive boundaries, truncation, no padding), textbook special cases of
Moller-Trumbore and even-odd crossing, and brute-force cross-checks against the
independent C restatement in ``oracle/rtdl_oracle.c``.
"""
from __future__ import annotations

import math
import os
import sys

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)

from rtdl_amd.records import (  # noqa: E402
    Point3D,
    Polygon,
    Ray2D,
    Ray3D,
    Triangle3D,
    normalize_records,
    project_rows,
)


# --------------------------------------------------------------------------
# fixed-radius neighbours / KNN (reference: src/rtdsl/reference.py:349-434)
# --------------------------------------------------------------------------


def _distance_sq(left, right) -> float:
    # reference: src/rtdsl/reference.py:882-890 (2-D records use z = 0.0)
    dx = right.x - left.x
    dy = right.y - left.y
    dz = getattr(right, "z", 0.0) - getattr(left, "z", 0.0)
    return dx * dx + dy * dy + dz * dz


def fixed_radius_neighbors_cpu(query_points, search_points, *, radius: float, k_max: int):
    # reference: src/rtdsl/reference.py:349-376
    rows = []
    radius_sq = radius * radius
    for query in query_points:
        found = []
        for point in search_points:
            d2 = _distance_sq(query, point)
            if d2 > radius_sq:
                continue
            found.append((math.sqrt(d2), point.id))
        found.sort(key=lambda item: (item[0], item[1]))
        for distance, neighbor_id in found[:k_max]:
            rows.append({"query_id": query.id, "neighbor_id": neighbor_id, "distance": distance})
    rows.sort(key=lambda row: row["query_id"])
    return tuple(rows)


def knn_rows_cpu(query_points, search_points, *, k: int):
    # reference: src/rtdsl/reference.py:379-403
    rows = []
    for query in query_points:
        found = [(math.sqrt(_distance_sq(query, point)), point.id) for point in search_points]
        found.sort(key=lambda item: (item[0], item[1]))
        for rank, (distance, neighbor_id) in enumerate(found[:k], start=1):
            rows.append(
                {
                    "query_id": query.id,
                    "neighbor_id": neighbor_id,
                    "distance": distance,
                    "neighbor_rank": rank,
                }
            )
    rows.sort(key=lambda row: row["query_id"])
    return tuple(rows)


def bounded_knn_rows_cpu(query_points, search_points, *, radius: float, k_max: int):
    # reference: src/rtdsl/reference.py:406-434
    rows = []
    radius_sq = radius * radius
    for query in query_points:
        found = []
        for point in search_points:
            d2 = _distance_sq(query, point)
            if d2 > radius_sq:
                continue
            found.append((math.sqrt(d2), point.id))
        found.sort(key=lambda item: (item[0], item[1]))
        for rank, (distance, neighbor_id) in enumerate(found[:k_max], start=1):
            rows.append(
                {
                    "query_id": query.id,
                    "neighbor_id": neighbor_id,
                    "distance": distance,
                    "neighbor_rank": rank,
                }
            )
    rows.sort(key=lambda row: row["query_id"])
    return tuple(rows)


# --------------------------------------------------------------------------
# ray / triangle (reference: src/rtdsl/reference.py:174-200, 248-266, 771-915)
# --------------------------------------------------------------------------


def _hit_t_3d(ray, tri):
    # reference: src/rtdsl/reference.py:810-846 (Moller-Trumbore, this op order)
    e1x = tri.x1 - tri.x0
    e1y = tri.y1 - tri.y0
    e1z = tri.z1 - tri.z0
    e2x = tri.x2 - tri.x0
    e2y = tri.y2 - tri.y0
    e2z = tri.z2 - tri.z0
    px = ray.dy * e2z - ray.dz * e2y
    py = ray.dz * e2x - ray.dx * e2z
    pz = ray.dx * e2y - ray.dy * e2x
    det = e1x * px + e1y * py + e1z * pz
    if abs(det) <= 1.0e-8:
        return None
    inv = 1.0 / det
    tx = ray.ox - tri.x0
    ty = ray.oy - tri.y0
    tz = ray.oz - tri.z0
    u = (tx * px + ty * py + tz * pz) * inv
    if u < 0.0 or u > 1.0:
        return None
    qx = ty * e1z - tz * e1y
    qy = tz * e1x - tx * e1z
    qz = tx * e1y - ty * e1x
    v = (ray.dx * qx + ray.dy * qy + ray.dz * qz) * inv
    if v < 0.0 or (u + v) > 1.0:
        return None
    t = (e2x * qx + e2y * qy + e2z * qz) * inv
    if t >= 0.0 and t <= ray.tmax:
        return t
    return None


def _segment_intersection(px, py, ex, ey, qx0, qy0, qx1, qy1):
    # reference: src/rtdsl/reference.py:693-714 (left = p -> e, right = q0 -> q1)
    rx = ex - px
    ry = ey - py
    sx = qx1 - qx0
    sy = qy1 - qy0
    denom = rx * sy - ry * sx
    if abs(denom) < 1.0e-7:
        return False
    qpx = qx0 - px
    qpy = qy0 - py
    t = (qpx * sy - qpy * sx) / denom
    u = (qpx * ry - qpy * rx) / denom
    return 0.0 <= t <= 1.0 and 0.0 <= u <= 1.0


def _point_in_triangle_2d(x, y, ax, ay, bx, by, cx, cy) -> bool:
    # reference: src/rtdsl/reference.py:893-915
    v0x = cx - ax
    v0y = cy - ay
    v1x = bx - ax
    v1y = by - ay
    v2x = x - ax
    v2y = y - ay
    dot00 = v0x * v0x + v0y * v0y
    dot01 = v0x * v1x + v0y * v1y
    dot02 = v0x * v2x + v0y * v2y
    dot11 = v1x * v1x + v1y * v1y
    dot12 = v1x * v2x + v1y * v2y
    denom = dot00 * dot11 - dot01 * dot01
    if abs(denom) < 1.0e-7:
        return False
    inv = 1.0 / denom
    u = (dot11 * dot02 - dot01 * dot12) * inv
    v = (dot00 * dot12 - dot01 * dot02) * inv
    return u >= 0.0 and v >= 0.0 and (u + v) <= 1.0


def _hits_2d(ray, tri) -> bool:
    # reference: src/rtdsl/reference.py:783-803
    ex = ray.ox + ray.dx * ray.tmax
    ey = ray.oy + ray.dy * ray.tmax
    verts = (tri.x0, tri.y0, tri.x1, tri.y1, tri.x2, tri.y2)
    if _point_in_triangle_2d(ray.ox, ray.oy, *verts):
        return True
    if _point_in_triangle_2d(ex, ey, *verts):
        return True
    edges = (
        (tri.x0, tri.y0, tri.x1, tri.y1),
        (tri.x1, tri.y1, tri.x2, tri.y2),
        (tri.x2, tri.y2, tri.x0, tri.y0),
    )
    return any(_segment_intersection(ray.ox, ray.oy, ex, ey, *edge) for edge in edges)


def _hit_t(ray, tri):
    # reference: src/rtdsl/reference.py:775-780
    if isinstance(ray, Ray3D) or isinstance(tri, Triangle3D):
        if not isinstance(ray, Ray3D) or not isinstance(tri, Triangle3D):
            raise ValueError(
                "ray_triangle_hit_count_cpu requires rays and triangles to both be 2D or both be 3D"
            )
        return _hit_t_3d(ray, tri)
    return 0.0 if _hits_2d(ray, tri) else None


def ray_triangle_any_hit_cpu(rays, triangles):
    # reference: src/rtdsl/reference.py:188-200
    rows = []
    for ray in rays:
        hit = 0
        for tri in triangles:
            if _hit_t(ray, tri) is not None:
                hit = 1
                break
        rows.append({"ray_id": ray.id, "any_hit": hit})
    return tuple(rows)


def ray_triangle_hit_count_cpu(rays, triangles):
    # reference: src/rtdsl/reference.py:174-186
    rows = []
    for ray in rays:
        count = sum(1 for tri in triangles if _hit_t(ray, tri) is not None)
        rows.append({"ray_id": ray.id, "hit_count": count})
    return tuple(rows)


def ray_triangle_closest_hit_cpu(rays, triangles):
    # reference: src/rtdsl/reference.py:248-266
    rows = []
    for ray in rays:
        best_t = None
        best_id = None
        for tri in triangles:
            t = _hit_t(ray, tri)
            if t is None:
                continue
            tid = int(tri.id)
            if best_t is None or t < best_t or (t == best_t and tid < int(best_id)):
                best_t = t
                best_id = tid
        if best_t is not None and best_id is not None:
            rows.append({"ray_id": int(ray.id), "triangle_id": best_id, "t": best_t})
    return tuple(rows)


# --------------------------------------------------------------------------
# visibility rows (reference: src/rtdsl/reference.py:275-346)
# --------------------------------------------------------------------------


def _dimension(record) -> int:
    # reference: src/rtdsl/reference.py:269-274 (points and triangles)
    return 3 if isinstance(record, (Point3D, Triangle3D)) else 2


def _visibility_ray(ray_id, observer, target):
    # reference: src/rtdsl/reference.py:275-293 (segment observer -> target, tmax 1)
    if _dimension(observer) != _dimension(target):
        raise ValueError("visibility_rows requires observers and targets with the same dimensionality")
    if isinstance(observer, Point3D):
        dx = target.x - observer.x
        dy = target.y - observer.y
        dz = target.z - observer.z
        if dx == 0.0 and dy == 0.0 and dz == 0.0:
            raise ValueError("visibility_rows requires distinct observer and target points")
        return Ray3D(ray_id, observer.x, observer.y, observer.z, dx, dy, dz, 1.0)
    dx = target.x - observer.x
    dy = target.y - observer.y
    if dx == 0.0 and dy == 0.0:
        raise ValueError("visibility_rows requires distinct observer and target points")
    return Ray2D(ray_id, observer.x, observer.y, dx, dy, 1.0)


def visibility_ray_pairs(observers, targets, blockers):
    # reference: src/rtdsl/reference.py:310-334 (observer-major ray ids)
    if not observers or not targets:
        return (), ()
    dim = _dimension(observers[0])
    if any(_dimension(p) != dim for p in (*observers, *targets)):
        raise ValueError("visibility_rows requires all observers and targets to share one dimensionality")
    if blockers and any(_dimension(t) != dim for t in blockers):
        raise ValueError("visibility_rows blocker triangles must match observer/target dimensionality")
    rays, pairs = [], []
    for observer in observers:
        for target in targets:
            rid = len(rays)
            rays.append(_visibility_ray(rid, observer, target))
            pairs.append((rid, int(observer.id), int(target.id)))
    return tuple(rays), tuple(pairs)


def visibility_rows_from_any_hit(ray_pairs, any_hit_by_ray):
    # reference: src/rtdsl/reference.py:337-346
    return tuple({"observer_id": o, "target_id": t, "visible": 0 if any_hit_by_ray.get(r, 0) else 1}
                 for r, o, t in ray_pairs)


def visibility_rows_cpu(observers, targets, blockers):
    # reference: src/rtdsl/reference.py:296-307
    rays, pairs = visibility_ray_pairs(observers, targets, blockers)
    hits = {int(r["ray_id"]): int(r["any_hit"]) for r in ray_triangle_any_hit_cpu(rays, blockers)}
    return visibility_rows_from_any_hit(pairs, hits)


# --------------------------------------------------------------------------
# point-in-polygon (reference: src/rtdsl/reference.py:103-130, 717-768)
# --------------------------------------------------------------------------


def _point_on_segment(px, py, ax, ay, bx, by) -> bool:
    # reference: src/rtdsl/reference.py:745-768
    eps = 1.0e-12
    length_sq = (bx - ax) ** 2 + (by - ay) ** 2
    if length_sq <= (eps ** 2):
        return abs(px - ax) <= eps and abs(py - ay) <= eps
    cross = (px - ax) * (by - ay) - (py - ay) * (bx - ax)
    length = length_sq ** 0.5
    if abs(cross) > eps * length:
        return False
    dot = (px - ax) * (bx - ax) + (py - ay) * (by - ay)
    along_eps = eps * length
    if dot < -along_eps:
        return False
    if dot - length_sq > along_eps:
        return False
    return True


def point_in_polygon_test(x: float, y: float, vertices, *, boundary_mode: str = "inclusive") -> bool:
    # reference: src/rtdsl/reference.py:717-742
    if boundary_mode != "inclusive":
        raise ValueError(
            "the current point-in-polygon implementation supports only boundary_mode='inclusive'"
        )
    n = len(vertices)
    for i in range(n):
        ax, ay = vertices[i]
        bx, by = vertices[(i + 1) % n]
        if _point_on_segment(x, y, ax, ay, bx, by):
            return True
    inside = False
    j = n - 1
    for i in range(n):
        xi, yi = vertices[i]
        xj, yj = vertices[j]
        if ((yi > y) != (yj > y)) and (x <= (xj - xi) * (y - yi) / ((yj - yi) or 1.0e-20) + xi):
            inside = not inside
        j = i
    return inside


def pip_cpu(points, polygons, *, boundary_mode: str = "inclusive", result_mode: str = "full_matrix"):
    # reference: src/rtdsl/reference.py:103-130
    if result_mode not in {"full_matrix", "positive_hits"}:
        raise ValueError("pip_cpu result_mode must be 'full_matrix' or 'positive_hits'")
    bounds = None
    if result_mode == "positive_hits":
        bounds = []
        for polygon in polygons:
            xs = [vx for vx, _ in polygon.vertices]
            ys = [vy for _, vy in polygon.vertices]
            bounds.append((min(xs), min(ys), max(xs), max(ys)))
    rows = []
    for point in points:
        for index, polygon in enumerate(polygons):
            if bounds is not None:
                x0, y0, x1, y1 = bounds[index]
                if point.x < x0 or point.x > x1 or point.y < y0 or point.y > y1:
                    continue
            inside = point_in_polygon_test(point.x, point.y, polygon.vertices, boundary_mode=boundary_mode)
            contains = 1 if inside else 0
            if result_mode == "positive_hits" and contains != 1:
                continue
            rows.append({"point_id": point.id, "polygon_id": polygon.id, "contains": contains})
    return tuple(rows)


# --------------------------------------------------------------------------
# kernel-level runner (reference: src/rtdsl/runtime.py:120-246)
# --------------------------------------------------------------------------


def _resolve(kernel_fn_or_compiled):
    from rtdl_amd.ir import CompiledKernel

    if isinstance(kernel_fn_or_compiled, CompiledKernel):
        return kernel_fn_or_compiled
    return kernel_fn_or_compiled()


def run_cpu_python_reference(kernel_fn_or_compiled, **inputs):
    compiled = _resolve(kernel_fn_or_compiled)
    if compiled.precision != "float_approx":
        raise ValueError("RTDL CPU simulator currently supports only precision='float_approx'")
    if compiled.candidates is None or compiled.refine_op is None or compiled.emit_op is None:
        raise ValueError("RTDL CPU simulator requires a fully compiled kernel with traverse/refine/emit")
    expected = {item.name: item for item in compiled.inputs}
    missing = [name for name in expected if name not in inputs]
    unexpected = [name for name in inputs if name not in expected]
    if missing:
        raise ValueError(f"missing RTDL simulator inputs: {', '.join(sorted(missing))}")
    if unexpected:
        raise ValueError(f"unexpected RTDL simulator inputs: {', '.join(sorted(unexpected))}")
    data = {
        name: normalize_records(name, expected[name].geometry.name, payload)
        for name, payload in inputs.items()
    }
    predicate = compiled.refine_op.predicate
    left = data[compiled.candidates.left.name]
    right = data[compiled.candidates.right.name]
    opts = predicate.options
    if predicate.name == "fixed_radius_neighbors":
        rows = fixed_radius_neighbors_cpu(left, right, radius=float(opts["radius"]), k_max=int(opts["k_max"]))
    elif predicate.name == "knn_rows":
        rows = knn_rows_cpu(left, right, k=int(opts["k"]))
    elif predicate.name == "bounded_knn_rows":
        rows = bounded_knn_rows_cpu(left, right, radius=float(opts["radius"]), k_max=int(opts["k_max"]))
    elif predicate.name == "ray_triangle_any_hit":
        rows = ray_triangle_any_hit_cpu(left, right)
    elif predicate.name == "ray_triangle_hit_count":
        rows = ray_triangle_hit_count_cpu(left, right)
    elif predicate.name == "ray_triangle_closest_hit":
        rows = ray_triangle_closest_hit_cpu(left, right)
    elif predicate.name == "point_in_polygon":
        rows = pip_cpu(
            left,
            right,
            boundary_mode=opts.get("boundary_mode", "inclusive"),
            result_mode=opts.get("result_mode", "full_matrix"),
        )
    else:
        raise ValueError(f"unsupported RTDL CPU simulator predicate: {predicate.name}")
    return project_rows(compiled, rows)


__all__ = [
    "Point3D",
    "Polygon",
    "bounded_knn_rows_cpu",
    "fixed_radius_neighbors_cpu",
    "knn_rows_cpu",
    "pip_cpu",
    "point_in_polygon_test",
    "ray_triangle_any_hit_cpu",
    "ray_triangle_closest_hit_cpu",
    "ray_triangle_hit_count_cpu",
    "run_cpu_python_reference",
]
