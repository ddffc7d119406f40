"""Pins the oracle before it is trusted as the checker.

* The reference's own authored cases and their hard-coded expectations
  (tests/goal198, goal204, goal548, goal552, goal632, goal633 in the reference).
* 2-D any-hit against an independent separating-axis decision on random
  segments (both signs of tmax) and triangles, plus special cases (a segment
  inside a triangle, an end on an edge, unbounded and negative tmax).
* Closed forms a plausible mistake would break: inclusive radius boundary,
  truncation to k_max, (distance, id) tie order, grouping by query id, no
  KNN padding, KNN rows as a superset of FRN rows at the k-th distance,
  Moller-Trumbore special cases (det threshold, u/v/t bounds, tmax), PIP on
  vertices and edges and the even-odd rule on a concave polygon.
* The two independent restatements (pure Python and C) agree bit for bit on
  random inputs, with and without the C grid filter.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import cpu_python_reference as ref
from oracle import native
from rtdl_amd.records import Point, Point3D, Polygon, Ray2D, Ray3D, Triangle, Triangle3D
from tests import cases

# ------------------------------------------------------------ authored cases --


def _authored_points():
    return (
        (Point(100, 0.0, 0.0), Point(101, 3.0, 0.0)),
        (Point(1, 0.0, 0.0), Point(2, 0.3, 0.0), Point(3, -0.3, 0.0), Point(4, 3.2, 0.0),
         Point(5, 4.0, 0.0)),
    )


def test_frn_authored_rows_goal198():
    q, s = _authored_points()
    rows = ref.fixed_radius_neighbors_cpu(q, s, radius=0.5, k_max=3)
    assert len(rows) == 4
    assert [r["query_id"] for r in rows] == [100, 100, 100, 101]
    assert [r["neighbor_id"] for r in rows] == [1, 2, 3, 4]
    assert rows[0]["distance"] == 0.0
    assert math.isclose(rows[1]["distance"], 0.3, rel_tol=1e-9)
    assert math.isclose(rows[2]["distance"], 0.3, rel_tol=1e-9)
    assert math.isclose(rows[3]["distance"], 0.2, rel_tol=1e-9)


def test_frn_grouping_and_zero_radius_goal198():
    rows = ref.fixed_radius_neighbors_cpu(
        (Point(20, 3.0, 0.0), Point(10, 0.0, 0.0)), (Point(1, 0.0, 0.0), Point(2, 3.0, 0.0)),
        radius=0.1, k_max=2)
    assert [r["query_id"] for r in rows] == [10, 20]
    rows = ref.fixed_radius_neighbors_cpu(
        (Point(10, 1.0, 1.0),), (Point(1, 1.0, 1.0), Point(2, 1.0, 1.0001)), radius=0.0, k_max=4)
    assert rows == ({"query_id": 10, "neighbor_id": 1, "distance": 0.0},)


def test_knn_authored_rows_goal204():
    q, s = _authored_points()
    rows = ref.knn_rows_cpu(q, s, k=3)
    assert len(rows) == 6
    assert [(r["neighbor_id"], r["neighbor_rank"]) for r in rows[:3]] == [(1, 1), (2, 2), (3, 3)]
    assert rows[3]["query_id"] == 101 and rows[3]["neighbor_rank"] == 1
    rows = ref.knn_rows_cpu((Point(20, 3.0, 0.0), Point(10, 0.0, 0.0)),
                            (Point(1, 0.0, 0.0), Point(2, 3.0, 0.0)), k=2)
    assert [r["query_id"] for r in rows] == [10, 10, 20, 20]
    rows = ref.knn_rows_cpu((Point(10, 1.0, 1.0),), (Point(1, 1.0, 1.0),), k=4)
    assert rows == ({"query_id": 10, "neighbor_id": 1, "distance": 0.0, "neighbor_rank": 1},)


def test_frn_3d_authored_goal548():
    q = (Point3D(10, 0.0, 0.0, 0.0), Point3D(20, 5.0, 0.0, 0.0))
    s = (Point3D(1, 0.0, 0.0, 0.0), Point3D(2, 0.0, 0.0, 0.6), Point3D(3, 0.0, 0.8, 0.0),
         Point3D(4, 5.0, 0.0, 0.9))
    rows = ref.fixed_radius_neighbors_cpu(q, s, radius=1.0, k_max=3)
    assert [(r["query_id"], r["neighbor_id"]) for r in rows] == [(10, 1), (10, 2), (10, 3), (20, 4)]
    assert [r["distance"] for r in rows] == [0.0, 0.6, 0.8, 0.9]


def test_any_hit_authored_goal632():
    rays = (Ray2D(1, 0.0, 0.0, 1.0, 0.0, 10.0), Ray2D(2, 0.0, 0.0, 0.0, 1.0, 2.0))
    tris = (Triangle(10, 2.0, -1.0, 3.0, 1.0, 4.0, -1.0), Triangle(11, 6.0, -1.0, 7.0, 1.0, 8.0, -1.0),
            Triangle(12, -1.0, 3.0, 1.0, 3.0, 0.0, 4.0))
    assert ref.ray_triangle_any_hit_cpu(rays, tris) == ({"ray_id": 1, "any_hit": 1}, {"ray_id": 2, "any_hit": 0})
    assert [r["hit_count"] for r in ref.ray_triangle_hit_count_cpu(rays, tris)] == [2, 0]
    rays3 = (Ray3D(1, 0.0, 0.0, 0.0, 10.0, 0.0, 0.0, 1.0), Ray3D(2, 0.0, 2.0, 0.0, 10.0, 0.0, 0.0, 1.0))
    tri3 = (Triangle3D(20, 5.0, -1.0, -1.0, 5.0, 1.0, -1.0, 5.0, 0.0, 1.0),)
    assert ref.ray_triangle_any_hit_cpu(rays3, tri3) == ({"ray_id": 1, "any_hit": 1}, {"ray_id": 2, "any_hit": 0})


def _square(pid, x0, y0, size):
    return Polygon(pid, ((x0, y0), (x0 + size, y0), (x0 + size, y0 + size), (x0, y0 + size)))


def test_pip_authored_squares_goal552():
    points = (Point(1, 0.0, 0.0), Point(2, 0.5, 0.5), Point(3, 3.0, 3.0))
    polys = (_square(10, -0.25, -0.25, 1.0), _square(20, 2.0, 2.0, 1.0))
    full = ref.pip_cpu(points, polys)
    assert [(r["point_id"], r["polygon_id"], r["contains"]) for r in full] == [
        (1, 10, 1), (1, 20, 0), (2, 10, 1), (2, 20, 0), (3, 10, 0), (3, 20, 1)]
    pos = ref.pip_cpu(points, polys, result_mode="positive_hits")
    assert [(r["point_id"], r["polygon_id"]) for r in pos] == [(1, 10), (2, 10), (3, 20)]


def test_bounded_knn_authored_rows_goal263():
    """The reference's authored bounded-KNN rows
    (tests/goal263_v0_5_bounded_knn_rows_surface_test.py:27-56): inclusive
    radius, ranks from 1, points beyond the radius dropped even below k_max."""
    q = (Point3D(10, 0.0, 0.0, 0.0),)
    s = (Point3D(1, 0.0, 0.0, 0.0), Point3D(2, 0.0, 0.0, 0.4), Point3D(3, 0.0, 0.0, 0.8))
    want = ({"query_id": 10, "neighbor_id": 1, "distance": 0.0, "neighbor_rank": 1},
            {"query_id": 10, "neighbor_id": 2, "distance": 0.4, "neighbor_rank": 2})
    assert ref.bounded_knn_rows_cpu(q, s, radius=0.5, k_max=3) == want
    cols = lambda pts: {"id": np.array([p.id for p in pts], np.uint32), "x": np.array([p.x for p in pts]),
                        "y": np.array([p.y for p in pts]), "z": np.array([p.z for p in pts])}
    got = native.neighbors("bounded_knn", cols(q), cols(s), radius=0.5, k=3)
    assert list(zip(got["qid"].tolist(), got["nid"].tolist(), got["dist"].tolist(), got["rank"].tolist())) == \
        [(10, 1, 0.0, 1), (10, 2, 0.4, 2)]
    s2 = (Point3D(1, 0.0, 0.0, 0.0), Point3D(2, 0.0, 0.0, 0.25), Point3D(3, 0.0, 0.0, 2.0))
    rows = ref.bounded_knn_rows_cpu(q, s2, radius=1.0, k_max=2)
    assert [r["neighbor_id"] for r in rows] == [1, 2]
    assert [r["neighbor_rank"] for r in rows] == [1, 2]
    got = native.neighbors("bounded_knn", cols(q), cols(s2), radius=1.0, k=2)
    assert got["nid"].tolist() == [1, 2] and got["rank"].tolist() == [1, 2]


# ------------------------------------------------------------ closed forms --


def test_frn_radius_is_inclusive_and_truncates_by_distance_then_id():
    q = (Point(0, 0.0, 0.0),)
    s = (Point(9, 0.5, 0.0), Point(3, 0.0, 0.5), Point(7, 0.0, -0.5), Point(1, 0.25, 0.0), Point(2, 0.6, 0.0))
    rows = ref.fixed_radius_neighbors_cpu(q, s, radius=0.5, k_max=3)
    assert [r["neighbor_id"] for r in rows] == [1, 3, 7]  # 0.25, then the 0.5 tie by id
    rows = ref.fixed_radius_neighbors_cpu(q, s, radius=0.5, k_max=10)
    assert [r["neighbor_id"] for r in rows] == [1, 3, 7, 9]  # 0.5 is inside, 0.6 is not


def test_knn_contains_frn_rows_at_kth_distance():
    rng = np.random.default_rng(3)
    pts = [Point3D(i, *map(float, rng.random(3))) for i in range(300)]
    q = pts[:20]
    knn = ref.knn_rows_cpu(q, pts, k=7)
    for query in q:
        mine = [r for r in knn if r["query_id"] == query.id]
        kth = mine[-1]["distance"]
        frn = ref.fixed_radius_neighbors_cpu((query,), pts, radius=kth, k_max=7)
        assert {r["neighbor_id"] for r in frn} <= {r["neighbor_id"] for r in mine}
        assert [r["neighbor_rank"] for r in mine] == list(range(1, 8))


@pytest.mark.parametrize("dim", [2, 3])
def test_bounded_knn_is_frn_with_ranks(dim):
    """bounded_knn_rows(r, k) (reference.py:406-434) selects exactly the rows
    of fixed_radius_neighbors(r, k) (pinned above by the authored cases) and
    numbers them 1..n within each query; duplicates and points on the radius
    make the tie order and the inclusive bound decide rows."""
    rng = np.random.default_rng(70 + dim)
    q = cases.points(rng, 120, dim, fp32=False, ids=rng.permutation(120))
    s = cases.points(rng, 300, dim, kind="mixture", fp32=False, ids=rng.permutation(300))
    for key in ("x", "y", "z")[:dim]:
        s[key][:30] = s[key][30:60]          # exact duplicates: equal distances, other ids
        s[key][60:70] = q[key][:10]          # neighbours at distance 0
    s["x"][70:80] = q["x"][10:20] + 0.125    # exactly on the radius below
    for key in ("y", "z")[:dim - 1]:
        s[key][70:80] = q[key][10:20]
    qr, sr = _as_records(q, dim), _as_records(s, dim)
    for radius, k in ((0.125, 4), (0.05, 1), (0.3, 40), (0.0, 3)):
        frn = ref.fixed_radius_neighbors_cpu(qr, sr, radius=radius, k_max=k)
        bkn = ref.bounded_knn_rows_cpu(qr, sr, radius=radius, k_max=k)
        assert [(r["query_id"], r["neighbor_id"], r["distance"]) for r in bkn] == \
            [(r["query_id"], r["neighbor_id"], r["distance"]) for r in frn]
        rank, last = [], None
        for r in frn:
            rank.append(rank[-1] + 1 if rank and r["query_id"] == last else 1)
            last = r["query_id"]
        assert [r["neighbor_rank"] for r in bkn] == rank
        got = native.neighbors("bounded_knn", q, s, radius=radius, k=k)
        assert got["rank"].tolist() == rank
        assert got["nid"].tolist() == [r["neighbor_id"] for r in frn]


def test_closest_hit_3d_exact_t_and_lower_id_tie():
    """reference.py:248-266 keeps the smallest t and, at equal t, the lower
    triangle id.  An axis ray (0.25, 0.25, 0) + t (0, 0, 1) crosses the
    triangle (0,0,c) (1,0,c) (0,1,c) at exactly t = c (det = -1, u = v = 0.25,
    every product exact), so t is known in closed form.  The nearest planes
    come after farther ones in input order and a coplanar duplicate with a
    lower id follows its twin, so neither "first hit" nor "last hit" passes."""
    def tri(ident, c):
        return Triangle3D(ident, 0.0, 0.0, c, 1.0, 0.0, c, 0.0, 1.0, c)

    tris = (tri(7, 2.0), tri(9, 3.0), tri(5, 1.0), tri(2, 1.0), tri(11, 1.0), tri(1, 5.0))
    rays = (Ray3D(40, 0.25, 0.25, 0.0, 0.0, 0.0, 1.0, 4.0),   # hits c = 1, 1, 1, 2, 3
            Ray3D(41, 0.25, 0.25, 0.0, 0.0, 0.0, 1.0, 1.0),   # t = tmax = 1 is a hit
            Ray3D(42, 0.25, 0.25, 0.0, 0.0, 0.0, 1.0, 0.5),   # no hit: no row
            Ray3D(43, 0.25, 0.25, 2.5, 0.0, 0.0, 1.0, 4.0))   # starts past c = 1, 2; hits 3 and 5
    want = ({"ray_id": 40, "triangle_id": 2, "t": 1.0}, {"ray_id": 41, "triangle_id": 2, "t": 1.0},
            {"ray_id": 43, "triangle_id": 9, "t": 0.5})
    assert ref.ray_triangle_closest_hit_cpu(rays, tris) == want
    assert [r["hit_count"] for r in ref.ray_triangle_hit_count_cpu(rays, tris)] == [5, 3, 0, 2]
    rm = np.array([[r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, r.tmax] for r in rays])
    tm = np.array([[t.x0, t.y0, t.z0, t.x1, t.y1, t.z1, t.x2, t.y2, t.z2] for t in tris])
    ids = np.array([t.id for t in tris], np.uint32)
    tid, tt, has = native.rays_3d("closest", rm, tm, ids)
    assert has.tolist() == [True, True, False, True]
    assert tid[has].tolist() == [2, 2, 9] and tt[has].tolist() == [1.0, 1.0, 0.5]
    assert native.rays_3d("count", rm, tm, ids).tolist() == [5, 3, 0, 2]


def test_fl32_key_tie_pair_is_ordered_by_id():
    """Two FP32-exact points whose FP64 squared distances round to FP32 keys
    one ulp apart share one FP64 square root: the reference orders them by id
    (reference.py:362-366, 388-389), not by squared distance.  The pair is the
    case a selection by FP32 keys gets wrong unless it treats keys within one
    ulp as tied."""
    (x, ya, yb, a, b), = cases.fl32_tie_pairs(1)
    assert a < b and math.sqrt(a) == math.sqrt(b)
    ka, kb = (int(np.float32(v).view(np.uint32)) for v in (a, b))
    assert kb - ka == 1
    assert x * x + ya * ya == a and x * x + yb * yb == b  # FP64 in the reference order
    for dim in (2, 3):
        pad = (0.0,) * (dim - 2)
        mk = Point3D if dim == 3 else Point
        q = (mk(0, 0.0, 0.0, *pad),)
        s = (mk(9, x, ya, *pad), mk(4, x, yb, *pad), mk(6, 0.5, 0.5, *pad))
        rows = ref.knn_rows_cpu(q, s, k=2)
        assert [r["neighbor_id"] for r in rows] == [4, 9]  # the farther d2 first: lower id
        assert ref.fixed_radius_neighbors_cpu(q, s, radius=0.02, k_max=1)[0]["neighbor_id"] == 4
        cols = {"id": np.array([9, 4, 6], np.uint32), "x": np.array([x, x, 0.5]), "y": np.array([ya, yb, 0.5])}
        qc = {"id": np.array([0], np.uint32), "x": np.zeros(1), "y": np.zeros(1)}
        if dim == 3:
            cols["z"], qc["z"] = np.zeros(3), np.zeros(1)
        assert native.neighbors("knn", qc, cols, k=2)["nid"].tolist() == [4, 9]


def test_moller_trumbore_special_cases():
    tri = Triangle3D(1, 0.0, 0.0, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0, 1.0)
    hit = ref._hit_t_3d
    assert hit(Ray3D(0, 0.2, 0.2, 0.0, 0.0, 0.0, 1.0, 5.0), tri) == 1.0
    assert hit(Ray3D(0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 5.0), tri) == 1.0  # vertex, u = v = 0
    assert hit(Ray3D(0, 0.5, 0.5, 0.0, 0.0, 0.0, 1.0, 5.0), tri) == 1.0  # hypotenuse, u + v = 1
    assert hit(Ray3D(0, 0.6, 0.6, 0.0, 0.0, 0.0, 1.0, 5.0), tri) is None
    assert hit(Ray3D(0, 0.2, 0.2, 0.0, 0.0, 0.0, 1.0, 1.0), tri) == 1.0  # t == tmax inclusive
    assert hit(Ray3D(0, 0.2, 0.2, 0.0, 0.0, 0.0, 1.0, 0.5), tri) is None
    assert hit(Ray3D(0, 0.2, 0.2, 2.0, 0.0, 0.0, 1.0, 5.0), tri) is None  # behind (t < 0)
    assert hit(Ray3D(0, 0.2, 0.2, 0.0, 1.0, 0.0, 0.0, 5.0), tri) is None  # parallel, det = 0
    # det just above / below the 1e-8 threshold: det = d_z * (e1 x e2)_z
    small = Triangle3D(1, 0.0, 0.0, 1.0, 1e-4, 0.0, 1.0, 0.0, 1e-4, 1.0)
    assert hit(Ray3D(0, 1e-5, 1e-5, 0.0, 0.0, 0.0, 1.0, 5.0), small) is None  # det = 1e-8 -> miss
    assert hit(Ray3D(0, 1e-5, 1e-5, 0.0, 0.0, 0.0, 2.0, 5.0), small) == 0.5


def test_pip_on_edges_vertices_and_concave_even_odd():
    concave = ((0.0, 0.0), (4.0, 0.0), (4.0, 4.0), (2.0, 1.0), (0.0, 4.0))
    test = ref.point_in_polygon_test
    assert test(1.0, 0.5, concave)
    assert not test(2.0, 2.0, concave)  # inside the notch
    assert test(2.0, 1.0, concave)  # on the notch vertex
    assert test(3.0, 2.5, concave)
    assert test(4.0, 2.0, concave)  # on an edge
    assert test(3.0, 1.5, concave)  # on the slanted edge (4,4)-(2,1)
    assert not test(-1e-9, 2.0, concave)
    assert test(-1e-13, 2.0, concave)  # within the 1e-12 edge tolerance


# --------------------------------------------- two restatements cross-check --


def _as_records(cols, dim):
    if dim == 3:
        return tuple(Point3D(int(i), float(x), float(y), float(z))
                     for i, x, y, z in zip(cols["id"], cols["x"], cols["y"], cols["z"]))
    return tuple(Point(int(i), float(x), float(y)) for i, x, y in zip(cols["id"], cols["x"], cols["y"]))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("brute", [False, True])
def test_c_oracle_matches_python_oracle_neighbors(dim, brute):
    rng = np.random.default_rng(40 + dim)
    q = cases.points(rng, 150, dim, fp32=False, ids=rng.permutation(150))
    s = cases.points(rng, 250, dim, kind="mixture", fp32=False)
    s["x"][:5] = q["x"][:5]
    s["y"][:5] = q["y"][:5]
    if dim == 3:
        s["z"][:5] = q["z"][:5]
    qr, sr = _as_records(q, dim), _as_records(s, dim)
    for radius, k in ((0.05, 3), (0.2, 1000), (0.0, 2)):
        want = ref.fixed_radius_neighbors_cpu(qr, sr, radius=radius, k_max=k)
        got = native.neighbors("frn", q, s, radius=radius, k=k, brute=brute)
        assert [(r["query_id"], r["neighbor_id"], r["distance"]) for r in want] == \
            list(zip(got["qid"].tolist(), got["nid"].tolist(), got["dist"].tolist()))
        want = ref.bounded_knn_rows_cpu(qr, sr, radius=radius, k_max=k)
        got = native.neighbors("bounded_knn", q, s, radius=radius, k=k, brute=brute)
        assert [r["neighbor_rank"] for r in want] == got["rank"].tolist()
    for k in (1, 9, 300):
        want = ref.knn_rows_cpu(qr, sr, k=k)
        got = native.neighbors("knn", q, s, k=k, brute=brute)
        assert [(r["query_id"], r["neighbor_id"], r["distance"], r["neighbor_rank"]) for r in want] == \
            list(zip(got["qid"].tolist(), got["nid"].tolist(), got["dist"].tolist(), got["rank"].tolist()))


@pytest.mark.parametrize("brute", [False, True])
def test_c_oracle_matches_python_oracle_rays(brute):
    rng = np.random.default_rng(50)
    t = cases.triangles(rng, 300, sigma=0.05, fp32=False)
    r = cases.rays(rng, 200, tmax=0.5, fp32=False)
    tm, rm = cases.tri_matrix(t), cases.ray_matrix(r)
    tris = tuple(Triangle3D(int(t["id"][i]), *map(float, tm[i])) for i in range(len(tm)))
    rays = tuple(Ray3D(int(r["id"][i]), *map(float, rm[i])) for i in range(len(rm)))
    assert [x["any_hit"] for x in ref.ray_triangle_any_hit_cpu(rays, tris)] == \
        native.rays_3d("any", rm, tm, t["id"], brute=brute).tolist()
    assert [x["hit_count"] for x in ref.ray_triangle_hit_count_cpu(rays, tris)] == \
        native.rays_3d("count", rm, tm, t["id"], brute=brute).tolist()
    tid, tt, has = native.rays_3d("closest", rm, tm, t["id"], brute=brute)
    want = ref.ray_triangle_closest_hit_cpu(rays, tris)
    assert [(x["ray_id"], x["triangle_id"], x["t"]) for x in want] == \
        [(int(i), int(tid[i]), float(tt[i])) for i in np.flatnonzero(has)]


@pytest.mark.parametrize("positive_only", [True, False])
@pytest.mark.parametrize("brute", [False, True])
def test_c_oracle_matches_python_oracle_pip(positive_only, brute):
    rng = np.random.default_rng(60)
    polys = cases.star_polygons(rng, 40, rmin=0.05, rmax=0.2, fp32=False)
    refs, v = polys["refs"], polys["vertices_xy"]
    pts = cases.points(rng, 400, 2, fp32=False)
    vv = v.reshape(-1, 2)
    pts["x"][:20], pts["y"][:20] = vv[:20, 0], vv[:20, 1]
    pts["x"][20:40] = (vv[20:40, 0] + vv[21:41, 0]) / 2
    pts["y"][20:40] = (vv[20:40, 1] + vv[21:41, 1]) / 2
    recs = tuple(Polygon(int(refs["id"][g]), tuple(map(tuple, vv[refs["vertex_offset"][g]:
                 refs["vertex_offset"][g] + refs["vertex_count"][g]].tolist()))) for g in range(len(refs)))
    probes = _as_records(pts, 2)
    mode = "positive_hits" if positive_only else "full_matrix"
    want = [(r["point_id"], r["polygon_id"]) for r in ref.pip_cpu(probes, recs, result_mode=mode) if r["contains"]]
    pairs = native.pip_pairs(pts["x"], pts["y"], refs["vertex_offset"], refs["vertex_count"], v,
                             positive_only=positive_only, brute=brute)
    assert want == [(int(pts["id"][i]), int(refs["id"][g])) for i, g in pairs]


def test_libm_pow_is_not_the_correctly_rounded_square():
    """Documents why the on-edge test can differ at its tolerance boundary:
    the reference's `x ** 2` is libm pow, which is not always x * x."""
    rng = np.random.default_rng(0)
    xs = rng.random(200000).tolist()
    diffs = sum(1 for x in xs if math.pow(x, 2.0) != x * x)
    assert diffs < len(xs) * 0.01


# ------------------------------------------------ 2-D any-hit and visibility --


def _sat_segment_triangle(seg, tri, margin=1e-6):
    """Separating-axis decision for a segment and a triangle in the plane:
    1 intersect, 0 disjoint, None when some axis is within `margin` of touching.
    Independent of the reference's barycentric / crossing formulation."""
    (ax, ay), (bx, by) = seg
    verts = np.array(tri).reshape(3, 2)
    sv = np.array([[ax, ay], [bx, by]])
    axes = [np.array([-(by - ay), bx - ax])]
    for i in range(3):
        e = verts[(i + 1) % 3] - verts[i]
        axes.append(np.array([-e[1], e[0]]))
    verdict = 1
    for n in axes:
        n = n / np.linalg.norm(n)
        tp, sp = verts @ n, sv @ n
        gap = max(tp.min() - sp.max(), sp.min() - tp.max())
        if abs(gap) < margin:
            return None
        if gap > 0:
            verdict = 0
    return verdict


def _random_2d_case(rng, n_rays, n_tris):
    rays = []
    for i in range(n_rays):
        ox, oy = rng.random(2)
        ang = rng.uniform(0, 2 * math.pi)
        tmax = rng.choice([rng.uniform(0.02, 0.4), -rng.uniform(0.02, 0.4)])
        rays.append(Ray2D(i, float(ox), float(oy), math.cos(ang), math.sin(ang), float(tmax)))
    tris = []
    while len(tris) < n_tris:
        c = rng.random(2)
        v = c + rng.normal(0, 0.08, (3, 2))
        area = abs((v[1, 0] - v[0, 0]) * (v[2, 1] - v[0, 1]) - (v[2, 0] - v[0, 0]) * (v[1, 1] - v[0, 1])) / 2
        if area > 1e-3:
            tris.append(Triangle(len(tris) + 100, *map(float, v.reshape(-1))))
    return tuple(rays), tuple(tris)


def test_2d_any_hit_matches_separating_axes():
    """Random segments (both signs of tmax) against random triangles: every
    pair that the separating-axis test decides clearly has the same verdict in
    the Python and C restatements."""
    rng = np.random.default_rng(633)
    rays, tris = _random_2d_case(rng, 60, 80)
    rm = np.array([[r.ox, r.oy, r.dx, r.dy, r.tmax] for r in rays])
    decided = hits = 0
    for r, row in zip(rays, rm):
        seg = ((r.ox, r.oy), (r.ox + r.dx * r.tmax, r.oy + r.dy * r.tmax))
        for t in tris:
            want = _sat_segment_triangle(seg, (t.x0, t.y0, t.x1, t.y1, t.x2, t.y2))
            if want is None:
                continue
            decided += 1
            hits += want
            got_py = ref.ray_triangle_any_hit_cpu((r,), (t,))[0]["any_hit"]
            tm = np.array([[t.x0, t.y0, t.x1, t.y1, t.x2, t.y2]])
            got_c = int(native.rays_2d_any(row[None], tm, brute=True)[0])
            assert got_py == got_c == want, (r, t)
    assert decided > 4000 and 50 < hits < decided - 50


def test_2d_any_hit_special_cases():
    tri = Triangle(7, 0.0, 0.0, 4.0, 0.0, 0.0, 4.0)
    cases_ = [
        (Ray2D(1, 0.5, 0.5, 1.0, 0.0, 1.0), 1),     # segment wholly inside: no edge crossing
        (Ray2D(2, 5.0, 5.0, 1.0, 1.0, 3.0), 0),     # outside, pointing away
        (Ray2D(3, 5.0, 1.0, 1.0, 0.0, -3.0), 1),    # negative tmax reaches back across the hypotenuse
        (Ray2D(4, 5.0, 1.0, 1.0, 0.0, 3.0), 0),     # the same ray forward misses
        (Ray2D(5, 1.0, 1.0, 0.0, 0.0, 0.0), 1),     # a point inside
        (Ray2D(6, -1.0, 1.0, 1.0, 0.0, 0.5), 0),    # stops before the left edge
        (Ray2D(7, -1.0, 1.0, 1.0, 0.0, 1.0), 1),    # ends on the left edge (inclusive)
        (Ray2D(8, 1.0, 1.0, 1.0, 0.0, math.inf), 1),   # origin inside, unbounded
        (Ray2D(9, -1.0, 1.0, 1.0, 0.0, math.inf), 0),  # outside origin, unbounded: no finite crossing
    ]
    for ray, want in cases_:
        assert ref.ray_triangle_any_hit_cpu((ray,), (tri,))[0]["any_hit"] == want, ray
        rm = np.array([[ray.ox, ray.oy, ray.dx, ray.dy, ray.tmax]])
        tm = np.array([[tri.x0, tri.y0, tri.x1, tri.y1, tri.x2, tri.y2]])
        for brute in (True, False):
            assert int(native.rays_2d_any(rm, tm, brute=brute)[0]) == want, (ray, brute)


@pytest.mark.parametrize("brute", [False, True])
def test_c_oracle_matches_python_oracle_rays_2d(brute):
    rng = np.random.default_rng(634)
    rays, tris = _random_2d_case(rng, 300, 400)
    perm = rng.permutation(len(tris))
    tris = tuple(Triangle(int(perm[i]) + 100, t.x0, t.y0, t.x1, t.y1, t.x2, t.y2) for i, t in enumerate(tris))
    rm = np.array([[r.ox, r.oy, r.dx, r.dy, r.tmax] for r in rays])
    tm = np.array([[t.x0, t.y0, t.x1, t.y1, t.x2, t.y2] for t in tris])
    want = [x["any_hit"] for x in ref.ray_triangle_any_hit_cpu(rays, tris)]
    assert native.rays_2d_any(rm, tm, brute=brute).tolist() == want
    assert 0 < sum(want) < len(want)
    ids = np.array([t.id for t in tris], dtype=np.uint32)
    counts = [x["hit_count"] for x in ref.ray_triangle_hit_count_cpu(rays, tris)]
    assert native.rays_2d("count", rm, tm, ids, brute=brute).tolist() == counts
    assert max(counts) > 1
    tid, tt, has = native.rays_2d("closest", rm, tm, ids, brute=brute)
    closest = ref.ray_triangle_closest_hit_cpu(rays, tris)
    assert [(x["ray_id"], x["triangle_id"], x["t"]) for x in closest] == \
        [(int(i), int(tid[i]), float(tt[i])) for i in np.flatnonzero(has)]


def test_2d_closest_hit_is_the_lowest_hit_id_at_t_zero():
    """2-D hits carry t = 0.0 (reference.py:775-780), so the (t, id) order of
    closest-hit reduces to the lowest id among the triangles hit one by one."""
    rng = np.random.default_rng(635)
    rays, tris = _random_2d_case(rng, 80, 120)
    perm = rng.permutation(len(tris))  # ids out of input order
    tris = tuple(Triangle(int(perm[i]) + 100, t.x0, t.y0, t.x1, t.y1, t.x2, t.y2) for i, t in enumerate(tris))
    rows = {x["ray_id"]: x for x in ref.ray_triangle_closest_hit_cpu(rays, tris)}
    for r in rays:
        hit_ids = [t.id for t in tris if ref.ray_triangle_any_hit_cpu((r,), (t,))[0]["any_hit"]]
        if hit_ids:
            assert rows[r.id]["triangle_id"] == min(hit_ids) and rows[r.id]["t"] == 0.0
        else:
            assert r.id not in rows


def test_visibility_authored_goal633():
    o, t = (Point(1, 0.0, 0.0),), (Point(2, 10.0, 0.0),)
    assert ref.visibility_rows_cpu(o, t, (Triangle(10, 5.0, -1.0, 5.0, 1.0, 6.0, 0.0),)) == \
        ({"observer_id": 1, "target_id": 2, "visible": 0},)
    assert ref.visibility_rows_cpu(o, t, (Triangle(11, 12.0, -1.0, 12.0, 1.0, 13.0, 0.0),)) == \
        ({"observer_id": 1, "target_id": 2, "visible": 1},)
    obs = (Point(1, 0.0, 0.0), Point(2, 0.0, 2.0))
    tgt = (Point(10, 10.0, 0.0), Point(11, 10.0, 2.0))
    assert ref.visibility_rows_cpu(obs, tgt, ()) == tuple(
        {"observer_id": a, "target_id": b, "visible": 1} for a in (1, 2) for b in (10, 11))
    o3, t3 = (Point3D(1, 0.0, 0.0, 0.0),), (Point3D(2, 10.0, 0.0, 0.0),)
    between = (Triangle3D(20, 5.0, -1.0, -1.0, 5.0, 1.0, -1.0, 5.0, 0.0, 1.0),)
    behind = (Triangle3D(21, 12.0, -1.0, -1.0, 12.0, 1.0, -1.0, 12.0, 0.0, 1.0),)
    assert ref.visibility_rows_cpu(o3, t3, between)[0]["visible"] == 0
    assert ref.visibility_rows_cpu(o3, t3, behind)[0]["visible"] == 1
    with pytest.raises(ValueError, match="one dimensionality"):
        ref.visibility_rows_cpu((Point(1, 0.0, 0.0),), (Point3D(2, 1.0, 0.0, 0.0),), ())
    with pytest.raises(ValueError, match="distinct observer and target"):
        ref.visibility_rows_cpu((Point(1, 0.0, 0.0),), (Point(2, 0.0, 0.0),), ())
    with pytest.raises(ValueError, match="blocker triangles must match"):
        ref.visibility_rows_cpu((Point(1, 0.0, 0.0),), (Point(2, 1.0, 0.0),), between)
