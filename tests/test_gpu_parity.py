"""GPU parity: the HIP backend against the oracle, through the C-ABI.

Neighbour ids, ranks, hit flags, counts and PIP rows must be bit-exact after
the reference's canonical ordering; FP64 distances and hit parameters are
compared exactly too (the contract's tolerance is 1e-6 relative, the backend
computes the reference's FP64 expression so it reproduces it bit for bit).
"""
from __future__ import annotations

import numpy as np
import pytest

import rtdl_amd as rt
from rtdl_amd import mi355x_runtime as m
from oracle import cpu_python_reference as ref
from oracle import native as oracle
from tests import cases

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


def _assert_nn_equal(got: np.ndarray, want: np.ndarray, *, ranked: bool) -> None:
    assert len(got) == len(want), (len(got), len(want))
    assert np.array_equal(got["query_id"], want["qid"])
    assert np.array_equal(got["neighbor_id"], want["nid"])
    assert np.array_equal(got["distance"], want["dist"])
    if ranked:
        assert np.array_equal(got["neighbor_rank"], want["rank"])


# ---------------------------------------------------------------- FP64 --


def test_device_fp64_sqrt_and_division_are_correctly_rounded(gpu_ready):
    _say("case: fp64 selftest")
    rng = np.random.default_rng(11)
    a = np.concatenate([
        rng.random(200000), rng.random(200000) * 1e-300, np.exp(rng.normal(0, 30, 200000)),
        np.arange(1, 20001, dtype=np.float64) ** 2,
        np.nextafter(np.arange(1, 20001, dtype=np.float64) ** 2, 0.0),
        np.nextafter(np.arange(1, 20001, dtype=np.float64) ** 2, np.inf),
    ])
    b = np.concatenate([rng.normal(size=len(a) - 1000) * 1e3, np.full(1000, 3.0)])
    s, d = m.selftest_fp64(a, b)
    assert np.array_equal(s, np.sqrt(a))
    assert np.array_equal(d, a / b)


# ---------------------------------------------- fixed-radius neighbours --


def test_authored_fixed_radius_case_matches_reference(gpu_ready):
    _say("case: authored fixed-radius 2-D (goal198)")

    @rt.kernel(backend="rtdl", precision="float_approx")
    def frn():
        q = rt.input("query_points", rt.Points, role="probe")
        s = rt.input("search_points", rt.Points, role="build")
        c = rt.traverse(q, s, accel="bvh")
        h = rt.refine(c, predicate=rt.fixed_radius_neighbors(radius=0.5, k_max=3))
        return rt.emit(h, fields=["query_id", "neighbor_id", "distance"])

    case = {
        "query_points": (rt.Point(100, 0.0, 0.0), rt.Point(101, 3.0, 0.0)),
        "search_points": (rt.Point(1, 0.0, 0.0), rt.Point(2, 0.3, 0.0), rt.Point(3, -0.3, 0.0),
                          rt.Point(4, 3.2, 0.0), rt.Point(5, 4.0, 0.0)),
    }
    got = rt.run_mi355x(frn, **case)
    assert got == ref.run_cpu_python_reference(frn, **case)
    assert [r["neighbor_id"] for r in got] == [1, 2, 3, 4]


def test_authored_fixed_radius_3d_case(gpu_ready):
    _say("case: authored fixed-radius 3-D (goal548)")

    @rt.kernel(backend="rtdl", precision="float_approx")
    def frn3():
        q = rt.input("query_points", rt.Points3D, layout=rt.Point3DLayout, role="probe")
        s = rt.input("search_points", rt.Points3D, layout=rt.Point3DLayout, role="build")
        c = rt.traverse(q, s, accel="bvh")
        h = rt.refine(c, predicate=rt.fixed_radius_neighbors(radius=1.0, k_max=3))
        return rt.emit(h, fields=["query_id", "neighbor_id", "distance"])

    case = {
        "query_points": (rt.Point3D(10, 0.0, 0.0, 0.0), rt.Point3D(20, 5.0, 0.0, 0.0)),
        "search_points": (rt.Point3D(1, 0.0, 0.0, 0.0), rt.Point3D(2, 0.0, 0.0, 0.6),
                          rt.Point3D(3, 0.0, 0.8, 0.0), rt.Point3D(4, 5.0, 0.0, 0.9)),
    }
    assert rt.run_mi355x(frn3, **case) == ref.run_cpu_python_reference(frn3, **case)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("radius,k_max", [(0.03, 1), (0.05, 3), (0.05, 16), (0.08, 64), (0.06, 100),
                                          (0.02, 4096), (0.0, 5)])
def test_fixed_radius_random_bit_exact(gpu_ready, dim, fp32, radius, k_max):
    _say(f"case: frn dim={dim} fp32={fp32} r={radius} k_max={k_max}")
    rng = np.random.default_rng(1000 + dim * 10 + int(fp32) + k_max)
    q = cases.points(rng, 3000, dim, fp32=fp32)
    s = cases.points(rng, 4000, dim, fp32=fp32)
    s_idx = rng.integers(0, 4000, 50)
    for key in ("x", "y", "z"):
        if key in q:
            q[key][:50] = s[key][s_idx]  # coincident points (distance 0)
    got = m.fixed_radius_neighbors_mi355x(q, s, radius=radius, k_max=k_max)
    want = oracle.neighbors("frn", q, s, radius=radius, k=k_max)
    _assert_nn_equal(got, want, ranked=False)


def test_fixed_radius_gaussian_mixture_large_k_truncates_like_reference(gpu_ready):
    _say("case: frn gaussian mixture truncation k_max=8")
    rng = np.random.default_rng(7)
    centres = rng.random((8, 3))
    q = cases.points(rng, 20000, 3, kind="mixture", centres=centres)
    s = cases.points(rng, 30000, 3, kind="mixture", centres=centres)
    got = m.fixed_radius_neighbors_mi355x(q, s, radius=0.05, k_max=8)
    want = oracle.neighbors("frn", q, s, radius=0.05, k=8)
    _assert_nn_equal(got, want, ranked=False)


def test_fixed_radius_heavy_queries_beyond_64(gpu_ready):
    _say("case: frn heavy path (k_max 500, dense)")
    rng = np.random.default_rng(8)
    q = cases.points(rng, 500, 2)
    s = cases.points(rng, 20000, 2)
    got = m.fixed_radius_neighbors_mi355x(q, s, radius=0.1, k_max=500)
    want = oracle.neighbors("frn", q, s, radius=0.1, k=500)
    _assert_nn_equal(got, want, ranked=False)


def test_rows_grouped_by_query_id_stably(gpu_ready):
    _say("case: frn unsorted and duplicate query ids")
    rng = np.random.default_rng(9)
    ids = rng.integers(0, 300, 3000).astype(np.uint32)  # duplicates, unsorted
    q = cases.points(rng, 3000, 3, ids=ids)
    s = cases.points(rng, 3000, 3, ids=rng.permutation(3000).astype(np.uint32))
    got = m.fixed_radius_neighbors_mi355x(q, s, radius=0.07, k_max=10)
    want = oracle.neighbors("frn", q, s, radius=0.07, k=10)
    _assert_nn_equal(got, want, ranked=False)


def test_fixed_radius_ties_order_by_lower_id(gpu_ready):
    _say("case: frn exact distance ties")
    # every search point sits on a lattice, so many distances tie exactly
    g = np.arange(0, 20, dtype=np.float64) * 0.25
    xx, yy = np.meshgrid(g, g)
    s = {"id": np.arange(xx.size, dtype=np.uint32)[::-1].copy(), "x": xx.ravel(), "y": yy.ravel()}
    q = {"id": np.arange(40, dtype=np.uint32), "x": np.repeat(g[::2], 4)[:40], "y": np.tile(g[:4], 10)}
    got = m.fixed_radius_neighbors_mi355x(q, s, radius=0.6, k_max=5)
    want = oracle.neighbors("frn", q, s, radius=0.6, k=5)
    _assert_nn_equal(got, want, ranked=False)


def test_empty_inputs(gpu_ready):
    _say("case: empty inputs")
    rng = np.random.default_rng(3)
    q = cases.points(rng, 10, 3)
    empty = cases.points(rng, 0, 3)
    assert len(m.fixed_radius_neighbors_mi355x(q, empty, radius=0.1, k_max=3)) == 0
    assert len(m.fixed_radius_neighbors_mi355x(empty, q, radius=0.1, k_max=3)) == 0
    assert len(m.knn_rows_mi355x(q, empty, k=3)) == 0
    r = cases.rays(rng, 5)
    t0 = cases.triangles(rng, 0)
    rows = m.ray_triangle_mi355x(r, t0, mode="any_hit")
    assert rows["ray_id"].tolist() == list(range(5)) and rows["any_hit"].tolist() == [0] * 5


# ------------------------------------------------------------------ KNN --


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("k", [1, 5, 16, 64])
def test_knn_random_bit_exact(gpu_ready, dim, k):
    _say(f"case: knn dim={dim} k={k}")
    rng = np.random.default_rng(2000 + dim + k)
    n = 2500
    q = cases.points(rng, n, dim, kind="mixture")
    s = cases.points(rng, n + 500, dim, kind="mixture")
    got = m.knn_rows_mi355x(q, s, k=k)
    want = oracle.neighbors("knn", q, s, k=k)
    _assert_nn_equal(got, want, ranked=True)


@pytest.mark.parametrize("k", [5, 13, 64])
def test_knn_lattice_ties_bit_exact(gpu_ready, k):
    """Points on a 1/16 lattice (exact in FP32 and FP64) and queries on the
    lattice and half-lattice: many neighbours share one distance, so the
    (distance, id) tie rule and the k-th-distance boundary decide the rows."""
    _say(f"case: knn lattice ties k={k}")
    rng = np.random.default_rng(3000 + k)
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(12), indexing="ij"), -1).reshape(-1, 3)
    ids = rng.permutation(len(g)).astype(np.uint32)  # ids unrelated to position
    s = {"id": ids, "x": g[:, 0] / 16.0, "y": g[:, 1] / 16.0, "z": g[:, 2] / 16.0}
    qg = rng.integers(0, 24, (900, 3)) / 32.0
    q = {"id": np.arange(len(qg), dtype=np.uint32), "x": qg[:, 0], "y": qg[:, 1], "z": qg[:, 2]}
    got = m.knn_rows_mi355x(q, s, k=k)
    want = oracle.neighbors("knn", q, s, k=k)
    _assert_nn_equal(got, want, ranked=True)


def test_knn_k_above_64_is_rejected_by_the_library(gpu_ready):
    """An unbounded search beyond the 64-entry list would gather every scene
    point per query; the C-ABI rejects it with a status, as HIPRT does."""
    _say("case: knn k > 64 rejected in the library")
    rng = np.random.default_rng(7)
    q = m.pack_points(cases.points(rng, 10, 3))[0]
    s = m.pack_points(cases.points(rng, 100, 3))[0]
    with pytest.raises(RuntimeError, match="k <= 64"):
        m._call_rows("rtdl_mi355x_run_knn_rows_3d", m.KNN_ROW_DTYPE, m._ptr(q), len(q), m._ptr(s), len(s), 80)


@pytest.mark.parametrize("k", [80, 200])
def test_bounded_knn_large_k_bit_exact(gpu_ready, k):
    """k_max above the 64-entry list with a finite radius: the radius bounds
    the full collection of the heavy queries."""
    _say(f"case: bounded knn k_max={k}")
    rng = np.random.default_rng(2100 + k)
    q = cases.points(rng, 1500, 3, kind="mixture")
    s = cases.points(rng, 2500, 3, kind="mixture")
    got = m.fixed_radius_neighbors_mi355x(q, s, radius=0.08, k_max=k, bounded_rank=True)
    want = oracle.neighbors("bounded_knn", q, s, radius=0.08, k=k)
    _assert_nn_equal(got, want, ranked=True)


def test_knn_short_result_is_not_padded(gpu_ready):
    _say("case: knn short result")
    q = {"id": np.array([10], np.uint32), "x": np.array([1.0]), "y": np.array([1.0])}
    s = {"id": np.array([1], np.uint32), "x": np.array([1.0]), "y": np.array([1.0])}
    rows = m.knn_rows_mi355x(q, s, k=4)
    assert rows.tolist() == [(10, 1, 0.0, 1)]


def test_bounded_knn_matches_reference(gpu_ready):
    _say("case: bounded knn")
    rng = np.random.default_rng(5)
    q = cases.points(rng, 2000, 3)
    s = cases.points(rng, 3000, 3)
    got = m.fixed_radius_neighbors_mi355x(q, s, radius=0.1, k_max=6, bounded_rank=True)
    want = oracle.neighbors("bounded_knn", q, s, radius=0.1, k=6)
    _assert_nn_equal(got, want, ranked=True)


# ----------------------------------------------------------------- rays --


def test_authored_3d_any_hit(gpu_ready):
    _say("case: authored 3-D any-hit (goal632)")

    @rt.kernel(backend="rtdl", precision="float_approx")
    def anyhit():
        r = rt.input("rays", rt.Rays3D, layout=rt.Ray3DLayout, role="probe")
        t = rt.input("triangles", rt.Triangles3D, layout=rt.Triangle3DLayout, role="build")
        c = rt.traverse(r, t, accel="bvh")
        h = rt.refine(c, predicate=rt.ray_triangle_any_hit(exact=False))
        return rt.emit(h, fields=["ray_id", "any_hit"])

    case = {
        "rays": (rt.Ray3D(1, 0.0, 0.0, 0.0, 10.0, 0.0, 0.0, 1.0),
                 rt.Ray3D(2, 0.0, 2.0, 0.0, 10.0, 0.0, 0.0, 1.0)),
        "triangles": (rt.Triangle3D(20, 5.0, -1.0, -1.0, 5.0, 1.0, -1.0, 5.0, 0.0, 1.0),),
    }
    got = rt.run_mi355x(anyhit, **case)
    assert got == ({"ray_id": 1, "any_hit": 1}, {"ray_id": 2, "any_hit": 0})
    assert got == ref.run_cpu_python_reference(anyhit, **case)


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("mode", ["any", "count", "closest"])
def test_rays_random_bit_exact(gpu_ready, fp32, mode):
    _say(f"case: rays mode={mode} fp32={fp32}")
    rng = np.random.default_rng(3000 + int(fp32))
    t = cases.triangles(rng, 20000, sigma=0.01, fp32=fp32)
    r = cases.rays(rng, 8000, tmax=0.3, fp32=fp32)
    want = oracle.rays_3d(mode, cases.ray_matrix(r), cases.tri_matrix(t), t["id"])
    api_mode = {"any": "any_hit", "count": "hit_count", "closest": "closest_hit"}[mode]
    got = m.ray_triangle_mi355x(r, t, mode=api_mode)
    if mode == "any":
        assert np.array_equal(got["ray_id"], r["id"])
        assert np.array_equal(got["any_hit"], want)
    elif mode == "count":
        assert np.array_equal(got["hit_count"], want)
    else:
        tid, tt, has = want
        assert np.array_equal(got["ray_id"], r["id"][has])
        assert np.array_equal(got["triangle_id"], tid[has])
        assert np.array_equal(got["t"], tt[has])


@pytest.mark.parametrize("mode", ["any", "count", "closest"])
def test_unbounded_rays_bit_exact(gpu_ready, mode):
    """tmax = inf: the walk bounds the ray by the scene box, the FP64 test
    still takes every t >= 0.  Origins inside and outside the scene, on a
    scene scaled and shifted away from the origin."""
    _say(f"case: unbounded rays mode={mode}")
    rng = np.random.default_rng(3100)
    t = cases.triangles(rng, 20000, sigma=0.01, fp32=False)
    for f in cases.TRI_FIELDS:
        t[f] = t[f] * 1000.0 + 250.0
    r = cases.rays(rng, 6000, tmax=np.inf, fp32=False)
    r["tmax"][:] = np.inf
    for f, base in (("ox", 250.0), ("oy", 250.0), ("oz", 250.0)):
        r[f] = r[f] * 1000.0 + base
    r["ox"][:2000] = -3000.0  # outside the scene, some pointing at it
    r["dx"][:1000] = np.abs(r["dx"][:1000]) + 0.5
    want = oracle.rays_3d(mode, cases.ray_matrix(r), cases.tri_matrix(t), t["id"])
    api_mode = {"any": "any_hit", "count": "hit_count", "closest": "closest_hit"}[mode]
    got = m.ray_triangle_mi355x(r, t, mode=api_mode)
    if mode == "any":
        assert np.array_equal(got["any_hit"], want)
        assert 500 < int(want.sum()) < len(want) - 500
    elif mode == "count":
        assert np.array_equal(got["hit_count"], want)
    else:
        tid, tt, has = want
        assert np.array_equal(got["ray_id"], r["id"][has])
        assert np.array_equal(got["triangle_id"], tid[has])
        assert np.array_equal(got["t"], tt[has])


def test_ray_edge_cases(gpu_ready):
    _say("case: ray edge cases (axis-aligned, zero tmax, negative tmax, edge hits)")
    tris = {"id": np.array([7, 8], np.uint32)}
    tv = np.array([[0, 0, 1, 1, 0, 1, 0, 1, 1], [0, 0, 2, 2, 0, 2, 0, 2, 2]], dtype=np.float64)
    for i, f in enumerate(cases.TRI_FIELDS):
        tris[f] = tv[:, i]
    rv = np.array([
        [0.25, 0.25, 0, 0, 0, 1, 5],     # hits both
        [0.0, 0.0, 0, 0, 0, 1, 5],       # through a shared vertex
        [0.5, 0.5, 0, 0, 0, 1, 5],       # through an edge
        [0.25, 0.25, 0, 0, 0, 1, 1.0],   # t == tmax exactly on the first
        [0.25, 0.25, 0, 0, 0, 1, 0.999], # stops short
        [0.25, 0.25, 0, 0, 0, 1, -1.0],  # negative tmax
        [0.25, 0.25, 1, 0, 0, 0, 5],     # parallel to the triangles (det = 0)
        [0.25, 0.25, 3, 0, 0, -1, 5],    # backwards direction
    ], dtype=np.float64)
    rays = {"id": np.arange(len(rv), dtype=np.uint32)}
    for i, f in enumerate(cases.RAY_FIELDS):
        rays[f] = rv[:, i]
    for mode, api_mode in (("any", "any_hit"), ("count", "hit_count"), ("closest", "closest_hit")):
        want = oracle.rays_3d(mode, rv, tv, tris["id"], brute=True)
        got = m.ray_triangle_mi355x(rays, tris, mode=api_mode)
        if mode == "closest":
            tid, tt, has = want
            assert got["triangle_id"].tolist() == tid[has].tolist()
            assert got["t"].tolist() == tt[has].tolist()
        else:
            assert got[got.dtype.names[1]].tolist() == want.tolist()


# ------------------------------------------------------------------ PIP --


def test_authored_pip_squares(gpu_ready):
    _say("case: authored PIP squares (goal552)")

    @rt.kernel(backend="rtdl", precision="float_approx")
    def pip():
        p = rt.input("points", rt.Points, role="probe")
        g = rt.input("polygons", rt.Polygons, role="build")
        c = rt.traverse(p, g, accel="bvh")
        h = rt.refine(c, predicate=rt.point_in_polygon(exact=False))
        return rt.emit(h, fields=["point_id", "polygon_id", "contains"])

    def square(pid, x0, y0, size):
        return rt.Polygon(pid, ((x0, y0), (x0 + size, y0), (x0 + size, y0 + size), (x0, y0 + size)))

    case = {
        "points": (rt.Point(1, 0.0, 0.0), rt.Point(2, 0.5, 0.5), rt.Point(3, 3.0, 3.0)),
        "polygons": (square(10, -0.25, -0.25, 1.0), square(20, 2.0, 2.0, 1.0)),
    }
    assert rt.run_mi355x(pip, **case) == ref.run_cpu_python_reference(pip, **case)


@pytest.mark.parametrize("positive_only", [True, False])
def test_pip_random_star_polygons(gpu_ready, positive_only):
    _say(f"case: pip star polygons positive_only={positive_only}")
    rng = np.random.default_rng(4000 + int(positive_only))
    polys = cases.star_polygons(rng, 1000 if positive_only else 200, rmin=0.01, rmax=0.05)
    n = 20000 if positive_only else 3000
    pts = cases.points(rng, n, 2)
    # put some probes exactly on vertices, on edge midpoints, and just off
    # edges inside and outside the 1e-12 on-edge tolerance
    v = polys["vertices_xy"].reshape(-1, 2)
    pts["x"][:100] = v[:100, 0]
    pts["y"][:100] = v[:100, 1]
    refs = polys["refs"]
    a = v[refs["vertex_offset"][:100]]
    b = v[refs["vertex_offset"][:100] + 1]
    mid = (a + b) / 2
    nrm = np.stack([a[:, 1] - b[:, 1], b[:, 0] - a[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    for j, off in enumerate((0.0, 5e-13, -5e-13, 3e-12, -3e-12)):
        sl = slice(100 + 100 * j, 200 + 100 * j)
        pts["x"][sl] = mid[:, 0] + off * nrm[:, 0]
        pts["y"][sl] = mid[:, 1] + off * nrm[:, 1]
    pairs = oracle.pip_pairs(pts["x"], pts["y"], refs["vertex_offset"], refs["vertex_count"],
                             polys["vertices_xy"], positive_only=positive_only)
    got = m.point_in_polygon_mi355x(pts, polys, positive_only=positive_only)
    if positive_only:
        assert got["point_id"].tolist() == pts["id"][pairs[:, 0]].tolist()
        assert got["polygon_id"].tolist() == refs["id"][pairs[:, 1]].tolist()
        assert np.all(got["contains"] == 1)
    else:
        full = np.zeros((n, len(refs)), dtype=np.uint32)
        full[pairs[:, 0], pairs[:, 1]] = 1
        assert len(got) == n * len(refs)
        assert np.array_equal(got["contains"], full.ravel())
        assert np.array_equal(got["point_id"], np.repeat(pts["id"], len(refs)))


# ------------------------------------------------------------- prepared --


def test_prepared_batches_match_per_call(gpu_ready):
    _say("case: prepared scene, several batches")

    @rt.kernel(backend="rtdl", precision="float_approx")
    def knn():
        q = rt.input("query_points", rt.Points3D, layout=rt.Point3DLayout, role="probe")
        s = rt.input("search_points", rt.Points3D, layout=rt.Point3DLayout, role="build")
        c = rt.traverse(q, s, accel="bvh")
        h = rt.refine(c, predicate=rt.knn_rows(k=8))
        return rt.emit(h, fields=["query_id", "neighbor_id", "distance", "neighbor_rank"])

    rng = np.random.default_rng(6)
    s = cases.points(rng, 50000, 3, kind="mixture")
    with rt.prepare_mi355x(knn, result_mode="raw", search_points=s) as prepared:
        for batch in range(3):
            q = cases.points(rng, 3000 + 1000 * batch, 3, kind="mixture")
            got = prepared.run(query_points=q)
            want = oracle.neighbors("knn", q, s, k=8)
            _assert_nn_equal(got, want, ranked=True)


def test_concurrent_threads_on_two_prepared_scenes(gpu_ready):
    """Two threads run per-call and prepared neighbour queries at once (the
    C-ABI releases the GIL).  Per-call temporaries come from one process-wide
    pool, so a block handed back while its stream still used it would show
    here as wrong rows in either thread."""
    import threading

    _say("case: concurrent per-call and prepared neighbour queries on two threads")
    rng = np.random.default_rng(77)
    work = []
    for t in range(2):
        q = cases.points(rng, 6000, 3, kind="mixture")
        s = cases.points(rng, 9000, 3, kind="mixture")
        work.append((q, s, oracle.neighbors("knn", q, s, k=16),
                     oracle.neighbors("frn", q, s, radius=0.03, k=100)))
    errors = []

    def run(t):
        q, s, want_knn, want_frn = work[t]
        arr, _ = m.pack_points(s)
        qa, _ = m.pack_points(q)
        handle = m._call_prepare("rtdl_mi355x_prepare_points_3d", m._ptr(arr), len(arr))
        try:
            for rep in range(4):
                got = m._call_rows("rtdl_mi355x_run_prepared_knn_rows", m.KNN_ROW_DTYPE, handle,
                                   m._ptr(qa), len(qa), 16)
                _assert_nn_equal(got, want_knn, ranked=True)
                got = m._call_rows("rtdl_mi355x_run_prepared_fixed_radius_neighbors", m.FRN_ROW_DTYPE,
                                   handle, m._ptr(qa), len(qa), 0.03, 100)
                _assert_nn_equal(got, want_frn, ranked=False)
                _assert_nn_equal(m.knn_rows_mi355x(q, s, k=16), want_knn, ranked=True)
        except Exception as exc:  # noqa: BLE001 - reported by the main thread
            errors.append((t, repr(exc)))
        finally:
            m._library().rtdl_mi355x_destroy_prepared(handle)

    threads = [threading.Thread(target=run, args=(t,)) for t in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=240)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors


def test_reference_shaped_prepared_fixed_radius_and_hitcount(gpu_ready):
    """The HIPRT-shaped prepared calls: fixed radius with the radius bound at
    prepare time and k_max per run (src/native/hiprt/rtdl_hiprt_api.cpp:439-486),
    and 3-D hit count (:15-79), each against the oracle over several runs."""
    _say("case: reference-shaped prepared fixed radius and 3-D hit count")
    rng = np.random.default_rng(4390)
    s = cases.points(rng, 20000, 3, kind="mixture", fp32=False)
    arr, _ = m.pack_points(s)
    lib = m._library()
    handle = m._call_prepare("rtdl_mi355x_prepare_fixed_radius_neighbors_3d", m._ptr(arr), len(arr), 0.03)
    try:
        for k_max in (1, 16, 64, 150):
            q = cases.points(rng, 4000, 3, kind="mixture", fp32=False)
            qa, _ = m.pack_points(q)
            got = m._call_rows("rtdl_mi355x_run_prepared_fixed_radius_neighbors_3d", m.FRN_ROW_DTYPE, handle,
                               m._ptr(qa), len(qa), k_max)
            _assert_nn_equal(got, oracle.neighbors("frn", q, s, radius=0.03, k=k_max), ranked=False)
    finally:
        lib.rtdl_mi355x_destroy_prepared_fixed_radius_neighbors_3d(handle)
    t = cases.triangles(rng, 30000, sigma=0.01, fp32=False)
    tris = m.pack_triangles(t)
    handle = m._call_prepare("rtdl_mi355x_prepare_ray_hitcount_3d", m._ptr(tris), len(tris))
    try:
        for n in (2000, 7000):
            r = cases.rays(rng, n, tmax=0.4, fp32=False)
            rays = m.pack_rays(r)
            got = m._call_rows("rtdl_mi355x_run_prepared_ray_hitcount_3d", m.HITCOUNT_ROW_DTYPE, handle,
                               m._ptr(rays), len(rays))
            want = oracle.rays_3d("count", cases.ray_matrix(r), cases.tri_matrix(t), t["id"])
            assert np.array_equal(got["ray_id"], r["id"])
            assert np.array_equal(got["hit_count"], want)
    finally:
        lib.rtdl_mi355x_destroy_prepared_ray_hitcount_3d(handle)
