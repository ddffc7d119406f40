"""GPU parity of the 2-D ray any-hit path and of the visibility rows built on
it, through the C-ABI, against the oracle (bit-exact flags).

2-D any-hit follows reference.py:783-803: the ray is the segment
o -> o + d * tmax for any tmax (negative, zero, unbounded and NaN included);
it hits a triangle when an end lies inside it or the segment crosses an edge.
Visibility rows follow visibility_runtime.py:44-155 / reference.py:275-346.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import rtdl_amd as rt
from oracle import cpu_python_reference as ref
from oracle import native as oracle
from rtdl_amd import mi355x_runtime as m

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


@rt.kernel(backend="rtdl", precision="float_approx")
def anyhit_2d():
    rays = rt.input("rays", rt.Rays, layout=rt.Ray2DLayout, role="probe")
    tris = rt.input("triangles", rt.Triangles, layout=rt.Triangle2DLayout, role="build")
    c = rt.traverse(rays, tris, accel="bvh")
    h = rt.refine(c, predicate=rt.ray_triangle_any_hit(exact=False))
    return rt.emit(h, fields=["ray_id", "any_hit"])


@rt.kernel(backend="rtdl", precision="float_approx")
def count_2d():
    rays = rt.input("rays", rt.Rays, layout=rt.Ray2DLayout, role="probe")
    tris = rt.input("triangles", rt.Triangles, layout=rt.Triangle2DLayout, role="build")
    c = rt.traverse(rays, tris, accel="bvh")
    h = rt.refine(c, predicate=rt.ray_triangle_hit_count(exact=False))
    return rt.emit(h, fields=["ray_id", "hit_count"])


@rt.kernel(backend="rtdl", precision="float_approx")
def closest_2d():
    rays = rt.input("rays", rt.Rays, layout=rt.Ray2DLayout, role="probe")
    tris = rt.input("triangles", rt.Triangles, layout=rt.Triangle2DLayout, role="build")
    c = rt.traverse(rays, tris, accel="bvh")
    h = rt.refine(c, predicate=rt.ray_triangle_closest_hit(exact=False))
    return rt.emit(h, fields=["ray_id", "triangle_id", "t"])


def _tris_2d(rng, n: int, *, sigma: float = 0.002, fp32: bool = False) -> dict:
    c = rng.random((n, 1, 2))
    v = (c + rng.normal(0, sigma, (n, 3, 2))).reshape(n, 6)
    if fp32:
        v = v.astype(np.float32).astype(np.float64)
    out = {"id": rng.permutation(n).astype(np.uint32) + 7}
    for j, name in enumerate(("x0", "y0", "x1", "y1", "x2", "y2")):
        out[name] = v[:, j].copy()
    return out


def _rays_2d(rng, n: int, *, fp32: bool = False) -> dict:
    ang = rng.uniform(0, 2 * math.pi, n)
    tmax = rng.uniform(-0.01, 0.02, n)
    cols = {"ox": rng.random(n), "oy": rng.random(n), "dx": np.cos(ang), "dy": np.sin(ang), "tmax": tmax}
    if fp32:
        cols = {k: v.astype(np.float32).astype(np.float64) for k, v in cols.items()}
    return {"id": np.arange(n, dtype=np.uint32), **cols}


def _mat(cols: dict, names) -> np.ndarray:
    return np.stack([cols[k] for k in names], 1)


_RN = ("ox", "oy", "dx", "dy", "tmax")
_TN = ("x0", "y0", "x1", "y1", "x2", "y2")


def test_authored_2d_any_hit_goal632(gpu_ready):
    _say("case: authored 2-D any-hit (goal632)")
    rays = (rt.Ray2D(1, 0.0, 0.0, 1.0, 0.0, 10.0), rt.Ray2D(2, 0.0, 0.0, 0.0, 1.0, 2.0))
    tris = (rt.Triangle(10, 2.0, -1.0, 3.0, 1.0, 4.0, -1.0), rt.Triangle(11, 6.0, -1.0, 7.0, 1.0, 8.0, -1.0),
            rt.Triangle(12, -1.0, 3.0, 1.0, 3.0, 0.0, 4.0))
    got = rt.run_mi355x(anyhit_2d, rays=rays, triangles=tris)
    assert got == ({"ray_id": 1, "any_hit": 1}, {"ray_id": 2, "any_hit": 0})
    assert got == ref.ray_triangle_any_hit_cpu(rays, tris)
    counts = rt.run_mi355x(count_2d, rays=rays, triangles=tris)
    assert [r["hit_count"] for r in counts] == [2, 0]
    assert rt.run_mi355x(closest_2d, rays=rays, triangles=tris) == ({"ray_id": 1, "triangle_id": 10, "t": 0.0},)
    assert rt.run_mi355x(anyhit_2d, rays=(), triangles=tris) == ()


def _check_mode(got: np.ndarray, want, mode: str, ray_ids: np.ndarray) -> None:
    if mode == "any":
        assert np.array_equal(got["ray_id"], ray_ids)
        assert np.array_equal(got["any_hit"], want), int((got["any_hit"] != want).sum())
    elif mode == "count":
        assert np.array_equal(got["ray_id"], ray_ids)
        assert np.array_equal(got["hit_count"], want)
    else:
        tid, tt, has = want
        assert np.array_equal(got["ray_id"], ray_ids[has])
        assert np.array_equal(got["triangle_id"], tid[has])
        assert np.array_equal(got["t"], tt[has])


_API = {"any": "any_hit", "count": "hit_count", "closest": "closest_hit"}


@pytest.mark.parametrize("mode", ["count", "closest"])
def test_random_2d_count_and_closest_match_the_oracle(gpu_ready, mode):
    _say(f"case: random 2-D {mode} 30000 rays x 40000 triangles")
    rng = np.random.default_rng(6325)
    t = _tris_2d(rng, 40000)
    r = _rays_2d(rng, 30000)
    r["tmax"][:4] = np.inf
    r["tmax"][4:6] = np.nan
    got = m.ray_triangle_mi355x(r, t, mode=_API[mode])
    want = oracle.rays_2d(mode, _mat(r, _RN), _mat(t, _TN), t["id"])
    _check_mode(got, want, mode, r["id"])
    if mode == "count":
        assert int(want.max()) > 3


@pytest.mark.parametrize("fp32", [False, True])
def test_random_2d_any_hit_matches_the_oracle(gpu_ready, fp32):
    _say(f"case: random 2-D any-hit 60000 rays x 40000 triangles (fp32 inputs: {fp32})")
    rng = np.random.default_rng(6320 + int(fp32))
    t = _tris_2d(rng, 40000, fp32=fp32)
    r = _rays_2d(rng, 60000, fp32=fp32)
    # edge cases: zero length, unbounded, NaN, overflowing end, origin on a vertex
    r["tmax"][:8] = 0.0
    r["tmax"][8:12] = np.inf
    r["tmax"][12:14] = np.nan
    r["tmax"][14:16] = 1e308
    r["ox"][16:24], r["oy"][16:24] = t["x0"][:8], t["y0"][:8]
    got = m.ray_triangle_mi355x(r, t, mode="any_hit")
    want = oracle.rays_2d_any(_mat(r, _RN), _mat(t, _TN))
    assert np.array_equal(got["ray_id"], r["id"])
    assert np.array_equal(got["any_hit"], want), int((got["any_hit"] != want).sum())
    assert 1000 < int(want.sum()) < len(want) - 1000


def test_2d_any_hit_empty_and_single_inputs(gpu_ready):
    _say("case: 2-D any-hit empty / single")
    tri = {"id": np.array([3], np.uint32), "x0": [0.0], "y0": [0.0], "x1": [1.0], "y1": [0.0],
           "x2": [0.0], "y2": [1.0]}
    tri = {k: np.asarray(v) for k, v in tri.items()}
    empty_r = {k: np.zeros(0) for k in _RN} | {"id": np.zeros(0, np.uint32)}
    empty_t = {k: np.zeros(0) for k in _TN} | {"id": np.zeros(0, np.uint32)}
    assert len(m.ray_triangle_mi355x(empty_r, tri)) == 0
    one = {"id": np.array([9], np.uint32), "ox": np.array([0.2]), "oy": np.array([0.2]),
           "dx": np.array([1.0]), "dy": np.array([0.0]), "tmax": np.array([0.1])}
    assert m.ray_triangle_mi355x(one, empty_t).tolist() == [(9, 0)]
    assert m.ray_triangle_mi355x(one, tri).tolist() == [(9, 1)]


def test_prepared_2d_any_hit_serves_several_batches(gpu_ready):
    _say("case: prepared 2-D any-hit, 3 batches")
    rng = np.random.default_rng(6330)
    t = _tris_2d(rng, 20000)
    tris = m.pack_triangles_2d(t)
    with rt.prepare_mi355x(anyhit_2d, result_mode="raw", triangles=tris) as prepared:
        for b in range(3):
            r = _rays_2d(rng, 5000 + 1000 * b)
            got = prepared.run(rays=m.pack_rays_2d(r))
            want = oracle.rays_2d_any(_mat(r, _RN), _mat(t, _TN))
            assert np.array_equal(got["any_hit"], want)
            assert np.array_equal(got, m.ray_triangle_mi355x(r, t))
    for mode, kernel in (("count", count_2d), ("closest", closest_2d)):
        with rt.prepare_mi355x(kernel, result_mode="raw", triangles=tris) as prepared:
            r = _rays_2d(rng, 4000)
            got = prepared.run(rays=m.pack_rays_2d(r))
            _check_mode(got, oracle.rays_2d(mode, _mat(r, _RN), _mat(t, _TN), t["id"]), mode, r["id"])


def _random_points(rng, n: int, dim: int, first_id: int):
    p = rng.random((n, dim))
    if dim == 2:
        return tuple(rt.Point(first_id + i, float(x), float(y)) for i, (x, y) in enumerate(p))
    return tuple(rt.Point3D(first_id + i, float(x), float(y), float(z)) for i, (x, y, z) in enumerate(p))


def _random_blockers(rng, n: int, dim: int, sigma: float = 0.06):
    c = rng.random((n, 1, dim))
    v = (c + rng.normal(0, sigma, (n, 3, dim))).reshape(n, 3 * dim)
    if dim == 2:
        return tuple(rt.Triangle(500 + i, *map(float, row)) for i, row in enumerate(v))
    return tuple(rt.Triangle3D(500 + i, *map(float, row)) for i, row in enumerate(v))


def test_visibility_authored_goal633(gpu_ready):
    _say("case: visibility rows authored (goal633)")
    o, t = (rt.Point(1, 0.0, 0.0),), (rt.Point(2, 10.0, 0.0),)
    assert rt.visibility_rows(o, t, (rt.Triangle(10, 5.0, -1.0, 5.0, 1.0, 6.0, 0.0),)) == \
        ({"observer_id": 1, "target_id": 2, "visible": 0},)
    assert rt.visibility_rows(o, t, (rt.Triangle(11, 12.0, -1.0, 12.0, 1.0, 13.0, 0.0),)) == \
        ({"observer_id": 1, "target_id": 2, "visible": 1},)
    obs = (rt.Point(1, 0.0, 0.0), rt.Point(2, 0.0, 2.0))
    tgt = (rt.Point(10, 10.0, 0.0), rt.Point(11, 10.0, 2.0))
    assert rt.visibility_rows(obs, tgt, ()) == tuple(
        {"observer_id": a, "target_id": b, "visible": 1} for a in (1, 2) for b in (10, 11))
    assert rt.visibility_pair_rows(((obs[0], tgt[0]), (obs[1], tgt[1])), ()) == (
        {"observer_id": 1, "target_id": 10, "visible": 1}, {"observer_id": 2, "target_id": 11, "visible": 1})
    o3, t3 = (rt.Point3D(1, 0.0, 0.0, 0.0),), (rt.Point3D(2, 10.0, 0.0, 0.0),)
    between = (rt.Triangle3D(20, 5.0, -1.0, -1.0, 5.0, 1.0, -1.0, 5.0, 0.0, 1.0),)
    behind = (rt.Triangle3D(21, 12.0, -1.0, -1.0, 12.0, 1.0, -1.0, 12.0, 0.0, 1.0),)
    assert rt.visibility_rows(o3, t3, between)[0]["visible"] == 0
    assert rt.visibility_rows(o3, t3, behind)[0]["visible"] == 1


@pytest.mark.parametrize("dim", [2, 3])
def test_visibility_rows_match_the_oracle(gpu_ready, dim):
    _say(f"case: visibility rows {dim}-D, 40 x 50 pairs")
    rng = np.random.default_rng(6340 + dim)
    obs = _random_points(rng, 40, dim, 1)
    tgt = _random_points(rng, 50, dim, 1000)
    blockers = _random_blockers(rng, 30 if dim == 2 else 100, dim, sigma=0.03 if dim == 2 else 0.1)
    got = rt.visibility_rows(obs, tgt, blockers)
    want = ref.visibility_rows_cpu(obs, tgt, blockers)
    assert got == want
    vis = sum(r["visible"] for r in want)
    assert 100 < vis < len(want) - 100
    raw = rt.visibility_rows(obs, tgt, blockers, result_mode="raw")
    assert [r["visible"] for r in want] == raw["visible"].tolist()


@pytest.mark.parametrize("dim", [2, 3])
def test_visibility_pair_rows_match_the_oracle(gpu_ready, dim):
    _say(f"case: visibility pair rows {dim}-D, 1500 pairs")
    rng = np.random.default_rng(6350 + dim)
    obs = _random_points(rng, 1500, dim, 1)
    tgt = _random_points(rng, 1500, dim, 5000)
    blockers = _random_blockers(rng, 30 if dim == 2 else 100, dim, sigma=0.03 if dim == 2 else 0.1)
    pairs = tuple(zip(obs, tgt))
    got = rt.visibility_pair_rows(pairs, blockers)
    hits = ref.ray_triangle_any_hit_cpu(
        tuple(ref._visibility_ray(i, a, b) for i, (a, b) in enumerate(pairs)), blockers)
    want = tuple({"observer_id": a.id, "target_id": b.id, "visible": 1 - h["any_hit"]}
                 for (a, b), h in zip(pairs, hits))
    assert got == want


def test_reference_shaped_prepared_any_hit_2d(gpu_ready):
    """rtdl_mi355x_prepare_ray_anyhit_2d / run_prepared_ray_anyhit_2d /
    destroy_prepared_ray_anyhit_2d with the HIPRT parameter lists
    (src/native/hiprt/rtdl_hiprt_api.cpp:81-181), against the oracle."""
    _say("case: reference-shaped prepared 2-D any-hit")
    rng = np.random.default_rng(6331)
    t = _tris_2d(rng, 15000)
    tris = m.pack_triangles_2d(t)
    handle = m._call_prepare("rtdl_mi355x_prepare_ray_anyhit_2d", m._ptr(tris), len(tris))
    try:
        for b in range(2):
            r = _rays_2d(rng, 3000 + 2000 * b)
            rays = m.pack_rays_2d(r)
            got = m._call_rows("rtdl_mi355x_run_prepared_ray_anyhit_2d", m.ANYHIT_ROW_DTYPE, handle,
                               m._ptr(rays), len(rays))
            assert np.array_equal(got["ray_id"], r["id"])
            assert np.array_equal(got["any_hit"], oracle.rays_2d_any(_mat(r, _RN), _mat(t, _TN)))
    finally:
        m._library().rtdl_mi355x_destroy_prepared_ray_anyhit_2d(handle)
