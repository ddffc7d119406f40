"""The allocation-free batch entry point (rtdl_mi355x_run_batch_noalloc)
against rtdl_mi355x_run_batch: the same jobs of every kind, two calls in a
row (the second smaller, so stale scratch contents of the first would show),
rows compared bit for bit after canonical ordering; batches beyond the limits
are rejected by the host before any device work."""
from __future__ import annotations

import numpy as np
import pytest

from rtdl_amd import batch as B
from rtdl_amd.device import DeviceBuffer
from tests import cases

pytestmark = pytest.mark.gpu

_KEYS = {
    B.JOB_FIXED_RADIUS: ["query_id", "neighbor_id"],
    B.JOB_KNN: ["query_id", "neighbor_rank"],
    B.JOB_BOUNDED_KNN: ["query_id", "neighbor_rank"],
    B.JOB_RAY_ANYHIT: ["ray_id"],
    B.JOB_RAY_HITCOUNT: ["ray_id"],
    B.JOB_RAY_CLOSEST: ["ray_id"],
    B.JOB_PIP_POSITIVE: ["point_id", "polygon_id"],
}


def _say(msg):
    print(msg, flush=True)


def _canon(rows, kind):
    keys = _KEYS[kind]
    return rows[np.lexsort(tuple(rows[k] for k in reversed(keys)))]


def _assert_rows_identical(a, b, kind):
    """Field by field, floats as bit patterns (row structs have padding that
    numpy does not copy when it reorders them)."""
    assert len(a) == len(b), (kind, len(a), len(b))
    for name in a.dtype.names:
        x, y = a[name], b[name]
        if x.dtype.kind == "f":
            x, y = x.view(np.uint64), y.view(np.uint64)
        bad = np.flatnonzero(x != y)
        assert bad.size == 0, (kind, name, int(bad.size), int(bad[0]))


def _scenes(rng):
    pts = cases.points(rng, 30000, 3, kind="mixture")
    p3 = np.stack([pts["x"], pts["y"], pts["z"]], axis=1)
    points = B.prepare_point_instances(DeviceBuffer.from_numpy(p3, dtype=np.float32),
                                       np.array([0, 20000, 30000]))
    tri = cases.tri_matrix(cases.triangles(rng, 20000, sigma=0.01))
    tris = B.prepare_triangle_instances(DeviceBuffer.from_numpy(tri, dtype=np.float32), np.array([0, 20000]))
    poly = cases.star_polygons(rng, 600, rmin=0.02, rmax=0.06)
    refs = poly["refs"]
    polys = B.prepare_polygon_instances(
        DeviceBuffer.from_numpy(refs["vertex_offset"], dtype=np.int32),
        DeviceBuffer.from_numpy(refs["vertex_count"], dtype=np.int32),
        DeviceBuffer.from_numpy(poly["vertices_xy"].reshape(-1, 2), dtype=np.float32),
        np.array([0, len(refs)]))
    return points, tris, polys


def _jobs(rng, scenes, scale):
    points, tris, polys = scenes
    out = []

    def q3(n):
        p = cases.points(rng, n, 3, kind="mixture")
        return DeviceBuffer.from_numpy(np.stack([p["x"], p["y"], p["z"]], axis=1), dtype=np.float32)

    for inst, kind, n, radius, k in [
        (0, B.JOB_FIXED_RADIUS, 2500, 0.02, 16), (1, B.JOB_FIXED_RADIUS, 1500, 0.05, 64),
        (0, B.JOB_KNN, 2000, 0.0, 8), (1, B.JOB_KNN, 700, 0.0, 33), (0, B.JOB_BOUNDED_KNN, 1200, 0.03, 5),
    ]:
        out.append(B.BatchJob(points, kind, q3(max(1, n // scale)), instance=inst, radius=radius, k=k))
    r = cases.ray_matrix(cases.rays(rng, max(1, 1900 // scale), tmax=0.3))
    rd = DeviceBuffer.from_numpy(r, dtype=np.float32)
    for kind in (B.JOB_RAY_ANYHIT, B.JOB_RAY_HITCOUNT, B.JOB_RAY_CLOSEST):
        out.append(B.BatchJob(tris, kind, rd))
    pp = rng.random((max(1, 2500 // scale), 2))
    out.append(B.BatchJob(polys, B.JOB_PIP_POSITIVE, DeviceBuffer.from_numpy(pp, dtype=np.float32)))
    return out


def _with_own_rows(jobs):
    twins = []
    for j in jobs:
        nq = int(j.queries.shape[0])
        cap = B.row_capacity(j.kind, nq, j.k)
        rows = DeviceBuffer((max(1, cap * B.ROW_DTYPES[j.kind].itemsize),), np.uint8)
        twins.append(B.BatchJob(j.scene, j.kind, j.queries, instance=j.instance, radius=j.radius, k=j.k,
                                rows=rows, row_capacity=cap))
    return twins


def test_noalloc_batch_matches_run_batch(gpu_ready):
    rng = np.random.default_rng(909)
    scenes = _scenes(rng)
    for call, scale in enumerate((1, 5)):
        jobs = _jobs(rng, scenes, scale)
        total = sum(int(j.queries.shape[0]) for j in jobs)
        assert total <= B.NOALLOC_MAX_QUERIES
        _say(f"case: run_batch_noalloc call {call + 1}: {len(jobs)} jobs, {total} queries")
        want, _ = B.run_batch(jobs)
        mine = _with_own_rows(jobs)
        got, _ = B.run_batch_noalloc(mine)
        for j_ref, j_new, w, g in zip(jobs, mine, want, got):
            assert g["status"] == w["status"] == 0, (j_ref.kind, g, w)
            assert g["rows"] == w["rows"], (j_ref.kind, g, w)
            a = _canon(B.job_rows(j_ref, w), j_ref.kind)
            b = _canon(B.job_rows(j_new, g), j_new.kind)
            _assert_rows_identical(a, b, j_ref.kind)


def test_noalloc_batch_rejects_batches_beyond_its_limits(gpu_ready):
    rng = np.random.default_rng(910)
    points, _, _ = _scenes(rng)
    big = DeviceBuffer.from_numpy(rng.random((B.NOALLOC_MAX_QUERIES + 1, 3)), dtype=np.float32)
    job = _with_own_rows([B.BatchJob(points, B.JOB_KNN, big, k=4)])
    with pytest.raises(RuntimeError, match="at most 16384 queries"):
        B.run_batch_noalloc(job)
    small = DeviceBuffer.from_numpy(rng.random((4, 3)), dtype=np.float32)
    many = _with_own_rows([B.BatchJob(points, B.JOB_KNN, small, k=1)] * (B.NOALLOC_MAX_JOBS + 1))
    with pytest.raises(RuntimeError, match="at most 128 jobs"):
        B.run_batch_noalloc(many)
