"""GPU tests of in-place rebuilds of instance scenes (moving primitives):
rtdl_mi355x_rebuild_instances, and hipGraphs that record the rebuild of the
scenes in front of the batch (build + traverse + emit replayed by one call).
Rows must equal those of scenes prepared from scratch from the same
coordinates, bit for bit, and sampled queries those of the C oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import native as oracle
from rtdl_amd import batch as B
from rtdl_amd.device import DeviceBuffer
from tests import cases

pytestmark = pytest.mark.gpu

# an empty and a one-point instance among ordinary ones
POINT_SIZES = [3000, 0, 1, 4500, 2000]
TRI_SIZES = [2500, 1, 3000]


def _say(msg: str) -> None:
    print(msg, flush=True)


def _offsets(sizes):
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)


class _Motion:
    """Seeded positions that move with time: (base + t * velocity) wrapped
    into the unit cube, clustered around a few centres; FP32."""

    def __init__(self, rng, n, per):
        centres = rng.random((8, per))
        self.base = centres[rng.integers(0, 8, n)] + rng.normal(0.0, 0.05, (n, per))
        self.vel = rng.normal(0.0, 0.4, (n, per))
        self.per = per

    def at(self, t: float) -> np.ndarray:
        return np.mod(self.base + t * self.vel, 1.0).astype(np.float32)


def _point_jobs(scene, queries):
    jobs = []
    for i, q in enumerate(queries):
        jobs.append(B.BatchJob(scene, B.JOB_FIXED_RADIUS, q, radius=0.03, k=32, instance=i))
        jobs.append(B.BatchJob(scene, B.JOB_KNN, q, k=8, instance=i))
    return jobs


def _sorted(rows, key):
    return rows[np.lexsort((rows["neighbor_id"], rows["distance"], rows[key]))]


def _same_rows(jobs_a, res_a, jobs_b, res_b, key="query_id"):
    assert [r["status"] for r in res_a] == [0] * len(res_a)
    assert [r["rows"] for r in res_a] == [r["rows"] for r in res_b]
    for ja, ra, jb, rb in zip(jobs_a, res_a, jobs_b, res_b):
        a, b = _sorted(B.job_rows(ja, ra), key), _sorted(B.job_rows(jb, rb), key)
        for name in a.dtype.names:
            assert np.array_equal(a[name], b[name]), name


def test_point_rebuild_eager_and_graph_match_fresh_scenes(gpu_ready):
    _say("case: point instances rebuilt in place (eager, then recorded in a hipGraph)")
    rng = np.random.default_rng(301)
    offs = _offsets(POINT_SIZES)
    n = int(offs[-1])
    motion = _Motion(rng, n, 3)
    motion.base[int(offs[3]):int(offs[4]), 2] = 0.0  # one 2-D instance (z = 0)
    motion.vel[int(offs[3]):int(offs[4]), 2] = 0.0
    queries = [DeviceBuffer.from_numpy(rng.random((700, 3)).astype(np.float32)) for _ in POINT_SIZES]

    def fresh(t):
        scene = B.prepare_point_instances(DeviceBuffer.from_numpy(motion.at(t)), offs, leaf_size=4)
        jobs = _point_jobs(scene, queries)
        res, _ = B.run_batch(jobs)
        return scene, jobs, res

    # eager: a scene built at t = 0 and rebuilt at t = 0.5
    coords = DeviceBuffer.from_numpy(motion.at(0.0))
    scene = B.prepare_point_instances(coords, offs, leaf_size=4)
    jobs = _point_jobs(scene, queries)
    coords.upload(motion.at(0.5))
    B.rebuild_instances(scene, coords)
    res, _ = B.run_batch(jobs)
    _, jobs_f, res_f = fresh(0.5)
    _same_rows(jobs, res, jobs_f, res_f)

    # graph: build + traverse + emit replayed with new contents each launch
    with B.BatchGraph(jobs, rebuild=[(scene, coords)]) as graph:
        for t in (0.9, 1.7, 1.7, 2.3):
            _say(f"  graph launch at t = {t}")
            coords.upload(motion.at(t))
            res, _ = graph.launch()
            scene_f, jobs_f, res_f = fresh(t)
            _same_rows(jobs, res, jobs_f, res_f)
            # sampled queries of instance 0 against the C oracle
            inst = motion.at(t)[: POINT_SIZES[0]].astype(np.float64)
            s = {"id": np.arange(POINT_SIZES[0], dtype=np.uint32), "x": inst[:, 0], "y": inst[:, 1],
                 "z": inst[:, 2]}
            qh = queries[0].download().view(np.float32).reshape(-1, 3)[:40].astype(np.float64)
            qs = {"id": np.arange(40, dtype=np.uint32), "x": qh[:, 0], "y": qh[:, 1], "z": qh[:, 2]}
            want = oracle.neighbors("knn", qs, s, k=8)
            got = B.job_rows(jobs[1], res[1])
            got = got[got["query_id"] < 40]
            got = got[np.lexsort((got["neighbor_rank"], got["query_id"]))]
            assert np.array_equal(got["neighbor_id"], want["nid"])
            assert np.array_equal(got["distance"], want["dist"])


def test_triangle_rebuild_in_a_graph_matches_fresh_scenes(gpu_ready):
    _say("case: triangle instances rebuilt in a hipGraph, ray any-hit and closest-hit")
    rng = np.random.default_rng(302)
    offs = _offsets(TRI_SIZES)
    n = int(offs[-1])
    centre = _Motion(rng, n, 3)
    shape = (rng.normal(0.0, 0.01, (n, 9))).astype(np.float64)

    def tris(t):
        c = centre.at(t).astype(np.float64)
        return (np.tile(c, (1, 3)) + shape).astype(np.float32)

    r = cases.rays(rng, 1500, tmax=0.6)
    ray_f32 = cases.ray_matrix(r).astype(np.float32)
    rays = DeviceBuffer.from_numpy(ray_f32)

    def ray_jobs(scene):
        jobs = []
        for i in range(len(TRI_SIZES)):
            jobs.append(B.BatchJob(scene, B.JOB_RAY_ANYHIT, rays, instance=i))
            jobs.append(B.BatchJob(scene, B.JOB_RAY_CLOSEST, rays, instance=i))
        return jobs

    coords = DeviceBuffer.from_numpy(tris(0.0))
    scene = B.prepare_triangle_instances(coords, offs)
    jobs = ray_jobs(scene)
    with B.BatchGraph(jobs, rebuild=[(scene, coords)]) as graph:
        for t in (0.4, 1.1):
            _say(f"  graph launch at t = {t}")
            coords.upload(tris(t))
            res, _ = graph.launch()
            assert all(x["status"] == 0 for x in res)
            fresh = B.prepare_triangle_instances(DeviceBuffer.from_numpy(tris(t)), offs)
            jobs_f = ray_jobs(fresh)
            res_f, _ = B.run_batch(jobs_f)
            assert [x["rows"] for x in res] == [x["rows"] for x in res_f]
            for ja, ra, jb, rb in zip(jobs, res, jobs_f, res_f):
                a, b = B.job_rows(ja, ra), B.job_rows(jb, rb)
                a, b = a[np.argsort(a["ray_id"], kind="stable")], b[np.argsort(b["ray_id"], kind="stable")]
                for name in a.dtype.names:
                    assert np.array_equal(a[name], b[name]), name
            # instance 0 any-hit against the C oracle
            t0 = tris(t)[: TRI_SIZES[0]].astype(np.float64)
            want = oracle.rays_3d("any", ray_f32.astype(np.float64), t0, np.arange(TRI_SIZES[0], dtype=np.uint32))
            got = B.job_rows(jobs[0], res[0])
            got = got[np.argsort(got["ray_id"], kind="stable")]
            assert np.array_equal(got["any_hit"].astype(bool), want.astype(bool))
            assert 0 < int(want.sum()) < len(want)


def test_rebuild_rejects_what_it_cannot_keep(gpu_ready):
    _say("case: rebuild error (NaN coordinate), then a good rebuild")
    rng = np.random.default_rng(303)
    pts = rng.random((500, 3)).astype(np.float32)
    scene = B.prepare_point_instances(DeviceBuffer.from_numpy(pts), [0, 500])
    bad = pts.copy()
    bad[17, 1] = np.nan
    with pytest.raises(RuntimeError, match="not exact"):
        B.rebuild_instances(scene, DeviceBuffer.from_numpy(bad))
    # the scene is usable again after a good rebuild
    B.rebuild_instances(scene, DeviceBuffer.from_numpy(pts))
    q = DeviceBuffer.from_numpy(rng.random((64, 3)).astype(np.float32))
    job = B.BatchJob(scene, B.JOB_KNN, q, k=4)
    (res,), _ = B.run_batch([job])
    assert res["status"] == 0 and res["rows"] == 256
