"""GPU parity at the shapes of the reference's benchmark configurations.

- Config [0]: fixed-radius rows, 4,096 x 4,096 2-D FP64 points, r = 0.02,
  against the pure-Python restatement of the reference at full size.
- Config [1]: ray any-hit, 2,000,000 triangles x 1,000,000 rays at full size;
  20,000 sampled rays checked against the C oracle over all triangles.
- Config [2]: KNN k = 16, 4,000,000 x 4,000,000 3-D Gaussian-mixture points at
  full size; 20,000 sampled queries checked against the C oracle over all
  search points.
- Config [3]: the batched fixed-radius + PIP mix at full size (2,048
  instances of 1e3-2.6e5 points, 2-D and 3-D, radius log-uniform in
  [0.005, 0.05], every 4th instance PIP against 1,000 star polygons),
  autotuned and replayed from a hipGraph (also with the point instances
  rebuilt in the graph); sampled queries of every job checked against the C
  oracle over the whole instance.
- Config [4]: the served stream at its scene scale (16 resident scenes of
  1e5-1e7 primitives) and request sizes (1-65,536 queries): half a second of
  Poisson arrivals at 20,000 requests/s; sampled queries of every 25th
  request checked against the C oracle over the whole scene, and a few
  captured requests replayed alone with the same rows as in the mix.

Every comparison is bit-exact: ids, flags and FP64 distances.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu_python_reference as ref
from oracle import native as oracle
from rtdl_amd import batch as B
from rtdl_amd import mi355x_runtime as m
from rtdl_amd.records import Point
from tests import cases

pytestmark = pytest.mark.gpu


_T0 = [0.0]


def _say(msg: str) -> None:
    import time

    if msg.startswith("case:"):
        _T0[0] = time.perf_counter()
    print(f"[{time.perf_counter() - _T0[0]:7.1f}s] {msg}", flush=True)


def test_config0_fixed_radius_matches_the_python_reference_at_full_size(gpu_ready):
    _say("case: config [0] fixed radius 4096 x 4096 2-D FP64, r = 0.02")
    rng = np.random.default_rng(20261203)
    q, s = rng.random((4096, 2)), rng.random((4096, 2))
    qd = {"id": np.arange(4096, dtype=np.uint32), "x": q[:, 0], "y": q[:, 1]}
    sd = {"id": np.arange(4096, dtype=np.uint32), "x": s[:, 0], "y": s[:, 1]}
    got = m.fixed_radius_neighbors_mi355x(qd, sd, radius=0.02, k_max=4096)
    want = ref.fixed_radius_neighbors_cpu([Point(i, float(x), float(y)) for i, (x, y) in enumerate(q)],
                                          [Point(i, float(x), float(y)) for i, (x, y) in enumerate(s)],
                                          radius=0.02, k_max=4096)
    assert len(got) == len(want) > 20000
    assert np.array_equal(got["query_id"], np.array([r["query_id"] for r in want], dtype=np.uint32))
    assert np.array_equal(got["neighbor_id"], np.array([r["neighbor_id"] for r in want], dtype=np.uint32))
    assert np.array_equal(got["distance"], np.array([r["distance"] for r in want]))


def test_config1_ray_anyhit_full_size_sampled_against_the_oracle(gpu_ready):
    _say("case: config [1] any-hit, 2M triangles x 1M rays")
    rng = np.random.default_rng(20261204)
    t = cases.triangles(rng, 2_000_000, sigma=0.005)
    r = cases.rays(rng, 1_000_000, tmax=0.5)
    got = m.ray_triangle_mi355x(r, t, mode="any_hit")
    assert np.array_equal(got["ray_id"], r["id"])
    sample = np.sort(rng.choice(1_000_000, 20_000, replace=False))
    rays = cases.ray_matrix(r)[sample]
    want = oracle.rays_3d("any", rays, cases.tri_matrix(t), t["id"])
    assert np.array_equal(got["any_hit"][sample], want)
    assert 0 < int(want.sum()) < len(want)


def test_config2_knn16_full_size_sampled_against_the_oracle(gpu_ready):
    _say("case: config [2] KNN k = 16, 4M x 4M mixture points")
    rng = np.random.default_rng(20261205)
    centres = rng.random((64, 3))
    s = cases.points(rng, 4_000_000, 3, kind="mixture", centres=centres)
    q = cases.points(rng, 4_000_000, 3, kind="mixture", centres=centres)
    got = m.knn_rows_mi355x(q, s, k=16)
    assert len(got) == 16 * 4_000_000
    sample = np.sort(rng.choice(4_000_000, 20_000, replace=False))
    qs = {key: v[sample] for key, v in q.items()}
    want = oracle.neighbors("knn", qs, s, k=16)
    rows = got.reshape(4_000_000, 16)[sample].reshape(-1)
    assert np.array_equal(rows["query_id"], want["qid"])
    assert np.array_equal(rows["neighbor_id"], want["nid"])
    assert np.array_equal(rows["distance"], want["dist"])
    assert np.array_equal(rows["neighbor_rank"], want["rank"])


def test_config3_batched_fixed_radius_and_pip_at_full_size(gpu_ready):
    """Config [3] at its stated workload: 2,048 instances of 1e3-2.6e5 points
    (2-D / 3-D 50/50, r log-uniform in [0.005, 0.05], k_max 64), every 4th a
    PIP instance against 1,000 star polygons, in one batch.  The batch is
    autotuned (launch size and point leaf size; every candidate's rows
    compared on 12 sampled jobs), then recorded into a hipGraph and replayed; every
    replay's row counts and sampled rows must equal the eager run's, and 8
    sampled queries of every job are checked against the C oracle over the
    whole instance."""
    import bench
    from rtdl_amd import autotune as AT

    _say("case: config [3] 2,048 instances at full size")
    rng = np.random.default_rng(20261206)
    ws = bench.config3_workload(rng)
    c3 = bench.Config3(B, ws)
    n_jobs = len(ws["sizes"])
    _say(f"  {int(ws['point_offsets'][-1])} fixed-radius + {int(ws['probe_offsets'][-1])} PIP queries")
    original = AT.get_tuning()
    try:
        check = sorted(rng.choice(n_jobs, 12, replace=False).tolist())
        _say("  autotune_batch")
        tuned = AT.autotune_batch(c3.build_scenes, c3.jobs, repeats=1, check_jobs=check,
                                  knobs={"blocks_per_cu": (4, 8), "point_leaf": (2, 4)})
        assert len(tuned["table"]) >= 3
        jobs = c3.jobs()
        _say("  eager run_batch")
        eager, _ = B.run_batch(jobs)
        assert all(r["status"] == 0 for r in eager)
        sample_rows = {i: bench._rows_by_query(B.job_rows(jobs[i], eager[i]),
                                               "query_id" if i < len(ws["frn_ids"]) else "point_id",
                                               np.arange(0))[0] for i in check[:4]}
        with B.BatchGraph(jobs) as graph:
            for rep in range(3):
                _say(f"  graph replay {rep}")
                res, _ = graph.launch()
                assert [r["rows"] for r in res] == [r["rows"] for r in eager]
                assert all(r["status"] == 0 for r in res)
                for i, want in sample_rows.items():
                    key = "query_id" if i < len(ws["frn_ids"]) else "point_id"
                    got = bench._rows_by_query(B.job_rows(jobs[i], res[i]), key, np.arange(0))[0]
                    for name in want.dtype.names:
                        assert np.array_equal(got[name], want[name]), (i, name)
        # build + traverse + emit: the point instances rebuilt from their
        # device coordinates in the graph (same coordinates, same rows)
        with B.BatchGraph(jobs, rebuild=[(c3.pscene, c3.dev_build)]) as graph:
            for rep in range(2):
                _say(f"  graph replay with rebuild {rep}")
                res, _ = graph.launch()
                assert [r["rows"] for r in res] == [r["rows"] for r in eager]
                for i, want in sample_rows.items():
                    key = "query_id" if i < len(ws["frn_ids"]) else "point_id"
                    got = bench._rows_by_query(B.job_rows(jobs[i], res[i]), key, np.arange(0))[0]
                    for name in want.dtype.names:
                        assert np.array_equal(got[name], want[name]), (i, name)
        _say("  oracle check, 8 queries of every job")
        rep = bench.config3_check(B, ws, jobs, eager, rng, jobs_checked=n_jobs, per_job=8)
        assert rep["jobs_checked"] == n_jobs
        assert rep["mismatched_queries"] == 0, rep
        assert rep["rows_checked"] > 8 * n_jobs
    finally:
        AT.set_tuning(original)


def test_config4_served_stream_at_scene_scale(gpu_ready):
    import bench

    _say("case: config [4] served stream, 16 scenes of 1e5-1e7 primitives, 0.5 s of requests")
    rng = np.random.default_rng(20261207)
    scenes, centres = bench.make_scenes(B, rng, 1e7)
    assert max(s["n"] for s in scenes) == 10_000_000 and min(s["n"] for s in scenes) == 100_000
    handles = [s["handle"] for s in scenes]
    pools_host, pools_dev = bench.make_pools(rng, centres)
    stream = bench.make_stream(rng, 0.5, 20000.0, scenes)
    assert stream["size"].max() > 30000  # requests up to the 65,536-query range
    offsets = bench.assign_windows(rng, stream, scenes)
    reqs = bench.requests_array(B, stream, offsets, {s["index"]: s["index"] for s in scenes}, capture_every=25)
    kw = dict(radius=bench.RADIUS, k_max=bench.K_MAX, k=bench.K_KNN, slots=4,
              point_queries=pools_dev["point"], ray_queries=pools_dev["ray"], pip_queries=pools_dev["pip"])
    _say(f"  serve_replay: {len(reqs)} requests, {int(stream['size'].sum())} queries")
    results, stats, captured = B.serve_replay(handles, reqs, **kw)
    assert np.all(results["status"] == 0)
    assert stats["queries"] == int(stream["size"].sum())
    assert len(captured) == len(range(0, len(reqs), 25))
    check_rng = np.random.default_rng(7)
    checks = [bench.sample_capture(check_rng, stream, offsets, i, rows) for i, rows in captured.items()]
    _say(f"  oracle check of {len(checks)} captured requests")
    rep = bench.verify_captures(checks, scenes, pools_host)
    assert rep["mismatched_queries"] == 0, rep
    assert rep["requests_checked"] == len(checks) and rep["served_rows_checked"] > 0
    kinds_seen = {int(stream["kind"][i]) for i in captured}
    assert kinds_seen == {0, 1, 2, 3}
    for kind in range(4):
        i = next(i for i in sorted(captured) if int(stream["kind"][i]) == kind)
        alone = reqs[i:i + 1].copy()
        alone["arrival_s"] = 0.0
        _say(f"  request {i} (kind {kind}, {int(alone['query_count'][0])} queries) alone")
        res1, _, cap1 = B.serve_replay(handles, alone, **kw)
        assert res1["rows"][0] == results["rows"][i]
        field = bench._QID_FIELD[kind]
        a = bench._rows_by_query(captured[i], field, np.arange(0))[0]
        b = bench._rows_by_query(cap1[0], field, np.arange(0))[0]
        for name in a.dtype.names:
            assert np.array_equal(a[name], b[name]), (kind, name)
