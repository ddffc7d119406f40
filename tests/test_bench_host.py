"""Host-side parts of bench.py (CPU only): the one-GPU contract and the
pure-Python reference subsample leg."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode != 0
    assert "--gpus 1" in proc.stderr
    assert proc.stdout.strip() == ""  # no JSON line for a run that was not made


def test_python_subsample_leg_reports_rates_and_labelled_extrapolations():
    sys.path.insert(0, ROOT)
    import bench

    out = bench.python_subsample_leg()
    ray = out["ray_anyhit_2M_triangles_1M_rays"]
    knn = out["knn16_4M_x_4M_mixture"]
    assert ray["rays_per_s_at_sample"] > 0 and ray["extrapolated_full_s_upper_bound"] > 0
    assert knn["queries_per_s_at_sample"] > 0 and knn["extrapolated_full_s"] > 0
    assert out["pip_positive_1000_polygons"]["points_per_s"] > 0
    frn = out["fixed_radius_r0.01_1e5_points"]
    assert frn["extrapolated_queries_per_s_at_1e7_points"] < frn["queries_per_s"]
    for v in out.values():
        assert "sample" in v


def _fake_rows(kind, n, sample, q, scene, bench, rtm, oracle):
    """Rows as the device writes them (request-local query ids), made from
    the oracle for every query of a small request."""
    import numpy as np

    if kind in (0, 1):
        h = scene["host"].astype(np.float64)
        sd = {"id": np.arange(len(h), dtype=np.uint32), "x": h[:, 0], "y": h[:, 1], "z": h[:, 2]}
        qd = {"id": np.arange(n, dtype=np.uint32), "x": q[:, 0], "y": q[:, 1], "z": q[:, 2]}
        w = oracle.neighbors("frn" if kind == 0 else "knn", qd, sd, radius=bench.RADIUS,
                             k=bench.K_MAX if kind == 0 else bench.K_KNN)
        rows = np.zeros(len(w), dtype=rtm.FRN_ROW_DTYPE if kind == 0 else rtm.KNN_ROW_DTYPE)
        rows["query_id"], rows["neighbor_id"], rows["distance"] = w["qid"], w["nid"], w["dist"]
        if kind == 1:
            rows["neighbor_rank"] = w["rank"]
        return rows
    if kind == 2:
        w = oracle.rays_3d("any", q, scene["host"].astype(np.float64), np.arange(scene["n"], dtype=np.uint32))
        rows = np.zeros(n, dtype=rtm.ANYHIT_ROW_DTYPE)
        rows["ray_id"], rows["any_hit"] = np.arange(n), w
        return rows
    voff, vcnt, verts = scene["host"]
    w = oracle.pip_pairs(q[:, 0], q[:, 1], voff, vcnt, verts.astype(np.float64).reshape(-1), positive_only=True)
    rows = np.zeros(len(w), dtype=rtm.PIP_ROW_DTYPE)
    rows["point_id"], rows["polygon_id"], rows["contains"] = w[:, 0], w[:, 1], 1
    return rows


def test_served_row_checker_accepts_oracle_rows_and_flags_planted_mistakes():
    """verify_captures (the bench line's served_rows_checked / mismatches)
    passes rows equal to the oracle's and counts a planted wrong distance, a
    swapped rank, a flipped hit flag and a dropped PIP pair."""
    import numpy as np

    sys.path.insert(0, ROOT)
    import bench
    from oracle import native as oracle
    from rtdl_amd import mi355x_runtime as rtm

    rng = np.random.default_rng(5)
    centres = rng.random((64, 3), dtype=np.float32)
    scenes = []
    for i, kind in enumerate(bench.KINDS):
        n = 3000
        if kind in ("frn", "knn"):
            host = rng.random((n, 3), dtype=np.float32)
        elif kind == "ray":
            c = rng.random((n, 3), dtype=np.float32)
            host = (c[:, None, :] + 0.02 * rng.standard_normal((n, 3, 3), dtype=np.float32)).reshape(n, 9)
        else:
            host = bench.star_polygons(rng, n, "uniform", centres)
        scenes.append({"index": i, "kind": kind, "dist": "uniform", "n": n, "host": host})
    pools_host, _ = {}, None
    pool_n = 5000
    pools_host["point"] = rng.random((2 * pool_n, 3), dtype=np.float32)
    pools_host["pip"] = rng.random((2 * pool_n, 2), dtype=np.float32)
    o, d = rng.random((pool_n, 3)), rng.standard_normal((pool_n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pools_host["ray"] = np.concatenate([o, d, np.full((pool_n, 1), 0.2)], 1).astype(np.float32)
    # fixed-radius scene points beside the queries, so sampled queries have rows
    scenes[0]["host"] = (pools_host["point"][:3000] + 0.004 * rng.standard_normal((3000, 3))).astype(np.float32)
    group = {0: "point", 1: "point", 2: "ray", 3: "pip"}
    checks = []
    for kind in range(4):
        n, off = 300, 100 * kind
        sample = np.sort(rng.choice(n, 40, replace=False))
        q = pools_host[group[kind]][off:off + n].astype(np.float64)
        rows = _fake_rows(kind, n, sample, q, scenes[kind], bench, rtm, oracle)
        # rows keep their written order per query; other queries' rows interleave
        order = np.argsort(rows[bench._QID_FIELD[kind]], kind="stable")
        rows = rows[order]
        keep = np.isin(rows[bench._QID_FIELD[kind]], sample)
        checks.append({"kind": kind, "scene": kind, "offset": off, "n": n, "sample": sample, "rows": rows[keep]})
    rep = bench.verify_captures(checks, scenes, pools_host)
    assert rep["mismatched_queries"] == 0 and rep["requests_checked"] == 4
    assert rep["queries_checked"] == 160 and rep["served_rows_checked"] > 160
    bad = [dict(c, rows=c["rows"].copy()) for c in checks]
    frn = bad[0]["rows"]
    frn["distance"][0] = np.nextafter(frn["distance"][0], 1.0)  # one FP64 ulp off
    knn = bad[1]["rows"]
    knn["neighbor_rank"][:2] = knn["neighbor_rank"][:2][::-1].copy()
    bad[2]["rows"]["any_hit"][0] ^= 1
    bad[3]["rows"] = bad[3]["rows"][1:]
    rep = bench.verify_captures(bad, scenes, pools_host)
    assert rep["mismatched_queries"] == 4 and rep["mismatched_requests"] == 4
