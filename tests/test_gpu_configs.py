"""GPU parity at the shapes of the reference's benchmark configurations.

- Config [0]: fixed-radius rows, 4,096 x 4,096 2-D FP64 points, r = 0.02,
  against the pure-Python restatement of the reference at full size.
- Config [1]: ray any-hit, 2,000,000 triangles x 1,000,000 rays at full size;
  20,000 sampled rays checked against the C oracle over all triangles.
- Config [2]: KNN k = 16, 4,000,000 x 4,000,000 3-D Gaussian-mixture points at
  full size; 20,000 sampled queries checked against the C oracle over all
  search points.
- Config [3]: the batched fixed-radius + PIP mix with 256 instances (the
  configuration has 2,048 of up to 2.6e5 points; the instance count and the
  size range are scaled down so the oracle finishes in seconds), 2-D and 3-D
  instances, radius log-uniform in [0.005, 0.05], every 4th instance PIP
  against 1,000 star polygons; every job checked.

Every comparison is bit-exact: ids, flags and FP64 distances.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import cpu_python_reference as ref
from oracle import native as oracle
from rtdl_amd import batch as B
from rtdl_amd import mi355x_runtime as m
from rtdl_amd.device import DeviceBuffer
from rtdl_amd.records import Point
from tests import cases

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


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


def test_config3_batched_fixed_radius_and_pip_instances(gpu_ready):
    _say("case: config [3] batched fixed radius + PIP, 256 instances")
    rng = np.random.default_rng(20261206)
    n_inst = 256
    sizes = np.round(np.exp(rng.uniform(math.log(1e3), math.log(2.6e4), n_inst))).astype(np.int64)
    dims = rng.integers(2, 4, n_inst)
    radii = np.exp(rng.uniform(math.log(0.005), math.log(0.05), n_inst))
    frn = [i for i in range(n_inst) if i % 4 != 3]
    pip = [i for i in range(n_inst) if i % 4 == 3]
    # point instances: build and query draws of the same size, z = 0 for 2-D
    coords, offs, inst = [], [0], {}
    for i in frn:
        p = rng.random((sizes[i], 3)).astype(np.float32)
        if dims[i] == 2:
            p[:, 2] = 0
        inst[i] = p
        coords.append(p)
        offs.append(offs[-1] + len(p))
    pscene = B.prepare_point_instances(DeviceBuffer.from_numpy(np.concatenate(coords)), offs)
    polys = {i: cases.star_polygons(rng, 1000) for i in pip}
    voff, vcnt, verts, poffs, base = [], [], [], [0], 0
    for i in pip:
        refs = polys[i]["refs"]
        voff.append(refs["vertex_offset"] + base)
        vcnt.append(refs["vertex_count"])
        verts.append(polys[i]["vertices_xy"].reshape(-1, 2))
        base += len(verts[-1])
        poffs.append(poffs[-1] + len(refs))
    gscene = B.prepare_polygon_instances(DeviceBuffer.from_numpy(np.concatenate(voff), dtype=np.int32),
                                         DeviceBuffer.from_numpy(np.concatenate(vcnt), dtype=np.int32),
                                         DeviceBuffer.from_numpy(np.concatenate(verts), dtype=np.float32),
                                         poffs)
    jobs, wants = [], []
    for slot, i in enumerate(frn):
        d = int(dims[i])
        q = rng.random((sizes[i], 3)).astype(np.float32)
        jobs.append(B.BatchJob(pscene, B.JOB_FIXED_RADIUS, DeviceBuffer.from_numpy(q[:, :d]),
                               instance=slot, radius=float(radii[i]), k=64))
        keys = "xyz"[:d]
        qd = {"id": np.arange(len(q), dtype=np.uint32), **{c: q[:, j].astype(np.float64) for j, c in enumerate(keys)}}
        sd = {"id": np.arange(len(inst[i]), dtype=np.uint32),
              **{c: inst[i][:, j].astype(np.float64) for j, c in enumerate(keys)}}
        wants.append(("frn", oracle.neighbors("frn", qd, sd, radius=float(radii[i]), k=64)))
    for slot, i in enumerate(pip):
        pts = cases.points(rng, int(sizes[i]), 2)
        jobs.append(B.BatchJob(gscene, B.JOB_PIP_POSITIVE,
                               DeviceBuffer.from_numpy(np.stack([pts["x"], pts["y"]], 1), dtype=np.float32),
                               instance=slot))
        refs = polys[i]["refs"]
        wants.append(("pip", oracle.pip_pairs(pts["x"], pts["y"], refs["vertex_offset"], refs["vertex_count"],
                                             polys[i]["vertices_xy"], positive_only=True)))
    _say(f"  run_batch: {len(jobs)} jobs, {int(sizes.sum())} queries")
    results, _ = B.run_batch(jobs)
    for job, res, (kind, want) in zip(jobs, results, wants):
        assert res["status"] == 0
        rows = B.job_rows(job, res)
        if kind == "frn":
            order = np.lexsort((rows["neighbor_id"], rows["query_id"]))
            rows = rows[order]
            w = want[np.lexsort((want["nid"], want["qid"]))]
            assert np.array_equal(rows["query_id"], w["qid"])
            assert np.array_equal(rows["neighbor_id"], w["nid"])
            assert np.array_equal(rows["distance"], w["dist"])
        else:
            order = np.lexsort((rows["polygon_id"], rows["point_id"]))
            assert np.array_equal(rows["point_id"][order], want[:, 0])
            assert np.array_equal(rows["polygon_id"][order], want[:, 1])
