"""GPU regression: neighbours whose FP32 selection keys differ by one ulp
while their FP64 distances are equal.

The neighbour lists rank candidates by fl32(d2).  Two squared distances on
either side of an FP32 rounding point get keys one ulp apart yet can share
one FP64 square root; the reference (reference.py:362-366, 388-389) then
orders them by id.  Three such pairs around a query at the origin, with the
ids inside each pair reversed against their squared distances, make every
k from 1 to 6 cut the list inside or between tied pairs.  Rows must equal the
oracle's bit for bit on the per-call, prepared and batched paths.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import native as oracle
from rtdl_amd import batch as B
from rtdl_amd import mi355x_runtime as m
from rtdl_amd.device import DeviceBuffer
from tests import cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _scene(dim: int):
    pairs = cases.fl32_tie_pairs(3)
    xs, ys, ids = [], [], []
    signs = [(1.0, 1.0), (-1.0, 1.0), (1.0, -1.0)]
    for j, ((x, ya, yb, _, _), (sx, sy)) in enumerate(zip(pairs, signs)):
        # A (smaller d2) gets the higher id: the reference puts B first
        xs += [sx * x, sx * x]
        ys += [sy * ya, sy * yb]
        ids += [100 + 2 * j + 1, 100 + 2 * j]
    rng = np.random.default_rng(11)
    far = 0.5 + 0.5 * rng.random((200, 2))
    xs = np.concatenate([xs, far[:, 0]]).astype(np.float32).astype(np.float64)
    ys = np.concatenate([ys, far[:, 1]]).astype(np.float32).astype(np.float64)
    ids = np.concatenate([ids, np.arange(200)]).astype(np.uint32)
    s = {"id": ids, "x": xs, "y": ys}
    q = {"id": np.arange(4, dtype=np.uint32), "x": np.zeros(4), "y": np.zeros(4)}
    if dim == 3:
        s["z"] = np.zeros(len(xs))
        q["z"] = np.zeros(4)
    return q, s


def _rows_equal(got, want, ranked):
    assert len(got) == len(want), (len(got), len(want))
    assert got["query_id"].tolist() == want["qid"].tolist()
    assert got["neighbor_id"].tolist() == want["nid"].tolist()
    assert np.array_equal(got["distance"], want["dist"])
    if ranked:
        assert got["neighbor_rank"].tolist() == want["rank"].tolist()


@pytest.mark.parametrize("dim", [2, 3])
def test_per_call_and_prepared_tie_pairs(gpu_ready, dim):
    q, s = _scene(dim)
    for k in range(1, 7):
        _say(f"case: fl32 key ties per call dim={dim} k={k}")
        want = oracle.neighbors("knn", q, s, k=k)
        first = want[want["qid"] == 0]["nid"][:2].tolist()
        assert first == [100, 101][: min(k, 2)]  # the tied pair: lower id first
        _rows_equal(m.knn_rows_mi355x(q, s, k=k), want, ranked=True)
        want = oracle.neighbors("frn", q, s, radius=0.02, k=k)
        _rows_equal(m.fixed_radius_neighbors_mi355x(q, s, radius=0.02, k_max=k), want, ranked=False)
        want = oracle.neighbors("bounded_knn", q, s, radius=0.02, k=k)
        _rows_equal(m.fixed_radius_neighbors_mi355x(q, s, radius=0.02, k_max=k, bounded_rank=True),
                    want, ranked=True)
    arr, _ = m.pack_points(s)
    qa, _ = m.pack_points(q)
    handle = m._call_prepare(f"rtdl_mi355x_prepare_points_{dim}d", m._ptr(arr), len(arr))
    try:
        for k in (1, 3, 5):
            _say(f"case: fl32 key ties prepared dim={dim} k={k}")
            got = m._call_rows("rtdl_mi355x_run_prepared_knn_rows", m.KNN_ROW_DTYPE, handle, m._ptr(qa),
                               len(qa), k)
            _rows_equal(got, oracle.neighbors("knn", q, s, k=k), ranked=True)
    finally:
        m._library().rtdl_mi355x_destroy_prepared(handle)


def test_batched_tie_pairs(gpu_ready):
    q, s = _scene(3)
    s3 = np.stack([s["x"], s["y"], s["z"]], axis=1)
    q3 = np.stack([q["x"], q["y"], q["z"]], axis=1)
    coords = DeviceBuffer.from_numpy(s3, dtype=np.float32)
    inst = B.prepare_point_instances(coords, np.array([0, len(s3)]))
    qd = DeviceBuffer.from_numpy(q3, dtype=np.float32)
    local = {"id": np.arange(len(s3), dtype=np.uint32), "x": s["x"], "y": s["y"], "z": s["z"]}
    jobs, wants = [], []
    for k in range(1, 7):
        jobs.append(B.BatchJob(inst, B.JOB_KNN, qd, k=k))
        wants.append((oracle.neighbors("knn", q, local, k=k), True))
        jobs.append(B.BatchJob(inst, B.JOB_FIXED_RADIUS, qd, radius=0.02, k=k))
        wants.append((oracle.neighbors("frn", q, local, radius=0.02, k=k), False))
        jobs.append(B.BatchJob(inst, B.JOB_BOUNDED_KNN, qd, radius=0.02, k=k))
        wants.append((oracle.neighbors("bounded_knn", q, local, radius=0.02, k=k), True))
    _say("case: fl32 key ties batched, 18 jobs")
    res, _ = B.run_batch(jobs)
    for job, r, (want, ranked) in zip(jobs, res, wants):
        assert r["status"] == 0
        rows = B.job_rows(job, r)
        # a query's rows are contiguous and in (distance, id) order; queries
        # are compacted in wavefront order, so group them by query id
        rows = rows[np.argsort(rows["query_id"], kind="stable")]
        if ranked:
            assert rows["neighbor_rank"].tolist() == want["rank"].tolist()
        assert len(rows) == len(want)
        assert rows["query_id"].tolist() == want["qid"].tolist()
        assert rows["neighbor_id"].tolist() == want["nid"].tolist()
        assert np.array_equal(rows["distance"], want["dist"])
