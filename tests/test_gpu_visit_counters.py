"""The in-kernel visit counters (nodes fetched, primitives tested) that
bench.py turns into algorithmic bytes equal a host replay of the same LBVH
walks (tests/walk_replay.py) on more than 1,000 sampled queries."""
from __future__ import annotations

import numpy as np
import pytest

from rtdl_amd import batch as B
from rtdl_amd.device import DeviceBuffer
from tests import cases
from tests import walk_replay as W

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


@pytest.mark.parametrize("kind", ["uniform", "mixture"])
def test_knn_visit_counters_equal_a_host_replay(gpu_ready, kind):
    _say(f"case: KNN k = 8 visit counters vs host replay ({kind})")
    rng = np.random.default_rng(96)
    centres = rng.random((64, 3))
    pts = cases.points(rng, 60000, 3, kind=kind, centres=centres)
    s3 = np.stack([pts["x"], pts["y"], pts["z"]], 1).astype(np.float32)
    q = cases.points(rng, 1200, 3, kind=kind, centres=centres)
    q3 = np.stack([q["x"], q["y"], q["z"]], 1).astype(np.float32)
    scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, len(s3)])
    job = B.BatchJob(scene, B.JOB_KNN, DeviceBuffer.from_numpy(q3), k=8)
    res, _, visits = B.run_batch_profiled([job], count_visits=True)
    assert res[0]["status"] == 0
    sc = W.export_scene(scene.handle)
    assert sc["exact"] is None  # FP32 device input: the FP32 copy is exact
    nodes = prims = 0
    for row in q3.astype(np.float64):
        n, p = W.knn_counts(sc, row, 8)
        nodes += n
        prims += p
    assert (nodes, prims) == visits[:2], ((nodes, prims), visits)


def test_ray_anyhit_visit_counters_equal_a_host_replay(gpu_ready):
    _say("case: ray any-hit visit counters vs host replay")
    rng = np.random.default_rng(97)
    t = cases.triangles(rng, 40000, sigma=0.01)
    r = cases.rays(rng, 1200, tmax=0.3)
    rays = cases.ray_matrix(r).astype(np.float32)
    rays[::7, 6] = np.inf  # some unbounded rays
    scene = B.prepare_triangle_instances(DeviceBuffer.from_numpy(cases.tri_matrix(t).astype(np.float32)),
                                         [0, len(t["id"])])
    job = B.BatchJob(scene, B.JOB_RAY_ANYHIT, DeviceBuffer.from_numpy(rays))
    res, _, visits = B.run_batch_profiled([job], count_visits=True)
    assert res[0]["status"] == 0
    sc = W.export_scene(scene.handle)
    nodes = prims = 0
    for row in rays.astype(np.float64):
        n, p = W.ray_anyhit_counts(sc, row)
        nodes += n
        prims += p
    assert (nodes, prims) == visits[:2], ((nodes, prims), visits)


def test_fixed_radius_visit_counters_equal_a_host_replay(gpu_ready):
    _say("case: fixed-radius count walk visit counters vs host replay")
    rng = np.random.default_rng(98)
    centres = rng.random((64, 3))
    pts = cases.points(rng, 60000, 3, kind="mixture", centres=centres)
    s3 = np.stack([pts["x"], pts["y"], pts["z"]], 1).astype(np.float32)
    # 300 copies of one point: queries next to it need more than one count round
    s3 = np.concatenate([s3, np.repeat(s3[:1], 300, axis=0)])
    q = cases.points(rng, 1100, 3, kind="mixture", centres=centres)
    q3 = np.stack([q["x"], q["y"], q["z"]], 1).astype(np.float32)
    q3 = np.concatenate([q3, s3[:1] + rng.normal(0.0, 1e-4, (20, 3)).astype(np.float32)])
    scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, len(s3)])
    sc = W.export_scene(scene.handle)
    assert sc["exact"] is None  # FP32 device input: the FP32 copy is exact
    for radius, k in ((0.05, 64), (0.01, 16), (0.02, 1)):
        _say(f"  radius {radius}, k_max {k}")
        job = B.BatchJob(scene, B.JOB_FIXED_RADIUS, DeviceBuffer.from_numpy(q3), radius=radius, k=k)
        # without the Morton reordering a work item is 64 consecutive queries,
        # one per lane, walked in lock step: the replay runs per item
        res, _, visits = B.run_batch_profiled([job], morton_sort=False, count_visits=True)
        assert res[0]["status"] == 0
        nodes = prims = steps = 0
        q64 = q3.astype(np.float64)
        for first in range(0, len(q64), 64):
            n, p, iters = W.frn_counts(sc, q64[first:first + 64], radius, k)
            nodes += n
            prims += p
            steps += iters
        assert (nodes, prims, steps) == visits, (radius, k, (nodes, prims, steps), visits)


def _check_boxes(sc, leaf_points):
    nodes = sc["nodes"]
    lo = nodes["lo"].astype(np.float32)
    hi = nodes["hi"].astype(np.float32)
    link = nodes["link"].astype(np.int64)
    leaf = (link & W.LEAF) != 0
    internal = np.flatnonzero(~leaf)
    left = internal + 1  # the next record
    right = link[internal]
    # the right child also follows the left subtree in preorder
    assert np.array_equal(nodes["escape"][left].astype(np.int64), right)
    assert np.array_equal(lo[internal], np.minimum(lo[left], lo[right]))
    assert np.array_equal(hi[internal], np.maximum(hi[left], hi[right]))
    first = link[leaf] & W.START_MASK
    count = ((link[leaf] >> 28) & 7) + 1
    assert count.sum() == len(leaf_points)
    llo, lhi = lo[leaf], hi[leaf]
    for j in range(int(count.max())):
        ok = j < count
        p = leaf_points[first[ok] + j]
        assert np.all(llo[ok] <= p) and np.all(lhi[ok] >= p)
    # in preorder the node after a leaf is the next record (the last leaf ends the walk)
    esc = nodes["escape"][leaf].astype(np.int64)
    idx = np.flatnonzero(leaf)
    assert np.all((esc == idx + 1) | (esc == 0xFFFFFFFF))
    if sc["qnodes"] is not None:
        _check_compact(sc, leaf)
    return int(len(internal))


def _check_compact(sc, leaf):
    """The compact nodes of the count walk: every dequantised box contains the
    node's FP32 box and is the tightest one of the frame's grid, the frame's
    steps are powers of two (so the dequantisation is exact), and the words
    are the leaves' links and the internal nodes' escapes."""
    nodes, fr = sc["nodes"], sc["qframe"]
    n = len(nodes)
    assert fr[6] == 0 and fr[7] == 0
    o, st = fr[:3].astype(np.float64), fr[3:6].astype(np.float64)
    m, e = np.frexp(st)
    assert np.all(m == 0.5), st  # powers of two
    assert np.all(np.floor(o / st) * st == o)
    lo = nodes["lo"].astype(np.float32)
    hi = nodes["hi"].astype(np.float32)
    qlo, qhi, word = W.compact_boxes(sc)
    assert np.all(qlo <= lo) and np.all(qhi >= hi)
    # one step tighter would no longer contain the box (exact in FP64: o + q * s
    # has at most 24 significant bits)
    q = sc["qnodes"]
    ql = np.stack([q[:, 0] & 0xFFFF, q[:, 0] >> 16, q[:, 1] & 0xFFFF], 1).astype(np.float64)
    qh = np.stack([q[:, 1] >> 16, q[:, 2] & 0xFFFF, q[:, 2] >> 16], 1).astype(np.float64)
    assert np.array_equal(o + ql * st, qlo.astype(np.float64))
    assert np.all((ql == 65535) | (o + (ql + 1) * st > lo))
    assert np.all((qh == 0) | (o + (qh - 1) * st < hi))
    link = nodes["link"].astype(np.int64)
    esc = nodes["escape"].astype(np.int64)
    want = np.where(leaf, link, np.where(esc == 0xFFFFFFFF, n, esc))
    assert np.array_equal(word.astype(np.int64), want)


def test_lbvh_node_boxes_are_the_unions_of_their_children(gpu_ready):
    _say("case: LBVH refit: internal boxes equal the union of the children, leaves hold their primitives")
    rng = np.random.default_rng(99)
    centres = rng.random((64, 3))
    for n, kind in ((100000, "uniform"), (250001, "mixture"), (1, "uniform"), (5, "uniform")):
        pts = cases.points(rng, n, 3, kind=kind, centres=centres)
        s3 = np.stack([pts["x"], pts["y"], pts["z"]], 1).astype(np.float32)
        s3[: n // 3] = s3[0]  # duplicate keys: index tie-breaks deepen the tree
        scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, n])
        sc = W.export_scene(scene.handle)
        _check_boxes(sc, sc["prims"][:, :3])
    # coordinates of both signs
    pts = cases.points(rng, 20000, 3, kind="uniform", centres=centres)
    s3 = (np.stack([pts["x"], pts["y"], pts["z"]], 1) * 3.0 - 2.0).astype(np.float32)
    scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, len(s3)])
    sc = W.export_scene(scene.handle)
    assert sc["qframe"][0] < 0 < sc["qframe"][0] + 65535 * sc["qframe"][3]
    _check_boxes(sc, sc["prims"][:, :3])
    # several instances far from the origin: one frame over all of them
    pts = cases.points(rng, 30000, 3, kind="uniform", centres=centres)
    s3 = (np.stack([pts["x"], pts["y"], pts["z"]], 1) * 3.0 + 1000.0).astype(np.float32)
    scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, 7000, 7001, 30000])
    sc = W.export_scene(scene.handle)
    assert sc["qnodes"] is not None
    _check_boxes(sc, sc["prims"][:, :3])
    t = cases.triangles(rng, 50000, sigma=0.01)
    scene = B.prepare_triangle_instances(DeviceBuffer.from_numpy(cases.tri_matrix(t).astype(np.float32)),
                                         [0, len(t["id"])])
    sc = W.export_scene(scene.handle)
    assert sc["qnodes"] is None  # compact nodes are for point scenes
    v = sc["prims"].reshape(-1, 3, 4)[:, :, :3]
    _check_boxes(sc, v.min(axis=1))
    _check_boxes(sc, v.max(axis=1))


def test_compact_nodes_fall_back_when_a_scene_box_is_not_finite(gpu_ready):
    """A point at infinity makes the quantisation frame unusable: the frame is
    flagged and the count walk reads the full nodes; the rows are those of the
    finite points (a point at infinity is never within a finite radius)."""
    from oracle import native as oracle

    _say("case: scene with a point at infinity: flagged frame, count walk on full nodes")
    rng = np.random.default_rng(101)
    s3 = rng.random((5000, 3)).astype(np.float32)
    s3[17] = (np.inf, 0.5, 0.5)
    scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, len(s3)])
    sc = W.export_scene(scene.handle)
    assert sc["qnodes"] is not None and sc["qframe"][6] == 1
    q3 = rng.random((400, 3)).astype(np.float32)
    job = B.BatchJob(scene, B.JOB_FIXED_RADIUS, DeviceBuffer.from_numpy(q3), radius=0.08, k=16)
    (res,), _ = B.run_batch([job])
    assert res["status"] == 0
    got = B.job_rows(job, res)
    got = got[np.lexsort((got["neighbor_id"], got["query_id"]))]
    keep = np.flatnonzero(np.isfinite(s3).all(axis=1))
    s = s3[keep].astype(np.float64)
    q = q3.astype(np.float64)
    want = oracle.neighbors("frn", {"id": np.arange(len(q)), "x": q[:, 0], "y": q[:, 1], "z": q[:, 2]},
                            {"id": keep, "x": s[:, 0], "y": s[:, 1], "z": s[:, 2]}, radius=0.08, k=16)
    want = want[np.lexsort((want["nid"], want["qid"]))]
    assert len(got) == len(want) > 0
    assert np.array_equal(got["query_id"], want["qid"])
    assert np.array_equal(got["neighbor_id"], want["nid"])
    assert np.array_equal(got["distance"], want["dist"])


def _sorted_rows(rows, key):
    order = np.lexsort((rows["neighbor_id"], rows["distance"], rows[key]))
    return rows[order]


def test_per_scene_leaf_size_changes_the_tree_not_the_rows(gpu_ready):
    """prepare_point_instances(leaf_size=L): every leaf of the scene holds at
    most L points (0 takes the process tuning), the boxes stay exact unions,
    and fixed-radius / KNN rows equal those of the default scene."""
    _say("case: per-scene leaf size 1 / 4 / 8 vs the default")
    from rtdl_amd import autotune as AT

    rng = np.random.default_rng(100)
    centres = rng.random((64, 3))
    pts = cases.points(rng, 60000, 3, kind="mixture", centres=centres)
    s3 = np.stack([pts["x"], pts["y"], pts["z"]], 1).astype(np.float32)
    q = cases.points(rng, 3000, 3, kind="mixture", centres=centres)
    q3 = DeviceBuffer.from_numpy(np.stack([q["x"], q["y"], q["z"]], 1).astype(np.float32))
    want = None
    for leaf in (0, 1, 4, 8):
        scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, len(s3)], leaf_size=leaf)
        sc = W.export_scene(scene.handle)
        link = sc["nodes"]["link"].astype(np.int64)
        counts = ((link[(link & W.LEAF) != 0] >> 28) & 7) + 1
        assert counts.max() <= (leaf or AT.get_tuning().point_leaf), (leaf, counts.max())
        _check_boxes(sc, sc["prims"][:, :3])
        jobs = [B.BatchJob(scene, B.JOB_FIXED_RADIUS, q3, radius=0.02, k=64), B.BatchJob(scene, B.JOB_KNN, q3, k=8)]
        res, _ = B.run_batch(jobs)
        assert all(r["status"] == 0 for r in res)
        got = (_sorted_rows(B.job_rows(jobs[0], res[0]), "query_id"),
               _sorted_rows(B.job_rows(jobs[1], res[1]), "query_id"))
        if want is None:
            want = got
            assert len(want[0]) > 3000
            continue
        for a, b in zip(got, want):
            for name in b.dtype.names:
                assert np.array_equal(a[name], b[name]), (leaf, name)
    with pytest.raises(RuntimeError, match="leaf_size"):
        B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, len(s3)], leaf_size=9)
