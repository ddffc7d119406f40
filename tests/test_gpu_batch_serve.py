"""GPU parity of the batched and served paths against the oracle.

Batch rows carry job-local query ids and are compacted per job in
wavefront order, so they are compared after the canonical (query, id)
ordering the parity contract prescribes; every value must match the oracle
bit for bit.  The serving test also checks that a request's rows are the
same when it runs alone and when it runs inside a mixed stream.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import native as oracle
from rtdl_amd import batch as B
from rtdl_amd.device import DeviceBuffer
from tests import cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _dev(a):
    return DeviceBuffer.from_numpy(a, dtype=np.float32)


def _dev_i32(a):
    return DeviceBuffer.from_numpy(a, dtype=np.int32)


def _canon(rows, keys):
    order = np.lexsort(tuple(rows[k] for k in reversed(keys)))
    return rows[order]


def _check_nn(rows, want, ranked):
    got = _canon(rows, ["query_id", "neighbor_id"])
    w = want[np.lexsort((want["nid"], want["qid"]))]
    assert len(got) == len(w), (len(got), len(w))
    assert np.array_equal(got["query_id"], w["qid"])
    assert np.array_equal(got["neighbor_id"], w["nid"])
    assert np.array_equal(got["distance"], w["dist"])
    if ranked:
        assert np.array_equal(got["neighbor_rank"], w["rank"])


def _points3(rng, n, kind="uniform", centres=None):
    p = cases.points(rng, n, 3, kind=kind, centres=centres)
    return p, np.stack([p["x"], p["y"], p["z"]], axis=1)


def _prepared_points(s3):
    import rtdl_amd as rt
    from rtdl_amd import mi355x_runtime as m

    arr, dim = m.pack_points({"id": np.arange(len(s3), dtype=np.uint32), "x": s3[:, 0], "y": s3[:, 1],
                              "z": s3[:, 2]})
    return B.Instances(m._call_prepare("rtdl_mi355x_prepare_points_3d", m._ptr(arr), len(arr)),
                       np.array([0, len(arr)]), "points")


@pytest.mark.parametrize("morton", [True, False])
def test_batch_mixed_jobs_match_oracle(gpu_ready, morton):
    _say(f"case: batch mixed jobs morton={morton}")
    rng = np.random.default_rng(77)
    centres = rng.random((64, 3))
    sdict, s3 = _points3(rng, 60000, "mixture", centres)
    scene = _prepared_points(s3)
    tris = cases.triangles(rng, 30000, sigma=0.01)
    _say("  prepare_triangle_instances")
    tri_scene = B.prepare_triangle_instances(_dev(cases.tri_matrix(tris)), [0, 30000])
    polys = cases.star_polygons(rng, 800, rmin=0.01, rmax=0.04)
    _say("  prepare_polygon_instances")
    refs = polys["refs"]
    poly_scene = B.prepare_polygon_instances(
        _dev_i32(refs["vertex_offset"]), _dev_i32(refs["vertex_count"]),
        _dev(polys["vertices_xy"].reshape(-1, 2)), [0, len(refs)])
    jobs, wants = [], []
    for kind, k, radius, n in [(B.JOB_FIXED_RADIUS, 64, 0.03, 3000), (B.JOB_FIXED_RADIUS, 5, 0.05, 100),
                               (B.JOB_KNN, 8, 0.0, 2500), (B.JOB_KNN, 16, 0.0, 777),
                               (B.JOB_BOUNDED_KNN, 6, 0.02, 1000)]:
        qd, q3 = _points3(rng, n, "mixture", centres)
        jobs.append(B.BatchJob(scene, kind, _dev(q3), radius=radius, k=k))
        mode = {B.JOB_FIXED_RADIUS: "frn", B.JOB_KNN: "knn", B.JOB_BOUNDED_KNN: "bounded_knn"}[kind]
        wants.append(("nn", kind != B.JOB_FIXED_RADIUS,
                      oracle.neighbors(mode, qd, sdict, radius=radius, k=k)))
    for kind, mode in [(B.JOB_RAY_ANYHIT, "any"), (B.JOB_RAY_HITCOUNT, "count"), (B.JOB_RAY_CLOSEST, "closest")]:
        r = cases.rays(rng, 4000, tmax=0.3)
        rm = cases.ray_matrix(r)
        jobs.append(B.BatchJob(tri_scene, kind, _dev(rm)))
        wants.append(("ray", mode, oracle.rays_3d(mode, rm, cases.tri_matrix(tris), tris["id"])))
    pts = cases.points(rng, 20000, 2)
    jobs.append(B.BatchJob(poly_scene, B.JOB_PIP_POSITIVE, _dev(np.stack([pts["x"], pts["y"]], axis=1))))
    wants.append(("pip", None, oracle.pip_pairs(pts["x"], pts["y"], refs["vertex_offset"], refs["vertex_count"],
                                                polys["vertices_xy"], positive_only=True)))
    _say("  run_batch: 9 jobs over 4 families")
    results, seconds = B.run_batch(jobs, morton_sort=morton)
    assert seconds > 0
    for job, res, (what, extra, want) in zip(jobs, results, wants):
        assert res["status"] == 0, res
        rows = B.job_rows(job, res)
        if what == "nn":
            _check_nn(rows, want, extra)
        elif what == "ray":
            if extra == "closest":
                tid, tt, has = want
                got = _canon(rows, ["ray_id"])
                assert np.array_equal(got["ray_id"], np.flatnonzero(has))
                assert np.array_equal(got["triangle_id"], tid[has])
                assert np.array_equal(got["t"], tt[has])
            else:
                got = _canon(rows, ["ray_id"])
                assert np.array_equal(got["ray_id"], np.arange(len(want)))
                assert np.array_equal(got[got.dtype.names[1]], want)
        else:
            got = _canon(rows, ["point_id", "polygon_id"])
            assert np.array_equal(got["point_id"], want[:, 0])
            assert np.array_equal(got["polygon_id"], refs["id"][want[:, 1]])


def test_multi_instance_points_and_polygons(gpu_ready):
    _say("case: multi-instance FRN (2-D and 3-D instances) and PIP instances")
    rng = np.random.default_rng(78)
    sizes = [1000, 37, 5000, 1, 2500, 0, 800]
    dims = [2, 3, 3, 2, 2, 3, 3]
    coords, offs, inst = [], [0], []
    for n, d in zip(sizes, dims):
        p = rng.random((n, 3)).astype(np.float32)
        if d == 2:
            p[:, 2] = 0
        coords.append(p)
        offs.append(offs[-1] + n)
        inst.append(p)
    _say("  prepare_point_instances (7 instances)")
    scene = B.prepare_point_instances(_dev(np.concatenate(coords)), offs)
    jobs, wants = [], []
    for i, (n, d) in enumerate(zip(sizes, dims)):
        q = rng.random((700, 3)).astype(np.float32)
        if d == 2:
            q[:, 2] = 0
        radius = 0.05
        jobs.append(B.BatchJob(scene, B.JOB_FIXED_RADIUS, _dev(q[:, :d]), instance=i, radius=radius, k=16))
        qd = {"id": np.arange(700, dtype=np.uint32), "x": q[:, 0].astype(np.float64), "y": q[:, 1].astype(np.float64)}
        sd = {"id": np.arange(n, dtype=np.uint32), "x": inst[i][:, 0].astype(np.float64),
              "y": inst[i][:, 1].astype(np.float64)}
        if d == 3:
            qd["z"] = q[:, 2].astype(np.float64)
            sd["z"] = inst[i][:, 2].astype(np.float64)
        wants.append(oracle.neighbors("frn", qd, sd, radius=radius, k=16))
    _say("  run_batch: 7 fixed-radius jobs, k_max 16")
    results, _ = B.run_batch(jobs)
    for job, res, want in zip(jobs, results, wants):
        assert res["status"] == 0
        _check_nn(B.job_rows(job, res), want, False)

    polys = [cases.star_polygons(rng, 300) for _ in range(3)]
    voff, vcnt, verts, poffs = [], [], [], [0]
    base = 0
    for pset in polys:
        voff.append(pset["refs"]["vertex_offset"] + base)
        vcnt.append(pset["refs"]["vertex_count"])
        verts.append(pset["vertices_xy"].reshape(-1, 2))
        base += len(verts[-1])
        poffs.append(poffs[-1] + len(pset["refs"]))
    _say("  prepare_polygon_instances (3 instances)")
    pscene = B.prepare_polygon_instances(_dev_i32(np.concatenate(voff)), _dev_i32(np.concatenate(vcnt)),
                                         _dev(np.concatenate(verts)), poffs)
    jobs, wants = [], []
    for i, pset in enumerate(polys):
        pts = cases.points(rng, 30000, 2)
        jobs.append(B.BatchJob(pscene, B.JOB_PIP_POSITIVE, _dev(np.stack([pts["x"], pts["y"]], 1)), instance=i))
        r = pset["refs"]
        wants.append(oracle.pip_pairs(pts["x"], pts["y"], r["vertex_offset"], r["vertex_count"],
                                      pset["vertices_xy"], positive_only=True))
    _say("  run_batch: 3 PIP jobs")
    results, _ = B.run_batch(jobs)
    for job, res, want in zip(jobs, results, wants):
        assert res["status"] == 0
        got = _canon(B.job_rows(job, res), ["point_id", "polygon_id"])
        assert np.array_equal(got["point_id"], want[:, 0])
        assert np.array_equal(got["polygon_id"], want[:, 1])


def test_serve_replay_results_same_alone_and_in_mix(gpu_ready):
    _say("case: serve replay, alone vs in mix")
    rng = np.random.default_rng(79)
    sdict, s3 = _points3(rng, 40000)
    pscene = _prepared_points(s3)
    tris = cases.triangles(rng, 20000, sigma=0.01)
    _say("  prepare_triangle_instances")
    tscene = B.prepare_triangle_instances(_dev(cases.tri_matrix(tris)), [0, 20000])
    polys = cases.star_polygons(rng, 500, rmin=0.01, rmax=0.04)
    refs = polys["refs"]
    gscene = B.prepare_polygon_instances(_dev_i32(refs["vertex_offset"]), _dev_i32(refs["vertex_count"]),
                                         _dev(polys["vertices_xy"].reshape(-1, 2)), [0, len(refs)])
    n_req = 400
    kinds = rng.choice([B.JOB_FIXED_RADIUS, B.JOB_KNN, B.JOB_RAY_ANYHIT, B.JOB_PIP_POSITIVE], n_req)
    sizes = np.floor(np.exp(rng.uniform(0, np.log(3000), n_req))).astype(np.int64)
    arrivals = np.cumsum(rng.exponential(1 / 20000.0, n_req))
    pq = rng.random((int(sizes.sum()), 3)).astype(np.float32)
    rq = np.concatenate([cases.ray_matrix(cases.rays(rng, int(sizes.sum()), tmax=0.3))], axis=0).astype(np.float32)
    gq = rng.random((int(sizes.sum()), 2)).astype(np.float32)
    reqs = np.zeros(n_req, dtype=B.SERVE_REQUEST_DTYPE)
    off = 0
    scene_of = {B.JOB_FIXED_RADIUS: 0, B.JOB_KNN: 0, B.JOB_RAY_ANYHIT: 1, B.JOB_PIP_POSITIVE: 2}
    for i in range(n_req):
        reqs[i] = (kinds[i], scene_of[kinds[i]], off, sizes[i], 1 if i % 40 == 3 else 0, arrivals[i])
        off += sizes[i]
    arenas = dict(point_queries=_dev(pq), ray_queries=_dev(rq), pip_queries=_dev(gq))
    scenes = [pscene, tscene, gscene]
    kw = dict(radius=0.02, k_max=64, k=8, slots=4, **arenas)
    _say("  serve_replay: 400 mixed requests")
    results, stats, captured = B.serve_replay(scenes, reqs, **kw)
    assert np.all(results["status"] == 0)
    assert np.all(results["done_s"] >= results["arrival_s"])
    assert stats["queries"] == sizes.sum()
    assert stats["batches"] >= 2
    for i, rows in captured.items():
        alone = reqs[i:i + 1].copy()
        alone["arrival_s"] = 0.0
        _say(f"  serve_replay: request {i} alone")
        res1, _, cap1 = B.serve_replay(scenes, alone, **kw)
        assert res1["rows"][0] == results["rows"][i]
        kind = int(reqs["kind"][i])
        keys = {B.JOB_FIXED_RADIUS: ["query_id", "neighbor_id"], B.JOB_KNN: ["query_id", "neighbor_id"],
                B.JOB_RAY_ANYHIT: ["ray_id"], B.JOB_PIP_POSITIVE: ["point_id", "polygon_id"]}[kind]
        a, b = _canon(rows, keys), _canon(cap1[0], keys)
        # field by field: a structured copy leaves the row dtype's padding bytes undefined
        assert len(a) == len(b)
        for name in a.dtype.names:
            assert np.array_equal(a[name], b[name]), name
        o, n = int(reqs["query_offset"][i]), int(reqs["query_count"][i])
        if kind in (B.JOB_FIXED_RADIUS, B.JOB_KNN):
            q = pq[o:o + n].astype(np.float64)
            qd = {"id": np.arange(n, dtype=np.uint32), "x": q[:, 0], "y": q[:, 1], "z": q[:, 2]}
            mode = "frn" if kind == B.JOB_FIXED_RADIUS else "knn"
            want = oracle.neighbors(mode, qd, sdict, radius=0.02, k=64 if mode == "frn" else 8)
            _check_nn(rows, want, mode == "knn")
        elif kind == B.JOB_RAY_ANYHIT:
            want = oracle.rays_3d("any", rq[o:o + n].astype(np.float64), cases.tri_matrix(tris), tris["id"])
            assert np.array_equal(a["any_hit"], want)
        else:
            q = gq[o:o + n].astype(np.float64)
            want = oracle.pip_pairs(q[:, 0], q[:, 1], refs["vertex_offset"], refs["vertex_count"],
                                    polys["vertices_xy"], positive_only=True)
            assert np.array_equal(a["point_id"], want[:, 0])
            assert np.array_equal(a["polygon_id"], refs["id"][want[:, 1]])


def test_serve_admission_rejects_oversized_requests(gpu_ready):
    _say("case: serve admission budget")
    rng = np.random.default_rng(80)
    _, s3 = _points3(rng, 1000)
    scene = _prepared_points(s3)
    reqs = np.zeros(3, dtype=B.SERVE_REQUEST_DTYPE)
    reqs[0] = (B.JOB_FIXED_RADIUS, 0, 0, 10, 0, 0.0)
    reqs[1] = (B.JOB_FIXED_RADIUS, 0, 10, 5000, 0, 0.0)  # 5000 * 64 rows * 16 B > budget / slots
    reqs[2] = (B.JOB_FIXED_RADIUS, 0, 0, 10, 0, 0.0)
    pq = _dev(rng.random((6000, 3)))
    res, stats, _ = B.serve_replay([scene], reqs, radius=0.05, k_max=64, k=1, point_queries=pq,
                                   slots=2, memory_budget_bytes=2 * 1024 * 1024)
    assert res["status"].tolist() == [0, B.STATUS_REJECTED, 0]


def test_batch_pip_points_inside_many_polygons(gpu_ready):
    _say("case: batched PIP, points inside up to 40 polygons (past the 16-hit register list)")
    rng = np.random.default_rng(81)
    # 40 nested squares around (0.5, 0.5) plus random star polygons
    verts, voff, vcnt = [], [], []
    for i in range(40):
        h = 0.01 + 0.01 * i
        voff.append(sum(vcnt))
        vcnt.append(4)
        verts += [(0.5 - h, 0.5 - h), (0.5 + h, 0.5 - h), (0.5 + h, 0.5 + h), (0.5 - h, 0.5 + h)]
    stars = cases.star_polygons(rng, 300, rmin=0.01, rmax=0.05)
    sv = stars["vertices_xy"].reshape(-1, 2)
    base = len(verts)
    for r in stars["refs"]:
        voff.append(base + int(r["vertex_offset"]))
        vcnt.append(int(r["vertex_count"]))
    verts = np.concatenate([np.asarray(verts, dtype=np.float64), sv])
    # FP32 on the device: the oracle gets the same FP32-rounded values
    verts32 = verts.astype(np.float32)
    voff = np.asarray(voff, dtype=np.int64)
    vcnt = np.asarray(vcnt, dtype=np.int64)
    scene = B.prepare_polygon_instances(_dev_i32(voff), _dev_i32(vcnt), _dev(verts32), [0, len(voff)])
    pts = np.concatenate([0.5 + 0.05 * rng.standard_normal((3000, 2)), rng.random((2000, 2))]).astype(np.float32)
    # the default budget is 16 rows per point on average; these points average more
    job = B.BatchJob(scene, B.JOB_PIP_POSITIVE, _dev(pts), row_capacity=64 * len(pts))
    _say("  run_batch: 1 PIP job, 5000 points")
    (res,), _ = B.run_batch([job])
    want = oracle.pip_pairs(pts[:, 0].astype(np.float64), pts[:, 1].astype(np.float64), voff, vcnt,
                            verts32.astype(np.float64).reshape(-1), positive_only=True)
    per_point = np.bincount(want[:, 0], minlength=len(pts))
    assert per_point.max() > 16  # the case exercises the overflow pass
    assert res["status"] == 0
    got = _canon(B.job_rows(job, res), ["point_id", "polygon_id"])
    assert np.array_equal(got["point_id"], want[:, 0])
    assert np.array_equal(got["polygon_id"], want[:, 1])


def test_batch_knn_lattice_ties(gpu_ready):
    _say("case: batched KNN on a lattice with many equal distances")
    rng = np.random.default_rng(82)
    g = np.stack(np.meshgrid(np.arange(14), np.arange(14), np.arange(14), indexing="ij"), -1).reshape(-1, 3)
    order = rng.permutation(len(g))
    g = g[order]  # prepared points carry ids 0..n-1 in input order
    s3 = (g / 16.0).astype(np.float64)
    sdict = {"id": np.arange(len(g), dtype=np.uint32), "x": s3[:, 0], "y": s3[:, 1], "z": s3[:, 2]}
    scene = _prepared_points(s3)
    jobs, wants = [], []
    for k in (1, 6, 8, 27, 64):
        qg = rng.integers(0, 28, (700, 3)) / 32.0
        qd = {"id": np.arange(len(qg), dtype=np.uint32), "x": qg[:, 0], "y": qg[:, 1], "z": qg[:, 2]}
        jobs.append(B.BatchJob(scene, B.JOB_KNN, _dev(qg), k=k))
        wants.append(oracle.neighbors("knn", qd, sdict, k=k))
    _say("  run_batch: 5 KNN jobs")
    results, _ = B.run_batch(jobs)
    for job, res, want in zip(jobs, results, wants):
        assert res["status"] == 0
        _check_nn(B.job_rows(job, res), want, True)


def _check_frn_order(rows):
    """Each query's rows are contiguous and in (distance, neighbour id) order."""
    q = rows["query_id"]
    starts = np.flatnonzero(np.r_[True, q[1:] != q[:-1]])
    assert len(np.unique(q)) == len(starts), "a query's rows are not contiguous"
    d, n = rows["distance"], rows["neighbor_id"]
    same = q[1:] == q[:-1]
    ok = (d[1:] > d[:-1]) | ((d[1:] == d[:-1]) & (n[1:] > n[:-1]))
    assert np.all(ok | ~same), "rows of a query are not in (distance, id) order"


def test_batch_fixed_radius_dense_heavy_and_light(gpu_ready):
    _say("case: batched fixed radius, queries with more and fewer neighbours than k")
    rng = np.random.default_rng(91)
    centres = rng.random((8, 3))
    sdict, s3 = _points3(rng, 120000, "mixture", centres)
    scene = _prepared_points(s3)
    jobs, wants = [], []
    for k, radius, kind in ((64, 0.02, "mixture"), (16, 0.01, "mixture"), (5, 0.03, "uniform"),
                            (64, 0.002, "uniform")):
        qd, q3 = _points3(rng, 3000, kind, centres)
        jobs.append(B.BatchJob(scene, B.JOB_FIXED_RADIUS, _dev(q3), radius=radius, k=k))
        wants.append(oracle.neighbors("frn", qd, sdict, radius=radius, k=k))
    _say("  run_batch: 4 fixed-radius jobs")
    results, _ = B.run_batch(jobs)
    for job, res, want in zip(jobs, results, wants):
        assert res["status"] == 0
        rows = B.job_rows(job, res)
        _check_frn_order(rows)
        _check_nn(rows, want, False)


def test_batch_fixed_radius_lattice_ties_and_duplicates(gpu_ready):
    _say("case: batched fixed radius on a lattice with duplicate points")
    rng = np.random.default_rng(92)
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(12), indexing="ij"), -1).reshape(-1, 3)
    # 300 copies of one lattice site: more than the candidate parking of one query holds
    g = np.concatenate([g, np.repeat(np.array([[5, 5, 5]]), 300, axis=0)])
    g = g[rng.permutation(len(g))]
    s3 = (g / 16.0).astype(np.float64)
    sdict = {"id": np.arange(len(g), dtype=np.uint32), "x": s3[:, 0], "y": s3[:, 1], "z": s3[:, 2]}
    scene = _prepared_points(s3)
    jobs, wants = [], []
    for k, radius in ((20, 2.5 / 16), (64, 3.2 / 16), (7, 1.0 / 16), (64, 0.0)):
        qg = np.concatenate([rng.integers(0, 24, (600, 3)) / 32.0, np.array([[5, 5, 5], [5, 5, 6]]) / 16.0])
        qd = {"id": np.arange(len(qg), dtype=np.uint32), "x": qg[:, 0], "y": qg[:, 1], "z": qg[:, 2]}
        jobs.append(B.BatchJob(scene, B.JOB_FIXED_RADIUS, _dev(qg), radius=radius, k=k))
        wants.append(oracle.neighbors("frn", qd, sdict, radius=radius, k=k))
    _say("  run_batch: 4 fixed-radius lattice jobs")
    results, _ = B.run_batch(jobs)
    for job, res, want in zip(jobs, results, wants):
        assert res["status"] == 0
        rows = B.job_rows(job, res)
        _check_frn_order(rows)
        _check_nn(rows, want, False)


def test_batch_fixed_radius_points_on_bucket_edges_and_radius(gpu_ready):
    _say("case: batched fixed radius, neighbours on distance-bucket edges and on the radius, +-1 ulp")
    rng = np.random.default_rng(93)
    radius = 0.25  # bucket edges j * radius / 32 = j / 128 are FP32 values
    q0 = np.array([0.5, 0.5, 0.5])
    pts = []
    for j in range(1, 33):
        e = j / 128.0
        for sign in (-1.0, 1.0):
            for axis in range(3):
                for off in (np.float32(e), np.nextafter(np.float32(e), np.float32(0)),
                            np.nextafter(np.float32(e), np.float32(1))):
                    p = q0.copy()
                    p[axis] = np.float32(q0[axis] + sign * np.float64(off))
                    pts.append(p)
    # diagonal points: irrational distances close to the edges
    for j in range(1, 33):
        d = j / 128.0 / np.sqrt(3.0)
        pts.append(np.float32(q0 + d))
        pts.append(np.float32(q0 - d))
    pts = np.concatenate([np.array(pts), rng.random((2000, 3))]).astype(np.float32).astype(np.float64)
    pts = pts[rng.permutation(len(pts))]
    sdict = {"id": np.arange(len(pts), dtype=np.uint32), "x": pts[:, 0], "y": pts[:, 1], "z": pts[:, 2]}
    scene = _prepared_points(pts)
    qg = np.concatenate([q0[None, :], q0[None, :] + np.array([[1 / 128.0, 0, 0]]), rng.random((300, 3))])
    qg = qg.astype(np.float32).astype(np.float64)
    qd = {"id": np.arange(len(qg), dtype=np.uint32), "x": qg[:, 0], "y": qg[:, 1], "z": qg[:, 2]}
    jobs, wants = [], []
    for k, r in ((64, radius), (8, radius), (1, radius), (64, 1.0 / 128.0)):
        jobs.append(B.BatchJob(scene, B.JOB_FIXED_RADIUS, _dev(qg), radius=r, k=k))
        wants.append(oracle.neighbors("frn", qd, sdict, radius=r, k=k))
    _say("  run_batch: 4 fixed-radius jobs on bucket edges")
    results, _ = B.run_batch(jobs)
    for job, res, want in zip(jobs, results, wants):
        assert res["status"] == 0
        rows = B.job_rows(job, res)
        _check_frn_order(rows)
        _check_nn(rows, want, False)


def test_batch_knn_small_scene_and_duplicates(gpu_ready):
    _say("case: batched KNN with k above the scene size and with duplicate points")
    rng = np.random.default_rng(93)
    jobs, wants = [], []
    tiny = rng.random((5, 3))
    tiny_d = {"id": np.arange(5, dtype=np.uint32), "x": tiny[:, 0], "y": tiny[:, 1], "z": tiny[:, 2]}
    tiny_scene = _prepared_points(tiny)
    base = rng.random((4000, 3))
    dup = np.concatenate([base, np.repeat(base[:1], 250, axis=0), np.repeat(base[1:2], 40, axis=0)])
    dup = dup[rng.permutation(len(dup))]
    dup_d = {"id": np.arange(len(dup), dtype=np.uint32), "x": dup[:, 0], "y": dup[:, 1], "z": dup[:, 2]}
    dup_scene = _prepared_points(dup)
    for scene, sd, kind, k, radius in ((tiny_scene, tiny_d, B.JOB_KNN, 8, 0.0),
                                       (tiny_scene, tiny_d, B.JOB_KNN, 64, 0.0),
                                       (dup_scene, dup_d, B.JOB_KNN, 64, 0.0),
                                       (dup_scene, dup_d, B.JOB_KNN, 3, 0.0),
                                       (dup_scene, dup_d, B.JOB_BOUNDED_KNN, 64, 0.05)):
        q = np.concatenate([rng.random((500, 3)), base[:2], base[:2] + 1e-3])
        q = q.astype(np.float32).astype(np.float64)  # batched queries are FP32
        qd = {"id": np.arange(len(q), dtype=np.uint32), "x": q[:, 0], "y": q[:, 1], "z": q[:, 2]}
        jobs.append(B.BatchJob(scene, kind, _dev(q), radius=radius, k=k))
        mode = "knn" if kind == B.JOB_KNN else "bounded_knn"
        wants.append(oracle.neighbors(mode, qd, sd, radius=radius, k=k))
    _say("  run_batch: 5 KNN jobs")
    results, _ = B.run_batch(jobs)
    for job, res, want in zip(jobs, results, wants):
        assert res["status"] == 0
        rows = B.job_rows(job, res)
        _check_frn_order(rows)
        _check_nn(rows, want, True)


def _rows_canon(job, res):
    rows = B.job_rows(job, res)
    names = rows.dtype.names
    keys = [n for n in ("query_id", "ray_id", "point_id", "neighbor_id", "polygon_id", "triangle_id")
            if n in names]
    return _canon(rows, keys)


def test_batch_graph_replays_match_eager_runs(gpu_ready):
    _say("case: batch captured into a hipGraph, replayed between eager batches")
    rng = np.random.default_rng(94)
    centres = rng.random((64, 3))
    _, s3 = _points3(rng, 40000, "mixture", centres)
    scene = _prepared_points(s3)
    tris = cases.triangles(rng, 20000, sigma=0.01)
    tri_scene = B.prepare_triangle_instances(_dev(cases.tri_matrix(tris)), [0, 20000])
    polys = cases.star_polygons(rng, 500, rmin=0.01, rmax=0.04)
    refs = polys["refs"]
    poly_scene = B.prepare_polygon_instances(
        _dev_i32(refs["vertex_offset"]), _dev_i32(refs["vertex_count"]),
        _dev(polys["vertices_xy"].reshape(-1, 2)), [0, len(refs)])

    def queries():
        _, q3 = _points3(rng, 3000, "mixture", centres)
        o = rng.random((2000, 3))
        d = rng.standard_normal((2000, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        rays = np.concatenate([o, d, np.full((2000, 1), 0.3)], 1).astype(np.float32)
        pts = rng.random((2500, 2)).astype(np.float32)
        return q3.astype(np.float32), rays, pts

    q3, rays, pts = queries()
    qbuf, rbuf, pbuf = _dev(q3), _dev(rays), _dev(pts)
    jobs = [B.BatchJob(scene, B.JOB_FIXED_RADIUS, qbuf, radius=0.02, k=32),
            B.BatchJob(scene, B.JOB_KNN, qbuf, k=8),
            B.BatchJob(scene, B.JOB_BOUNDED_KNN, qbuf, radius=0.01, k=4),
            B.BatchJob(tri_scene, B.JOB_RAY_ANYHIT, rbuf),
            B.BatchJob(tri_scene, B.JOB_RAY_CLOSEST, rbuf),
            B.BatchJob(poly_scene, B.JOB_PIP_POSITIVE, pbuf)]
    other = [B.BatchJob(scene, B.JOB_KNN, _dev(q3[::-1].copy()), k=5)]

    def eager():
        res, _ = B.run_batch(jobs)
        return res, [_rows_canon(j, r) for j, r in zip(jobs, res)]

    want_res, want_rows = eager()
    with B.BatchGraph(jobs) as graph:
        for rep in range(4):
            if rep == 2:
                # new query contents in the same buffers: the replay reads them
                q3, rays, pts = queries()
                qbuf.upload(q3)
                rbuf.upload(rays)
                pbuf.upload(pts)
                want_res, want_rows = eager()
            if rep == 1:
                B.run_batch(other)  # eager work between replays
            _say(f"  replay {rep}")
            res, _ = graph.launch()
            for j, (job, r, w, wr) in enumerate(zip(jobs, res, want_res, want_rows)):
                assert r == w, (rep, j, r, w)
                got = _rows_canon(job, r)
                for name in got.dtype.names:
                    assert np.array_equal(got[name], wr[name]), (rep, j, name)


def test_autotune_picks_the_fastest_and_every_candidate_gives_the_same_rows(gpu_ready):
    from rtdl_amd import autotune as T

    _say("case: autotuner over leaf sizes, launch sizes, KNN stack depth and walk workgroup size, "
         "rows checked per candidate")
    rng = np.random.default_rng(95)
    centres = rng.random((64, 3))
    _, s3 = _points3(rng, 30000, "mixture", centres)
    tris = cases.triangles(rng, 15000, sigma=0.01)
    tri_m = cases.tri_matrix(tris)
    _, q3 = _points3(rng, 4000, "mixture", centres)
    o = rng.random((3000, 3))
    d = rng.standard_normal((3000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.concatenate([o, d, np.full((3000, 1), 0.3)], 1).astype(np.float32)
    qbuf, rbuf = _dev(q3), _dev(rays)
    pts_buf, tri_buf = _dev(s3), _dev(tri_m)

    def build():
        return (B.prepare_point_instances(pts_buf, [0, len(s3)]),
                B.prepare_triangle_instances(tri_buf, [0, len(tri_m)]))

    def jobs(scenes):
        pts, tri = scenes
        return [B.BatchJob(pts, B.JOB_FIXED_RADIUS, qbuf, radius=0.02, k=64),
                B.BatchJob(pts, B.JOB_KNN, qbuf, k=8),
                B.BatchJob(tri, B.JOB_RAY_CLOSEST, rbuf)]

    original = T.get_tuning()
    try:
        out = T.autotune_batch(build, jobs, repeats=2,
                               knobs={"blocks_per_cu": (4, 16), "point_leaf": (4, 8), "triangle_leaf": (2, 4),
                                      "knn_stack": (16, 48), "walk_block": (64, 128)})
        table = out["table"]
        assert len(table) == 11  # the start and two values of each knob, every one with the start's rows
        assert {r["knn_stack"] for r in table} == {16, 32, 48}
        assert {r["walk_block"] for r in table} == {64, 128, 256}
        fastest = min(table, key=lambda r: r["seconds"])
        assert out["best"] == T.Tuning(**{k: fastest[k] for k in ("point_leaf", "triangle_leaf", "polygon_leaf",
                                                                  "blocks_per_cu", "knn_stack", "walk_block")})
        assert T.get_tuning() == out["best"]
    finally:
        T.set_tuning(original)
    assert T.get_tuning() == original


def test_serve_admission_waits_for_memory_under_a_small_budget(gpu_ready):
    """The device budget binds: 40 fixed-radius requests of 1,000-1,600
    queries (1.0-1.6 MB of rows each) arrive at once against a 4 MB budget
    and 4 slots.  Rows in flight never exceed the budget, requests wait for
    memory instead of being rejected, and every request's rows equal the
    oracle's."""
    _say("case: serve admission bound by a 4 MB budget")
    rng = np.random.default_rng(82)
    sdict, s3 = _points3(rng, 30000)
    scene = _prepared_points(s3)
    sizes = rng.integers(1000, 1601, 40)
    pq = rng.random((int(sizes.sum()), 3)).astype(np.float32)
    reqs = np.zeros(40, dtype=B.SERVE_REQUEST_DTYPE)
    off = 0
    for i in range(40):
        reqs[i] = (B.JOB_FIXED_RADIUS, 0, off, sizes[i], 1, 0.0)
        off += sizes[i]
    budget = 4 << 20
    # k_max 64: 64 rows x 16 B = 1 KB per query
    res, stats, captured = B.serve_replay([scene], reqs, radius=0.03, k_max=64, k=1, point_queries=_dev(pq),
                                          slots=4, memory_budget_bytes=budget)
    assert np.all(res["status"] == 0)
    assert stats["region_bytes"] <= budget
    assert stats["peak_inflight_bytes"] <= budget
    assert stats["peak_inflight_bytes"] > budget // 2  # the budget was used, not a smaller cap
    assert stats["batches"] >= 40 * 1.0e6 // budget    # admission waited for completions
    for i in range(40):
        o, n = int(reqs["query_offset"][i]), int(reqs["query_count"][i])
        q = pq[o:o + n].astype(np.float64)
        qd = {"id": np.arange(n, dtype=np.uint32), "x": q[:, 0], "y": q[:, 1], "z": q[:, 2]}
        _check_nn(captured[i], oracle.neighbors("frn", qd, sdict, radius=0.03, k=64), False)


def _frn_requests(rng, n_req, sizes_hi, scene_of=None, capture=True):
    sizes = rng.integers(200, sizes_hi, n_req)
    reqs = np.zeros(n_req, dtype=B.SERVE_REQUEST_DTYPE)
    off = 0
    for i in range(n_req):
        reqs[i] = (B.JOB_FIXED_RADIUS, 0 if scene_of is None else scene_of(i), off, sizes[i], int(capture), 0.0)
        off += sizes[i]
    return reqs, int(sizes.sum())


def _check_frn_request(captured, reqs, i, pq, sdict, radius, k):
    o, n = int(reqs["query_offset"][i]), int(reqs["query_count"][i])
    q = pq[o:o + n].astype(np.float64)
    qd = {"id": np.arange(n, dtype=np.uint32), "x": q[:, 0], "y": q[:, 1], "z": q[:, 2]}
    _check_nn(captured[i], oracle.neighbors("frn", qd, sdict, radius=radius, k=k), False)


def test_serve_failures_stay_with_their_requests(gpu_ready):
    """A launch failure in a batch (forced on the host: RTDL_MI355X_FAULT
    "launch:3") sends the batch's requests through again one at a time, and
    requests naming a missing scene or a scene of the wrong type are marked
    invalid; every other request is served with exact rows."""
    from rtdl_amd import mi355x_runtime as m

    _say("case: serve failure isolation (forced launch failure, invalid requests)")
    rng = np.random.default_rng(83)
    sdict, s3 = _points3(rng, 20000)
    pscene = _prepared_points(s3)
    tris = cases.triangles(rng, 2000, sigma=0.01)
    tscene = B.prepare_triangle_instances(_dev(cases.tri_matrix(tris)), [0, 2000])
    bad = {5: 7, 9: 1}  # request -> scene: 7 does not exist, 1 holds triangles
    reqs, total = _frn_requests(rng, 24, 1500, scene_of=lambda i: bad.get(i, 0))
    pq = rng.random((total, 3)).astype(np.float32)
    m.set_fault("launch:3")
    try:
        res, stats, captured = B.serve_replay([pscene, tscene], reqs, radius=0.03, k_max=16, k=1,
                                              point_queries=_dev(pq), slots=2, realtime=False)
    finally:
        m.set_fault("")
    for i in range(len(reqs)):
        if i in bad:
            assert res["status"][i] == B.STATUS_INVALID and i not in captured
        else:
            assert res["status"][i] == 0, (i, res["status"][i])
            _check_frn_request(captured, reqs, i, pq, sdict, 0.03, 16)
    assert stats["retried"] >= 1 and stats["failed"] == 0


def test_serve_forced_timeout_marks_only_the_late_batch(gpu_ready):
    """"timeout:1" treats the first batch to complete as past its deadline:
    its requests carry RTDL_STATUS_TIMEOUT (their rows are complete, only
    late); every other request is on time."""
    from rtdl_amd import mi355x_runtime as m

    _say("case: serve forced timeout")
    rng = np.random.default_rng(84)
    sdict, s3 = _points3(rng, 20000)
    scene = _prepared_points(s3)
    reqs, total = _frn_requests(rng, 12, 900)
    pq = rng.random((total, 3)).astype(np.float32)
    m.set_fault("timeout:1")
    try:
        res, stats, captured = B.serve_replay([scene], reqs, radius=0.03, k_max=16, k=1, point_queries=_dev(pq),
                                              slots=1, realtime=False, max_batch_queries=2000)
    finally:
        m.set_fault("")
    first = res["batch"] == res["batch"].min()
    assert stats["batches"] >= 3 and first.any() and not first.all()
    assert np.all(res["status"][first] == B.STATUS_TIMEOUT)
    assert np.all(res["status"][~first] == 0)
    for i in range(len(reqs)):
        _check_frn_request(captured, reqs, i, pq, sdict, 0.03, 16)


def test_forced_allocation_failure_leaves_the_library_usable(gpu_ready):
    """An allocation failure (forced: "alloc") surfaces as an error of that
    call; the next call gives exact rows."""
    from rtdl_amd import mi355x_runtime as m

    _say("case: forced allocation failure, then a normal call")
    rng = np.random.default_rng(85)
    q = cases.points(rng, 3000, 3)
    s = cases.points(rng, 5000, 3)
    m.set_fault("alloc")
    try:
        with pytest.raises(RuntimeError, match="injected fault"):
            m.knn_rows_mi355x(q, s, k=4)
    finally:
        m.set_fault("")
    got = m.knn_rows_mi355x(q, s, k=4)
    want = oracle.neighbors("knn", q, s, k=4)
    assert np.array_equal(got["neighbor_id"], want["nid"]) and np.array_equal(got["distance"], want["dist"])
