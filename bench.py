#!/usr/bin/env python3
"""Benchmark: served request stream vs the plain per-call backend on 1 MI355X.

Workload (the served request mix): Poisson arrivals at 20,000 requests/s for
60 s; each request is a fixed-radius (r = 0.01, k_max = 64), KNN (k = 8), 3-D
ray any-hit (tmax = 0.5) or point-in-polygon (positive hits) query set, 25 %
each, with batch size floor(exp(U[0, ln 65536])), against 16 resident scenes
whose sizes are log-spaced in [1e5, 1e7] primitives (4 per request type,
alternating uniform / 64-component Gaussian mixture, sigma 0.02, FP32).  PIP
scenes are star polygons (8-32 vertices) whose radii U[0.005, 0.02] are
scaled by sqrt(1000 / N) so that a point lies in about as many polygons as
with 1,000 polygons.  Admission is bounded by a 200 GB device-memory budget.

The 60 s stream is replayed in one continuous serving loop and timed as
`--steps` steps of 60 / steps seconds, after `--warmup` untimed steps of
stream.  Query sets are windows of seeded per-distribution query pools
resident in HBM, so the whole stream runs with bounded host and device
memory.  `value` is the number of queries the served backend completed
divided by the wall time of the replay (inputs resident in HBM before the
clock starts).
Beside it: p50/p99 latency (arrival to rows ready in device memory), the
rows of every 1,000th request checked against the C FP64 oracle on a sample
of their queries after the timed region, the same requests replayed
saturated (all arrived at once: capacity), the plain per-call C-ABI backend
on a prefix of the stream (host records in, host rows out, scene built per
call), a per-request resident-scene run (one synchronous batch per request),
the HBM roofline of the dominant traversal kernel on a batch drawn from its
own seed, GPU lines for the other benchmark configurations (the 4,096-point
FP64 fixed-radius case, the per-call 2M x 1M any-hit and 4M x 4M KNN calls,
the 2,048-instance fixed-radius + PIP batch) and the multi-threaded C oracle
and the single-threaded pure-Python reference on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "queries/s + p99 latency, batched/served vs plain per-call backend, 1 MI355X"
SEED = 20261203 + 4
ROOF_SEED = 20261203 + 40  # the roofline batch's own stream, independent of --steps
STREAM_SECONDS = 60.0
POOL_QUERIES = 8 << 20      # queries per (request group, distribution) pool
CAPTURE_EVERY = 1000        # served requests whose rows are checked
CHECK_PER_REQUEST = 64      # sampled queries checked per captured request
RADIUS = 0.01
K_MAX = 64
K_KNN = 8
FRN_LEAF = 4   # points per leaf of the fixed-radius scenes (KNN scenes keep the default 2)
TMAX = 0.5
HBM_PEAK = 8.0e12
KINDS = ("frn", "knn", "ray", "pip")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# --------------------------------------------------------------- scenes --


def _points(rng, count, dim, dist, centres):
    if dist == "uniform":
        return rng.random((count, dim), dtype=np.float32)
    pick = rng.integers(0, 64, count)
    return (centres[pick, :dim] + 0.02 * rng.standard_normal((count, dim), dtype=np.float32)).astype(np.float32)


def star_polygons(rng, n, dist, centres):
    """n star polygons with 8-32 vertices at increasing angles (one jittered
    vertex per equal sector, so no sort is needed); radii U[0.005, 0.02]
    scaled by sqrt(1000 / n)."""
    counts = rng.integers(8, 33, n).astype(np.int64)
    offsets = np.cumsum(counts) - counts
    total = int(counts.sum())
    seg = np.repeat(np.arange(n, dtype=np.int32), counts)
    j = (np.arange(total, dtype=np.int64) - offsets[seg]).astype(np.float32)
    ang = (j + np.float32(0.9) * rng.random(total, dtype=np.float32)) * (
        np.float32(2.0 * math.pi) / counts.astype(np.float32))[seg]
    rad = (np.float32(0.005) + np.float32(0.015) * rng.random(total, dtype=np.float32)) * np.float32(
        math.sqrt(1000.0 / n))
    cen = _points(rng, n, 2, dist, centres)
    verts = np.empty((total, 2), dtype=np.float32)
    verts[:, 0] = cen[seg, 0] + rad * np.cos(ang)
    verts[:, 1] = cen[seg, 1] + rad * np.sin(ang)
    return offsets.astype(np.int32), counts.astype(np.int32), verts


def make_scenes(B, rng, max_prims: float, n_scenes: int = 16):
    """Generates the scenes on the host (seeded) and builds them resident on
    the device; returns a list of dicts holding the host copy as well."""
    from rtdl_amd.device import DeviceBuffer

    centres = rng.random((64, 3), dtype=np.float32)
    scenes = []
    for i in range(n_scenes):
        n = int(round(1e5 * (max_prims / 1e5) ** (i / (n_scenes - 1))))
        kind = KINDS[i % 4]
        dist = "uniform" if (i // 4) % 2 == 0 else "mixture"
        if kind in ("frn", "knn"):
            host = _points(rng, n, 3, dist, centres)
        elif kind == "ray":
            c = _points(rng, n, 3, dist, centres)
            host = (c[:, None, :] + 0.005 * rng.standard_normal((n, 3, 3), dtype=np.float32))
            host = host.reshape(n, 9).astype(np.float32)
        else:
            host = star_polygons(rng, n, dist, centres)
        t0 = time.perf_counter()
        if kind in ("frn", "knn"):
            dev = DeviceBuffer.from_numpy(host)
            handle = B.prepare_point_instances(dev, [0, n], leaf_size=FRN_LEAF if kind == "frn" else 0)
        elif kind == "ray":
            dev = DeviceBuffer.from_numpy(host)
            handle = B.prepare_triangle_instances(dev, [0, n])
        else:
            dev = tuple(DeviceBuffer.from_numpy(a) for a in host)
            handle = B.prepare_polygon_instances(*dev, [0, n])
        scenes.append({"index": i, "kind": kind, "dist": dist, "n": n, "handle": handle, "host": host,
                       "build_s": time.perf_counter() - t0})
        del dev  # the scene keeps its own copy
        log(f"scene {i:2d}: {kind:3s} {dist:7s} n={n:>9d} uploaded + built in {scenes[-1]['build_s']:.2f}s")
    return scenes, centres


# ------------------------------------------------------- request stream --


def make_stream(rng, seconds: float, rate: float, scenes, t_offset: float = 0.0):
    n = rng.poisson(rate * seconds)
    arrivals = np.sort(rng.uniform(0.0, seconds, n)) + t_offset
    kinds = rng.integers(0, 4, n)
    sizes = np.floor(np.exp(rng.uniform(0.0, math.log(65536.0), n))).astype(np.int64)
    sizes = np.clip(sizes, 1, 65536)
    by_kind = {k: [s for s in scenes if s["kind"] == KINDS[k]] for k in range(4)}
    scene_idx = np.array([by_kind[k][rng.integers(0, len(by_kind[k]))]["index"] for k in kinds])
    return {"arrival": arrivals, "kind": kinds, "size": sizes, "scene": scene_idx}


def make_pools(rng, centres, pool_n: int = POOL_QUERIES):
    """Seeded query pools, host copy + device buffer per request group:
    point (3 FP32) and PIP (2 FP32) pools hold a uniform half [0, pool_n) and
    a 64-Gaussian-mixture half [pool_n, 2 pool_n); the ray pool holds pool_n
    rays (origin U[0,1)^3, direction uniform on the sphere, tmax).  A
    request's queries are a window of the pool of its scene's distribution."""
    from rtdl_amd.device import DeviceBuffer

    host, dev = {}, {}
    for group, dim in (("point", 3), ("pip", 2)):
        u = rng.random((pool_n, dim), dtype=np.float32)
        g = (centres[rng.integers(0, 64, pool_n), :dim]
             + 0.02 * rng.standard_normal((pool_n, dim), dtype=np.float32)).astype(np.float32)
        host[group] = np.concatenate([u, g])
    o = rng.random((pool_n, 3))
    d = rng.standard_normal((pool_n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    host["ray"] = np.concatenate([o, d, np.full((pool_n, 1), TMAX)], 1).astype(np.float32)
    for group, q in host.items():
        dev[group] = DeviceBuffer.from_numpy(q)
    return host, dev


def assign_windows(rng, stream, scenes, pool_n: int = POOL_QUERIES):
    """Per-request query offsets into its group's pool (random windows)."""
    n = len(stream["kind"])
    sizes = stream["size"]
    start = (rng.random(n) * (pool_n - sizes + 1)).astype(np.int64)
    mixture = np.array([scenes[int(s)]["dist"] == "mixture" for s in stream["scene"]], dtype=bool)
    base = np.where((stream["kind"] != 2) & mixture, pool_n, 0)
    return base + start


JOB_OF_KIND = None


def requests_array(B, stream, offsets, scene_slot, capture_every: int = 0):
    kind_map = {0: B.JOB_FIXED_RADIUS, 1: B.JOB_KNN, 2: B.JOB_RAY_ANYHIT, 3: B.JOB_PIP_POSITIVE}
    n = len(stream["kind"])
    reqs = np.zeros(n, dtype=B.SERVE_REQUEST_DTYPE)
    reqs["kind"] = [kind_map[int(k)] for k in stream["kind"]]
    reqs["scene"] = [scene_slot[int(s)] for s in stream["scene"]]
    reqs["query_offset"] = offsets
    reqs["query_count"] = stream["size"]
    reqs["arrival_s"] = stream["arrival"]
    if capture_every:
        reqs["capture"][::capture_every] = 1
    return reqs


# ------------------------------------------------------------- roofline --


def roofline_leg(B, scenes, stream, offsets, arenas, limit_queries: int, kinds=range(4)):
    """One batch per predicate family over the first requests of the stream.
    Kernel seconds from events on the launch stream (timed run), node and
    primitive visits from a separate counting run (untimed)."""
    kind_map = {0: B.JOB_FIXED_RADIUS, 1: B.JOB_KNN, 2: B.JOB_RAY_ANYHIT, 3: B.JOB_PIP_POSITIVE}
    group_of = {0: "point", 1: "point", 2: "ray", 3: "pip"}
    stride = {0: 3, 1: 3, 2: 7, 3: 2}
    probe_bytes = {0: 12, 1: 12, 2: 28, 3: 8}
    prim_bytes = {0: 12, 1: 12, 2: 36, 3: None}
    row_bytes = {0: 16, 1: 24, 2: 1, 3: 8}
    out = {}
    for k in kinds:
        idx = np.flatnonzero(stream["kind"] == k)
        jobs, q_total = [], 0
        for i in idx:
            if q_total >= limit_queries:
                break
            n = int(stream["size"][i])
            o = int(offsets[i])
            s = scenes[int(stream["scene"][i])]
            jobs.append(B.BatchJob(s["handle"], kind_map[k], arenas[group_of[k]][o:o + n],
                                   radius=RADIUS, k=K_KNN if k == 1 else K_MAX))
            q_total += n
        if not jobs:
            continue
        try:
            B.run_batch_profiled(jobs)  # warm
            res, fam, _ = B.run_batch_profiled(jobs)
            _, _, visits = B.run_batch_profiled(jobs, count_visits=True)
        except RuntimeError as exc:
            log(f"kernel {KINDS[k]}: {exc}")
            out[KINDS[k]] = {"error": str(exc)}
            continue
        fname = {0: "neighbors", 1: "neighbors", 2: "ray_any_hit", 3: "pip"}[k]
        t = fam[fname]
        rows = sum(r["rows"] for r in res)
        nodes, prims, steps = visits
        used = {int(stream["scene"][i]) for i in idx[:len(jobs)]}
        if k == 3:
            # polygon refs (16 B) plus 8 B per vertex, at the mean vertex count
            # of the polygons of the scenes this batch queries
            counts = np.concatenate([scenes[i]["host"][1] for i in used])
            mean_v = float(counts.mean())
            pb = 16 + 8 * mean_v
        else:
            pb = prim_bytes[k]
        w = q_total * probe_bytes[k] + 32 * nodes + pb * prims + row_bytes[k] * rows
        # implementation-independent floor: every probe, every primitive of
        # the scenes used and every row once
        n_prims = sum(scenes[i]["n"] for i in used)
        w_min = q_total * probe_bytes[k] + pb * n_prims + row_bytes[k] * rows
        out[KINDS[k]] = {
            "kernel": {0: "nbr_count/emit/sort_kernel", 1: f"batch_points_kernel<{K_KNN}>",
                       2: "batch_rays_kernel<0>", 3: "batch_pip_kernel"}[k],
            "queries": q_total, "jobs": len(jobs), "kernel_s": t, "node_visits": nodes,
            "prim_tests": prims, "wave_steps": steps,
            "nodes_per_lane_step": nodes / (64.0 * steps) if steps else 0.0,
            "rows": rows, "algorithmic_bytes": w, "w_min_bytes": w_min,
            "prim_record_bytes": pb,
            "achieved_GBps": w / t / 1e9 if t > 0 else 0.0,
            "queries_per_s": q_total / t if t > 0 else 0.0,
        }
        log(f"kernel {KINDS[k]}: {q_total} q in {t*1e3:.2f} ms ({q_total/t/1e6:.1f} Mq/s), "
            f"{nodes/q_total:.1f} nodes/q, {prims/q_total:.1f} prims/q, "
            f"{nodes / (64.0 * max(steps, 1)):.2f} nodes per lane step, W/t = {w/t/1e9:.0f} GB/s")
    return out


def measured_traffic(kind: str, measured: dict, args):
    """HBM bytes per launch of the roofline batch, from the committed
    rocprofv3 --pmc measurements of this same batch (FETCH_SIZE + WRITE_SIZE,
    each scaled by the counter's reading for a copy of known size), when one
    matches it: the batch is drawn from its own seed, so it depends only on
    the kind, the scene sizes and the query limit; otherwise None."""
    path = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return None
    for rec in recs if isinstance(recs, list) else [recs]:
        if (rec.get("kind") == kind and rec.get("queries") == measured["queries"]
                and rec.get("roofline_seed") == ROOF_SEED and rec.get("max_prims") == args.max_prims
                and rec.get("roofline_queries") == args.roofline_queries):
            return rec.get("bytes_per_launch")
    return None


# ------------------------------------------------------------ baselines --


def host_scene_records(rtm, s):
    if "records" in s:
        return s["records"]
    if s["kind"] in ("frn", "knn"):
        d = s["host"].astype(np.float64)
        rec = np.zeros(len(d), dtype=rtm.POINT3D_DTYPE)
        rec["id"] = np.arange(len(d))
        rec["x"], rec["y"], rec["z"] = d[:, 0], d[:, 1], d[:, 2]
    elif s["kind"] == "ray":
        d = s["host"].astype(np.float64)
        rec = np.zeros(len(d), dtype=rtm.TRIANGLE3D_DTYPE)
        rec["id"] = np.arange(len(d))
        for j, f in enumerate(["x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2"]):
            rec[f] = d[:, j]
    else:
        voff, vcnt, verts = s["host"]
        refs = np.zeros(len(voff), dtype=rtm.POLYGON_REF_DTYPE)
        refs["id"] = np.arange(len(voff))
        refs["vertex_offset"] = voff
        refs["vertex_count"] = vcnt
        rec = (refs, verts.astype(np.float64).reshape(-1))
    s["records"] = rec
    return rec


def host_queries(rtm, host, kind, o, n):
    if kind in (0, 1):
        q = host["point"][o:o + n].astype(np.float64)
        rec = np.zeros(n, dtype=rtm.POINT3D_DTYPE)
        rec["id"] = np.arange(n)
        rec["x"], rec["y"], rec["z"] = q[:, 0], q[:, 1], q[:, 2]
    elif kind == 2:
        q = host["ray"][o:o + n].astype(np.float64)
        rec = np.zeros(n, dtype=rtm.RAY3D_DTYPE)
        rec["id"] = np.arange(n)
        for j, f in enumerate(["ox", "oy", "oz", "dx", "dy", "dz", "tmax"]):
            rec[f] = q[:, j]
    else:
        q = host["pip"][o:o + n].astype(np.float64)
        rec = np.zeros(n, dtype=rtm.POINT2D_DTYPE)
        rec["id"] = np.arange(n)
        rec["x"], rec["y"] = q[:, 0], q[:, 1]
    return rec


def plain_per_call_leg(rtm, scenes, stream, offsets, host, budget_s: float):
    """Plain per-call C-ABI: host records in, scene built per call, host rows out."""
    p = rtm._ptr
    done_q, spent, calls = 0, 0.0, 0
    for i in range(len(stream["kind"])):
        if spent >= budget_s:
            break
        k = int(stream["kind"][i])
        n = int(stream["size"][i])
        s = scenes[int(stream["scene"][i])]
        scene_rec = host_scene_records(rtm, s)
        q = host_queries(rtm, host, k, int(offsets[i]), n)
        t0 = time.perf_counter()
        if k == 0:
            rtm._call_rows("rtdl_mi355x_run_fixed_radius_neighbors_3d", rtm.FRN_ROW_DTYPE, p(q), n,
                           p(scene_rec), len(scene_rec), RADIUS, K_MAX)
        elif k == 1:
            rtm._call_rows("rtdl_mi355x_run_knn_rows_3d", rtm.KNN_ROW_DTYPE, p(q), n, p(scene_rec),
                           len(scene_rec), K_KNN)
        elif k == 2:
            rtm._call_rows("rtdl_mi355x_run_ray_anyhit_3d", rtm.ANYHIT_ROW_DTYPE, p(q), n, p(scene_rec),
                           len(scene_rec))
        else:
            refs, verts = scene_rec
            rtm._call_rows("rtdl_mi355x_run_point_in_polygon", rtm.PIP_ROW_DTYPE, p(q), n, p(refs),
                           len(refs), p(verts), len(verts), 1)
        spent += time.perf_counter() - t0
        done_q += n
        calls += 1
    return {"queries": done_q, "seconds": spent, "calls": calls,
            "queries_per_s": done_q / spent if spent else 0.0}


def per_call_configs_leg(rtm):
    """The plain per-call backend at the two single-GPU per-call shapes of the
    reference's benchmark list: 3-D ray any-hit, 2,000,000 triangles x
    1,000,000 rays (tmax 0.5), and KNN k = 16, 4,000,000 x 4,000,000 3-D
    Gaussian-mixture points; FP32 values, seeds 20261203 + 1 and + 2.  Each
    runs once untimed, once timed (host records in, LBVH built, walked, rows
    copied back to the host, with the library's phase timings and the walk's
    kernel time from HIP events) and once counting node and primitive visits,
    which give the walk's algorithmic bytes W (32 B per node record, 28 / 12 B
    per ray / point probe, 36 / 12 B per triangle / point tested, 1 / 24 B
    per any-hit / KNN row) and its HBM roofline fraction."""
    out = {}
    p = rtm._ptr
    rng = np.random.default_rng(20261203 + 1)
    c = rng.random((2_000_000, 3))
    v = (c[:, None, :] + rng.normal(0.0, 0.005, (2_000_000, 3, 3))).reshape(-1, 9)
    tris = np.zeros(2_000_000, dtype=rtm.TRIANGLE3D_DTYPE)
    tris["id"] = np.arange(2_000_000)
    for j, f in enumerate(["x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2"]):
        tris[f] = v[:, j].astype(np.float32)
    o = rng.random((1_000_000, 3))
    d = rng.normal(size=(1_000_000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.zeros(1_000_000, dtype=rtm.RAY3D_DTYPE)
    rays["id"] = np.arange(1_000_000)
    for j, f in enumerate(["ox", "oy", "oz"]):
        rays[f] = o[:, j].astype(np.float32)
    for j, f in enumerate(["dx", "dy", "dz"]):
        rays[f] = d[:, j].astype(np.float32)
    rays["tmax"] = 0.5
    fn = ("rtdl_mi355x_run_ray_anyhit_3d", rtm.ANYHIT_ROW_DTYPE)
    args = (p(rays), len(rays), p(tris), len(tris))
    rtm._call_rows(*fn, *args)  # warm
    rows, wall, ph, t, nodes, prims = per_call_walk(rtm, fn, args)
    q, r, n = len(rays), len(rays), len(tris)
    w = q * 28 + 32 * nodes + 36 * prims + 1 * r
    out["ray_anyhit_2M_triangles_1M_rays"] = {
        "wall_s": wall, "rays_per_s": q / wall, "hits": int(rows["any_hit"].sum()),
        "phases_ms": {k: x * 1e3 for k, x in ph.items()}, "device_rays_per_s": q / t,
        "node_visits": nodes, "prim_tests": prims,
        "roofline": {"kernel": "ray_kernel<0>", **_roof(w, t, q * 28 + 36 * n + r)}}
    log(f"per-call ray any-hit 2M x 1M: {wall*1e3:.1f} ms wall, walk {t*1e3:.2f} ms, "
        f"W/t {w/t/1e9:.0f} GB/s, phases {out['ray_anyhit_2M_triangles_1M_rays']['phases_ms']}")
    del rows, tris, rays, v, c, o, d

    rng = np.random.default_rng(20261203 + 2)
    centres = rng.random((64, 3))
    pts = []
    for _ in range(2):
        xyz = (centres[rng.integers(0, 64, 4_000_000)] + rng.normal(0.0, 0.02, (4_000_000, 3))).astype(np.float32)
        rec = np.zeros(4_000_000, dtype=rtm.POINT3D_DTYPE)
        rec["id"] = np.arange(4_000_000)
        rec["x"], rec["y"], rec["z"] = xyz[:, 0], xyz[:, 1], xyz[:, 2]
        pts.append(rec)
    search, query = pts
    fn = ("rtdl_mi355x_run_knn_rows_3d", rtm.KNN_ROW_DTYPE)
    args = (p(query), len(query), p(search), len(search), 16)
    rtm._call_rows(*fn, *args)  # warm
    rows, wall, ph, t, nodes, prims = per_call_walk(rtm, fn, args)
    q, r, n = len(query), int(len(rows)), len(search)
    w = q * 12 + 32 * nodes + 12 * prims + 24 * r
    out["knn16_4M_x_4M_mixture"] = {
        "wall_s": wall, "queries_per_s": q / wall, "rows": r,
        "phases_ms": {k: x * 1e3 for k, x in ph.items()}, "device_queries_per_s": q / t,
        "node_visits": nodes, "prim_tests": prims,
        "roofline": {"kernel": "collect_packed_kernel<16> + exact re-run", **_roof(w, t, q * 12 + 12 * n + 24 * r)}}
    log(f"per-call KNN k=16 4M x 4M: {wall*1e3:.1f} ms wall, walk {t*1e3:.2f} ms, W/t {w/t/1e9:.0f} GB/s, "
        f"phases {out['knn16_4M_x_4M_mixture']['phases_ms']}")
    return out


def per_request_resident_leg(B, scenes, stream, offsets, arenas, budget_s: float):
    """Resident scenes, one synchronous single-job batch per request."""
    kind_map = {0: B.JOB_FIXED_RADIUS, 1: B.JOB_KNN, 2: B.JOB_RAY_ANYHIT, 3: B.JOB_PIP_POSITIVE}
    group_of = {0: "point", 1: "point", 2: "ray", 3: "pip"}
    done_q, calls = 0, 0
    t0 = time.perf_counter()
    for i in range(len(stream["kind"])):
        if time.perf_counter() - t0 >= budget_s:
            break
        k = int(stream["kind"][i])
        n = int(stream["size"][i])
        o = int(offsets[i])
        s = scenes[int(stream["scene"][i])]
        job = B.BatchJob(s["handle"], kind_map[k], arenas[group_of[k]][o:o + n], radius=RADIUS,
                         k=K_KNN if k == 1 else K_MAX)
        B.run_batch([job], morton_sort=True)
        done_q += n
        calls += 1
    spent = time.perf_counter() - t0
    return {"queries": done_q, "seconds": spent, "calls": calls,
            "queries_per_s": done_q / spent if spent else 0.0}


def cpu_baseline_leg(rtm, scenes, stream, offsets, host, budget_s: float):
    """The C oracle (FP64, multi-threaded over all host cores) on a bounded
    sample of the same requests, grouped per scene (one index per scene)."""
    from oracle import native as oracle

    threads = oracle.default_threads()
    done_q, spent = 0, 0.0
    # take requests in stream order; each scene's pending requests run as one
    # call (one index build per scene) once they reach `flush_at` queries
    flush_at = 16384
    pending: dict[int, list[int]] = {}
    batches = []
    for i in range(len(stream["kind"])):
        sidx = int(stream["scene"][i])
        pending.setdefault(sidx, []).append(i)
        if sum(int(stream["size"][j]) for j in pending[sidx]) >= flush_at:
            batches.append((sidx, pending.pop(sidx)))
    batches.extend(pending.items())
    for sidx, reqs in batches:
        if spent >= budget_s:
            break
        if True:
            s = scenes[sidx]
            k = int(stream["kind"][reqs[0]])
            qs = [host_queries(rtm, host, k, int(offsets[i]), int(stream["size"][i])) for i in reqs]
            q = np.concatenate(qs)
            rec = host_scene_records(rtm, s)
            t0 = time.perf_counter()
            if k in (0, 1):
                qd = {"id": np.arange(len(q), dtype=np.uint32), "x": q["x"], "y": q["y"], "z": q["z"]}
                sd = {"id": rec["id"], "x": rec["x"], "y": rec["y"], "z": rec["z"]}
                if k == 0:
                    oracle.neighbors("frn", qd, sd, radius=RADIUS, k=K_MAX, threads=threads)
                else:
                    oracle.neighbors("knn", qd, sd, k=K_KNN, threads=threads)
            elif k == 2:
                rm = np.stack([q[f] for f in ["ox", "oy", "oz", "dx", "dy", "dz", "tmax"]], 1)
                tm = np.stack([rec[f] for f in ["x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2"]], 1)
                oracle.rays_3d("any", rm, tm, rec["id"], threads=threads)
            else:
                refs, verts = rec
                oracle.pip_pairs(q["x"], q["y"], refs["vertex_offset"], refs["vertex_count"], verts,
                                 positive_only=True, threads=threads)
            spent += time.perf_counter() - t0
            done_q += len(q)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": done_q / spent if spent else 0.0, "unit": "queries/s", "cores": threads,
            "kind": "port", "cpu_model": model,
            "sample": f"{done_q} queries of the timed stream, grouped per scene (grid index built per "
                      f"group, included), C FP64 oracle, {spent:.1f} s"}


def python_subsample_leg():
    """The pure-Python FP64 restatement of the reference, single-threaded, on
    fixed-seed subsamples of the other benchmark shapes.  Its per-query
    work is a scan over every primitive, so each full-size figure is a
    linear extrapolation in the primitive count (labelled as such): exact
    for KNN up to its sort's log factor, an upper bound for any-hit (a
    ray stops at its first hit in input order)."""
    from oracle import cpu_python_reference as ref
    from rtdl_amd.records import Point, Point3D, Polygon, Ray3D, Triangle3D

    out = {}

    def timed(fn):
        t0 = time.perf_counter()
        res = fn()
        return res, time.perf_counter() - t0

    # ray any-hit, 2M triangles x 1M rays: 16 rays x 20,000 triangles
    rng = np.random.default_rng(20261203 + 1)
    c = rng.random((20000, 3))
    v = (c[:, None, :] + rng.normal(0.0, 0.005, (20000, 3, 3))).reshape(-1, 9).astype(np.float32).astype(float)
    tris = [Triangle3D(i, *map(float, row)) for i, row in enumerate(v)]
    o = rng.random((16, 3))
    d = rng.normal(size=(16, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = [Ray3D(i, *map(float, np.float32(o[i])), *map(float, np.float32(d[i])), 0.5) for i in range(16)]
    rows, dt = timed(lambda: ref.ray_triangle_any_hit_cpu(rays, tris))
    per_ray = dt / len(rays)
    out["ray_anyhit_2M_triangles_1M_rays"] = {
        "sample": f"16 rays x 20,000 triangles: {sum(r['any_hit'] for r in rows)} hits in {dt:.2f} s",
        "rays_per_s_at_sample": 1.0 / per_ray,
        "extrapolated_full_s_upper_bound": per_ray * (2_000_000 / 20000) * 1_000_000}
    # KNN k = 16, 4M x 4M mixture points: 8 queries x 50,000 points
    rng = np.random.default_rng(20261203 + 2)
    centres = rng.random((64, 3))
    sp = (centres[rng.integers(0, 64, 50000)] + rng.normal(0.0, 0.02, (50000, 3))).astype(np.float32).astype(float)
    qp = (centres[rng.integers(0, 64, 8)] + rng.normal(0.0, 0.02, (8, 3))).astype(np.float32).astype(float)
    search = [Point3D(i, *row) for i, row in enumerate(sp.tolist())]
    queries = [Point3D(i, *row) for i, row in enumerate(qp.tolist())]
    _, dt = timed(lambda: ref.knn_rows_cpu(queries, search, k=16))
    per_q = dt / len(queries)
    out["knn16_4M_x_4M_mixture"] = {
        "sample": f"8 queries x 50,000 points in {dt:.2f} s",
        "queries_per_s_at_sample": 1.0 / per_q,
        "extrapolated_full_s": per_q * (4_000_000 / 50000) * 4_000_000}
    # PIP positive hits against one instance's 1,000 star polygons: 400 points
    rng = np.random.default_rng(20261203 + 3)
    polys = []
    for i in range(1000):
        nv = int(rng.integers(8, 33))
        ang = np.sort(rng.random(nv) * 2.0 * np.pi)
        rad = rng.uniform(0.005, 0.02, nv)
        cx, cy = rng.random(2)
        xy = np.stack([cx + rad * np.cos(ang), cy + rad * np.sin(ang)], 1).astype(np.float32).astype(float)
        polys.append(Polygon(i, tuple(map(tuple, xy.tolist()))))
    pts = [Point(i, float(x), float(y)) for i, (x, y) in enumerate(rng.random((400, 2)).astype(np.float32).astype(float))]
    rows, dt = timed(lambda: ref.pip_cpu(pts, polys, result_mode="positive_hits"))
    out["pip_positive_1000_polygons"] = {
        "sample": f"400 points x 1,000 star polygons: {len(rows)} hits in {dt:.2f} s",
        "points_per_s": len(pts) / dt}
    # fixed radius r = 0.01 on a 1e5-point uniform scene (the smallest served scene): 16 queries
    rng = np.random.default_rng(20261203 + 4)
    sp = rng.random((100000, 3)).astype(np.float32).astype(float)
    search = [Point3D(i, *row) for i, row in enumerate(sp.tolist())]
    queries = [Point3D(i, *row) for i, row in enumerate(rng.random((16, 3)).astype(np.float32).astype(float).tolist())]
    rows, dt = timed(lambda: ref.fixed_radius_neighbors_cpu(queries, search, radius=0.01, k_max=64))
    out["fixed_radius_r0.01_1e5_points"] = {
        "sample": f"16 queries x 100,000 points: {len(rows)} rows in {dt:.2f} s",
        "queries_per_s": len(queries) / dt,
        "extrapolated_queries_per_s_at_1e7_points": len(queries) / dt / 100.0}
    return out


# ------------------------------------------------------ served stream --


def replay_stream(B, handles, scenes, pools_dev, rng_stream, rng_win, seconds: float, rate: float,
                  serve_kw: dict, capture_every: int = 0, rng_check=None):
    """Replays `seconds` of the request stream in one continuous serving loop
    (query windows drawn from the resident pools) and returns the totals, the
    latencies, the sampled captured rows and the stream itself."""
    st = make_stream(rng_stream, seconds, rate, scenes)
    offs = assign_windows(rng_win, st, scenes)
    reqs = requests_array(B, st, offs, {s["index"]: s["index"] for s in scenes}, capture_every)
    log(f"replay: {len(reqs)} requests, {int(st['size'].sum())} queries over {seconds:.1f} s")
    results, stats, captured = B.serve_replay(handles, reqs, point_queries=pools_dev["point"],
                                              ray_queries=pools_dev["ray"], pip_queries=pools_dev["pip"],
                                              **serve_kw)
    ok = results["status"] == 0
    out = {"served_q": int(st["size"][ok].sum()), "wall_s": stats["wall_s"], "batches": stats["batches"],
           "rows": stats["rows"], "rejected": int((~ok).sum()), "requests": len(reqs),
           "queries": int(st["size"].sum()), "peak_inflight_bytes": stats["peak_inflight_bytes"],
           "region_bytes": stats["region_bytes"],
           "latency_ms": ((results["done_s"] - results["arrival_s"]) * 1e3)[ok], "latency_kind": st["kind"][ok],
           "arrival_s": st["arrival"][ok]}
    checks = []
    if rng_check is not None:
        for i, rows in captured.items():
            checks.append(sample_capture(rng_check, st, offs, i, rows))
    log(f"replay done: wall {stats['wall_s']:.3f}s, {stats['batches']} batches, rejected {out['rejected']}")
    return out, checks, (st, offs, reqs)


_QID_FIELD = {0: "query_id", 1: "query_id", 2: "ray_id", 3: "point_id"}


def _rows_by_query(rows, field, ids, n_queries: int | None = None):
    """Rows grouped by query id (stable: each query's rows keep the order they
    were written in) and the [start, end) row span of each id in `ids`.  With
    `n_queries`, only the rows of `ids` are kept (a lookup table over the
    query ids, so large row sets are filtered before they are sorted)."""
    if n_queries is not None:
        lut = np.zeros(n_queries, dtype=bool)
        lut[ids] = True
        rows = rows[lut[rows[field]]]
    rows = rows[np.argsort(rows[field], kind="stable")]
    key = rows[field]
    return rows, np.searchsorted(key, ids, "left"), np.searchsorted(key, ids, "right")


def sample_capture(rng, stream, offsets, i, rows):
    """Keeps the rows of up to CHECK_PER_REQUEST sampled queries of captured
    request i (rows carry request-local query ids)."""
    n = int(stream["size"][i])
    k = int(stream["kind"][i])
    sample = np.sort(rng.choice(n, min(n, CHECK_PER_REQUEST), replace=False))
    keep = np.isin(rows[_QID_FIELD[k]], sample)
    return {"kind": k, "scene": int(stream["scene"][i]), "offset": int(offsets[i]), "n": n,
            "sample": sample, "rows": rows[keep]}


def verify_captures(checks, scenes, pools_host):
    """Checks the sampled queries of the captured served requests against the
    C FP64 oracle over the whole scene, bit for bit (one oracle call per
    scene).  Returns counts of queries, rows and mismatches."""
    from oracle import native as oracle

    rep = {"requests_checked": 0, "queries_checked": 0, "served_rows_checked": 0, "mismatched_queries": 0,
           "mismatched_requests": 0}
    by_scene: dict[int, list] = {}
    for c in checks:
        by_scene.setdefault(c["scene"], []).append(c)
    group = {0: "point", 1: "point", 2: "ray", 3: "pip"}
    for sidx, cs in by_scene.items():
        s = scenes[sidx]
        k = cs[0]["kind"]
        q = np.concatenate([pools_host[group[k]][c["offset"] + c["sample"]] for c in cs]).astype(np.float64)
        base = np.cumsum([0] + [len(c["sample"]) for c in cs])
        if k in (0, 1):
            h = s["host"].astype(np.float64)
            sd = {"id": np.arange(len(h), dtype=np.uint32), "x": h[:, 0], "y": h[:, 1], "z": h[:, 2]}
            qd = {"id": np.arange(len(q), dtype=np.uint32), "x": q[:, 0], "y": q[:, 1], "z": q[:, 2]}
            want = oracle.neighbors("frn" if k == 0 else "knn", qd, sd, radius=RADIUS, k=K_MAX if k == 0 else K_KNN)
            w_start = np.searchsorted(want["qid"], np.arange(len(q) + 1))
        elif k == 2:
            want = oracle.rays_3d("any", q, s["host"].astype(np.float64), np.arange(s["n"], dtype=np.uint32))
        else:
            voff, vcnt, verts = s["host"]
            want = oracle.pip_pairs(q[:, 0], q[:, 1], voff, vcnt, verts.astype(np.float64).reshape(-1),
                                    positive_only=True)
            w_start = np.searchsorted(want[:, 0], np.arange(len(q) + 1))
        for ci, c in enumerate(cs):
            bad = 0
            rows, lo, hi = _rows_by_query(c["rows"], _QID_FIELD[k], c["sample"])
            for j, local in enumerate(c["sample"]):
                g = base[ci] + j
                mine = rows[lo[j]:hi[j]]  # the query's rows, in the order written
                if k in (0, 1):
                    w = want[w_start[g]:w_start[g + 1]]
                    same = (len(mine) == len(w) and np.array_equal(mine["neighbor_id"], w["nid"])
                            and np.array_equal(mine["distance"].view(np.uint64), w["dist"].view(np.uint64)))
                    if same and k == 1:
                        same = np.array_equal(mine["neighbor_rank"], w["rank"])
                elif k == 2:
                    same = len(mine) == 1 and int(mine["any_hit"][0]) == int(want[g])
                else:
                    w = want[w_start[g]:w_start[g + 1], 1]
                    same = len(mine) == len(w) and np.array_equal(np.sort(mine["polygon_id"]), np.sort(w))
                bad += not same
                rep["served_rows_checked"] += len(mine)
            rep["queries_checked"] += len(c["sample"])
            rep["mismatched_queries"] += bad
            rep["mismatched_requests"] += bad > 0
            rep["requests_checked"] += 1
    return rep


# --------------------------------------------------- other configurations --


def config0_leg(rtm):
    """The reference's CPU-runnable case on the GPU: fixed-radius rows,
    4,096 x 4,096 2-D FP64 points uniform in [0, 1)^2 (seed 20261203),
    r = 0.02, k_max = 4,096 (never truncates), plain per-call backend (host
    records in, LBVH built, rows copied back), best of five calls; its rows
    against the pure-Python reference's rows on the same points, whose time
    is the reference's own CPU baseline."""
    from oracle import cpu_python_reference as ref
    from rtdl_amd.records import Point

    rng = np.random.default_rng(20261203)
    q = rng.random((4096, 2))
    sp = rng.random((4096, 2))
    qd = {"id": np.arange(4096, dtype=np.uint32), "x": q[:, 0], "y": q[:, 1]}
    sd = {"id": np.arange(4096, dtype=np.uint32), "x": sp[:, 0], "y": sp[:, 1]}
    walls = []
    for _ in range(5):
        t0 = time.perf_counter()
        got = rtm.fixed_radius_neighbors_mi355x(qd, sd, radius=0.02, k_max=4096)
        walls.append(time.perf_counter() - t0)
    ph = rtm.last_phase_timings()
    t0 = time.perf_counter()
    want = ref.fixed_radius_neighbors_cpu([Point(i, float(x), float(y)) for i, (x, y) in enumerate(q)],
                                          [Point(i, float(x), float(y)) for i, (x, y) in enumerate(sp)],
                                          radius=0.02, k_max=4096)
    py_s = time.perf_counter() - t0
    match = (len(got) == len(want)
             and got["query_id"].tolist() == [r["query_id"] for r in want]
             and got["neighbor_id"].tolist() == [r["neighbor_id"] for r in want]
             and got["distance"].tolist() == [r["distance"] for r in want])
    wall = min(walls)
    return {"workload": "fixed-radius rows, 4,096 x 4,096 2-D FP64 points, r = 0.02, k_max = 4,096, per call",
            "wall_s": wall, "queries_per_s": 4096 / wall, "rows": int(len(got)),
            "phases_ms": {k: v * 1e3 for k, v in ph.items()},
            "rows_match_python_reference": bool(match),
            "cpu_python_reference": {"value": 4096 / py_s, "unit": "queries/s", "cores": 1, "kind": "port",
                                     "sample": f"the same 4,096 x 4,096 case, {len(want)} rows in {py_s:.1f} s"},
            "speedup_vs_python_reference": py_s / wall,
            "roofline": "not graded: about 10 MB of algorithmic bytes, bound by launch and copy latency"}


def per_call_walk(rtm, fn, args):
    """One timed call and one counting call of a per-call entry point:
    (rows, wall seconds, phases, walk kernel seconds, node visits, prim tests)."""
    t0 = time.perf_counter()
    rows = rtm._call_rows(*fn, *args)
    wall = time.perf_counter() - t0
    ph = rtm.last_phase_timings()
    walk = rtm.last_walk_stats()
    del rows
    rtm.set_walk_counting(True)
    try:
        rows = rtm._call_rows(*fn, *args)
        counted = rtm.last_walk_stats()
    finally:
        rtm.set_walk_counting(False)
    return rows, wall, ph, walk["kernel_s"], counted["node_visits"], counted["prim_tests"]


def _roof(w, t, w_min=None):
    out = {"algorithmic_bytes": int(w), "kernel_s": t, "achieved_GBps": w / t / 1e9 if t else 0.0,
           "peak_GBps": HBM_PEAK / 1e9, "frac": w / t / HBM_PEAK if t else 0.0}
    if w_min is not None:
        out["w_min_bytes"] = int(w_min)
    return out


# ------------------------------------------------------------ config [3] --


def config3_workload(rng, n_inst: int = 2048, lo: float = 1e3, hi: float = 2.6e5, n_polys: int = 1000):
    """2,048 independent instances, point counts round(exp(U[ln 1e3, ln 2.6e5]))
    with as many query points (independent draws), 2-D or 3-D 50/50, FP32
    uniform in [0, 1)^dim, radius log-uniform in [0.005, 0.05]; every 4th
    instance is point-in-polygon against 1,000 star polygons (8-32 vertices
    at sorted angles U[0, 2 pi), radii U[0.005, 0.02], centres U[0, 1)^2),
    positive hits.  Returns host arrays: build points (z = 0 in 2-D) and
    query points per fixed-radius instance, polygons and probe points per
    PIP instance, each concatenated with per-instance offsets."""
    sizes = np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), n_inst))).astype(np.int64)
    dims = rng.integers(2, 4, n_inst)
    radii = np.exp(rng.uniform(math.log(0.005), math.log(0.05), n_inst))
    pip = np.arange(n_inst) % 4 == 3
    frn_ids = np.flatnonzero(~pip)
    pip_ids = np.flatnonzero(pip)
    n_f = sizes[frn_ids]
    tot = int(n_f.sum())
    build = rng.random((tot, 3), dtype=np.float32)
    query = rng.random((tot, 3), dtype=np.float32)
    two_d = np.repeat(dims[frn_ids] == 2, n_f)
    build[two_d, 2] = 0.0
    query[two_d, 2] = 0.0
    # polygons: n_polys per PIP instance
    ng = n_polys * len(pip_ids)
    counts = rng.integers(8, 33, ng).astype(np.int64)
    seg = np.repeat(np.arange(ng), counts)
    total_v = int(counts.sum())
    ang = rng.random(total_v) * (2.0 * math.pi)
    ang = ang[np.lexsort((ang, seg))]  # sorted within each polygon (seg is already sorted)
    rad = rng.uniform(0.005, 0.02, total_v)
    cen = rng.random((ng, 2))
    verts = np.stack([cen[seg, 0] + rad * np.cos(ang), cen[seg, 1] + rad * np.sin(ang)], 1).astype(np.float32)
    voff_local = np.cumsum(counts) - counts  # into verts (one vertex array for all instances)
    n_p = sizes[pip_ids]
    probes = rng.random((int(n_p.sum()), 2), dtype=np.float32)
    return {
        "sizes": sizes, "dims": dims, "radii": radii, "frn_ids": frn_ids, "pip_ids": pip_ids,
        "build": build, "query": query, "point_offsets": np.concatenate([[0], np.cumsum(n_f)]),
        "voff": voff_local.astype(np.int32), "vcnt": counts.astype(np.int32), "verts": verts,
        "poly_offsets": np.arange(len(pip_ids) + 1, dtype=np.int64) * n_polys,
        "probes": probes, "probe_offsets": np.concatenate([[0], np.cumsum(n_p)]),
    }


def config3_row_budget(n: int, dim: int, r: float, k: int = 64) -> int:
    """Rows a fixed-radius job of n uniform queries against n uniform points
    in [0, 1)^dim is given: 1.5 x the expected neighbours in the ball (capped
    at k) plus 4, per query; a job that ran out would report
    RTDL_STATUS_ROW_OVERFLOW."""
    ball = math.pi * r * r if dim == 2 else 4.0 / 3.0 * math.pi * r ** 3
    return n * int(min(k, math.ceil(1.5 * n * ball) + 4))


class Config3:
    """The config [3] batch resident on the device: one multi-instance point
    scene (the fixed-radius instances), one multi-instance polygon scene (the
    PIP instances), query arenas and one row arena, and the BatchJobs."""

    def __init__(self, B, ws):
        from rtdl_amd.device import DeviceBuffer

        self.B, self.ws = B, ws
        self.dev_build = DeviceBuffer.from_numpy(ws["build"])
        self.dev_voff = DeviceBuffer.from_numpy(ws["voff"])
        self.dev_vcnt = DeviceBuffer.from_numpy(ws["vcnt"])
        self.dev_verts = DeviceBuffer.from_numpy(ws["verts"])
        q, po = ws["query"], ws["point_offsets"]
        dims = ws["dims"][ws["frn_ids"]]
        # 2-D instances take (x, y) queries (stride 2), 3-D ones (x, y, z)
        self.q2 = DeviceBuffer.from_numpy(np.concatenate(
            [q[po[j]:po[j + 1], :2] for j in range(len(dims)) if dims[j] == 2] or [np.zeros((1, 2), np.float32)]))
        self.q3 = DeviceBuffer.from_numpy(np.concatenate(
            [q[po[j]:po[j + 1]] for j in range(len(dims)) if dims[j] == 3] or [np.zeros((1, 3), np.float32)]))
        self.probes = DeviceBuffer.from_numpy(ws["probes"])
        caps = []
        for j, i in enumerate(ws["frn_ids"]):
            n = int(po[j + 1] - po[j])
            caps.append(config3_row_budget(n, int(ws["dims"][i]), float(ws["radii"][i])) * 16)
        pp = ws["probe_offsets"]
        caps += [int(pp[j + 1] - pp[j]) * B.PIP_HIT_CAP * 12 for j in range(len(ws["pip_ids"]))]
        self.row_bytes = [(c + 255) // 256 * 256 for c in caps]
        self.rows = DeviceBuffer((max(1, sum(self.row_bytes)),), np.uint8)
        self.build_scenes()

    def build_scenes(self):
        """(Re)builds both multi-instance scenes under the current tuning."""
        B, ws = self.B, self.ws
        self.pscene = B.prepare_point_instances(self.dev_build, ws["point_offsets"])
        self.gscene = B.prepare_polygon_instances(self.dev_voff, self.dev_vcnt, self.dev_verts, ws["poly_offsets"])
        return self

    def jobs(self, scenes=None):
        B, ws = self.B, self.ws
        pscene, gscene = (scenes.pscene, scenes.gscene) if scenes is not None else (self.pscene, self.gscene)
        po, pp = ws["point_offsets"], ws["probe_offsets"]
        out, at, o2, o3 = [], 0, 0, 0
        for j, i in enumerate(ws["frn_ids"]):
            n = int(po[j + 1] - po[j])
            d = int(ws["dims"][i])
            if d == 2:
                q, o2 = self.q2[o2:o2 + n], o2 + n
            else:
                q, o3 = self.q3[o3:o3 + n], o3 + n
            nb = self.row_bytes[len(out)]
            out.append(B.BatchJob(pscene, B.JOB_FIXED_RADIUS, q, instance=j, radius=float(ws["radii"][i]), k=64,
                                  rows=self.rows[at:at + nb], row_capacity=nb // 16))
            at += nb
        for j in range(len(ws["pip_ids"])):
            n = int(pp[j + 1] - pp[j])
            nb = self.row_bytes[len(out)]
            out.append(B.BatchJob(gscene, B.JOB_PIP_POSITIVE, self.probes[int(pp[j]):int(pp[j]) + n], instance=j,
                                  rows=self.rows[at:at + nb], row_capacity=nb // 12))
            at += nb
        return out


def config3_check(B, ws, jobs, results, rng, jobs_checked: int, per_job: int):
    """Bit-exact check of `per_job` sampled queries of `jobs_checked` sampled
    jobs against the C FP64 oracle over each whole instance."""
    from oracle import native as oracle

    po, pp = ws["point_offsets"], ws["probe_offsets"]
    n_frn = len(ws["frn_ids"])
    rep = {"jobs_checked": 0, "queries_checked": 0, "rows_checked": 0, "mismatched_queries": 0}
    pick = np.sort(rng.choice(len(jobs), min(jobs_checked, len(jobs)), replace=False))
    for ji in pick:
        job, res = jobs[ji], results[ji]
        if res["status"] != 0:
            rep["mismatched_queries"] += 1
            continue
        rows = B.job_rows(job, res)
        n = int(job.queries.shape[0])
        sample = np.sort(rng.choice(n, min(n, per_job), replace=False))
        if ji < n_frn:
            i = int(ws["frn_ids"][ji])
            d = int(ws["dims"][i])
            b = ws["build"][po[ji]:po[ji + 1]].astype(np.float64)
            q = ws["query"][po[ji]:po[ji + 1]][sample].astype(np.float64)
            keys = "xyz"[:d]
            sd = {"id": np.arange(len(b), dtype=np.uint32), **{c: b[:, a] for a, c in enumerate(keys)}}
            qd = {"id": np.arange(len(q), dtype=np.uint32), **{c: q[:, a] for a, c in enumerate(keys)}}
            want = oracle.neighbors("frn", qd, sd, radius=float(ws["radii"][i]), k=64)
            ws_ = np.searchsorted(want["qid"], np.arange(len(q) + 1))
            rows, lo, hi = _rows_by_query(rows, "query_id", sample, n)
            for a, local in enumerate(sample):
                mine = rows[lo[a]:hi[a]]
                w = want[ws_[a]:ws_[a + 1]]
                ok = (len(mine) == len(w) and np.array_equal(mine["neighbor_id"], w["nid"])
                      and np.array_equal(mine["distance"].view(np.uint64), w["dist"].view(np.uint64)))
                rep["mismatched_queries"] += not ok
                rep["rows_checked"] += len(mine)
        else:
            j = ji - n_frn
            g0, g1 = int(ws["poly_offsets"][j]), int(ws["poly_offsets"][j + 1])
            vo, vc = ws["voff"][g0:g1].astype(np.int64), ws["vcnt"][g0:g1].astype(np.int64)
            first = int(vo[0])
            last = int(vo[-1] + vc[-1])
            verts = ws["verts"][first:last].astype(np.float64).reshape(-1)
            p = ws["probes"][pp[j]:pp[j + 1]][sample].astype(np.float64)
            want = oracle.pip_pairs(p[:, 0], p[:, 1], vo - first, vc, verts, positive_only=True)
            ws_ = np.searchsorted(want[:, 0], np.arange(len(p) + 1))
            rows, lo, hi = _rows_by_query(rows, "point_id", sample, n)
            for a, local in enumerate(sample):
                mine = np.sort(rows["polygon_id"][lo[a]:hi[a]])
                w = np.sort(want[ws_[a]:ws_[a + 1], 1])
                rep["mismatched_queries"] += not (len(mine) == len(w) and np.array_equal(mine, w))
                rep["rows_checked"] += len(mine)
        rep["jobs_checked"] += 1
        rep["queries_checked"] += len(sample)
    return rep


def config3_leg(B, autotune: bool, seed: int = 20261203 + 3, jobs_checked: int = 256, per_job: int = 16):
    """Config [3]: the 2,048-instance fixed-radius (k_max = 64) + PIP batch
    in one persistent launch per family, eager and replayed from a hipGraph,
    optionally autotuned first (launch size and leaf sizes, rows compared
    between candidates on sampled jobs), with W from counting runs and
    sampled queries checked against the C oracle."""
    from dataclasses import asdict

    from rtdl_amd import autotune as AT

    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    ws = config3_workload(rng)
    c3 = Config3(B, ws)
    setup_s = time.perf_counter() - t0
    n_frn = len(ws["frn_ids"])
    q_frn = int(ws["point_offsets"][-1])
    q_pip = int(ws["probe_offsets"][-1])
    log(f"config [3]: {len(ws['sizes'])} instances, {q_frn} fixed-radius and {q_pip} PIP queries, "
        f"set up in {setup_s:.1f}s")
    original = AT.get_tuning()
    tuned = None
    try:
        if autotune:
            knobs = {"blocks_per_cu": (4, 8, 16), "point_leaf": (2, 4), "polygon_leaf": (1, 2)}
            check = sorted(rng.choice(len(ws["sizes"]), 12, replace=False).tolist())
            t0 = time.perf_counter()
            res = AT.autotune_batch(c3.build_scenes, c3.jobs, knobs=knobs, repeats=1, check_jobs=check)
            tuned = {"best": asdict(res["best"]), "seconds": time.perf_counter() - t0,
                     "table": res["table"], "rows_compared_on_jobs": len(check)}
            log(f"config [3] autotune: {tuned['best']} in {tuned['seconds']:.1f}s")
        jobs = c3.jobs()
        B.run_batch_profiled(jobs)  # warm
        res, fam, _ = B.run_batch_profiled(jobs)
        _, eager_s = B.run_batch(jobs)
        if any(r["status"] for r in res):
            raise RuntimeError(f"config [3]: job status {sorted({r['status'] for r in res})}")
        rows_frn = sum(r["rows"] for r in res[:n_frn])
        rows_pip = sum(r["rows"] for r in res[n_frn:])
        with B.BatchGraph(jobs) as graph:
            replays = [graph.launch() for _ in range(4)]
        graph_s = float(np.median([s for _, s in replays[1:]]))
        graph_same = all([r["rows"] for r in g] == [r["rows"] for r in res] for g, _ in replays)
        # the same batch with the point instances rebuilt from their device
        # coordinates in front of it at every launch (build + traverse + emit)
        with B.BatchGraph(jobs, rebuild=[(c3.pscene, c3.dev_build)]) as graph:
            rebuilt = [graph.launch() for _ in range(4)]
        graph_rebuild_s = float(np.median([s for _, s in rebuilt[1:]]))
        graph_same = graph_same and all([r["rows"] for r in g] == [r["rows"] for r in res] for g, _ in rebuilt)
        # W from counting runs, per family and (fixed radius) per dimension
        frn2 = [j for j, i in zip(jobs[:n_frn], ws["frn_ids"]) if ws["dims"][i] == 2]
        frn3 = [j for j, i in zip(jobs[:n_frn], ws["frn_ids"]) if ws["dims"][i] == 3]
        w_frn, w_pip = 0.0, 0.0
        for sub, sp in ((frn2, 8), (frn3, 12)):
            if sub:
                _, _, (v, t, _) = B.run_batch_profiled(sub, count_visits=True)
                q = sum(int(j.queries.shape[0]) for j in sub)
                w_frn += q * sp + 32 * v + sp * t
        w_frn += 16 * rows_frn
        _, _, (v_pip, t_pip, _) = B.run_batch_profiled(jobs[n_frn:], count_visits=True)
        mean_v = float(ws["vcnt"].mean())
        w_pip = q_pip * 8 + 32 * v_pip + (16 + 8 * mean_v) * t_pip + 8 * rows_pip
        checked = config3_check(B, ws, jobs, res, rng, jobs_checked, per_job)
    finally:
        AT.set_tuning(original)
    total_q = q_frn + q_pip
    t_frn, t_pip_s = fam["neighbors"], fam["pip"]
    log(f"config [3]: eager {eager_s*1e3:.1f} ms, graph replay {graph_s*1e3:.1f} ms "
        f"({graph_rebuild_s*1e3:.1f} ms with the point instances rebuilt in the graph), "
        f"fixed radius {t_frn*1e3:.1f} ms ({w_frn/t_frn/1e9:.0f} GB/s), PIP {t_pip_s*1e3:.1f} ms "
        f"({w_pip/t_pip_s/1e9:.0f} GB/s), check {checked}")
    return {
        "workload": "2,048 instances, 1e3-2.6e5 points (log-uniform), 2-D/3-D 50/50, r log-uniform "
                    "[0.005, 0.05], k_max 64; every 4th instance PIP vs 1,000 star polygons, positive hits",
        "instances": int(len(ws["sizes"])), "fixed_radius_queries": q_frn, "pip_queries": q_pip,
        "rows": int(rows_frn + rows_pip), "eager_s": eager_s, "graph_replay_s": graph_s,
        "graph_rebuild_replay_s": graph_rebuild_s,
        "graph_rows_match_eager": bool(graph_same),
        "queries_per_s": total_q / graph_s, "eager_queries_per_s": total_q / eager_s,
        "families": {"fixed_radius": _roof(w_frn, t_frn), "pip": _roof(w_pip, t_pip_s)},
        "autotune": tuned, "checked": checked,
    }


def embree_status() -> str:
    import ctypes.util

    return "available" if ctypes.util.find_library("embree4") else "unavailable (no libembree4 on this host)"


# ----------------------------------------------------------------- main --


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="steps the 60 s request stream is timed as (one step = 60 / steps seconds of it)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps of stream replayed before it")
    ap.add_argument("--rate", type=float, default=20000.0)
    ap.add_argument("--max-prims", type=float, default=1e7)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-batch-queries", type=int, default=4 << 20)
    ap.add_argument("--saturated-steps", type=int, default=4,
                    help="timed steps whose requests are replayed again, all arrived at once (capacity)")
    ap.add_argument("--percall-seconds", type=float, default=8.0)
    ap.add_argument("--resident-seconds", type=float, default=4.0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--roofline-queries", type=int, default=2 << 20)
    ap.add_argument("--skip-baselines", action="store_true")
    ap.add_argument("--skip-configs", action="store_true", help="omit the other configurations' GPU lines")
    ap.add_argument("--no-config3-autotune", action="store_true")
    ap.add_argument("--roofline-only", choices=KINDS + ("all",), default=None,
                    help="only build the scenes and measure this kind's roofline batch (the batch the "
                         "bench line's roofline is quoted on), or every kind's in turn ('all'), for "
                         "profiling those batches alone")
    ap.add_argument("--per-call-only", action="store_true",
                    help="run only the per-call 2M x 1M ray any-hit and 4M x 4M KNN k=16 calls")
    ap.add_argument("--config3-only", action="store_true", help="run only the 2,048-instance batch")
    ap.add_argument("--fp64-calibration-mib", type=int, default=0,
                    help="run only fp64_selftest_kernel over two inputs of this many MiB (counter calibration)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (the reference kernel "
                         "for scaling FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every query set is independent, multi-GPU sharding is a "
              "design note only; run with --gpus 1", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        return 2

    from rtdl_amd import batch as B
    from rtdl_amd import mi355x_runtime as rtm

    if args.copy_calibration_mib:
        from rtdl_amd.device import DeviceBuffer

        n = args.copy_calibration_mib << 20
        src, dst = DeviceBuffer((n,), np.uint8), DeviceBuffer((n,), np.uint8)
        dst.copy_from(src)  # first touch
        dst.copy_from(src)  # the dispatch the counters are read from
        print(json.dumps({"copy_bytes": n}), flush=True)
        return 0
    if args.fp64_calibration_mib:
        # a second counter calibration with a kernel of this library:
        # fp64_selftest_kernel streams two FP64 inputs and writes two outputs
        # of the same size (2 x n x 8 bytes each way)
        n = (args.fp64_calibration_mib << 20) // 8
        a = np.random.default_rng(1).random(n) + 1.0
        for _ in range(2):  # the second dispatch is the one read
            rtm.selftest_fp64(a, a)
        print(json.dumps({"fp64_selftest_read_bytes": 16 * n, "fp64_selftest_write_bytes": 16 * n}), flush=True)
        return 0

    info = rtm.mi355x_context_probe()
    log(f"device {info}")
    if args.per_call_only:
        print(json.dumps({"per_call_configs": per_call_configs_leg(rtm)}), flush=True)
        return 0
    if args.config3_only:
        print(json.dumps({"config3": config3_leg(B, not args.no_config3_autotune)}), flush=True)
        return 0
    rng = np.random.default_rng(SEED)
    scenes, centres = make_scenes(B, rng, args.max_prims)
    handles = [s["handle"] for s in scenes]
    pools_host, pools_dev = make_pools(np.random.default_rng(SEED + 1), centres)
    serve_kw = dict(radius=RADIUS, k_max=K_MAX, k=K_KNN, slots=args.slots,
                    memory_budget_bytes=int(args.budget_gb * (1 << 30)),
                    max_batch_queries=args.max_batch_queries)

    # the roofline batch: its own seeded stream, the same for any --steps
    rng_roof = np.random.default_rng(ROOF_SEED)
    roof_stream = make_stream(rng_roof, 0.25, args.rate, scenes)
    roof_offsets = assign_windows(rng_roof, roof_stream, scenes)
    if args.roofline_only:
        only = list(KINDS) if args.roofline_only == "all" else [args.roofline_only]
        roof = roofline_leg(B, scenes, roof_stream, roof_offsets, pools_dev, args.roofline_queries,
                            kinds=[KINDS.index(k) for k in only])
        for k in only:
            print(json.dumps({"roofline_only": k, "roofline_seed": ROOF_SEED,
                              "max_prims": args.max_prims, "roofline_queries": args.roofline_queries,
                              **roof[k]}), flush=True)
        return 0

    step_s = STREAM_SECONDS / args.steps
    if args.warmup:
        log(f"warmup: {args.warmup} steps ({args.warmup * step_s:.1f} s of stream)")
        replay_stream(B, handles, scenes, pools_dev, np.random.default_rng(SEED + 5),
                      np.random.default_rng(SEED + 6), args.warmup * step_s, args.rate, serve_kw)
    log(f"timed: {args.steps} steps of {step_s:.2f} s ({STREAM_SECONDS:.0f} s of stream, one serving loop)")
    served, checks, (stream, offsets, reqs) = replay_stream(
        B, handles, scenes, pools_dev, np.random.default_rng(SEED + 2), np.random.default_rng(SEED + 3),
        STREAM_SECONDS, args.rate, serve_kw, CAPTURE_EVERY, np.random.default_rng(SEED + 4))
    value = served["served_q"] / served["wall_s"]
    lat = served["latency_ms"]
    log(f"served: {served['served_q']} queries in {served['wall_s']:.3f}s ({value/1e6:.2f} Mq/s), "
        f"p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms, "
        f"{served['batches']} batches, rejected {served['rejected']}")
    per_kind = {}
    for k in range(4):
        sel = served["latency_kind"] == k
        if sel.any():
            per_kind[KINDS[k]] = {"requests": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                                  "p99_ms": float(np.percentile(lat[sel], 99))}
    per_step_p99 = [float(np.percentile(lat[(served["arrival_s"] >= i * step_s) & (served["arrival_s"] < (i + 1) * step_s)], 99))
                    for i in range(args.steps)]
    t0 = time.perf_counter()
    verified = verify_captures(checks, scenes, pools_host)
    verified["seconds"] = time.perf_counter() - t0
    log(f"served rows checked against the oracle: {verified}")

    # capacity: the requests of the first steps again, all arrived at time 0
    n_sat = int(np.searchsorted(stream["arrival"], args.saturated_steps * step_s))
    sat = reqs[:n_sat].copy()
    sat["capture"] = 0
    sat["arrival_s"] = 0.0
    res, sst, _ = B.serve_replay(handles, sat, point_queries=pools_dev["point"], ray_queries=pools_dev["ray"],
                                 pip_queries=pools_dev["pip"], **serve_kw)
    capacity = int(stream["size"][:n_sat][res["status"] == 0].sum()) / sst["wall_s"]
    log(f"saturated replay of {args.saturated_steps} steps' requests: {capacity/1e6:.2f} Mq/s")

    roof = roofline_leg(B, scenes, roof_stream, roof_offsets, pools_dev, args.roofline_queries)
    measured = {k: v for k, v in roof.items() if "kernel_s" in v}
    # the kernel family that takes the largest share of the stream's device time
    stream_q = {k: int(stream["size"][stream["kind"] == KINDS.index(k)].sum()) for k in measured}
    dominant = max(measured, key=lambda k: measured[k]["kernel_s"] / max(1, measured[k]["queries"]) * stream_q[k])
    d = measured[dominant]

    first_stream, first_offsets = stream, offsets
    baselines = {}
    cpu = None
    if not args.skip_baselines:
        baselines["plain_per_call"] = plain_per_call_leg(rtm, scenes, first_stream, first_offsets, pools_host,
                                                         args.percall_seconds)
        log(f"plain per-call: {baselines['plain_per_call']}")
        baselines["per_request_resident"] = per_request_resident_leg(B, scenes, first_stream, first_offsets,
                                                                      pools_dev, args.resident_seconds)
        log(f"per-request resident: {baselines['per_request_resident']}")
        cpu = cpu_baseline_leg(rtm, scenes, first_stream, first_offsets, pools_host, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
        baselines["cpu_python_reference_subsamples"] = python_subsample_leg()
        log(f"python reference subsamples: {baselines['cpu_python_reference_subsamples']}")
        baselines["embree"] = embree_status()
    configs = {}
    if not args.skip_configs:
        configs["config0_fixed_radius_4096_2d_fp64"] = config0_leg(rtm)
        log(f"config [0]: {configs['config0_fixed_radius_4096_2d_fp64']}")
        baselines["cpu_python_reference"] = configs["config0_fixed_radius_4096_2d_fp64"]["cpu_python_reference"]
        configs.update(per_call_configs_leg(rtm))
        configs["config3_batched_fixed_radius_pip_2048"] = config3_leg(B, not args.no_config3_autotune)

    plain = baselines.get("plain_per_call", {}).get("queries_per_s")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": served["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "served request stream: Poisson 20k req/s for 60 s, 25% each fixed-radius r=0.01 "
                        "k_max=64 / KNN k=8 / 3-D ray any-hit tmax=0.5 / PIP positive hits, batch log-uniform "
                        "[1, 65536], 16 resident scenes of 1e5-1e7 FP32 primitives (uniform / 64-Gaussian "
                        "mixture), 200 GB admission budget",
            "stream_seconds": STREAM_SECONDS,
            "step_seconds": step_s,
            "requests": served["requests"],
            "queries": served["queries"],
            "slots": args.slots,
            "memory_budget_gb": args.budget_gb,
            "row_arena_gb": served["region_bytes"] / (1 << 30),
            "max_prims": args.max_prims,
        },
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "rows_per_s": served["rows"] / served["wall_s"],
        "served": {"wall_s": served["wall_s"], "batches": served["batches"], "rows": served["rows"],
                   "rejected": served["rejected"], "peak_inflight_bytes": served["peak_inflight_bytes"],
                   "per_kind": per_kind, "p99_ms_per_step": per_step_p99},
        "served_rows_checked": verified["served_rows_checked"],
        "mismatches": verified["mismatched_queries"],
        "served_check": verified,
        "saturated_queries_per_s": capacity,
        "offered_queries_per_s": served["queries"] / STREAM_SECONDS,
        "baselines": baselines,
        "configs": configs,
        "speedup_vs_plain_per_call": (value / plain) if plain else None,
        "speedup_vs_per_request_resident": (value / baselines["per_request_resident"]["queries_per_s"])
        if baselines.get("per_request_resident", {}).get("queries_per_s") else None,
        "roofline": {
            "bound": "hbm",
            "kernel": f"{d['kernel']} ({dominant})",
            "achieved": d["achieved_GBps"],
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": d["achieved_GBps"] * 1e9 / HBM_PEAK,
            "traffic": measured_traffic(dominant, d, args),
            "roofline_seed": ROOF_SEED,
            "per_kind": roof,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
