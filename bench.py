#!/usr/bin/env python3
"""Benchmark: served request stream vs the plain per-call backend on 1 MI355X.

Workload (the served request mix): Poisson arrivals at 20,000 requests/s;
each request is a fixed-radius (r = 0.01, k_max = 64), KNN (k = 8), 3-D ray
any-hit (tmax = 0.5) or point-in-polygon (positive hits) query set, 25 % each,
with batch size floor(exp(U[0, ln 65536])), against 16 resident scenes whose
sizes are log-spaced in [1e5, 1e7] primitives (4 per request type,
alternating uniform / 64-component Gaussian mixture, sigma 0.02, FP32).
PIP scenes are star polygons (8-32 vertices) whose radii U[0.005, 0.02] are
scaled by sqrt(1000 / N) so that a point lies in about as many polygons as
with 1,000 polygons.

One step = one second of the request stream.  `value` is the number of
queries completed by the served backend divided by the wall time from the
start of the timed replay until its last request completed (inputs resident
in HBM before the clock starts).  Beside it: p50/p99 latency (arrival to rows
ready in device memory), the plain per-call C-ABI backend on a prefix of the
same stream (host records in, host rows out, scene built per call), a
per-request resident-scene run (one synchronous batch per request), the HBM
roofline of the dominant traversal kernel and the multi-threaded C oracle
on the host cores.
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
RADIUS = 0.01
K_MAX = 64
K_KNN = 8
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
            handle = B.prepare_point_instances(dev, [0, n])
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


def make_arenas(stream, scenes, centres, seed: int):
    """Query arenas (host copies + device buffers); returns (host arenas,
    device arenas, per-request offsets, totals)."""
    from rtdl_amd.device import DeviceBuffer

    rng = np.random.default_rng(seed)
    n = len(stream["kind"])
    offsets = np.zeros(n, dtype=np.int64)
    # neighbour requests share the point arena, drawn per request from its
    # scene's distribution; rays are uniform in the unit box, PIP points
    # follow their scene's distribution
    dists = np.array([scenes[s]["dist"] == "mixture" for s in stream["scene"]], dtype=bool)
    host, dev, totals = {}, {}, {}
    for group, kinds in (("point", (0, 1)), ("ray", (2,)), ("pip", (3,))):
        sel = np.isin(stream["kind"], kinds)
        idx = np.flatnonzero(sel)
        sizes = stream["size"][idx]
        offsets[idx] = np.concatenate([[0], np.cumsum(sizes)[:-1]]) if len(idx) else []
        total = int(sizes.sum())
        totals[group] = total
        if group in ("point", "pip"):
            dim = 3 if group == "point" else 2
            q = rng.random((total, dim), dtype=np.float32)
            mix = np.repeat(dists[idx], sizes)
            pick = rng.integers(0, 64, total)
            g = centres[pick, :dim] + 0.02 * rng.standard_normal((total, dim), dtype=np.float32)
            q = np.where(mix[:, None], g, q).astype(np.float32)
        else:
            o = rng.random((total, 3))
            d = rng.standard_normal((total, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            q = np.concatenate([o, d, np.full((total, 1), TMAX)], 1).astype(np.float32)
        host[group] = q
        dev[group] = DeviceBuffer.from_numpy(q if total else np.zeros((1, q.shape[1]), np.float32))
    return host, dev, offsets, totals


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
        nodes, prims = visits
        if k == 3:
            # polygon refs (16 B) plus 8 B per vertex; the mean star polygon has 20 vertices
            pb = 16 + 8 * 20
        else:
            pb = prim_bytes[k]
        w = q_total * probe_bytes[k] + 32 * nodes + pb * prims + row_bytes[k] * rows
        # implementation-independent floor: every probe, every primitive of
        # the scenes used and every row once
        used = {int(stream["scene"][i]) for i in idx[:len(jobs)]}
        n_prims = sum(scenes[i]["n"] for i in used)
        w_min = q_total * probe_bytes[k] + pb * n_prims + row_bytes[k] * rows
        out[KINDS[k]] = {
            "queries": q_total, "jobs": len(jobs), "kernel_s": t, "node_visits": nodes,
            "prim_tests": prims, "rows": rows, "algorithmic_bytes": w, "w_min_bytes": w_min,
            "achieved_GBps": w / t / 1e9 if t > 0 else 0.0,
            "queries_per_s": q_total / t if t > 0 else 0.0,
        }
        log(f"kernel {KINDS[k]}: {q_total} q in {t*1e3:.2f} ms ({q_total/t/1e6:.1f} Mq/s), "
            f"{nodes/q_total:.1f} nodes/q, {prims/q_total:.1f} prims/q, W/t = {w/t/1e9:.0f} GB/s")
    return out


def measured_traffic(kind: str, measured: dict, args):
    """HBM bytes per launch of the roofline batch, from the committed
    rocprofv3 --pmc measurements of this same workload (FETCH_SIZE +
    WRITE_SIZE, each scaled by the counter's reading for a copy of known size),
    when one matches this run's batch; otherwise None."""
    path = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return None
    for rec in recs if isinstance(recs, list) else [recs]:
        if (rec.get("kind") == kind and rec.get("queries") == measured["queries"]
                and rec.get("steps") == args.steps and rec.get("max_prims") == args.max_prims
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
    runs once untimed, then once timed: host records in, LBVH built, walked
    and rows copied back to the host, with the library's phase timings."""
    out = {}
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
    p = rtm._ptr
    for rep in range(2):
        t0 = time.perf_counter()
        rows = rtm._call_rows("rtdl_mi355x_run_ray_anyhit_3d", rtm.ANYHIT_ROW_DTYPE, p(rays), len(rays),
                              p(tris), len(tris))
        wall = time.perf_counter() - t0
    ph = rtm.last_phase_timings()
    out["ray_anyhit_2M_triangles_1M_rays"] = {
        "wall_s": wall, "rays_per_s": len(rays) / wall, "hits": int(rows["any_hit"].sum()),
        "phases_ms": {k: v * 1e3 for k, v in ph.items()},
        "device_rays_per_s": len(rays) / ph["traverse"] if ph.get("traverse") else None}
    log(f"per-call ray any-hit 2M x 1M: {wall*1e3:.1f} ms wall, phases {out['ray_anyhit_2M_triangles_1M_rays']['phases_ms']}")
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
    for rep in range(2):
        t0 = time.perf_counter()
        rows = rtm._call_rows("rtdl_mi355x_run_knn_rows_3d", rtm.KNN_ROW_DTYPE, p(query), len(query),
                              p(search), len(search), 16)
        wall = time.perf_counter() - t0
    ph = rtm.last_phase_timings()
    out["knn16_4M_x_4M_mixture"] = {
        "wall_s": wall, "queries_per_s": len(query) / wall, "rows": int(len(rows)),
        "phases_ms": {k: v * 1e3 for k, v in ph.items()},
        "device_queries_per_s": len(query) / ph["traverse"] if ph.get("traverse") else None}
    log(f"per-call KNN k=16 4M x 4M: {wall*1e3:.1f} ms wall, phases {out['knn16_4M_x_4M_mixture']['phases_ms']}")
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


def python_baseline_leg():
    """The pure-Python FP64 restatement of the reference, single-threaded, on
    the reference's own CPU-runnable case: 4,096 query and 4,096 search
    points, 2-D, uniform in [0, 1)^2, radius 0.02, k_max = 4,096 (never
    truncates)."""
    from oracle import cpu_python_reference as ref
    from rtdl_amd.records import Point

    rng = np.random.default_rng(20261203)
    q = rng.random((4096, 2))
    sp = rng.random((4096, 2))
    qs = [Point(i, float(x), float(y)) for i, (x, y) in enumerate(q)]
    ss = [Point(i, float(x), float(y)) for i, (x, y) in enumerate(sp)]
    t0 = time.perf_counter()
    rows = ref.fixed_radius_neighbors_cpu(qs, ss, radius=0.02, k_max=4096)
    dt = time.perf_counter() - t0
    return {"value": len(qs) / dt, "unit": "queries/s", "cores": 1, "kind": "port",
            "sample": f"fixed-radius 4096 x 4096 2-D points, r = 0.02: {len(rows)} rows in {dt:.1f} s"}


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


def embree_status() -> str:
    import ctypes.util

    return "available" if ctypes.util.find_library("embree4") else "unavailable (no libembree4 on this host)"


# ----------------------------------------------------------------- main --


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="seconds of request stream timed")
    ap.add_argument("--warmup", type=int, default=1, help="seconds of request stream replayed untimed")
    ap.add_argument("--rate", type=float, default=20000.0)
    ap.add_argument("--max-prims", type=float, default=1e7)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-batch-queries", type=int, default=4 << 20)
    ap.add_argument("--percall-seconds", type=float, default=8.0)
    ap.add_argument("--resident-seconds", type=float, default=4.0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--roofline-queries", type=int, default=2 << 20)
    ap.add_argument("--skip-baselines", action="store_true")
    ap.add_argument("--roofline-only", choices=KINDS, default=None,
                    help="only build the scenes and streams of a full run and measure this kind's "
                         "roofline batch (the batch the bench line's roofline is quoted on), for "
                         "profiling that batch alone")
    ap.add_argument("--per-call-only", action="store_true",
                    help="run only the per-call 2M x 1M ray any-hit and 4M x 4M KNN k=16 calls")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (the reference kernel "
                         "for scaling FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every query set is independent, multi-GPU sharding is a "
              "design note only; run with --gpus 1", file=sys.stderr)
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

    info = rtm.mi355x_context_probe()
    log(f"device {info}")
    if args.per_call_only:
        print(json.dumps({"per_call_configs": per_call_configs_leg(rtm)}), flush=True)
        return 0
    rng = np.random.default_rng(SEED)
    scenes, centres = make_scenes(B, rng, args.max_prims)
    scene_slot = {s["index"]: s["index"] for s in scenes}
    handles = [s["handle"] for s in scenes]

    warm = make_stream(rng, args.warmup, args.rate, scenes)
    timed = make_stream(rng, args.steps, args.rate, scenes)
    log(f"stream: warmup {len(warm['kind'])} requests, timed {len(timed['kind'])} requests, "
        f"{int(timed['size'].sum())} queries")
    serve_kw = dict(radius=RADIUS, k_max=K_MAX, k=K_KNN, slots=args.slots,
                    memory_budget_bytes=int(args.budget_gb * (1 << 30)),
                    max_batch_queries=args.max_batch_queries)
    per_slot = min(serve_kw["memory_budget_bytes"] // args.slots,
                   (args.max_batch_queries + 65536) * K_MAX * 24 + (64 << 20))
    log(f"admission budget {args.budget_gb:.0f} GiB over {args.slots} slots; "
        f"{per_slot * args.slots / (1 << 30):.1f} GiB of row buffers resident")

    if args.roofline_only:
        host, arenas, offsets, totals = make_arenas(timed, scenes, centres, SEED + 2)
        roof = roofline_leg(B, scenes, timed, offsets, arenas, args.roofline_queries,
                            kinds=[KINDS.index(args.roofline_only)])
        print(json.dumps({"roofline_only": args.roofline_only, "steps": args.steps,
                          "max_prims": args.max_prims, "roofline_queries": args.roofline_queries,
                          **roof[args.roofline_only]}), flush=True)
        return 0

    if len(warm["kind"]):
        _, wa, wo, _ = make_arenas(warm, scenes, centres, SEED + 1)
        wreq = requests_array(B, warm, wo, scene_slot)
        log(f"warmup replay: {len(wreq)} requests")
        B.serve_replay(handles, wreq, point_queries=wa["point"], ray_queries=wa["ray"],
                       pip_queries=wa["pip"], **serve_kw)
        del wa
        log("warmup replay done")

    host, arenas, offsets, totals = make_arenas(timed, scenes, centres, SEED + 2)
    reqs = requests_array(B, timed, offsets, scene_slot)
    log(f"timed replay: {len(reqs)} requests")
    t0 = time.perf_counter()
    results, stats, _ = B.serve_replay(handles, reqs, point_queries=arenas["point"],
                                       ray_queries=arenas["ray"], pip_queries=arenas["pip"], **serve_kw)
    wall = time.perf_counter() - t0
    ok = results["status"] == 0
    lat = (results["done_s"] - results["arrival_s"])[ok] * 1e3
    served_q = int(timed["size"][ok].sum())
    value = served_q / stats["wall_s"]
    log(f"served: {served_q} queries in {stats['wall_s']:.3f}s ({value/1e6:.2f} Mq/s), "
        f"p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms, "
        f"{stats['batches']} batches, rejected {int((~ok).sum())}")
    per_kind = {}
    for k in range(4):
        sel = (timed["kind"] == k) & ok
        if sel.any():
            lk = (results["done_s"] - results["arrival_s"])[sel] * 1e3
            per_kind[KINDS[k]] = {"requests": int(sel.sum()), "queries": int(timed["size"][sel].sum()),
                                  "p50_ms": float(np.percentile(lk, 50)), "p99_ms": float(np.percentile(lk, 99))}

    log(f"timed replay wall (host clock) {wall:.3f}s")
    # capacity: the same requests, all arrived at time 0 (a saturated replay)
    sat = reqs.copy()
    sat["arrival_s"] = 0.0
    _, sat_stats, _ = B.serve_replay(handles, sat, point_queries=arenas["point"], ray_queries=arenas["ray"],
                                     pip_queries=arenas["pip"], **serve_kw)
    capacity = served_q / sat_stats["wall_s"]
    log(f"saturated replay (every request arrived at 0): {capacity/1e6:.2f} Mq/s, "
        f"{sat_stats['batches']} batches")
    roof = roofline_leg(B, scenes, timed, offsets, arenas, args.roofline_queries)
    measured = {k: v for k, v in roof.items() if "kernel_s" in v}
    # the kernel family that takes the largest share of the stream's device time
    dominant = max(measured, key=lambda k: measured[k]["kernel_s"] / max(1, measured[k]["queries"])
                   * max(1, int(timed["size"][timed["kind"] == KINDS.index(k)].sum())))
    d = measured[dominant]

    baselines = {}
    cpu = None
    if not args.skip_baselines:
        baselines["plain_per_call"] = plain_per_call_leg(rtm, scenes, timed, offsets, host,
                                                         args.percall_seconds)
        log(f"plain per-call: {baselines['plain_per_call']}")
        baselines["per_request_resident"] = per_request_resident_leg(B, scenes, timed, offsets,
                                                                      arenas, args.resident_seconds)
        log(f"per-request resident: {baselines['per_request_resident']}")
        cpu = cpu_baseline_leg(rtm, scenes, timed, offsets, host, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
        baselines["cpu_python_reference"] = python_baseline_leg()
        log(f"python reference: {baselines['cpu_python_reference']}")
        baselines["cpu_python_reference_subsamples"] = python_subsample_leg()
        log(f"python reference subsamples: {baselines['cpu_python_reference_subsamples']}")
        baselines["embree"] = embree_status()
        baselines["per_call_configs"] = per_call_configs_leg(rtm)

    plain = baselines.get("plain_per_call", {}).get("queries_per_s")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": stats["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "served request stream: Poisson 20k req/s, 25% each fixed-radius r=0.01 k_max=64 / "
                        "KNN k=8 / 3-D ray any-hit tmax=0.5 / PIP positive hits, batch log-uniform [1, 65536], "
                        "16 resident scenes of 1e5-1e7 FP32 primitives (uniform / 64-Gaussian mixture)",
            "stream_seconds": args.steps,
            "requests": int(len(timed["kind"])),
            "queries": int(timed["size"].sum()),
            "slots": args.slots,
            "memory_budget_gb": args.budget_gb,
            "row_region_gb_per_slot": stats["region_bytes"] / (1 << 30),
            "max_prims": args.max_prims,
        },
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "rows_per_s": stats["rows"] / stats["wall_s"],
        "served": {"wall_s": stats["wall_s"], "batches": stats["batches"], "rows": stats["rows"],
                   "rejected": int((~ok).sum()), "peak_inflight_bytes": stats["peak_inflight_bytes"],
                   "per_kind": per_kind},
        "saturated_queries_per_s": capacity,
        "offered_queries_per_s": int(timed["size"].sum()) / args.steps,
        "baselines": baselines,
        "speedup_vs_plain_per_call": (value / plain) if plain else None,
        "speedup_vs_per_request_resident": (value / baselines["per_request_resident"]["queries_per_s"])
        if baselines.get("per_request_resident", {}).get("queries_per_s") else None,
        "roofline": {
            "bound": "hbm",
            "kernel": {"frn": "nbr_count/emit/sort_kernel", "knn": "nbr_count/emit/sort_kernel",
                       "ray": "batch_rays_kernel", "pip": "batch_pip_kernel"}[dominant] + f" ({dominant})",
            "achieved": d["achieved_GBps"],
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": d["achieved_GBps"] * 1e9 / HBM_PEAK,
            "traffic": measured_traffic(dominant, d, args),
            "per_kind": roof,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
