"""Autotuner for the batched path (no counterpart in the reference).

It picks the process-wide build and launch choices of the library (the LBVH
leaf size per primitive type, the persistent launch size, the KNN walk's
stack depth and the ray / PIP walks' workgroup size, the last two as kernel
template variants; see ``RtdlTuning`` in include/rtdl_mi355x.h) for a
caller's workload. It times
candidates on that workload, one knob at a time (coordinate descent from the
current values), and checks on the way that every candidate gives the same
rows. The chosen values can be cached in a JSON file under a caller-chosen
workload key, so later processes apply them without timing anything.
"""
from __future__ import annotations

import ctypes
import json
import os
import statistics
from dataclasses import asdict, dataclass, replace
from typing import Callable

import numpy as np

from . import batch as B
from . import mi355x_runtime as rt_native


@dataclass(frozen=True)
class Tuning:
    point_leaf: int = 2
    triangle_leaf: int = 1
    polygon_leaf: int = 1
    blocks_per_cu: int = 8
    knn_stack: int = 32
    walk_block: int = 256


class _CTuning(ctypes.Structure):
    _fields_ = [("point_leaf", ctypes.c_uint32), ("triangle_leaf", ctypes.c_uint32),
                ("polygon_leaf", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("knn_stack", ctypes.c_uint32), ("walk_block", ctypes.c_uint32)]


# values tried per knob, in the order the descent visits the knobs
KNOBS: dict[str, tuple[int, ...]] = {
    "blocks_per_cu": (2, 4, 8, 16),
    "point_leaf": (2, 4, 6, 8),
    "triangle_leaf": (1, 2, 4),
    "polygon_leaf": (1, 2),
    "knn_stack": (16, 32, 48),
    "walk_block": (64, 128, 256),
}


def get_tuning() -> Tuning:
    out = _CTuning()
    err = ctypes.create_string_buffer(1024)
    st = rt_native._library().rtdl_mi355x_get_tuning(ctypes.byref(out), err, 1024)
    if st != 0:
        raise RuntimeError(f"rtdl_mi355x_get_tuning failed: {err.value.decode(errors='replace')}")
    return Tuning(out.point_leaf, out.triangle_leaf, out.polygon_leaf, out.blocks_per_cu, out.knn_stack,
                  out.walk_block)


def set_tuning(t: Tuning) -> None:
    c = _CTuning(t.point_leaf, t.triangle_leaf, t.polygon_leaf, t.blocks_per_cu, t.knn_stack, t.walk_block)
    err = ctypes.create_string_buffer(1024)
    st = rt_native._library().rtdl_mi355x_set_tuning(ctypes.byref(c), err, 1024)
    if st != 0:
        raise ValueError(f"rtdl_mi355x_set_tuning failed: {err.value.decode(errors='replace')}")


def coordinate_descent(evaluate: Callable[[Tuning], float], start: Tuning,
                       knobs: dict[str, tuple[int, ...]] = KNOBS) -> tuple[Tuning, list[dict]]:
    """Minimises evaluate(tuning) (seconds) one knob at a time: each knob's
    values are tried with every other knob at its best value so far.
    Returns the best tuning and one table row per evaluated tuning."""
    seen: dict[Tuning, float] = {}

    def score(t: Tuning) -> float:
        if t not in seen:
            seen[t] = float(evaluate(t))
        return seen[t]

    best, best_s = start, score(start)
    for name, values in knobs.items():
        for v in values:
            t = replace(best, **{name: v})
            s = score(t)
            if s < best_s:
                best, best_s = t, s
    table = [{**asdict(t), "seconds": s} for t, s in seen.items()]
    return best, table


def _canonical_rows(job: B.BatchJob, result: dict) -> dict[str, np.ndarray]:
    rows = B.job_rows(job, result)
    names = list(rows.dtype.names)
    order = np.lexsort(tuple(rows[n] for n in reversed(names)))
    return {n: rows[n][order] for n in names}


def autotune_batch(build_scenes: Callable[[], object], make_jobs: Callable[[object], list],
                   *, knobs: dict[str, tuple[int, ...]] | None = None, repeats: int = 3,
                   morton_sort: bool = True, check_rows: bool = True, check_jobs: list[int] | None = None,
                   cache_path: str | None = None, key: str | None = None) -> dict:
    """Tunes the library for one workload and leaves the best tuning set.

    build_scenes() builds the workload's scenes (under the tuning being
    tried); make_jobs(scenes) returns its BatchJobs. Each candidate is timed
    as the median of `repeats` runs after a warm run. With `check_rows`,
    every candidate's rows must equal the first candidate's, field by field
    after the canonical order, or RuntimeError is raised and the original
    tuning restored; `check_jobs` limits the comparison to those job indices
    (for batches whose full row set is too large to compare on the host).
    With `cache_path` and `key`, a cached choice is applied without timing,
    and a new choice is stored."""
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            best = Tuning(**cache[key])
            set_tuning(best)
            return {"best": best, "table": [], "cached": True}
    original = get_tuning()
    reference: list | None = None

    def evaluate(t: Tuning) -> float:
        nonlocal reference
        set_tuning(t)
        scenes = build_scenes()
        jobs = make_jobs(scenes)
        res, _ = B.run_batch(jobs, morton_sort=morton_sort)
        if any(r["status"] for r in res):
            raise RuntimeError(f"tuning {t}: a job reported status {[r['status'] for r in res]}")
        if check_rows:
            picked = range(len(jobs)) if check_jobs is None else check_jobs
            rows = [_canonical_rows(jobs[i], res[i]) for i in picked]
            if reference is None:
                reference = rows
            else:
                for i, (got, want) in enumerate(zip(rows, reference)):
                    for name in want:
                        if not np.array_equal(got[name], want[name]):
                            raise RuntimeError(f"tuning {t} changed field {name} of job {i}")
        times = [B.run_batch(jobs, morton_sort=morton_sort)[1] for _ in range(max(1, repeats))]
        return statistics.median(times)

    try:
        best, table = coordinate_descent(evaluate, original, knobs or KNOBS)
    except BaseException:
        set_tuning(original)
        raise
    set_tuning(best)
    if cache_path and key:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = asdict(best)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return {"best": best, "table": table, "cached": False}
