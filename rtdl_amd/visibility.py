"""Visibility rows on the MI355X any-hit path.

Mirrors the reference's visibility helpers (src/rtdsl/visibility_runtime.py:44-155
and their row rules in src/rtdsl/reference.py:275-346):

- every observer/target pair becomes one ray from the observer to the target,
  d = target - observer and tmax = 1.0, so the ray is the closed segment
  between the two points (2-D rays are tested as segments, 3-D rays accept
  hits at t in [0, 1]);
- ray ids run observer-major (observer 0 with every target, then observer 1 ...),
  and the rows come back in that order;
- ``visible`` is 1 - any_hit of that ray against the blocker triangles.

The rays are built with numpy on the host and traced by the native any-hit
entry points (``rtdl_mi355x_run_ray_anyhit_2d`` / ``_3d``); there is no CPU
path. Errors follow the reference's messages: mixed dimensions, blockers of
the other dimension, and an observer equal to its target raise ValueError.
"""
from __future__ import annotations

from typing import Iterable

import numpy as np

from . import mi355x_runtime as m
from .records import Point3D, Triangle3D, normalize_records

VISIBILITY_ROW_DTYPE = np.dtype([("observer_id", "<u4"), ("target_id", "<u4"), ("visible", "<u4")])

_BACKENDS = {"mi355x", "mi355", "hip", "rtdl_amd"}


def _check_backend(backend: str, what: str) -> None:
    name = backend.lower().replace("-", "_")
    if name not in _BACKENDS:
        raise ValueError(
            f"{what} backend must be 'mi355x' (this package has no CPU path; "
            f"the reference's visibility_rows_cpu is the CPU baseline), got {backend!r}"
        )


_MIXED = "visibility_rows requires all observers and targets to share one dimensionality"


def _points(payload, name: str, mixed: str = _MIXED) -> tuple[np.ndarray, np.ndarray, int]:
    """(ids uint32, coordinates (n, dim) float64, dim) of a point input; dim 0 when empty."""
    if isinstance(payload, np.ndarray) and payload.dtype.names or hasattr(payload, "keys"):
        arr, dim = m.pack_points(payload, name=name)
    else:
        records = normalize_records(name, "points", payload)
        if not records:
            return np.zeros(0, np.uint32), np.zeros((0, 2)), 0
        dims = {3 if isinstance(r, Point3D) else 2 for r in records}
        if len(dims) > 1:
            raise ValueError(mixed)
        arr, dim = m.pack_points(records, name=name)
    if len(arr) == 0:
        return np.zeros(0, np.uint32), np.zeros((0, 2)), 0
    keys = ("x", "y", "z")[:dim]
    return arr["id"].astype(np.uint32), np.stack([arr[k] for k in keys], 1), dim


def _blockers(payload, dim: int) -> np.ndarray:
    """Packed blocker triangles of dimension `dim` (RtdlTriangle or RtdlTriangle3D records)."""
    if isinstance(payload, np.ndarray) and payload.dtype.names or hasattr(payload, "keys"):
        names = set(payload.dtype.names if isinstance(payload, np.ndarray) else payload.keys())
        got = 3 if "z0" in names else 2
        n = len(payload["id"])
    else:
        payload = normalize_records("blockers", "triangles", payload)
        n = len(payload)
        dims = {3 if isinstance(t, Triangle3D) else 2 for t in payload}
        if len(dims) > 1:
            raise ValueError("visibility_rows blocker triangles must match observer/target dimensionality")
        got = dims.pop() if dims else dim
    if n and got != dim:
        raise ValueError("visibility_rows blocker triangles must match observer/target dimensionality")
    return m.pack_triangles(payload, name="blockers") if dim == 3 else m.pack_triangles_2d(payload, name="blockers")


def _trace(origin: np.ndarray, delta: np.ndarray, blockers: np.ndarray, dim: int) -> np.ndarray:
    """any_hit flags (uint32) of the segments origin -> origin + delta, in input order."""
    n = len(origin)
    if dim == 3:
        rays = np.zeros(n, dtype=m.RAY3D_DTYPE)
        rays["oz"], rays["dz"] = origin[:, 2], delta[:, 2]
    else:
        rays = np.zeros(n, dtype=m.RAY2D_DTYPE)
    rays["id"] = np.arange(n, dtype=np.uint32)
    rays["ox"], rays["oy"] = origin[:, 0], origin[:, 1]
    rays["dx"], rays["dy"] = delta[:, 0], delta[:, 1]
    rays["tmax"] = 1.0
    name = "rtdl_mi355x_run_ray_anyhit_3d" if dim == 3 else "rtdl_mi355x_run_ray_anyhit_2d"
    rows = m._call_rows(name, m.ANYHIT_ROW_DTYPE, m._ptr(rays), n, m._ptr(blockers), len(blockers))
    if len(rows) != n or not np.array_equal(rows["ray_id"], rays["id"]):
        raise RuntimeError("any-hit rows do not match the rays one to one")
    return rows["any_hit"]


def _rows(observer_ids, target_ids, any_hit, result_mode: str):
    out = np.zeros(len(any_hit), dtype=VISIBILITY_ROW_DTYPE)
    out["observer_id"], out["target_id"] = observer_ids, target_ids
    out["visible"] = (any_hit == 0).astype(np.uint32)
    if result_mode == "raw":
        return out
    if result_mode != "dict":
        raise ValueError("result_mode must be 'dict' or 'raw'")
    return tuple({"observer_id": int(o), "target_id": int(t), "visible": int(v)}
                 for o, t, v in zip(out["observer_id"].tolist(), out["target_id"].tolist(),
                                    out["visible"].tolist()))


def _check_distinct(delta: np.ndarray) -> None:
    if np.any(np.all(delta == 0.0, axis=1)):
        raise ValueError("visibility_rows requires distinct observer and target points")


def visibility_rows(observers, targets, blockers, *, backend: str = "mi355x", native_only: bool = False,
                    result_mode: str = "dict"):
    """Rows {observer_id, target_id, visible} for the full observer x target
    matrix, observer-major (reference: src/rtdsl/visibility_runtime.py:44-93).
    `native_only` is accepted for signature parity; every call is native.
    result_mode "raw" returns a numpy structured array instead of dicts."""
    _check_backend(backend, "visibility_rows")
    oid, oxy, odim = _points(observers, "observers")
    tid, txy, tdim = _points(targets, "targets")
    if odim == 0 or tdim == 0:
        return () if result_mode == "dict" else np.zeros(0, dtype=VISIBILITY_ROW_DTYPE)
    if odim != tdim:
        raise ValueError(_MIXED)
    tris = _blockers(blockers, odim)
    no, nt = len(oid), len(tid)
    origin = np.repeat(oxy, nt, axis=0)
    delta = np.tile(txy, (no, 1)) - origin
    _check_distinct(delta)
    hits = _trace(origin, delta, tris, odim)
    return _rows(np.repeat(oid, nt), np.tile(tid, no), hits, result_mode)


def visibility_pair_rows(observer_target_pairs: Iterable, blockers, *, backend: str = "mi355x",
                         native_only: bool = False, result_mode: str = "dict"):
    """Rows for explicit (observer, target) pairs in the given order
    (reference: src/rtdsl/visibility_runtime.py:96-155)."""
    _check_backend(backend, "visibility_pair_rows")
    pairs = tuple(observer_target_pairs)
    if not pairs:
        return () if result_mode == "dict" else np.zeros(0, dtype=VISIBILITY_ROW_DTYPE)
    mixed = "visibility_pair_rows requires all points to share one dimensionality"
    oid, oxy, odim = _points([p[0] for p in pairs], "observers", mixed)
    tid, txy, tdim = _points([p[1] for p in pairs], "targets", mixed)
    if odim != tdim:
        raise ValueError(mixed)
    delta = txy - oxy
    _check_distinct(delta)
    try:
        tris = _blockers(blockers, odim)
    except ValueError:
        raise ValueError("visibility_pair_rows blocker triangles must match observer/target dimensionality") from None
    hits = _trace(oxy, delta, tris, odim)
    return _rows(oid, tid, hits, result_mode)


__all__ = ["VISIBILITY_ROW_DTYPE", "visibility_pair_rows", "visibility_rows"]
