"""Seeded input generators shared by the tests and the benchmark.

Distributions follow the workloads the backend is measured on: uniform
points, 64-component Gaussian mixtures (sigma 0.02), jittered small
triangles with unit-direction rays, and star-shaped polygons.
"""
from __future__ import annotations

import numpy as np

TRI_FIELDS = ["x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2"]
RAY_FIELDS = ["ox", "oy", "oz", "dx", "dy", "dz", "tmax"]


def f32(a: np.ndarray) -> np.ndarray:
    """Round to FP32 and back, so values are exact in both precisions."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def points(rng, n: int, dim: int, *, kind: str = "uniform", fp32: bool = True,
           ids: np.ndarray | None = None, centres: np.ndarray | None = None) -> dict:
    if kind == "uniform":
        xyz = rng.random((n, dim))
    else:
        if centres is None:
            centres = rng.random((64, dim))
        pick = rng.integers(0, len(centres), n)
        xyz = centres[pick] + rng.normal(0.0, 0.02, (n, dim))
    if fp32:
        xyz = f32(xyz)
    out = {"id": np.arange(n, dtype=np.uint32) if ids is None else ids.astype(np.uint32),
           "x": xyz[:, 0].copy(), "y": xyz[:, 1].copy()}
    if dim == 3:
        out["z"] = xyz[:, 2].copy()
    return out


def triangles(rng, n: int, *, sigma: float = 0.005, fp32: bool = True) -> dict:
    c = rng.random((n, 3))
    v = (c[:, None, :] + rng.normal(0.0, sigma, (n, 3, 3))).reshape(n, 9)
    if fp32:
        v = f32(v)
    out = {"id": np.arange(n, dtype=np.uint32)}
    for i, f in enumerate(TRI_FIELDS):
        out[f] = v[:, i].copy()
    return out


def rays(rng, n: int, *, tmax: float = 0.5, fp32: bool = True) -> dict:
    o = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = np.concatenate([o, d, np.full((n, 1), tmax)], axis=1)
    if fp32:
        r = f32(r)
    out = {"id": np.arange(n, dtype=np.uint32)}
    for i, f in enumerate(RAY_FIELDS):
        out[f] = r[:, i].copy()
    return out


def star_polygons(rng, n: int, *, vmin: int = 8, vmax: int = 32, rmin: float = 0.005,
                  rmax: float = 0.02, fp32: bool = True) -> dict:
    counts = rng.integers(vmin, vmax + 1, n)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]])
    verts = []
    for i in range(n):
        ang = np.sort(rng.random(counts[i]) * 2.0 * np.pi)
        rad = rng.uniform(rmin, rmax, counts[i])
        c = rng.random(2)
        verts.append(np.stack([c[0] + rad * np.cos(ang), c[1] + rad * np.sin(ang)], axis=1))
    v = np.concatenate(verts).reshape(-1)
    if fp32:
        v = f32(v)
    refs = np.zeros(n, dtype=[("id", "<u4"), ("vertex_offset", "<u4"), ("vertex_count", "<u4")])
    refs["id"] = np.arange(n, dtype=np.uint32)
    refs["vertex_offset"] = offsets
    refs["vertex_count"] = counts
    return {"refs": refs, "vertices_xy": v}


def ray_matrix(r: dict) -> np.ndarray:
    return np.stack([r[f] for f in RAY_FIELDS], axis=1)


def tri_matrix(t: dict) -> np.ndarray:
    return np.stack([t[f] for f in TRI_FIELDS], axis=1)


def subset(cols: dict, idx) -> dict:
    return {k: v[idx] for k, v in cols.items()}


def fl32_tie_pairs(count: int = 1, start: float = 0.01) -> list[tuple[float, float, float, float, float]]:
    """FP32-exact points A = (x, ya) and B = (x, yb) around a query at the
    origin whose FP64 squared distances a < b (reference order: x*x + y*y)
    round to FP32 keys one ulp apart, while sqrt(a) == sqrt(b): the two are
    tied in the reference's (distance, id) order although their FP32 keys are
    not.  Returns `count` tuples (x, ya, yb, a, b), searched deterministically
    upwards from `start` (a few dozen FP32 steps per pair)."""
    import math

    def as_f32(v: float) -> float:
        return float(np.float32(v))

    out = []
    x = np.float32(start)
    while len(out) < count:
        x = np.nextafter(x, np.float32(1.0))
        xx = float(x) * float(x)  # exact: 24-bit by 24-bit
        lo = np.float32(xx)
        if float(lo) > xx:
            lo = np.nextafter(lo, np.float32(0.0))
        mid = (float(lo) + float(np.nextafter(lo, np.float32(1.0)))) / 2.0  # FP32 rounding point
        gap = mid - xx
        if not 0.0 < gap < 1e-13:  # y*y must bridge the gap at FP64 resolution
            continue
        y = np.float32(math.sqrt(gap))
        ys = [y]
        lo_y = hi_y = y
        for _ in range(40):
            lo_y = np.nextafter(lo_y, np.float32(0.0))
            hi_y = np.nextafter(hi_y, np.float32(1.0))
            ys += [lo_y, hi_y]
        d2 = {}
        for yv in ys:
            d2.setdefault(float(x) * float(x) + float(yv) * float(yv), float(yv))
        keys = sorted(d2)
        pair = next(((a, b) for a in keys for b in keys
                     if a < b and as_f32(a) != as_f32(b) and math.sqrt(a) == math.sqrt(b)), None)
        if pair:
            a, b = pair
            out.append((float(x), d2[a], d2[b], a, b))
    return out
