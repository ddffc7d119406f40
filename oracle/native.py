"""ORACLE — test infrastructure only, never a product path.

ctypes wrapper over ``oracle/librtdl_oracle.so``, the independent C
restatement of the FP64 reference predicates (see ``oracle/rtdl_oracle.c``).
Inputs are numpy arrays; outputs are numpy arrays in the reference's row
order (rows grouped by a stable sort on query id, each group ordered by
(distance, id); PIP pairs in point-then-polygon input order).

Pinned by tests/test_oracle_pins.py against the pure-Python restatement
(``oracle/cpu_python_reference.py``) on random inputs, with and without the
grid filter.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NN_ROW_DTYPE = np.dtype(
    [("qid", "<u4"), ("nid", "<u4"), ("dist", "<f8"), ("rank", "<u4"), ("pad", "<u4")]
)


def library() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librtdl_oracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run `make -C oracle`")
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        up = ctypes.POINTER(ctypes.c_uint32)
        lib.oracle_neighbors.restype = ctypes.c_int64
        lib.oracle_neighbors.argtypes = [
            ctypes.c_int, ctypes.c_int,
            ctypes.c_size_t, dp, dp, dp, up,
            ctypes.c_size_t, dp, dp, dp, up,
            ctypes.c_double, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
            ctypes.POINTER(ctypes.c_void_p),
        ]
        lib.oracle_rays_3d.restype = ctypes.c_int
        lib.oracle_rays_3d.argtypes = [
            ctypes.c_int, ctypes.c_size_t, dp, ctypes.c_size_t, dp, up,
            ctypes.c_int, ctypes.c_int, up, dp, ctypes.POINTER(ctypes.c_uint8),
        ]
        lib.oracle_pip.restype = ctypes.c_int64
        lib.oracle_pip.argtypes = [
            ctypes.c_size_t, dp, dp, ctypes.c_size_t, up, up, dp,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
        ]
        lib.oracle_free.restype = None
        lib.oracle_free.argtypes = [ctypes.c_void_p]
        _LIB = lib
    return _LIB


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _up(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def neighbors(mode: str, queries: dict, search: dict, *, radius: float = 0.0, k: int = 1,
              threads: int | None = None, brute: bool = False) -> np.ndarray:
    """mode: 'frn' | 'knn' | 'bounded_knn'.  Point dicts hold x, y, (z), id arrays."""
    dim = 3 if "z" in queries else 2
    q = {key: (_u32(v) if key == "id" else _f64(v)) for key, v in queries.items()}
    s = {key: (_u32(v) if key == "id" else _f64(v)) for key, v in search.items()}
    out = ctypes.c_void_p()
    code = {"frn": 0, "knn": 1, "bounded_knn": 2}[mode]
    n = library().oracle_neighbors(
        code, dim,
        len(q["x"]), _dp(q["x"]), _dp(q["y"]), _dp(q.get("z")), _up(q["id"]),
        len(s["x"]), _dp(s["x"]), _dp(s["y"]), _dp(s.get("z")), _up(s["id"]),
        float(radius), int(k), int(threads or default_threads()), int(bool(brute)),
        ctypes.byref(out),
    )
    try:
        rows = np.frombuffer(
            (ctypes.c_char * (n * NN_ROW_DTYPE.itemsize)).from_address(out.value), dtype=NN_ROW_DTYPE
        ).copy() if n else np.zeros(0, dtype=NN_ROW_DTYPE)
    finally:
        library().oracle_free(out)
    order = np.argsort(rows["qid"], kind="stable")
    return rows[order]


def rays_3d(mode: str, rays: np.ndarray, tris: np.ndarray, tri_ids: np.ndarray, *,
            threads: int | None = None, brute: bool = False):
    """rays: (n,7) ox oy oz dx dy dz tmax; tris: (m,9).  mode 'any'|'count'|'closest'."""
    r = _f64(rays).reshape(-1, 7)
    t = _f64(tris).reshape(-1, 9)
    ids = _u32(tri_ids)
    n = r.shape[0]
    out_u = np.zeros(n, dtype=np.uint32)
    out_t = np.zeros(n, dtype=np.float64)
    out_h = np.zeros(n, dtype=np.uint8)
    library().oracle_rays_3d(
        {"any": 0, "count": 1, "closest": 2}[mode], n, _dp(r), t.shape[0], _dp(t), _up(ids),
        int(threads or default_threads()), int(bool(brute)),
        _up(out_u), _dp(out_t), out_h.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
    )
    if mode == "closest":
        return out_u, out_t, out_h.astype(bool)
    return out_u


def rays_2d_any(rays: np.ndarray, tris: np.ndarray, *, threads: int | None = None,
                brute: bool = False) -> np.ndarray:
    """2-D any-hit flags.  rays: (n,5) ox oy dx dy tmax; tris: (m,6) x0 y0 x1 y1 x2 y2."""
    return rays_2d("any", rays, tris, np.arange(len(tris), dtype=np.uint32), threads=threads, brute=brute)


def rays_2d(mode: str, rays: np.ndarray, tris: np.ndarray, tri_ids: np.ndarray, *,
            threads: int | None = None, brute: bool = False):
    """2-D rays: (n,5) ox oy dx dy tmax; tris: (m,6).  mode 'any'|'count'|'closest'
    (closest: the lowest hit id, t = 0.0), returned as rays_3d does."""
    r2 = _f64(rays).reshape(-1, 5)
    t2 = _f64(tris).reshape(-1, 6)
    r = np.zeros((r2.shape[0], 7))
    r[:, [0, 1, 3, 4, 6]] = r2
    t = np.zeros((t2.shape[0], 9))
    t[:, [0, 1, 3, 4, 6, 7]] = t2
    n = r.shape[0]
    out_u = np.zeros(n, dtype=np.uint32)
    out_t = np.zeros(n, dtype=np.float64)
    out_h = np.zeros(n, dtype=np.uint8)
    ids = _u32(tri_ids)
    library().oracle_rays_3d(
        {"any": 3, "count": 4, "closest": 5}[mode], n, _dp(r), t.shape[0], _dp(t), _up(ids),
        int(threads or default_threads()), int(bool(brute)),
        _up(out_u), _dp(out_t), out_h.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
    )
    if mode == "closest":
        return out_u, out_t, out_h.astype(bool)
    return out_u


def pip_pairs(px, py, offsets, counts, vertices_xy, *, positive_only: bool,
              threads: int | None = None, brute: bool = False) -> np.ndarray:
    """Return an (n, 2) array of (point index, polygon index) with contains = 1."""
    px, py = _f64(px), _f64(py)
    off, cnt = _u32(offsets), _u32(counts)
    verts = _f64(vertices_xy)
    out = ctypes.c_void_p()
    n = library().oracle_pip(
        len(px), _dp(px), _dp(py), len(off), _up(off), _up(cnt), _dp(verts),
        int(bool(positive_only)), int(threads or default_threads()), int(bool(brute)), ctypes.byref(out),
    )
    try:
        pairs = np.frombuffer(
            (ctypes.c_char * (n * 8)).from_address(out.value), dtype=np.uint32
        ).reshape(-1, 2).copy() if n else np.zeros((0, 2), dtype=np.uint32)
    finally:
        library().oracle_free(out)
    return pairs
