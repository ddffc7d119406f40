"""The ``mi355x`` backend: ``run_mi355x`` / ``prepare_mi355x`` over ctypes.

Mirrors the reference's HIPRT backend surface (reference:
src/rtdsl/hiprt_runtime.py:2433-2755 validation and ``run_hiprt``,
:2758-2948 prepared kernels, :2950-3061 ``prepare_hiprt``) for the hot-path
predicates: fixed-radius neighbours, KNN rows, bounded KNN rows, 3-D ray /
triangle any-hit, hit-count and closest-hit, and 2-D point-in-polygon.

Kernels are validated before the native library is loaded, with the
reference's exception types (``ValueError`` for malformed kernels and inputs,
``NotImplementedError`` for recognised but unsupported shapes; no CPU
fallback exists).  Inputs are the reference's record tuples (mappings or
dataclasses) or, for large inputs, columnar numpy data: a dict of arrays
(``{"id", "x", "y"[, "z"]}`` and the triangle / ray field names) or a numpy
structured array with those fields.  Rows come back as a tuple of dicts
(``result_mode="dict"``, the reference's contract) or as a numpy structured
array in the C row layout (``result_mode="raw"``).
"""
from __future__ import annotations

import ctypes
import ctypes.util
import functools
import math
import os
import weakref
from collections.abc import Mapping

import numpy as np

from .api import compile_kernel
from .ir import CompiledKernel
from .records import Point, Point3D, Polygon, Ray2D, Ray3D, Triangle, Triangle3D, normalize_records, project_rows

# --------------------------------------------------------------------------
# C record and row layouts (reference: src/rtdsl/hiprt_runtime.py:108-267,
# src/rtdsl/embree_runtime.py:420-479)
# --------------------------------------------------------------------------

POINT2D_DTYPE = np.dtype({"names": ["id", "x", "y"], "formats": ["<u4", "<f8", "<f8"],
                          "offsets": [0, 8, 16], "itemsize": 24})
POINT3D_DTYPE = np.dtype({"names": ["id", "x", "y", "z"], "formats": ["<u4", "<f8", "<f8", "<f8"],
                          "offsets": [0, 8, 16, 24], "itemsize": 32})
_TRI_FIELDS = ["x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2"]
TRIANGLE3D_DTYPE = np.dtype({"names": ["id", *_TRI_FIELDS], "formats": ["<u4"] + ["<f8"] * 9,
                             "offsets": [0] + [4 + 8 * i for i in range(9)], "itemsize": 76})
_RAY_FIELDS = ["ox", "oy", "oz", "dx", "dy", "dz", "tmax"]
RAY3D_DTYPE = np.dtype({"names": ["id", *_RAY_FIELDS], "formats": ["<u4"] + ["<f8"] * 7,
                        "offsets": [0] + [4 + 8 * i for i in range(7)], "itemsize": 60})
_TRI2_FIELDS = ["x0", "y0", "x1", "y1", "x2", "y2"]
TRIANGLE2D_DTYPE = np.dtype({"names": ["id", *_TRI2_FIELDS], "formats": ["<u4"] + ["<f8"] * 6,
                             "offsets": [0] + [8 + 8 * i for i in range(6)], "itemsize": 56})
_RAY2_FIELDS = ["ox", "oy", "dx", "dy", "tmax"]
RAY2D_DTYPE = np.dtype({"names": ["id", *_RAY2_FIELDS], "formats": ["<u4"] + ["<f8"] * 5,
                        "offsets": [0] + [4 + 8 * i for i in range(5)], "itemsize": 44})
POLYGON_REF_DTYPE = np.dtype([("id", "<u4"), ("vertex_offset", "<u4"), ("vertex_count", "<u4")])

FRN_ROW_DTYPE = np.dtype({"names": ["query_id", "neighbor_id", "distance"],
                          "formats": ["<u4", "<u4", "<f8"], "offsets": [0, 4, 8], "itemsize": 16})
KNN_ROW_DTYPE = np.dtype({"names": ["query_id", "neighbor_id", "distance", "neighbor_rank"],
                          "formats": ["<u4", "<u4", "<f8", "<u4"], "offsets": [0, 4, 8, 16],
                          "itemsize": 24})
ANYHIT_ROW_DTYPE = np.dtype([("ray_id", "<u4"), ("any_hit", "<u4")])
HITCOUNT_ROW_DTYPE = np.dtype([("ray_id", "<u4"), ("hit_count", "<u4")])
CLOSEST_ROW_DTYPE = np.dtype({"names": ["ray_id", "triangle_id", "t"],
                              "formats": ["<u4", "<u4", "<f8"], "offsets": [0, 4, 8], "itemsize": 16})
PIP_ROW_DTYPE = np.dtype([("point_id", "<u4"), ("polygon_id", "<u4"), ("contains", "<u4")])

_IMPLEMENTED = {
    "fixed_radius_neighbors",
    "knn_rows",
    "bounded_knn_rows",
    "ray_triangle_any_hit",
    "ray_triangle_hit_count",
    "ray_triangle_closest_hit",
    "point_in_polygon",
}

# --------------------------------------------------------------------------
# library discovery (reference: src/rtdsl/hiprt_runtime.py:308-332)
# --------------------------------------------------------------------------

_PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))


def _candidate_paths() -> list[str]:
    paths = []
    env = os.environ.get("RTDL_MI355X_LIB")
    if env:
        paths.append(env)
    paths.append(os.path.join(_PACKAGE_DIR, "librtdl_mi355x.so"))
    found = ctypes.util.find_library("rtdl_mi355x")
    if found:
        paths.append(found)
    return paths


@functools.lru_cache(maxsize=1)
def _library() -> ctypes.CDLL:
    errors = []
    for path in _candidate_paths():
        try:
            return _bind(ctypes.CDLL(path))
        except OSError as exc:
            errors.append(f"{path}: {exc}")
    message = (
        "RTDL MI355X backend library not found; build it with "
        "`python -c 'import __graft_entry__ as g; g.build()'` or set RTDL_MI355X_LIB"
    )
    if errors:
        message += " (" + "; ".join(errors) + ")"
    raise FileNotFoundError(message)


def library_path() -> str:
    lib = _library()
    return lib._name


_VP = ctypes.c_void_p
_SZ = ctypes.c_size_t
_ERR = [ctypes.c_char_p, ctypes.c_size_t]
_ROWS = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t)]

_SIGNATURES = {
    "rtdl_mi355x_get_version": ([ctypes.POINTER(ctypes.c_int)] * 3, ctypes.c_int),
    "rtdl_mi355x_context_probe": (
        [ctypes.c_char_p, _SZ, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), *_ERR],
        ctypes.c_int,
    ),
    "rtdl_mi355x_free_rows": ([_VP], None),
    "rtdl_mi355x_get_last_phase_timings": ([ctypes.POINTER(ctypes.c_double), _SZ], ctypes.c_int),
    "rtdl_mi355x_run_fixed_radius_neighbors_2d": ([_VP, _SZ, _VP, _SZ, ctypes.c_double, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_fixed_radius_neighbors_3d": ([_VP, _SZ, _VP, _SZ, ctypes.c_double, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_knn_rows_2d": ([_VP, _SZ, _VP, _SZ, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_knn_rows_3d": ([_VP, _SZ, _VP, _SZ, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_bounded_knn_rows_2d": ([_VP, _SZ, _VP, _SZ, ctypes.c_double, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_bounded_knn_rows_3d": ([_VP, _SZ, _VP, _SZ, ctypes.c_double, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_ray_anyhit_2d": ([_VP, _SZ, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_ray_hitcount_2d": ([_VP, _SZ, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_ray_closest_hit_2d": ([_VP, _SZ, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_ray_anyhit_3d": ([_VP, _SZ, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_ray_hitcount_3d": ([_VP, _SZ, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_ray_closest_hit_3d": ([_VP, _SZ, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_point_in_polygon": ([_VP, _SZ, _VP, _SZ, _VP, _SZ, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_prepare_points_2d": ([_VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_prepare_points_3d": ([_VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_fixed_radius_neighbors": ([_VP, _VP, _SZ, ctypes.c_double, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_knn_rows": ([_VP, _VP, _SZ, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_bounded_knn_rows": ([_VP, _VP, _SZ, ctypes.c_double, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_prepare_triangles_2d": ([_VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_ray_anyhit_2d": ([_VP, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_ray_hitcount_2d": ([_VP, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_ray_closest_hit_2d": ([_VP, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_prepare_triangles_3d": ([_VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_ray_anyhit_3d": ([_VP, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_ray_hitcount_3d": ([_VP, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_ray_closest_hit_3d": ([_VP, _VP, _SZ, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_prepare_polygons": ([_VP, _SZ, _VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_prepared_point_in_polygon": ([_VP, _VP, _SZ, ctypes.c_uint32, *_ROWS, *_ERR], ctypes.c_int),
    "rtdl_mi355x_destroy_prepared": ([_VP], None),
    "rtdl_mi355x_set_walk_counting": ([ctypes.c_uint32, *_ERR], ctypes.c_int),
    "rtdl_mi355x_set_fault": ([ctypes.c_char_p, *_ERR], ctypes.c_int),
    "rtdl_mi355x_get_device": ([ctypes.POINTER(ctypes.c_int), *_ERR], ctypes.c_int),
    "rtdl_mi355x_get_last_walk_stats": ([ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64), *_ERR],
                                        ctypes.c_int),
    # reference-shaped prepared handles (the HIPRT parameter lists)
    "rtdl_mi355x_prepare_fixed_radius_neighbors_3d": ([_VP, _SZ, ctypes.c_double, ctypes.POINTER(_VP), *_ERR],
                                                      ctypes.c_int),
    "rtdl_mi355x_run_prepared_fixed_radius_neighbors_3d": ([_VP, _VP, _SZ, ctypes.c_uint32, *_ROWS, *_ERR],
                                                           ctypes.c_int),
    "rtdl_mi355x_destroy_prepared_fixed_radius_neighbors_3d": ([_VP], None),
    "rtdl_mi355x_prepare_ray_hitcount_3d": ([_VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_destroy_prepared_ray_hitcount_3d": ([_VP], None),
    "rtdl_mi355x_prepare_ray_anyhit_2d": ([_VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_destroy_prepared_ray_anyhit_2d": ([_VP], None),
    "rtdl_mi355x_selftest_fp64": ([_VP, _VP, _SZ, _VP, _VP, *_ERR], ctypes.c_int),
    # batched / served execution (structs passed by address; see rtdl_amd.batch)
    "rtdl_mi355x_run_batch": ([_VP, _SZ, ctypes.c_uint32, _VP, ctypes.POINTER(ctypes.c_double), *_ERR], ctypes.c_int),
    "rtdl_mi355x_run_batch_noalloc": ([_VP, _SZ, ctypes.c_uint32, _VP, ctypes.POINTER(ctypes.c_double), *_ERR],
                                      ctypes.c_int),
    "rtdl_mi355x_run_batch_profiled": ([_VP, _SZ, ctypes.c_uint32, _VP, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_uint64), *_ERR], ctypes.c_int),
    "rtdl_mi355x_prepare_point_instances": ([_VP, _SZ, _VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_prepare_point_instances_leaf": ([_VP, _SZ, _VP, _SZ, ctypes.c_uint32, ctypes.POINTER(_VP), *_ERR],
                                                 ctypes.c_int),
    "rtdl_mi355x_prepare_triangle_instances": ([_VP, _SZ, _VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_prepare_polygon_instances": ([_VP, _VP, _SZ, _VP, _SZ, _VP, _SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_serve_replay": ([_VP, _SZ, _VP, _SZ, _VP, _VP, _VP, *_ERR], ctypes.c_int),
    "rtdl_mi355x_batch_graph_create": ([_VP, _SZ, ctypes.c_uint32, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_batch_graph_launch": ([_VP, _VP, ctypes.POINTER(ctypes.c_double), *_ERR], ctypes.c_int),
    "rtdl_mi355x_batch_graph_destroy": ([_VP, *_ERR], ctypes.c_int),
    "rtdl_mi355x_rebuild_instances": ([_VP, _VP, *_ERR], ctypes.c_int),
    "rtdl_mi355x_batch_graph_create_rebuild": ([_VP, _SZ, ctypes.c_uint32, _VP, _VP, _SZ, ctypes.POINTER(_VP),
                                                *_ERR], ctypes.c_int),
    "rtdl_mi355x_scene_export": ([_VP, _VP, _VP, _VP, _VP, _VP, _VP, *_ERR], ctypes.c_int),
    "rtdl_mi355x_scene_export_compact": ([_VP, _VP, _VP, _VP, *_ERR], ctypes.c_int),
    "rtdl_mi355x_get_tuning": ([_VP, *_ERR], ctypes.c_int),
    "rtdl_mi355x_set_tuning": ([_VP, *_ERR], ctypes.c_int),
    # device buffers (see rtdl_amd.device)
    "rtdl_mi355x_device_alloc": ([_SZ, ctypes.POINTER(_VP), *_ERR], ctypes.c_int),
    "rtdl_mi355x_device_free": ([_VP, *_ERR], ctypes.c_int),
    "rtdl_mi355x_copy_to_device": ([_VP, _VP, _SZ, *_ERR], ctypes.c_int),
    "rtdl_mi355x_copy_to_host": ([_VP, _VP, _SZ, *_ERR], ctypes.c_int),
    "rtdl_mi355x_copy_device": ([_VP, _VP, _SZ, *_ERR], ctypes.c_int),
}


def _bind(lib: ctypes.CDLL) -> ctypes.CDLL:
    for name, (argtypes, restype) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    return lib


def exported_symbols() -> tuple[str, ...]:
    return tuple(_SIGNATURES)


# --------------------------------------------------------------------------
# native call helpers
# --------------------------------------------------------------------------


def _ptr(array: np.ndarray | None):
    if array is None or array.size == 0:
        return None
    return ctypes.c_void_p(array.ctypes.data)


def _call_rows(fn_name: str, row_dtype: np.dtype, *args) -> np.ndarray:
    lib = _library()
    rows_ptr = ctypes.c_void_p()
    count = ctypes.c_size_t()
    error = ctypes.create_string_buffer(4096)
    status = getattr(lib, fn_name)(*args, ctypes.byref(rows_ptr), ctypes.byref(count), error,
                                   ctypes.sizeof(error))
    if status != 0:
        detail = error.value.decode("utf-8", errors="replace")
        raise RuntimeError(f"{fn_name} failed with status {status}: {detail}")
    n = count.value
    if n == 0:
        lib.rtdl_mi355x_free_rows(rows_ptr)
        return np.zeros(0, dtype=row_dtype)
    # Zero-copy: the array views the library's row buffer, which is freed
    # with rtdl_mi355x_free_rows once the array and every view of it are gone.
    buf = (ctypes.c_char * (n * row_dtype.itemsize)).from_address(rows_ptr.value)
    weakref.finalize(buf, lib.rtdl_mi355x_free_rows, ctypes.c_void_p(rows_ptr.value))
    return np.frombuffer(buf, dtype=row_dtype)


def _call_prepare(fn_name: str, *args) -> ctypes.c_void_p:
    lib = _library()
    handle = ctypes.c_void_p()
    error = ctypes.create_string_buffer(4096)
    status = getattr(lib, fn_name)(*args, ctypes.byref(handle), error, ctypes.sizeof(error))
    if status != 0:
        detail = error.value.decode("utf-8", errors="replace")
        raise RuntimeError(f"{fn_name} failed with status {status}: {detail}")
    return handle


def set_fault(spec: str) -> None:
    """Arms a host-forced fault ("alloc", "launch" or "timeout", optionally
    ":N" for the N-th occurrence from now); "" disarms.  For testing failure
    handling: the library raises the failure on the host, no kernel faults."""
    err = ctypes.create_string_buffer(512)
    if _library().rtdl_mi355x_set_fault(spec.encode(), err, 512) != 0:
        raise ValueError(err.value.decode())


def set_walk_counting(on: bool) -> None:
    """Count node / primitive visits in this thread's per-call traversals."""
    err = ctypes.create_string_buffer(512)
    if _library().rtdl_mi355x_set_walk_counting(int(bool(on)), err, 512) != 0:
        raise RuntimeError(err.value.decode())


def last_walk_stats() -> dict:
    """Traversal kernel seconds of this thread's last per-call traversal and,
    when counting was on, its node and primitive visits and wavefront walk
    steps (the busiest lane's node visits per wavefront, summed)."""
    t = ctypes.c_double()
    v = (ctypes.c_uint64 * 3)()
    err = ctypes.create_string_buffer(512)
    if _library().rtdl_mi355x_get_last_walk_stats(ctypes.byref(t), v, err, 512) != 0:
        raise RuntimeError(err.value.decode())
    return {"kernel_s": t.value, "node_visits": int(v[0]), "prim_tests": int(v[1]), "wave_steps": int(v[2])}


def last_phase_timings() -> dict[str, float]:
    out = (ctypes.c_double * 6)()
    n = _library().rtdl_mi355x_get_last_phase_timings(out, 6)
    names = ("upload", "build", "traverse", "emit", "copy_back", "total")
    return {names[i]: out[i] for i in range(n)}


def mi355x_context_probe() -> dict[str, object]:
    name = ctypes.create_string_buffer(256)
    device_type = ctypes.c_int()
    api_version = ctypes.c_int()
    error = ctypes.create_string_buffer(4096)
    status = _library().rtdl_mi355x_context_probe(
        name, ctypes.sizeof(name), ctypes.byref(device_type), ctypes.byref(api_version), error,
        ctypes.sizeof(error),
    )
    if status != 0:
        raise RuntimeError(error.value.decode("utf-8", errors="replace"))
    return {
        "device_name": name.value.decode(),
        "device_type": device_type.value,
        "api_version": api_version.value,
    }


def selftest_fp64(a: np.ndarray, b: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Device FP64 sqrt(a) and a / b, for rounding checks."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    s = np.empty_like(a)
    d = np.empty_like(a)
    error = ctypes.create_string_buffer(4096)
    status = _library().rtdl_mi355x_selftest_fp64(_ptr(a), _ptr(b), len(a), _ptr(s), _ptr(d), error,
                                                   ctypes.sizeof(error))
    if status != 0:
        raise RuntimeError(error.value.decode("utf-8", errors="replace"))
    return s, d


def mi355x_version() -> tuple[int, int, int]:
    vals = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int()]
    _library().rtdl_mi355x_get_version(*(ctypes.byref(v) for v in vals))
    return tuple(v.value for v in vals)


# --------------------------------------------------------------------------
# packing (records or columnar numpy -> C record arrays)
# --------------------------------------------------------------------------


def _columns(payload, names):
    """Return a dict of numpy columns when `payload` is columnar, else None."""
    if isinstance(payload, np.ndarray) and payload.dtype.names:
        if all(name in payload.dtype.names for name in names):
            return {name: payload[name] for name in names}
        return None
    if isinstance(payload, Mapping) and all(name in payload for name in names):
        cols = {name: np.asarray(payload[name]) for name in names}
        if all(col.ndim == 1 for col in cols.values()):
            return cols
    return None


def _pack(dtype: np.dtype, cols: dict) -> np.ndarray:
    n = len(cols["id"])
    out = np.zeros(n, dtype=dtype)
    for name in dtype.names:
        out[name] = cols[name]
    return out


def pack_points(payload, *, name: str = "points") -> tuple[np.ndarray, int]:
    """Pack point input into RtdlPoint / RtdlPoint3D records; returns (array, dim)."""
    cols = _columns(payload, ("id", "x", "y", "z"))
    if cols is not None:
        return _pack(POINT3D_DTYPE, cols), 3
    cols = _columns(payload, ("id", "x", "y"))
    if cols is not None:
        return _pack(POINT2D_DTYPE, cols), 2
    records = normalize_records(name, "points", payload)
    dims = {3 if isinstance(r, Point3D) else 2 for r in records}
    if len(dims) > 1:
        raise ValueError(f"input `{name}` mixes 2-D and 3-D points")
    dim = dims.pop() if dims else 2
    if dim == 3:
        arr = np.array([(r.id, r.x, r.y, r.z) for r in records],
                       dtype=[("id", "<u4"), ("x", "<f8"), ("y", "<f8"), ("z", "<f8")])
        return _pack(POINT3D_DTYPE, {k: arr[k] for k in ("id", "x", "y", "z")}), 3
    arr = np.array([(r.id, r.x, r.y) for r in records], dtype=[("id", "<u4"), ("x", "<f8"), ("y", "<f8")])
    return _pack(POINT2D_DTYPE, {k: arr[k] for k in ("id", "x", "y")}), 2


def pack_triangles(payload, *, name: str = "triangles") -> np.ndarray:
    cols = _columns(payload, ("id", *_TRI_FIELDS))
    if cols is not None:
        return _pack(TRIANGLE3D_DTYPE, cols)
    records = normalize_records(name, "triangles", payload)
    if any(not isinstance(r, Triangle3D) for r in records):
        raise NotImplementedError(
            "MI355X ray queries support only Triangle3D build inputs; No CPU fallback is used."
        )
    arr = np.array([(r.id, *(getattr(r, f) for f in _TRI_FIELDS)) for r in records],
                   dtype=[("id", "<u4")] + [(f, "<f8") for f in _TRI_FIELDS])
    return _pack(TRIANGLE3D_DTYPE, {k: arr[k] for k in arr.dtype.names})


def pack_rays(payload, *, name: str = "rays") -> np.ndarray:
    cols = _columns(payload, ("id", *_RAY_FIELDS))
    if cols is not None:
        return _pack(RAY3D_DTYPE, cols)
    records = normalize_records(name, "rays", payload)
    if any(not isinstance(r, Ray3D) for r in records):
        raise NotImplementedError(
            "MI355X ray queries support only Ray3D probe inputs; No CPU fallback is used."
        )
    arr = np.array([(r.id, *(getattr(r, f) for f in _RAY_FIELDS)) for r in records],
                   dtype=[("id", "<u4")] + [(f, "<f8") for f in _RAY_FIELDS])
    return _pack(RAY3D_DTYPE, {k: arr[k] for k in arr.dtype.names})


def pack_triangles_2d(payload, *, name: str = "triangles") -> np.ndarray:
    """Pack 2-D triangles into RtdlTriangle records (56 bytes, natural alignment)."""
    cols = _columns(payload, ("id", *_TRI2_FIELDS))
    if cols is not None:
        return _pack(TRIANGLE2D_DTYPE, cols)
    records = normalize_records(name, "triangles", payload)
    if any(not isinstance(r, Triangle) for r in records):
        raise ValueError(f"input `{name}` mixes 2-D and 3-D triangles")
    arr = np.array([(r.id, *(getattr(r, f) for f in _TRI2_FIELDS)) for r in records],
                   dtype=[("id", "<u4")] + [(f, "<f8") for f in _TRI2_FIELDS])
    return _pack(TRIANGLE2D_DTYPE, {k: arr[k] for k in arr.dtype.names})


def pack_rays_2d(payload, *, name: str = "rays") -> np.ndarray:
    """Pack 2-D rays into packed RtdlRay2D records (44 bytes)."""
    cols = _columns(payload, ("id", *_RAY2_FIELDS))
    if cols is not None:
        return _pack(RAY2D_DTYPE, cols)
    records = normalize_records(name, "rays", payload)
    if any(not isinstance(r, Ray2D) for r in records):
        raise ValueError(f"input `{name}` mixes 2-D and 3-D rays")
    arr = np.array([(r.id, *(getattr(r, f) for f in _RAY2_FIELDS)) for r in records],
                   dtype=[("id", "<u4")] + [(f, "<f8") for f in _RAY2_FIELDS])
    return _pack(RAY2D_DTYPE, {k: arr[k] for k in arr.dtype.names})


def _ray_input_dim(payload, geometry: str, name: str) -> int | None:
    """2 for Ray2D / Triangle payloads, 3 for Ray3D / Triangle3D ones, None for
    an empty record sequence."""
    if isinstance(payload, np.ndarray) and payload.dtype.names:
        names = set(payload.dtype.names)
    elif isinstance(payload, Mapping):
        names = set(payload)
    else:
        records = normalize_records(name, geometry, payload)
        if not records:
            return None
        return 2 if isinstance(records[0], (Ray2D, Triangle)) else 3
    return 3 if ("oz" in names or "z0" in names) else 2


def pack_polygons(payload, *, name: str = "polygons") -> tuple[np.ndarray, np.ndarray]:
    """Return (RtdlPolygonRef records, interleaved FP64 vertices)."""
    if isinstance(payload, Mapping) and {"refs", "vertices_xy"} <= set(payload):
        refs = np.asarray(payload["refs"])
        verts = np.ascontiguousarray(payload["vertices_xy"], dtype=np.float64).reshape(-1)
        out = np.zeros(len(refs), dtype=POLYGON_REF_DTYPE)
        for key in POLYGON_REF_DTYPE.names:
            out[key] = refs[key]
        return out, verts
    records = normalize_records(name, "polygons", payload)
    refs = np.zeros(len(records), dtype=POLYGON_REF_DTYPE)
    counts = np.array([len(r.vertices) for r in records], dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]) if len(records) else np.zeros(0, np.int64)
    refs["id"] = [r.id for r in records]
    refs["vertex_offset"] = offsets
    refs["vertex_count"] = counts
    verts = np.array([c for r in records for v in r.vertices for c in v], dtype=np.float64)
    return refs, verts


def pack_probe_points(payload, *, name: str = "points") -> np.ndarray:
    arr, dim = pack_points(payload, name=name)
    if dim != 2:
        raise NotImplementedError(
            "MI355X point_in_polygon supports only Point2D probes; No CPU fallback is used."
        )
    return arr


# --------------------------------------------------------------------------
# row conversion
# --------------------------------------------------------------------------


def rows_to_dicts(rows: np.ndarray) -> tuple[dict[str, object], ...]:
    names = rows.dtype.names
    float_fields = {n for n in names if rows.dtype.fields[n][0].kind == "f"}
    columns = [rows[n].tolist() for n in names]
    out = []
    for values in zip(*columns):
        out.append({n: (float(v) if n in float_fields else int(v)) for n, v in zip(names, values)})
    return tuple(out)


# --------------------------------------------------------------------------
# direct entry points (numpy in, numpy rows out)
# --------------------------------------------------------------------------


def fixed_radius_neighbors_mi355x(query_points, search_points, *, radius: float, k_max: int,
                                  bounded_rank: bool = False) -> np.ndarray:
    queries, qdim = pack_points(query_points, name="query_points")
    search, sdim = pack_points(search_points, name="search_points")
    if len(queries) and len(search) and qdim != sdim:
        raise ValueError("query and search points must have the same dimension")
    dim = sdim if len(search) else qdim
    if len(search) == 0 and dim != qdim:
        dim = qdim
    if bounded_rank:
        name = f"rtdl_mi355x_run_bounded_knn_rows_{dim}d"
        dtype = KNN_ROW_DTYPE
    else:
        name = f"rtdl_mi355x_run_fixed_radius_neighbors_{dim}d"
        dtype = FRN_ROW_DTYPE
    return _call_rows(name, dtype, _ptr(queries), len(queries), _ptr(search), len(search),
                      float(radius), int(k_max))


# The KNN walk keeps each query's k nearest in registers; like the
# reference's HIPRT backend (src/rtdsl/hiprt_runtime.py:1307-1308) it takes
# k <= 64.  Fixed-radius and bounded KNN take larger k_max: their radius
# bounds the candidates gathered beyond 64.
KNN_MAX_K = 64


def _check_knn_k(k: int) -> None:
    if int(k) <= 0:
        raise ValueError("MI355X knn_rows k must be positive")
    if int(k) > KNN_MAX_K:
        raise ValueError(f"MI355X knn_rows currently supports k <= {KNN_MAX_K}")


def knn_rows_mi355x(query_points, search_points, *, k: int) -> np.ndarray:
    _check_knn_k(k)
    queries, qdim = pack_points(query_points, name="query_points")
    search, sdim = pack_points(search_points, name="search_points")
    if len(queries) and len(search) and qdim != sdim:
        raise ValueError("query and search points must have the same dimension")
    dim = sdim if len(search) else qdim
    return _call_rows(f"rtdl_mi355x_run_knn_rows_{dim}d", KNN_ROW_DTYPE, _ptr(queries), len(queries),
                      _ptr(search), len(search), int(k))


_RAY_MODES = {
    "any_hit": ("ray_anyhit", ANYHIT_ROW_DTYPE),
    "hit_count": ("ray_hitcount", HITCOUNT_ROW_DTYPE),
    "closest_hit": ("ray_closest_hit", CLOSEST_ROW_DTYPE),
}


def ray_triangle_mi355x(rays, triangles, *, mode: str = "any_hit", dim: int | None = None) -> np.ndarray:
    """Per-call ray queries; `dim` (2 or 3) is taken from the inputs when not given."""
    rdim = _ray_input_dim(rays, "rays", "rays")
    tdim = _ray_input_dim(triangles, "triangles", "triangles")
    if mode not in _RAY_MODES:
        raise ValueError(f"unknown ray mode {mode!r}")
    fn, dtype = _RAY_MODES[mode]
    dims = {d for d in (rdim, tdim, dim) if d is not None}
    if len(dims) > 1:
        raise ValueError("ray_triangle queries require rays and triangles to both be 2D or both be 3D")
    if dims == {2}:
        r2 = pack_rays_2d(rays)
        t2 = pack_triangles_2d(triangles)
        return _call_rows(f"rtdl_mi355x_run_{fn}_2d", dtype, _ptr(r2), len(r2), _ptr(t2), len(t2))
    r = pack_rays(rays)
    t = pack_triangles(triangles)
    return _call_rows(f"rtdl_mi355x_run_{fn}_3d", dtype, _ptr(r), len(r), _ptr(t), len(t))


def point_in_polygon_mi355x(points, polygons, *, positive_only: bool) -> np.ndarray:
    pts = pack_probe_points(points)
    refs, verts = pack_polygons(polygons)
    return _call_rows("rtdl_mi355x_run_point_in_polygon", PIP_ROW_DTYPE, _ptr(pts), len(pts),
                      _ptr(refs), len(refs), _ptr(verts), len(verts), int(bool(positive_only)))


# --------------------------------------------------------------------------
# kernel validation (reference: src/rtdsl/hiprt_runtime.py:2433-2630)
# --------------------------------------------------------------------------


def _resolve_kernel(kernel_fn_or_compiled) -> CompiledKernel:
    if isinstance(kernel_fn_or_compiled, CompiledKernel):
        return kernel_fn_or_compiled
    return compile_kernel(kernel_fn_or_compiled)


def _unsupported(predicate: str, detail: str) -> NotImplementedError:
    return NotImplementedError(
        f"MI355X backend recognizes `{predicate}`, but {detail}. No CPU fallback is used."
    )


def _validate_kernel(compiled: CompiledKernel) -> tuple[str, tuple[str, str], dict]:
    if compiled.precision != "float_approx":
        raise ValueError("RTDL MI355X backend currently supports only precision='float_approx'")
    if compiled.candidates is None or compiled.refine_op is None or compiled.emit_op is None:
        raise ValueError("RTDL MI355X backend requires a fully compiled kernel with traverse/refine/emit")
    predicate = compiled.refine_op.predicate.name
    if predicate not in _IMPLEMENTED:
        raise ValueError(f"unsupported predicate for MI355X backend: {predicate!r}")
    left = compiled.candidates.left
    right = compiled.candidates.right
    options = dict(compiled.refine_op.predicate.options)
    if predicate in {"fixed_radius_neighbors", "knn_rows", "bounded_knn_rows"}:
        if left.geometry.name != "points" or right.geometry.name != "points":
            raise _unsupported(predicate, "only Point2D/Point3D inputs are implemented")
        if left.role != "probe" or right.role != "build":
            raise ValueError(
                f"MI355X {predicate} requires query points role='probe' and search points role='build'"
            )
        pair = (left.layout.name, right.layout.name)
        if pair not in {("Point2D", "Point2D"), ("Point3D", "Point3D")}:
            raise _unsupported(predicate, "only matching Point2DLayout or Point3DLayout inputs are implemented")
        options["layout_pair"] = pair
        if predicate == "knn_rows":
            _check_knn_k(options["k"])
        return predicate, (left.name, right.name), options
    if predicate == "point_in_polygon":
        if left.geometry.name != "points" or right.geometry.name != "polygons":
            raise ValueError("MI355X point_in_polygon requires point probe input followed by polygon build input")
        if left.role != "probe" or right.role != "build":
            raise ValueError("MI355X point_in_polygon requires points role='probe' and polygons role='build'")
        if left.layout.name != "Point2D" or right.layout.name != "Polygon2DRef":
            raise _unsupported(predicate, "only Point2DLayout and Polygon2DRef inputs are implemented")
        if options.get("boundary_mode", "inclusive") != "inclusive":
            raise ValueError("MI355X point_in_polygon currently supports only boundary_mode='inclusive'")
        return predicate, (left.name, right.name), options
    # ray family
    if left.geometry.name != "rays" or right.geometry.name != "triangles":
        raise ValueError(f"MI355X {predicate} requires ray probe input followed by triangle build input")
    if left.role != "probe" or right.role != "build":
        raise ValueError(f"MI355X {predicate} requires rays role='probe' and triangles role='build'")
    pair = (left.layout.name, right.layout.name)
    if pair not in {("Ray3D", "Triangle3D"), ("Ray2D", "Triangle2D")}:
        raise _unsupported(predicate, "only matching Ray3D/Triangle3D or Ray2D/Triangle2D layouts are implemented")
    options["layout_pair"] = pair
    return predicate, (left.name, right.name), options


def _check_inputs(compiled: CompiledKernel, inputs: dict, *, what: str) -> None:
    expected = {item.name for item in compiled.inputs}
    missing = [name for name in expected if name not in inputs]
    unexpected = [name for name in inputs if name not in expected]
    if missing:
        raise ValueError(f"missing RTDL MI355X {what}: {', '.join(sorted(missing))}")
    if unexpected:
        raise ValueError(f"unexpected RTDL MI355X {what}: {', '.join(sorted(unexpected))}")


def _finish(compiled: CompiledKernel, rows: np.ndarray, result_mode: str):
    if result_mode == "raw":
        return rows
    return project_rows(compiled, rows_to_dicts(rows))


def run_mi355x(kernel_fn_or_compiled, *, result_mode: str = "dict", **inputs):
    """Plain per-call execution: upload, build, traverse, emit, copy back."""
    if result_mode not in {"dict", "raw"}:
        raise ValueError("MI355X result_mode supports only 'dict' or 'raw'")
    compiled = _resolve_kernel(kernel_fn_or_compiled)
    predicate, (probe, build), options = _validate_kernel(compiled)
    _check_inputs(compiled, inputs, what="inputs")
    if predicate == "fixed_radius_neighbors":
        rows = fixed_radius_neighbors_mi355x(inputs[probe], inputs[build], radius=float(options["radius"]),
                                             k_max=int(options["k_max"]))
    elif predicate == "bounded_knn_rows":
        rows = fixed_radius_neighbors_mi355x(inputs[probe], inputs[build], radius=float(options["radius"]),
                                             k_max=int(options["k_max"]), bounded_rank=True)
    elif predicate == "knn_rows":
        rows = knn_rows_mi355x(inputs[probe], inputs[build], k=int(options["k"]))
    elif predicate == "point_in_polygon":
        rows = point_in_polygon_mi355x(
            inputs[probe], inputs[build],
            positive_only=options.get("result_mode", "full_matrix") == "positive_hits",
        )
    else:
        mode = {"ray_triangle_any_hit": "any_hit", "ray_triangle_hit_count": "hit_count",
                "ray_triangle_closest_hit": "closest_hit"}[predicate]
        rows = ray_triangle_mi355x(inputs[probe], inputs[build], mode=mode,
                                   dim=2 if options["layout_pair"][0] == "Ray2D" else 3)
    return _finish(compiled, rows, result_mode)


# --------------------------------------------------------------------------
# prepared (scene-resident) execution
# --------------------------------------------------------------------------


class PreparedMi355xKernel:
    """A device-resident build side; `.run(**probe_inputs)` per query batch."""

    def __init__(self, compiled: CompiledKernel, predicate: str, probe_name: str, build_name: str,
                 options: dict, handle: ctypes.c_void_p, dim: int, result_mode: str) -> None:
        self.compiled = compiled
        self.predicate = predicate
        self.probe_name = probe_name
        self.build_name = build_name
        self.options = options
        self.dim = dim
        self.result_mode = result_mode
        self._handle = handle

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._handle is None:
            raise RuntimeError("prepared MI355X kernel is closed")
        return self._handle

    def close(self) -> None:
        if self._handle is not None:
            _library().rtdl_mi355x_destroy_prepared(self._handle)
            self._handle = None

    def __enter__(self) -> "PreparedMi355xKernel":
        return self

    def __exit__(self, exc_type, exc, tb) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def run_raw(self, **inputs) -> np.ndarray:
        missing = [self.probe_name] if self.probe_name not in inputs else []
        unexpected = [name for name in inputs if name != self.probe_name]
        if missing:
            raise ValueError(f"missing RTDL MI355X prepared query inputs: {', '.join(missing)}")
        if unexpected:
            raise ValueError(
                f"unexpected RTDL MI355X prepared query inputs: {', '.join(sorted(unexpected))}"
            )
        payload = inputs[self.probe_name]
        p = self.predicate
        opts = self.options
        if p in {"fixed_radius_neighbors", "knn_rows", "bounded_knn_rows"}:
            queries, dim = pack_points(payload, name=self.probe_name)
            if len(queries) and dim != self.dim:
                raise ValueError("query points must match the prepared search point dimension")
            if p == "fixed_radius_neighbors":
                return _call_rows("rtdl_mi355x_run_prepared_fixed_radius_neighbors", FRN_ROW_DTYPE,
                                  self.handle, _ptr(queries), len(queries), float(opts["radius"]),
                                  int(opts["k_max"]))
            if p == "bounded_knn_rows":
                return _call_rows("rtdl_mi355x_run_prepared_bounded_knn_rows", KNN_ROW_DTYPE,
                                  self.handle, _ptr(queries), len(queries), float(opts["radius"]),
                                  int(opts["k_max"]))
            _check_knn_k(opts["k"])
            return _call_rows("rtdl_mi355x_run_prepared_knn_rows", KNN_ROW_DTYPE, self.handle,
                              _ptr(queries), len(queries), int(opts["k"]))
        if p == "point_in_polygon":
            pts = pack_probe_points(payload, name=self.probe_name)
            positive = opts.get("result_mode", "full_matrix") == "positive_hits"
            return _call_rows("rtdl_mi355x_run_prepared_point_in_polygon", PIP_ROW_DTYPE, self.handle,
                              _ptr(pts), len(pts), int(positive))
        fn, dtype = _RAY_MODES[{"ray_triangle_any_hit": "any_hit", "ray_triangle_hit_count": "hit_count",
                                "ray_triangle_closest_hit": "closest_hit"}[p]]
        rays = (pack_rays_2d if self.dim == 2 else pack_rays)(payload, name=self.probe_name)
        return _call_rows(f"rtdl_mi355x_run_prepared_{fn}_{self.dim}d", dtype, self.handle, _ptr(rays), len(rays))

    def run(self, **inputs):
        return _finish(self.compiled, self.run_raw(**inputs), self.result_mode)


def prepare_mi355x(kernel_fn_or_compiled, *, result_mode: str = "dict", **inputs) -> PreparedMi355xKernel:
    """Build the kernel's build-side input into a device-resident LBVH scene."""
    if result_mode not in {"dict", "raw"}:
        raise ValueError("MI355X result_mode supports only 'dict' or 'raw'")
    compiled = _resolve_kernel(kernel_fn_or_compiled)
    predicate, (probe, build), options = _validate_kernel(compiled)
    missing = [build] if build not in inputs else []
    unexpected = [name for name in inputs if name != build]
    if missing:
        raise ValueError(f"missing RTDL MI355X prepare inputs: {', '.join(missing)}")
    if unexpected:
        raise ValueError(
            "unexpected RTDL MI355X prepare inputs: " + ", ".join(sorted(unexpected))
            + f"; pass query inputs later to prepared.run({probe}=...)"
        )
    payload = inputs[build]
    dim = 3
    if predicate in {"fixed_radius_neighbors", "knn_rows", "bounded_knn_rows"}:
        search, dim = pack_points(payload, name=build)
        if len(search) == 0:
            dim = 3 if options["layout_pair"][1] == "Point3D" else 2
        handle = _call_prepare(f"rtdl_mi355x_prepare_points_{dim}d", _ptr(search), len(search))
    elif predicate == "point_in_polygon":
        refs, verts = pack_polygons(payload, name=build)
        handle = _call_prepare("rtdl_mi355x_prepare_polygons", _ptr(refs), len(refs), _ptr(verts), len(verts))
        dim = 2
    else:
        if options["layout_pair"] == ("Ray2D", "Triangle2D"):
            tris2 = pack_triangles_2d(payload, name=build)
            handle = _call_prepare("rtdl_mi355x_prepare_triangles_2d", _ptr(tris2), len(tris2))
            dim = 2
        else:
            tris = pack_triangles(payload, name=build)
            handle = _call_prepare("rtdl_mi355x_prepare_triangles_3d", _ptr(tris), len(tris))
    return PreparedMi355xKernel(compiled, predicate, probe, build, options, handle, dim, result_mode)


def mi355x_available() -> bool:
    try:
        mi355x_context_probe()
        return True
    except Exception:
        return False


__all__ = [
    "PreparedMi355xKernel",
    "fixed_radius_neighbors_mi355x",
    "knn_rows_mi355x",
    "last_phase_timings",
    "mi355x_available",
    "mi355x_context_probe",
    "mi355x_version",
    "pack_points",
    "pack_polygons",
    "pack_rays",
    "pack_rays_2d",
    "pack_triangles",
    "pack_triangles_2d",
    "point_in_polygon_mi355x",
    "prepare_mi355x",
    "ray_triangle_mi355x",
    "run_mi355x",
]
