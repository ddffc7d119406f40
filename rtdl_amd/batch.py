"""Batched and served execution over device-resident inputs.

``run_batch`` runs many independent jobs (query sets against resident
scenes or instances of a multi-instance scene) in one persistent launch per
predicate family; ``serve_replay`` replays a timed request stream against
resident scenes with continuous batching and memory-budget admission.  Both
sit on the C-ABI (``rtdl_mi355x_run_batch`` / ``rtdl_mi355x_serve_replay``).
Device inputs and row buffers are ``rtdl_amd.device.DeviceBuffer`` objects
(allocated by the library itself) or any object exposing ``data_ptr()``,
``shape`` and ``numel()`` for device memory; every computation runs in the
library's HIP kernels.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import mi355x_runtime as rt_native
from .device import DeviceBuffer

JOB_FIXED_RADIUS = 0
JOB_KNN = 1
JOB_BOUNDED_KNN = 2
JOB_RAY_ANYHIT = 3
JOB_RAY_HITCOUNT = 4
JOB_RAY_CLOSEST = 5
JOB_PIP_POSITIVE = 6

STATUS_ROW_OVERFLOW = 1
STATUS_REJECTED = 4
STATUS_INTERNAL = 8
STATUS_FAILED = 16
STATUS_TIMEOUT = 32
STATUS_INVALID = 64
PIP_HIT_CAP = 16  # default PIP row budget per point (average); one point may exceed it

ROW_DTYPES = {
    JOB_FIXED_RADIUS: rt_native.FRN_ROW_DTYPE,
    JOB_KNN: rt_native.KNN_ROW_DTYPE,
    JOB_BOUNDED_KNN: rt_native.KNN_ROW_DTYPE,
    JOB_RAY_ANYHIT: rt_native.ANYHIT_ROW_DTYPE,
    JOB_RAY_HITCOUNT: rt_native.HITCOUNT_ROW_DTYPE,
    JOB_RAY_CLOSEST: rt_native.CLOSEST_ROW_DTYPE,
    JOB_PIP_POSITIVE: rt_native.PIP_ROW_DTYPE,
}
QUERY_STRIDE = {JOB_FIXED_RADIUS: 3, JOB_KNN: 3, JOB_BOUNDED_KNN: 3, JOB_RAY_ANYHIT: 7,
                JOB_RAY_HITCOUNT: 7, JOB_RAY_CLOSEST: 7, JOB_PIP_POSITIVE: 2}


class _BatchJob(ctypes.Structure):
    _fields_ = [
        ("scene", ctypes.c_void_p), ("instance", ctypes.c_uint32), ("kind", ctypes.c_uint32),
        ("queries", ctypes.c_void_p), ("query_count", ctypes.c_uint32),
        ("query_stride", ctypes.c_uint32), ("radius", ctypes.c_double), ("k", ctypes.c_uint32),
        ("pad", ctypes.c_uint32), ("rows", ctypes.c_void_p), ("row_capacity", ctypes.c_uint64),
    ]


class _BatchJobResult(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_uint64), ("status", ctypes.c_uint32), ("ambiguous", ctypes.c_uint32)]


class _ServeConfig(ctypes.Structure):
    _fields_ = [
        ("radius", ctypes.c_double), ("k_max", ctypes.c_uint32), ("k", ctypes.c_uint32),
        ("slots", ctypes.c_uint32), ("morton_sort", ctypes.c_uint32),
        ("memory_budget_bytes", ctypes.c_uint64), ("max_batch_queries", ctypes.c_uint64),
        ("point_queries", ctypes.c_void_p), ("ray_queries", ctypes.c_void_p),
        ("pip_queries", ctypes.c_void_p), ("realtime", ctypes.c_uint32), ("timeout_ms", ctypes.c_uint32),
    ]


class _ServeStats(ctypes.Structure):
    _fields_ = [
        ("wall_s", ctypes.c_double), ("batches", ctypes.c_uint64), ("queries", ctypes.c_uint64),
        ("rows", ctypes.c_uint64), ("peak_inflight_bytes", ctypes.c_uint64),
        ("region_bytes", ctypes.c_uint64), ("retried", ctypes.c_uint64), ("failed", ctypes.c_uint64),
    ]


SERVE_REQUEST_DTYPE = np.dtype({
    "names": ["kind", "scene", "query_offset", "query_count", "capture", "arrival_s"],
    "formats": ["<u4", "<u4", "<u8", "<u4", "<u4", "<f8"],
    "offsets": [0, 4, 8, 16, 20, 24], "itemsize": 32,
})
SERVE_RESULT_DTYPE = np.dtype({
    "names": ["arrival_s", "dispatch_s", "done_s", "rows", "status", "batch", "captured_rows"],
    "formats": ["<f8", "<f8", "<f8", "<u8", "<u4", "<u4", "<u8"],
    "offsets": [0, 8, 16, 24, 32, 36, 40], "itemsize": 48,
})

def _lib():
    return rt_native._library()


def _check(status: int, error, what: str) -> None:
    if status != 0:
        raise RuntimeError(f"{what} failed with status {status}: "
                           + error.value.decode("utf-8", errors="replace"))


def _ptr(t) -> int | None:
    if t is None:
        return None
    return int(t.data_ptr()) if t.numel() else None


def _sync_foreign() -> None:
    """Device arrays made by torch are written on torch's streams; when torch
    is in use in this process, its work is completed before a native call.
    (torch is never imported here: buffers from the library need no sync.)"""
    import sys

    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        torch.cuda.synchronize()


class Instances:
    """A multi-instance scene: one LBVH with one subtree per instance."""

    def __init__(self, handle: ctypes.c_void_p, instance_offsets: np.ndarray, kind: str):
        self._handle = handle
        self.offsets = np.asarray(instance_offsets, dtype=np.uint64)
        self.kind = kind

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._handle is None:
            raise RuntimeError("instances scene is closed")
        return self._handle

    def close(self) -> None:
        if self._handle is not None:
            rt_native._library().rtdl_mi355x_destroy_prepared(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prepare_point_instances(coords, instance_offsets, *, leaf_size: int = 0) -> Instances:
    """coords: float32 device tensor (n, 3) (z = 0 for 2-D instances).
    `leaf_size`: points per leaf of this scene's trees (1-8), 0 for the
    process tuning (autotune.Tuning.point_leaf)."""
    offs = np.ascontiguousarray(instance_offsets, dtype=np.uint64)
    assert coords.shape[-1] == 3
    _sync_foreign()
    handle = ctypes.c_void_p()
    err = ctypes.create_string_buffer(4096)
    st = _lib().rtdl_mi355x_prepare_point_instances_leaf(
        _ptr(coords), int(coords.shape[0]), offs.ctypes.data, len(offs) - 1, int(leaf_size),
        ctypes.byref(handle), err, 4096)
    _check(st, err, "rtdl_mi355x_prepare_point_instances_leaf")
    return Instances(handle, offs, "points")


def rebuild_instances(inst: Instances, coords) -> None:
    """Rebuilds a point or triangle instance scene in place from new FP32
    device coordinates of the same primitives (same shape and instance
    offsets as at prepare time); ids, handles and jobs stay valid."""
    per = 3 if inst.kind == "points" else 9
    assert coords.shape[-1] == per and int(coords.shape[0]) == int(inst.offsets[-1]), "coords shape"
    err = ctypes.create_string_buffer(4096)
    _sync_foreign()
    st = _lib().rtdl_mi355x_rebuild_instances(inst.handle, _ptr(coords), err, 4096)
    _check(st, err, "rtdl_mi355x_rebuild_instances")


def prepare_triangle_instances(coords, instance_offsets) -> Instances:
    """coords: float32 device tensor (n, 9): three vertices per triangle."""
    offs = np.ascontiguousarray(instance_offsets, dtype=np.uint64)
    assert coords.shape[-1] == 9
    _sync_foreign()
    handle = ctypes.c_void_p()
    err = ctypes.create_string_buffer(4096)
    st = _lib().rtdl_mi355x_prepare_triangle_instances(
        _ptr(coords), int(coords.shape[0]), offs.ctypes.data, len(offs) - 1, ctypes.byref(handle), err, 4096)
    _check(st, err, "rtdl_mi355x_prepare_triangle_instances")
    return Instances(handle, offs, "triangles")


def prepare_polygon_instances(vertex_offsets, vertex_counts, vertices_xy, instance_offsets) -> Instances:
    """vertex_offsets / vertex_counts: int32 device arrays; vertices_xy: float32 (nv, 2)."""
    offs = np.ascontiguousarray(instance_offsets, dtype=np.uint64)
    _sync_foreign()
    handle = ctypes.c_void_p()
    err = ctypes.create_string_buffer(4096)
    st = _lib().rtdl_mi355x_prepare_polygon_instances(
        _ptr(vertex_offsets), _ptr(vertex_counts), int(vertex_offsets.shape[0]), _ptr(vertices_xy),
        int(vertices_xy.shape[0]), offs.ctypes.data, len(offs) - 1, ctypes.byref(handle), err, 4096)
    _check(st, err, "rtdl_mi355x_prepare_polygon_instances")
    return Instances(handle, offs, "polygons")


@dataclass
class BatchJob:
    scene: object          # PreparedMi355xKernel, Instances or a raw handle
    kind: int
    queries: object        # float32 device tensor (n, stride)
    instance: int = 0
    radius: float = 0.0
    k: int = 1
    rows: object = None    # uint8 device tensor; allocated when None
    row_capacity: int = 0


def _handle_of(scene):
    if hasattr(scene, "handle"):
        return scene.handle
    return scene


def _raw(handle) -> int:
    return handle.value if isinstance(handle, ctypes.c_void_p) else int(handle)


def row_capacity(kind: int, nq: int, k: int) -> int:
    if kind in (JOB_FIXED_RADIUS, JOB_KNN, JOB_BOUNDED_KNN):
        return nq * k
    if kind == JOB_PIP_POSITIVE:
        return nq * PIP_HIT_CAP
    return nq


def _pack_jobs(jobs: list[BatchJob]):
    arr = (_BatchJob * max(1, len(jobs)))()
    for i, j in enumerate(jobs):
        nq = int(j.queries.shape[0])
        cap = j.row_capacity or row_capacity(j.kind, nq, j.k)
        if j.rows is None:
            j.rows = DeviceBuffer((max(1, cap * ROW_DTYPES[j.kind].itemsize),), np.uint8)
            j.row_capacity = cap
        arr[i].scene = _raw(_handle_of(j.scene))
        arr[i].instance = j.instance
        arr[i].kind = j.kind
        arr[i].queries = _ptr(j.queries)
        arr[i].query_count = nq
        arr[i].query_stride = int(j.queries.shape[1])
        arr[i].radius = float(j.radius)
        arr[i].k = int(j.k)
        arr[i].rows = _ptr(j.rows)
        arr[i].row_capacity = int(j.row_capacity or cap)
    return arr, (_BatchJobResult * max(1, len(jobs)))()


def run_batch(jobs: list[BatchJob], *, morton_sort: bool = True):
    """Runs the jobs as one batch; returns (per-job result dicts, device seconds)."""
    arr, res = _pack_jobs(jobs)
    seconds = ctypes.c_double()
    err = ctypes.create_string_buffer(4096)
    _sync_foreign()
    st = _lib().rtdl_mi355x_run_batch(ctypes.addressof(arr), len(jobs), int(bool(morton_sort)),
                                      ctypes.addressof(res), ctypes.byref(seconds), err, 4096)
    _check(st, err, "rtdl_mi355x_run_batch")
    out = [{"rows": int(res[i].rows), "status": int(res[i].status), "ambiguous": int(res[i].ambiguous)}
           for i in range(len(jobs))]
    return out, seconds.value


NOALLOC_MAX_JOBS = 128
NOALLOC_MAX_QUERIES = 16384


def run_batch_noalloc(jobs: list[BatchJob], *, morton_sort: bool = True):
    """run_batch through the allocation-free entry point: the library
    allocates nothing for the batch (its scratch is a static device array).
    Row buffers must already exist (pass `rows`, or run the jobs through
    run_batch once, which allocates them).  At most NOALLOC_MAX_JOBS jobs and
    NOALLOC_MAX_QUERIES queries.  Returns (per-job result dicts, wall seconds)."""
    if any(j.rows is None for j in jobs):
        raise ValueError("run_batch_noalloc needs the jobs' row buffers allocated beforehand")
    arr, res = _pack_jobs(jobs)
    seconds = ctypes.c_double()
    err = ctypes.create_string_buffer(4096)
    _sync_foreign()
    st = _lib().rtdl_mi355x_run_batch_noalloc(ctypes.addressof(arr), len(jobs), int(bool(morton_sort)),
                                              ctypes.addressof(res), ctypes.byref(seconds), err, 4096)
    _check(st, err, "rtdl_mi355x_run_batch_noalloc")
    out = [{"rows": int(res[i].rows), "status": int(res[i].status), "ambiguous": int(res[i].ambiguous)}
           for i in range(len(jobs))]
    return out, seconds.value


class BatchGraph:
    """A batch recorded once into a hipGraph and replayed by `launch()`.

    The jobs' scenes and device buffers are read at every launch and must
    stay alive while the graph does (the graph keeps references to the
    jobs); their contents may change between launches.  Rows, status and
    counts of every launch are those `run_batch` would give for the same
    jobs and contents."""

    def __init__(self, jobs: list[BatchJob], *, morton_sort: bool = True, rebuild=()):
        """`rebuild`: (Instances, coords) pairs recorded in front of the
        batch: every launch rebuilds each point / triangle instance scene
        from the current contents of its FP32 device coordinate buffer (same
        shape as at prepare time), then runs the batch."""
        self._jobs = list(jobs)
        self._rebuild = [(inst, coords) for inst, coords in rebuild]
        self._arr, self._res = _pack_jobs(self._jobs)
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(4096)
        _sync_foreign()
        if self._rebuild:
            n = len(self._rebuild)
            scenes = (ctypes.c_void_p * n)(*[inst.handle.value for inst, _ in self._rebuild])
            ptrs = (ctypes.c_void_p * n)(*[_ptr(c) for _, c in self._rebuild])
            for inst, c in self._rebuild:
                per = 3 if inst.kind == "points" else 9
                assert c.shape[-1] == per and int(c.shape[0]) == int(inst.offsets[-1]), "coords shape"
            st = _lib().rtdl_mi355x_batch_graph_create_rebuild(
                ctypes.addressof(self._arr), len(self._jobs), int(bool(morton_sort)), scenes, ptrs, n,
                ctypes.byref(handle), err, 4096)
            _check(st, err, "rtdl_mi355x_batch_graph_create_rebuild")
        else:
            st = _lib().rtdl_mi355x_batch_graph_create(ctypes.addressof(self._arr), len(self._jobs),
                                                       int(bool(morton_sort)), ctypes.byref(handle), err, 4096)
            _check(st, err, "rtdl_mi355x_batch_graph_create")
        self._handle = handle

    def launch(self):
        """Replays the batch; returns (per-job result dicts, device seconds)."""
        if self._handle is None:
            raise RuntimeError("BatchGraph is closed")
        seconds = ctypes.c_double()
        err = ctypes.create_string_buffer(4096)
        _sync_foreign()
        st = _lib().rtdl_mi355x_batch_graph_launch(self._handle, ctypes.addressof(self._res),
                                                   ctypes.byref(seconds), err, 4096)
        _check(st, err, "rtdl_mi355x_batch_graph_launch")
        res = self._res
        out = [{"rows": int(res[i].rows), "status": int(res[i].status), "ambiguous": int(res[i].ambiguous)}
               for i in range(len(self._jobs))]
        return out, seconds.value

    def close(self):
        if self._handle is not None:
            err = ctypes.create_string_buffer(4096)
            st = _lib().rtdl_mi355x_batch_graph_destroy(self._handle, err, 4096)
            self._handle = None
            _check(st, err, "rtdl_mi355x_batch_graph_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception as exc:  # noqa: BLE001 - interpreter shutdown or a failed destroy
            print(f"BatchGraph: {exc}")


FAMILIES = ("neighbors", "ray_any_hit", "ray_hit_count", "ray_closest_hit", "pip")


def run_batch_profiled(jobs: list[BatchJob], *, morton_sort: bool = True, count_visits: bool = False):
    """Like run_batch, returning per-family kernel seconds and, when
    `count_visits`, the (node records, primitive records, wavefront walk
    steps) visited; node records / (64 * walk steps) is the walks' lane
    efficiency."""
    arr, res = _pack_jobs(jobs)
    fam = (ctypes.c_double * 5)()
    visits = (ctypes.c_uint64 * 3)()
    err = ctypes.create_string_buffer(4096)
    _sync_foreign()
    st = _lib().rtdl_mi355x_run_batch_profiled(ctypes.addressof(arr), len(jobs), int(bool(morton_sort)),
                                               ctypes.addressof(res), fam,
                                               visits if count_visits else None, err, 4096)
    _check(st, err, "rtdl_mi355x_run_batch_profiled")
    out = [{"rows": int(res[i].rows), "status": int(res[i].status), "ambiguous": int(res[i].ambiguous)}
           for i in range(len(jobs))]
    seconds = {FAMILIES[f]: fam[f] for f in range(5)}
    return out, seconds, (tuple(int(v) for v in visits) if count_visits else None)


def job_rows(job: BatchJob, result: dict) -> np.ndarray:
    """Copies a job's rows to the host as a numpy structured array."""
    dtype = ROW_DTYPES[job.kind]
    n = min(int(result["rows"]), int(job.row_capacity))
    if hasattr(job.rows, "download"):  # DeviceBuffer or DeviceView
        raw = job.rows.download(n * dtype.itemsize)
    else:
        raw = job.rows[: n * dtype.itemsize].cpu().numpy()
    return raw.view(dtype).copy()


def serve_replay(scenes, requests: np.ndarray, *, radius: float, k_max: int, k: int,
                 point_queries=None, ray_queries=None, pip_queries=None, slots: int = 4,
                 morton_sort: bool = True, memory_budget_bytes: int = 200 << 30,
                 max_batch_queries: int = 4 << 20, realtime: bool = True, timeout_ms: int = 0):
    """Replays `requests` (SERVE_REQUEST_DTYPE) against `scenes`.

    Returns (results as SERVE_RESULT_DTYPE, stats dict, captured rows by request index)."""
    reqs = np.ascontiguousarray(requests, dtype=SERVE_REQUEST_DTYPE)
    handles = (ctypes.c_void_p * max(1, len(scenes)))(*[_raw(_handle_of(s)) for s in scenes])
    cfg = _ServeConfig()
    cfg.radius = float(radius)
    cfg.k_max = int(k_max)
    cfg.k = int(k)
    cfg.slots = int(slots)
    cfg.morton_sort = int(bool(morton_sort))
    cfg.memory_budget_bytes = int(memory_budget_bytes)
    cfg.max_batch_queries = int(max_batch_queries)
    cfg.point_queries = _ptr(point_queries)
    cfg.ray_queries = _ptr(ray_queries)
    cfg.pip_queries = _ptr(pip_queries)
    cfg.realtime = int(bool(realtime))
    cfg.timeout_ms = int(timeout_ms)
    results = np.zeros(len(reqs), dtype=SERVE_RESULT_DTYPE)
    stats = _ServeStats()
    err = ctypes.create_string_buffer(4096)
    _sync_foreign()
    st = _lib().rtdl_mi355x_serve_replay(ctypes.addressof(handles), len(scenes), reqs.ctypes.data,
                                         len(reqs), ctypes.addressof(cfg), results.ctypes.data,
                                         ctypes.addressof(stats), err, 4096)
    _check(st, err, "rtdl_mi355x_serve_replay")
    captured = {}
    lib = rt_native._library()
    for i in np.flatnonzero(results["captured_rows"]):
        dtype = ROW_DTYPES[int(reqs["kind"][i])]
        n = int(results["rows"][i])
        addr = int(results["captured_rows"][i])
        buf = (ctypes.c_char * (n * dtype.itemsize)).from_address(addr) if n else b""
        captured[int(i)] = np.frombuffer(buf, dtype=dtype).copy() if n else np.zeros(0, dtype)
        lib.rtdl_mi355x_free_rows(ctypes.c_void_p(addr))
        results["captured_rows"][i] = 0
    stats_d = {name: getattr(stats, name) for name, _ in _ServeStats._fields_}
    return results, stats_d, captured
