"""Device-resident row seam: DLPack export and paged row streams.

Batched and served rows stay in device row buffers (``DeviceBuffer`` /
``DeviceView``).  This module hands them to other GPU frameworks without a
host copy, and to host code a page at a time:

* ``to_dlpack(buf)`` / ``DeviceBuffer.__dlpack__``: a DLPack capsule
  ("dltensor") over the buffer's memory, device type ROCm, so that for
  instance ``torch.from_dlpack(view)`` wraps the rows in place.  Row structs
  are exported as ``uint8`` tensors of shape (rows, row bytes); fields are
  reinterpreted by the consumer.  The capsule keeps the buffer alive until
  the consumer releases it.  The library's batch calls return only once their
  rows are written, so the data is ready on any consumer stream.
* ``device_rows(job, result)``: the job's rows as a (rows, row bytes) uint8
  device view, ready for ``__dlpack__``.
* ``row_pages(job, result, page_rows)``: the rows as consecutive pages of at
  most ``page_rows`` rows, each copied to the host when it is reached (a
  stream over a row set larger than host staging), with the page's offset,
  capacity, a continuation token and the job's overflow flag, after the
  segmented row pages of the reference
  (src/rtdsl/segmented_row_stream.py:34-120).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Iterator

import numpy as np

from . import mi355x_runtime as rt_native

# ---------------------------------------------------------------- DLPack --

KDL_ROCM = 10
_KDL_CODE = {"i": 0, "u": 1, "f": 2}


class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DELETER = ctypes.CFUNCTYPE(None, ctypes.POINTER(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p), ("deleter", _DELETER)]

# exports alive until their consumer (or the unconsumed capsule) releases them
_LIVE: dict[int, tuple] = {}


def _release(ptr) -> None:
    _LIVE.pop(ctypes.addressof(ptr.contents), None)


_DELETER_FN = _DELETER(_release)

_PyCapsule_New = ctypes.pythonapi.PyCapsule_New
_PyCapsule_New.restype = ctypes.py_object
_PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
# The capsule destructor runs while the capsule is being deallocated: it
# must not take a reference to it, so it sees the raw pointer and calls the
# capsule API through function pointers typed with c_void_p (item access
# gives fresh function objects, apart from the attribute-cached ones).
_IsValidRaw = ctypes.pythonapi["PyCapsule_IsValid"]
_IsValidRaw.restype = ctypes.c_int
_IsValidRaw.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
_GetPointerRaw = ctypes.pythonapi["PyCapsule_GetPointer"]
_GetPointerRaw.restype = ctypes.c_void_p
_GetPointerRaw.argtypes = [ctypes.c_void_p, ctypes.c_char_p]

_CAPSULE_DESTRUCTOR = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


@_CAPSULE_DESTRUCTOR
def _capsule_destructor(capsule) -> None:
    # a capsule still named "dltensor" was never consumed: release it here
    # (a consumer renames it to "used_dltensor" and calls the deleter itself)
    if _IsValidRaw(capsule, b"dltensor"):
        ptr = _GetPointerRaw(capsule, b"dltensor")
        managed = ctypes.cast(ptr, ctypes.POINTER(_DLManagedTensor))
        managed.contents.deleter(managed)


def current_device() -> int:
    dev = ctypes.c_int()
    err = ctypes.create_string_buffer(512)
    if rt_native._library().rtdl_mi355x_get_device(ctypes.byref(dev), err, 512) != 0:
        raise RuntimeError(err.value.decode())
    return dev.value


def dlpack_device(buf) -> tuple[int, int]:
    return (KDL_ROCM, getattr(buf, "device", 0))


def to_dlpack(buf):
    """A DLPack capsule over a DeviceBuffer / DeviceView (no copy)."""
    dtype = np.dtype(buf.dtype)
    if dtype.kind not in _KDL_CODE or dtype.fields:
        raise TypeError(f"DLPack export of {dtype} (structured rows are exported as uint8; see device_rows)")
    shape = (ctypes.c_int64 * len(buf.shape))(*buf.shape)
    managed = _DLManagedTensor()
    t = managed.dl_tensor
    t.data = ctypes.c_void_p(buf.data_ptr())
    t.device = _DLDevice(KDL_ROCM, getattr(buf, "device", 0))
    t.ndim = len(buf.shape)
    t.dtype = _DLDataType(_KDL_CODE[dtype.kind], dtype.itemsize * 8, 1)
    t.shape = shape
    t.strides = None  # compact, row-major
    t.byte_offset = 0
    managed.manager_ctx = None
    managed.deleter = _DELETER_FN
    _LIVE[ctypes.addressof(managed)] = (managed, shape, buf)  # keeps the memory alive
    return _PyCapsule_New(ctypes.addressof(managed), b"dltensor", ctypes.cast(_capsule_destructor, ctypes.c_void_p))


# ------------------------------------------------------------------ rows --


def device_rows(job, result):
    """The job's rows in place: a (rows, row bytes) uint8 device view."""
    from .batch import ROW_DTYPES
    from .device import DeviceView

    itemsize = ROW_DTYPES[job.kind].itemsize
    n = min(int(result["rows"]), int(job.row_capacity))
    base = getattr(job.rows, "base", job.rows)
    view = DeviceView(base, job.rows.data_ptr(), (n, itemsize))
    view.dtype = np.dtype(np.uint8)
    return view


@dataclass(frozen=True)
class RowPage:
    """One page of a job's rows (cf. SegmentedRowPage)."""

    stream_id: str
    page_index: int
    row_offset: int
    rows: np.ndarray
    page_capacity: int
    next_token: str | None
    overflowed: bool

    @property
    def valid_count(self) -> int:
        return len(self.rows)

    @property
    def end_offset(self) -> int:
        return self.row_offset + self.valid_count


def row_pages(job, result, page_rows: int, *, stream_id: str = "rows", start_token: str | None = None
              ) -> Iterator[RowPage]:
    """Pages of at most `page_rows` rows, in row order, each copied to the
    host only when reached.  `start_token` (a page's next_token) resumes a
    stream.  `overflowed` carries the job's RTDL_STATUS_ROW_OVERFLOW: the
    rows of an overflowing job are incomplete."""
    from .batch import ROW_DTYPES, STATUS_ROW_OVERFLOW

    if page_rows < 1:
        raise ValueError("page_rows must be positive")
    dtype = ROW_DTYPES[job.kind]
    total = min(int(result["rows"]), int(job.row_capacity))
    overflowed = bool(int(result["status"]) & STATUS_ROW_OVERFLOW)
    first = 0
    if start_token is not None:
        name, _, offset = start_token.rpartition(":")
        if name != stream_id or not offset.isdigit():
            raise ValueError(f"token {start_token!r} does not continue stream {stream_id!r}")
        first = int(offset)
    view = device_rows(job, result)
    for page, offset in enumerate(range(first, total, page_rows), start=first // page_rows):
        n = min(page_rows, total - offset)
        raw = view[offset:offset + n].download()
        end = offset + n
        yield RowPage(stream_id, page, offset, raw.view(dtype).copy(), page_rows,
                      f"{stream_id}:{end}" if end < total else None, overflowed)
