"""Device-resident buffers allocated by the MI355X library itself.

The batched and served entry points take device pointers (query arenas,
instance inputs, row buffers).  ``DeviceBuffer`` holds such an allocation
through the library's own C-ABI (``rtdl_mi355x_device_alloc`` /
``_free`` / ``copy_to_device`` / ``copy_to_host``), so a process needs no
other GPU framework to drive the batched path: the library, its kernels and
its HIP runtime are the only things that touch the device.

Anything else that exposes ``data_ptr()``, ``shape`` and ``numel()`` for a
device allocation (a torch tensor, for instance) can stand in for a
``DeviceBuffer`` in the batch API; the caller then owns its synchronisation.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import mi355x_runtime as rt_native


def _check(status: int, error, what: str) -> None:
    if status != 0:
        raise RuntimeError(f"{what} failed with status {status}: "
                           + error.value.decode("utf-8", errors="replace"))


class DeviceBuffer:
    """A hipMalloc allocation with a numpy-like shape and dtype."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        ptr = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        lib = rt_native._library()
        _check(lib.rtdl_mi355x_device_alloc(self.nbytes, ctypes.byref(ptr), err, 1024), err,
               "rtdl_mi355x_device_alloc")
        self._ptr = ptr.value
        from .rows import current_device

        self.device = current_device()

    @classmethod
    def from_numpy(cls, array, dtype=None) -> "DeviceBuffer":
        host = np.ascontiguousarray(array, dtype=dtype)
        buf = cls(host.shape, host.dtype)
        buf.upload(host)
        return buf

    def upload(self, array) -> None:
        host = np.ascontiguousarray(array, dtype=self.dtype)
        if host.nbytes != self.nbytes:
            raise ValueError(f"upload of {host.nbytes} bytes into a {self.nbytes}-byte buffer")
        err = ctypes.create_string_buffer(1024)
        _check(rt_native._library().rtdl_mi355x_copy_to_device(self.ptr, host.ctypes.data, host.nbytes,
                                                                err, 1024),
               err, "rtdl_mi355x_copy_to_device")

    def download(self, nbytes: int | None = None) -> np.ndarray:
        """Copies the first `nbytes` (default all) to a uint8 host array."""
        n = self.nbytes if nbytes is None else int(nbytes)
        if n < 0 or n > self.nbytes:
            raise ValueError(f"download of {n} bytes from a {self.nbytes}-byte buffer")
        out = np.empty(n, dtype=np.uint8)
        if n:
            err = ctypes.create_string_buffer(1024)
            _check(rt_native._library().rtdl_mi355x_copy_to_host(out.ctypes.data, self.ptr, n, err, 1024),
                   err, "rtdl_mi355x_copy_to_host")
        return out

    def copy_from(self, other: "DeviceBuffer") -> None:
        """Device-to-device copy of `other` (same byte size) into this buffer."""
        if other.nbytes != self.nbytes:
            raise ValueError(f"copy of {other.nbytes} bytes into a {self.nbytes}-byte buffer")
        err = ctypes.create_string_buffer(1024)
        _check(rt_native._library().rtdl_mi355x_copy_device(self.ptr, other.ptr, self.nbytes, err, 1024),
               err, "rtdl_mi355x_copy_device")

    def to_numpy(self) -> np.ndarray:
        return self.download().view(self.dtype).reshape(self.shape)

    @property
    def ptr(self) -> int:
        if self._ptr is None:
            raise RuntimeError("device buffer is freed")
        return self._ptr

    # the interface the batch API reads from device arrays
    def data_ptr(self) -> int:
        return self.ptr

    def numel(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    def close(self) -> None:
        if getattr(self, "_ptr", None) is not None:
            ptr, self._ptr = self._ptr, None
            err = ctypes.create_string_buffer(1024)
            _check(rt_native._library().rtdl_mi355x_device_free(ptr, err, 1024), err,
                   "rtdl_mi355x_device_free")

    def __del__(self):
        try:
            self.close()
        except Exception as exc:  # a finaliser cannot raise; report instead
            import sys

            print(f"rtdl_amd.device: {exc}", file=sys.stderr)

    # DLPack: zero-copy hand-off to other frameworks (torch.from_dlpack, ...)
    def __dlpack__(self, stream=None):
        from .rows import to_dlpack

        return to_dlpack(self)

    def __dlpack_device__(self):
        from .rows import dlpack_device

        return dlpack_device(self)

    def __getitem__(self, rows) -> "DeviceView":
        """A view of a contiguous range of rows (first axis): buf[a:b]."""
        if not isinstance(rows, slice) or rows.step not in (None, 1):
            raise TypeError("device buffers support contiguous row slices only")
        start, stop, _ = rows.indices(self.shape[0])
        stop = max(start, stop)
        row_bytes = self.nbytes // max(1, self.shape[0])
        return DeviceView(self, self.ptr + start * row_bytes, (stop - start, *self.shape[1:]))


class DeviceView:
    """Rows [a, b) of a DeviceBuffer; keeps the buffer alive and never frees
    anything itself (only the DeviceBuffer owns the allocation)."""

    def __init__(self, base: DeviceBuffer, ptr: int, shape: tuple):
        self.base = base
        self._ptr = ptr
        self.shape = shape
        self.dtype = base.dtype
        self.device = getattr(base, "device", 0)

    def __dlpack__(self, stream=None):
        from .rows import to_dlpack

        return to_dlpack(self)

    def __dlpack_device__(self):
        from .rows import dlpack_device

        return dlpack_device(self)

    def data_ptr(self) -> int:
        return self._ptr

    def numel(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def nbytes(self) -> int:
        return self.numel() * self.dtype.itemsize

    def download(self, nbytes: int | None = None) -> np.ndarray:
        """Copies the view's first `nbytes` (default all) to a uint8 host array."""
        n = self.nbytes if nbytes is None else int(nbytes)
        if n < 0 or n > self.nbytes:
            raise ValueError(f"download of {n} bytes from a {self.nbytes}-byte view")
        out = np.empty(n, dtype=np.uint8)
        if n:
            err = ctypes.create_string_buffer(1024)
            _check(rt_native._library().rtdl_mi355x_copy_to_host(out.ctypes.data, self._ptr, n, err, 1024),
                   err, "rtdl_mi355x_copy_to_host")
        return out

    def __getitem__(self, rows) -> "DeviceView":
        if not isinstance(rows, slice) or rows.step not in (None, 1):
            raise TypeError("device views support contiguous row slices only")
        start, stop, _ = rows.indices(self.shape[0])
        stop = max(start, stop)
        row_bytes = self.nbytes // max(1, self.shape[0])
        view = DeviceView(self.base, self._ptr + start * row_bytes, (stop - start, *self.shape[1:]))
        view.dtype = self.dtype
        return view


def to_device(array, dtype=None) -> DeviceBuffer:
    return DeviceBuffer.from_numpy(array, dtype=dtype)
