"""Ownership rules of rtdl_amd.device (CPU only, against a recording fake of
the library's device-buffer entry points): a DeviceBuffer frees its
allocation exactly once, with the pointer it was given; row views never free
anything and keep their buffer alive."""
from __future__ import annotations

import ctypes
import gc

import numpy as np
import pytest

from rtdl_amd import device as dev
from rtdl_amd import mi355x_runtime as rt


class _FakeLib:
    def __init__(self):
        self.next = 0x10000
        self.live = {}
        self.freed = []

    def rtdl_mi355x_device_alloc(self, nbytes, ptr_ref, err, size):
        ptr_ref._obj.value = self.next
        self.live[self.next] = nbytes
        self.next += max(256, (nbytes + 255) // 256 * 256) + 256
        return 0

    def rtdl_mi355x_device_free(self, ptr, err, size):
        if ptr not in self.live:  # what hipFree rejects: not an allocation base
            return 1
        del self.live[ptr]
        self.freed.append(ptr)
        return 0

    def rtdl_mi355x_copy_to_device(self, dst, src, nbytes, err, size):
        return 0

    def rtdl_mi355x_copy_to_host(self, dst, src, nbytes, err, size):
        ctypes.memset(dst, 0, nbytes)
        return 0

    def rtdl_mi355x_get_device(self, device_ref, err, size):
        device_ref._obj.value = 0
        return 0


@pytest.fixture
def fake(monkeypatch):
    lib = _FakeLib()
    monkeypatch.setattr(rt, "_library", lambda: lib)
    return lib


def test_buffer_frees_its_base_pointer_once(fake):
    buf = dev.DeviceBuffer((100, 3), np.float32)
    base = buf.ptr
    buf.close()
    buf.close()
    del buf
    gc.collect()
    assert fake.freed == [base]
    assert not fake.live


def test_views_never_free_and_keep_the_buffer_alive(fake):
    buf = dev.DeviceBuffer((1000, 3), np.float32)
    base = buf.ptr
    views = [buf[0:10], buf[10:500], buf[999:1000], buf[5:5]]
    assert views[0].data_ptr() == base
    assert views[1].data_ptr() == base + 10 * 12
    assert views[1].shape == (490, 3) and views[3].numel() == 0
    assert not hasattr(views[0], "close")
    del buf
    gc.collect()
    assert fake.freed == []  # the views still hold the buffer
    del views
    gc.collect()
    assert fake.freed == [base]


def test_view_slicing_rejects_steps(fake):
    buf = dev.DeviceBuffer((10, 2), np.float32)
    with pytest.raises(TypeError):
        buf[0:10:2]


def test_dlpack_capsule_describes_the_rows_in_place(fake):
    """A DeviceBuffer / row view exports a "dltensor" capsule: ROCm device,
    the buffer's own address (no copy), its shape and a compact layout; the
    capsule keeps the buffer alive until it is released."""
    from rtdl_amd import rows

    buf = dev.DeviceBuffer((100, 24), np.uint8)
    view = buf[10:30]
    assert view.__dlpack_device__() == (rows.KDL_ROCM, 0)
    cap = view.__dlpack__()
    ptr = ctypes.pythonapi.PyCapsule_GetPointer
    ptr.restype = ctypes.c_void_p
    ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    managed = ctypes.cast(ptr(cap, b"dltensor"), ctypes.POINTER(rows._DLManagedTensor)).contents
    t = managed.dl_tensor
    assert t.data == buf.ptr + 10 * 24
    assert (t.device.device_type, t.device.device_id) == (rows.KDL_ROCM, 0)
    assert t.ndim == 2 and [t.shape[0], t.shape[1]] == [20, 24]
    assert (t.dtype.code, t.dtype.bits, t.dtype.lanes) == (1, 8, 1)
    assert not t.strides and t.byte_offset == 0
    n_live = len(rows._LIVE)
    del view, buf
    gc.collect()
    assert fake.freed == []  # the capsule still holds the buffer
    del cap, managed, t
    gc.collect()
    assert len(rows._LIVE) == n_live - 1
    assert len(fake.freed) == 1
