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
