"""The C-ABI library: it loads, exports every symbol the header declares, and
rejects invalid arguments before touching a device (so these run on CPU)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from rtdl_amd import mi355x_runtime as m

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rtdl_mi355x.h")


def _declared_symbols() -> set[str]:
    text = open(HEADER).read()
    return set(re.findall(r"\b(rtdl_mi355x_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    lib = m._library()
    declared = _declared_symbols()
    assert len(declared) >= 25
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert missing == []
    assert declared == set(m.exported_symbols())


def test_record_layouts_match_the_reference_ctypes_structs():
    assert m.POINT2D_DTYPE.itemsize == 24 and m.POINT3D_DTYPE.itemsize == 32
    assert m.TRIANGLE3D_DTYPE.itemsize == 76 and m.RAY3D_DTYPE.itemsize == 60
    assert m.POLYGON_REF_DTYPE.itemsize == 12
    assert m.FRN_ROW_DTYPE.itemsize == 16 and m.KNN_ROW_DTYPE.itemsize == 24
    assert m.CLOSEST_ROW_DTYPE.itemsize == 16 and m.PIP_ROW_DTYPE.itemsize == 12


def test_version():
    assert m.mi355x_version() == (0, 1, 0)


def _call(name, *args):
    rows = ctypes.c_void_p()
    n = ctypes.c_size_t(123)
    err = ctypes.create_string_buffer(512)
    status = getattr(m._library(), name)(*args, ctypes.byref(rows), ctypes.byref(n), err, 512)
    return status, rows.value, n.value, err.value.decode()


def test_null_inputs_with_counts_are_rejected_without_a_device():
    status, rows, n, err = _call("rtdl_mi355x_run_fixed_radius_neighbors_3d", None, 5, None, 5, 1.0, 3)
    assert status != 0 and rows is None and n == 0
    assert "must not be null" in err
    status, _, _, err = _call("rtdl_mi355x_run_ray_anyhit_3d", None, 2, None, 0)
    assert status != 0 and "must not be null" in err
    status, _, _, err = _call("rtdl_mi355x_run_knn_rows_2d", None, 0, None, 0, 0)
    assert status != 0 and "k must be positive" in err


def test_invalid_prepared_handle_is_rejected():
    status, _, _, err = _call("rtdl_mi355x_run_prepared_knn_rows", None, None, 0, 3)
    assert status != 0 and "handle" in err


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setenv("RTDL_MI355X_LIB", "/nonexistent/librtdl_mi355x.so")
    monkeypatch.setattr(m, "_PACKAGE_DIR", "/nonexistent")
    m._library.cache_clear()
    try:
        with pytest.raises(FileNotFoundError, match="library not found"):
            m._library()
    finally:
        m._library.cache_clear()


def test_reference_shaped_prepared_calls_validate_before_the_device():
    """The HIPRT-shaped prepared calls take the HIPRT parameter lists
    (src/native/hiprt/rtdl_hiprt_api.cpp:15-79, 81-181, 439-486) and reject
    what rtdl_hiprt_core.cpp:1048-1063 rejects before any device work."""
    lib = m._library()
    handle = ctypes.c_void_p(7)
    err = ctypes.create_string_buffer(512)
    st = lib.rtdl_mi355x_prepare_fixed_radius_neighbors_3d(None, 3, 0.5, ctypes.byref(handle), err, 512)
    assert st != 0 and handle.value is None
    st = lib.rtdl_mi355x_prepare_fixed_radius_neighbors_3d(None, 0, -1.0, ctypes.byref(handle), err, 512)
    assert st != 0 and b"non-negative" in err.value
    st = lib.rtdl_mi355x_prepare_fixed_radius_neighbors_3d(None, 0, 0.5, ctypes.byref(handle), err, 512)
    assert st != 0 and b"at least one search point" in err.value
    status, rows, n, msg = _call("rtdl_mi355x_run_prepared_fixed_radius_neighbors_3d", None, None, 0, 4)
    assert status != 0 and rows is None and n == 0 and "handle" in msg
    st = lib.rtdl_mi355x_prepare_ray_hitcount_3d(None, 2, ctypes.byref(handle), err, 512)
    assert st != 0 and b"must not be null" in err.value
    st = lib.rtdl_mi355x_prepare_ray_anyhit_2d(None, 2, ctypes.byref(handle), err, 512)
    assert st != 0 and b"must not be null" in err.value
    for name in ("rtdl_mi355x_destroy_prepared_fixed_radius_neighbors_3d",
                 "rtdl_mi355x_destroy_prepared_ray_hitcount_3d", "rtdl_mi355x_destroy_prepared_ray_anyhit_2d"):
        getattr(lib, name)(None)  # a null handle is ignored, as delete nullptr is


def test_fault_hook_parses_specs_and_fails_an_allocation_before_the_device():
    """RTDL_MI355X_FAULT / rtdl_mi355x_set_fault: a bad spec is an error; an
    armed "alloc" fault fails the next device allocation on the host, once,
    with a message that names the hook; "" disarms."""
    lib = m._library()
    for bad in ("bogus", "alloc:0", "launch:x", "timeout:"):
        with pytest.raises(ValueError, match="fault spec"):
            m.set_fault(bad)
    m.set_fault("alloc:2")
    ptr = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    try:
        st = lib.rtdl_mi355x_device_alloc(64, ctypes.byref(ptr), err, 512)
        first = err.value.decode()
        assert "injected fault" not in first  # the first allocation is not the armed one
        st = lib.rtdl_mi355x_device_alloc(64, ctypes.byref(ptr), err, 512)
        assert st != 0 and ptr.value is None
        assert "injected fault (RTDL_MI355X_FAULT=alloc)" in err.value.decode()
    finally:
        m.set_fault("")
