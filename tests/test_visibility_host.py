"""Host-side checks of the visibility helpers that raise before any device
call (reference: tests/goal633_visibility_rows_test.py:86-101 for the
messages); no GPU is needed."""
from __future__ import annotations

import numpy as np
import pytest

import rtdl_amd as rt
from rtdl_amd import visibility as V


def test_mixed_dimensions_are_rejected():
    with pytest.raises(ValueError, match="one dimensionality"):
        rt.visibility_rows((rt.Point(1, 0.0, 0.0),), (rt.Point3D(2, 1.0, 0.0, 0.0),), ())
    with pytest.raises(ValueError, match="one dimensionality"):
        rt.visibility_rows((rt.Point(1, 0.0, 0.0), rt.Point3D(3, 0.0, 0.0, 1.0)), (rt.Point(2, 1.0, 0.0),), ())
    with pytest.raises(ValueError, match="share one dimensionality"):
        rt.visibility_pair_rows(((rt.Point(1, 0.0, 0.0), rt.Point3D(2, 1.0, 0.0, 0.0)),), ())


def test_zero_length_and_blocker_dimension_are_rejected():
    with pytest.raises(ValueError, match="distinct observer and target"):
        rt.visibility_rows((rt.Point(1, 0.0, 0.0),), (rt.Point(2, 0.0, 0.0),), ())
    with pytest.raises(ValueError, match="distinct observer and target"):
        rt.visibility_pair_rows(((rt.Point3D(1, 1.0, 2.0, 3.0), rt.Point3D(2, 1.0, 2.0, 3.0)),), ())
    tri3 = (rt.Triangle3D(20, 0.5, -1.0, -1.0, 0.5, 1.0, -1.0, 0.5, 0.0, 1.0),)
    with pytest.raises(ValueError, match="blocker triangles must match"):
        rt.visibility_rows((rt.Point(1, 0.0, 0.0),), (rt.Point(2, 1.0, 0.0),), tri3)
    with pytest.raises(ValueError, match="blocker triangles must match"):
        rt.visibility_pair_rows(((rt.Point(1, 0.0, 0.0), rt.Point(2, 1.0, 0.0)),), tri3)


def test_empty_inputs_and_backends():
    assert rt.visibility_rows((), (rt.Point(2, 1.0, 0.0),), ()) == ()
    assert rt.visibility_pair_rows((), ()) == ()
    raw = rt.visibility_rows((rt.Point(1, 0.0, 0.0),), (), (), result_mode="raw")
    assert raw.dtype == V.VISIBILITY_ROW_DTYPE and len(raw) == 0
    with pytest.raises(ValueError, match="backend must be 'mi355x'"):
        rt.visibility_rows((rt.Point(1, 0.0, 0.0),), (rt.Point(2, 1.0, 0.0),), (), backend="cpu")


def test_two_d_records_pack_to_the_reference_layouts():
    from rtdl_amd import mi355x_runtime as m

    assert m.RAY2D_DTYPE.itemsize == 44 and m.TRIANGLE2D_DTYPE.itemsize == 56
    r = m.pack_rays_2d((rt.Ray2D(5, 1.0, 2.0, 3.0, 4.0, 0.5),))
    assert r.tobytes() == np.array([5], "<u4").tobytes() + np.array([1.0, 2.0, 3.0, 4.0, 0.5]).tobytes()
    t = m.pack_triangles_2d((rt.Triangle(9, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0),))
    assert t.tobytes() == np.array([9, 0], "<u4").tobytes() + np.arange(1.0, 7.0).tobytes()
    with pytest.raises(ValueError, match="both be 2D or both be 3D"):
        m.ray_triangle_mi355x((rt.Ray2D(5, 1.0, 2.0, 3.0, 4.0, 0.5),),
                              (rt.Triangle3D(1, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0),), mode="hit_count")
