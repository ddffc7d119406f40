"""The ``@rt.kernel`` contract, lowering, validation and the execution facade.

Mirrors the reference's language / lowering / dispatch tests
(tests/goal198, goal204, goal548, goal632, goal543 in the reference): same
names, same plan fields, same exceptions raised before any library loads.
"""
from __future__ import annotations

import pytest

import rtdl_amd as rt
from rtdl_amd import lowering
from rtdl_amd.parity import compare_rows


@rt.kernel(backend="rtdl", precision="float_approx")
def frn_kernel():
    q = rt.input("query_points", rt.Points, role="probe")
    s = rt.input("search_points", rt.Points, role="build")
    c = rt.traverse(q, s, accel="bvh")
    h = rt.refine(c, predicate=rt.fixed_radius_neighbors(radius=0.5, k_max=3))
    return rt.emit(h, fields=["query_id", "neighbor_id", "distance"])


@rt.kernel(backend="rtdl", precision="float_approx")
def knn_kernel():
    q = rt.input("query_points", rt.Points3D, layout=rt.Point3DLayout, role="probe")
    s = rt.input("search_points", rt.Points3D, layout=rt.Point3DLayout, role="build")
    c = rt.traverse(q, s, accel="bvh")
    h = rt.refine(c, predicate=rt.knn_rows(k=16))
    return rt.emit(h, fields=["query_id", "neighbor_id", "distance", "neighbor_rank"])


@rt.kernel(backend="rtdl", precision="float_approx")
def anyhit_kernel():
    r = rt.input("rays", rt.Rays3D, layout=rt.Ray3DLayout, role="probe")
    t = rt.input("triangles", rt.Triangles3D, layout=rt.Triangle3DLayout, role="build")
    c = rt.traverse(r, t, accel="bvh")
    h = rt.refine(c, predicate=rt.ray_triangle_any_hit(exact=False))
    return rt.emit(h, fields=["ray_id", "any_hit"])


@rt.kernel(backend="rtdl", precision="float_approx")
def pip_kernel():
    p = rt.input("points", rt.Points, role="probe")
    g = rt.input("polygons", rt.Polygons, role="build")
    c = rt.traverse(p, g, accel="bvh")
    h = rt.refine(c, predicate=rt.point_in_polygon(exact=False, result_mode="positive_hits"))
    return rt.emit(h, fields=["point_id", "polygon_id", "contains"])


def test_compile_records_inputs_roles_and_predicate():
    compiled = rt.compile_kernel(frn_kernel)
    assert compiled.name == "frn_kernel"
    assert [(i.name, i.role, i.layout.name) for i in compiled.inputs] == [
        ("query_points", "probe", "Point2D"), ("search_points", "build", "Point2D")]
    assert compiled.refine_op.predicate.options == {"radius": 0.5, "k_max": 3}
    assert compiled.emit_op.fields == ("query_id", "neighbor_id", "distance")
    assert "Lowering Plan" in compiled.format()


@pytest.mark.parametrize("kernel,workload,record", [
    (frn_kernel, "fixed_radius_neighbors", "FixedRadiusNeighborRecord"),
    (knn_kernel, "knn_rows", "KnnNeighborRecord"),
    (anyhit_kernel, "ray_tri_anyhit", "RayAnyHitRecord"),
    (pip_kernel, "pip", "PointInPolygonRecord"),
])
def test_lowering_produces_the_reference_plan_fields(kernel, workload, record):
    plan = rt.lower_to_execution_plan(rt.compile_kernel(kernel))
    assert plan.workload_kind == workload
    assert plan.output_record.name == record
    d = plan.to_dict()
    assert d["schema_version"] == "v1alpha1"
    assert d["emit_fields"] == list(plan.emit_fields)
    assert [p["role"] for p in d["launch_params"]][:3] == ["rt_accel", "device_input_build", "device_input_probe"]
    assert "RTDL Backend Plan" in plan.format()


def test_any_hit_plan_mentions_early_termination():
    plan = rt.lower_to_execution_plan(rt.compile_kernel(anyhit_kernel))
    assert "terminate traversal after the first accepted triangle hit" in " ".join(plan.host_steps)
    assert plan.ray_spec.tmax == "probe.tmax"


def test_emit_field_types_follow_reference_table():
    plan = rt.lower_to_execution_plan(rt.compile_kernel(knn_kernel))
    assert [(f.name, f.c_type) for f in plan.output_record.fields] == [
        ("query_id", "uint32_t"), ("neighbor_id", "uint32_t"), ("distance", "float"),
        ("neighbor_rank", "uint32_t")]


def test_role_choice_without_explicit_roles():
    @rt.kernel(backend="rtdl", precision="float_approx")
    def implicit():
        p = rt.input("points", rt.Points)
        g = rt.input("polygons", rt.Polygons)
        c = rt.traverse(p, g, accel="bvh")
        h = rt.refine(c, predicate=rt.point_in_polygon(exact=False))
        return rt.emit(h, fields=["point_id", "polygon_id", "contains"])

    build, probe = lowering._choose_roles(rt.compile_kernel(implicit))
    assert (build.name, probe.name) == ("polygons", "points")


def test_predicate_option_validation_messages():
    with pytest.raises(ValueError, match="radius must be non-negative"):
        rt.fixed_radius_neighbors(radius=-1.0, k_max=1)
    with pytest.raises(ValueError, match="k_max must be positive"):
        rt.fixed_radius_neighbors(radius=1.0, k_max=0)
    with pytest.raises(ValueError, match="knn_rows k must be positive"):
        rt.knn_rows(k=0)
    with pytest.raises(ValueError, match="result_mode"):
        rt.point_in_polygon(result_mode="bogus")


def test_dsl_guards():
    with pytest.raises(RuntimeError, match="inside @rt.kernel"):
        rt.input("x", rt.Points)

    @rt.kernel(backend="rtdl", precision="float_approx")
    def no_emit():
        rt.input("a", rt.Points)

    with pytest.raises(TypeError, match="must return rt.emit"):
        no_emit()

    @rt.kernel(backend="rtdl", precision="float_approx")
    def duplicate():
        rt.input("a", rt.Points)
        rt.input("a", rt.Points)

    with pytest.raises(ValueError, match="duplicate input name"):
        duplicate()


def test_lowering_rejects_unsupported_shapes():
    @rt.kernel(backend="rtdl", precision="exact")
    def exact_kernel():
        q = rt.input("q", rt.Points, role="probe")
        s = rt.input("s", rt.Points, role="build")
        c = rt.traverse(q, s, accel="bvh")
        h = rt.refine(c, predicate=rt.knn_rows(k=1))
        return rt.emit(h, fields=["query_id"])

    with pytest.raises(ValueError, match="precision='float_approx'"):
        rt.lower_to_execution_plan(exact_kernel())
    with pytest.raises(ValueError, match="unsupported emitted field"):
        lowering._build_output_record("X", ("bogus",))


def test_validation_happens_before_the_library_loads(monkeypatch):
    from rtdl_amd import mi355x_runtime as m

    def boom():
        raise AssertionError("library must not load during validation")

    monkeypatch.setattr(m, "_library", boom)
    with pytest.raises(ValueError, match="missing RTDL MI355X inputs"):
        rt.run_mi355x(frn_kernel, query_points=())
    with pytest.raises(ValueError, match="unexpected RTDL MI355X inputs"):
        rt.run_mi355x(frn_kernel, query_points=(), search_points=(), extra=())
    with pytest.raises(ValueError, match="result_mode"):
        rt.run_mi355x(frn_kernel, result_mode="bogus", query_points=(), search_points=())

    @rt.kernel(backend="rtdl", precision="float_approx")
    def wrong_roles():
        q = rt.input("query_points", rt.Points, role="build")
        s = rt.input("search_points", rt.Points, role="probe")
        c = rt.traverse(q, s, accel="bvh")
        h = rt.refine(c, predicate=rt.knn_rows(k=1))
        return rt.emit(h, fields=["query_id"])

    with pytest.raises(ValueError, match="role='probe'"):
        rt.run_mi355x(wrong_roles, query_points=(), search_points=())
    with pytest.raises(ValueError, match="missing RTDL MI355X prepare inputs"):
        rt.prepare_mi355x(knn_kernel)


def test_execution_policy_and_registry():
    assert rt.ExecutionPolicy().backend == "mi355x"
    with pytest.raises(ValueError, match="backend must be one of"):
        rt.ExecutionPolicy(backend="cuda")
    with pytest.raises(ValueError, match="no registered runner"):
        rt.run(frn_kernel, {}, execution=rt.ExecutionPolicy(backend="embree"))
    assert "mi355x" in rt.registered_backends()


def test_execution_facade_reports_registered_backend():
    from oracle.cpu_python_reference import run_cpu_python_reference

    rt.register_backend("cpu_python_reference",
                        lambda k, result_mode="dict", **inputs: run_cpu_python_reference(k, **inputs))
    case = {
        "query_points": (rt.Point(100, 0.0, 0.0),),
        "search_points": (rt.Point(1, 0.0, 0.0), rt.Point(2, 0.3, 0.0)),
    }
    result = rt.run(frn_kernel, case, execution=rt.ExecutionPolicy(backend="cpu_python_reference"))
    assert [r["neighbor_id"] for r in result.rows] == [1, 2]
    report = result.execution_report.to_dict()
    assert report["selected_backend"] == "cpu_python_reference"
    assert report["primitive_family"] == "nearest_neighbor"
    assert report["claim_boundary"]["fallback_used"] is False


def test_compare_rows_contract():
    a = ({"query_id": 1, "neighbor_id": 2, "distance": 0.5},)
    assert compare_rows(a, ({"query_id": 1, "neighbor_id": 2, "distance": 0.5 + 1e-9},))
    assert not compare_rows(a, ({"query_id": 1, "neighbor_id": 3, "distance": 0.5},))
    assert not compare_rows(a, ({"query_id": 1, "neighbor_id": 2, "distance": 0.5 + 1e-9},), exact=True)
    assert not compare_rows(a, ())


def test_records_normalisation():
    from rtdl_amd.records import normalize_records

    pts = normalize_records("p", "points", [{"id": 1, "x": 1, "y": 2}, {"id": 2, "x": 1, "y": 2, "z": 3}])
    assert type(pts[0]).__name__ == "Point" and type(pts[1]).__name__ == "Point3D"
    with pytest.raises(ValueError, match="at least 3 vertices"):
        normalize_records("g", "polygons", [{"id": 1, "vertices": [(0, 0), (1, 1)]}])
    with pytest.raises(ValueError, match="missing fields"):
        normalize_records("p", "points", [{"id": 1, "x": 1}])
    with pytest.raises(ValueError, match="iterable of records"):
        normalize_records("p", "points", "abc")
