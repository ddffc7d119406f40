"""Lowering ``CompiledKernel -> RTExecutionPlan`` for the hot-path predicates.

Restates the reference lowering (reference: src/rtdsl/lowering.py:41-90)
for fixed-radius neighbours (:435-497), KNN rows (:500-561), bounded KNN
rows (:564-626), ray/triangle any-hit (:702-771), hit-count and closest-hit
(:629-700) and point-in-polygon (:165-239), together with the role choice
(:976-1007) and the emitted record schema (:13-38, :1010-1021).

The plan's contract fields (workload kind, predicate, refine mode, payload
registers, launch parameters, buffers, output record, ray spec) keep the
reference's values so a plan produced here is interchangeable with the
reference's.  The descriptive ``host_steps`` say what the MI355X backend does.
"""
from __future__ import annotations

from .ir import (
    BufferSpec,
    CompiledKernel,
    LaunchParam,
    OutputField,
    OutputRecord,
    PayloadRegister,
    RaySpec,
    RTExecutionPlan,
)

# reference: src/rtdsl/lowering.py:13-38
_EMIT_FIELD_TYPES = {
    "left_id": "uint32_t",
    "right_id": "uint32_t",
    "intersection_point_x": "float",
    "intersection_point_y": "float",
    "point_id": "uint32_t",
    "polygon_id": "uint32_t",
    "contains": "uint32_t",
    "left_polygon_id": "uint32_t",
    "right_polygon_id": "uint32_t",
    "requires_lsi": "uint32_t",
    "requires_pip": "uint32_t",
    "intersection_area": "uint32_t",
    "left_area": "uint32_t",
    "right_area": "uint32_t",
    "union_area": "uint32_t",
    "jaccard_similarity": "float",
    "ray_id": "uint32_t",
    "hit_count": "uint32_t",
    "any_hit": "uint32_t",
    "segment_id": "uint32_t",
    "query_id": "uint32_t",
    "neighbor_id": "uint32_t",
    "distance": "float",
    "neighbor_rank": "uint32_t",
    "triangle_id": "uint32_t",
    "t": "float",
}


def lower_to_execution_plan(kernel: CompiledKernel) -> RTExecutionPlan:
    if kernel.backend not in {"rtdl", "rayjoin"}:
        raise ValueError(f"unsupported backend for RTDL lowering: {kernel.backend}")
    if kernel.candidates is None or kernel.refine_op is None or kernel.emit_op is None:
        raise ValueError("kernel is incomplete and cannot be lowered")
    if kernel.candidates.accel != "bvh":
        raise ValueError(
            "the current RTDL lowering supports only accel='bvh' "
            f"(got accel={kernel.candidates.accel!r})"
        )
    if kernel.precision != "float_approx":
        raise ValueError(
            "the current RTDL lowering supports only precision='float_approx'; "
            "precision='exact' is not implemented yet"
        )
    build_input, probe_input = _choose_roles(kernel)
    name = kernel.refine_op.predicate.name
    lower = _LOWERERS.get(name)
    if lower is None:
        raise ValueError(f"unsupported predicate for current RTDL lowering: {name}")
    return lower(kernel, build_input, probe_input)


def lower_to_rayjoin(kernel: CompiledKernel) -> RTExecutionPlan:
    """Legacy alias kept for interface parity with the reference."""
    return lower_to_execution_plan(kernel)


# --------------------------------------------------------------------------
# shared pieces
# --------------------------------------------------------------------------


def _choose_roles(kernel: CompiledKernel):
    """Return ``(build, probe)``; explicit roles win, then geometry pairs."""
    left = kernel.candidates.left
    right = kernel.candidates.right
    if left.role is not None and left.role == right.role:
        raise ValueError("candidate inputs cannot share the same explicit role")
    if left.role == "build" or right.role == "probe":
        return left, right
    if right.role == "build" or left.role == "probe":
        return right, left
    pair = (left.geometry.name, right.geometry.name)
    if pair in {("polygons", "points"), ("triangles", "rays")}:
        return left, right
    return right, left


def _build_output_record(record_name: str, emit_fields: tuple[str, ...]) -> OutputRecord:
    if not emit_fields:
        raise ValueError("emit schema must contain at least one field")
    fields = []
    for name in emit_fields:
        c_type = _EMIT_FIELD_TYPES.get(name)
        if c_type is None:
            raise ValueError(f"unsupported emitted field for current RTDL lowering: {name}")
        fields.append(OutputField(name=name, c_type=c_type))
    return OutputRecord(name=record_name, fields=tuple(fields))


def _input_buffer_name(geometry_input) -> str:
    if geometry_input.name == geometry_input.geometry.name:
        return geometry_input.name
    return f"{geometry_input.name}_{geometry_input.geometry.name}"


def _io(build_input, probe_input, record: OutputRecord):
    build_name = _input_buffer_name(build_input)
    probe_name = _input_buffer_name(probe_input)
    params = (
        LaunchParam(name="traversable", c_type="OptixTraversableHandle", role="rt_accel"),
        LaunchParam(name=build_name, c_type=f"const {build_input.layout.name}*", role="device_input_build"),
        LaunchParam(name=probe_name, c_type=f"const {probe_input.layout.name}*", role="device_input_probe"),
        LaunchParam(name="output_records", c_type=f"{record.name}*", role="device_output"),
        LaunchParam(name="output_count", c_type="uint32_t*", role="device_counter"),
        LaunchParam(name="output_capacity", c_type="uint32_t", role="device_limit"),
        LaunchParam(name="probe_count", c_type="uint32_t", role="launch_size"),
    )
    buffers = (
        BufferSpec(name=build_name, element=build_input.layout.name, role="device_input_build"),
        BufferSpec(name=probe_name, element=probe_input.layout.name, role="device_input_probe"),
        BufferSpec(name="output_records", element=record.name, role="device_output"),
        BufferSpec(name="output_count", element="uint32_t", role="device_counter"),
        BufferSpec(name="output_capacity", element="uint32_t", role="device_limit"),
    )
    return params, buffers


def _registers(*names_and_encodings: tuple[str, str]) -> tuple[PayloadRegister, ...]:
    return tuple(
        PayloadRegister(index=i, name=name, encoding=enc)
        for i, (name, enc) in enumerate(names_and_encodings)
    )


def _point_ray_spec(description: str) -> RaySpec:
    return RaySpec(
        origin=("probe.x", "probe.y", "0.0f"),
        direction=("0.0f", "1.0f", "0.0f"),
        tmin="0.0f",
        tmax="FLT_MAX",
        description=description,
    )


def _require_points(build_input, probe_input, what: str) -> None:
    if build_input.geometry.name != "points" or probe_input.geometry.name != "points":
        raise ValueError(f"{what} lowering requires point build input and point probe input")


def _require_float_predicate(kernel: CompiledKernel, message: str) -> None:
    if kernel.refine_op.predicate.options.get("exact") is not False:
        raise ValueError(message)


# --------------------------------------------------------------------------
# nearest-neighbour family
# --------------------------------------------------------------------------


def _neighbor_plan(
    kernel: CompiledKernel,
    build_input,
    probe_input,
    *,
    workload: str,
    record_name: str,
    fourth_register: str,
    selection_step: str,
    description: str,
) -> RTExecutionPlan:
    record = _build_output_record(record_name, kernel.emit_op.fields)
    params, buffers = _io(build_input, probe_input, record)
    return RTExecutionPlan(
        kernel_name=kernel.name,
        workload_kind=workload,
        backend="rtdl",
        precision=kernel.precision,
        build_input=build_input,
        probe_input=probe_input,
        accel_kind="native_loop",
        predicate=workload,
        exact_refine_mode=f"analytic_float_{workload}",
        emit_fields=kernel.emit_op.fields,
        payload_registers=_registers(
            ("query_index", "u32"),
            ("neighbor_index", "u32"),
            ("distance_bits", "f32_bits"),
            (fourth_register, "u32"),
        ),
        launch_params=params,
        host_steps=(
            f"Upload `{build_input.name}` search points and `{probe_input.name}` query points.",
            "Build an LBVH over the search points (Morton codes, LDS radix sort, Karras hierarchy, bottom-up refit).",
            selection_step,
            "Compact per-query rows with a wavefront prefix-sum into a device row buffer grouped by ascending query id.",
        ),
        device_programs=(
            f"__raygen__rtdl_{workload}",
            "__miss__rtdl_miss",
            f"__closesthit__rtdl_{workload}_refine",
        ),
        buffers=buffers,
        output_record=record,
        ray_spec=_point_ray_spec(description),
        bvh_policy=f"build over `{build_input.name}`, probe with `{probe_input.name}`",
    )


def _lower_fixed_radius_neighbors(kernel, build_input, probe_input) -> RTExecutionPlan:
    _require_points(build_input, probe_input, "fixed_radius_neighbors")
    opts = kernel.refine_op.predicate.options
    return _neighbor_plan(
        kernel,
        build_input,
        probe_input,
        workload="fixed_radius_neighbors",
        record_name="FixedRadiusNeighborRecord",
        fourth_register="query_kind",
        selection_step=(
            f"Apply inclusive radius filtering with radius={opts['radius']!r}, sort per query by "
            f"distance then neighbor id, and truncate to k_max={opts['k_max']!r}."
        ),
        description="Fixed-radius neighbours over point sets via LBVH traversal with an FP64 refine.",
    )


def _lower_knn_rows(kernel, build_input, probe_input) -> RTExecutionPlan:
    _require_points(build_input, probe_input, "knn_rows")
    k = kernel.refine_op.predicate.options["k"]
    return _neighbor_plan(
        kernel,
        build_input,
        probe_input,
        workload="knn_rows",
        record_name="KnnNeighborRecord",
        fourth_register="neighbor_rank",
        selection_step=(
            "Sort each query's candidate set by distance then neighbor id, assign 1-based "
            f"neighbor_rank, and emit at most k={k!r} rows."
        ),
        description="KNN over point sets via LBVH collect-k traversal with an FP64 refine.",
    )


def _lower_bounded_knn_rows(kernel, build_input, probe_input) -> RTExecutionPlan:
    _require_points(build_input, probe_input, "bounded_knn_rows")
    opts = kernel.refine_op.predicate.options
    return _neighbor_plan(
        kernel,
        build_input,
        probe_input,
        workload="bounded_knn_rows",
        record_name="BoundedKnnNeighborRecord",
        fourth_register="neighbor_rank",
        selection_step=(
            f"Filter neighbors by inclusive radius={opts['radius']!r}, sort each query by distance "
            f"then neighbor id, assign 1-based neighbor_rank, and emit at most k_max={opts['k_max']!r} rows."
        ),
        description="Bounded KNN over point sets via LBVH collect-k traversal with an FP64 refine.",
    )


# --------------------------------------------------------------------------
# ray / triangle family
# --------------------------------------------------------------------------


def _ray_plan(
    kernel: CompiledKernel,
    build_input,
    probe_input,
    *,
    workload: str,
    record_name: str,
    refine_mode: str,
    registers: tuple[tuple[str, str], ...],
    steps: tuple[str, ...],
    programs: tuple[str, ...],
    description: str,
    bvh_note: str,
) -> RTExecutionPlan:
    predicate = kernel.refine_op.predicate
    record = _build_output_record(record_name, kernel.emit_op.fields)
    params, buffers = _io(build_input, probe_input, record)
    return RTExecutionPlan(
        kernel_name=kernel.name,
        workload_kind=workload,
        backend="rtdl",
        precision=kernel.precision,
        build_input=build_input,
        probe_input=probe_input,
        accel_kind=kernel.candidates.accel,
        predicate=predicate.name,
        exact_refine_mode=refine_mode,
        emit_fields=kernel.emit_op.fields,
        payload_registers=_registers(*registers),
        launch_params=params,
        host_steps=(
            f"Upload `{build_input.name}` triangles and `{probe_input.name}` rays using "
            f"`{build_input.layout.name}` / `{probe_input.layout.name}` layouts.",
            f"Build an LBVH over `{build_input.name}` triangles in device memory.",
        )
        + steps,
        device_programs=programs,
        buffers=buffers,
        output_record=record,
        ray_spec=RaySpec(
            origin=("probe.ox", "probe.oy", "0.0f"),
            direction=("probe.dx", "probe.dy", "0.0f"),
            tmin="0.0f",
            tmax="probe.tmax",
            description=description,
        ),
        bvh_policy=f"build over `{build_input.name}`, probe with `{probe_input.name}`{bvh_note}",
    )


def _require_rays(build_input, probe_input, what: str) -> None:
    if build_input.geometry.name != "triangles" or probe_input.geometry.name != "rays":
        raise ValueError(
            f"the current {what} lowering requires triangle build input and ray probe input"
        )


def _lower_ray_triangle_any_hit(kernel, build_input, probe_input) -> RTExecutionPlan:
    _require_rays(build_input, probe_input, "ray any-hit")
    _require_float_predicate(
        kernel,
        "the current ray any-hit lowering supports only float-based ray_triangle_any_hit predicates; "
        "use ray_triangle_any_hit(exact=False)",
    )
    return _ray_plan(
        kernel,
        build_input,
        probe_input,
        workload="ray_tri_anyhit",
        record_name="RayAnyHitRecord",
        refine_mode="analytic_float_ray_triangle_any_hit",
        registers=(("ray_index", "u32"), ("triangle_index", "u32"), ("any_hit", "u32"), ("hit_kind", "u32")),
        steps=(
            "Launch one finite ray per probe lane over FP32 boxes inflated by the stated bound.",
            "Refine candidate triangles with FP64 Moller-Trumbore and terminate traversal after the first accepted triangle hit.",
            "Emit exactly one row per ray with any_hit=1 for a blocker and any_hit=0 otherwise.",
        ),
        programs=(
            "__raygen__rtdl_ray_anyhit",
            "__miss__rtdl_miss",
            "__anyhit__rtdl_triangle_terminate",
            "__intersection__rtdl_triangles",
        ),
        description="Trace each finite ray against the build-side triangle BVH and stop after the first accepted hit.",
        bvh_note="; any-hit traversal is bounded by early termination",
    )


def _lower_ray_triangle_hitcount(kernel, build_input, probe_input) -> RTExecutionPlan:
    _require_rays(build_input, probe_input, "ray hit-count")
    _require_float_predicate(
        kernel,
        "the current ray hit-count lowering supports only float-based ray_triangle_hit_count predicates; "
        "use ray_triangle_hit_count(exact=False)",
    )
    return _ray_plan(
        kernel,
        build_input,
        probe_input,
        workload="ray_tri_hitcount",
        record_name="RayHitCountRecord",
        refine_mode="analytic_float_ray_triangle_hit_count",
        registers=(("ray_index", "u32"), ("triangle_index", "u32"), ("hit_count", "u32"), ("hit_kind", "u32")),
        steps=(
            "Launch one finite ray per probe lane and visit every overlapping leaf (no early exit).",
            "Count FP64 Moller-Trumbore hits per ray and emit one row per ray.",
        ),
        programs=(
            "__raygen__rtdl_ray_hitcount",
            "__miss__rtdl_miss",
            "__anyhit__rtdl_triangle_count",
            "__intersection__rtdl_triangles",
        ),
        description="Trace each finite ray against the build-side triangle BVH and count every accepted hit.",
        bvh_note="",
    )


def _lower_ray_triangle_closest_hit(kernel, build_input, probe_input) -> RTExecutionPlan:
    _require_rays(build_input, probe_input, "ray closest-hit")
    _require_float_predicate(
        kernel,
        "the current ray closest-hit lowering supports only float-based ray_triangle_closest_hit predicates; "
        "use ray_triangle_closest_hit(exact=False)",
    )
    return _ray_plan(
        kernel,
        build_input,
        probe_input,
        workload="ray_tri_closest_hit",
        record_name="RayClosestHitRecord",
        refine_mode="analytic_float_ray_triangle_closest_hit",
        registers=(("ray_index", "u32"), ("triangle_index", "u32"), ("t_bits", "f64_bits"), ("hit_kind", "u32")),
        steps=(
            "Launch one finite ray per probe lane and shrink tmax as closer FP64 hits are accepted.",
            "Emit one row per ray that hits, with the lowest t and the lower triangle id on ties.",
        ),
        programs=(
            "__raygen__rtdl_ray_closest_hit",
            "__miss__rtdl_miss",
            "__closesthit__rtdl_triangle_closest",
            "__intersection__rtdl_triangles",
        ),
        description="Trace each finite ray against the build-side triangle BVH and keep the nearest accepted hit.",
        bvh_note="",
    )


# --------------------------------------------------------------------------
# point-in-polygon
# --------------------------------------------------------------------------


def _lower_pip(kernel, build_input, probe_input) -> RTExecutionPlan:
    predicate = kernel.refine_op.predicate
    if build_input.geometry.name != "polygons" or probe_input.geometry.name != "points":
        raise ValueError("the current PIP lowering requires polygon build input and point probe input")
    _require_float_predicate(
        kernel,
        "the current RTDL lowering supports only float-based point_in_polygon predicates; "
        "use point_in_polygon(exact=False)",
    )
    if predicate.options.get("boundary_mode") != "inclusive":
        raise ValueError("the current PIP lowering supports only boundary_mode='inclusive'")
    if predicate.options.get("result_mode", "full_matrix") not in {"full_matrix", "positive_hits"}:
        raise ValueError(
            "the current PIP lowering supports only result_mode='full_matrix' or 'positive_hits'"
        )
    record = _build_output_record("PointInPolygonRecord", kernel.emit_op.fields)
    params, buffers = _io(build_input, probe_input, record)
    return RTExecutionPlan(
        kernel_name=kernel.name,
        workload_kind="pip",
        backend="rtdl",
        precision=kernel.precision,
        build_input=build_input,
        probe_input=probe_input,
        accel_kind=kernel.candidates.accel,
        predicate=predicate.name,
        exact_refine_mode="analytic_float_point_in_polygon",
        emit_fields=kernel.emit_op.fields,
        payload_registers=_registers(
            ("point_index", "u32"), ("polygon_face_index", "u32"), ("winding_hits", "u32"), ("hit_kind", "u32")
        ),
        launch_params=params,
        host_steps=(
            f"Upload `{build_input.name}` polygon refs and `{probe_input.name}` points using "
            f"`{build_input.layout.name}` / `{probe_input.layout.name}` layouts.",
            f"Build an LBVH over polygon bounding boxes from `{build_input.name}`.",
            f"Traverse one `{probe_input.name}` point per lane and refine candidate polygons with the FP64 "
            "inclusive edge test and even-odd crossing rule before emitting records.",
        ),
        device_programs=(
            "__raygen__rtdl_pip_probe",
            "__miss__rtdl_miss",
            "__closesthit__rtdl_pip_refine",
            "__intersection__rtdl_polygon_refs",
        ),
        buffers=buffers,
        output_record=record,
        ray_spec=RaySpec(
            origin=("probe.x", "probe.y", "0.0f"),
            direction=("0.0f", "1.0f", "0.0f"),
            tmin="0.0f",
            tmax="FLT_MAX",
            description="Cast one upward parity ray from each probe point over the build-side polygon BVH.",
        ),
        bvh_policy=f"build over `{build_input.name}`, probe with `{probe_input.name}`",
    )


_LOWERERS = {
    "point_in_polygon": _lower_pip,
    "ray_triangle_any_hit": _lower_ray_triangle_any_hit,
    "ray_triangle_hit_count": _lower_ray_triangle_hitcount,
    "ray_triangle_closest_hit": _lower_ray_triangle_closest_hit,
    "fixed_radius_neighbors": _lower_fixed_radius_neighbors,
    "knn_rows": _lower_knn_rows,
    "bounded_knn_rows": _lower_bounded_knn_rows,
}
