"""The ``@rt.kernel`` DSL front end: ``input -> traverse -> refine -> emit``.

Restates the reference front end (reference: src/rtdsl/api.py:16-279) for
the predicates this backend executes.  A kernel body runs once, inside a
compilation context, and records one ``GeometryInput`` per ``rt.input``, a
``CandidateSet`` for ``rt.traverse``, a ``RefineOp`` carrying the predicate
and an ``EmitOp`` naming the output fields; calling the decorated function
returns the ``CompiledKernel``.

Predicate constructors validate their options with the reference's messages
(reference: src/rtdsl/api.py:122-200) so that invalid kernels fail the same
way before any backend is touched.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from functools import wraps

from .ir import CandidateSet, CompiledKernel, EmitOp, GeometryInput, Predicate, RefineOp
from .layout_types import GeometryType, Layout

_LOWERING_STEPS = (
    "Normalize geometry roles, layouts, and kernel constraints.",
    "Select the `{precision}` precision policy for `{backend}`.",
    "Choose acceleration ownership, launch parameter schema, and payload register mapping.",
    "Lower traversal and refine stages into RTDL backend-plan operations.",
    "Materialize the declared output record schema.",
)


@dataclass
class _Recording:
    """State collected while one kernel body executes."""

    name: str
    backend: str
    precision: str
    inputs: list[GeometryInput] = field(default_factory=list)
    candidates: CandidateSet | None = None
    refine_op: RefineOp | None = None
    emit_op: EmitOp | None = None

    def finish(self) -> CompiledKernel:
        steps = tuple(
            step.format(precision=self.precision, backend=self.backend) for step in _LOWERING_STEPS
        )
        return CompiledKernel(
            name=self.name,
            backend=self.backend,
            precision=self.precision,
            inputs=tuple(self.inputs),
            candidates=self.candidates,
            refine_op=self.refine_op,
            emit_op=self.emit_op,
            lowering_plan=steps,
        )


_recordings: list[_Recording] = []


def _active() -> _Recording:
    if not _recordings:
        raise RuntimeError("RTDL operations must run inside @rt.kernel compilation")
    return _recordings[-1]


def kernel(*, backend: str, precision: str):
    """Decorate a kernel body; calling the result compiles it."""

    def decorate(body):
        @wraps(body)
        def compile_body():
            recording = _Recording(name=body.__name__, backend=backend, precision=precision)
            _recordings.append(recording)
            try:
                returned = body()
                if not isinstance(returned, EmitOp):
                    raise TypeError("kernel function must return rt.emit(...)")
                return recording.finish()
            finally:
                _recordings.pop()

        compile_body._rtdsl_kernel = True
        return compile_body

    return decorate


def input(  # noqa: A001 - the DSL name is part of the contract
    name: str,
    geometry: GeometryType,
    *,
    layout: Layout | None = None,
    role: str | None = None,
) -> GeometryInput:
    recording = _active()
    if role is not None and role not in {"build", "probe"}:
        raise ValueError("input role must be one of: build, probe")
    if any(item.name == name for item in recording.inputs):
        raise ValueError(f"duplicate input name in kernel: {name}")
    chosen = layout or geometry.default_layout
    if geometry.required_fields:
        chosen.require_fields(geometry.required_fields)
    node = GeometryInput(name=name, geometry=geometry, layout=chosen, role=role)
    recording.inputs.append(node)
    return node


def traverse(
    left: GeometryInput,
    right: GeometryInput,
    *,
    accel: str = "bvh",
    mode: str | None = None,
) -> CandidateSet:
    recording = _active()
    if mode is not None and mode not in {"graph_expand", "graph_intersect", "db_scan", "db_group"}:
        raise ValueError(
            "traverse mode must be one of: graph_expand, graph_intersect, db_scan, db_group"
        )
    node = CandidateSet(left=left, right=right, accel=accel, mode=mode)
    recording.candidates = node
    return node


def refine(candidates: CandidateSet, *, predicate: Predicate) -> RefineOp:
    recording = _active()
    node = RefineOp(candidates=candidates, predicate=predicate)
    recording.refine_op = node
    return node


def emit(source: RefineOp, *, fields: list[str]) -> EmitOp:
    recording = _active()
    node = EmitOp(source=source, fields=tuple(fields))
    recording.emit_op = node
    return node


def compile_kernel(fn) -> CompiledKernel:
    if not callable(fn):
        raise TypeError("compile_kernel expects a kernel function")
    return fn()


# --------------------------------------------------------------------------
# Predicates (reference: src/rtdsl/api.py:122-200)
# --------------------------------------------------------------------------


def point_in_polygon(
    *,
    exact: bool = False,
    boundary_mode: str = "inclusive",
    result_mode: str = "full_matrix",
) -> Predicate:
    if result_mode not in {"full_matrix", "positive_hits"}:
        raise ValueError("point_in_polygon result_mode must be 'full_matrix' or 'positive_hits'")
    return Predicate(
        name="point_in_polygon",
        options={"exact": exact, "boundary_mode": boundary_mode, "result_mode": result_mode},
    )


def contains(
    *,
    exact: bool = False,
    boundary_mode: str = "inclusive",
    result_mode: str = "full_matrix",
) -> Predicate:
    """Alias for point_in_polygon."""
    return point_in_polygon(exact=exact, boundary_mode=boundary_mode, result_mode=result_mode)


def ray_triangle_any_hit(*, exact: bool = False) -> Predicate:
    return Predicate(name="ray_triangle_any_hit", options={"exact": exact})


def ray_triangle_hit_count(*, exact: bool = False) -> Predicate:
    return Predicate(name="ray_triangle_hit_count", options={"exact": exact})


def ray_triangle_closest_hit(*, exact: bool = False) -> Predicate:
    return Predicate(name="ray_triangle_closest_hit", options={"exact": exact})


def fixed_radius_neighbors(*, radius: float, k_max: int) -> Predicate:
    if radius < 0.0:
        raise ValueError("fixed_radius_neighbors radius must be non-negative")
    if k_max <= 0:
        raise ValueError("fixed_radius_neighbors k_max must be positive")
    return Predicate(
        name="fixed_radius_neighbors", options={"radius": float(radius), "k_max": int(k_max)}
    )


def knn_rows(*, k: int) -> Predicate:
    if k <= 0:
        raise ValueError("knn_rows k must be positive")
    return Predicate(name="knn_rows", options={"k": int(k)})


def bounded_knn_rows(*, radius: float, k_max: int) -> Predicate:
    if radius < 0.0:
        raise ValueError("bounded_knn_rows radius must be non-negative")
    if k_max <= 0:
        raise ValueError("bounded_knn_rows k_max must be positive")
    return Predicate(
        name="bounded_knn_rows", options={"radius": float(radius), "k_max": int(k_max)}
    )
