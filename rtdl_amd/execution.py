"""Contract-first execution facade: ``rt.run(kernel, inputs, execution=...)``.

Restates the reference facade (reference: src/rtdsl/execution.py:16-316):
``ExecutionPolicy``, ``ExecutionReport``, ``ExecutionResult`` and ``run``.
The reference's dispatch is a closed if-chain over backend names
(:169-204); here backends are looked up in a registry in which ``mi355x`` is
the one built-in backend.  Other names (``cpu_python_reference`` and so on)
are accepted by the policy so reference code constructs unchanged, but they
execute only when a runner has been registered for them with
``register_backend`` — the package itself ships no CPU execution path, and
``allow_fallback`` therefore has nothing to fall back to unless a runner for
``cpu_python_reference`` was registered.
"""
from __future__ import annotations

import os
import platform
import subprocess
from dataclasses import asdict, dataclass, field
from typing import Any, Callable, Mapping

from .api import compile_kernel
from .ir import CompiledKernel

_SUPPORTED_BACKENDS = (
    "auto",
    "cpu_python_reference",
    "cpu",
    "embree",
    "optix",
    "vulkan",
    "hiprt",
    "apple_rt",
    "mi355x",
)

_RUNNERS: dict[str, Callable[..., tuple]] = {}


def register_backend(name: str, runner: Callable[..., tuple]) -> None:
    """Register ``runner(kernel, result_mode=..., **inputs) -> rows`` for `name`."""
    if name not in _SUPPORTED_BACKENDS or name == "auto":
        raise ValueError("backend must be one of: " + ", ".join(_SUPPORTED_BACKENDS[1:]))
    _RUNNERS[name] = runner


def registered_backends() -> tuple[str, ...]:
    return tuple(sorted(_RUNNERS))


def _run_mi355x(kernel_fn_or_compiled, *, result_mode: str = "dict", **inputs):
    from .mi355x_runtime import run_mi355x

    return run_mi355x(kernel_fn_or_compiled, result_mode=result_mode, **inputs)


_RUNNERS["mi355x"] = _run_mi355x


@dataclass(frozen=True)
class ExecutionPolicy:
    """Explicit execution request for the contract-first RTDL runner."""

    backend: str = "mi355x"
    partner: str | None = None
    allow_fallback: bool = False
    require_rt_core: bool = False
    explain: bool = True
    result_mode: str = "dict"

    def __post_init__(self) -> None:
        if self.backend not in _SUPPORTED_BACKENDS:
            raise ValueError("backend must be one of: " + ", ".join(_SUPPORTED_BACKENDS))
        if self.partner is not None and self.partner not in {"numpy", "cupy", "torch"}:
            raise ValueError("partner must be one of: numpy, cupy, torch")
        if self.require_rt_core and self.backend not in {"auto", "optix"}:
            raise ValueError(
                "require_rt_core is only supported with backend='auto' or backend='optix'"
            )


@dataclass(frozen=True)
class ExecutionReport:
    """Reproducible explanation of one RTDL execution decision."""

    requested_backend: str
    selected_backend: str
    requested_partner: str | None
    selected_partner: str | None
    fallback_backend: str | None
    fallback_reason: str | None
    primitive_family: str
    predicate: str | None
    output_schema: tuple[str, ...]
    exact_mode: str
    memory_status: str
    copy_status: str
    rt_core_status: str
    cuda_core_partner_status: str
    cpu_status: str
    claim_boundary: Mapping[str, bool]
    reproducibility: Mapping[str, str]
    notes: tuple[str, ...] = field(default_factory=tuple)

    def to_dict(self) -> dict[str, Any]:
        payload = asdict(self)
        payload["output_schema"] = list(self.output_schema)
        payload["notes"] = list(self.notes)
        payload["claim_boundary"] = dict(self.claim_boundary)
        payload["reproducibility"] = dict(self.reproducibility)
        return payload


@dataclass(frozen=True)
class ExecutionResult:
    """Rows plus the report that explains how they were produced."""

    rows: tuple[dict[str, object], ...]
    execution_report: ExecutionReport

    def to_dict(self) -> dict[str, object]:
        return {"rows": list(self.rows), "execution_report": self.execution_report.to_dict()}


def run(
    kernel_fn_or_compiled,
    inputs: Mapping[str, object] | None = None,
    *,
    execution: ExecutionPolicy | None = None,
    **input_kwargs,
) -> ExecutionResult:
    """Run a kernel through an explicit, explainable execution policy."""
    policy = execution or ExecutionPolicy()
    merged = dict(inputs or {})
    merged.update(input_kwargs)
    compiled = _compiled_or_none(kernel_fn_or_compiled)
    selected = _select_backend(policy)
    fallback_backend = None
    fallback_reason = None
    notes: list[str] = []
    try:
        rows = _dispatch(selected, kernel_fn_or_compiled, policy, merged)
    except Exception as exc:
        if (
            not policy.allow_fallback
            or selected == "cpu_python_reference"
            or "cpu_python_reference" not in _RUNNERS
        ):
            raise
        fallback_backend = "cpu_python_reference"
        fallback_reason = f"{selected} failed: {type(exc).__name__}: {exc}"
        notes.append("Execution used the registered CPU Python reference fallback.")
        rows = _dispatch(fallback_backend, kernel_fn_or_compiled, policy, merged)
    report = _build_report(
        policy=policy,
        compiled=compiled,
        report_backend=fallback_backend or selected,
        fallback_backend=fallback_backend,
        fallback_reason=fallback_reason,
        notes=tuple(notes),
    )
    if policy.result_mode == "raw" and not isinstance(rows, tuple):
        return ExecutionResult(rows=rows, execution_report=report)
    return ExecutionResult(rows=tuple(rows), execution_report=report)


def _compiled_or_none(kernel_fn_or_compiled) -> CompiledKernel | None:
    try:
        if isinstance(kernel_fn_or_compiled, CompiledKernel):
            return kernel_fn_or_compiled
        return compile_kernel(kernel_fn_or_compiled)
    except Exception:
        return None


def _select_backend(policy: ExecutionPolicy) -> str:
    if policy.backend == "auto":
        return "mi355x"
    return policy.backend


def _dispatch(backend: str, kernel_fn_or_compiled, policy: ExecutionPolicy, inputs):
    runner = _RUNNERS.get(backend)
    if runner is None:
        raise ValueError(
            f"backend {backend!r} has no registered runner in this package "
            f"(registered: {', '.join(registered_backends())})"
        )
    return runner(kernel_fn_or_compiled, result_mode=policy.result_mode, **dict(inputs))


def _primitive_family(predicate: str) -> str:
    if predicate in {"ray_triangle_any_hit", "ray_triangle_hit_count", "ray_triangle_closest_hit"}:
        return "traversal"
    if predicate in {"fixed_radius_neighbors", "knn_rows", "bounded_knn_rows"}:
        return "nearest_neighbor"
    if predicate == "point_in_polygon":
        return "spatial_relation"
    return "generic_refine"


def _exact_mode(options: Mapping[str, object]) -> str:
    if "exact" in options:
        return "exact" if bool(options["exact"]) else "float_approx"
    return "contract_defined"


def _build_report(*, policy, compiled, report_backend, fallback_backend, fallback_reason, notes):
    predicate = None
    schema: tuple[str, ...] = ()
    exact_mode = "unknown"
    family = "unknown"
    if compiled is not None and compiled.refine_op is not None:
        predicate = compiled.refine_op.predicate.name
        family = _primitive_family(predicate)
        exact_mode = _exact_mode(compiled.refine_op.predicate.options)
    if compiled is not None and compiled.emit_op is not None:
        schema = tuple(compiled.emit_op.fields)
    partner = policy.partner
    partner_status = "not_requested"
    if partner in {"cupy", "torch"}:
        partner_status = f"{partner}_requested_not_proven_by_runner"
    elif partner == "numpy":
        partner_status = "numpy_cpu_partner_requested"
    return ExecutionReport(
        requested_backend=policy.backend,
        selected_backend=report_backend,
        requested_partner=policy.partner,
        selected_partner=partner,
        fallback_backend=fallback_backend,
        fallback_reason=fallback_reason,
        primitive_family=family,
        predicate=predicate,
        output_schema=schema,
        exact_mode=exact_mode,
        memory_status="device_resident_lbvh" if report_backend == "mi355x" else "not_reported_by_runtime",
        copy_status="host_rows_copied_back" if report_backend == "mi355x" else "not_reported_by_runtime",
        rt_core_status="not_used",
        cuda_core_partner_status=partner_status,
        cpu_status="selected" if report_backend in {"cpu", "cpu_python_reference", "embree"} else "not_selected",
        claim_boundary={
            "rt_core_speedup_claim_authorized": False,
            "whole_app_speedup_claim_authorized": False,
            "zero_copy_claim_authorized": False,
            "fallback_used": fallback_backend is not None,
        },
        reproducibility=_reproducibility(),
        notes=notes,
    )


def _reproducibility() -> dict[str, str]:
    commit = "unknown"
    try:
        commit = subprocess.check_output(
            ["git", "rev-parse", "--short", "HEAD"],
            text=True,
            stderr=subprocess.DEVNULL,
            cwd=os.path.dirname(os.path.abspath(__file__)),
        ).strip()
    except Exception:
        pass
    return {
        "git_commit": commit,
        "python": platform.python_version(),
        "platform": platform.platform(),
        "RTDL_MI355X_LIB": os.environ.get("RTDL_MI355X_LIB", ""),
    }


__all__ = [
    "ExecutionPolicy",
    "ExecutionReport",
    "ExecutionResult",
    "register_backend",
    "registered_backends",
    "run",
]
