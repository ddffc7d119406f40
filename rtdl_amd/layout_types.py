"""Geometry types and field layouts of the RTDL kernel contract.

Restates the public layout vocabulary of the reference DSL
(reference: src/rtdsl/layout_types.py:63-229) for the geometry families the
MI355X backend serves: 2-D/3-D points, polygon references, 2-D/3-D triangles
and 2-D/3-D rays.  Names, field orders and scalar types are kept identical so
that kernels written against the reference compile unchanged.
"""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class ScalarType:
    name: str
    c_type: str
    cuda_type: str
    size: int


@dataclass(frozen=True)
class Field:
    name: str
    dtype: ScalarType

    def to_dict(self) -> dict[str, object]:
        return {
            "name": self.name,
            "scalar_type": self.dtype.name,
            "c_type": self.dtype.c_type,
            "cuda_type": self.dtype.cuda_type,
            "size": self.dtype.size,
        }


@dataclass(frozen=True)
class Layout:
    name: str
    fields: tuple[Field, ...]

    def field_names(self) -> tuple[str, ...]:
        return tuple(item.name for item in self.fields)

    def require_fields(self, names: tuple[str, ...]) -> None:
        present = set(self.field_names())
        absent = [name for name in names if name not in present]
        if absent:
            raise ValueError(
                f"layout `{self.name}` is missing required fields: {', '.join(absent)}"
            )


@dataclass(frozen=True)
class GeometryType:
    name: str
    dimension: int
    default_layout: Layout
    required_fields: tuple[str, ...] = ()


def field(name: str, dtype: ScalarType) -> Field:
    return Field(name=name, dtype=dtype)


def layout(name: str, *fields_in_order: Field) -> Layout:
    if not fields_in_order:
        raise ValueError("layout must define at least one field")
    return Layout(name=name, fields=tuple(fields_in_order))


# The C-side scalar names are part of the plan schema (ir.RTExecutionPlan.to_dict).
f32 = ScalarType(name="f32", c_type="float", cuda_type="float", size=4)
u32 = ScalarType(name="u32", c_type="uint32_t", cuda_type="uint32_t", size=4)


def _f32_fields(*names: str) -> tuple[Field, ...]:
    return tuple(field(name, f32) for name in names)


Segment2DLayout = layout("Segment2D", *_f32_fields("x0", "y0", "x1", "y1"), field("id", u32))
Point2DLayout = layout("Point2D", *_f32_fields("x", "y"), field("id", u32))
Point3DLayout = layout("Point3D", *_f32_fields("x", "y", "z"), field("id", u32))
Polygon2DLayout = layout(
    "Polygon2DRef",
    field("vertex_offset", u32),
    field("vertex_count", u32),
    field("id", u32),
)
Triangle2DLayout = layout(
    "Triangle2D", *_f32_fields("x0", "y0", "x1", "y1", "x2", "y2"), field("id", u32)
)
Triangle3DLayout = layout(
    "Triangle3D",
    *_f32_fields("x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2"),
    field("id", u32),
)
Ray2DLayout = layout("Ray2D", *_f32_fields("ox", "oy", "dx", "dy", "tmax"), field("id", u32))
Ray3DLayout = layout(
    "Ray3D", *_f32_fields("ox", "oy", "oz", "dx", "dy", "dz", "tmax"), field("id", u32)
)


def _geometry(name: str, dimension: int, default: Layout) -> GeometryType:
    # Every field of the default layout is required, as in the reference.
    return GeometryType(
        name=name,
        dimension=dimension,
        default_layout=default,
        required_fields=default.field_names(),
    )


Segments = _geometry("segments", 1, Segment2DLayout)
Points = _geometry("points", 0, Point2DLayout)
Points3D = _geometry("points", 0, Point3DLayout)
Polygons = GeometryType(
    name="polygons",
    dimension=2,
    default_layout=Polygon2DLayout,
    required_fields=("vertex_offset", "vertex_count", "id"),
)
Triangles = _geometry("triangles", 2, Triangle2DLayout)
Triangles3D = _geometry("triangles", 2, Triangle3DLayout)
Rays = _geometry("rays", 1, Ray2DLayout)
Rays3D = _geometry("rays", 1, Ray3DLayout)
