"""Logical input records, their normalisation, and output-row projection.

The record types restate the reference's frozen dataclasses
(reference: src/rtdsl/reference.py:7-81).  ``normalize_records`` restates the
reference input coercion (reference: src/rtdsl/runtime.py:263-457): each input
is an iterable of records given either as mappings or as attribute objects;
points/triangles/rays are 3-D exactly when the 3-D field set is present;
polygons carry an ``id`` and at least three 2-D ``vertices``.
``project_rows`` restates reference: src/rtdsl/runtime.py:249-260.
"""
from __future__ import annotations

from collections.abc import Iterable, Mapping
from dataclasses import dataclass


@dataclass(frozen=True)
class Point:
    id: int
    x: float
    y: float


@dataclass(frozen=True)
class Point3D:
    id: int
    x: float
    y: float
    z: float


@dataclass(frozen=True)
class Polygon:
    id: int
    vertices: tuple[tuple[float, float], ...]


@dataclass(frozen=True)
class Triangle:
    id: int
    x0: float
    y0: float
    x1: float
    y1: float
    x2: float
    y2: float


@dataclass(frozen=True)
class Triangle3D:
    id: int
    x0: float
    y0: float
    z0: float
    x1: float
    y1: float
    z1: float
    x2: float
    y2: float
    z2: float


@dataclass(frozen=True)
class Ray2D:
    id: int
    ox: float
    oy: float
    dx: float
    dy: float
    tmax: float


@dataclass(frozen=True)
class Ray3D:
    id: int
    ox: float
    oy: float
    oz: float
    dx: float
    dy: float
    dz: float
    tmax: float


# (3-D record type, its fields), (2-D record type, its fields) per geometry.
_FIELD_SETS = {
    "points": ((Point3D, ("id", "x", "y", "z")), (Point, ("id", "x", "y"))),
    "triangles": (
        (Triangle3D, ("id", "x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2")),
        (Triangle, ("id", "x0", "y0", "x1", "y1", "x2", "y2")),
    ),
    "rays": (
        (Ray3D, ("id", "ox", "oy", "oz", "dx", "dy", "dz", "tmax")),
        (Ray2D, ("id", "ox", "oy", "dx", "dy", "tmax")),
    ),
}


def _has_fields(record, names) -> bool:
    if isinstance(record, Mapping):
        return all(name in record for name in names)
    return all(hasattr(record, name) for name in names)


def _fields_of(input_name: str, record, names, *, hint: str | None = None) -> dict:
    if isinstance(record, Mapping):
        absent = [name for name in names if name not in record]
        if absent:
            message = f"simulator input `{input_name}` record is missing fields: {', '.join(absent)}"
            raise ValueError(f"{message}. {hint}" if hint else message)
        return {name: record[name] for name in names}
    if all(hasattr(record, name) for name in names):
        return {name: getattr(record, name) for name in names}
    message = f"simulator input `{input_name}` record must be a mapping or dataclass-like object"
    raise ValueError(f"{message}. {hint}" if hint else message)


def _coerce_flat(input_name: str, geometry_name: str, record):
    (type3, names3), (type2, names2) = _FIELD_SETS[geometry_name]
    if isinstance(record, (type3, type2)):
        return record
    record_type, names = (type3, names3) if _has_fields(record, names3) else (type2, names2)
    data = _fields_of(input_name, record, names)
    values = {name: (int(data[name]) if name == "id" else float(data[name])) for name in names}
    return record_type(**values)


def _coerce_polygon(input_name: str, record) -> Polygon:
    if isinstance(record, Polygon):
        return record
    data = _fields_of(
        input_name,
        record,
        ("id", "vertices"),
        hint="CPU simulator polygons must provide logical polygon records with `id` and `vertices`.",
    )
    raw = data["vertices"]
    if isinstance(raw, (str, bytes)) or not isinstance(raw, Iterable):
        raise ValueError(f"simulator polygon input `{input_name}` must provide iterable `vertices`")
    vertices = []
    for vertex in raw:
        if isinstance(vertex, (str, bytes)) or not isinstance(vertex, Iterable):
            raise ValueError(f"simulator polygon input `{input_name}` has invalid vertex: {vertex!r}")
        coords = tuple(vertex)
        if len(coords) != 2:
            raise ValueError(f"simulator polygon input `{input_name}` vertex must have 2 coordinates")
        vertices.append((float(coords[0]), float(coords[1])))
    if len(vertices) < 3:
        raise ValueError(f"simulator polygon input `{input_name}` requires at least 3 vertices")
    return Polygon(id=int(data["id"]), vertices=tuple(vertices))


def normalize_records(input_name: str, geometry_name: str, payload):
    """Coerce one kernel input into a tuple of frozen records."""
    if isinstance(payload, (str, bytes)) or not isinstance(payload, Iterable):
        raise ValueError(f"simulator input `{input_name}` must be an iterable of records")
    if geometry_name == "polygons":
        return tuple(_coerce_polygon(input_name, item) for item in payload)
    if geometry_name in _FIELD_SETS:
        return tuple(_coerce_flat(input_name, geometry_name, item) for item in payload)
    raise ValueError(f"unsupported RTDL simulator geometry: {geometry_name}")


def project_rows(compiled, rows) -> tuple[dict[str, object], ...]:
    """Keep exactly the emitted fields of each row, in emit order."""
    emit_fields = compiled.emit_op.fields
    projected = []
    for row in rows:
        absent = [name for name in emit_fields if name not in row]
        if absent:
            raise ValueError("CPU simulator row is missing emitted fields: " + ", ".join(absent))
        projected.append({name: row[name] for name in emit_fields})
    return tuple(projected)
