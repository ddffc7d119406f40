"""Row-set comparison contract of the reference's parity harness.

Restates ``compare_baseline_rows`` / ``_row_sort_key`` (reference:
src/rtdsl/baseline_contracts.py:408-448) and its tolerances (:19-21): both
row sets are put in canonical order (rows sorted on all fields, floats
rounded to 9 digits), integer fields must match exactly and float fields
within ``isclose(rel_tol=1e-6, abs_tol=1e-6)`` (``exact=False``) or exactly
(``exact=True``, the PIP contract).
"""
from __future__ import annotations

from math import isclose

BASELINE_FLOAT_ABS_TOL = 1e-6
BASELINE_FLOAT_REL_TOL = 1e-6


def row_sort_key(row: dict[str, object]) -> tuple:
    key = []
    for name, value in sorted(row.items()):
        key.append((name, round(value, 9) if isinstance(value, float) else value))
    return tuple(key)


def compare_rows(expected, actual, *, exact: bool = False) -> bool:
    left = sorted(expected, key=row_sort_key)
    right = sorted(actual, key=row_sort_key)
    if len(left) != len(right):
        return False
    for a, b in zip(left, right):
        if set(a) != set(b):
            return False
        for name in a:
            x, y = a[name], b[name]
            if exact:
                if x != y:
                    return False
            elif isinstance(x, float) or isinstance(y, float):
                if not isclose(float(x), float(y), rel_tol=BASELINE_FLOAT_REL_TOL,
                               abs_tol=BASELINE_FLOAT_ABS_TOL):
                    return False
            elif x != y:
                return False
    return True
