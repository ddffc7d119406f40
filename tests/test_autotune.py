"""CPU tests of the autotuner's search (no library calls)."""
from __future__ import annotations

from rtdl_amd.autotune import KNOBS, Tuning, coordinate_descent


def test_coordinate_descent_finds_the_separable_minimum_and_evaluates_each_tuning_once():
    calls = []

    def cost(t: Tuning) -> float:
        calls.append(t)
        return ((t.blocks_per_cu - 4) ** 2 + (t.point_leaf - 6) ** 2 + (t.triangle_leaf - 1) ** 2
                + (t.polygon_leaf - 2) ** 2)

    best, table = coordinate_descent(cost, Tuning())
    assert best == Tuning(point_leaf=6, triangle_leaf=1, polygon_leaf=2, blocks_per_cu=4)
    assert len(calls) == len(set(calls)) == len(table)
    assert min(table, key=lambda r: r["seconds"])["seconds"] == 0
    # the start plus each knob's other values
    assert len(table) == 1 + sum(len(v) for v in KNOBS.values()) - len(KNOBS)


def test_coordinate_descent_keeps_the_start_when_nothing_is_faster():
    best, table = coordinate_descent(lambda t: 0.0 if t == Tuning() else 1.0, Tuning(),
                                     {"blocks_per_cu": (2, 16)})
    assert best == Tuning()
    assert {r["blocks_per_cu"] for r in table} == {2, 8, 16}
