"""Polygon x polygon relations on the CPU: the rational oracle against the
interval formulas of axis-aligned boxes (many exact touches and shared
edges), hand-built cases with holes and multipolygons, and the reference's
own sindex.query vectors for polygon predicates."""

from __future__ import annotations

import itertools
import random

import pytest

from oracle import relate_ref as R
from vibespatial_amd.geometry import MultiPolygon, Polygon, PolygonArray, box

PREDS = ("intersects", "within", "contains", "covers", "covered_by", "contains_properly", "touches", "overlaps",
         "crosses")


def _box_truth(a, b):
    ax0, ay0, ax1, ay1 = a
    bx0, by0, bx1, by1 = b
    inter = ax0 <= bx1 and bx0 <= ax1 and ay0 <= by1 and by0 <= ay1
    meet = ax0 < bx1 and bx0 < ax1 and ay0 < by1 and by0 < ay1
    a_in_b = bx0 <= ax0 and ax1 <= bx1 and by0 <= ay0 and ay1 <= by1
    b_in_a = ax0 <= bx0 and bx1 <= ax1 and ay0 <= by0 and by1 <= ay1
    b_in_int_a = ax0 < bx0 and bx1 < ax1 and ay0 < by0 and by1 < ay1
    return {"intersects": inter, "within": a_in_b, "covered_by": a_in_b, "contains": b_in_a, "covers": b_in_a,
            "contains_properly": b_in_int_a, "touches": inter and not meet,
            "overlaps": meet and not a_in_b and not b_in_a, "crosses": False}


def test_boxes_against_interval_formulas():
    rng = random.Random(5)
    boxes = []
    for _ in range(28):
        x0, y0 = rng.randint(0, 6), rng.randint(0, 6)
        boxes.append((x0, y0, x0 + rng.randint(1, 4), y0 + rng.randint(1, 4)))
    P = PolygonArray.from_geometries([box(*b) for b in boxes])
    for i, j in itertools.product(range(len(boxes)), repeat=2):
        truth = _box_truth(boxes[i], boxes[j])
        for pred in PREDS:
            assert R.predicate(pred, P, i, P, j) == truth[pred], (boxes[i], boxes[j], pred)


def _one(*geoms):
    return PolygonArray.from_geometries(list(geoms))


def test_holes_and_multipolygons():
    donut = Polygon([(0, 0), (10, 0), (10, 10), (0, 10), (0, 0)], [[(3, 3), (7, 3), (7, 7), (3, 7), (3, 3)]])
    hole_fill = box(3, 3, 7, 7)
    inside_hole = box(4, 4, 6, 6)
    in_ring = box(0.5, 0.5, 2, 2)
    big = box(-1, -1, 11, 11)
    P = _one(donut, hole_fill, inside_hole, in_ring, big)
    # the filled hole shares the hole ring: touches, not within
    assert R.predicate("touches", P, 1, P, 0) and not R.predicate("within", P, 1, P, 0)
    assert not R.predicate("intersects", P, 2, P, 0)  # strictly inside the hole
    assert R.predicate("within", P, 3, P, 0) and R.predicate("contains_properly", P, 0, P, 3)
    assert R.predicate("within", P, 0, P, 4) and not R.predicate("within", P, 4, P, 0)
    # the big box contains the donut but the donut does not contain the big box
    assert R.predicate("contains", P, 4, P, 0) and R.predicate("overlaps", P, 4, P, 0) is False
    # a box over part of the hole and part of the ring overlaps the donut
    Q = _one(box(2, 2, 5, 5))
    assert R.predicate("overlaps", Q, 0, P, 0) and not R.predicate("within", Q, 0, P, 0)
    # identical polygons: within both ways, not properly, no overlap, no touch
    assert R.predicate("within", P, 0, P, 0) and R.predicate("contains", P, 0, P, 0)
    assert not R.predicate("contains_properly", P, 0, P, 0) and not R.predicate("overlaps", P, 0, P, 0)
    assert not R.predicate("touches", P, 0, P, 0)
    # a multipolygon whose parts touch the hole filler on two sides
    M = _one(MultiPolygon([box(0, 3, 3, 7), box(7, 3, 10, 7)]), hole_fill, box(0, 3, 10, 7))
    assert R.predicate("touches", M, 0, M, 1)
    assert R.predicate("within", M, 0, M, 2) and R.predicate("within", M, 1, M, 2)
    assert not R.predicate("within", M, 2, M, 0)
    # a triangle sharing a diagonal edge with a square's half
    T = _one(Polygon([(0, 0), (4, 0), (4, 4), (0, 0)]), box(0, 0, 4, 4), Polygon([(0, 0), (4, 4), (0, 4), (0, 0)]))
    assert R.predicate("within", T, 0, T, 1) and R.predicate("touches", T, 0, T, 2)
    assert not R.predicate("overlaps", T, 0, T, 2) and R.predicate("within", T, 2, T, 1)


@pytest.mark.parametrize("pred,query,expected", [
    # tests/upstream/geopandas/tests/test_sindex.py:297-369, the polygon rows
    # of the tree (box(10, 10, 20, 20)) against polygon queries
    ("within", (0, 0, 10, 10), False),    # touches at the corner only
    ("within", (11, 11, 12, 12), True),
    ("contains", (10, 10, 20, 20), True),
    ("contains_properly", (10, 10, 20, 20), False),
    ("touches", (0, 0, 10, 10), True),
    ("intersects", (0, 0, 10, 10), True),
])
def test_reference_query_vectors(pred, query, expected):
    P = _one(box(*query), box(10, 10, 20, 20))
    assert R.predicate(pred, P, 0, P, 1) == expected
