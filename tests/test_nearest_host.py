"""Nearest on the CPU: the oracle's fp64 distances against exact rational
distances and closed forms, and the oracle's nearest selection against the
reference's own sindex.nearest vectors."""

from __future__ import annotations

import math
import random
from fractions import Fraction as F

import numpy as np
import pytest

import oracle
from oracle import fraction_ref as FR
from vibespatial_amd.geometry import PointArray, PolygonArray, box


def test_segment_distance_closed_forms():
    for ox, oy in [(0.0, 0.0), (5e5, 3e6)]:
        # perpendicular foot inside: (2, 6) to y = 1 from x = -1 .. 9 is 5
        assert oracle.segment_dist_f64(ox + 2, oy + 6, ox - 1, oy + 1, ox + 9, oy + 1) == 5.0
        # beyond the end: distance to (9, 1) from (12, 5) is 5
        assert oracle.segment_dist_f64(ox + 12, oy + 5, ox - 1, oy + 1, ox + 9, oy + 1) == 5.0
        # before the start, and a degenerate segment
        assert oracle.segment_dist_f64(ox - 4, oy - 3, ox - 1, oy + 1, ox + 9, oy + 1) == 5.0
        assert oracle.segment_dist_f64(ox + 3, oy + 4, ox, oy, ox, oy) == 5.0
    # oblique: (3, 4) to the segment (-8, 6)-(8, -6) through the origin
    assert oracle.segment_dist_f64(3.0, 4.0, -8.0, 6.0, 8.0, -6.0) == 5.0


def test_segment_distance_against_rationals():
    rng = random.Random(7)
    for _ in range(3000):
        s = rng.choice([1.0, 1e3, 5e5])
        ax, ay = rng.uniform(-1, 1) * s, rng.uniform(-1, 1) * s
        bx, by = ax + rng.uniform(-50, 50), ay + rng.uniform(-50, 50)
        px, py = ax + rng.uniform(-80, 80), ay + rng.uniform(-80, 80)
        d = float(oracle.segment_dist_f64(px, py, ax, ay, bx, by))
        exact = FR.dist2_point_segment(px, py, ax, ay, bx, by)
        # a few roundings of relative size 2^-53 on a well-conditioned value
        assert abs(F(d) * F(d) - exact) <= exact * F(1, 2 ** 44) + F(1, 2 ** 100), (px, py, ax, ay, bx, by)


def test_point_polygon_and_polygon_distances():
    P = PolygonArray.from_geometries([box(0, 0, 10, 10), box(20, 0, 30, 10), box(10, 13, 12, 15)])
    pts = PointArray(np.array([15.0, 5.0, 10.0, 13.0]), np.array([5.0, 5.0, 13.0, 14.0]))
    assert oracle.distance_f64(0, 1, pts, P, 0, 0) == 5.0
    assert oracle.distance_f64(0, 1, pts, P, 1, 0) == 0.0  # inside
    assert oracle.distance_f64(0, 1, pts, P, 2, 0) == 3.0
    assert oracle.distance_f64(0, 1, pts, P, 3, 0) == math.sqrt(9.0 + 16.0)  # corner (10, 10)
    assert oracle.distance_f64(1, 1, P, P, 0, 1) == 10.0
    assert oracle.distance_f64(1, 1, P, P, 0, 2) == 3.0
    assert oracle.distance_f64(1, 1, P, P, 0, 0) == 0.0
    assert oracle.distance_f64(0, 0, pts, pts, 0, 0, exclusive=True) == math.inf


def test_oracle_exclusive_polygons_is_topological_equality():
    """exclusive drops equal polygons whatever their vertex order (start vertex,
    orientation, a collinear extra vertex), and only those."""
    from vibespatial_amd.geometry import Polygon

    sq = [(0, 0), (10, 0), (10, 10), (0, 10)]
    P = PolygonArray.from_geometries([
        Polygon(sq),
        Polygon(sq[2:] + sq[:2]),            # other start vertex
        Polygon(sq[::-1]),                   # clockwise
        Polygon([(0, 0), (5, 0), (10, 0), (10, 10), (0, 10)]),  # collinear extra vertex
        box(0, 0, 10, 10.5),                 # not equal: larger
        box(0, 0, 9, 10),                    # not equal: smaller, inside
        box(20, 0, 30, 10),                  # apart
    ])
    for j in range(4):
        assert oracle.distance_f64(1, 1, P, P, 0, j, exclusive=True) == math.inf
        assert oracle.distance_f64(1, 1, P, P, 0, j) == 0.0
    assert oracle.distance_f64(1, 1, P, P, 0, 4, exclusive=True) == 0.0
    assert oracle.distance_f64(1, 1, P, P, 0, 5, exclusive=True) == 0.0
    assert oracle.distance_f64(1, 1, P, P, 0, 6, exclusive=True) == 10.0
    l, r, d = oracle.nearest_pairs(1, 1, P, P, exclusive=True)
    assert list(zip(l.tolist(), r.tolist())) == [(0, 4), (0, 5), (1, 4), (1, 5), (2, 4), (2, 5), (3, 4), (3, 5),
                                                 (4, 0), (4, 1), (4, 2), (4, 3), (4, 5), (5, 0), (5, 1), (5, 2),
                                                 (5, 3), (5, 4), (6, 0), (6, 1), (6, 2), (6, 3), (6, 4)]


_GRID = PointArray(np.arange(10.0), np.arange(10.0))


@pytest.mark.parametrize("return_all", [True, False])
@pytest.mark.parametrize("xy,expected", [
    # tests/upstream/geopandas/tests/test_sindex.py:806-824
    ([(0.25, 0.25)], [[0], [0]]),
    ([(0.75, 0.75)], [[0], [1]]),
    # :826-855
    ([(1, 1), (0, 0)], [[0, 1], [1, 0]]),
    ([(1, 1), (0.25, 1)], [[0, 1], [1, 1]]),
])
def test_oracle_reference_nearest_vectors(xy, expected, return_all):
    q = PointArray(np.array([p[0] for p in xy], float), np.array([p[1] for p in xy], float))
    l, r, _ = oracle.nearest_pairs(0, 0, q, _GRID, return_all=return_all)
    assert [l.tolist(), r.tolist()] == expected


@pytest.mark.parametrize("return_all,max_distance,expected", [
    # tests/upstream/geopandas/tests/test_sindex.py:871-880
    (True, None, ([[0, 0, 1], [0, 1, 5]], [math.sqrt(0.5), math.sqrt(0.5), math.sqrt(50)])),
    (False, None, ([[0, 1], [0, 5]], [math.sqrt(0.5), math.sqrt(50)])),
    (True, 1, ([[0, 0], [0, 1]], [math.sqrt(0.5), math.sqrt(0.5)])),
    (False, 1, ([[0], [0]], [math.sqrt(0.5)])),
])
def test_oracle_reference_max_distance_vectors(return_all, max_distance, expected):
    q = PointArray(np.array([0.5, 0.0]), np.array([0.5, 10.0]))
    l, r, d = oracle.nearest_pairs(0, 0, q, _GRID, max_distance=max_distance, return_all=return_all)
    assert [l.tolist(), r.tolist()] == expected[0]
    assert d.tolist() == expected[1]


@pytest.mark.parametrize("return_all,max_distance,exclusive,expected", [
    # tests/upstream/geopandas/tests/test_sindex.py:908-916
    (False, None, False, ([[0, 1, 2, 3, 4], [0, 1, 2, 3, 4]], 5 * [0])),
    (False, None, True, ([[0, 1, 2, 3, 4], [1, 0, 1, 2, 3]], 5 * [math.sqrt(2)])),
    (True, None, False, ([[0, 1, 2, 3, 4], [0, 1, 2, 3, 4]], 5 * [0])),
    (True, None, True, ([[0, 1, 1, 2, 2, 3, 3, 4], [1, 0, 2, 1, 3, 2, 4, 3]], 8 * [math.sqrt(2)])),
    (False, 1.1, True, ([[1, 2, 5], [5, 5, 1]], 3 * [1])),
    (True, 1.1, True, ([[1, 2, 5, 5], [5, 5, 1, 2]], 4 * [1])),
])
def test_oracle_reference_exclusive_vectors(return_all, max_distance, exclusive, expected):
    x, y = np.arange(5.0), np.arange(5.0)
    if max_distance:
        x, y = np.append(x, 1.0), np.append(y, 2.0)
    g = PointArray(x, y)
    l, r, d = oracle.nearest_pairs(0, 0, g, g, max_distance=max_distance, return_all=return_all,
                                   exclusive=exclusive)
    assert [l.tolist(), r.tolist()] == expected[0]
    assert d.tolist() == expected[1]
