"""dwithin decisions on the CPU: the exact-rational oracle against closed forms,
and the library's host-compiled exact predicates (the same code the GPU
refine runs) against the oracle on near-tie inputs."""

import math
import random
from fractions import Fraction as F

import numpy as np
import pytest

from oracle import fraction_ref as FR
from vibespatial_amd import _lib


@pytest.fixture(scope="module")
def L():
    return _lib.lib()


def test_oracle_distances_closed_forms():
    # 3-4-5 triangles at several offsets (exactly representable)
    for ox, oy in [(0.0, 0.0), (5e5, 3e6), (-1e4, 2.5e3)]:
        assert FR.dist2_point_point(ox, oy, ox + 3, oy + 4) == 25
        # axis-aligned segment: perpendicular distance |py - ay| inside the span
        assert FR.dist2_point_segment(ox + 2, oy + 7, ox, oy + 1, ox + 10, oy + 1) == 36
        # beyond an end point: distance to that end point
        assert FR.dist2_point_segment(ox + 13, oy + 5, ox, oy + 1, ox + 10, oy + 1) == 9 + 16
        assert FR.dist2_point_segment(ox - 3, oy - 3, ox, oy + 1, ox + 10, oy + 1) == 9 + 16
        # degenerate segment = point
        assert FR.dist2_point_segment(ox + 3, oy + 4, ox, oy, ox, oy) == 25
    # diagonal segment y = x, point (0, 2): distance sqrt(2)
    assert FR.dist2_point_segment(0.0, 2.0, -5.0, -5.0, 5.0, 5.0) == 2
    assert FR.within_distance(F(25), 5.0) and not FR.within_distance(F(25), math.nextafter(5.0, 0.0))
    assert not FR.within_distance(F(0), float("nan")) and not FR.within_distance(F(0), -1.0)
    assert FR.within_distance(F(10) ** 40, float("inf"))


def test_oracle_segment_distance_matches_line_form():
    # independent form: min(|p-a|^2, |p-b|^2, and cross^2 / |u|^2 when the
    # foot of the perpendicular lies inside the segment)
    rng = random.Random(3)
    for _ in range(400):
        a = (rng.randint(-50, 50) / 4, rng.randint(-50, 50) / 4)
        b = (rng.randint(-50, 50) / 4, rng.randint(-50, 50) / 4)
        p = (rng.randint(-80, 80) / 8, rng.randint(-80, 80) / 8)
        ux, uy = F(b[0]) - F(a[0]), F(b[1]) - F(a[1])
        wx, wy = F(p[0]) - F(a[0]), F(p[1]) - F(a[1])
        cands = [FR.dist2_point_point(*p, *a), FR.dist2_point_point(*p, *b)]
        l2 = ux * ux + uy * uy
        if l2 and 0 <= ux * wx + uy * wy <= l2:
            cands.append((ux * wy - uy * wx) ** 2 / l2)
        assert FR.dist2_point_segment(*p, *a, *b) == min(cands)


def _near_tie_cases(rng, n):
    for _ in range(n):
        s = rng.choice([1.0, 1e5, 5e5, 3e6])
        ax, ay = rng.uniform(-1, 1) * s, rng.uniform(-1, 1) * s
        bx, by = ax + rng.uniform(-100, 100), ay + rng.uniform(-100, 100)
        if rng.random() < 0.1:
            bx, by = ax, ay
        t = rng.uniform(-0.5, 1.5)
        px = ax + t * (bx - ax) + rng.uniform(-50, 50)
        py = ay + t * (by - ay) + rng.uniform(-50, 50)
        dd = math.sqrt(float(FR.dist2_point_segment(px, py, ax, ay, bx, by)))
        r = rng.random()
        if r < 0.3:
            d = math.nextafter(dd, math.inf)
        elif r < 0.6:
            d = math.nextafter(dd, -math.inf)
        elif r < 0.8:
            d = dd
        else:
            d = dd * (1 + rng.choice([1e-15, -1e-15, 1e-9, -1e-9]))
        yield px, py, ax, ay, bx, by, d


def test_host_point_segment_within_matches_oracle(L):
    rng = random.Random(11)
    for px, py, ax, ay, bx, by, d in _near_tie_cases(rng, 6000):
        exp = FR.within_distance(FR.dist2_point_segment(px, py, ax, ay, bx, by), d)
        assert bool(L.vs_host_point_segment_within(px, py, ax, ay, bx, by, d)) == exp, (px, py, ax, ay, bx, by, d)


def test_host_point_point_within_matches_oracle(L):
    rng = random.Random(12)
    for px, py, ax, ay, _, _, d in _near_tie_cases(rng, 6000):
        exp = FR.within_distance(FR.dist2_point_point(px, py, ax, ay), d)
        assert bool(L.vs_host_point_point_within(px, py, ax, ay, d)) == exp


def test_host_exact_ties_and_special_values(L):
    below = math.nextafter(5.0, 0.0)
    for ox, oy in [(0.0, 0.0), (5e5, 3e6)]:
        assert L.vs_host_point_point_within(ox, oy, ox + 3, oy + 4, 5.0) == 1
        assert L.vs_host_point_point_within(ox, oy, ox + 3, oy + 4, below) == 0
        # perpendicular foot inside the segment, distance exactly 5
        assert L.vs_host_point_segment_within(ox + 2, oy + 6, ox - 1, oy + 1, ox + 9, oy + 1, 5.0) == 1
        assert L.vs_host_point_segment_within(ox + 2, oy + 6, ox - 1, oy + 1, ox + 9, oy + 1, below) == 0
        # oblique segment (-8, 6)-(8, -6) through the origin: p = (3, 4) is
        # perpendicular to it ((3, 4).(8, -6) = 0), foot at the origin, distance 5
        assert L.vs_host_point_segment_within(ox + 3, oy + 4, ox - 8, oy + 6, ox + 8, oy - 6, 5.0) == 1
        assert L.vs_host_point_segment_within(ox + 3, oy + 4, ox - 8, oy + 6, ox + 8, oy - 6, below) == 0
    nan, inf = float("nan"), float("inf")
    assert L.vs_host_point_point_within(0, 0, 1, 1, nan) == 0
    assert L.vs_host_point_point_within(0, 0, 1, 1, -1.0) == 0
    assert L.vs_host_point_point_within(nan, 0, 1, 1, 10.0) == 0
    assert L.vs_host_point_point_within(0, 0, 1e300, 1, inf) == 1
    assert L.vs_host_point_point_within(0, 0, 0, 0, 0.0) == 1
    assert L.vs_host_point_segment_within(0, 0, 1, 1, 2, 2, -0.5) == 0
    assert L.vs_host_point_segment_within(1.5, 1.5, 1, 1, 2, 2, 0.0) == 1  # on the segment
    assert L.vs_host_point_segment_within(0, 0, nan, 1, 2, 2, 1e9) == 0


def test_oracle_dwithin_pair_polygon_cases():
    import oracle
    from vibespatial_amd.geometry import PointArray, PolygonArray, box

    P = PolygonArray.from_geometries([box(0, 0, 10, 10), box(20, 0, 30, 10)])
    pts = PointArray(np.array([15.0, 5.0, 10.0, 15.0]), np.array([5.0, 5.0, 13.0, 20.0]))
    # (15, 5): 5 from both boxes; (5, 5): inside box 0; (10, 13): 3 from box 0;
    # (15, 20): sqrt(125) from box 0's corner (10, 10)
    assert oracle.dwithin_pair(0, 1, pts, P, 0, 0, 5.0) and oracle.dwithin_pair(0, 1, pts, P, 0, 1, 5.0)
    assert not oracle.dwithin_pair(0, 1, pts, P, 0, 0, math.nextafter(5.0, 0.0))
    assert oracle.dwithin_pair(0, 1, pts, P, 1, 0, 0.0)
    assert oracle.dwithin_pair(0, 1, pts, P, 2, 0, 3.0) and not oracle.dwithin_pair(0, 1, pts, P, 2, 0, 2.999)
    assert oracle.dwithin_pair(1, 0, P, pts, 0, 3, math.sqrt(125) + 1e-9)
    assert not oracle.dwithin_pair(1, 0, P, pts, 0, 3, 11.0)
    # polygon x polygon: the two boxes are 10 apart
    assert oracle.dwithin_pair(1, 1, P, P, 0, 1, 10.0) and not oracle.dwithin_pair(1, 1, P, P, 0, 1, 9.999)
    assert oracle.dwithin_pair(1, 1, P, P, 0, 0, 0.0)
