"""GPU parity of predicate='dwithin' (exact distance decisions) against the
exact-rational oracle, through sindex.query and sjoin, for every family pair,
with per-row distances, exact ties and the argument errors."""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from vibespatial_amd import GeoDataFrame, GeoSeries, sjoin
from vibespatial_amd import synthetic as S
from vibespatial_amd.geometry import Point, PointArray, PolygonArray, box

pytestmark = pytest.mark.gpu


def _points(rng, n, lo, hi, grid=None):
    if grid:  # integer lattice: exact distances such as 3-4-5 occur
        return PointArray(rng.integers(lo, hi, n).astype(np.float64) * grid,
                          rng.integers(lo, hi, n).astype(np.float64) * grid)
    return PointArray(rng.uniform(lo, hi, n), rng.uniform(lo, hi, n))


def _query(tree, q, d):
    res = GeoSeries(tree).sindex.query(GeoSeries(q).values, predicate="dwithin", distance=d)
    return res[0], res[1]


def _bounds(kind, g):
    if kind == 0:
        return np.stack([g.x, g.y, g.x, g.y], axis=1)
    return oracle.polygon_bounds(g)


def _expect(lk, rk, Lg, Rg, d):
    d = np.broadcast_to(np.asarray(d, dtype=np.float64), (len(Lg.x) if lk == 0 else len(Lg),)).copy()
    return oracle.dwithin_pairs(lk, rk, Lg, Rg, d, _bounds(lk, Lg), _bounds(rk, Rg))


@pytest.mark.parametrize("offset", [0.0, 5e5])
def test_points_points_with_exact_ties(offset):
    print(f"dwithin points x points offset {offset}", flush=True)
    rng = np.random.default_rng(1)
    q = _points(rng, 1500, 0, 60, grid=1.0)
    t = _points(rng, 2000, 0, 60, grid=1.0)
    q = PointArray(q.x + offset, q.y + offset)
    t = PointArray(t.x + offset, t.y + offset)
    for d in (5.0, math.nextafter(5.0, 0.0), 0.0):
        l, r = _query(t, q, d)
        el, er = _expect(0, 0, q, t, d)
        assert np.array_equal(l, el) and np.array_equal(r, er), d
    # per-row distances, including rows that never match (NaN, negative)
    d = rng.choice([0.0, 1.0, 2.5, 5.0, -1.0, np.nan], len(q.x))
    l, r = _query(t, q, d)
    el, er = _expect(0, 0, q, t, d)
    assert np.array_equal(l, el) and np.array_equal(r, er)
    assert not np.isin(l, np.nonzero(~(d >= 0))[0]).any()


def test_points_polygons_both_directions():
    print("dwithin points x polygons", flush=True)
    rng = np.random.default_rng(2)
    P = S.star_polygons(rng, rng.uniform(0, 200, 300), rng.uniform(0, 200, 300), rng.uniform(2, 8, 300),
                        rng.integers(4, 24, 300))
    q = _points(rng, 3000, -10, 210)
    # some points exactly on vertices and at exact offsets from axis-aligned boxes
    q.x[:200], q.y[:200] = P.x[:200], P.y[:200]
    B = PolygonArray.from_geometries([box(0, 0, 10, 10), box(30, 0, 40, 10)])
    qb = PointArray(np.array([15.0, 5.0, 10.0, 20.0, 45.0]), np.array([5.0, 5.0, 13.0, 5.0, 14.0]))
    for d in (0.0, 1.5, 4.0):
        l, r = _query(P, q, d)
        el, er = _expect(0, 1, q, P, d)
        assert np.array_equal(l, el) and np.array_equal(r, er), d
        l, r = _query(q, P, d)  # polygons query tree points
        el, er = _expect(1, 0, P, q, d)
        assert np.array_equal(l, el) and np.array_equal(r, er), d
    for d in (5.0, math.nextafter(5.0, 0.0), 3.0, 10.0):
        l, r = _query(B, qb, d)
        el, er = _expect(0, 1, qb, B, d)
        assert np.array_equal(l, el) and np.array_equal(r, er), d
    # (15, 5) is 5 from box 0, (5, 5) inside it, (10, 13) 3 above it; (20, 5)
    # is 10 from both boxes, (45, 14) sqrt(41) from box 1's corner
    l, r = _query(B, qb, 5.0)
    assert set(zip(l.tolist(), r.tolist())) == {(0, 0), (1, 0), (2, 0)}
    l, r = _query(B, qb, 10.0)
    assert set(zip(l.tolist(), r.tolist())) == {(0, 0), (1, 0), (2, 0), (3, 0), (3, 1), (4, 1)}


def test_polygons_polygons():
    print("dwithin polygons x polygons", flush=True)
    rng = np.random.default_rng(3)
    A = S.star_polygons(rng, rng.uniform(0, 150, 400), rng.uniform(0, 150, 400), rng.uniform(1, 5, 400),
                        rng.integers(4, 16, 400))
    B = S.star_polygons(rng, rng.uniform(0, 150, 250), rng.uniform(0, 150, 250), rng.uniform(2, 7, 250),
                        rng.integers(4, 20, 250))
    for d in (0.0, 0.75, 3.0):
        l, r = _query(B, A, d)
        el, er = _expect(1, 1, A, B, d)
        assert np.array_equal(l, el) and np.array_equal(r, er), d
    C = PolygonArray.from_geometries([box(0, 0, 10, 10), box(20, 0, 30, 10), box(10, 13, 12, 15)])
    for d, exp in ((10.0, {(0, 1), (1, 0)}), (math.nextafter(10.0, 0.0), set())):
        l, r = _query(C, C, d)
        got = {(a, b) for a, b in zip(l.tolist(), r.tolist()) if a != b and {a, b} != {0, 2} and {a, b} != {1, 2}}
        assert got == exp, d


def test_sjoin_dwithin_and_argument_errors():
    left = GeoDataFrame({"a": np.arange(50), "geometry": [Point(float(i), 0.0) for i in range(50)]})
    right = GeoDataFrame({"b": np.arange(3), "geometry": [box(10, 2, 12, 4), box(30, -1, 31, 1), None]})
    j = sjoin(left, right, predicate="dwithin", distance=2.0)
    pairs = sorted(zip(j.index.tolist(), j["index_right"].tolist()))
    # points (i, 0): within 2 of box(10, 2, 12, 4) for i in 10..12; inside / near box 1 for i in 28..33
    assert pairs == [(i, 0) for i in (10, 11, 12)] + [(i, 1) for i in range(28, 34)]
    with pytest.raises(ValueError):
        right.sindex.query(left.geometry.values, predicate="dwithin")
    with pytest.raises(ValueError):
        right.sindex.query(left.geometry.values, predicate="intersects", distance=1.0)
    with pytest.raises(ValueError):
        right.sindex.query(left.geometry.values, predicate="dwithin", distance=np.ones(7))
    # scalar query geometry
    res = right.sindex.query(Point(11.0, 6.0), predicate="dwithin", distance=2.0)
    assert res.tolist() == [0]


def test_null_rows_keep_their_row_numbers():
    # family indices are mapped back to column rows when nulls sit between
    tree = GeoSeries([None, Point(0.0, 0.0), None, Point(5.0, 5.0)])
    assert tree.sindex.query(Point(0.0, 0.0), predicate="intersects").tolist() == [1]
    assert tree.sindex.query(Point(5.0, 5.0), predicate="dwithin", distance=0.0).tolist() == [3]
    res = tree.sindex.query([None, Point(5.0, 4.0)], predicate="dwithin", distance=1.0)
    assert res.tolist() == [[1], [3]]


def test_polygons_with_holes():
    print("dwithin holed polygons", flush=True)
    rng = np.random.default_rng(23)
    holes = rng.integers(0, 3, 80)
    nh = int(holes.sum())
    P = S.star_polygons(rng, rng.uniform(0, 150, 80), rng.uniform(0, 150, 80), rng.uniform(4, 10, 80),
                        rng.integers(6, 20, 80), holes=holes, hole_vertices=rng.integers(4, 9, nh),
                        hole_frac=rng.uniform(0.2, 0.5, nh))
    q = _points(rng, 2000, -5, 155)
    for d in (0.0, 0.5, 2.0):
        l, r = _query(P, q, d)
        el, er = _expect(0, 1, q, P, d)
        assert np.array_equal(l, el) and np.array_equal(r, er), d
