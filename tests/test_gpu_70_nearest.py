"""GPU nearest (sindex.nearest / sjoin_nearest) against the brute-force
oracle: the reference's own sindex.nearest vectors, lattice points with exact
ties, points x polygons both ways, polygons x polygons, max_distance,
return_all, exclusive, null rows and far-away queries that need many search
rounds.  Indices and fp64 distances bit-exact."""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from vibespatial_amd import GeoDataFrame, GeoSeries, sjoin_nearest
from vibespatial_amd import synthetic as S
from vibespatial_amd.geometry import Point, PointArray, Polygon, PolygonArray, box

pytestmark = pytest.mark.gpu


def _nearest(tree, q, **kw):
    (l, r), d = GeoSeries(tree).sindex.nearest(GeoSeries(q).values, return_distance=True, **kw)
    return l, r, d


def _check(lk, rk, q, tree, **kw):
    l, r, d = _nearest(tree, q, **kw)
    el, er, ed = oracle.nearest_pairs(lk, rk, q, tree, **kw)
    assert np.array_equal(l, el) and np.array_equal(r, er), kw
    assert np.array_equal(d.view(np.int64), ed.view(np.int64)), kw  # bit-exact distances


def test_reference_vectors():
    print("nearest reference vectors", flush=True)
    s = GeoSeries(PointArray(np.arange(10.0), np.arange(10.0)))
    for ra in (True, False):
        # tests/upstream/geopandas/tests/test_sindex.py:806-880
        assert s.sindex.nearest(Point(0.25, 0.25), return_all=ra).tolist() == [[0], [0]]
        assert s.sindex.nearest(Point(0.75, 0.75), return_all=ra).tolist() == [[0], [1]]
        assert s.sindex.nearest([Point(1, 1), Point(0, 0)], return_all=ra).tolist() == [[0, 1], [1, 0]]
        assert s.sindex.nearest([Point(1, 1), Point(0.25, 1)], return_all=ra).tolist() == [[0, 1], [1, 1]]
        assert s.sindex.nearest(None, return_all=ra).tolist() == [[], []]
        assert s.sindex.nearest([None], return_all=ra).tolist() == [[], []]
    ps = [Point(0.5, 0.5), Point(0, 10)]
    idx, d = s.sindex.nearest(ps, return_distance=True)
    assert idx.tolist() == [[0, 0, 1], [0, 1, 5]] and d.tolist() == [math.sqrt(0.5)] * 2 + [math.sqrt(50)]
    idx, d = s.sindex.nearest(ps, return_all=False, max_distance=1, return_distance=True)
    assert idx.tolist() == [[0], [0]] and d.tolist() == [math.sqrt(0.5)]
    # :908-944, exclusive
    g = GeoSeries(PointArray(np.append(np.arange(5.0), 1.0), np.append(np.arange(5.0), 2.0)))
    idx, d = g.sindex.nearest(g.values, return_all=True, max_distance=1.1, exclusive=True, return_distance=True)
    assert idx.tolist() == [[1, 2, 5, 5], [5, 5, 1, 2]] and d.tolist() == [1.0] * 4
    g5 = GeoSeries(PointArray(np.arange(5.0), np.arange(5.0)))
    assert g5.sindex.nearest(g5.values, exclusive=True).tolist() == [[0, 1, 1, 2, 2, 3, 3, 4],
                                                                      [1, 0, 2, 1, 3, 2, 4, 3]]
    assert g5.sindex.nearest(g5.values, exclusive=True, return_all=False).tolist() == [[0, 1, 2, 3, 4],
                                                                                        [1, 0, 1, 2, 3]]


@pytest.mark.parametrize("offset", [0.0, 5e5])
def test_points_points_lattice_ties(offset):
    print(f"nearest points x points offset {offset}", flush=True)
    rng = np.random.default_rng(4)
    t = PointArray(rng.integers(0, 40, 400).astype(float) + offset, rng.integers(0, 40, 400).astype(float) + offset)
    q = PointArray(rng.integers(-5, 45, 300).astype(float) + offset + rng.choice([0, 0.5], 300),
                   rng.integers(-5, 45, 300).astype(float) + offset)
    for kw in ({}, {"return_all": False}, {"max_distance": 1.0}, {"max_distance": 0.5, "return_all": False},
               {"exclusive": True}):
        _check(0, 0, q, t, **kw)


def test_far_queries_need_many_rounds():
    print("nearest far queries", flush=True)
    rng = np.random.default_rng(5)
    t = PointArray(rng.normal(0, 1, 3000), rng.normal(0, 1, 3000))
    q = PointArray(np.array([1e4, -3e5, 0.1, 7.0, 2e6]), np.array([0.0, 2.0, 0.1, -7.0, -2e6]))
    _check(0, 0, q, t)
    _check(0, 0, q, t, max_distance=10.0)


def test_points_polygons_both_ways():
    print("nearest points x polygons", flush=True)
    rng = np.random.default_rng(6)
    P = S.star_polygons(rng, rng.uniform(0, 200, 80), rng.uniform(0, 200, 80), rng.uniform(2, 8, 80),
                        rng.integers(4, 24, 80))
    q = PointArray(rng.uniform(-20, 220, 250), rng.uniform(-20, 220, 250))
    q.x[:30], q.y[:30] = P.x[:30], P.y[:30]  # on vertices: distance 0
    _check(0, 1, q, P)
    _check(0, 1, q, P, return_all=False, max_distance=3.0)
    # polygons query point trees
    t = PointArray(rng.uniform(0, 200, 400), rng.uniform(0, 200, 400))
    _check(1, 0, P, t)
    B = PolygonArray.from_geometries([box(0, 0, 10, 10), box(20, 0, 30, 10)])
    qb = PointArray(np.array([15.0, 5.0, 10.0, 25.0]), np.array([5.0, 5.0, 13.0, 12.0]))
    l, r, d = _nearest(B, qb)
    assert list(zip(l.tolist(), r.tolist(), d.tolist())) == [(0, 0, 5.0), (0, 1, 5.0), (1, 0, 0.0), (2, 0, 3.0),
                                                             (3, 1, 2.0)]


def test_polygons_polygons():
    print("nearest polygons x polygons", flush=True)
    rng = np.random.default_rng(7)
    A = S.star_polygons(rng, rng.uniform(0, 150, 60), rng.uniform(0, 150, 60), rng.uniform(1, 5, 60),
                        rng.integers(4, 16, 60))
    B = S.star_polygons(rng, rng.uniform(0, 150, 50), rng.uniform(0, 150, 50), rng.uniform(2, 7, 50),
                        rng.integers(4, 20, 50))
    _check(1, 1, A, B)
    _check(1, 1, A, B, max_distance=4.0)


def test_polygons_polygons_exclusive_equal_rows():
    """exclusive between polygon columns: tree rows topologically equal to
    the query row (same ring with another start vertex or orientation) are
    left out, and the next nearest rows returned, as the oracle decides."""
    print("nearest polygons x polygons exclusive", flush=True)
    rng = np.random.default_rng(8)
    B = S.star_polygons(rng, rng.uniform(0, 150, 50), rng.uniform(0, 150, 50), rng.uniform(2, 7, 50),
                        rng.integers(4, 20, 50))
    geoms = []
    for i in range(40):
        g = B.scalar(i)
        ring = np.asarray(g.parts[0][0])[:-1]
        k = i % 3
        if k == 1:
            ring = np.roll(ring, 2, axis=0)
        elif k == 2:
            ring = ring[::-1]
        geoms.append(Polygon([tuple(p) for p in ring]))
    A = PolygonArray.from_geometries(geoms + [B.scalar(45), box(0, 0, 3, 3)])
    _check(1, 1, A, B, exclusive=True)
    _check(1, 1, A, B, exclusive=True, return_all=False)
    _check(1, 1, A, B, exclusive=True, max_distance=2.0)


def test_null_rows_and_sjoin_nearest():
    print("sjoin_nearest", flush=True)
    left = GeoDataFrame({"a": np.arange(6), "geometry": [Point(0, 0), None, Point(5, 5), Point(), Point(9, 1),
                                                         Point(30, 30)]})
    right = GeoDataFrame({"b": [10, 11, 12, 13], "geometry": [None, box(4, 4, 6, 6), Point(8, 0), Point(10, 2)]})
    j = sjoin_nearest(left, right, distance_col="dist")
    got = list(zip(j.index.tolist(), j["index_right"].tolist(), j["dist"].tolist()))
    # (0,0): box corner (4,4) at sqrt(32), (8,0) at 8 -> box; (5,5) inside the box;
    # (9,1): (8,0) and (10,2) both at sqrt(2); (30,30): box corner (6,6) at sqrt(1152)
    assert got == [(0, 1, math.sqrt(32)), (2, 1, 0.0), (4, 2, math.sqrt(2)), (4, 3, math.sqrt(2)),
                   (5, 1, math.sqrt(1152))]
    assert j["b"].tolist() == [11, 11, 12, 13, 11]
    jl = sjoin_nearest(left, right, how="left", max_distance=2.0, distance_col="dist")
    assert jl.index.tolist() == [0, 1, 2, 3, 4, 4, 5]
    assert np.isnan(jl["index_right"].tolist()[0]) and jl["index_right"].tolist()[2] == 1
    jr = sjoin_nearest(left, right, how="right", distance_col="dist")
    # right rows query the left points: box -> (5,5) at 0; (8,0) -> (9,1); (10,2) -> (9,1); null row kept
    assert jr.index.tolist() == [0, 1, 2, 3]
    assert np.isnan(jr["index_left"].tolist()[0]) and jr["index_left"].tolist()[1:] == [2, 4, 4]
    assert jr["dist"].tolist()[1:] == [0.0, math.sqrt(2), math.sqrt(2)]


def test_mixed_tree_ties_and_edge_options():
    print("nearest mixed tree", flush=True)
    # a tree of points and polygons: ties across the two families, the
    # least tree row kept when return_all is off, max_distance 0 keeps exact
    # contacts only, empty / NaN rows never match
    tree = GeoSeries([Point(3, 0), box(-1, 2, 1, 4), None, Point(), box(10, 10, 12, 12), Point(0, -2),
                      Polygon([])])
    q = GeoSeries([Point(0, 0), Point(11, 11), Point(1, 4), Point(50, 50)])
    (l, r), d = tree.sindex.nearest(q.values, return_distance=True)
    got = list(zip(l.tolist(), r.tolist(), d.tolist()))
    # (0, 0): box(-1, 2, 1, 4) at 2 and (0, -2) at 2 tie; (3, 0) at 3
    assert got[:2] == [(0, 1, 2.0), (0, 5, 2.0)]
    assert (1, 4, 0.0) in got and (2, 1, 0.0) in got
    assert [g for g in got if g[0] == 3] == [(3, 4, math.sqrt(38.0 * 38.0 + 38.0 * 38.0))]
    (l, r), d = tree.sindex.nearest(q.values, return_all=False, return_distance=True)
    assert list(zip(l.tolist(), r.tolist())) == [(0, 1), (1, 4), (2, 1), (3, 4)]
    res = tree.sindex.nearest(q.values, max_distance=0.0)
    assert res.tolist() == [[1, 2], [4, 1]]


def _holed(rng, n, lo, hi, extent):
    holes = rng.integers(0, 3, n)
    nh = int(holes.sum())
    return S.star_polygons(rng, rng.uniform(0, extent, n), rng.uniform(0, extent, n), rng.uniform(lo, hi, n),
                           rng.integers(6, 20, n), holes=holes, hole_vertices=rng.integers(4, 9, nh),
                           hole_frac=rng.uniform(0.2, 0.5, nh))


def test_polygons_with_holes():
    print("nearest holed polygons", flush=True)
    rng = np.random.default_rng(17)
    P = _holed(rng, 60, 4, 10, 150)
    # points at hole centres (inside a hole: distance to the hole ring) and random ones
    q = PointArray(np.concatenate([P.x[P.ring_offsets[1:40]], rng.uniform(0, 150, 200)]),
                   np.concatenate([P.y[P.ring_offsets[1:40]], rng.uniform(0, 150, 200)]))
    _check(0, 1, q, P)
    _check(0, 1, q, P, max_distance=2.0, return_all=False)
    A = _holed(rng, 40, 2, 6, 150)
    _check(1, 1, A, P)


def test_pair_distances_cap_and_outside_flag():
    print("pair distances cap / outside", flush=True)
    import torch

    from vibespatial_amd import join as J
    from vibespatial_amd.device import DevicePoints, DevicePolygons

    rng = np.random.default_rng(23)
    P = _holed(rng, 30, 3, 8, 100)
    q = PointArray(rng.uniform(-10, 110, 400), rng.uniform(-10, 110, 400))
    left = np.repeat(np.arange(400), 30)
    right = np.tile(np.arange(30), 400)
    expect = np.array([oracle.distance_f64(0, 1, q, P, int(a), int(b)) for a, b in zip(left, right)])
    Q, DP = DevicePoints.from_host(q), DevicePolygons.from_host(P)
    dl, dr = (torch.as_tensor(a, device="cuda") for a in (left, right))
    full = J.pair_distances(dl, dr, Q, DP).cpu().numpy()
    assert np.array_equal(full.view(np.int64), expect.view(np.int64))
    cap = rng.uniform(0, 30, 400)
    cap[::7] = np.nan  # no cap
    capped = J.pair_distances(dl, dr, Q, DP, cap=torch.as_tensor(cap, device="cuda")).cpu().numpy()
    c = cap[left]
    want = np.where(expect > c, np.inf, expect)
    assert np.array_equal(capped.view(np.int64), want.view(np.int64))
    # points outside every polygon: skipping the location gives the same distances
    out = expect > 0
    o = torch.as_tensor(np.flatnonzero(out), device="cuda")
    got = J.pair_distances(dl[o], dr[o], Q, DP, outside=True).cpu().numpy()
    assert np.array_equal(got.view(np.int64), expect[out].view(np.int64))
