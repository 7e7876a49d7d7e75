"""GPU polygon x polygon predicates (within / contains / covers / covered_by /
contains_properly / touches / overlaps / crosses) against the rational
oracle: integer boxes with shared edges and corners, holes and
multipolygons, snapped star polygons, and the reference's own sindex.query
vectors through the public API."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import relate_ref as R
from vibespatial_amd import GeoSeries
from vibespatial_amd import synthetic as S
from vibespatial_amd.geometry import MultiPolygon, Point, Polygon, PolygonArray, box

pytestmark = pytest.mark.gpu

PREDS = ("intersects", "within", "contains", "covers", "covered_by", "contains_properly", "touches", "overlaps",
         "crosses")


def _pairs(tree, query, pred):
    res = GeoSeries(tree).sindex.query(GeoSeries(query).values, predicate=pred)
    return set(zip(res[0].tolist(), res[1].tolist()))


def _expect(tree, query, pred):
    qb, tb = query.bounds(), tree.bounds()
    out = set()
    for i in range(len(query)):
        for j in range(len(tree)):
            if qb[i, 0] <= tb[j, 2] and tb[j, 0] <= qb[i, 2] and qb[i, 1] <= tb[j, 3] and tb[j, 1] <= qb[i, 3]:
                if R.predicate(pred, query, i, tree, j):
                    out.add((i, j))
    return out


def test_integer_boxes():
    print("relate integer boxes", flush=True)
    rng = np.random.default_rng(3)
    x0, y0 = rng.integers(0, 8, 30), rng.integers(0, 8, 30)
    w, h = rng.integers(1, 5, 30), rng.integers(1, 5, 30)
    P = PolygonArray.from_geometries([box(a, b, a + c, b + d) for a, b, c, d in zip(x0, y0, w, h)])
    for pred in PREDS:
        assert _pairs(P, P, pred) == _expect(P, P, pred), pred


def test_holes_and_multipolygons():
    print("relate holes / multipolygons", flush=True)
    donut = Polygon([(0, 0), (10, 0), (10, 10), (0, 10), (0, 0)], [[(3, 3), (7, 3), (7, 7), (3, 7), (3, 3)]])
    geoms = [donut, box(3, 3, 7, 7), box(4, 4, 6, 6), box(0.5, 0.5, 2, 2), box(-1, -1, 11, 11), box(2, 2, 5, 5),
             MultiPolygon([box(0, 3, 3, 7), box(7, 3, 10, 7)]), box(0, 3, 10, 7),
             Polygon([(0, 0), (4, 0), (4, 4), (0, 0)]), box(0, 0, 4, 4), Polygon([(0, 0), (4, 4), (0, 4), (0, 0)])]
    P = PolygonArray.from_geometries(geoms)
    for pred in PREDS:
        assert _pairs(P, P, pred) == _expect(P, P, pred), pred


def _simple(ring):
    """No two non-adjacent edges of the closed ring meet (exact)."""
    e = list(zip(ring[:-1], ring[1:]))
    n = len(e)
    for i in range(n):
        for j in range(i + 1, n):
            if j == i + 1 or (i == 0 and j == n - 1):
                continue
            if R._touch(*e[i], *e[j]):
                return False
    return True


def _snapped_stars(rng, n, lo, hi):
    """Single-ring star polygons snapped to a grid of 0.5 (so vertices and
    edges of different rows coincide now and then), kept only when the
    snapped ring is still simple."""
    from fractions import Fraction as F

    out = []
    while len(out) < n:
        P = S.star_polygons(rng, rng.uniform(0, 60, 1), rng.uniform(0, 60, 1), rng.uniform(lo, hi, 1),
                            rng.integers(4, 9, 1))
        xs, ys = np.round(P.x[P.ring_offsets[0]:P.ring_offsets[1]] * 2) / 2, \
            np.round(P.y[P.ring_offsets[0]:P.ring_offsets[1]] * 2) / 2
        ring = [(F(float(x)), F(float(y))) for x, y in zip(xs, ys)]
        if len(set(ring)) >= 3 and _simple(ring) and R._area2(ring) != 0:
            out.append(Polygon(list(zip(xs.tolist(), ys.tolist()))))
    return PolygonArray.from_geometries(out)


def test_snapped_stars():
    print("relate snapped stars", flush=True)
    rng = np.random.default_rng(9)
    A = _snapped_stars(rng, 40, 3, 8)
    B = _snapped_stars(rng, 30, 4, 12)
    for pred in ("within", "contains", "touches", "overlaps", "contains_properly"):
        assert _pairs(B, A, pred) == _expect(B, A, pred), pred


@pytest.mark.parametrize("pred,query,expected", [
    # tests/upstream/geopandas/tests/test_sindex.py:297-369 (tree: points
    # (i, i) for i < 5 and box(10, 10, 20, 20))
    ("within", box(0.25, 0.28, 0.75, 0.75), []),
    ("within", box(0, 0, 10, 10), []),
    ("within", box(11, 11, 12, 12), [5]),
    ("contains", box(0, 0, 1, 1), []),
    ("contains", box(0, 0, 1.001, 1.001), [1]),
    ("contains", box(0.5, 0.5, 1.5, 1.5), [1]),
    ("contains", box(-1, -1, 2, 2), [0, 1]),
    ("touches", box(-1, -1, 0, 0), [0]),
    ("touches", box(-0.5, -0.5, 1.5, 1.5), []),
    ("contains", box(10, 10, 20, 20), [5]),
    ("covers", box(-0.5, -0.5, 1, 1), [0, 1]),
    ("covers", box(0.001, 0.001, 0.99, 0.99), []),
    ("covers", box(0, 0, 1, 1), [0, 1]),
    ("contains_properly", box(0, 0, 1, 1), []),
    ("contains_properly", box(0, 0, 1.001, 1.001), [1]),
    ("contains_properly", box(0.5, 0.5, 1.001, 1.001), [1]),
    ("contains_properly", box(0.5, 0.5, 1.5, 1.5), [1]),
    ("contains_properly", box(-1, -1, 2, 2), [0, 1]),
    ("contains_properly", box(10, 10, 20, 20), []),
])
def test_reference_query_vectors(pred, query, expected):
    tree = GeoSeries([Point(i, i) for i in range(5)] + [box(10, 10, 20, 20)])
    assert tree.sindex.query(query, predicate=pred).tolist() == expected


def test_shared_chains_and_multipolygon_parts():
    print("relate shared chains", flush=True)
    # an L-shaped polygon and the square completing it share a two-edge chain;
    # a multipolygon whose two parts touch at a corner; identical multipolygons
    L = Polygon([(0, 0), (4, 0), (4, 2), (2, 2), (2, 4), (0, 4), (0, 0)])
    sq = box(2, 2, 4, 4)
    full = box(0, 0, 4, 4)
    M = MultiPolygon([box(0, 0, 2, 2), box(2, 2, 4, 4)])
    geoms = [L, sq, full, M, MultiPolygon([box(0, 0, 2, 2), box(2, 2, 4, 4)]), box(1, 1, 3, 3)]
    P = PolygonArray.from_geometries(geoms)
    for pred in PREDS:
        assert _pairs(P, P, pred) == _expect(P, P, pred), pred
    # spot checks of the expected relations themselves
    assert (0, 1) in _expect(P, P, "touches") and (0, 2) in _expect(P, P, "within")
    assert (3, 4) in _expect(P, P, "within") and (3, 4) in _expect(P, P, "contains")
    assert (5, 3) in _expect(P, P, "overlaps")
