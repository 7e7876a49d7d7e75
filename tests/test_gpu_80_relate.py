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


# ---------------------------------------------------------------------------
# heavy pairs: one wave per pair (split 4 / 2 / 1 lanes per edge, rings over
# 64 vertices, holes, multipolygon parts, shared collinear chains)
# ---------------------------------------------------------------------------
def _simple_int(xy):
    """Closed ring of integer coordinates: no two non-adjacent edges meet."""
    a = xy[:-1]; b = xy[1:]
    n = len(a)
    i, j = np.triu_indices(n, 2)
    keep = ~((i == 0) & (j == n - 1))
    i, j = i[keep], j[keep]
    def o(p, q, r):
        return np.sign((q[..., 0] - p[..., 0]) * (r[..., 1] - p[..., 1]) - (q[..., 1] - p[..., 1]) * (r[..., 0] - p[..., 0]))
    A, B, C, D = a[i], b[i], a[j], b[j]
    o1, o2, o3, o4 = o(A, B, C), o(A, B, D), o(C, D, A), o(C, D, B)
    proper = (o1 * o2 < 0) & (o3 * o4 < 0)
    def on(p, q, r, oo):
        return (oo == 0) & (np.minimum(p[..., 0], q[..., 0]) <= r[..., 0]) & (r[..., 0] <= np.maximum(p[..., 0], q[..., 0])) \
            & (np.minimum(p[..., 1], q[..., 1]) <= r[..., 1]) & (r[..., 1] <= np.maximum(p[..., 1], q[..., 1]))
    touch = on(A, B, C, o1) | on(A, B, D, o2) | on(C, D, A, o3) | on(C, D, B, o4)
    return not (proper | touch).any()


def grid_star(rng, cx, cy, r, n):
    """Star ring on a grid of 0.5 (doubled to integers for the checks)."""
    while True:
        t = np.sort(rng.uniform(0, 2 * np.pi, n))
        rad = r * (1 + rng.uniform(-0.3, 0.3, n))
        xy = np.round(np.column_stack([cx + rad * np.cos(t), cy + rad * np.sin(t)]) * 2).astype(np.int64)
        keep = np.r_[True, np.any(xy[1:] != xy[:-1], axis=1)]
        xy = xy[keep]
        if len(xy) > 3 and np.all(xy[-1] == xy[0]):
            xy = xy[:-1]
        if len(xy) < 3:
            continue
        ring = np.vstack([xy, xy[:1]])
        if _simple_int(ring):
            return [tuple(v) for v in (ring[:-1] / 2.0).tolist()]


def comb(x0, y0, x1, y1, nb, nt):
    """Rectangle whose bottom / top edges carry nb / nt collinear vertices."""
    bot = [(x0 + (x1 - x0) * k / nb, y0) for k in range(nb)]
    top = [(x1 - (x1 - x0) * k / nt, y1) for k in range(nt)]
    return Polygon(bot + [(x1, y0)] + top + [(x0, y1)])


def _heavy_rows(rng):
    A, B = [], []
    # split 4 (rows <= 16 positions) x split 2 (<= 32), holes on the larger
    for k in range(6):
        cx, cy = 40.0 * k, 0.0
        A.append(Polygon(grid_star(rng, cx, cy, 6, int(rng.integers(8, 14)))))
        h = [(cx - 2, cy - 2), (cx - 2, cy + 2), (cx + 2, cy + 2), (cx + 2, cy - 2)]
        B.append(Polygon(grid_star(rng, cx + 4, cy, 12, int(rng.integers(22, 30))), [h]))
    # split 1 rows (> 32 positions) and rings over 64 vertices, small partners
    for k in range(5):
        cx, cy = 40.0 * k, 60.0
        A.append(Polygon(grid_star(rng, cx, cy, 14, int(rng.integers(66, 90)))))
        B.append(Polygon(grid_star(rng, cx + 8, cy + 3, 7, int(rng.integers(10, 16)))))
        B.append(box(cx - 3, cy - 3, cx + 3, cy + 3))  # inside the big star
    # multipolygon rows against stars
    for k in range(4):
        cx, cy = 40.0 * k, 120.0
        A.append(MultiPolygon([Polygon(grid_star(rng, cx - 6, cy, 5, 12)), Polygon(grid_star(rng, cx + 9, cy, 5, 14))]))
        B.append(Polygon(grid_star(rng, cx, cy, 11, int(rng.integers(30, 40)))))
    # shared collinear chains, cut at different points (touches / overlaps)
    for k in range(4):
        x = 40.0 * k
        A.append(comb(x, 180, x + 20, 190, 20 + k, 3))
        B.append(comb(x + 5 * (k % 2), 170, x + 20, 180, 3, 17 + 2 * k))   # below, sharing y=180
        B.append(comb(x + 2, 181, x + 18, 189, 16, 16))                  # strictly inside
    # rows inside one another touching the boundary from inside (covers but
    # not contains_properly), and query rows inside tree rows (within)
    for k in range(3):
        x = 40.0 * k
        A.append(comb(x, 220, x + 20, 230, 18, 4))
        B.append(comb(x + 2, 225, x + 10, 230, 12, 4))
        A.append(Polygon(grid_star(rng, x + 31, 225, 2, 12)))
        B.append(Polygon(grid_star(rng, x + 31, 225, 9, int(rng.integers(24, 40)))))
    # equal rows
    A.append(Polygon(grid_star(np.random.default_rng(77), 200, 200, 15, 70)))
    B.append(Polygon(grid_star(np.random.default_rng(77), 200, 200, 15, 70)))
    return PolygonArray.from_geometries(A), PolygonArray.from_geometries(B)



def test_heavy_pairs_every_predicate_against_oracle_and_both_routes():
    """Every candidate pair here multiplies past the 2^8 coordinate-product
    threshold, so the default route refines them one wave per pair
    (refine_heavy_kernel, then the wave relation for undecided ones).  The
    same pairs through the one-thread route (threshold 2^63) and with every
    pair heavy (2^1) give the same pairs, and all equal the rational oracle."""
    import torch

    from vibespatial_amd import api

    print("relate heavy pairs", flush=True)
    A, B = _heavy_rows(np.random.default_rng(5))
    qa, tb = A.bounds(), B.bounds()
    pairs = [(i, j) for i in range(len(A)) for j in range(len(B))
             if qa[i, 0] <= tb[j, 2] and tb[j, 0] <= qa[i, 2] and qa[i, 1] <= tb[j, 3] and tb[j, 1] <= qa[i, 3]]
    ca = np.diff(A.ring_offsets[A.part_offsets[A.geom_offsets]])
    cb = np.diff(B.ring_offsets[B.part_offsets[B.geom_offsets]])
    assert all(ca[i] * cb[j] > 256 for i, j in pairs)
    assert max(ca) > 64 and min(ca) <= 16 and any(16 < c <= 32 for c in ca)
    l = torch.tensor([p[0] for p in pairs], dtype=torch.int64, device="cuda")
    r = torch.tensor([p[1] for p in pairs], dtype=torch.int64, device="cuda")
    rel = {p: (R.relate(A, p[0], B, p[1]), R.intersects(A, p[0], B, p[1])) for p in pairs}
    seen = set()
    for pred in PREDS:
        expect = set()
        for p, ((meet, a_in_b, b_in_a, touch), inter) in rel.items():
            v = {"intersects": inter, "crosses": False, "within": a_in_b, "covered_by": a_in_b,
                 "contains": b_in_a, "covers": b_in_a, "contains_properly": b_in_a and not touch,
                 "touches": not meet and inter, "overlaps": meet and not a_in_b and not b_in_a}[pred]
            if v:
                expect.add(p)
        got = {}
        for hl in (0, 1, 63):
            kl, kr = api._refine(1, pred, l, r, None, A, B, heavy_log2=hl)
            got[hl] = set(zip(kl.cpu().tolist(), kr.cpu().tolist()))
        assert got[0] == got[1] == got[63], pred
        assert got[0] == expect, pred
        if expect:
            seen.add(pred)
    # the inputs exercise every relation
    assert seen >= {"intersects", "within", "contains", "contains_properly", "touches", "overlaps"}
