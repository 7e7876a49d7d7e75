"""Host logic of the public API (no GPU): outer-join indexers and frames, the
rectangle test behind the box-query route, the strategy table, vectorised
geometry gathers."""

from __future__ import annotations

import numpy as np
import pytest

from vibespatial_amd import api
from vibespatial_amd import synthetic as S
from vibespatial_amd.geometry import GeometryArray, MultiPolygon, Point, Polygon, PolygonArray, box


def _outer_loop(pairs, n_left, n_right):
    """Outer-join rows by definition: each left row with its matches in right
    order (or one row with no match), then unmatched right rows ascending."""
    rows = []
    for i in range(n_left):
        m = sorted(r for l, r in pairs if l == i)
        rows += [(i, r) for r in m] if m else [(i, -1)]
    hit = {r for _, r in pairs}
    rows += [(-1, j) for j in range(n_right) if j not in hit]
    return rows


@pytest.mark.parametrize("seed", range(6))
def test_outer_indexers_match_definition(seed):
    rng = np.random.default_rng(seed)
    nl, nr = int(rng.integers(0, 12)), int(rng.integers(1, 12))
    k = int(rng.integers(0, 30)) if nl else 0
    pairs = sorted({(int(rng.integers(0, nl)), int(rng.integers(0, nr))) for _ in range(k)})
    perm = rng.permutation(len(pairs))
    l = np.array([pairs[i][0] for i in perm], dtype=np.int64)
    r = np.array([pairs[i][1] for i in perm], dtype=np.int64)
    lo, ro, pos = api.outer_indexers(l, r, nl, nr)
    assert list(zip(lo.tolist(), ro.tolist())) == _outer_loop(pairs, nl, nr)
    # every input pair lands on its own output row
    assert np.array_equal(lo[pos], l) and np.array_equal(ro[pos], r)


def _dfs():
    """The upstream fixture (tests/upstream/geopandas/tools/tests/test_sjoin.py:27-96)."""
    polys1 = [Polygon([(0, 0), (5, 0), (5, 5), (0, 5)]), Polygon([(5, 5), (6, 5), (6, 6), (5, 6)]),
              Polygon([(6, 0), (9, 0), (9, 3), (6, 3)])]
    polys2 = [Polygon([(1, 1), (4, 1), (4, 4), (1, 4)]), Polygon([(4, 4), (7, 4), (7, 7), (4, 7)]),
              Polygon([(7, 7), (10, 7), (10, 10), (7, 10)])]
    df1 = api.GeoDataFrame({"geometry": polys1, "df1": [0, 1, 2]})
    df2 = api.GeoDataFrame({"geometry": polys2, "df2": [3, 4, 5]})
    return df1, df2


# (left, right) rows of the upstream expected outer frames for each predicate
UPSTREAM_OUTER = {
    "intersects": ([(0, 0), (0, 1), (1, 1)], [(0, 0), (0, 1), (1, 1), (2, -1), (-1, 2)]),
    "contains": ([(0, 0)], [(0, 0), (1, -1), (2, -1), (-1, 1), (-1, 2)]),
    "within": ([(1, 1)], [(0, -1), (1, 1), (2, -1), (-1, 0), (-1, 2)]),
}


@pytest.mark.parametrize("predicate", sorted(UPSTREAM_OUTER))
def test_outer_frame_matches_upstream_expectation(predicate):
    df1, df2 = _dfs()
    pairs, rows = UPSTREAM_OUTER[predicate]
    l = np.array([p[0] for p in pairs])
    r = np.array([p[1] for p in pairs])
    out = api._assemble(df1, df2, l, r, "outer", "left", "right", None)
    assert list(out.columns) == ["index_left", "df1", "index_right", "df2", "geometry"]
    assert list(out.index) == list(range(len(rows)))
    il = out["index_left"].to_numpy()
    ir = out["index_right"].to_numpy()
    exp_l = np.array([np.nan if a < 0 else a for a, _ in rows], dtype=float)
    exp_r = np.array([np.nan if b < 0 else b for _, b in rows], dtype=float)
    np.testing.assert_array_equal(il, exp_l)
    np.testing.assert_array_equal(ir, exp_r)
    np.testing.assert_array_equal(out["df1"].to_numpy(dtype=float), np.where(np.isnan(exp_l), np.nan, exp_l))
    np.testing.assert_array_equal(out["df2"].to_numpy(dtype=float), np.where(np.isnan(exp_r), np.nan, exp_r + 3))
    # geometry: the left row's, else the right row's
    g = out.geometry
    assert g.notna().all()
    for k, (a, b) in enumerate(rows):
        src = df1.geometry[a] if a >= 0 else df2.geometry[b]
        assert np.array_equal(g[k].exterior, src.exterior)


def test_left_and_inner_frames_keep_left_order_and_labels():
    df1, df2 = _dfs()
    df1._df.index = ["a", "b", "c"]
    df2._df.index = ["d", "e", "f"]
    l, r = np.array([0, 0, 1]), np.array([0, 1, 1])
    left = api._assemble(df1, df2, l, r, "left", "left", "right", None)
    assert list(left.index) == ["a", "a", "b", "c"]
    assert left["index_right"].tolist()[:3] == ["d", "e", "e"] and left["index_right"].isna().tolist()[3]
    inner = api._assemble(df1, df2, l, r, "inner", "left", "right", None)
    assert list(inner.index) == ["a", "a", "b"] and inner["index_right"].tolist() == ["d", "e", "e"]


def test_all_rectangles():
    boxes = PolygonArray.from_geometries([box(0, 0, 1, 2), box(3, 3, 4, 5, ccw=False), None, Polygon()])
    assert api.all_rectangles(boxes)
    star = S.c1(n_points=1, n_polys=3).polygons
    assert not api.all_rectangles(star)
    holed = PolygonArray.from_geometries([Polygon([(0, 0), (4, 0), (4, 4), (0, 4)], [[(1, 1), (2, 1), (2, 2), (1, 2)]])])
    assert not api.all_rectangles(holed)
    flat = PolygonArray.from_geometries([Polygon([(0, 0), (4, 0), (4, 0), (0, 0)])])
    assert not api.all_rectangles(flat)
    multi = PolygonArray.from_geometries([MultiPolygon([box(0, 0, 1, 1), box(2, 2, 3, 3)])])
    assert not api.all_rectangles(multi)
    skew = PolygonArray.from_geometries([Polygon([(0, 0), (4, 0), (5, 4), (0, 4)])])
    assert not api.all_rectangles(skew)


def test_strategy_table():
    rect = PolygonArray.from_geometries([box(0, 0, 1, 1)])
    star = S.c1(n_points=1, n_polys=2).polygons
    st = api._strategy
    for p in ("within", "intersects", "covered_by", "touches"):
        assert st("point", "polygon", p, None) == "learned_join"
    assert st("point", "polygon", None, None) == "tree"
    assert st("point", "polygon", "dwithin", None) == "tree"
    for p in (None, "intersects", "within", "contains"):
        assert st("point", "point", p, None) == "learned_boxes"
    assert st("polygon", "point", "intersects", rect) == "learned_boxes"
    assert st("polygon", "point", "covers", rect) == "learned_boxes"
    assert st("polygon", "point", None, star) == "learned_boxes"
    for p in ("contains", "contains_properly", "touches"):
        assert st("polygon", "point", p, rect) == "learned_polygons"
    assert st("polygon", "point", "intersects", star) == "learned_polygons"
    for p in (None, "intersects", "within", "overlaps", "touches"):
        assert st("polygon", "polygon", p, star) == "tree"


def test_geometry_take_and_concat_match_scalars():
    w = S.c2(n_points=50, n_polys=40)
    P = w.polygons
    idx = np.array([5, 0, 39, 5, 12])
    A = P.take(idx)
    B = PolygonArray.from_geometries([P.scalar(int(i)) for i in idx])
    for f in ("x", "y", "ring_offsets", "part_offsets", "geom_offsets"):
        assert np.array_equal(getattr(A, f), getattr(B, f)), f
    g = GeometryArray.from_scalars([Point(1, 2), None, box(0, 0, 1, 1), MultiPolygon([box(0, 0, 1, 1), box(2, 2, 3, 3)]),
                                    Point(), Polygon()])
    sel = np.array([3, -1, 0, 2, 4, 5, 1, 3])
    t = g.take(sel)
    for k, i in enumerate(sel):
        a, b = t[k], (g[int(i)] if i >= 0 else None)
        assert repr(a) == repr(b)
    c = GeometryArray.concat([g, t])
    assert len(c) == len(g) + len(t)
    for k in range(len(c)):
        assert repr(c[k]) == repr(g[k] if k < len(g) else t[k - len(g)])
