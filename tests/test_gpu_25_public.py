"""The public API reaches the learned + FP32 form, and its frames follow the
reference's join modes: results through sjoin / sindex.query compared with
the oracle and with the plain packed-tree form, the strategy each call took,
outer joins and on_attribute filtering on the reference's own cases."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN

import oracle
from vibespatial_amd import GeoDataFrame, GeoSeries, sjoin
from vibespatial_amd import join as J
from vibespatial_amd import synthetic as S
from vibespatial_amd.device import DevicePoints
from vibespatial_amd.geometry import Point, Polygon, PolygonArray, box_array
from vibespatial_amd.io import read_file

pytestmark = pytest.mark.gpu

NTHREADS = min(16, len(os.sched_getaffinity(0)))


def _frames(points, polygons):
    left = GeoDataFrame({"pid": np.arange(len(points))}, geometry=GeoSeries(points))
    right = GeoDataFrame({"qid": np.arange(len(polygons))}, geometry=GeoSeries(polygons))
    return left, right


# ---------------------------------------------------------------------------
# sjoin(points, polygons) -> learned join
# ---------------------------------------------------------------------------
def test_public_sjoin_within_full_c2_equals_oracle():
    """The benchmark workload through the public sjoin: 10M points x 100k
    polygons, within, frame rows equal to the oracle's pairs."""
    print("public sjoin c2 full", flush=True)
    w = S.c2()
    left, right = _frames(w.points, w.polygons)
    out = sjoin(left, right, how="inner", predicate="within")
    assert right.sindex.last_strategy == [("point", "polygon", "learned_join")]
    el, er = oracle.sjoin_points_polygons(w.points.x, w.points.y, w.polygons, "within", nthreads=NTHREADS)
    assert np.array_equal(out.index.to_numpy(), el)
    assert np.array_equal(out["index_right"].to_numpy(), er)
    assert np.array_equal(out["pid"].to_numpy(), el) and np.array_equal(out["qid"].to_numpy(), er)
    g = out.geometry.values
    assert np.array_equal(g.points.x, w.points.x[el]) and np.array_equal(g.points.y, w.points.y[el])


@pytest.mark.parametrize("predicate", ["within", "intersects", "covered_by", "touches"])
def test_public_sindex_query_point_predicates_learned(predicate):
    w = S.c1()
    rng = np.random.default_rng(3)
    px, py = w.points.x.copy(), w.points.y.copy()
    k = rng.integers(0, len(w.polygons.x), 800)
    px[:800], py[:800] = w.polygons.x[k], w.polygons.y[k]
    px[9] = np.nan
    from vibespatial_amd.geometry import PointArray

    tree = GeoSeries(w.polygons)
    res = tree.sindex.query(GeoSeries(PointArray(px, py)).values, predicate=predicate, return_device=True)
    assert tree.sindex.last_strategy == [("point", "polygon", "learned_join")]
    el, er = oracle.sjoin_points_polygons(px, py, w.polygons, predicate, nthreads=NTHREADS)
    assert np.array_equal(res[0].cpu().numpy(), el) and np.array_equal(res[1].cpu().numpy(), er)


# ---------------------------------------------------------------------------
# point trees: learned box query / learned polygon query
# ---------------------------------------------------------------------------
def test_public_sindex_query_c4_shape_equals_tree_form():
    """Rectangles against a point tree through sindex.query: the learned box
    query alone (no refine), equal to the packed-tree form and to brute
    force on a sample."""
    print("public sindex.query c4-shaped 4M x 20k", flush=True)
    w = S.c4(n_points=4_000_000, n_boxes=20_000)
    boxes = w.meta["boxes"]
    tree = GeoSeries(w.points)
    q = GeoSeries(box_array(boxes))
    res = tree.sindex.query(q.values, predicate="intersects", return_device=True)
    assert tree.sindex.last_strategy == [("polygon", "point", "learned_boxes")]
    Q = DevicePoints.from_host(w.points)
    t = J.build_tree(torch.stack([Q.x, Q.y, Q.x, Q.y], dim=1).contiguous())
    l0, r0, _ = J.query_boxes(torch.from_numpy(boxes).cuda(), t)
    assert torch.equal(res[0], l0) and torch.equal(res[1], r0)
    # predicate None: the same boxes are the polygons' bounds
    res_none = tree.sindex.query(q.values, predicate=None, return_device=True)
    assert torch.equal(res_none[0], l0) and torch.equal(res_none[1], r0)
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(len(boxes), 300, replace=False))
    el, er = oracle.box_query_points(w.points.x, w.points.y, boxes[sample], nthreads=NTHREADS)
    ln, rn = res[0].cpu().numpy(), res[1].cpu().numpy()
    m = np.isin(ln, sample)
    assert np.array_equal(ln[m], sample[el]) and np.array_equal(rn[m], er)


@pytest.mark.parametrize("predicate,point_pred", [("contains", "within"), ("contains_properly", "within"),
                                                  ("covers", "intersects"), ("intersects", "intersects"),
                                                  ("touches", "touches")])
def test_public_polygons_over_point_tree(predicate, point_pred):
    """Star polygons (holes, boundary points) and rectangles (points on their
    edges) over a point tree: the learned polygon query, exact."""
    w = S.c2(n_points=200_000, n_polys=3_000, extent=2e4)
    rng = np.random.default_rng(7)
    px, py = w.points.x.copy(), w.points.y.copy()
    k = rng.integers(0, len(w.polygons.x), 3000)
    px[:3000], py[:3000] = w.polygons.x[k], w.polygons.y[k]
    from vibespatial_amd.geometry import PointArray

    pts = PointArray(px, py)
    tree = GeoSeries(pts)
    res = tree.sindex.query(GeoSeries(w.polygons).values, predicate=predicate, return_device=True)
    assert tree.sindex.last_strategy == [("polygon", "point", "learned_polygons")]
    el, er = oracle.sjoin_points_polygons(px, py, w.polygons, point_pred, nthreads=NTHREADS)
    o = np.lexsort((el, er))
    assert np.array_equal(res[0].cpu().numpy(), er[o]) and np.array_equal(res[1].cpu().numpy(), el[o])
    # rectangles whose edges pass through tree points
    c = rng.integers(0, len(px), 500)
    half = rng.uniform(5, 200, (500, 2))
    b = np.column_stack([px[c] - half[:, 0], py[c] - half[:, 1], px[c] + half[:, 0], py[c] + half[:, 1]])
    b[:100, 0] = px[c[:100]]  # a tree point on the left edge
    rects = box_array(b)
    res = tree.sindex.query(GeoSeries(rects).values, predicate=predicate, return_device=True)
    expect_route = "learned_boxes" if predicate in ("intersects", "covers") else "learned_polygons"
    assert tree.sindex.last_strategy == [("polygon", "point", expect_route)]
    el, er = oracle.sjoin_points_polygons(px, py, rects, point_pred, nthreads=NTHREADS)
    o = np.lexsort((el, er))
    assert np.array_equal(res[0].cpu().numpy(), er[o]) and np.array_equal(res[1].cpu().numpy(), el[o])


def test_public_points_over_point_tree_equality():
    rng = np.random.default_rng(4)
    tx = np.round(rng.uniform(0, 100, 50_000), 1)
    ty = np.round(rng.uniform(0, 100, 50_000), 1)
    qx = np.concatenate([tx[:2000], rng.uniform(0, 100, 2000), [np.nan]])
    qy = np.concatenate([ty[:2000], rng.uniform(0, 100, 2000), [1.0]])
    from vibespatial_amd.geometry import PointArray

    tree = GeoSeries(PointArray(tx, ty))
    res = tree.sindex.query(GeoSeries(PointArray(qx, qy)).values, predicate="intersects")
    assert tree.sindex.last_strategy == [("point", "point", "learned_boxes")]
    pos = {}
    for j, key in enumerate(zip(tx.tolist(), ty.tolist())):
        pos.setdefault(key, []).append(j)
    exp = [(i, j) for i, key in enumerate(zip(qx.tolist(), qy.tolist())) for j in pos.get(key, [])]
    assert [tuple(p) for p in res.T.tolist()] == exp


# ---------------------------------------------------------------------------
# join modes: outer and on_attribute
# ---------------------------------------------------------------------------
def _nybb_frames():
    polydf = read_file(os.path.join(GOLDEN, "upstream_data", "nybb_16a.zip"))
    b = polydf.geometry.values.bounds()
    tb = [int(v) for v in (np.nanmin(b[:, 0]), np.nanmin(b[:, 1]), np.nanmax(b[:, 2]), np.nanmax(b[:, 3]))]
    n = 20
    xs = range(tb[0], tb[2], int((tb[2] - tb[0]) / n))
    ys = range(tb[1], tb[3], int((tb[3] - tb[1]) / n))
    pts = [Point(x, y) for x, y in zip(xs, ys)]
    pointdf = GeoDataFrame({"pointattr1": [p.x + p.y for p in pts], "pointattr2": [p.x - p.y for p in pts],
                            "geometry": pts})
    return pointdf, polydf


def test_outer_join_nybb_upstream_invariants():
    """tests/upstream/geopandas/tools/tests/test_sjoin.py:989-998."""
    pointdf, polydf = _nybb_frames()
    df = sjoin(pointdf, polydf, how="outer")
    left = sjoin(pointdf, polydf, how="left")
    right = sjoin(pointdf, polydf, how="right")
    unmatched_left = int(left["index_right"].isna().sum())
    unmatched_right = int(right["index_left"].isna().sum())
    assert len(df) == len(left) + unmatched_right == 22
    assert int(df["index_left"].isna().sum()) == unmatched_right
    assert int(df["index_right"].isna().sum()) == unmatched_left
    assert df.geometry.notna().all()
    assert list(df.columns) == ["index_left", "pointattr1", "pointattr2", "index_right", "BoroCode", "BoroName",
                                "Shape_Leng", "Shape_Area", "geometry"]


UPSTREAM_OUTER = {  # (left, right) rows of test_sjoin.py:27-96's expected frames
    "intersects": [(0, 0), (0, 1), (1, 1), (2, -1), (-1, 2)],
    "contains": [(0, 0), (1, -1), (2, -1), (-1, 1), (-1, 2)],
    "within": [(0, -1), (1, 1), (2, -1), (-1, 0), (-1, 2)],
}


@pytest.mark.parametrize("predicate", sorted(UPSTREAM_OUTER))
def test_outer_join_upstream_fixture(predicate):
    """tests/upstream/geopandas/tools/tests/test_sjoin.py:384-417 (default index)."""
    polys1 = [Polygon([(0, 0), (5, 0), (5, 5), (0, 5)]), Polygon([(5, 5), (6, 5), (6, 6), (5, 6)]),
              Polygon([(6, 0), (9, 0), (9, 3), (6, 3)])]
    polys2 = [Polygon([(1, 1), (4, 1), (4, 4), (1, 4)]), Polygon([(4, 4), (7, 4), (7, 7), (4, 7)]),
              Polygon([(7, 7), (10, 7), (10, 10), (7, 10)])]
    df1 = GeoDataFrame({"geometry": polys1, "df1": [0, 1, 2]})
    df2 = GeoDataFrame({"geometry": polys2, "df2": [3, 4, 5]})
    res = sjoin(df1, df2, how="outer", predicate=predicate)
    rows = UPSTREAM_OUTER[predicate]
    il = res["index_left"].to_numpy()
    ir = res["index_right"].to_numpy()
    np.testing.assert_array_equal(il, [np.nan if a < 0 else a for a, _ in rows])
    np.testing.assert_array_equal(ir, [np.nan if b < 0 else b for _, b in rows])
    assert list(res.columns) == ["index_left", "df1", "index_right", "df2", "geometry"]
    for k, (a, b) in enumerate(rows):
        src = polys1[a] if a >= 0 else polys2[b]
        assert np.array_equal(res.geometry[k].exterior, src.exterior)


def test_on_attribute_world_cities():
    """tests/upstream/geopandas/tools/tests/test_sjoin.py:513-526."""
    d = os.path.join(GOLDEN, "upstream_data")
    world = read_file(os.path.join(d, "naturalearth_lowres"))
    cities = read_file(os.path.join(d, "naturalearth_cities"))
    countries = GeoDataFrame({"country": world["name"].to_numpy(), "geometry": world.geometry})
    countries._df["firstLetter"] = countries["country"].astype(str).str[0].to_numpy()
    cities._df["firstLetter"] = cities["name"].astype(str).str[0].to_numpy()
    result = sjoin(cities, countries, on_attribute="firstLetter")
    assert (result["country"].astype(str).str[0] == result["name"].astype(str).str[0]).all()
    assert result.shape == (23, 5)


def _shared_attribute_frames():
    """tests/upstream/geopandas/tools/tests/test_sjoin.py:99-142."""
    pts = [Point(i, i) for i in range(8)]
    left = GeoDataFrame({"geometry": pts, "attr_tracker": list("ABCDEFGH"), "duplicate_column": list(range(8)),
                         "attr1": [True] * 8, "attr2": [True] * 8})
    right = GeoDataFrame({"geometry": list(pts), "duplicate_column": list(range(8)),
                          "attr1": [True, True, False, False, True, True, False, False],
                          "attr2": [True, True, False, False, False, False, False, False]})
    return left, right


@pytest.mark.parametrize("mapping", [{True: "merge", False: "no_merge"}, {True: 2, False: 1},
                                     {True: True, False: False}])
def test_on_attribute_multiple_columns(mapping):
    """test_sjoin.py:528-563: joined on two attributes keeps rows A and B."""
    left, right = _shared_attribute_frames()
    for f in (left, right):
        for c in ("attr1", "attr2"):
            f._df[c] = f._df[c].map(mapping)
    joined = sjoin(left, right, on_attribute=("attr1", "attr2"))
    assert (["A", "B"] == joined["attr_tracker"].values).all()


def test_on_attribute_check_header():
    """test_sjoin.py:565-570."""
    left, right = _shared_attribute_frames()
    joined = sjoin(left, right, on_attribute=["attr1"])
    assert (["A", "B", "E", "F"] == joined["attr_tracker"].values).all()
    assert {"attr2_left", "attr2_right", "attr1"}.issubset(set(joined.columns))
    assert "attr1_left" not in joined
