"""GPU parity of the join path against the oracle: bit-exact pair sets."""

from __future__ import annotations

import os
import zipfile

import numpy as np
import pytest
import torch

from conftest import GOLDEN, decode_geom

import oracle
from vibespatial_amd import GeoDataFrame, GeoSeries, sjoin
from vibespatial_amd import join as J
from vibespatial_amd import synthetic as S
from vibespatial_amd.device import DevicePoints, DevicePolygons
from vibespatial_amd.geometry import PointArray, PolygonArray
from vibespatial_amd.io import read_file

pytestmark = pytest.mark.gpu

NTHREADS = min(16, len(os.sched_getaffinity(0)))


def _gpu_join(px, py, polys, predicate, capacity=None):
    Q = DevicePoints(torch.from_numpy(np.ascontiguousarray(px)).cuda(), torch.from_numpy(np.ascontiguousarray(py)).cuda())
    P = DevicePolygons.from_host(polys)
    tree = J.build_tree(J.polygon_bounds(P))
    l, r, st = J.sjoin_points_polygons(Q, P, tree, predicate, capacity=capacity)
    return l.cpu().numpy(), r.cpu().numpy(), st


def _with_boundary_points(px, py, polys, rng, k):
    """Overwrite k points with vertices and k with dyadic edge midpoints."""
    px, py = px.copy(), py.copy()
    n = len(px)
    vi = rng.integers(0, len(polys.x), k)
    idx = rng.choice(n, 2 * k, replace=False)
    px[idx[:k]], py[idx[:k]] = polys.x[vi], polys.y[vi]
    vj = rng.integers(0, len(polys.x) - 1, k)
    px[idx[k:]] = (polys.x[vj] + polys.x[vj + 1]) * 0.5
    py[idx[k:]] = (polys.y[vj] + polys.y[vj + 1]) * 0.5
    return px, py


@pytest.mark.parametrize("predicate", ["within", "intersects", "covered_by", "touches"])
def test_c1_bit_exact(predicate):
    print(f"c1 {predicate}", flush=True)
    w = S.c1()
    px, py = _with_boundary_points(w.points.x, w.points.y, w.polygons, np.random.default_rng(1), 500)
    px[17] = np.nan
    l, r, st = _gpu_join(px, py, w.polygons, predicate)
    el, er = oracle.sjoin_points_polygons(px, py, w.polygons, predicate, nthreads=NTHREADS)
    assert len(el) > 0
    assert np.array_equal(l, el) and np.array_equal(r, er)


def test_c1_default_points_within():
    w = S.c1()
    l, r, st = _gpu_join(w.points.x, w.points.y, w.polygons, "within")
    el, er = oracle.sjoin_points_polygons(w.points.x, w.points.y, w.polygons, "within", nthreads=NTHREADS)
    assert np.array_equal(l, el) and np.array_equal(r, er)
    assert st.candidates >= len(l)


def test_c2_scaled_bit_exact_with_holes():
    print("c2 1M x 100k", flush=True)
    w = S.c2(n_points=1_000_000, n_polys=100_000)
    px, py = _with_boundary_points(w.points.x, w.points.y, w.polygons, np.random.default_rng(2), 2000)
    for pred in ("within", "intersects"):
        l, r, st = _gpu_join(px, py, w.polygons, pred)
        el, er = oracle.sjoin_points_polygons(px, py, w.polygons, pred, nthreads=NTHREADS)
        assert np.array_equal(l, el) and np.array_equal(r, er), pred


def test_c2_full_size_bit_exact():
    """The benchmark workload itself: 10M points x 100k polygons, within."""
    print("c2 full 10M x 100k", flush=True)
    w = S.c2()
    l, r, st = _gpu_join(w.points.x, w.points.y, w.polygons, "within")
    assert np.all(np.diff(l) >= 0) and np.all((np.diff(l) > 0) | (np.diff(r) > 0))  # strictly lexsorted
    el, er = oracle.sjoin_points_polygons(w.points.x, w.points.y, w.polygons, "within", nthreads=NTHREADS)
    assert np.array_equal(l, el) and np.array_equal(r, er)


def test_edge_cases():
    w = S.c1(n_points=1000, n_polys=50)
    P = w.polygons
    # no points
    l, r, _ = _gpu_join(np.empty(0), np.empty(0), P, "within")
    assert len(l) == 0
    # all NaN points
    l, r, _ = _gpu_join(np.full(100, np.nan), np.full(100, np.nan), P, "intersects")
    assert len(l) == 0
    # no polygons
    empty = PolygonArray.from_geometries([])
    l, r, _ = _gpu_join(w.points.x, w.points.y, empty, "within")
    assert len(l) == 0
    # single polygon, single point at its vertex: boundary
    one = PolygonArray.from_geometries([P.scalar(0)])
    vx, vy = one.x[:1], one.y[:1]
    assert len(_gpu_join(vx, vy, one, "within")[0]) == 0
    assert len(_gpu_join(vx, vy, one, "intersects")[0]) == 1
    assert len(_gpu_join(vx, vy, one, "touches")[0]) == 1
    # empty rows among polygons and a multipolygon with a hole part
    from vibespatial_amd.geometry import MultiPolygon, Polygon, box

    geoms = [None, box(0, 0, 4, 4), Polygon(), MultiPolygon([box(10, 10, 12, 12),
                                                             Polygon([(20, 20), (30, 20), (30, 30), (20, 30)],
                                                                     [[(22, 22), (28, 22), (28, 28), (22, 28)]])])]
    PA = PolygonArray.from_geometries(geoms)
    px = np.array([2, 4, 11, 25, 21, 22, 100, 0.0])
    py = np.array([2, 2, 11, 25, 21, 25, 100, 0.0])
    for pred in ("within", "intersects", "touches"):
        l, r, _ = _gpu_join(px, py, PA, pred)
        el, er = oracle.sjoin_points_polygons(px, py, PA, pred)
        assert np.array_equal(l, el) and np.array_equal(r, er), pred


def test_capacity_overflow_retries_to_the_same_answer():
    w = S.c1(n_points=20_000, n_polys=300)
    l, r, st = _gpu_join(w.points.x, w.points.y, w.polygons, "intersects", capacity=7)
    assert st.retries == 1
    el, er = oracle.sjoin_points_polygons(w.points.x, w.points.y, w.polygons, "intersects")
    assert np.array_equal(l, el) and np.array_equal(r, er)


def test_many_overlapping_polygons_per_point():
    # 300 nested squares around one spot: long per-point runs exercise the
    # per-point sort of polygon ids
    from vibespatial_amd.geometry import box

    geoms = [box(-s, -s, s, s) for s in np.linspace(1, 300, 300)[::-1]]
    PA = PolygonArray.from_geometries(geoms)
    rng = np.random.default_rng(4)
    px, py = rng.uniform(-310, 310, 5000), rng.uniform(-310, 310, 5000)
    l, r, _ = _gpu_join(px, py, PA, "within")
    el, er = oracle.sjoin_points_polygons(px, py, PA, "within")
    assert np.array_equal(l, el) and np.array_equal(r, er)


# ---------------------------------------------------------------------------
# reference literal vectors through the public API
# ---------------------------------------------------------------------------
def test_golden_pip_through_sindex(golden):
    for case in golden["pip"]:
        polys = [decode_geom(g) for g in case["polygons"]]
        points = [decode_geom(g) for g in case["points"]]
        for poly, p, exp in zip(polys, points, case["expected"]):
            if exp is None:
                continue
            res = GeoSeries([p]).sindex.query(poly, predicate="covers")
            assert (len(res) == 1) == exp, case["source"]
            res2 = GeoSeries([poly]).sindex.query(p, predicate="covered_by")
            assert (len(res2) == 1) == exp, case["source"]


def test_golden_queries(golden):
    for case in golden["query"]:
        tree = GeoSeries([decode_geom(g) for g in case["tree"]])
        if "query_scalar" in case:
            res = tree.sindex.query(decode_geom(case["query_scalar"]), predicate=case["predicate"], sort=True)
            assert sorted(res.tolist()) == case["expected"], case["source"]
        else:
            q = [decode_geom(g) for g in case["query"]]
            res = tree.sindex.query(q, predicate=case["predicate"], sort=True)
            assert res.tolist() == case["expected"], case["source"]


def test_golden_sindex_table(golden):
    table = golden["sindex_table"]
    tree = GeoSeries([decode_geom(g) for g in table["tree"]])
    from vibespatial_amd.geometry import box

    for row in table["rows"]:
        q = [box(*b) for b in row["query"]]
        arg = q[0] if row["scalar"] else q
        # rows marked polygon_de9im need the polygon x polygon relations
        res = tree.sindex.query(arg, predicate=row["predicate"])
        assert res.tolist() == row["expected"], (row["source"], row["predicate"], row["query"])


def test_golden_polygon_intersects(golden):
    for case in golden["polygon_intersects"]:
        A = GeoSeries([decode_geom(g) for g in case["left"]])
        B = GeoSeries([decode_geom(g) for g in case["right"]])
        res = B.sindex.query(A.values, predicate="intersects")
        got = set(zip(res[0].tolist(), res[1].tolist()))
        for (i, j), exp in zip(case["pairs"], case["expected"]):
            assert ((i, j) in got) == exp, case["source"]


def test_polygon_join_matches_oracle():
    rng = np.random.default_rng(7)
    A = S.star_polygons(rng, rng.uniform(0, 300, 3000), rng.uniform(0, 300, 3000), rng.uniform(1, 6, 3000),
                        rng.integers(4, 30, 3000))
    B = S.star_polygons(rng, rng.uniform(0, 300, 800), rng.uniform(0, 300, 800), rng.uniform(2, 9, 800),
                        rng.integers(4, 40, 800))
    res = GeoSeries(B).sindex.query(GeoSeries(A).values, predicate="intersects")
    el, er = oracle.sjoin_polygons_polygons(A, B, nthreads=NTHREADS)
    assert np.array_equal(res[0], el) and np.array_equal(res[1], er)


def test_polygons_contain_tree_points_matches_oracle():
    w = S.c1(n_points=20_000, n_polys=400)
    tree = GeoSeries(w.points)
    for pred, opred in (("contains", "within"), ("covers", "intersects"), ("touches", "touches")):
        res = tree.sindex.query(GeoSeries(w.polygons).values, predicate=pred)
        el, er = oracle.sjoin_points_polygons(w.points.x, w.points.y, w.polygons, opred)
        o = np.lexsort((el, er))
        assert np.array_equal(res[0], er[o]) and np.array_equal(res[1], el[o]), pred


# ---------------------------------------------------------------------------
# upstream GeoPandas sjoin counts (GEOS results recorded by the reference)
# ---------------------------------------------------------------------------
def test_upstream_nybb_sjoin_shapes():
    polydf = read_file(os.path.join(GOLDEN, "upstream_data", "nybb_16a.zip"))
    b = polydf.geometry.values.bounds()
    tb = [int(v) for v in (np.nanmin(b[:, 0]), np.nanmin(b[:, 1]), np.nanmax(b[:, 2]), np.nanmax(b[:, 3]))]
    N = 20
    xs = range(tb[0], tb[2], int((tb[2] - tb[0]) / N))
    ys = range(tb[1], tb[3], int((tb[3] - tb[1]) / N))
    from vibespatial_amd.geometry import Point

    pts = [Point(x, y) for x, y in zip(xs, ys)]
    pointdf = GeoDataFrame({"pointattr1": [p.x + p.y for p in pts], "pointattr2": [p.x - p.y for p in pts],
                            "geometry": pts})
    assert sjoin(pointdf, polydf, how="inner").shape == (11, 8)
    assert sjoin(pointdf, polydf, how="left").shape == (21, 8)
    assert sjoin(pointdf, polydf, how="right").shape == (12, 8)
    df = sjoin(pointdf, polydf, how="left", predicate="within")
    assert df.loc[1]["BoroName"] == "Staten Island"
    df = sjoin(pointdf, polydf, how="left", predicate="contains")
    assert df.shape == (21, 8) and np.isnan(df.loc[1]["Shape_Area"])


def test_upstream_world_cities_shape():
    d = os.path.join(GOLDEN, "upstream_data")
    world = read_file(os.path.join(d, "naturalearth_lowres"))
    cities = read_file(os.path.join(d, "naturalearth_cities"))
    countries = GeoDataFrame({"country": world["name"].to_numpy(), "geometry": world.geometry})
    res = sjoin(cities, countries, how="inner", predicate="intersects")
    assert res.shape == (213, 4)


def test_box_query_output_sized_from_previous_queries():
    print("box query sizing", flush=True)
    # one tree, queries whose hits per box go up and down: the output size
    # follows the last query (too small -> a second pass, too large -> slack)
    rng = np.random.default_rng(31)
    c = rng.uniform(0, 1000, (5000, 2))
    s = rng.uniform(1, 20, (5000, 2))
    tb = np.column_stack([c - s, c + s])
    tree = J.build_tree(torch.as_tensor(tb, device="cuda"))
    expect_retries = []
    for half in (30.0, 1.0, 200.0, 0.5, 30.0):
        q = rng.uniform(0, 1000, (3000, 2))
        qb = np.column_stack([q - half, q + half])
        l, r, st = J.query_boxes(torch.as_tensor(qb, device="cuda"), tree)
        el, er = oracle.bbox_pairs(qb, tb)
        order = np.lexsort((er, el))
        assert np.array_equal(l.cpu().numpy(), el[order]) and np.array_equal(r.cpu().numpy(), er[order]), half
        expect_retries.append(st.retries)
    assert expect_retries[2] == 1 and expect_retries[3] == 0  # grown past the hint, then slack


def test_box_query_both_traversal_shapes_agree():
    """Box queries run 64 boxes per wave when the caller expects few hits per
    box (capacity below 32 per box) and one box per wave otherwise; the same
    queries through both shapes, on 16- and 64-wide point trees, give the
    oracle's pairs (NaN and empty-hit boxes included)."""
    print("box query traversal shapes", flush=True)
    rng = np.random.default_rng(33)
    p = rng.uniform(0, 1000, (20_000, 2))
    tb = np.column_stack([p, p])
    q = rng.uniform(0, 1000, (2_000, 2))
    half = rng.uniform(1, 15, (2_000, 1))
    qb = np.column_stack([q - half, q + half])
    qb[7] = np.nan
    qb[11] = [2000.0, 2000.0, 2001.0, 2001.0]
    el, er = oracle.bbox_pairs(qb, tb)
    o = np.lexsort((er, el))
    n_hits = len(el)
    assert n_hits < 16 * len(qb)  # so capacity 16 per box is enough for the 64-box shape
    for fanout in (16, 64):
        tree = J.build_tree(torch.as_tensor(tb, device="cuda"), fanout=fanout)
        for per_box in (16, 48):  # below / above the one-box threshold
            l, r, st = J.query_boxes(torch.as_tensor(qb, device="cuda"), tree, capacity=per_box * len(qb))
            assert st.retries == 0 and st.pairs == n_hits, (fanout, per_box)
            assert np.array_equal(l.cpu().numpy(), el[o]) and np.array_equal(r.cpu().numpy(), er[o]), (fanout, per_box)


def test_box_query_hint_does_not_oversize_a_large_query():
    """A query with about 5,000 hits per box followed by 400k point boxes:
    the second query's output is sized from at most HINT_SLOTS_PER_QUERY
    slots per box, not from the first query's hits per box; queries given a
    capacity (nearest rounds) and dwithin leave the plain hint alone."""
    print("box query hint clamp", flush=True)
    rng = np.random.default_rng(32)
    c = rng.uniform(0, 1000, (5000, 2))
    tb = np.column_stack([c - 2, c + 2])
    tree = J.build_tree(torch.as_tensor(tb, device="cuda"))
    big = np.array([[-10.0, -10.0, 1010.0, 1010.0]] * 64)
    l, r, st = J.query_boxes(torch.as_tensor(big, device="cuda"), tree)
    assert st.pairs == 64 * 5000
    assert tree.hits_per_query["bbox"] == 5000
    q = rng.uniform(0, 1000, (400_000, 2))
    qb = np.column_stack([q, q])
    l, r, st = J.query_boxes(torch.as_tensor(qb, device="cuda"), tree)
    assert st.capacity <= J.HINT_SLOTS_PER_QUERY * 400_000
    el, er = oracle.bbox_pairs(qb[:20_000], tb)
    o = np.lexsort((er, el))
    ln, rn = l.cpu().numpy(), r.cpu().numpy()
    m = ln < 20_000
    assert np.array_equal(ln[m], el[o]) and np.array_equal(rn[m], er[o])
    # an explicit capacity (as the nearest search passes) does not move the hint
    before = dict(tree.hits_per_query)
    J.query_boxes(torch.as_tensor(big, device="cuda"), tree, capacity=64 * 5000 + 10)
    assert tree.hits_per_query == before
