"""Writes tests/golden/reference_vectors.json.

The cases are transcribed literally from the reference's own tests (inputs
and the expected values stated there); nothing from the reference is imported
or run.  Geometries are stored as {"type": ..., "parts": [[ring, ...], ...]}
with rings as [x, y] lists, points as {"type": "Point", "xy": [x, y]}, None
for a null geometry.  Floats that a test computes with numpy are computed here
the same way and stored exactly (float.hex).

    python tests/golden/make_reference_vectors.py
"""

from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def box(minx, miny, maxx, maxy):
    # shapely.box vertex order (ccw=True)
    return {"type": "Polygon", "parts": [[[[maxx, miny], [maxx, maxy], [minx, maxy], [minx, miny], [maxx, miny]]]]}


def poly(shell, holes=()):
    def close(r):
        r = [list(map(float, p)) for p in r]
        return r if r[0] == r[-1] else r + [r[0]]

    return {"type": "Polygon", "parts": [[close(shell)] + [close(h) for h in holes]]}


def mpoly(*polys):
    return {"type": "MultiPolygon", "parts": [p["parts"][0] for p in polys]}


def pt(x, y):
    return {"type": "Point", "xy": [float(x).hex(), float(y).hex()]}


EMPTY_POLY = {"type": "Polygon", "parts": []}
EMPTY_MPOLY = {"type": "MultiPolygon", "parts": []}
EMPTY_POINT = {"type": "Point", "xy": None}


def pip_cases():
    """polygon.covers(point) row-aligned, null -> None, empty -> False
    (src/vibespatial/testing/oracle.py:276-289)."""
    cases = []
    # tests/test_point_in_polygon.py:52-70
    cases.append({
        "source": "tests/test_point_in_polygon.py:52-70",
        "polygons": [box(0, 0, 2, 2), EMPTY_POLY, None, box(5, 1, 9, 3), poly([(0, 0), (4, 0), (4, 4), (0, 4)])],
        "points": [None, EMPTY_POINT, pt(1, 1), pt(8, 2), pt(4, 2)],
        "expected": [None, False, None, True, True],
    })
    # tests/test_point_in_polygon.py:73-88
    cases.append({
        "source": "tests/test_point_in_polygon.py:73-88",
        "polygons": [mpoly(box(0, 0, 2, 2), box(10, 10, 12, 12)), mpoly(box(0, 0, 2, 2)), EMPTY_MPOLY],
        "points": [pt(1, 1), pt(3, 3), pt(0, 0)],
        "expected": [True, False, False],
    })
    # tests/test_point_in_polygon.py:116-131 (donut: Point(2,2) sits in the hole)
    cases.append({
        "source": "tests/test_point_in_polygon.py:116-131",
        "polygons": [
            box(0, 0, 2, 2),
            mpoly(box(5, 5, 8, 8), box(10, 10, 12, 12)),
            poly([(0, 0), (4, 0), (4, 4), (0, 4)], holes=[[(1, 1), (3, 1), (3, 3), (1, 3), (1, 1)]]),
            EMPTY_MPOLY,
        ],
        "points": [pt(1, 1), pt(11, 11), pt(2, 2), pt(0, 0)],
        "expected": [True, True, False, False],
    })
    # tests/test_point_in_polygon.py:135-161 (1 microdegree outside a diagonal)
    tri = poly([(-113.0, 35.0), (-110.0, 37.0), (-110.0, 35.0), (-113.0, 35.0)])
    cases.append({
        "source": "tests/test_point_in_polygon.py:135-161",
        "polygons": [tri, tri, tri],
        "points": [pt(-111.5000005547002, 36.0000008320503), pt(-179.0, -80.0), pt(179.0, 80.0)],
        "expected": [False, False, False],
    })
    # tests/test_point_in_polygon.py:165-189 (1e-5 outside a short edge)
    edge_midpoint = np.asarray([0.0001, 0.00005])
    outward_normal = np.asarray([-0.0001, 0.0002])
    outward_normal /= np.linalg.norm(outward_normal)
    p = edge_midpoint + outward_normal * 1e-5
    cases.append({
        "source": "tests/test_point_in_polygon.py:165-189",
        "polygons": [poly([(0.0, 0.0), (0.0002, 0.0001), (0.0002, -0.0001), (0.0, 0.0)])],
        "points": [pt(p[0], p[1])],
        "expected": [False],
    })
    # tests/test_point_in_polygon.py:477-501 (UTM-scale midpoint, covers True)
    p1 = (362084.1914861024, 3075556.618107719)
    p2 = (362173.59818072635, 3076109.62579047)
    p3 = (362090.5054597832, 3076110.5708671827)
    mx = p2[0] + (p3[0] - p2[0]) * 0.5
    my = p2[1] + (p3[1] - p2[1]) * 0.5
    cases.append({
        "source": "tests/test_point_in_polygon.py:477-501",
        "polygons": [poly([p1, p2, p3])],
        "points": [pt(mx, my)],
        "expected": [True],
    })
    return cases


def query_cases():
    """(query geometries, tree geometries, predicate, expected (2, n) or (n,))."""
    cases = []
    cases.append({
        "source": "tests/test_spatial_query.py:245-253",
        "tree": [box(0, 0, 1, 1), box(10, 10, 11, 11), box(20, 20, 21, 21)],
        "query": [box(0.5, 0.5, 1.5, 1.5), box(30, 30, 31, 31)],
        "predicate": "intersects",
        "expected": [[0], [0]],
    })
    cases.append({
        "source": "tests/test_spatial_query.py:840-862",
        "tree": [box(0, 0, 1, 1), box(1, 0, 2, 1), box(0, 1, 1, 2)],
        "query": [pt(1, 1), pt(1.5, 1.5)],
        "predicate": "intersects",
        "expected": [[0, 0, 0], [0, 1, 2]],
    })
    cases.append({
        "source": "tests/test_spatial_query.py:276-292 (membership; scalar query)",
        "tree": [pt(5, 5), pt(2, 2), pt(4, 4), pt(0, 0), pt(3, 3), pt(1, 1)],
        "query_scalar": box(0, 0, 2, 2),
        "predicate": "intersects",
        "expected": [1, 3, 5],
    })
    # tests/test_polygon_predicates.py:809-823
    cases.append({
        "source": "tests/test_polygon_predicates.py:809-823",
        "tree": [box(0, 0, 2, 2), box(1, 1, 3, 3), box(5, 5, 7, 7), box(10, 10, 12, 12)],
        "query": [box(0, 0, 1.5, 1.5)],
        "predicate": "intersects",
        "expected": [[0, 0], [0, 1]],
    })
    # tests/test_polygon_predicates.py:825-844: 10 boxes [i, i+1.5] x [0, 1.5],
    # self-join intersects: every pair with |i - j| <= 1
    boxes = [box(i, 0, i + 1.5, 1.5) for i in range(10)]
    pairs = [(i, j) for i in range(10) for j in range(10) if abs(i - j) <= 1]
    cases.append({
        "source": "tests/test_polygon_predicates.py:825-844",
        "tree": boxes,
        "query": boxes,
        "predicate": "intersects",
        "expected": [[p[0] for p in pairs], [p[1] for p in pairs]],
    })
    return cases


def sindex_table():
    """Upstream GeoPandas sindex tables: tree = 5 points on the diagonal plus
    box(10, 10, 20, 20) (tests/upstream/geopandas/tests/test_sindex.py:242-249).
    "requires" names what a row needs beyond point/polygon location and
    polygon intersects (rows with LineString queries are left out)."""
    rows = []
    scalar = [  # test_sindex.py:274-390, (predicate, query box, expected tree rows)
        (None, (-1, -1, -0.5, -0.5), []),
        (None, (-0.5, -0.5, 0.5, 0.5), [0]),
        (None, (0, 0, 1, 1), [0, 1]),
        ("intersects", (-1, -1, -0.5, -0.5), []),
        ("intersects", (-0.5, -0.5, 0.5, 0.5), [0]),
        ("intersects", (0, 0, 1, 1), [0, 1]),
        ("within", (0.25, 0.28, 0.75, 0.75), []),
        ("within", (0, 0, 10, 10), []),
        ("within", (11, 11, 12, 12), [5]),
        ("contains", (0, 0, 1, 1), []),
        ("contains", (0, 0, 1.001, 1.001), [1]),
        ("contains", (0.5, 0.5, 1.5, 1.5), [1]),
        ("contains", (-1, -1, 2, 2), [0, 1]),
        ("touches", (-1, -1, 0, 0), [0]),
        ("touches", (-0.5, -0.5, 1.5, 1.5), []),
        ("contains", (10, 10, 20, 20), [5]),
        ("covers", (-0.5, -0.5, 1, 1), [0, 1]),
        ("covers", (0.001, 0.001, 0.99, 0.99), []),
        ("covers", (0, 0, 1, 1), [0, 1]),
        ("contains_properly", (0, 0, 1, 1), []),
        ("contains_properly", (0, 0, 1.001, 1.001), [1]),
        ("contains_properly", (0.5, 0.5, 1.001, 1.001), [1]),
        ("contains_properly", (0.5, 0.5, 1.5, 1.5), [1]),
        ("contains_properly", (-1, -1, 2, 2), [0, 1]),
        ("contains_properly", (10, 10, 20, 20), []),
    ]
    for pred, b, exp in scalar:
        rows.append({"source": "tests/upstream/geopandas/tests/test_sindex.py:274-390", "predicate": pred,
                     "query": [b], "scalar": True, "expected": exp,
                     "requires": _requires(pred, [b])})
    bulk = [  # test_sindex.py:566-690
        (None, [(-1, -1, -0.5, -0.5)], [[], []]),
        (None, [(-0.5, -0.5, 0.5, 0.5)], [[0], [0]]),
        (None, [(0, 0, 1, 1)], [[0, 0], [0, 1]]),
        ("intersects", [(-1, -1, -0.5, -0.5)], [[], []]),
        ("intersects", [(-0.5, -0.5, 0.5, 0.5)], [[0], [0]]),
        ("intersects", [(0, 0, 1, 1)], [[0, 0], [0, 1]]),
        ("intersects", [(-1, -1, -0.5, -0.5), (-0.5, -0.5, 0.5, 0.5)], [[1], [0]]),
        ("intersects", [(-1, -1, 1, 1), (-0.5, -0.5, 0.5, 0.5)], [[0, 0, 1], [0, 1, 0]]),
        ("within", [(0.25, 0.28, 0.75, 0.75)], [[], []]),
        ("within", [(0, 0, 10, 10)], [[], []]),
        ("within", [(11, 11, 12, 12)], [[0], [5]]),
        ("contains", [(0, 0, 1, 1)], [[], []]),
        ("contains", [(0, 0, 1.001, 1.001)], [[0], [1]]),
        ("contains", [(0.5, 0.5, 1.001, 1.001)], [[0], [1]]),
        ("contains", [(0.5, 0.5, 1.5, 1.5)], [[0], [1]]),
        ("contains", [(-1, -1, 2, 2)], [[0, 0], [0, 1]]),
        ("contains", [(10, 10, 20, 20)], [[0], [5]]),
        ("touches", [(-1, -1, 0, 0)], [[0], [0]]),
        ("touches", [(-0.5, -0.5, 1.5, 1.5)], [[], []]),
        ("covers", [(-0.5, -0.5, 1, 1)], [[0, 0], [0, 1]]),
        ("covers", [(0.001, 0.001, 0.99, 0.99)], [[], []]),
        ("covers", [(0, 0, 1, 1)], [[0, 0], [0, 1]]),
        ("contains_properly", [(0, 0, 1, 1)], [[], []]),
        ("contains_properly", [(0, 0, 1.001, 1.001)], [[0], [1]]),
        ("contains_properly", [(0.5, 0.5, 1.001, 1.001)], [[0], [1]]),
        ("contains_properly", [(0.5, 0.5, 1.5, 1.5)], [[0], [1]]),
        ("contains_properly", [(-1, -1, 2, 2)], [[0, 0], [0, 1]]),
        ("contains_properly", [(10, 10, 20, 20)], [[], []]),
    ]
    for pred, bs, exp in bulk:
        rows.append({"source": "tests/upstream/geopandas/tests/test_sindex.py:566-690", "predicate": pred,
                     "query": bs, "scalar": False, "expected": exp, "requires": _requires(pred, bs)})
    return {"tree": [pt(i, i) for i in range(5)] + [box(10, 10, 20, 20)], "rows": rows}


def _requires(pred, boxes):
    # a query box against the tree box(10,10,20,20) needs polygon x polygon
    # DE-9IM for any predicate but None / intersects when their boxes meet
    if pred in (None, "intersects"):
        return None
    for b in boxes:
        if b[0] <= 20 and b[2] >= 10 and b[1] <= 20 and b[3] >= 10:
            return "polygon_de9im"
    return None


def polygon_intersects_cases():
    """Row pairs and textbook answers (tests/test_polygon_predicates.py:732-807;
    the reference computes the expectation with shapely at test time, so these
    are the exact geometric answers)."""
    polys = [box(0, 0, 1, 1), box(2, 2, 3, 3), box(0, 0, 5, 5), box(10, 10, 11, 11),
             poly([(0, 0), (10, 0), (10, 10), (0, 10)], holes=[[(3, 3), (7, 3), (7, 7), (3, 7)]]),
             box(4, 4, 6, 6)]
    return [
        {"source": "tests/test_polygon_predicates.py:732-749", "left": polys, "right": polys,
         "pairs": [[0, 1], [0, 2], [2, 3], [3, 0]], "expected": [False, True, False, False]},
        {"source": "tests/test_polygon_predicates.py:751-780 (box strictly inside the hole)", "left": polys,
         "right": polys, "pairs": [[0, 1], [0, 2], [2, 3], [3, 0], [4, 5], [2, 5]],
         "expected": [False, True, False, False, False, True]},
        {"source": "tests/test_polygon_predicates.py:782-807",
         "left": [mpoly(box(0, 0, 1, 1), box(5, 5, 6, 6)), mpoly(box(10, 10, 11, 11))],
         "right": [box(0.5, 0.5, 2, 2), box(3, 3, 4, 4)],
         "pairs": [[0, 0], [0, 1], [1, 0]], "expected": [True, False, False]},
    ]


def upstream_joins():
    """Counts pinned by the vendored GeoPandas sjoin tests (GEOS results)."""
    return [
        {"source": "tests/upstream/geopandas/tools/tests/test_sjoin.py:839-898",
         "dataset": "nybb_16a.zip", "points": "diagonal_21",
         "expect": {"inner_intersects_rows": 11, "left_rows": 21, "right_rows": 12,
                    "within_point1_boro": "Staten Island"}},
        {"source": "tests/upstream/geopandas/tools/tests/test_sjoin.py:1039-1044",
         "dataset": "naturalearth_cities x naturalearth_lowres",
         "expect": {"inner_intersects_shape": [213, 4]}},
    ]


def main():
    out = {
        "pip": pip_cases(),
        "query": query_cases(),
        "sindex_table": sindex_table(),
        "polygon_intersects": polygon_intersects_cases(),
        "upstream_joins": upstream_joins(),
    }
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
