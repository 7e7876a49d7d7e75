"""Argument checks of the public API (no GPU needed: they raise before any
device work), following api/tools/sjoin.py:151-164, 268-321 and
api/sindex.py:122-142."""

from __future__ import annotations

import warnings

import numpy as np
import pytest

from vibespatial_amd import GeoDataFrame, Point, box, sjoin
from vibespatial_amd.geometry import GeometryArray, PointArray, PolygonArray, as_geometry_array


def _frames():
    left = GeoDataFrame({"a": [1, 2], "geometry": [Point(0, 0), Point(1, 1)]})
    right = GeoDataFrame({"b": [3], "geometry": [box(0, 0, 2, 2)]})
    return left, right


def test_sjoin_rejects_unknown_kwargs():
    left, right = _frames()
    with pytest.raises(TypeError, match="unexpected keyword argument 'foo'"):
        sjoin(left, right, foo=1)


def test_sjoin_rejects_bad_how():
    left, right = _frames()
    with pytest.raises(ValueError, match="`how` was"):
        sjoin(left, right, how="cross")


def test_sjoin_rejects_non_frames():
    left, right = _frames()
    with pytest.raises(ValueError, match="'left_df' should be GeoDataFrame"):
        sjoin(left.to_pandas(), right)
    with pytest.raises(ValueError, match="'right_df' should be GeoDataFrame"):
        sjoin(left, right.to_pandas())


def test_sjoin_on_attribute_checks():
    left, right = _frames()
    with pytest.raises(ValueError, match="missing from both"):
        sjoin(left, right, on_attribute="zzz")
    with pytest.raises(ValueError, match="missing from the right"):
        sjoin(left, right, on_attribute="a")


def test_sjoin_crs_mismatch_warns_then_validates_predicate():
    left, right = _frames()
    left.crs = "EPSG:4326"
    right.crs = "EPSG:3857"
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        with pytest.raises(ValueError, match="predicate"):
            sjoin(left, right, predicate="nope")
    assert any("CRS mismatch" in str(x.message) for x in w)


def test_sindex_predicate_and_distance_validation():
    _, right = _frames()
    si = right.sindex
    with pytest.raises(ValueError, match="`predicate` must be one of"):
        si.query(box(0, 0, 1, 1), predicate="equals_exact")
    with pytest.raises(ValueError, match="'distance' parameter is required"):
        si.query(box(0, 0, 1, 1), predicate="dwithin")
    with pytest.raises(ValueError, match="only supported in combination"):
        si.query(box(0, 0, 1, 1), predicate="intersects", distance=1.0)
    with pytest.raises(ValueError, match="Invalid output_format"):
        si.query(box(0, 0, 1, 1), output_format="dataarray")
    with pytest.raises(TypeError):
        si.query("notavalidgeom")
    assert si.valid_query_predicates >= {None, "intersects", "within", "contains", "covers", "touches"}


def test_geometry_containers_roundtrip():
    ga = as_geometry_array([Point(1, 2), None, box(0, 0, 1, 1), Point()])
    assert list(ga.tags) == [0, -1, 1, 0]
    assert ga[0] == Point(1, 2) and ga[1] is None
    b = ga.bounds()
    assert np.allclose(b[2], [0, 0, 1, 1]) and np.isnan(b[1]).all() and np.isnan(b[3]).all()
    pa = PolygonArray.from_geometries([box(0, 0, 1, 1), None, box(2, 2, 3, 4)])
    assert len(pa) == 3 and np.isnan(pa.bounds()[1]).all()
    assert np.allclose(pa.bounds()[2], [2, 2, 3, 4])
    assert isinstance(as_geometry_array(PointArray([0.0], [1.0])), GeometryArray)
