"""WKB / GeoParquet on the CPU: the WKB oracle against the reference's own WKB
vectors and hand-built records, the host encoder against the oracle, and the
GeoArrow-native and writer paths of the GeoParquet layer (no GPU needed)."""

from __future__ import annotations

import json
import struct

import numpy as np
import pytest

from oracle import wkb_ref as W
from vibespatial_amd.geometry import GeometryArray, MultiPolygon, Point, Polygon, box
from vibespatial_amd.io import _geoarrow_native, encode_wkb

# (hex, expected) vectors held by the reference's tests and docstrings
GOLDEN = [
    # tests/upstream/geopandas/tests/test_array.py:157
    ("0101000000CDCCCCCCCCCC1440CDCCCCCCCC0C4A40", ("point", (5.2, 52.1))),
    # src/vibespatial/api/geodataframe.py:4310-4311 (to_arrow example)
    ("0101000000000000000000F03F0000000000000040", ("point", (1.0, 2.0))),
    ("01010000000000000000000040000000000000F03F", ("point", (2.0, 1.0))),
    # tests/test_csv_gpu.py:981, :989
    ("01010000000000000000000840000000000000F03F", ("point", (3.0, 1.0))),
    ("010100000000000000000008400000000000001040", ("point", (3.0, 4.0))),
    # src/vibespatial/api/geoseries.py:1967-1968 (to_wkb example): POINT (0 0), POLYGON EMPTY
    ("010100000000000000000000000000000000000000", ("point", (0.0, 0.0))),
    ("010300000000000000", ("polygon", [])),
]


@pytest.mark.parametrize("hexrec,expected", GOLDEN)
def test_oracle_decodes_reference_vectors(hexrec, expected):
    kind, body = W.decode_record(bytes.fromhex(hexrec))
    assert kind == expected[0]
    if kind == "point":
        assert tuple(body) == expected[1]
    else:
        assert body == expected[1]


def test_oracle_rejects_reference_invalid_linestring():
    # tests/upstream/geopandas/tests/test_array.py:165: a one-point LineString;
    # LineStrings are outside the point / polygon model here
    with pytest.raises(W.WkbError) as e:
        W.decode_record(bytes.fromhex("01020000000100000000000000000008400000000000000840"))
    assert e.value.code == W.TYPE


def _le_point(x, y, t=1):
    return b"\x01" + struct.pack("<Idd", t, x, y)


def test_oracle_hand_built_records():
    # big endian point is the byte reverse of each field
    assert W.decode_record(b"\x00" + struct.pack(">Idd", 1, -7.25, 1e300)) == ("point", (-7.25, 1e300))
    # EWKB SRID-tagged point: type | 0x20000000 then the SRID word
    rec = b"\x01" + struct.pack("<IIdd", 0x20000001, 4326, 10.5, -3.0)
    assert W.decode_record(rec) == ("point", (10.5, -3.0))
    # polygon with a hole, big endian
    shell = [(0, 0), (4, 0), (4, 4), (0, 4), (0, 0)]
    hole = [(1, 1), (2, 1), (2, 2), (1, 1)]
    body = struct.pack(">I", 2)
    for ring in (shell, hole):
        body += struct.pack(">I", len(ring)) + b"".join(struct.pack(">dd", *p) for p in ring)
    kind, parts = W.decode_record(b"\x00" + struct.pack(">I", 3) + body)
    assert kind == "polygon" and len(parts) == 1 and len(parts[0]) == 2
    assert parts[0][1][0].tolist() == [1, 2, 2, 1] and parts[0][1][1].tolist() == [1, 1, 2, 1]
    # multipolygon whose members use the other byte order than the outer record
    member = b"\x00" + struct.pack(">I", 3) + body
    kind, parts = W.decode_record(b"\x01" + struct.pack("<II", 6, 2) + member + member)
    assert kind == "polygon" and len(parts) == 2 and parts[0][0][0].tolist() == [0, 4, 4, 0, 0]


@pytest.mark.parametrize("rec,code", [
    (_le_point(1, 2)[:-1], W.TRUNCATED),
    (b"\x02" + _le_point(1, 2)[1:], W.BYTE_ORDER),
    (_le_point(1, 2, t=1001), W.TYPE),  # ISO Point Z
    (_le_point(1, 2, t=0x80000001), W.TYPE),  # EWKB Point Z
    (_le_point(1, 2, t=7), W.TYPE),  # GeometryCollection
    (b"\x01" + struct.pack("<II", 6, 1) + _le_point(1, 2), W.MEMBER),
    (_le_point(1, 2) + b"\x00", W.TRAILING),
    (b"\x01" + struct.pack("<III", 3, 1, 5) + b"\x00" * 64, W.TRUNCATED),
    (b"\x01", W.TRUNCATED),
])
def test_oracle_error_codes(rec, code):
    with pytest.raises(W.WkbError) as e:
        W.decode_record(rec)
    assert e.value.code == code


def _mixed(rng, n_pts=200, n_polys=120):
    geoms = []
    for _ in range(n_pts + n_polys):
        r = rng.random()
        if r < 0.08:
            geoms.append(None)
        elif r < 0.5:
            geoms.append(Point(*rng.uniform(-1e6, 1e6, 2)))
        elif r < 0.55:
            geoms.append(Polygon([]))
        elif r < 0.75:
            x0, y0 = rng.uniform(0, 100, 2)
            geoms.append(Polygon([(x0, y0), (x0 + 3, y0), (x0 + 3, y0 + 3), (x0, y0 + 3), (x0, y0)],
                                 [[(x0 + 1, y0 + 1), (x0 + 2, y0 + 1), (x0 + 2, y0 + 2), (x0 + 1, y0 + 1)]]))
        else:
            k = int(rng.integers(2, 5))
            geoms.append(MultiPolygon([box(*(c := rng.uniform(0, 100, 2)), *(c + rng.uniform(0.5, 2, 2)))
                                       for _ in range(k)]))
    return GeometryArray.from_scalars(geoms)


def _expected_buffers(ga):
    P = ga.polygons
    return {"point_x": ga.points.x, "point_y": ga.points.y, "point_rows": ga.point_rows,
            "x": P.x, "y": P.y, "ring_offsets": P.ring_offsets, "part_offsets": P.part_offsets,
            "geom_offsets": P.geom_offsets, "polygon_rows": ga.polygon_rows}


@pytest.mark.parametrize("little_endian", [True, False])
def test_encoder_round_trips_through_oracle(little_endian):
    rng = np.random.default_rng(5)
    ga = _mixed(rng)
    payload, offsets, valid = encode_wkb(ga, little_endian)
    assert offsets[0] == 0 and offsets[-1] == len(payload)
    assert np.array_equal(valid, (ga.tags >= 0).astype(np.uint8))
    got = W.decode_column(payload, offsets, valid)
    for k, v in _expected_buffers(ga).items():
        assert np.array_equal(got[k], v), k
    # every record starts with its byte-order flag
    starts = offsets[:-1][valid == 1]
    assert np.all(payload[starts] == (1 if little_endian else 0))


def test_oracle_column_reports_first_bad_row():
    rng = np.random.default_rng(6)
    ga = _mixed(rng, 30, 20)
    payload, offsets, valid = encode_wkb(ga)
    rows = np.flatnonzero(valid)
    bad = bytearray(payload.tobytes())
    bad[offsets[rows[7]] + 1] = 2  # LineString type word
    bad[offsets[rows[12]]] = 5  # bad flag, later row
    assert W.decode_column(np.frombuffer(bytes(bad), np.uint8), offsets, valid) == ("error", int(rows[7]), W.TYPE)


def _write_geoarrow(tmp_path, arr, encoding, name="geom"):
    import pyarrow as pa
    import pyarrow.parquet as pq

    t = pa.table({"id": np.arange(len(arr)), name: arr})
    meta = {"version": "1.1.0", "primary_column": name, "columns": {name: {"encoding": encoding}}}
    t = t.replace_schema_metadata({b"geo": json.dumps(meta).encode()})
    path = tmp_path / f"{encoding}.parquet"
    pq.write_table(t, path)
    return path


def test_geoarrow_native_columns(tmp_path):
    import pyarrow as pa

    from vibespatial_amd.io import read_parquet

    xy = pa.struct([("x", pa.float64()), ("y", pa.float64())])
    pts = pa.array([{"x": 1.0, "y": 2.0}, None, {"x": -3.5, "y": 4.0}], type=xy)
    ga = read_parquet(_write_geoarrow(tmp_path, pts, "point")).geometry.values
    assert ga.tags.tolist() == [0, -1, 0]
    assert ga.points.x.tolist() == [1.0, -3.5] and ga.point_rows.tolist() == [0, 2]

    sq = [{"x": 0.0, "y": 0.0}, {"x": 1.0, "y": 0.0}, {"x": 1.0, "y": 1.0}, {"x": 0.0, "y": 0.0}]
    polys = pa.array([[sq], None, [], [sq, sq]], type=pa.list_(pa.list_(xy)))
    ga = read_parquet(_write_geoarrow(tmp_path, polys, "polygon")).geometry.values
    assert ga.tags.tolist() == [1, -1, 1, 1]
    P = ga.polygons
    assert P.geom_offsets.tolist() == [0, 1, 1, 2] and P.part_offsets.tolist() == [0, 1, 3]
    assert P.ring_offsets.tolist() == [0, 4, 8, 12] and ga.polygon_rows.tolist() == [0, 2, 3]

    mp = pa.array([[[sq], [sq, sq]], None, [[sq]]], type=pa.list_(pa.list_(pa.list_(xy))))
    ga = read_parquet(_write_geoarrow(tmp_path, mp, "multipolygon")).geometry.values
    P = ga.polygons
    assert ga.tags.tolist() == [1, -1, 1]
    assert P.geom_offsets.tolist() == [0, 2, 3] and P.part_offsets.tolist() == [0, 1, 3, 4]
    assert P.ring_offsets.tolist() == [0, 4, 8, 12, 16]
    # interleaved coordinates give the same buffers
    inter = pa.array([[[[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 0.0]]]], type=pa.list_(pa.list_(
        pa.list_(pa.float64(), 2))))
    g2 = _geoarrow_native(inter, "polygon")
    assert g2.polygons.x.tolist() == [0, 1, 1, 0] and g2.polygons.y.tolist() == [0, 0, 1, 0]


def test_writer_emits_geoparquet_wkb(tmp_path):
    import pyarrow.parquet as pq

    from vibespatial_amd import GeoDataFrame

    rng = np.random.default_rng(8)
    ga = _mixed(rng, 40, 30)
    geoms = [ga[i] for i in range(len(ga))]
    gdf = GeoDataFrame({"v": np.arange(len(ga)), "geometry": geoms})
    path = tmp_path / "w.parquet"
    gdf.to_parquet(path)
    t = pq.read_table(path)
    meta = json.loads(t.schema.metadata[b"geo"])
    assert meta["primary_column"] == "geometry" and meta["columns"]["geometry"]["encoding"] == "WKB"
    col = t.column("geometry").combine_chunks()
    payload, offsets, valid = encode_wkb(gdf.geometry.values)
    for r in range(len(ga)):
        rec = col[r].as_py()
        if valid[r]:
            assert rec == payload[offsets[r]:offsets[r + 1]].tobytes()
        else:
            assert rec is None
    assert t.column("v").to_pylist() == list(range(len(ga)))
