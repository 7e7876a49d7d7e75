"""GPU WKB ingest (vs_wkb_scan / vs_wkb_fill) against the WKB oracle: mixed
columns in both byte orders with nulls, empties, holes and multipolygons
bit-exact; SRID-tagged and mixed-order records; the first bad record of a
column; GeoParquet round trips feeding sjoin; a large column checked by the
encode -> decode identity."""

from __future__ import annotations

import struct

import numpy as np
import pytest

from oracle import wkb_ref as W
from vibespatial_amd import GeoDataFrame, sjoin
from vibespatial_amd import synthetic as S
from vibespatial_amd.geometry import GeometryArray, MultiPolygon, Point, PointArray, Polygon, box
from vibespatial_amd.io import decode_wkb, encode_wkb, read_parquet

pytestmark = pytest.mark.gpu

FIELDS = ("point_x", "point_y", "point_rows", "x", "y", "ring_offsets", "part_offsets", "geom_offsets",
          "polygon_rows")


def _buffers(ga: GeometryArray):
    P, Q = ga.polygons, ga.points
    e64, ei = np.zeros(0, np.float64), np.zeros(0, np.int64)
    return {"point_x": Q.x if Q else e64, "point_y": Q.y if Q else e64,
            "point_rows": ga.point_rows if Q else ei, "x": P.x if P else e64, "y": P.y if P else e64,
            "ring_offsets": P.ring_offsets if P else np.zeros(1, np.int32),
            "part_offsets": P.part_offsets if P else np.zeros(1, np.int32),
            "geom_offsets": P.geom_offsets if P else np.zeros(1, np.int32),
            "polygon_rows": ga.polygon_rows if P else ei}


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float64:  # bit patterns, so NaN payloads and -0.0 count
        return a.shape == b.shape and np.array_equal(a.view(np.int64), np.asarray(b, np.float64).view(np.int64))
    return np.array_equal(a.astype(np.int64), b.astype(np.int64))


def _check(payload, offsets, validity=None):
    got = _buffers(decode_wkb(payload, offsets, validity))
    exp = W.decode_column(payload, offsets, validity)
    for k in FIELDS:
        assert _same(got[k], exp[k]), k
    return got


def _mixed(rng, n):
    geoms = []
    for _ in range(n):
        r = rng.random()
        if r < 0.08:
            geoms.append(None)
        elif r < 0.45:
            geoms.append(Point(*rng.uniform(-1e6, 1e6, 2)))
        elif r < 0.48:
            geoms.append(Point())  # POINT EMPTY: NaN, NaN
        elif r < 0.53:
            geoms.append(Polygon([]))
        elif r < 0.75:
            x0, y0 = rng.uniform(0, 100, 2)
            k = int(rng.integers(3, 40))
            t = np.sort(rng.uniform(0, 2 * np.pi, k))
            ring = list(zip(x0 + 3 * np.cos(t), y0 + 3 * np.sin(t)))
            geoms.append(Polygon(ring + ring[:1], [[(x0 - .5, y0 - .5), (x0 + .5, y0 - .5), (x0, y0 + .5)]]))
        else:
            geoms.append(MultiPolygon([box(*(c := rng.uniform(0, 100, 2)), *(c + rng.uniform(0.5, 2, 2)))
                                       for _ in range(int(rng.integers(2, 6)))]))
    return GeometryArray.from_scalars(geoms)


@pytest.mark.parametrize("little_endian", [True, False])
@pytest.mark.parametrize("n", [1, 255, 256, 257, 4000])
def test_mixed_column_matches_oracle(little_endian, n):
    print(f"wkb mixed column n={n} little_endian={little_endian}", flush=True)
    rng = np.random.default_rng(n + little_endian)
    ga = _mixed(rng, n)
    payload, offsets, valid = encode_wkb(ga, little_endian)
    got = _check(payload, offsets, valid)
    for k, v in _buffers(ga).items():
        assert _same(got[k], v), k
    # without the validity array the null rows are the zero-length records
    _check(payload, offsets, None)


def test_srid_mixed_byte_orders_and_sliced_offsets():
    print("wkb srid / mixed byte order", flush=True)
    recs = [
        b"\x01" + struct.pack("<IIdd", 0x20000001, 4326, 10.5, -3.0),
        b"\x00" + struct.pack(">Idd", 1, -7.25, 1e300),
    ]
    shell = [(0, 0), (4, 0), (4, 4), (0, 4), (0, 0)]
    body = struct.pack(">II", 1, len(shell)) + b"".join(struct.pack(">dd", *p) for p in shell)
    member = b"\x00" + struct.pack(">I", 3) + body
    recs.append(b"\x01" + struct.pack("<III", 0x20000006, 3857, 2) + member + member)
    recs.append(b"\x00" + struct.pack(">II", 0x20000003, 4326) + body)
    recs.append(b"")  # null
    recs.append(b"\x01" + struct.pack("<II", 6, 0))  # MULTIPOLYGON EMPTY
    pad = b"\xee" * 7  # records need not start at byte 0
    payload = np.frombuffer(pad + b"".join(recs), np.uint8)
    offsets = np.concatenate([[0], np.cumsum([len(r) for r in recs])]) + len(pad)
    got = _check(payload, offsets)
    assert got["point_rows"].tolist() == [0, 1] and got["polygon_rows"].tolist() == [2, 3, 5]
    assert got["geom_offsets"].tolist() == [0, 2, 3, 3]


@pytest.mark.parametrize("code,mutate", [
    (W.TYPE, lambda b, o: b.__setitem__(o + 1, 2)),
    (W.BYTE_ORDER, lambda b, o: b.__setitem__(o, 7)),
    (W.TYPE, lambda b, o: b.__setitem__(o + 2, 0x03)),  # type word 769 (point) / 771 (polygon): unknown
])
def test_first_bad_record_is_reported(code, mutate):
    print(f"wkb bad record code={code}", flush=True)
    rng = np.random.default_rng(9)
    ga = _mixed(rng, 3000)
    payload, offsets, valid = encode_wkb(ga)
    rows = np.flatnonzero(valid)
    bad = bytearray(payload.tobytes())
    mutate(bad, int(offsets[rows[1700]]))
    mutate(bad, int(offsets[rows[2500]]))
    data = np.frombuffer(bytes(bad), np.uint8)
    exp = W.decode_column(data, offsets, valid)
    assert exp[0] == "error" and exp[1] == rows[1700]
    with pytest.raises(ValueError, match=f"WKB record {rows[1700]}:"):
        decode_wkb(data, offsets, valid)


@pytest.mark.parametrize("rec", [
    b"\x01" + struct.pack("<Idd", 1, 1.0, 2.0)[:-3],  # truncated
    b"\x01" + struct.pack("<Idd", 1001, 1.0, 2.0),  # ISO Point Z
    b"\x01" + struct.pack("<Idd", 0x40000001, 1.0, 2.0),  # EWKB Point M
    b"\x01" + struct.pack("<II", 6, 1) + b"\x01" + struct.pack("<Idd", 1, 1.0, 2.0),  # member not polygon
    b"\x01" + struct.pack("<Idd", 1, 1.0, 2.0) + b"\x00",  # trailing byte
    b"\x01" + struct.pack("<III", 3, 1, 0x7FFFFFFF),  # ring count past the record
    b"\x01" + struct.pack("<II", 6, 0xFFFFFFFF),  # member count past the record
])
def test_single_bad_record_after_good_ones(rec):
    good = b"\x01" + struct.pack("<Idd", 1, 3.0, 4.0)
    recs = [good] * 300 + [rec] + [good] * 5
    payload = np.frombuffer(b"".join(recs), np.uint8)
    offsets = np.concatenate([[0], np.cumsum([len(r) for r in recs])])
    exp = W.decode_column(payload, offsets)
    assert exp[:2] == ("error", 300)
    with pytest.raises(ValueError, match="WKB record 300:"):
        decode_wkb(payload, offsets)


def test_empty_and_all_null_columns():
    ga = decode_wkb(np.zeros(0, np.uint8), np.zeros(1, np.int64))
    assert len(ga) == 0 and ga.points is None and ga.polygons is None
    payload, offsets, valid = encode_wkb(_mixed(np.random.default_rng(1), 50))
    ga = decode_wkb(payload, offsets, np.zeros(50, np.uint8))
    assert (ga.tags == -1).all() and ga.points is None and ga.polygons is None


def test_large_column_round_trip():
    # 2M point records + 20k star polygons, interleaved: encode -> GPU decode
    # is the identity on every buffer
    print("wkb large round trip", flush=True)
    rng = np.random.default_rng(12)
    n_pts, n_poly = 2_000_000, 20_000
    pts = PointArray(rng.uniform(-180, 180, n_pts), rng.uniform(-90, 90, n_pts))
    P = S.star_polygons(rng, rng.uniform(0, 100, n_poly), rng.uniform(0, 100, n_poly), rng.uniform(1, 3, n_poly),
                        rng.integers(3, 64, n_poly))
    rows = rng.permutation(n_pts + n_poly)
    tags = np.zeros(n_pts + n_poly, np.int8)
    prow, grow = np.sort(rows[:n_pts]), np.sort(rows[n_pts:])
    tags[grow] = 1
    ga = GeometryArray(tags, pts, prow, P, grow)
    payload, offsets, valid = encode_wkb(ga)
    out = decode_wkb(payload, offsets, valid)
    exp = _buffers(ga)
    got = _buffers(out)
    for k in FIELDS:
        assert _same(got[k], exp[k]), k
    # decoded arrays keep their device buffers
    assert out.points._device is not None and out.polygons._device is not None


def test_geoparquet_round_trip_feeds_sjoin(tmp_path):
    print("geoparquet round trip + sjoin", flush=True)
    rng = np.random.default_rng(21)
    P = S.star_polygons(rng, rng.uniform(0, 100, 300), rng.uniform(0, 100, 300), rng.uniform(1, 6, 300),
                        rng.integers(4, 30, 300))
    polys = [GeometryArray.from_polygons(P)[i] for i in range(len(P))]
    polys[7] = None
    pts = [Point(*rng.uniform(-5, 105, 2)) for _ in range(20000)]
    pts[3] = None
    left = GeoDataFrame({"a": np.arange(len(pts)), "geometry": pts})
    right = GeoDataFrame({"b": np.arange(len(polys)) * 2, "geometry": polys})
    for le in (True, False):
        left.to_parquet(tmp_path / "l.parquet", little_endian=le)
        right.to_parquet(tmp_path / "r.parquet", little_endian=le)
        l2, r2 = read_parquet(tmp_path / "l.parquet"), read_parquet(tmp_path / "r.parquet")
        assert l2["a"].tolist() == left["a"].tolist() and r2["b"].tolist() == right["b"].tolist()
        for pred in ("intersects", "within"):
            j1 = sjoin(left, right, predicate=pred)
            j2 = sjoin(l2, r2, predicate=pred)
            assert j1.index.tolist() == j2.index.tolist() and j1["index_right"].tolist() == j2["index_right"].tolist()
            assert j2["b"].tolist() == j1["b"].tolist()


def test_read_parquet_builds_the_index(tmp_path):
    # build_index=True: the trees exist on return, built from the decoded
    # device buffers, and answer queries like a lazily built index
    rng = np.random.default_rng(31)
    P = S.star_polygons(rng, rng.uniform(0, 100, 200), rng.uniform(0, 100, 200), rng.uniform(1, 5, 200),
                        rng.integers(4, 20, 200))
    polys = [GeometryArray.from_polygons(P)[i] for i in range(len(P))] + [Point(5.0, 5.0), None]
    GeoDataFrame({"v": np.arange(len(polys)), "geometry": polys}).to_parquet(tmp_path / "m.parquet")
    g1 = read_parquet(tmp_path / "m.parquet", build_index=True)
    assert set(g1.sindex._index._trees) == {"point", "polygon"}
    g2 = read_parquet(tmp_path / "m.parquet")
    q = [Point(*rng.uniform(0, 100, 2)) for _ in range(500)]
    assert g1.sindex.query(q, predicate="intersects").tolist() == g2.sindex.query(q, predicate="intersects").tolist()


def test_device_offsets_outside_the_payload_are_reported():
    """decode_wkb_device takes offsets as they are: each record is bounded by
    the payload's size in the kernels (reported as truncated), and the
    column's first / last offsets are checked against it."""
    import torch

    from vibespatial_amd.io import decode_wkb_device

    ga = GeometryArray.from_scalars([Point(1, 2), box(0, 0, 1, 1), Point(3, 4)])
    payload, offsets, _ = encode_wkb(ga)
    d_pay = torch.from_numpy(payload).cuda()
    ok = decode_wkb_device(d_pay, torch.from_numpy(offsets).cuda())
    assert ok.points is not None and len(ok.points) == 2
    n = len(payload)
    cases = [
        (np.array([0, 21, n + 4096, n], np.int64), "record 1"),  # a middle record reaching past the payload
        (np.array([0, 21, 10, n], np.int64), "record 1"),  # an end before its start
        (np.array([0, 21, 114, n + 8], np.int64), "offsets out of the payload"),
        (np.array([-16, 21, 114, n], np.int64), "offsets out of the payload"),
    ]
    for off, msg in cases:
        with pytest.raises(ValueError, match=msg):
            decode_wkb_device(d_pay, torch.from_numpy(off).cuda())
