"""Readers into owned SoA arrays: GeoParquet (WKB columns decoded on the GPU,
GeoArrow native point / polygon / multipolygon columns by buffer views) and a
minimal ESRI shapefile reader (points and polygons).

GeoParquet follows the reference's read path (io/geoparquet.py:3430
read_geoparquet: pyarrow scan, the geometry column decoded to owned buffers
without per-row host objects; io/wkb.py:5049 decode_wkb_owned); the writer
emits WKB with GeoParquet 1.0 "geo" metadata.

Shapefiles: minimal ESRI shapefile reader (points and polygons) into owned SoA arrays.

Enough of the format to load the reference's upstream test datasets
(tests/upstream/geopandas/tests/data/: nybb_16a.zip, naturalearth_*), which
the reference reads through pyogrio (src/vibespatial/io/file.py).  Ring
grouping follows the shapefile convention: a clockwise ring starts a new
part (shell), counter-clockwise rings that follow it are its holes.
Attributes come from the .dbf (character, numeric, float, logical, date).
"""

from __future__ import annotations

import io as _io
import os
import struct
import zipfile

import numpy as np
import pandas as pd

from .geometry import GeometryArray, PointArray, PolygonArray

__all__ = ["read_file", "read_shapefile_bytes", "read_parquet", "to_parquet", "decode_wkb", "decode_wkb_device",
           "encode_wkb"]


def _signed_area2(x: np.ndarray, y: np.ndarray) -> float:
    return float(np.dot(x[:-1], y[1:]) - np.dot(x[1:], y[:-1]))


def _read_shp(data: bytes):
    if len(data) < 100 or struct.unpack(">i", data[:4])[0] != 9994:
        raise ValueError("not a shapefile")
    shape_type = struct.unpack("<i", data[32:36])[0]
    pos = 100
    records = []
    while pos + 8 <= len(data):
        _, clen = struct.unpack(">ii", data[pos:pos + 8])
        rec = data[pos + 8: pos + 8 + 2 * clen]
        pos += 8 + 2 * clen
        st = struct.unpack("<i", rec[:4])[0]
        if st == 0:
            records.append(None)
        elif st in (1, 11, 21):
            records.append(("point", struct.unpack("<2d", rec[4:20])))
        elif st in (5, 15, 25):
            n_parts, n_points = struct.unpack("<2i", rec[36:44])
            parts = np.frombuffer(rec, dtype="<i4", count=n_parts, offset=44)
            pts = np.frombuffer(rec, dtype="<f8", count=2 * n_points, offset=44 + 4 * n_parts).reshape(-1, 2)
            bounds = list(parts) + [n_points]
            rings = [pts[bounds[k]: bounds[k + 1]] for k in range(n_parts)]
            records.append(("polygon", rings))
        else:
            raise NotImplementedError(f"shape type {st}")
    return shape_type, records


def _read_dbf(data: bytes) -> pd.DataFrame:
    n_rec, hdr_len, rec_len = struct.unpack("<IHH", data[4:12])
    fields = []
    pos = 32
    while data[pos] != 0x0D:
        name = data[pos:pos + 11].split(b"\x00")[0].decode("latin-1")
        ftype = chr(data[pos + 11])
        flen, fdec = data[pos + 16], data[pos + 17]
        fields.append((name, ftype, flen, fdec))
        pos += 32
    cols = {f[0]: [] for f in fields}
    for r in range(n_rec):
        rec = data[hdr_len + r * rec_len: hdr_len + (r + 1) * rec_len]
        off = 1
        for name, ftype, flen, fdec in fields:
            raw = rec[off:off + flen].decode("utf-8", errors="replace").strip()
            off += flen
            if ftype in "NF":
                try:
                    v = float(raw) if (fdec or ftype == "F" or "." in raw) else int(raw)
                except ValueError:
                    v = np.nan
            elif ftype == "L":
                v = raw.upper() in ("Y", "T")
            else:
                v = raw
            cols[name].append(v)
    return pd.DataFrame(cols)


def read_shapefile_bytes(shp: bytes, dbf: bytes | None = None):
    """(GeometryArray, attribute DataFrame) from raw .shp/.dbf bytes."""
    _, records = _read_shp(shp)
    kinds = {r[0] for r in records if r is not None}
    if kinds == {"point"}:
        xy = np.array([r[1] if r is not None else (np.nan, np.nan) for r in records], dtype=np.float64)
        ga = GeometryArray.from_points(PointArray(xy[:, 0], xy[:, 1]))
    else:
        xs, ys, ring_off, part_off, geom_off = [], [], [0], [0], [0]
        n = 0
        for r in records:
            if r is not None and r[0] == "polygon":
                for k, ring in enumerate(r[1]):
                    x = np.ascontiguousarray(ring[:, 0])
                    y = np.ascontiguousarray(ring[:, 1])
                    if len(x) and (x[0] != x[-1] or y[0] != y[-1]):
                        x = np.append(x, x[0])
                        y = np.append(y, y[0])
                    clockwise = _signed_area2(x, y) < 0
                    if k > 0 and clockwise:
                        part_off.append(len(ring_off) - 1)
                    xs.append(x)
                    ys.append(y)
                    n += len(x)
                    ring_off.append(n)
                if r[1]:
                    part_off.append(len(ring_off) - 1)
            geom_off.append(len(part_off) - 1)
        pa = PolygonArray(np.concatenate(xs) if xs else np.empty(0), np.concatenate(ys) if ys else np.empty(0),
                          np.asarray(ring_off), np.asarray(part_off), np.asarray(geom_off))
        ga = GeometryArray.from_polygons(pa)
    attrs = _read_dbf(dbf) if dbf else pd.DataFrame(index=range(len(records)))
    return ga, attrs


def build_read_index(gdf) -> None:
    """Build the frame's spatial index now, on the device, from the decoded
    buffers (io/file.py:1786-1853 _attach_gpu_spatial_index, the reference's
    build_index=True): one packed Hilbert tree per geometry family present."""
    idx = gdf.sindex
    ga = gdf.geometry.values
    for fam, arr in (("point", ga.points), ("polygon", ga.polygons)):
        if arr is not None:
            idx._index.tree(fam)


def read_file(path: str, crs=None, build_index: bool = False):
    """GeoDataFrame from a .shp path, a directory holding one, or a .zip of one;
    build_index=True also builds its spatial index on the device."""
    from .api import GeoDataFrame, GeoSeries

    if path.endswith(".zip"):
        with zipfile.ZipFile(path) as z:
            names = z.namelist()
            shp_name = next(n for n in names if n.lower().endswith(".shp"))
            stem = shp_name[:-4]
            shp = z.read(shp_name)
            dbf = z.read(stem + ".dbf") if stem + ".dbf" in names else None
    else:
        if os.path.isdir(path):
            shp_name = next(f for f in sorted(os.listdir(path)) if f.lower().endswith(".shp"))
            path = os.path.join(path, shp_name)
        with open(path, "rb") as f:
            shp = f.read()
        dbf_path = path[:-4] + ".dbf"
        dbf = open(dbf_path, "rb").read() if os.path.exists(dbf_path) else None
    ga, attrs = read_shapefile_bytes(shp, dbf)
    gdf = GeoDataFrame(attrs, geometry=GeoSeries(ga, crs=crs), crs=crs)
    if build_index:
        build_read_index(gdf)
    return gdf


_ = _io


# ---------------------------------------------------------------------------
# WKB
# ---------------------------------------------------------------------------
_WKB_ERRORS = {1: "truncated record", 2: "unsupported geometry type or dimension (Point, Polygon, MultiPolygon "
               "in 2D are read)", 3: "bad byte-order byte", 4: "MultiPolygon member is not a Polygon",
               5: "count over 2^31", 6: "trailing bytes after the geometry"}


def encode_wkb(ga: GeometryArray, little_endian: bool = True):
    """(payload uint8, offsets int64 (n + 1), validity uint8) of ISO WKB records:
    Point, Polygon (one part; no rings when empty) or MultiPolygon (several
    parts); null rows are empty records with validity 0.  Point records are
    written in bulk; polygon records one row at a time."""
    bo = "<" if little_endian else ">"
    flag = b"\x01" if little_endian else b"\x00"
    n = len(ga)
    lens = np.zeros(n, dtype=np.int64)
    valid = np.zeros(n, dtype=np.uint8)
    P = ga.polygons
    recs = {}
    if P is not None:
        def poly_body(rings):
            out = [struct.pack(bo + "I", len(rings))]
            for b0, b1 in rings:
                out.append(struct.pack(bo + "I", b1 - b0))
                xy = np.empty(2 * (b1 - b0), dtype=bo + "f8")
                xy[0::2], xy[1::2] = P.x[b0:b1], P.y[b0:b1]
                out.append(xy.tobytes())
            return b"".join(out)

        ro, po, go = P.ring_offsets, P.part_offsets, P.geom_offsets
        for g, row in enumerate(ga.polygon_rows.tolist()):
            parts = [[(int(ro[r]), int(ro[r + 1])) for r in range(po[p], po[p + 1])]
                     for p in range(go[g], go[g + 1])]
            if len(parts) <= 1:
                rec = flag + struct.pack(bo + "I", 3) + poly_body(parts[0] if parts else [])
            else:
                rec = flag + struct.pack(bo + "II", 6, len(parts)) + b"".join(
                    flag + struct.pack(bo + "I", 3) + poly_body(r) for r in parts)
            recs[row] = rec
            lens[row] = len(rec)
            valid[row] = 1
    if ga.points is not None:
        lens[ga.point_rows] = 21
        valid[ga.point_rows] = 1
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offsets[1:])
    payload = np.empty(int(offsets[-1]), dtype=np.uint8)
    if ga.points is not None and len(ga.point_rows):
        rec = np.zeros(len(ga.point_rows), dtype=np.dtype([("f", "u1"), ("t", bo + "u4"), ("x", bo + "f8"),
                                                           ("y", bo + "f8")]))
        rec["f"] = 1 if little_endian else 0
        rec["t"] = 1
        rec["x"], rec["y"] = ga.points.x, ga.points.y
        payload[offsets[ga.point_rows][:, None] + np.arange(21)] = rec.view(np.uint8).reshape(-1, 21)
    for row, rec in recs.items():
        payload[offsets[row]: offsets[row + 1]] = np.frombuffer(rec, dtype=np.uint8)
    return payload, offsets, valid


class DeviceWkbColumn:
    """Decoded WKB column on the device: point x / y with their record
    numbers, polygon SoA buffers with theirs (None / empty when absent)."""

    def __init__(self, n_rows, points, point_rows, polygons, polygon_rows):
        self.n_rows = n_rows
        self.points = points
        self.point_rows = point_rows
        self.polygons = polygons
        self.polygon_rows = polygon_rows


def decode_wkb_device(d_payload, d_offsets, d_validity=None, workspace=None, timing=None) -> DeviceWkbColumn:
    """Decode WKB records already in device memory: d_payload uint8,
    d_offsets int64 (n + 1), d_validity uint8 (n) or None.  Two passes
    (vs_wkb_scan, then vs_wkb_fill into buffers sized from its totals); the
    host reads the eight-word result in between.  ``workspace`` may pass a
    reusable uint8 device tensor of at least vs_wkb_workspace_bytes(n); a
    ``timing`` list receives the (start, end) device-event pairs of the scan
    and of the fill, recorded on the launch stream."""
    import torch

    from ._lib import check, lib
    from .device import DevicePoints, DevicePolygons, ptr, stream

    L = lib()
    n = int(d_offsets.numel()) - 1
    if n < 0:
        raise ValueError("offsets must hold n + 1 entries")
    if d_payload.dtype != torch.uint8 or d_offsets.dtype != torch.int64:
        raise TypeError("payload must be uint8 and offsets int64")
    if not (d_payload.is_cuda and d_offsets.is_cuda and d_payload.is_contiguous() and d_offsets.is_contiguous()):
        raise ValueError("payload and offsets must be contiguous device tensors")
    if d_validity is not None and (d_validity.dtype != torch.uint8 or d_validity.numel() != n):
        raise ValueError("validity must be uint8 with one entry per record")
    dev = d_payload.device
    need = max(int(L.vs_wkb_workspace_bytes(n)), 1)
    ws = workspace if workspace is not None and workspace.numel() >= need else \
        torch.empty(need, dtype=torch.uint8, device=dev)
    res = torch.empty(8, dtype=torch.int64, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)] if timing is not None else None
    if ev:
        ev[0].record()
    # the kernels bound every record by the payload's size (a record outside
    # it is reported as truncated), so device offsets need no host check
    nbytes = int(d_payload.numel())
    check(L.vs_wkb_scan(ptr(d_payload), nbytes, ptr(d_offsets), ptr(d_validity), n, ptr(res), ptr(ws), ws.numel(),
                        stream()), "vs_wkb_scan")
    if ev:
        ev[1].record()
    # the result words and the column's first and last offsets in one read
    r = torch.cat([res, d_offsets[:1], d_offsets[n:]]).cpu().tolist() if n else res.cpu().tolist() + [0, 0]
    if r[8] < 0 or r[9] > nbytes:
        raise ValueError(f"offsets out of the payload: first {r[8]}, last {r[9]}, payload {nbytes} bytes")
    if r[5] >= 0:
        raise ValueError(f"WKB record {r[5]}: {_WKB_ERRORS.get(int(r[6]), 'invalid record')}")
    if r[7]:
        raise OverflowError("WKB column holds more than 2^31-1 parts, rings or coordinates")
    n_pts, n_poly, n_parts, n_rings, n_coords = (int(v) for v in r[:5])
    f64 = dict(dtype=torch.float64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    i64 = dict(dtype=torch.int64, device=dev)
    px, py = torch.empty(max(n_pts, 1), **f64), torch.empty(max(n_pts, 1), **f64)
    prow = torch.empty(max(n_pts, 1), **i64)
    x, y = torch.empty(max(n_coords, 1), **f64), torch.empty(max(n_coords, 1), **f64)
    geom = torch.empty(n_poly + 1, **i32)
    part = torch.empty(n_parts + 1, **i32)
    ring = torch.empty(n_rings + 1, **i32)
    grow = torch.empty(max(n_poly, 1), **i64)
    if ev:
        ev[2].record()
    check(L.vs_wkb_fill(ptr(d_payload), nbytes, ptr(d_offsets), ptr(d_validity), n, ptr(ws), ws.numel(), ptr(px),
                        ptr(py), ptr(prow), ptr(x), ptr(y), ptr(geom), ptr(part), ptr(ring), ptr(grow), stream()),
          "vs_wkb_fill")
    if ev:
        ev[3].record()
        timing.append(((ev[0], ev[1]), (ev[2], ev[3])))
    points = DevicePoints(px[:n_pts], py[:n_pts]) if n_pts else None
    polygons = DevicePolygons(x[:n_coords], y[:n_coords], ring, part, geom) if n_poly else None
    return DeviceWkbColumn(n, points, prow[:n_pts], polygons, grow[:n_poly])


def decode_wkb(payload, offsets, validity=None) -> GeometryArray:
    """GeometryArray from WKB records decoded on the GPU: record r is
    payload[offsets[r]:offsets[r + 1]]; validity 0 or an empty record is a null
    row.  The decoded buffers stay on the device (cached on the arrays) and are
    also copied to the host arrays (io/wkb.py:5049 decode_wkb_owned)."""
    import torch

    from .device import require_gpu

    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n = len(off) - 1
    if n < 0:
        raise ValueError("offsets must hold n + 1 entries")
    data = np.frombuffer(payload, dtype=np.uint8) if isinstance(payload, (bytes, bytearray)) else \
        np.ascontiguousarray(payload, dtype=np.uint8)
    if n and (off[0] < 0 or off[-1] > len(data) or np.any(np.diff(off) < 0)):
        raise ValueError("offsets out of the payload or not non-decreasing")
    if validity is not None and len(validity) != n:
        raise ValueError("validity must hold one entry per record")
    dev = require_gpu()
    d_data = torch.from_numpy(data.copy() if len(data) else np.zeros(1, np.uint8)).to(dev)
    d_off = torch.from_numpy(off).to(dev)
    d_valid = None if validity is None else torch.from_numpy(np.ascontiguousarray(validity, dtype=np.uint8)).to(dev)
    col = decode_wkb_device(d_data, d_off, d_valid)
    tags = np.full(n, -1, dtype=np.int8)
    points = polygons = prows = grows = None
    if col.points is not None:
        Q = col.points
        points = PointArray(Q.x.cpu().numpy(), Q.y.cpu().numpy())
        points._device = Q
        prows = col.point_rows.cpu().numpy()
        tags[prows] = 0
    if col.polygons is not None:
        D = col.polygons
        polygons = PolygonArray(D.x.cpu().numpy(), D.y.cpu().numpy(), D.ring_offsets.cpu().numpy(),
                                D.part_offsets.cpu().numpy(), D.geom_offsets.cpu().numpy())
        polygons._device = D
        grows = col.polygon_rows.cpu().numpy()
        tags[grows] = 1
    return GeometryArray(tags, points, prows, polygons, grows)


# ---------------------------------------------------------------------------
# GeoParquet
# ---------------------------------------------------------------------------
def _lists(arr):
    """(offsets rebased to 0, child values) of an Arrow list array."""
    off = np.asarray(arr.offsets, dtype=np.int64)
    return off - off[0], arr.values.slice(int(off[0]), int(off[-1] - off[0]))


def _xy(arr):
    """x, y float64 of GeoArrow coordinates (separated struct or interleaved)."""
    import pyarrow as pa

    if pa.types.is_struct(arr.type):
        return (np.asarray(arr.field("x").fill_null(np.nan), dtype=np.float64),
                np.asarray(arr.field("y").fill_null(np.nan), dtype=np.float64))
    flat = np.asarray(arr.flatten(), dtype=np.float64).reshape(-1, arr.type.list_size)
    return np.ascontiguousarray(flat[:, 0]), np.ascontiguousarray(flat[:, 1])


def _geoarrow_native(col, encoding: str) -> GeometryArray:
    """GeoArrow native point, polygon or multipolygon column -> GeometryArray
    (host buffer views).  Null rows hold empty lists and map to no family."""
    n = len(col)
    rows = np.flatnonzero(np.asarray(col.is_valid()))
    tags = np.full(n, -1, np.int8)
    if encoding == "point":
        xs, ys = _xy(col)
        tags[rows] = 0
        return GeometryArray(tags, PointArray(xs[rows], ys[rows]), rows, None, None)
    if encoding == "polygon":
        geom_rings, rings = _lists(col)
        ring_coords, coords = _lists(rings)
        nr = np.diff(geom_rings)
        if np.any(nr[np.asarray(col.is_null())] != 0):
            raise ValueError("null polygon rows with rings")
        has = nr[rows] > 0  # an empty polygon has no parts
        geom = np.concatenate([[0], np.cumsum(has)])
        part = np.concatenate([geom_rings[rows][has], geom_rings[-1:]])
    else:
        geom_parts, parts = _lists(col)
        if np.any(np.diff(geom_parts)[np.asarray(col.is_null())] != 0):
            raise ValueError("null multipolygon rows with parts")
        part, rings = _lists(parts)
        ring_coords, coords = _lists(rings)
        geom = np.concatenate([geom_parts[rows], geom_parts[-1:]])
    xs, ys = _xy(coords)
    tags[rows] = 1
    return GeometryArray(tags, None, None, PolygonArray(xs, ys, ring_coords, part, geom), rows)


def read_parquet(path, columns=None, crs=None, build_index: bool = False):
    """GeoDataFrame from a GeoParquet file: WKB geometry decoded on the GPU,
    GeoArrow native point / polygon / multipolygon columns by buffer views;
    other columns through pandas (io/geoparquet.py:3430 read_geoparquet).
    build_index=True also builds the spatial index on the device from the
    decoded buffers before returning."""
    import json

    import pyarrow as pa
    import pyarrow.parquet as pq

    from .api import GeoDataFrame, GeoSeries

    table = pq.read_table(path, columns=columns)
    raw = (table.schema.metadata or {}).get(b"geo")
    if raw is None:
        raise ValueError("no GeoParquet 'geo' metadata in the file")
    meta = json.loads(raw)
    name = meta["primary_column"]
    cmeta = meta["columns"][name]
    encoding = str(cmeta.get("encoding", "WKB"))
    col = table.column(name).combine_chunks()
    if encoding.upper() == "WKB":
        if isinstance(col, pa.ChunkedArray):
            col = col.chunk(0) if col.num_chunks == 1 else pa.concat_arrays(col.chunks)
        col = col.cast(pa.large_binary())
        _, off_buf, data_buf = col.buffers()
        off = np.frombuffer(off_buf, dtype=np.int64)[col.offset: col.offset + len(col) + 1]
        data = np.frombuffer(data_buf, dtype=np.uint8) if data_buf is not None else np.zeros(0, np.uint8)
        ga = decode_wkb(data, off, np.asarray(col.is_valid(), dtype=np.uint8))
    elif encoding in ("point", "polygon", "multipolygon"):
        ga = _geoarrow_native(col, encoding)
    else:
        raise NotImplementedError(f"GeoParquet encoding {encoding!r}")
    rest = table.drop([name])
    attrs = rest.to_pandas() if rest.num_columns else pd.DataFrame(index=pd.RangeIndex(len(ga)))
    crs = crs if crs is not None else cmeta.get("crs")
    gdf = GeoDataFrame(attrs, geometry=GeoSeries(ga, crs=crs), crs=crs)
    gdf._geom_name = name
    if build_index:
        build_read_index(gdf)
    return gdf


def to_parquet(gdf, path, little_endian: bool = True):
    """Write a GeoDataFrame as GeoParquet 1.0 with a WKB geometry column."""
    import json

    import pyarrow as pa
    import pyarrow.parquet as pq

    ga = gdf.geometry.values
    payload, offsets, valid = encode_wkb(ga, little_endian)
    geom = pa.LargeBinaryArray.from_buffers(pa.large_binary(), len(ga), [
        pa.py_buffer(np.packbits(valid.astype(bool), bitorder="little").tobytes()),
        pa.py_buffer(offsets.tobytes()), pa.py_buffer(payload.tobytes())], null_count=int((valid == 0).sum()))
    attrs = pa.Table.from_pandas(gdf._df.reset_index(drop=True), preserve_index=False)
    name = gdf._geom_name
    table = attrs.append_column(name, geom)
    types = sorted({{0: "Point", 1: "Polygon"}[int(t)] for t in np.unique(ga.tags) if t >= 0})
    meta = {"version": "1.0.0", "primary_column": name,
            "columns": {name: {"encoding": "WKB", "geometry_types": types}}}
    table = table.replace_schema_metadata({**(table.schema.metadata or {}), b"geo": json.dumps(meta).encode()})
    pq.write_table(table, path)
