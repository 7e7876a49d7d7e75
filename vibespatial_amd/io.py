"""Minimal ESRI shapefile reader (points and polygons) into owned SoA arrays.

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

__all__ = ["read_file", "read_shapefile_bytes"]


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


def read_file(path: str, crs=None):
    """GeoDataFrame from a .shp path, a directory holding one, or a .zip of one."""
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
    return GeoDataFrame(attrs, geometry=GeoSeries(ga, crs=crs), crs=crs)


_ = _io
