"""WKB decode restated in plain Python -- TEST INFRASTRUCTURE ONLY (the
checker for the GPU WKB ingest; the product never imports it).

Follows the reference's host WKB rules:
  * header scan, io/wkb.py:3757 _scan_wkb_value: byte-order flag 0 or 1, type
    word read in that order; EWKB Z (0x80000000) / M (0x40000000) and ISO
    1000/2000/3000 variants are outside the 2D model; an EWKB SRID flag
    (0x20000000) carries a 4-byte SRID after the type word;
  * polygon body, io/wkb.py:3954 _decode_polygon_wkb_payload: ring count, then
    per ring a point count and x, y pairs; a buffer overrun or trailing bytes
    is an error;
  * multipolygon, io/wkb.py:4043 _decode_multipolygon_wkb_payload: member count,
    each member a full Polygon record with its own header.
The reference decodes little-endian ISO records natively and sends big-endian
and SRID-tagged ones through a shapely bridge; both give the same geometry, so
here every record is decoded by the same rules.  Only the families of the
point / polygon geometry model (Point, Polygon, MultiPolygon) are accepted;
the rest are errors with the codes of vsjoin.h's vs_wkb_scan.
"""

from __future__ import annotations

import struct

import numpy as np

TRUNCATED, TYPE, BYTE_ORDER, MEMBER, COUNT, TRAILING = 1, 2, 3, 4, 5, 6


class WkbError(Exception):
    def __init__(self, code):
        super().__init__(code)
        self.code = code


class _Cur:
    def __init__(self, b: bytes):
        self.b, self.pos, self.bo = b, 0, "<"

    def take(self, fmt):
        k = struct.calcsize(fmt)
        if self.pos + k > len(self.b):
            raise WkbError(TRUNCATED)
        v = struct.unpack_from(self.bo + fmt, self.b, self.pos)
        self.pos += k
        return v

    def header(self) -> int:
        if self.pos + 1 > len(self.b):
            raise WkbError(TRUNCATED)
        flag = self.b[self.pos]
        self.pos += 1
        if flag not in (0, 1):
            raise WkbError(BYTE_ORDER)
        self.bo = "<" if flag == 1 else ">"
        (t,) = self.take("I")
        if t & 0xC0000000:
            raise WkbError(TYPE)
        if t & 0x20000000:
            self.take("I")  # SRID
        base = t & 0x1FFFFFFF
        if base >= 1000:
            raise WkbError(TYPE)
        return base


def _polygon(c: _Cur):
    (nr,) = c.take("I")
    rings = []
    for _ in range(nr):
        (npt,) = c.take("I")
        if c.pos + 16 * npt > len(c.b):
            raise WkbError(TRUNCATED)
        xy = np.frombuffer(c.b, dtype=c.bo + "f8", count=2 * npt, offset=c.pos).astype(np.float64)
        c.pos += 16 * npt
        rings.append((xy[0::2].copy(), xy[1::2].copy()))
    return rings


def decode_record(b: bytes):
    """('point', (x, y)) or ('polygon', [parts], parts = [rings], ring = (x, y))."""
    c = _Cur(bytes(b))
    t = c.header()
    if t == 1:
        out = ("point", c.take("dd"))
    elif t == 3:
        rings = _polygon(c)
        out = ("polygon", [rings] if rings else [])
    elif t == 6:
        (k,) = c.take("I")
        parts = []
        for _ in range(k):
            if c.header() != 3:
                raise WkbError(MEMBER)
            parts.append(_polygon(c))
        out = ("polygon", parts)
    else:
        raise WkbError(TYPE)
    if c.pos != len(c.b):
        raise WkbError(TRAILING)
    return out


def decode_column(payload, offsets, validity=None):
    """Column decode: dict of the SoA buffers the GPU fill writes, or
    ('error', first bad row, code)."""
    payload = bytes(np.asarray(payload, dtype=np.uint8))
    px, py, prow = [], [], []
    x, y, ring_off, part_off, geom_off, grow = [], [], [0], [0], [0], []
    for r in range(len(offsets) - 1):
        b0, b1 = int(offsets[r]), int(offsets[r + 1])
        if (validity is not None and not validity[r]) or b1 <= b0:
            continue
        try:
            kind, body = decode_record(payload[b0:b1])
        except WkbError as e:
            return ("error", r, e.code)
        if kind == "point":
            px.append(body[0])
            py.append(body[1])
            prow.append(r)
            continue
        for rings in body:
            for rx, ry in rings:
                x.extend(rx.tolist())
                y.extend(ry.tolist())
                ring_off.append(len(x))
            part_off.append(len(ring_off) - 1)
        geom_off.append(len(part_off) - 1)
        grow.append(r)
    return {"point_x": np.array(px, np.float64), "point_y": np.array(py, np.float64),
            "point_rows": np.array(prow, np.int64), "x": np.array(x, np.float64), "y": np.array(y, np.float64),
            "ring_offsets": np.array(ring_off, np.int32), "part_offsets": np.array(part_off, np.int32),
            "geom_offsets": np.array(geom_off, np.int32), "polygon_rows": np.array(grow, np.int64)}
