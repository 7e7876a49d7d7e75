"""vibespatial_amd -- MI355X-native spatial-join hot path of vibeSpatial.

``sjoin(left, right, predicate=...)`` and ``GeoSeries.sindex.query`` over a
Hilbert-packed spatial index (or a learned rank index over Hilbert-sorted
points), with exact point-in-polygon, polygon relations and distance
decisions; ``sjoin_nearest`` / ``sindex.nearest``; GeoParquet / WKB ingest
decoded on the GPU.  Computed by hand-written gfx950 HIP kernels in
``libvsjoin.so`` (C-ABI in ``include/vsjoin.h``, bound with ctypes).
"""

from .geometry import (  # noqa: F401
    GeometryArray,
    MultiPolygon,
    Point,
    PointArray,
    Polygon,
    PolygonArray,
    box,
    points_from_xy,
)

__version__ = "0.1.0"


def __getattr__(name):
    # the GPU-facing modules import torch and the library lazily
    if name in ("sjoin", "sjoin_nearest", "GeoSeries", "GeoDataFrame", "SpatialIndex", "PREDICATES"):
        from . import api

        return getattr(api, name)
    if name in ("read_parquet", "read_file", "decode_wkb", "encode_wkb"):
        from . import io

        return getattr(io, name)
    raise AttributeError(name)
