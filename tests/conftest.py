"""Test configuration.

Markers: ``gpu`` -- needs a ROCm GPU (runs on the MI355X box with
``pytest -m gpu``); everything else runs on the CPU.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X)")


def pytest_collection_modifyitems(config, items):
    try:
        import torch

        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no ROCm GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)


def decode_geom(g):
    """Fixture geometry -> vibespatial_amd scalar geometry (None for null)."""
    from vibespatial_amd.geometry import MultiPolygon, Point, Polygon

    if g is None:
        return None
    if g["type"] == "Point":
        if g["xy"] is None:
            return Point()
        return Point(float.fromhex(g["xy"][0]), float.fromhex(g["xy"][1]))
    parts = g["parts"]
    if g["type"] == "Polygon":
        if not parts:
            return Polygon()
        rings = parts[0]
        return Polygon(rings[0], rings[1:])
    return MultiPolygon([Polygon(r[0], r[1:]) for r in parts])


def geom_parts(g):
    """Scalar geometry -> list of parts (list of rings of (x, y) tuples)."""
    return [[list(map(tuple, np.asarray(r))) for r in rings] for rings in g.parts]


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle
