"""The batched join's autotune cache on disk (no GPU): entries written by one
cache object are read by a fresh one (a later process), keyed by device and
shape class; unreadable files and invalid tiles are ignored."""

from __future__ import annotations

import json
import os

from vibespatial_amd import join as J


def test_tile_cache_persists_across_instances(tmp_path, monkeypatch):
    monkeypatch.setenv("VIBESPATIAL_AMD_CACHE_DIR", str(tmp_path))
    a = J.TileCache()
    assert a.get("gfx950", (13, 24, 21)) is None
    a.put("gfx950", (13, 24, 21), (2, 128, 1))
    a.put("other", (13, 24, 21), (4, 256, 0))
    with open(os.path.join(tmp_path, J.TileCache.FILE)) as f:
        data = json.load(f)
    assert data == {"gfx950|13|24|21": [2, 128, 1], "other|13|24|21": [4, 256, 0]}
    b = J.TileCache()
    assert b.get("gfx950", (13, 24, 21)) == (2, 128, 1)
    assert b.get("other", (13, 24, 21)) == (4, 256, 0)
    assert b.get("gfx950", (12, 24, 21), (4, 256)) == (4, 256)


def test_tile_cache_ignores_bad_files_and_tiles(tmp_path, monkeypatch):
    monkeypatch.setenv("VIBESPATIAL_AMD_CACHE_DIR", str(tmp_path))
    p = os.path.join(tmp_path, J.TileCache.FILE)
    with open(p, "w") as f:
        f.write("{not json")
    assert J.TileCache().get("gfx950", (1, 2, 3)) is None
    with open(p, "w") as f:
        json.dump({"gfx950|1|2|3": [3, 100, 0], "gfx950|1|2|4": [2, 128, 1], "gfx950|1|2|5": [2, 128], "bad": [2, 128, 0]}, f)
    c = J.TileCache()
    assert c.get("gfx950", (1, 2, 3)) is None  # (3, 100, 0) is no candidate tile
    assert c.get("gfx950", (1, 2, 5)) is None  # an entry without the candidate order
    assert c.get("gfx950", (1, 2, 4)) == (2, 128, 1)
    # an unwritable directory keeps the entry in memory
    monkeypatch.setenv("VIBESPATIAL_AMD_CACHE_DIR", os.path.join(p, "sub"))
    d = J.TileCache()
    d.put("gfx950", (5, 5, 5), (4, 128, 1))
    assert d.get("gfx950", (5, 5, 5)) == (4, 128, 1)
