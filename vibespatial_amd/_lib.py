"""ctypes binding of libvsjoin.so (declared in include/vsjoin.h).

torch is imported first so that its bundled HIP runtime (libamdhip64.so.7) is
the one the library binds to: device pointers and stream handles from torch
are then valid inside the library.  There is no fallback: if the library is
missing the import raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime the library links against)

__all__ = ["lib", "check", "VsjoinError", "TreeLayout", "LIB_PATH", "PRED"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libvsjoin.so")

MAX_LEVELS = 16
NOALLOC_SCRATCH_BYTES = 64 << 20  # VS_NOALLOC_SCRATCH_BYTES

PRED = {
    "within": 0,
    "intersects": 1,
    "covered_by": 2,
    "touches": 3,
    "contains": 4,
    "covers": 5,
    "overlaps": 6,
    "crosses": 7,
    "contains_properly": 8,
}


class VsjoinError(RuntimeError):
    """A libvsjoin entry point returned an error code."""


class TreeLayout(ctypes.Structure):
    _fields_ = [
        ("fanout", ctypes.c_int32),
        ("n_levels", ctypes.c_int32),
        ("n_items", ctypes.c_int64),
        ("n_nodes", ctypes.c_int64),
        ("level_start", ctypes.c_int64 * MAX_LEVELS),
        ("level_count", ctypes.c_int64 * MAX_LEVELS),
    ]


_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_d = ctypes.c_double
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_LP = ctypes.POINTER(TreeLayout)

# name -> (restype, argtypes); every symbol include/vsjoin.h declares.
SIGNATURES = {
    "vs_version": (ctypes.c_int, []),
    "vs_last_error": (ctypes.c_int, [ctypes.c_char_p, _i64]),
    "vs_polygon_bounds": (ctypes.c_int, [_p, _p, _p, _p, _p, _i64, _p, _p]),
    "vs_point_boxes": (ctypes.c_int, [_p, _p, _i64, _p, _p]),
    "vs_tree_plan": (ctypes.c_int, [_i64, _i32, _LP]),
    "vs_index_workspace_bytes": (_i64, [_i64]),
    "vs_index_build": (ctypes.c_int, [_p, _i64, _LP, _p, _p, _p, _p, _i64, _p]),
    "vs_sjoin_points_polygons_workspace_bytes": (_i64, [_i64, _i64, _i64, _i64]),
    "vs_sjoin_points_polygons": (
        ctypes.c_int,
        [_p, _p, _i64, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _LP, _p, _p, _i32, _i64, _p, _p, _p, _p, _i64,
         _p],
    ),
    "vs_sjoin_points_polygons_noalloc": (
        ctypes.c_int,
        [_p, _p, _i64, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _LP, _p, _p, _i32, _i64, _p, _p, _p, _p],
    ),
    "vs_query_boxes_workspace_bytes": (_i64, [_i64, _i64]),
    "vs_query_boxes": (ctypes.c_int, [_p, _i64, _LP, _p, _p, _i64, _i32, _p, _p, _p, _p, _i64, _p]),
    "vs_refine_pairs_workspace_bytes": (_i64, [_i64]),
    "vs_refine_pairs": (
        ctypes.c_int,
        [_i32, _i32, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _p],
    ),
    "vs_expand_boxes": (ctypes.c_int, [_p, _i64, _p, _p, _p]),
    "vs_refine_dwithin_workspace_bytes": (_i64, [_i64]),
    "vs_refine_dwithin": (
        ctypes.c_int,
        [_i32, _i32, _p, _p, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _p],
    ),
    "vs_lexsort_pairs_workspace_bytes": (_i64, [_i64]),
    "vs_lexsort_pairs": (ctypes.c_int, [_p, _p, _i64, _p, _i64, _p]),
    "vs_learned_leaf_count": (_i32, [_i64, _i32]),
    "vs_learned_index_bytes": (_i64, [_i64, _i32, _i32]),
    "vs_learned_index_layout": (ctypes.c_int, [_i64, _i32, _i32, _p]),
    "vs_learned_index_build": (ctypes.c_int, [_p, _p, _i64, _i32, _i32, _i32, ctypes.c_float, _p, _i64, _p]),
    "vs_learned_lower_bound": (ctypes.c_int, [_p, _i64, _i32, _i32, _p, _i64, _i32, _p, _p, _p]),
    "vs_mlp_loss_grad": (ctypes.c_int, [_p, _p, _p, _i64, _p, _p, _p]),
    "vs_learned_query_boxes_workspace_bytes": (_i64, [_i64, _i64]),
    "vs_learned_query_boxes": (ctypes.c_int,
                               [_p, _i64, _i32, _i32, _p, _i64, _i32, _i64, _i32, _p, _p, _p, _p, _i64, _p]),
    "vs_learned_query_polygons_workspace_bytes": (_i64, [_i64, _i64, _i64]),
    "vs_learned_query_polygons": (
        ctypes.c_int,
        [_p, _i64, _i32, _i32, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _p, _i32, _i32, _i64, _i32, _p, _p, _p,
         _p, _i64, _p],
    ),
    "vs_sjoin_points_polygons_learned_workspace_bytes": (_i64, [_i64, _i64, _i64, _i64]),
    "vs_sjoin_points_polygons_learned": (
        ctypes.c_int,
        [_p, _p, _i64, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _p, _i32, _i32, _i32, _i64, _p, _p, _p, _p, _i64,
         _p],
    ),
    "vs_batched_intersects_workspace_bytes": (_i64, [_i64, _i64, _i64, _i64, _i64]),
    "vs_batched_intersects": (ctypes.c_int, [_p, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _i64, _i64, _p, _p, _i64,
                                             _i32, _i64, _p, _p, _p, _p, _i64, _p]),
    "vs_wkb_workspace_bytes": (_i64, [_i64]),
    "vs_wkb_scan": (ctypes.c_int, [_p, _i64, _p, _p, _i64, _p, _p, _i64, _p]),
    "vs_wkb_fill": (ctypes.c_int, [_p, _i64, _p, _p, _i64, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "vs_pair_distances_workspace_bytes": (_i64, [_i64]),
    "vs_pair_distances": (_i32, [_i32, _i32, _p, _p, _i64, _i64, _i32, _p] + [_p] * 10 + [_p, _p, _i64, _p]),
    "vs_nearest_select_workspace_bytes": (_i64, [_i64, _i64]),
    "vs_nearest_select": (_i32, [_p, _p, _p, _i64, _i64, _p, _p, _d, _d, _d, _d, _i32, _p, _p, _p, _p, _p, _i64,
                                 _p]),
    "vs_nearest_init": (_i32, [_p, _i64, _d, _p, _p, _p]),
    "vs_nearest_open_workspace_bytes": (_i64, [_i64]),
    "vs_nearest_open": (_i32, [_p, _p, _i64, _p, _p, _p, _p, _i64, _p]),
    "vs_map_left": (_i32, [_p, _p, _i64, _p]),
    "vs_nearest_finish_workspace_bytes": (_i64, [_i64]),
    "vs_nearest_finish": (_i32, [_p, _p, _p, _i64, _i64, _p, _p, _p, _p, _i64, _p]),
    "vs_nearest_bound_workspace_bytes": (_i64, [_i64]),
    "vs_nearest_bound": (_i32, [_p, _p, _i64, _p, _LP, _p, _p, _i32, _p] + [_p] * 5 + [_d, _p, _p, _p, _i64, _p]),
    "vs_nearest_rep_points": (_i32, [_i32, _p, _i64, _p, _p, _p, _p, _p, _p, _p]),
    "vs_scan_workspace_bytes": (_i64, [_i64]),
    "vs_exclusive_scan_i32": (ctypes.c_int, [_p, _p, _i64, _p, _p, _i64, _p]),
    "vs_sort_workspace_bytes": (_i64, [_i64, _i32]),
    "vs_sort_pairs_u64": (ctypes.c_int, [_p, _p, _i64, _i32, _i32, _p, _i64, _p]),
    "vs_sort_pairs_u32": (ctypes.c_int, [_p, _p, _i64, _i32, _i32, _p, _i64, _p]),
    "vs_point_order_workspace_bytes": (_i64, [_i64]),
    "vs_point_order": (ctypes.c_int, [_p, _p, _i64, _i32, _p, _p, _p, _p, _i64, _p]),
    "vs_set_kernel_timing": (ctypes.c_int, [_i32]),
    "vs_last_kernel_ms": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float)]),
    "vs_host_orient2d": (ctypes.c_int, [_d] * 6),
    "vs_host_point_location": (ctypes.c_int, [_d, _d, _p, _p, _p, _p, _p, _i64]),
    "vs_host_segments_touch": (ctypes.c_int, [_d] * 8),
    "vs_host_point_point_within": (ctypes.c_int, [_d] * 5),
    "vs_host_point_segment_within": (ctypes.c_int, [_d] * 7),
    "vs_host_hilbert64": (_u64, [_u32, _u32]),
    "vs_host_cover_ranges": (ctypes.c_int, [_d, _d, _d, _d, _p, _i32, _p]),
    "vs_host_cell_grid": (ctypes.c_int, [_p, _p, _i64, _p, _p, _p]),
    "vs_host_cell_decide": (ctypes.c_int, [_p, _p, _p, _p, _i32]),
    "vs_host_cell_span": (ctypes.c_int, [_p, _i32, _d, _p]),
}

_lock = threading.Lock()
_lib = None


def lib():
    """Load libvsjoin.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m vibespatial_amd.build` "
                    "(there is no CPU fallback)"
                )
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().vs_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise VsjoinError(f"{what} failed ({rc}): {last_error()}")
