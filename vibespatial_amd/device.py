"""Device residency: torch ROCm tensors as raw buffers for libvsjoin.

PyTorch is used only for device memory and the current stream; every kernel
comes from libvsjoin.so.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .geometry import PointArray, PolygonArray

__all__ = [
    "require_gpu",
    "ptr",
    "stream",
    "to_device",
    "DevicePoints",
    "DevicePolygons",
    "device_points",
    "device_polygons",
]


def require_gpu() -> torch.device:
    """The reference raises when GPU execution is requested but unavailable
    (src/vibespatial/runtime/_runtime.py:55-56); there is no CPU path here."""
    if not torch.cuda.is_available():
        raise RuntimeError("vibespatial_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def to_device(a: np.ndarray, dtype=None) -> torch.Tensor:
    dev = require_gpu()
    a = np.ascontiguousarray(a if dtype is None else a.astype(dtype, copy=False))
    if a.size == 0:
        return torch.empty(a.shape, dtype=torch.from_numpy(a[:0].copy()).dtype, device=dev)
    return torch.from_numpy(a).to(dev, non_blocking=False)


class DevicePoints:
    """x, y fp64 device tensors (NaN = null/empty)."""

    def __init__(self, x: torch.Tensor, y: torch.Tensor):
        assert x.dtype == torch.float64 and y.dtype == torch.float64
        self.x = x.contiguous()
        self.y = y.contiguous()

    def __len__(self) -> int:
        return int(self.x.numel())

    @classmethod
    def from_host(cls, pa: PointArray) -> "DevicePoints":
        return cls(to_device(pa.x), to_device(pa.y))


class DevicePolygons:
    """x, y fp64 and int32 ring/part/geom offsets on the device."""

    def __init__(self, x, y, ring_offsets, part_offsets, geom_offsets):
        self.x = x.contiguous()
        self.y = y.contiguous()
        self.ring_offsets = ring_offsets.contiguous()
        self.part_offsets = part_offsets.contiguous()
        self.geom_offsets = geom_offsets.contiguous()

    def __len__(self) -> int:
        return int(self.geom_offsets.numel()) - 1

    @property
    def n_parts(self) -> int:
        return int(self.part_offsets.numel()) - 1

    @property
    def n_rings(self) -> int:
        return int(self.ring_offsets.numel()) - 1

    @property
    def n_coords(self) -> int:
        return int(self.x.numel())

    @classmethod
    def from_host(cls, pa: PolygonArray) -> "DevicePolygons":
        return cls(
            to_device(pa.x),
            to_device(pa.y),
            to_device(pa.ring_offsets),
            to_device(pa.part_offsets),
            to_device(pa.geom_offsets),
        )

    def args(self):
        return [ptr(self.x), ptr(self.y), ptr(self.geom_offsets), ptr(self.part_offsets), ptr(self.ring_offsets)]


def device_points(pa: PointArray) -> DevicePoints:
    """Device copy cached on the host array (H2D once, like the reference's residency)."""
    if getattr(pa, "_device", None) is None:
        pa._device = DevicePoints.from_host(pa)
    return pa._device


def device_polygons(pa: PolygonArray) -> DevicePolygons:
    if getattr(pa, "_device", None) is None:
        pa._device = DevicePolygons.from_host(pa)
    return pa._device
