"""Grid types of the drop-in API.

Mirrors the reference's Grid / UniformGrid classes (include/grid/grid.cuh:10-35,
include/grid/uniform.cuh, src/grid/uniform.cu:8-83, bound at
src/isoext_ext.cu:111-168).  Values live in one torch tensor on the GPU;
positions, the default fill and extraction run in libisoext_amd.

Differences from the reference, on purpose:
* grids with more than 2^31-1 points are accepted (64-bit indexing; the
  reference rejects them, src/grid/uniform.cu:13-17);
* ``get_cells`` works on a UniformGrid (the reference binds it through the
  base class with no converter for its return type).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib

FLT_MAX = float(np.finfo(np.float32).max)


def _shape3(shape, name="shape"):
    try:
        vals = [int(s) for s in shape]
    except TypeError as e:
        raise TypeError(f"{name} must be a sequence of 3 non-negative integers") from e
    if len(vals) != 3 or any(v < 0 for v in vals):
        raise TypeError(f"{name} must be a sequence of 3 non-negative integers")
    return vals


def _vec3(v, name):
    try:
        vals = [float(x) for x in v]
    except TypeError as e:
        raise TypeError(f"{name} must be a sequence of 3 floats") from e
    if len(vals) != 3:
        raise TypeError(f"{name} must be a sequence of 3 floats")
    # round through float32 exactly as the reference's std::array<float, 3>
    out = [float(np.float32(x)) for x in vals]
    if not all(np.isfinite(out)):
        raise ValueError(f"{name} must be finite")
    return out


def check_cuda_f32(t, ndim=None, last=None, name="tensor"):
    """The reference's nanobind caster: torch, float32, CUDA, C-contiguous
    (src/isoext_ext.cu:21-29).  A mismatch is a TypeError."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    if t.device.type != "cuda":
        raise TypeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise TypeError(f"{name} must be C-contiguous")
    if ndim is not None and t.dim() != ndim:
        raise TypeError(f"{name} must have {ndim} dimensions, got {t.dim()}")
    if last is not None and (t.dim() == 0 or t.shape[-1] != last):
        raise TypeError(f"{name} must have shape (..., {last})")


class Grid:
    """Abstract base of UniformGrid and SparseGrid (include/grid/grid.cuh:10-35)."""

    def get_num_cells(self) -> int:
        raise NotImplementedError

    def get_num_points(self) -> int:
        raise NotImplementedError

    def get_points(self) -> torch.Tensor:
        raise NotImplementedError

    def get_values(self) -> torch.Tensor:
        raise NotImplementedError

    def set_values(self, new_values: torch.Tensor) -> None:
        raise NotImplementedError

    def get_cells(self) -> torch.Tensor:
        raise NotImplementedError


class UniformGrid(Grid):
    """A dense uniform grid of scalar values.

    ``shape`` counts grid POINTS per axis (cells = shape - 1), as in the
    reference (src/grid/uniform.cu:11-12).  Values are stored (X, Y, Z) in C
    order, z fastest.
    """

    def __init__(self, shape, aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0),
                 default_value=FLT_MAX, device=None):
        self.shape = _shape3(shape)
        self.aabb_min = _vec3(aabb_min, "aabb_min")
        self.aabb_max = _vec3(aabb_max, "aabb_max")
        self.default_value = float(np.float32(default_value))
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        X, Y, Z = self.shape
        self.values = torch.empty((X, Y, Z), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().isoext_fill_f32(self.values.data_ptr(), self.values.numel(),
                                                  self.default_value, _lib.stream_ptr()),
                       "UniformGrid")

    @classmethod
    def wrap(cls, values: torch.Tensor, aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0)):
        """A grid over an existing (X, Y, Z) float32 CUDA tensor, without a
        copy (the grid borrows the tensor)."""
        check_cuda_f32(values, ndim=3, name="values")
        g = cls.__new__(cls)
        g.shape = [int(s) for s in values.shape]
        g.aabb_min = _vec3(aabb_min, "aabb_min")
        g.aabb_max = _vec3(aabb_max, "aabb_max")
        g.default_value = float(np.float32(FLT_MAX))
        g.device = values.device
        g.values = values if values.is_contiguous() else values.contiguous()
        return g

    # -- counts -----------------------------------------------------------
    def get_num_cells(self) -> int:
        X, Y, Z = self.shape
        if X == 0 or Y == 0 or Z == 0:
            return 0
        return (X - 1) * (Y - 1) * (Z - 1)

    def get_num_points(self) -> int:
        X, Y, Z = self.shape
        return X * Y * Z

    def get_shape(self):
        return tuple(self.shape)

    # -- data ---------------------------------------------------------------
    def get_points(self) -> torch.Tensor:
        """(X, Y, Z, 3) float32 positions (src/grid/uniform.cu:22-30)."""
        X, Y, Z = self.shape
        out = torch.empty((X, Y, Z, 3), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().isoext_uniform_points(_lib.i64x3(self.shape),
                                                        _lib.f32x3(self.aabb_min),
                                                        _lib.f32x3(self.aabb_max),
                                                        out.data_ptr(), _lib.stream_ptr()),
                       "get_points")
        return out

    def get_values(self) -> torch.Tensor:
        """A copy of the (X, Y, Z) values (src/grid/uniform.cu:32-35)."""
        return self.values.clone()

    def set_values(self, new_values: torch.Tensor) -> None:
        """Copy (X, Y, Z) float32 values into the grid (src/grid/uniform.cu:37-40)."""
        check_cuda_f32(new_values, ndim=3, name="new_values")
        if list(new_values.shape) != self.shape:
            raise RuntimeError("Cannot set values with different shapes")
        self.values.copy_(new_values)

    def get_cells(self) -> torch.Tensor:
        """(X-1, Y-1, Z-1, 8) int64 corner point ids, Morton corner order
        (idx_to_cell_op, include/utils.cuh:32-60)."""
        X, Y, Z = self.shape
        dev = self.device
        x = torch.arange(max(X - 1, 0), device=dev).view(-1, 1, 1, 1)
        y = torch.arange(max(Y - 1, 0), device=dev).view(1, -1, 1, 1)
        z = torch.arange(max(Z - 1, 0), device=dev).view(1, 1, -1, 1)
        k = torch.arange(8, device=dev).view(1, 1, 1, 8)
        return ((x + (k >> 2)) * Y + (y + ((k >> 1) & 1))) * Z + (z + (k & 1))

    def get_cell_indices(self) -> torch.Tensor:
        return torch.arange(self.get_num_cells(), device=self.device, dtype=torch.int64)

    def __repr__(self):
        return f"UniformGrid(shape={self.shape}, aabb_min={self.aabb_min}, aabb_max={self.aabb_max})"
