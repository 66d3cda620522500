"""Differentiable marching cubes (no counterpart in the reference).

``marching_cubes(grid, values, level, method)`` returns (v, f) where v is
differentiable with respect to `values`: the grid's (X, Y, Z) values for a
UniformGrid or its (N, 8) per-cell corner values for a SparseGrid.  The
backward kernel (libisoext_amd, isoext_mc_backward) pushes dL/dv through the
edge interpolation t = (L - va) / (vb - va) of each vertex's representative
edge.  Faces are not differentiable.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .dc import IsoGrid, grid_struct
from .grid import UniformGrid
from .mc import _method_id, mc_dense_raw
from .sparse import SparseGrid, mc_sparse

_lib.declare("isoext_mc_backward", ctypes.c_int, [ctypes.POINTER(IsoGrid), ctypes.c_float, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                                  ctypes.c_void_p])


class _View:
    """A grid with its values replaced (shares geometry and sparse ids)."""

    def __init__(self, grid, values):
        self.shape = grid.shape
        self.aabb_min = grid.aabb_min
        self.aabb_max = grid.aabb_max
        self.values = values
        self.device = values.device
        if isinstance(grid, SparseGrid):
            self.cell_ids = grid.cell_ids

    def get_num_cells(self):
        return int(self.cell_ids.numel())


class MarchingCubesFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, grid, level, method):
        vals = values.detach().contiguous()
        view = _View(grid, vals)
        if isinstance(grid, UniformGrid):
            v, f, info = mc_dense_raw(vals, grid.shape, grid.aabb_min, grid.aabb_max, level, method,
                                      want_keys=True)
            src = info["keys"]
        else:
            v, f, src = mc_sparse(view, level, method, want_src=True)
            if v is None:
                v = torch.empty((0, 3), dtype=torch.float32, device=vals.device)
                f = torch.empty((0, 3), dtype=torch.int32, device=vals.device)
                src = torch.empty(0, dtype=torch.int64, device=vals.device)
        ctx.view = view
        ctx.level = float(level)
        ctx.save_for_backward(vals, src)
        ctx.mark_non_differentiable(f)
        return v, f

    @staticmethod
    def backward(ctx, grad_v, grad_f):
        vals, src = ctx.saved_tensors
        grad = torch.zeros_like(vals)
        if grad_v is not None and src.numel():
            gv = grad_v.contiguous().float()
            g = grid_struct(ctx.view)
            dev = vals.device
            with torch.cuda.device(dev):
                _lib.check(_lib.lib().isoext_mc_backward(ctypes.byref(g), ctx.level, gv.data_ptr(), src.data_ptr(),
                                                         src.numel(), grad.data_ptr(), _lib.stream_ptr(dev)),
                           "marching_cubes backward")
        return grad, None, None, None


def marching_cubes(grid, values: torch.Tensor | None = None, level: float = 0.0, method: str = "nagae"):
    """Differentiable marching cubes.  `values` defaults to the grid's values
    (then no gradient flows to the caller)."""
    m = _method_id(method)
    if values is None:
        values = grid.values
    expect = tuple(grid.values.shape)
    if tuple(values.shape) != expect:
        raise ValueError(f"values must have shape {expect}")
    if values.dtype != torch.float32 or values.device.type != "cuda":
        raise TypeError("values must be a float32 CUDA tensor")
    return MarchingCubesFunction.apply(values, grid, float(level), m)
