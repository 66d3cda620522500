"""Differentiable marching cubes and dual contouring (no counterpart in the
reference).

``marching_cubes(grid, values, level, method)`` returns (v, f) where v is
differentiable with respect to `values`: the grid's (X, Y, Z) values for a
UniformGrid or its (N, 8) per-cell corner values for a SparseGrid.  The
backward kernel (libisoext_amd, isoext_mc_backward) pushes dL/dv through the
edge interpolation t = (L - va) / (vb - va) of each vertex's representative
edge.

``dual_contouring(grid, values, normals, level, reg, svd_tol)`` does the same
for dual contouring: the backward kernel (isoext_dc_backward) goes through
each cell's regularised QEF solve to its intersection points and their edge
interpolation, and gives dL/dnormals when `normals` is a tensor.  With
normals computed from the values (trilinear, normals=None), their own
dependence on the values is not differentiated.  Faces are not
differentiable.
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


class DualContouringFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, normals, grid, level, reg, svd_tol):
        from .dc import compute_normals, dc_raw, get_intersection
        vals = values.detach().contiguous()
        view = _View(grid, vals)
        its = get_intersection(view, level, False)
        dev = vals.device
        if its.points.shape[0] == 0:
            ctx.empty = True
            ctx.save_for_backward(vals)
            e = torch.empty((0, 3), dtype=torch.float32, device=dev)
            return e, torch.empty((0, 3), dtype=torch.int32, device=dev)
        ctx.empty = False
        n = normals.detach().contiguous().float() if normals is not None else compute_normals(view, its)
        v, f, info = dc_raw(view, its, n, reg, svd_tol, want_vcell=True)
        ctx.view, ctx.its, ctx.normals = view, its, n
        ctx.level, ctx.reg, ctx.svd_tol = float(level), float(reg), float(svd_tol)
        ctx.has_normals = normals is not None
        ctx.save_for_backward(vals, info["vcell"])
        ctx.mark_non_differentiable(f)
        return v, f

    @staticmethod
    def backward(ctx, grad_v, grad_f):
        if ctx.empty:
            (vals,) = ctx.saved_tensors
            return torch.zeros_like(vals), None, None, None, None, None
        vals, vcell = ctx.saved_tensors
        grad = torch.zeros_like(vals)
        gn = torch.zeros_like(ctx.normals) if ctx.has_normals else None
        if grad_v is not None and vcell.numel():
            gv = grad_v.contiguous().float()
            g = grid_struct(ctx.view)
            s = ctx.its._struct(normals=ctx.normals)
            dev = vals.device
            with torch.cuda.device(dev):
                _lib.check(_lib.lib().isoext_dc_backward(ctypes.byref(g), ctypes.byref(s), ctx.reg, ctx.svd_tol,
                                                         ctx.level, vcell.numel(), vcell.data_ptr(), gv.data_ptr(),
                                                         grad.data_ptr(), gn.data_ptr() if gn is not None else None,
                                                         _lib.stream_ptr(dev)),
                           "dual_contouring backward")
        return grad, gn, None, None, None, None


def dual_contouring(grid, values: torch.Tensor | None = None, normals: torch.Tensor | None = None,
                    level: float = 0.0, reg: float = 1e-2, svd_tol: float = 1e-6):
    """Differentiable dual contouring.  `normals` (one per intersection of
    get_intersection(grid, level)) defaults to trilinear normals of the
    values; when given it receives a gradient too."""
    if values is None:
        values = grid.values
    expect = tuple(grid.values.shape)
    if tuple(values.shape) != expect:
        raise ValueError(f"values must have shape {expect}")
    if values.dtype != torch.float32 or values.device.type != "cuda":
        raise TypeError("values must be a float32 CUDA tensor")
    return DualContouringFunction.apply(values, normals, grid, float(level), float(reg), float(svd_tol))


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
