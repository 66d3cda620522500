"""SparseGrid -- drop-in for the reference's sparse grid (include/grid/sparse.cuh,
src/grid/sparse.cu:9-246, bound at src/isoext_ext.cu:170-303).

A sparse grid holds a sorted, unique set of linear cell ids over
(X-1, Y-1, Z-1) cells and an (N, 8) block of per-cell corner values (corners
in the reference's Morton order).  Cell ids are kept as int64, so grids with
more than 2^31 points (e.g. 2048^3) are accepted; int32 id tensors are
accepted everywhere and returned while the ids fit in int32.

Same semantics as the reference, including its quirks:
* get_potential_cell_indices enumerates X*Y*Z ids (the point count), not the
  cell count (src/grid/sparse.cu:128-142);
* add_cells / remove_cells reset every value to default_value
  (src/grid/sparse.cu:94-95,119-120).
Both are available without the quirk: get_potential_cell_indices(...,
cells_only=True) enumerates cells, add_cells / remove_cells(...,
keep_values=True) keep the values of surviving cells.  The id sets are built
on the GPU by the library's sorted-key union / difference kernels.
Empty results are returned as None, as the reference's empty ndarray.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, workspace
from .grid import FLT_MAX, Grid, _shape3, _vec3, check_cuda_f32

vp, i64, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
_lib.declare("isoext_sparse_corner_points", ctypes.c_int, [vp, i64, _lib.I64x3, _lib.F32x3, _lib.F32x3, vp, vp])
_lib.declare("isoext_sparse_crossing_flags", ctypes.c_int, [vp, i64, f32, vp, vp])
_lib.declare("isoext_mc_sparse_workspace_bytes", i64, [i64, i64])
_lib.declare("isoext_sparse_setop_scratch_bytes", i64, [i64, i64])
_lib.declare("isoext_sparse_union", ctypes.c_int, [vp, i64, vp, i64, vp, i64, vp, vp, vp, vp])
_lib.declare("isoext_sparse_difference", ctypes.c_int, [vp, i64, vp, i64, vp, i64, vp, vp, vp, vp])
_lib.declare("isoext_sparse_gather_values", ctypes.c_int, [vp, vp, i64, f32, vp, vp])

INT32_MAX = 2**31 - 1


def _ids_in(t, name="cell_indices") -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype not in (torch.int32, torch.int64):
        raise TypeError(f"{name} must be int32 or int64, got {t.dtype}")
    if t.device.type != "cuda":
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.dim() != 1:
        raise TypeError(f"{name} must be 1-dimensional")
    return t.long()


def _ids_out(t: torch.Tensor) -> torch.Tensor:
    if t.numel() == 0 or int(t.max()) <= INT32_MAX:
        return t.int()
    return t


class SparseGrid(Grid):
    """A sparse grid of cells, each with its own 8 corner values."""

    def __init__(self, shape, aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0), default_value=FLT_MAX,
                 device=None):
        self.shape = _shape3(shape)
        self.aabb_min = _vec3(aabb_min, "aabb_min")
        self.aabb_max = _vec3(aabb_max, "aabb_max")
        self.default_value = float(np.float32(default_value))
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.cell_ids = torch.empty(0, dtype=torch.int64, device=self.device)
        self.values = torch.empty((0, 8), dtype=torch.float32, device=self.device)

    # -- counts -----------------------------------------------------------
    def get_num_cells(self) -> int:
        return int(self.cell_ids.numel())

    def get_num_points(self) -> int:
        return 8 * self.get_num_cells()

    def get_shape(self):
        return tuple(self.shape)

    # -- geometry -----------------------------------------------------------
    def _corner_points(self, ids: torch.Tensor) -> torch.Tensor:
        out = torch.empty((ids.numel(), 8, 3), dtype=torch.float32, device=self.device)
        if ids.numel():
            with torch.cuda.device(self.device):
                _lib.check(_lib.lib().isoext_sparse_corner_points(
                    ids.contiguous().data_ptr(), ids.numel(), _lib.i64x3(self.shape), _lib.f32x3(self.aabb_min),
                    _lib.f32x3(self.aabb_max), out.data_ptr(), _lib.stream_ptr(self.device)), "corner points")
        return out

    def get_points(self):
        """(N, 8, 3) corner positions of the active cells."""
        if self.get_num_cells() == 0:
            return None
        return self._corner_points(self.cell_ids)

    def get_points_by_cell_indices(self, cell_indices: torch.Tensor):
        """(N, 8, 3) corner positions of the given cells."""
        ids = _ids_in(cell_indices)
        if ids.numel() == 0:
            return None
        return self._corner_points(ids)

    # -- values ---------------------------------------------------------------
    def get_values(self):
        """A copy of the (N, 8) corner values."""
        if self.get_num_cells() == 0:
            return None
        return self.values.clone()

    def set_values(self, new_values: torch.Tensor) -> None:
        check_cuda_f32(new_values, ndim=2, last=8, name="new_values")
        if new_values.numel() != self.get_num_points():
            raise RuntimeError("New values size does not match number of points")
        self.values.copy_(new_values)

    def get_cells(self):
        """(N, 8) point indices: the identity map into the (N, 8) values."""
        n = self.get_num_cells()
        if n == 0:
            return None
        return torch.arange(8 * n, device=self.device, dtype=torch.int64).view(n, 8).int()

    def get_cell_indices(self):
        if self.get_num_cells() == 0:
            return None
        return _ids_out(self.cell_ids.clone())

    # -- population -------------------------------------------------------------
    def _reset_values(self):
        self.values = torch.full((self.cell_ids.numel(), 8), self.default_value, dtype=torch.float32,
                                 device=self.device)

    def _setop(self, ids: torch.Tensor, difference: bool, keep_values: bool):
        """Sorted-key union (add) or difference (remove) on the GPU; values
        reset to default_value as in the reference, or carried over for the
        surviving cells with keep_values."""
        L = _lib.lib()
        a = self.cell_ids.contiguous()
        b = torch.sort(ids).values.contiguous()
        na, nb = a.numel(), b.numel()
        cap = na if difference else na + nb
        out = torch.empty(max(cap, 1), dtype=torch.int64, device=self.device)
        slot = torch.empty(max(cap, 1), dtype=torch.int64, device=self.device) if keep_values else None
        n_out = torch.zeros(1, dtype=torch.int64, device=self.device)
        scratch = workspace.get("sparse_setop", L.isoext_sparse_setop_scratch_bytes(na, nb), self.device)
        fn = L.isoext_sparse_difference if difference else L.isoext_sparse_union
        with _lib.on_device(self.device):
            stream = _lib.stream_ptr(self.device)
            _lib.check(fn(a.data_ptr() if na else None, na, b.data_ptr() if nb else None, nb, scratch.data_ptr(),
                          scratch.numel(), out.data_ptr(), slot.data_ptr() if slot is not None else None,
                          n_out.data_ptr(), stream), "remove_cells" if difference else "add_cells")
            n = int(n_out.item())
            if keep_values:
                vals = torch.empty((n, 8), dtype=torch.float32, device=self.device)
                _lib.check(L.isoext_sparse_gather_values(self.values.data_ptr() if na else None,
                                                         slot.data_ptr(), n, float(self.default_value),
                                                         vals.data_ptr(), stream), "cell values")
        self.cell_ids = out[:n].clone() if cap > 2 * n + 65536 else out[:n]
        if keep_values:
            self.values = vals
        else:
            self._reset_values()

    def add_cells(self, new_cell_indices: torch.Tensor, keep_values: bool = False) -> None:
        """Union with the given cells (any order, duplicates allowed).  Resets
        all values to default_value like the reference
        (src/grid/sparse.cu:71-96); keep_values=True keeps the values of the
        cells already present (new cells get default_value)."""
        ids = _ids_in(new_cell_indices, "new_cell_indices")
        self._setop(ids, False, keep_values)

    def remove_cells(self, new_cell_indices: torch.Tensor, keep_values: bool = False) -> None:
        """Set difference with the given cells (src/grid/sparse.cu:98-121);
        values reset as in the reference unless keep_values=True."""
        ids = _ids_in(new_cell_indices, "new_cell_indices")
        self._setop(ids, True, keep_values)

    def get_potential_cell_indices(self, chunk_size: int, cells_only: bool = False):
        """Chunks of consecutive ids covering [0, X*Y*Z) -- the reference's
        range, which counts points (src/grid/sparse.cu:128-142); with
        cells_only=True the cell range [0, (X-1)(Y-1)(Z-1)) instead."""
        chunk_size = int(chunk_size)
        if chunk_size <= 0:
            raise ValueError("chunk_size must be positive")
        X, Y, Z = self.shape
        total = (X - 1) * (Y - 1) * (Z - 1) if cells_only else X * Y * Z
        dt = torch.int32 if total - 1 <= INT32_MAX else torch.int64
        return [torch.arange(s, min(s + chunk_size, total), device=self.device, dtype=dt)
                for s in range(0, total, chunk_size)]

    def filter_cell_indices(self, cell_indices: torch.Tensor, values: torch.Tensor, level: float = 0.0):
        """The given cells whose 8 values straddle `level`."""
        ids = _ids_in(cell_indices)
        check_cuda_f32(values, ndim=2, last=8, name="values")
        if values.shape[0] != ids.numel():
            raise RuntimeError("values must have one row per cell index")
        flags = torch.empty(ids.numel(), dtype=torch.uint8, device=self.device)
        if ids.numel():
            with torch.cuda.device(self.device):
                _lib.check(_lib.lib().isoext_sparse_crossing_flags(values.data_ptr(), ids.numel(), float(level),
                                                                   flags.data_ptr(), _lib.stream_ptr(self.device)),
                           "filter_cell_indices")
        out = ids[flags.bool()]
        return out.int() if cell_indices.dtype == torch.int32 else out

    def __repr__(self):
        return f"SparseGrid(shape={self.shape}, cells={self.get_num_cells()})"


def grid_struct_of(grid):
    from .dc import grid_struct
    return grid_struct(grid)


def mc_sparse(grid: SparseGrid, level: float, method: int, want_src: bool = False):
    """Marching cubes over a sparse grid; returns (v, f[, vsrc])."""
    L = _lib.lib()
    dev = grid.device
    n = grid.get_num_cells()
    empty = (None, None, None) if want_src else (None, None)
    if n == 0:
        return empty
    g = grid_struct_of(grid)
    scratch = workspace.get("sparse_count", 8 * (n // 256 + 32), dev)
    ninst = ctypes.c_int64(0)
    with torch.cuda.device(dev):
        stream = _lib.stream_ptr(dev)
        _lib.check(L.isoext_mc_sparse_count(ctypes.byref(g), float(level), scratch.data_ptr(), ctypes.byref(ninst),
                                            stream), "marching_cubes")
        ni = int(ninst.value)
        if ni == 0:
            return empty
        ws = workspace.get("sparse_mc", L.isoext_mc_sparse_workspace_bytes(n, ni), dev)
        cap_v, cap_f = ni, 5 * n
        counts = (ctypes.c_int64 * 2)()
        for _ in range(2):
            v = torch.empty((max(cap_v, 1), 3), dtype=torch.float32, device=dev)
            f = torch.empty((max(cap_f, 1), 3), dtype=torch.int32, device=dev)
            src = torch.empty(max(cap_v, 1), dtype=torch.int64, device=dev) if want_src else None
            rc = L.isoext_mc_sparse(ctypes.byref(g), float(level), int(method), ni, ws.data_ptr(), ws.numel(),
                                    v.data_ptr(), cap_v, f.data_ptr(), cap_f,
                                    src.data_ptr() if src is not None else None, counts, stream)
            if rc == _lib.E_CAPACITY:
                cap_v, cap_f = int(counts[0]), int(counts[1])
                continue
            _lib.check(rc, "marching_cubes")
            break
    nv, nf = int(counts[0]), int(counts[1])
    if nf == 0:
        return empty
    v, f = v[:nv], f[:nf]
    if want_src:
        return v, f, src[:nv]
    return v, f
