"""Intersection, get_intersection and dual_contouring -- drop-ins for the
reference's DC front end and solver.

Reference: Intersection (include/its.cuh:6-26, bound at
src/isoext_ext.cu:305-327), get_intersection (src/its.cu:93-159, bound at
src/isoext_ext.cu:329-343), dual_contouring (src/dc.cu:161-218, bound at
src/isoext_ext.cu:345-378 with its default-normals logic).

Differences on purpose:
* get_points()/get_normals() return copies; the reference hands over its
  buffer on the first call and leaves the object empty
  (src/isoext_ext.cu:67-76).
* Quads whose four cells do not all exist are skipped; the reference indexes
  past the grid's upper faces (include/utils.cuh:130-134).
* The QEF is solved per cell in FP32 centred at the mass point and scaled by
  the cell size, refined once, and re-solved in FP64 exactly as the
  reference poses it when the condition number exceeds min(1e6, 0.5/svd_tol).
* Vertex order is active-cell order (the reference sorts positions); faces
  keep the reference's quad order.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib, workspace
from .grid import check_cuda_f32

P = ctypes.POINTER


class IsoGrid(ctypes.Structure):
    _fields_ = [("values", ctypes.c_void_p), ("sparse_cell_ids", ctypes.c_void_p), ("n_sparse", ctypes.c_int64),
                ("shape", ctypes.c_int64 * 3), ("aabb_min", ctypes.c_float * 3), ("aabb_max", ctypes.c_float * 3)]


class IsoIts(ctypes.Structure):
    _fields_ = [("ncells", ctypes.c_int64), ("npts", ctypes.c_int64), ("cell_indices", ctypes.c_void_p),
                ("cell_dense", ctypes.c_void_p), ("cell_offsets", ctypes.c_void_p), ("points", ctypes.c_void_p),
                ("normals", ctypes.c_void_p), ("edge_local", ctypes.c_void_p), ("is_out", ctypes.c_void_p),
                ("entry_cell", ctypes.c_void_p), ("amask_tab", ctypes.c_void_p), ("apre_tab", ctypes.c_void_p)]


vp, i64, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
_lib.declare("isoext_its_scratch_bytes", i64, [P(IsoGrid)])
_lib.declare("isoext_its_count", ctypes.c_int, [P(IsoGrid), f32, vp, i64, i64 * 2, vp])
_lib.declare("isoext_its_table_entries", i64, [i64 * 3])
_lib.declare("isoext_its_emit", ctypes.c_int, [P(IsoGrid), f32, vp, i64, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                               vp])
_lib.declare("isoext_its_normals", ctypes.c_int, [P(IsoGrid), P(IsoIts), vp, vp])
_lib.declare("isoext_dc_workspace_bytes", i64, [i64, i64])
_lib.declare("isoext_dc", ctypes.c_int, [P(IsoGrid), P(IsoIts), f32, f32, vp, i64, vp, i64, vp, i64, vp, vp,
                                         _lib.I64x4, vp])
_lib.declare("isoext_dc_backward", ctypes.c_int, [P(IsoGrid), P(IsoIts), f32, f32, f32, i64, vp, vp, vp, vp, vp])
_lib.declare("isoext_mc_sparse_count", ctypes.c_int, [P(IsoGrid), f32, vp, P(i64), vp])
_lib.declare("isoext_mc_sparse", ctypes.c_int, [P(IsoGrid), f32, ctypes.c_int, i64, vp, i64, vp, i64, vp, i64, vp,
                                                i64 * 2, vp])


def grid_struct(grid) -> IsoGrid:
    g = IsoGrid()
    g.values = grid.values.data_ptr() if grid.values.numel() else None
    ids = getattr(grid, "cell_ids", None)
    if ids is not None:
        g.sparse_cell_ids = ids.data_ptr() if ids.numel() else None
        g.n_sparse = ids.numel()
    else:
        g.sparse_cell_ids = None
        g.n_sparse = 0
    for i in range(3):
        g.shape[i] = int(grid.shape[i])
        g.aabb_min[i] = float(grid.aabb_min[i])
        g.aabb_max[i] = float(grid.aabb_max[i])
    return g


class Intersection:
    """Edge/surface intersections of a grid (points, optional normals)."""

    def __init__(self, grid, level, cell_indices, cell_dense, cell_offsets, points, normals, edge_local, is_out,
                 entry_cell, tables=None):
        self._grid = grid
        self._level = float(level)
        self.cell_indices = cell_indices
        self.cell_dense = cell_dense
        self.cell_offsets = cell_offsets
        self.points = points
        self.normals = normals if normals is not None else torch.zeros_like(points)
        self.edge_local = edge_local
        self.is_out = is_out
        self.entry_cell = entry_cell
        # dense grids: per (cell row, 64-cell word) active masks and first
        # active indices, the neighbour lookups of dual contouring's quads
        self._tables = tables
        self._has_normals = normals is not None

    def get_points(self) -> torch.Tensor:
        """(N, 3) float32 intersection points (a copy)."""
        return self.points.clone()

    def get_normals(self) -> torch.Tensor:
        """(N, 3) float32 normals (a copy; zeros until set or computed)."""
        return self.normals.clone()

    def has_normals(self) -> bool:
        return self._has_normals

    def set_normals(self, new_normals: torch.Tensor) -> None:
        """Copy (N, 3) float32 normals in (src/isoext_ext.cu:318-327)."""
        check_cuda_f32(new_normals, ndim=2, last=3, name="new_normals")
        if new_normals.shape != self.points.shape:
            raise RuntimeError("Cannot set values with different shapes")
        self.normals = new_normals.detach().clone()
        self._has_normals = True

    def __len__(self):
        return self.points.shape[0]

    def _struct(self, normals=None, grid=None) -> IsoIts:
        s = IsoIts()
        s.ncells = self.cell_indices.numel()
        s.npts = self.points.shape[0]
        s.cell_indices = self.cell_indices.data_ptr()
        s.cell_dense = self.cell_dense.data_ptr()
        s.cell_offsets = self.cell_offsets.data_ptr()
        s.points = self.points.data_ptr()
        n = self.normals if normals is None else normals
        s.normals = n.data_ptr()
        s.edge_local = self.edge_local.data_ptr()
        s.is_out = self.is_out.data_ptr()
        s.entry_cell = self.entry_cell.data_ptr()
        # the tables index cell rows of the grid shape they were built for
        same_shape = grid is None or [int(x) for x in grid.shape] == [int(x) for x in self._grid.shape]
        if self._tables is not None and same_shape and getattr(grid, "cell_ids", None) is None:
            s.amask_tab = self._tables[0].data_ptr()
            s.apre_tab = self._tables[1].data_ptr()
        return s


def _device(grid):
    return grid.values.device


def get_intersection(grid, level: float = 0.0, compute_normals: bool = False) -> Intersection:
    """Crossing-edge intersection points of the grid at `level`
    (linear interpolation, src/its.cu:37-90); with compute_normals, trilinear
    central-difference normals (src/its.cu:186-268)."""
    dev = _device(grid)
    L = _lib.lib()
    g = grid_struct(grid)
    nbytes = L.isoext_its_scratch_bytes(ctypes.byref(g))
    scratch = workspace.get("its", nbytes, dev)
    counts = (ctypes.c_int64 * 2)()
    with torch.cuda.device(dev):
        stream = _lib.stream_ptr(dev)
        _lib.check(L.isoext_its_count(ctypes.byref(g), float(level), scratch.data_ptr(), scratch.numel(), counts,
                                      stream), "get_intersection")
        nc, npt = int(counts[0]), int(counts[1])
        i64t = dict(dtype=torch.int64, device=dev)
        cell_indices = torch.empty(nc, **i64t)
        cell_dense = torch.empty(nc, **i64t)
        cell_offsets = torch.empty(nc + 1, **i64t)
        points = torch.empty((npt, 3), dtype=torch.float32, device=dev)
        normals = torch.empty((npt, 3), dtype=torch.float32, device=dev) if compute_normals else None
        edge_local = torch.empty(npt, dtype=torch.uint8, device=dev)
        is_out = torch.empty(npt, dtype=torch.uint8, device=dev)
        entry_cell = torch.empty(npt, dtype=torch.int32, device=dev)
        tables = None
        if not g.sparse_cell_ids and nc > 0:
            nt = int(L.isoext_its_table_entries(g.shape))
            tables = (torch.empty(nt, dtype=torch.int64, device=dev), torch.empty(nt, dtype=torch.int32, device=dev))
        _lib.check(L.isoext_its_emit(ctypes.byref(g), float(level), scratch.data_ptr(), nc, npt,
                                     cell_indices.data_ptr(), cell_dense.data_ptr(), cell_offsets.data_ptr(),
                                     points.data_ptr() if npt else None,
                                     normals.data_ptr() if (normals is not None and npt) else None,
                                     edge_local.data_ptr() if npt else None, is_out.data_ptr() if npt else None,
                                     entry_cell.data_ptr() if npt else None,
                                     tables[0].data_ptr() if tables else None,
                                     tables[1].data_ptr() if tables else None, stream), "get_intersection")
    return Intersection(grid, level, cell_indices, cell_dense, cell_offsets, points, normals, edge_local, is_out,
                        entry_cell, tables)


def compute_normals(grid, its: Intersection) -> torch.Tensor:
    """Trilinear central-difference normals for an intersection set."""
    dev = _device(grid)
    out = torch.empty_like(its.points)
    if its.points.shape[0] == 0:
        return out
    g = grid_struct(grid)
    s = its._struct(normals=out)
    with torch.cuda.device(dev):
        _lib.check(_lib.lib().isoext_its_normals(ctypes.byref(g), ctypes.byref(s), out.data_ptr(),
                                                 _lib.stream_ptr(dev)), "compute normals")
    return out


def dc_raw(grid, its: Intersection, normals: torch.Tensor, reg: float, svd_tol: float, want_dual: bool = False,
           want_vcell: bool = False):
    """Run the DC kernels; returns (v, f, info[, dual_v]); with want_vcell,
    info["vcell"] holds the active cell of each output vertex."""
    dev = _device(grid)
    L = _lib.lib()
    g = grid_struct(grid)
    s = its._struct(normals=normals, grid=grid)
    nc, npt = int(s.ncells), int(s.npts)
    ws = workspace.get("dc", L.isoext_dc_workspace_bytes(nc, npt), dev)
    out_v = torch.empty((max(nc, 1), 3), dtype=torch.float32, device=dev)
    out_f = torch.empty((max(2 * npt, 1), 3), dtype=torch.int32, device=dev)
    dual = torch.empty((max(nc, 1), 3), dtype=torch.float32, device=dev) if want_dual else None
    vcell = torch.empty(max(nc, 1), dtype=torch.int32, device=dev) if want_vcell else None
    counts = _lib.I64x4()
    with torch.cuda.device(dev):
        _lib.check(L.isoext_dc(ctypes.byref(g), ctypes.byref(s), float(reg), float(svd_tol), ws.data_ptr(),
                               ws.numel(), out_v.data_ptr(), out_v.shape[0], out_f.data_ptr(), out_f.shape[0],
                               dual.data_ptr() if dual is not None else None,
                               vcell.data_ptr() if vcell is not None else None, counts, _lib.stream_ptr(dev)),
                   "dual_contouring")
    nv, nf = int(counts[0]), int(counts[1])
    info = {"fp64_resolves": int(counts[2]), "quads": int(counts[3])}
    if want_vcell:
        info["vcell"] = vcell[:nv]
    v, f = out_v[:nv], out_f[:nf]
    if out_v.shape[0] > 2 * nv + 65536:
        v = v.clone()
    if out_f.shape[0] > 2 * nf + 65536:
        f = f.clone()
    if want_dual:
        return v, f, info, dual[:nc]
    return v, f, info


def dual_contouring(grid, level: float = 0.0, intersection: Intersection | None = None, reg: float = 1e-2,
                    svd_tol: float = 1e-6):
    """Extract an iso-surface with Dual Contouring.

    Without `intersection`, intersections and normals are computed at
    `level`; an intersection without normals gets trilinear normals.
    Returns (vertices (N, 3) float32, faces (M, 3) int32) or (None, None).

    On a UniformGrid without `intersection` no intersection set is stored:
    each active cell's crossing edges, points and trilinear normals are
    recomputed from its corner values (the same floats, in the same order, as
    get_intersection(grid, level, True) holds), which gives the same mesh.
    """
    from .grid import UniformGrid
    if intersection is None and isinstance(grid, UniformGrid):
        from .batch import GridJob, dc_batch
        job = GridJob(grid.values, float(level), 0, tuple(grid.aabb_min), tuple(grid.aabb_max))
        return dc_batch([job], reg=reg, svd_tol=svd_tol, ws_kind="dc_grid")[0]
    its = intersection if intersection is not None else get_intersection(grid, level, True)
    if its.points.shape[0] == 0:
        return None, None
    normals = its.normals if its.has_normals() else compute_normals(grid, its)
    v, f, _ = dc_raw(grid, its, normals.contiguous(), reg, svd_tol)
    if f.shape[0] == 0:
        return None, None
    return v, f
