"""SDF programs: an isoext.sdf tree (spheres, tori, cuboids, unions, smooth
unions, intersections, negations, translations, rotations --
src/isoext/sdf.py:70-221) compiled into the flat register program the HIP
kernels evaluate (isoext_sdf_upload in include/isoext_amd.h), so that a
field is produced where the extraction needs it instead of being written to
and read back from a grid.

    prog = SdfProgram(sdf)                 # validates and uploads once
    prog.grid(shape, aabb_min, aabb_max)   # the stored grid: set_values(sdf(p))
    prog.points(p, normals=True)           # values and analytic unit normals
    fused.marching_cubes_sdf(prog, shape)  # marching cubes without the grid

The program keeps the torch code's op order; its stored grid and the fused
extractions agree bit for bit.  Values differ from a torch evaluation of the
same tree only by the rounding of torch's own reductions (a few ulps).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import sdf as S

SPHERE, SPHERE_AT, TORUS, CUBOID, CYLINDER_Z, VALUE, SMOOTH_MIN, NEG, TRANSLATE, ROTATE = range(1, 11)
STORE, ACC_MIN, ACC_MAX = 0, 1, 2
MAX_OPS, MAX_PARAMS, VAL_SLOTS, PT_SLOTS = 256, 1024, 8, 4

_PRIMS = (S.SphereSDF, S.TorusSDF, S.CuboidSDF, S.CylinderSDF)


class _Compiler:
    def __init__(self):
        self.ops: list[tuple[int, int]] = []
        self.params: list[float] = []

    def _param(self, vals) -> int:
        o = len(self.params)
        self.params.extend(float(np.float32(v)) for v in vals)
        return o

    def _op(self, op, dst=0, src=0, acc=STORE, prm=0):
        if not (0 <= dst < VAL_SLOTS and 0 <= src < VAL_SLOTS):
            raise ValueError("SDF tree too deep for the program's register slots")
        self.ops.append((op | dst << 8 | src << 12 | acc << 16, prm))

    def emit(self, node, d: int, pt: int, acc: int = STORE):
        """Evaluate `node` at point slot pt into value slot d (combined by acc)."""
        if d >= VAL_SLOTS:
            raise ValueError("SDF tree nests deeper than 8 value slots")
        if isinstance(node, S.TranslationOp) and isinstance(node.sdf, S.SphereSDF):
            o = self._param(list(node.offset) + [node.sdf.radius])
            self._op(SPHERE_AT, d, pt, acc, o)
        elif isinstance(node, S.SphereSDF):
            self._op(SPHERE, d, pt, acc, self._param([node.radius]))
        elif isinstance(node, S.TorusSDF):
            self._op(TORUS, d, pt, acc, self._param([node.R, node.r]))
        elif isinstance(node, S.CuboidSDF):
            half = (torch.tensor(node.size, dtype=torch.float32) / 2).tolist()   # as CuboidSDF computes it
            self._op(CUBOID, d, pt, acc, self._param(half))
        elif isinstance(node, S.CylinderSDF):
            self._op(CYLINDER_Z, d, pt, acc, self._param([node.radius]))
        elif isinstance(node, (S.TranslationOp, S.RotationOp)):
            if pt + 1 >= PT_SLOTS:
                raise ValueError("SDF tree nests more than 3 transforms")
            if isinstance(node, S.TranslationOp):
                self._op(TRANSLATE, pt + 1, pt, STORE, self._param(node.offset))
            else:
                R = node.R.to(torch.float32).reshape(3, 3)
                self._op(ROTATE, pt + 1, pt, STORE, self._param(R.flatten().tolist()))
            self.emit(node.sdf, d, pt + 1, acc)
        elif isinstance(node, (S.UnionOp, S.IntersectionOp)):
            kids = list(node.sdf_list)
            if not kids:
                raise ValueError("empty union / intersection")
            comb = ACC_MIN if isinstance(node, S.UnionOp) else ACC_MAX
            if acc == STORE or acc == comb:
                # fold straight into slot d: min(min(a, b), c) = min over all
                self.emit(kids[0], d, pt, acc)
                for k in kids[1:]:
                    self._child(k, d, pt, comb)
            else:
                self.emit(node, d + 1, pt, STORE)
                self._op(VALUE, d, d + 1, acc)
        elif isinstance(node, S.SmoothUnionOp):
            kids = list(node.sdf_list)
            if not kids:
                raise ValueError("empty smooth union")
            t = d if acc == STORE else d + 1
            self.emit(kids[0], t, pt, STORE)
            kp = self._param([node.k])
            for k in kids[1:]:
                self.emit(k, t + 1, pt, STORE)
                self._op(SMOOTH_MIN, t, t + 1, STORE, kp)
            if t != d:
                self._op(VALUE, d, t, acc)
        elif isinstance(node, S.NegationOp):
            t = d if acc == STORE else d + 1
            self.emit(node.sdf, t, pt, STORE)
            self._op(NEG, t, 0, STORE)
            if t != d:
                self._op(VALUE, d, t, acc)
        else:
            raise TypeError(f"cannot compile {type(node).__name__} into an SDF program")

    def _child(self, node, d, pt, comb):
        """Fold a union/intersection child into slot d with comb."""
        prim = isinstance(node, _PRIMS) or (isinstance(node, (S.TranslationOp, S.RotationOp)))
        if prim or type(node) is (S.UnionOp if comb == ACC_MIN else S.IntersectionOp):
            self.emit(node, d, pt, comb)
        else:
            self.emit(node, d + 1, pt, STORE)
            self._op(VALUE, d, d + 1, comb)


def compile_sdf(sdf) -> tuple[np.ndarray, np.ndarray]:
    """(ops (n, 2) int32, params (m,) float32) of an isoext.sdf tree."""
    c = _Compiler()
    c.emit(sdf, 0, 0)
    if len(c.ops) > MAX_OPS:
        raise ValueError(f"SDF program has {len(c.ops)} ops (at most {MAX_OPS})")
    if len(c.params) > MAX_PARAMS:
        raise ValueError(f"SDF program has {len(c.params)} parameters (at most {MAX_PARAMS})")
    return np.asarray(c.ops, dtype=np.int32).reshape(-1, 2), np.asarray(c.params, dtype=np.float32)


_decl_done = False


def _decl():
    global _decl_done
    L = _lib.lib()
    if not _decl_done:
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.isoext_sdf_program_bytes.restype = i64
        L.isoext_sdf_program_bytes.argtypes = []
        L.isoext_sdf_upload.restype = ctypes.c_int
        L.isoext_sdf_upload.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, vp, vp]
        L.isoext_sdf_grid.restype = ctypes.c_int
        L.isoext_sdf_grid.argtypes = [vp, _lib.I64x3, _lib.F32x3, _lib.F32x3, vp, vp]
        L.isoext_sdf_points.restype = ctypes.c_int
        L.isoext_sdf_points.argtypes = [vp, vp, i64, vp, vp, vp]
        L.isoext_mc_sdf.restype = ctypes.c_int
        L.isoext_mc_sdf.argtypes = [vp, _lib.I64x3, _lib.F32x3, _lib.F32x3, ctypes.c_float, ctypes.c_int,
                                    ctypes.c_int, vp, i64, i64, vp, i64, vp, i64, _lib.I64x6, vp]
        _decl_done = True
    return L


class SdfProgram:
    """An SDF tree compiled and uploaded to one device."""

    def __init__(self, sdf, device="cuda"):
        self.sdf = sdf
        self.ops, self.params = compile_sdf(sdf)
        L = _decl()
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        nbytes = int(L.isoext_sdf_program_bytes())
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=dev)   # 256-byte aligned by the allocator
        ops = np.ascontiguousarray(self.ops)
        prm = np.ascontiguousarray(self.params) if self.params.size else np.zeros(1, np.float32)
        with torch.cuda.device(dev):
            _lib.check(L.isoext_sdf_upload(ops.ctypes.data, ops.shape[0], prm.ctypes.data, int(self.params.size),
                                           self.buf.data_ptr(), _lib.stream_ptr(dev)), "SDF program")

    @property
    def ptr(self) -> int:
        return self.buf.data_ptr()

    def grid(self, shape, aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0), out=None) -> torch.Tensor:
        """(X, Y, Z) float32 values at the grid points of (shape, aabb)."""
        shape = tuple(int(s) for s in shape)
        if out is None:
            out = torch.empty(shape, dtype=torch.float32, device=self.device)
        L = _decl()
        with torch.cuda.device(self.device):
            _lib.check(L.isoext_sdf_grid(self.ptr, _lib.i64x3(shape), _lib.f32x3(aabb_min), _lib.f32x3(aabb_max),
                                         out.data_ptr(), _lib.stream_ptr(self.device)), "SDF grid")
        return out

    def points(self, p: torch.Tensor, normals: bool = False):
        """Values at points p (..., 3); with normals, (values, unit normals)
        as get_sdf_normal (src/isoext/sdf.py:39-50)."""
        shp = p.shape[:-1]
        q = p.reshape(-1, 3).to(device=self.device, dtype=torch.float32).contiguous()
        n = q.shape[0]
        vals = torch.empty(n, dtype=torch.float32, device=self.device)
        nrm = torch.empty((n, 3), dtype=torch.float32, device=self.device) if normals else None
        L = _decl()
        with torch.cuda.device(self.device):
            _lib.check(L.isoext_sdf_points(self.ptr, q.data_ptr() if n else None, n, vals.data_ptr() if n else None,
                                           nrm.data_ptr() if (nrm is not None and n) else None,
                                           _lib.stream_ptr(self.device)), "SDF points")
        if normals:
            return vals.reshape(shp), nrm.reshape(*shp, 3)
        return vals.reshape(shp)


def as_program(sdf_or_prog, device="cuda") -> SdfProgram:
    return sdf_or_prog if isinstance(sdf_or_prog, SdfProgram) else SdfProgram(sdf_or_prog, device)
