"""marching_cubes -- drop-in for isoext.marching_cubes.

Reference: the nanobind binding src/isoext_ext.cu:95-109 over
mc::marching_cubes (src/mc/mc.cu:17-68) and the variant registry
(include/mc/base.cuh:12-41, src/mc/base.cu:19-27: an unknown name raises
RuntimeError("Unknown method: <name>")).  Returns (vertices (V, 3) float32,
faces (F, 3) int32) on the grid's device, or (None, None) for an empty mesh
(src/isoext_ext.cu:47-49).

Vertex order differs from the reference (which sorts positions
lexicographically): vertices come in grid-edge order, faces in tile/cell
order.  The vertex set, the face set and every face's corner order are the
reference's (compare after canonical ordering).
"""

from __future__ import annotations

import threading

import torch

from . import _lib, autotune, workspace
from .grid import Grid, UniformGrid

METHODS = {"nagae": 0, "lorensen": 1}

# provisional output sizes seen last time for a grid shape (sizing hint), and
# the weld-candidate capacity a shape needed when above the default
_size_hint: dict = {}
_special_hint: dict = {}
_hint_lock = threading.Lock()
DEFAULT_SPECIAL_CAP = 1 << 16
PROGRAM_CFG = 5   # (8, 8, 1) emit tiles for SDF-program fields


def _method_id(method: str) -> int:
    if method not in METHODS:
        raise RuntimeError(f"Unknown method: {method}")
    return METHODS[method]


def _initial_caps(shape):
    X, Y, Z = shape
    hint = _size_hint.get(tuple(shape))   # a single dict read needs no lock
    if hint is not None:
        return hint
    # surface-like estimate: a few times the largest cross-section, at least 64K
    cross = max(X * Y, Y * Z, X * Z)
    nv = max(1 << 16, 4 * cross)
    return nv, 2 * nv + 1024


_ws_bytes_cache: dict = {}
_abi_args: dict = {}   # (shape, aabb_min, aabb_max) -> the ctypes arrays of a call


def _abi_arrays(shape, aabb_min, aabb_max):
    key = (shape, tuple(aabb_min), tuple(aabb_max))
    a = _abi_args.get(key)
    if a is None:
        if len(_abi_args) > 4096:
            _abi_args.clear()
        a = _abi_args[key] = (_lib.i64x3(shape), _lib.f32x3(aabb_min), _lib.f32x3(aabb_max))
    return a


def mc_dense_raw(values: torch.Tensor | None, shape, aabb_min, aabb_max, level: float, method: int,
                 cfg: int | None = None, want_keys: bool = False, spheres: torch.Tensor | None = None,
                 program=None):
    """Run dense marching cubes on a (X, Y, Z) float32 CUDA tensor, or, with
    `spheres` ((K, 4) float32 (cx, cy, cz, r) on the device; values None), on
    the sphere-union field evaluated inside the kernels (never stored), or,
    with `program` (a csg.SdfProgram; values None), on the program's field.

    Returns (v, f, info) with v (V, 3) float32 and f (F, 3) int32 tensors
    (possibly zero rows) and an info dict (provisional counts, tile config;
    with want_keys, info["keys"] holds each vertex's grid edge).
    """
    if program is not None:
        from .csg import _decl as _csg_decl
        _csg_decl()
        dev = program.device
    else:
        dev = values.device if values is not None else spheres.device
    shape = tuple(int(s) for s in shape)
    if cfg is None:
        # program fields: small tiles, so each tile's distinct corner points
        # fit its LDS cache (every corner evaluated once)
        cfg = PROGRAM_CFG if program is not None else autotune.mc_config(shape)
    L = _lib.lib()
    special_cap = max(DEFAULT_SPECIAL_CAP, _special_hint.get(shape, 0))
    cap_v, cap_f = _initial_caps(shape)
    counts = _lib.I64x6()
    shape_c, amin_c, amax_c = _abi_arrays(shape, aabb_min, aabb_max)
    stream = _lib.stream_ptr(dev)
    for attempt in range(4):
        key = (shape, cfg, special_cap)
        ws_bytes = _ws_bytes_cache.get(key)
        if ws_bytes is None:
            ws_bytes = L.isoext_mc_dense_workspace_bytes(shape_c, cfg, special_cap)
            _ws_bytes_cache[key] = ws_bytes
        if ws_bytes < 0:
            raise ValueError(f"invalid tile config {cfg}")
        ws = workspace.get("mc_dense", ws_bytes, dev, stream)
        out_v = torch.empty((cap_v, 3), dtype=torch.float32, device=dev)
        out_f = torch.empty((cap_f, 3), dtype=torch.int32, device=dev)
        keys = torch.empty(cap_v, dtype=torch.int64, device=dev) if want_keys else None
        with _lib.on_device(dev):
            if program is not None:
                rc = L.isoext_mc_sdf(program.ptr, shape_c, amin_c, amax_c, float(level), int(method), int(cfg),
                                     ws.data_ptr(), ws.numel(), special_cap, out_v.data_ptr(), cap_v,
                                     out_f.data_ptr(), cap_f, counts, stream)
            elif spheres is not None:
                rc = L.isoext_mc_spheres(spheres.data_ptr(), int(spheres.shape[0]), shape_c, amin_c, amax_c,
                                         float(level), int(method), int(cfg), ws.data_ptr(), ws.numel(), special_cap,
                                         out_v.data_ptr(), cap_v, out_f.data_ptr(), cap_f, counts, stream)
            elif want_keys:
                rc = L.isoext_mc_dense_keyed(values.data_ptr(), shape_c, amin_c, amax_c, float(level), int(method),
                                             int(cfg), ws.data_ptr(), ws.numel(), special_cap, out_v.data_ptr(),
                                             cap_v, out_f.data_ptr(), cap_f, keys.data_ptr(), counts, stream)
            else:
                rc = L.isoext_mc_dense(values.data_ptr(), shape_c, amin_c, amax_c, float(level), int(method),
                                       int(cfg), ws.data_ptr(), ws.numel(), special_cap, out_v.data_ptr(), cap_v,
                                       out_f.data_ptr(), cap_f, counts, stream)
        if rc == _lib.E_CAPACITY:
            cap_v, cap_f = int(counts[0]) + 1, int(counts[1]) + 1
            continue
        if rc == _lib.E_WORKSPACE and counts[2] > special_cap:
            special_cap = int(counts[2]) + 1024
            continue
        _lib.check(rc, "marching_cubes")
        break
    else:
        raise _lib.IsoextError("marching_cubes: output sizing did not converge")
    hint = (max(int(counts[4]), 1), max(int(counts[5]), 1))
    if _size_hint.get(shape) != hint or special_cap > DEFAULT_SPECIAL_CAP:
        with _hint_lock:
            _size_hint[shape] = hint
            if special_cap > DEFAULT_SPECIAL_CAP:   # fields with many corner-exact crossings keep their need
                _special_hint[shape] = max(special_cap, _special_hint.get(shape, 0))
    nv, nf = int(counts[0]), int(counts[1])
    v, f = out_v[:nv], out_f[:nf]
    # do not pin a much larger allocation behind a small mesh
    if cap_v > 2 * nv + (1 << 16):
        v = v.clone()
    if cap_f > 2 * nf + (1 << 16):
        f = f.clone()
    info = {"cfg": cfg, "special": int(counts[2]), "err": int(counts[3])}
    if want_keys:
        info["keys"] = keys[:nv]
    return v, f, info


def marching_cubes(grid: Grid, level: float = 0.0, method: str = "nagae"):
    """Extract an iso-surface with Marching Cubes.

    Args:
        grid: a UniformGrid or SparseGrid.
        level: iso-value (default 0.0).
        method: "nagae" (default) or "lorensen".

    Returns:
        (vertices (N, 3) float32, faces (M, 3) int32), or (None, None) when
        no surface crosses the grid.
    """
    m = _method_id(method)
    if isinstance(grid, UniformGrid):
        v, f, _ = mc_dense_raw(grid.values, grid.shape, grid.aabb_min, grid.aabb_max, level, m)
    else:
        from .sparse import SparseGrid, mc_sparse
        if not isinstance(grid, SparseGrid):
            raise TypeError("grid must be a UniformGrid or SparseGrid")
        v, f = mc_sparse(grid, level, m)
    if v is None or f is None or f.shape[0] == 0:
        return None, None
    return v, f


class MarchingCubesPlan:
    """Repeated marching cubes of one grid tensor whose values change in
    place (an optimisation loop, an animation): the pipeline's kernel
    launches are captured once into a HIP graph (isoext_mc_plan_create) and
    each run() replays it, reads the counts back and welds when needed.  The
    plan owns its workspace and output buffers, sized from a first
    extraction with headroom; a run that outgrows them rebuilds the plan.
    Results equal marching_cubes on the tensor's current values."""

    def __init__(self, values: torch.Tensor, level: float = 0.0, method: str = "nagae",
                 aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0), cfg: int | None = None,
                 headroom: float = 1.25, special_cap: int | None = None):
        """special_cap: capacity of the special-vertex lists (weld candidates
        and their cells) the plan starts with; by default the larger of the
        library default and twice the first extraction's need."""
        import ctypes
        from .grid import check_cuda_f32
        check_cuda_f32(values, 3, name="values")
        if not values.is_contiguous():
            raise TypeError("values must be C-contiguous")
        self.values = values
        self.level = float(level)
        self.method = _method_id(method)
        self.aabb_min, self.aabb_max = tuple(aabb_min), tuple(aabb_max)
        self.shape = tuple(int(s) for s in values.shape)
        self.cfg = autotune.mc_config(self.shape) if cfg is None else int(cfg)
        if self.cfg == 6:
            self.cfg = 3   # plans take the four-stage configurations
        self.headroom = float(headroom)
        self._h = None
        self._L = _lib.lib()
        L = self._L
        L.isoext_mc_plan_create.restype = ctypes.c_int
        L.isoext_mc_plan_create.argtypes = [ctypes.c_void_p, _lib.I64x3, _lib.F32x3, _lib.F32x3, ctypes.c_float,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                            ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]
        L.isoext_mc_plan_run.restype = ctypes.c_int
        L.isoext_mc_plan_run.argtypes = [ctypes.c_void_p, _lib.I64x6, ctypes.c_void_p]
        L.isoext_mc_plan_destroy.restype = None
        L.isoext_mc_plan_destroy.argtypes = [ctypes.c_void_p]
        v, f, info = mc_dense_raw(values, self.shape, self.aabb_min, self.aabb_max, self.level, self.method,
                                  self.cfg)
        if special_cap is None:
            with _hint_lock:
                special = max(DEFAULT_SPECIAL_CAP, _special_hint.get(self.shape, 0), 2 * int(info["special"]))
        else:
            special = max(int(special_cap), 1)
        self._build(int(v.shape[0]), int(f.shape[0]), special)

    def _build(self, nv: int, nf: int, special: int):
        import ctypes
        self.close()
        dev = self.values.device
        self.special = int(special)
        self.cap_v = max(int(nv * self.headroom) + 1024, 1 << 12)
        self.cap_f = max(int(nf * self.headroom) + 2048, 1 << 13)
        shape_c = _lib.i64x3(self.shape)
        ws_bytes = int(self._L.isoext_mc_dense_workspace_bytes(shape_c, self.cfg, self.special))
        if ws_bytes < 0:
            raise ValueError(f"invalid tile config {self.cfg}")
        self.ws = torch.zeros(ws_bytes, dtype=torch.uint8, device=dev)   # this plan's alone: the graph holds it
        self.out_v = torch.empty((self.cap_v, 3), dtype=torch.float32, device=dev)
        self.out_f = torch.empty((self.cap_f, 3), dtype=torch.int32, device=dev)
        h = ctypes.c_void_p()
        with _lib.on_device(dev):
            _lib.check(self._L.isoext_mc_plan_create(self.values.data_ptr(), shape_c, _lib.f32x3(self.aabb_min),
                                                     _lib.f32x3(self.aabb_max), self.level, self.method, self.cfg,
                                                     self.ws.data_ptr(), ws_bytes, self.special,
                                                     self.out_v.data_ptr(), self.cap_v, self.out_f.data_ptr(),
                                                     self.cap_f, ctypes.byref(h)), "marching cubes plan")
        self._h = h

    def run(self, copy: bool = True):
        """Extract from the tensor's current values.  Returns (vertices,
        faces) or (None, None); with copy=False, views of the plan's buffers
        that the next run overwrites."""
        dev = self.values.device
        for _ in range(3):
            counts = _lib.I64x6()
            with _lib.on_device(dev):
                rc = self._L.isoext_mc_plan_run(self._h, counts, _lib.stream_ptr(dev))
            if rc == _lib.E_CAPACITY:
                self._build(int(counts[0]), int(counts[1]), self.special)
                continue
            if rc == _lib.E_WORKSPACE and counts[2] > self.special:
                nv, nf = self.cap_v, self.cap_f
                self._build(int(nv / self.headroom), int(nf / self.headroom), int(counts[2]) + 1024)
                continue
            _lib.check(rc, "marching cubes plan")
            break
        else:
            raise _lib.IsoextError("marching cubes plan: output sizing did not converge")
        nv, nf = int(counts[0]), int(counts[1])
        if nf == 0:
            return None, None
        v, f = self.out_v[:nv], self.out_f[:nf]
        return (v.clone(), f.clone()) if copy else (v, f)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.isoext_mc_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
