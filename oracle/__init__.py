"""CPU oracle for the iso-surface hot path -- TEST INFRASTRUCTURE ONLY.

A C++ restatement of the reference algorithms (GuangyanCai/isoext v0.5.1):
marching cubes on uniform and sparse grids (src/mc/mc.cu:17-68,
src/mc/nagae.cu:34-80, src/mc/lorensen.cu:34-80, src/utils.cu:32-59),
edge intersections and trilinear normals (src/its.cu:13-284) and dual
contouring with an FP64 QEF solve (src/dc.cu:14-218, truncation rule of
src/batched_la.cu:151-178).  Loaded with ctypes from ``oracle/liboracle.so``
(``make -C oracle``).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, and only as the checker / CPU baseline -- the
product path (isoext_amd) never calls it.

Parity pins (what checks each function against something independent):

* LUTs: closed-form crossing masks and structural invariants
  (tests/test_luts.py).
* ``mc_dense``: notebook known answers (vertex/face counts from
  doc/*.ipynb), brute-force single-cube cases against the closed-form
  case rule, Euler characteristic / manifoldness / consistent winding on
  closed surfaces, vertices on their edges (tests/test_oracle.py).
* ``mc_sparse``: equality with ``mc_dense`` on the same field, and the
  notebook's 1024^3 sparse counts.
* ``intersection``: the notebook cuboid's 24,576 points; points on their
  edges.  ``cell_normal`` (trilinear central differences): within 1e-5 of
  a/|a| on linear fields a.p + c; the closed-form derivative (y, x, 0) of
  the bilinear field x*y at the clamped cell coordinates (0.01 / 0.99) of
  edge points and of interior points; the (0, 0, 1) fallback at a saddle's
  centre; anisotropic cells (tests/test_oracle_pins.py).
* ``pinv_solve`` / ``dual_contouring``: notebook cuboid counts (6,146 V /
  12,288 F); ``numpy.linalg.pinv(A, rcond=tol)`` on random symmetric PSD
  systems whose eigenvalues straddle the truncation threshold; with reg = 0,
  a plane z = d gives the origin-centred minimum-norm vertex (0, 0, d)
  before the clip, two planes x = a, y = b give (a, b, 0), three planes
  give their corner (tests/test_oracle_pins.py, tests/test_oracle.py).
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

c_i64 = ctypes.c_int64
c_f = ctypes.c_float
P = ctypes.POINTER


class _Mesh(ctypes.Structure):
    _fields_ = [("nv", c_i64), ("nf", c_i64), ("v", P(c_f)), ("f", P(ctypes.c_int32)),
                ("n_degenerate", ctypes.c_uint64)]


class _Its(ctypes.Structure):
    _fields_ = [("ncells", c_i64), ("npts", c_i64), ("cell_indices", P(c_i64)),
                ("cell_offsets", P(c_i64)), ("points", P(c_f)), ("normals", P(c_f)),
                ("edges", P(c_i64)), ("is_out", P(ctypes.c_uint8))]


class _DC(ctypes.Structure):
    _fields_ = [("mesh", P(_Mesh)), ("ncells", c_i64), ("dual_v64", P(ctypes.c_double)),
                ("cond", P(ctypes.c_double)), ("nquads", c_i64), ("quads", P(c_i64)),
                ("split02", P(ctypes.c_uint8))]


def build() -> Path:
    """Compile ``liboracle.so`` with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _HERE / "liboracle.so"


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        fp = P(c_f)
        L.oracle_mc_dense.restype = P(_Mesh)
        L.oracle_mc_dense.argtypes = [fp, c_i64, c_i64, c_i64, fp, fp, c_f, ctypes.c_int, ctypes.c_int]
        L.oracle_mc_sparse.restype = P(_Mesh)
        L.oracle_mc_sparse.argtypes = [P(c_i64), fp, c_i64, c_i64, c_i64, c_i64, fp, fp, c_f,
                                       ctypes.c_int]
        L.oracle_count_active.restype = c_i64
        L.oracle_count_active.argtypes = [fp, c_i64, c_i64, c_i64, c_f, ctypes.c_int]
        L.oracle_intersection.restype = P(_Its)
        L.oracle_intersection.argtypes = [fp, P(c_i64), c_i64, ctypes.c_int, c_i64, c_i64, c_i64,
                                          fp, fp, c_f, ctypes.c_int]
        L.oracle_dual_contouring.restype = P(_DC)
        L.oracle_dual_contouring.argtypes = [fp, fp, P(c_i64), P(ctypes.c_uint8), c_i64,
                                             P(c_i64), P(c_i64), c_i64, ctypes.c_int, P(c_i64),
                                             c_i64, c_i64, c_i64, c_i64, fp, fp, c_f, c_f]
        L.oracle_cell_normal.restype = None
        L.oracle_cell_normal.argtypes = [fp, fp, fp, fp]
        dp = P(ctypes.c_double)
        L.oracle_pinv_solve.restype = ctypes.c_double
        L.oracle_pinv_solve.argtypes = [dp, dp, ctypes.c_double, dp]
        for name, t in (("oracle_free_mesh", _Mesh), ("oracle_free_its", _Its),
                        ("oracle_free_dc", _DC)):
            getattr(L, name).argtypes = [P(t)]
            getattr(L, name).restype = None
        _LIB = L
    return _LIB


def _fp(a):
    return a.ctypes.data_as(P(c_f))


def _ip(a):
    return a.ctypes.data_as(P(c_i64))


def _aabb(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(3))


def _take_mesh(m):
    r = m.contents
    if r.nf == 0:
        lib().oracle_free_mesh(m)
        return None, None
    v = np.ctypeslib.as_array(r.v, shape=(r.nv, 3)).copy()
    f = np.ctypeslib.as_array(r.f, shape=(r.nf, 3)).copy()
    lib().oracle_free_mesh(m)
    return v, f


METHODS = {"nagae": 0, "lorensen": 1}


def mc_dense(values, level=0.0, method="nagae", aabb_min=(-1, -1, -1), aabb_max=(1, 1, 1),
             threads=None, return_degenerate=False):
    """Reference-ordered marching cubes on a dense (X, Y, Z) float32 array.

    Returns (V, F) numpy arrays -- vertices in lexicographic order, faces in
    ascending active-cell order -- or (None, None) for an empty mesh.
    """
    vals = np.ascontiguousarray(values, dtype=np.float32)
    nx, ny, nz = vals.shape
    th = threads or min(16, os.cpu_count() or 1)
    m = lib().oracle_mc_dense(_fp(vals), nx, ny, nz, _fp(_aabb(aabb_min)), _fp(_aabb(aabb_max)),
                              float(level), METHODS[method], th)
    ndeg = m.contents.n_degenerate
    out = _take_mesh(m)
    return (out + (ndeg,)) if return_degenerate else out


def count_active(values, level=0.0, threads=None):
    vals = np.ascontiguousarray(values, dtype=np.float32)
    nx, ny, nz = vals.shape
    th = threads or min(16, os.cpu_count() or 1)
    return int(lib().oracle_count_active(_fp(vals), nx, ny, nz, float(level), th))


def mc_sparse(cell_ids, values, shape, level=0.0, method="nagae", aabb_min=(-1, -1, -1),
              aabb_max=(1, 1, 1)):
    ids = np.ascontiguousarray(cell_ids, dtype=np.int64)
    vals = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 8)
    m = lib().oracle_mc_sparse(_ip(ids), _fp(vals), len(ids), *map(int, shape),
                               _fp(_aabb(aabb_min)), _fp(_aabb(aabb_max)), float(level),
                               METHODS[method])
    return _take_mesh(m)


def intersection(values, level=0.0, aabb_min=(-1, -1, -1), aabb_max=(1, 1, 1),
                 compute_normals=False, cell_ids=None, shape=None):
    """get_intersection on a dense grid, or on a sparse grid when cell_ids is given."""
    sparse = cell_ids is not None
    vals = np.ascontiguousarray(values, dtype=np.float32)
    if sparse:
        ids = np.ascontiguousarray(cell_ids, dtype=np.int64)
        nx, ny, nz = shape
        r = lib().oracle_intersection(_fp(vals), _ip(ids), len(ids), 1, nx, ny, nz,
                                      _fp(_aabb(aabb_min)), _fp(_aabb(aabb_max)), float(level),
                                      int(compute_normals))
    else:
        nx, ny, nz = vals.shape
        dummy = np.zeros(1, np.int64)
        r = lib().oracle_intersection(_fp(vals), _ip(dummy), 0, 0, nx, ny, nz,
                                      _fp(_aabb(aabb_min)), _fp(_aabb(aabb_max)), float(level),
                                      int(compute_normals))
    c = r.contents
    n, k = c.npts, c.ncells
    out = {
        "cell_indices": np.ctypeslib.as_array(c.cell_indices, shape=(max(k, 1),))[:k].copy(),
        "cell_offsets": np.ctypeslib.as_array(c.cell_offsets, shape=(k + 1,)).copy(),
        "points": np.ctypeslib.as_array(c.points, shape=(max(n, 1), 3))[:n].copy(),
        "normals": np.ctypeslib.as_array(c.normals, shape=(max(n, 1), 3))[:n].copy(),
        "edges": np.ctypeslib.as_array(c.edges, shape=(max(n, 1), 2))[:n].copy(),
        "is_out": np.ctypeslib.as_array(c.is_out, shape=(max(n, 1),))[:n].astype(bool),
    }
    lib().oracle_free_its(r)
    return out


def dual_contouring(its, shape, reg=1e-2, svd_tol=1e-6, aabb_min=(-1, -1, -1),
                    aabb_max=(1, 1, 1), slot_cell=None):
    """FP64-QEF dual contouring from an intersection dict (see ``intersection``).

    Returns a dict with the welded mesh (V, F), the FP64 dual vertices per
    active cell, condition numbers and the oriented quads.
    """
    pts = np.ascontiguousarray(its["points"], np.float32)
    nrm = np.ascontiguousarray(its["normals"], np.float32)
    edg = np.ascontiguousarray(its["edges"], np.int64)
    out = np.ascontiguousarray(its["is_out"], np.uint8)
    ci = np.ascontiguousarray(its["cell_indices"], np.int64)
    off = np.ascontiguousarray(its["cell_offsets"], np.int64)
    sparse = slot_cell is not None
    sc = np.ascontiguousarray(slot_cell if sparse else np.zeros(1), np.int64)
    nx, ny, nz = shape
    r = lib().oracle_dual_contouring(_fp(pts), _fp(nrm), _ip(edg),
                                     out.ctypes.data_as(P(ctypes.c_uint8)), len(out), _ip(ci),
                                     _ip(off), len(ci), int(sparse), _ip(sc), len(sc), nx, ny, nz,
                                     _fp(_aabb(aabb_min)), _fp(_aabb(aabb_max)), float(reg),
                                     float(svd_tol))
    c = r.contents
    k, q = c.ncells, c.nquads
    res = {
        "dual_v64": np.ctypeslib.as_array(c.dual_v64, shape=(max(k, 1), 3))[:k].copy(),
        "cond": np.ctypeslib.as_array(c.cond, shape=(max(k, 1),))[:k].copy(),
        "quads": np.ctypeslib.as_array(c.quads, shape=(max(q, 1), 4))[:q].copy(),
        "split02": np.ctypeslib.as_array(c.split02, shape=(max(q, 1),))[:q].astype(bool),
    }
    addr = ctypes.cast(c.mesh, ctypes.c_void_p).value
    c.mesh = None
    res["V"], res["F"] = _take_mesh(ctypes.cast(addr, P(_Mesh)))
    lib().oracle_free_dc(r)
    return res


def cell_normal(corner_pos, corner_val, point):
    """compute_normals_op (src/its.cu:186-268) at one point of one cell:
    corner_pos (8, 3), corner_val (8,) in the reference's corner order."""
    cp = np.ascontiguousarray(corner_pos, np.float32).reshape(8, 3)
    cv = np.ascontiguousarray(corner_val, np.float32).reshape(8)
    p = np.ascontiguousarray(point, np.float32).reshape(3)
    n = np.zeros(3, np.float32)
    lib().oracle_cell_normal(_fp(cp), _fp(cv), _fp(p), _fp(n))
    return n


def pinv_solve(A, b, tol):
    """The truncated pseudo-inverse solve of src/batched_la.cu:151-178 on one
    symmetric 3x3 system (FP64).  Returns (x, condition number)."""
    a = np.ascontiguousarray(A, np.float64).reshape(9)
    bb = np.ascontiguousarray(b, np.float64).reshape(3)
    x = np.zeros(3, np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    cond = lib().oracle_pinv_solve(a.ctypes.data_as(dp), bb.ctypes.data_as(dp), float(tol),
                                   x.ctypes.data_as(dp))
    return x, cond


# ---- canonical mesh comparison ----------------------------------------------

def canonical(v, f):
    """Vertices sorted lexicographically; faces remapped, rotated so the
    smallest index comes first (orientation kept) and sorted."""
    if v is None or f is None or len(f) == 0:
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int64)
    v = np.asarray(v, np.float32)
    f = np.asarray(f, np.int64)
    order = np.lexsort((v[:, 2], v[:, 1], v[:, 0]))
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    vs = v[order]
    fs = inv[f]
    r = np.argmin(fs, axis=1)
    idx = (np.arange(3)[None, :] + r[:, None]) % 3
    fs = np.take_along_axis(fs, idx, axis=1)
    fo = np.lexsort((fs[:, 2], fs[:, 1], fs[:, 0]))
    return vs, fs[fo]
