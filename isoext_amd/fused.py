"""Extraction from fields that are never stored: sphere unions and SDF programs.

The reference's users build a grid with ``UniformGrid.set_values(sdf(p))``
and then extract (the step before the path).  For the union of spheres --
the field of the benchmark workloads -- ``marching_cubes_spheres`` skips the
grid: the classify kernel evaluates, per block of grid points, only the
spheres that can reach below the level there and writes inside bits, and the
emit kernel evaluates the corners of its active cells over the spheres that
can be the minimum in its tile.  The mesh is identical (same vertices, same
order) to ``marching_cubes`` on the grid that ``workloads.spheres_field``
writes for the same spheres, without the 4 bytes per point written and read.

``marching_cubes_sdf`` and ``dual_contouring_sdf`` do the same for any SDF
tree of the reference's primitives and operators (src/isoext/sdf.py:70-221)
compiled into a program (csg.py): classify evaluates the program at the
centres of point blocks and decides whole blocks by the field's Lipschitz
bound, then evaluates the points of the undecided ones; emit and the QEF
pass evaluate the corners of their active cells.  The meshes equal
``marching_cubes`` / ``dual_contouring`` on the grid ``SdfProgram.grid``
writes.  Dual contouring can take the program's analytic normals (the
normals ``get_sdf_normal`` gives at the intersection points) in place of
trilinear ones.
"""

from __future__ import annotations

import numpy as np
import torch

from .csg import as_program
from .mc import METHODS, mc_dense_raw


def sphere_tensor(centers, radii, device="cuda") -> torch.Tensor:
    """(K, 4) float32 (cx, cy, cz, r) on the device."""
    c = torch.as_tensor(np.asarray(centers, np.float32)).reshape(-1, 3)
    r = torch.as_tensor(np.asarray(radii, np.float32)).reshape(-1, 1)
    if c.shape[0] != r.shape[0]:
        raise ValueError("centers and radii differ in length")
    return torch.cat([c, r], 1).contiguous().to(device)


def marching_cubes_spheres(shape, centers, radii, level: float = 0.0, method: str = "nagae",
                           aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0), device="cuda",
                           cfg: int | None = None):
    """Marching cubes of v(p) = min_k |p - c_k| - r_k (1 to 256 spheres) on
    the grid points of (shape, aabb).  Returns (vertices (N, 3) float32,
    faces (M, 3) int32) or (None, None) when no surface crosses the grid."""
    if method not in METHODS:
        raise RuntimeError(f"Unknown method: {method}")
    shape = tuple(int(s) for s in shape)
    if len(shape) != 3:
        raise ValueError("shape must have 3 entries")
    sp = sphere_tensor(centers, radii, device)
    if not 1 <= sp.shape[0] <= 256:
        raise ValueError("between 1 and 256 spheres")
    v, f, _ = mc_dense_raw(None, shape, aabb_min, aabb_max, level, METHODS[method], cfg, spheres=sp)
    if f.shape[0] == 0:
        return None, None
    return v, f


def marching_cubes_sdf(sdf, shape, level: float = 0.0, method: str = "nagae", aabb_min=(-1.0, -1.0, -1.0),
                       aabb_max=(1.0, 1.0, 1.0), device="cuda", cfg: int | None = None):
    """Marching cubes of an SDF tree (or csg.SdfProgram) on the grid points
    of (shape, aabb) without storing the grid.  Returns (vertices, faces) or
    (None, None)."""
    if method not in METHODS:
        raise RuntimeError(f"Unknown method: {method}")
    shape = tuple(int(s) for s in shape)
    if len(shape) != 3:
        raise ValueError("shape must have 3 entries")
    prog = as_program(sdf, device)
    v, f, _ = mc_dense_raw(None, shape, aabb_min, aabb_max, level, METHODS[method], cfg, program=prog)
    if f.shape[0] == 0:
        return None, None
    return v, f


def dual_contouring_sdf(sdf, shape, level: float = 0.0, reg: float = 1e-2, svd_tol: float = 1e-6,
                        normals: str = "analytic", aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0),
                        device="cuda"):
    """Dual contouring of an SDF tree (or csg.SdfProgram) without storing
    the grid.  normals: "analytic" (the program's unit gradients at the
    intersection points) or "trilinear" (from the corner values, as
    dual_contouring on the stored grid).  Returns (vertices, faces) or
    (None, None)."""
    if normals not in ("analytic", "trilinear"):
        raise ValueError('normals must be "analytic" or "trilinear"')
    from .batch import GridJob, dc_batch
    prog = as_program(sdf, device)
    job = GridJob(None, float(level), 0, tuple(aabb_min), tuple(aabb_max), sdf=prog,
                  shape=tuple(int(s) for s in shape), sdf_normals=normals == "analytic")
    return dc_batch([job], reg=reg, svd_tol=svd_tol, ws_kind="dc_grid")[0]
