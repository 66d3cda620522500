"""Synthetic inputs of the benchmark workloads, generated on the GPU in fp32.

* ``multi_sphere``: min over n spheres, centres U[-0.8, 0.8]^3 and radii
  U[rmin, rmax] drawn in FP64 from numpy.random.default_rng(seed) and cast
  to fp32; each distance is sqrt((dx*dx + dy*dy) + dz*dz) - r in fp32.
* ``csg_box_cylinder``: min(max(box_0.6(p), -(|p_xy| - 0.3)), min_k sphere_k)
  with its analytic gradient (the dual-contouring workload).
* ``noisy_sphere``: sphere r plus a 3-octave seeded Fourier noise.
Grid point coordinates follow get_vtx_pos_op (include/utils.cuh:62-80).
"""

from __future__ import annotations

import numpy as np
import torch


def axis_coords(n: int, amin: float = -1.0, amax: float = 1.0, device="cuda") -> torch.Tensor:
    i = torch.arange(n, device=device, dtype=torch.float32)
    size = torch.tensor(np.float32(amax) - np.float32(amin), device=device)
    return torch.tensor(np.float32(amin), device=device) + (i / np.float32(n - 1)) * size


def sphere_params(n: int, seed: int, rmin=0.05, rmax=0.25, lo=-0.8, hi=0.8):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    r = rng.uniform(rmin, rmax, size=n).astype(np.float32)
    return c, r


@torch.no_grad()
def spheres_field(shape, centers, radii, device="cuda", out: torch.Tensor | None = None,
                  aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0)) -> torch.Tensor:
    """(X, Y, Z) fp32 min-of-spheres field evaluated in one HIP kernel."""
    import ctypes

    from . import _lib
    if out is None:
        out = torch.empty(tuple(int(s) for s in shape), dtype=torch.float32, device=device)
    sp = torch.cat([torch.as_tensor(np.asarray(centers, np.float32)).reshape(-1, 3),
                    torch.as_tensor(np.asarray(radii, np.float32)).reshape(-1, 1)], 1).contiguous().to(out.device)
    L = _lib.lib()
    fn = L.isoext_sdf_spheres
    fn.restype = ctypes.c_int
    fn.argtypes = [_lib.I64x3, _lib.F32x3, _lib.F32x3, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                   ctypes.c_void_p]
    with torch.cuda.device(out.device):
        rc = fn(_lib.i64x3(shape), _lib.f32x3(aabb_min), _lib.f32x3(aabb_max), sp.data_ptr(), sp.shape[0],
                out.data_ptr(), _lib.stream_ptr(out.device))
    if rc != 0:
        raise RuntimeError("isoext_sdf_spheres failed")
    return out


@torch.no_grad()
def multi_sphere(shape, n_spheres=64, seed=0, rmin=0.05, rmax=0.25, device="cuda",
                 out: torch.Tensor | None = None, slab: int = 32, fused: bool = True) -> torch.Tensor:
    """(X, Y, Z) fp32 field: min over seeded spheres."""
    if fused:
        c, r = sphere_params(n_spheres, seed, rmin, rmax)
        return spheres_field(shape, c, r, device=device, out=out)
    X, Y, Z = shape
    c, r = sphere_params(n_spheres, seed, rmin, rmax)
    ax, ay, az = (axis_coords(n, device=device) for n in shape)
    if out is None:
        out = torch.empty((X, Y, Z), dtype=torch.float32, device=device)
    ct = torch.from_numpy(c).to(device)
    rt = torch.from_numpy(r).to(device)
    for x0 in range(0, X, slab):
        xs = ax[x0:x0 + slab].view(-1, 1, 1)
        acc = None
        for k in range(n_spheres):
            dx = xs - ct[k, 0]
            dy = ay.view(1, -1, 1) - ct[k, 1]
            dz = az.view(1, 1, -1) - ct[k, 2]
            d = ((dx * dx + dy * dy) + dz * dz).sqrt() - rt[k]
            acc = d if acc is None else torch.minimum(acc, d)
        out[x0:x0 + slab] = acc
    return out


def csg_box_cylinder_sdf(n_spheres=16, seed=0):
    """The dual-contouring workload as an isoext.sdf tree: box (half-extent
    0.6) minus a z-cylinder (r 0.3), union seeded spheres r in [0.05, 0.15]
    (the same floats, in the same op order, as csg_eval)."""
    from . import sdf as S
    c, r = sphere_params(n_spheres, seed, 0.05, 0.15)
    body = S.IntersectionOp([S.CuboidSDF([1.2, 1.2, 1.2]), S.NegationOp(S.CylinderSDF(0.3))])
    return S.UnionOp([body] + [S.TranslationOp(S.SphereSDF(float(r[k])), [float(x) for x in c[k]])
                               for k in range(n_spheres)])


@torch.no_grad()
def csg_box_cylinder(shape, n_spheres=16, seed=0, device="cuda", program: bool = True):
    """(X, Y, Z) field of the dual-contouring workload, written by the SDF
    program evaluator (program=True; the grid the fused extractions
    reproduce) or by the torch ops of csg_eval."""
    if program:
        from .csg import SdfProgram
        return SdfProgram(csg_box_cylinder_sdf(n_spheres, seed), device).grid(shape)
    ax, ay, az = (axis_coords(n, device=device) for n in shape)
    p = torch.stack(torch.meshgrid(ax, ay, az, indexing="ij"), -1)
    return csg_eval(p, n_spheres, seed)


def csg_eval(p: torch.Tensor, n_spheres=16, seed=0):
    q = p.abs() - 0.6
    box = torch.norm(q.clamp(min=0), dim=-1) + q.max(dim=-1).values.clamp(max=0)
    cyl = p[..., :2].norm(dim=-1) - 0.3
    d = torch.maximum(box, -cyl)
    c, r = sphere_params(n_spheres, seed, 0.05, 0.15)
    ct, rt = torch.from_numpy(c).to(p.device), torch.from_numpy(r).to(p.device)
    for k in range(n_spheres):
        d = torch.minimum(d, (p - ct[k]).norm(dim=-1) - rt[k])
    return d


def csg_normals(p: torch.Tensor, n_spheres=16, seed=0):
    """Analytic-gradient normals of the CSG field at points p (via autograd of
    the active branch)."""
    p = p.detach().clone().requires_grad_(True)
    d = csg_eval(p, n_spheres, seed)
    (g,) = torch.autograd.grad(d.sum(), p)
    return torch.nn.functional.normalize(g, dim=-1).detach()


def noisy_sphere_eval(p: torch.Tensor, r=0.7, amp=0.02, seed=0, octaves=3):
    """Sphere plus seeded Fourier noise: sum_o amp/2^o * sin(k_o . p + phi_o)."""
    rng = np.random.default_rng(seed)
    d = p.norm(dim=-1) - r
    for o in range(octaves):
        k = torch.tensor(rng.normal(size=3) * (4.0 * 2 ** o), dtype=p.dtype, device=p.device)
        phi = float(rng.uniform(0, 2 * np.pi))
        d = d + (amp / 2 ** o) * torch.sin((p * k).sum(-1) + phi)
    return d


@torch.no_grad()
def narrow_band_ids(n, band_h, sdf, device="cuda", block=16):
    """Sorted int64 ids of the cells of an n^3-point grid whose centre lies
    within band_h cells of the surface: coarse blocks of block^3 cells are
    kept when their centre is within the band plus the block's half-diagonal
    (times a Lipschitz margin of 2), then their cells are tested."""
    import torch
    nc = n - 1
    h = 2.0 / nc
    nb = (nc + block - 1) // block
    bi = torch.arange(nb, device=device, dtype=torch.float32)
    bc = -1.0 + (bi * block + block / 2.0) * h
    keep = []
    margin = band_h * h + 2.0 * (3 ** 0.5) * block * h / 2
    for x in range(nb):
        P = torch.stack(torch.meshgrid(bc[x:x + 1], bc, bc, indexing="ij"), -1).reshape(-1, 3)
        d = sdf(P).abs()
        sel = torch.nonzero(d < margin).flatten()
        keep.append(sel + x * nb * nb)
    blocks = torch.cat(keep)
    ids = []
    off = torch.arange(block, device=device)
    lx, ly, lz = torch.meshgrid(off, off, off, indexing="ij")
    lx, ly, lz = lx.reshape(-1), ly.reshape(-1), lz.reshape(-1)
    for chunk in blocks.split(4096):
        bx, r = chunk // (nb * nb), chunk % (nb * nb)
        by, bz = r // nb, r % nb
        cx = (bx[:, None] * block + lx[None]).reshape(-1)
        cy = (by[:, None] * block + ly[None]).reshape(-1)
        cz = (bz[:, None] * block + lz[None]).reshape(-1)
        ok = (cx < nc) & (cy < nc) & (cz < nc)
        cx, cy, cz = cx[ok], cy[ok], cz[ok]
        ctr = torch.stack([-1.0 + (cx + 0.5) * h, -1.0 + (cy + 0.5) * h, -1.0 + (cz + 0.5) * h], -1).float()
        sel = sdf(ctr).abs() < band_h * h
        ids.append(((cx * nc + cy) * nc + cz)[sel])
    return torch.sort(torch.cat(ids)).values
