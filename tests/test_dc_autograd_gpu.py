"""Backward of dual contouring vs torch autodiff of an FP64 restatement of
the regularised QEF on the same intersections (normals as inputs), plus
FP64 central differences on a few grid values."""

import numpy as np
import pytest
import torch

from conftest import multi_sphere_field

pytestmark = pytest.mark.gpu

EA = np.array([0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 3, 2])
EB = np.array([1, 3, 3, 2, 5, 7, 7, 6, 4, 5, 7, 6])


def coords(i, n):
    return -1.0 + (i / (n - 1)) * 2.0


def reference(vals64, normals64, its, shape, vcell, reg, level=0.0):
    """FP64 dual vertices x[vcell] as a differentiable function of the grid
    values and the normals."""
    X, Y, Z = shape
    cd = its.cell_dense.cpu().numpy()
    off = its.cell_offsets.cpu().numpy()
    ent = its.entry_cell.cpu().numpy().astype(np.int64)
    el = its.edge_local.cpu().numpy().astype(np.int64)
    c = cd[ent]
    z, r = c % (Z - 1), c // (Z - 1)
    y, x = r % (Y - 1), r // (Y - 1)

    def pid(k):
        return ((x + (k >> 2)) * Y + (y + ((k >> 1) & 1))) * Z + (z + (k & 1))

    def pos(k):
        return np.stack([coords(x + (k >> 2), X), coords(y + ((k >> 1) & 1), Y), coords(z + (k & 1), Z)], 1)
    ka, kb = EA[el], EB[el]
    ia, ib = torch.from_numpy(pid(ka)), torch.from_numpy(pid(kb))
    pa, pb = torch.from_numpy(pos(ka)), torch.from_numpy(pos(kb))
    va, vb = vals64[ia], vals64[ib]
    t = (level - va) / (vb - va)
    p = (1 - t)[:, None] * pa + t[:, None] * pb
    n = normals64
    nc = len(cd)
    cell = torch.from_numpy(ent)
    cnt = torch.from_numpy(np.diff(off)).double()
    M = torch.zeros((nc, 3, 3), dtype=torch.float64).index_add(0, cell, n[:, :, None] * n[:, None, :])
    M = M + reg * torch.eye(3, dtype=torch.float64)
    cm = torch.zeros((nc, 3), dtype=torch.float64).index_add(0, cell, p) / cnt[:, None]
    b = torch.zeros((nc, 3), dtype=torch.float64).index_add(0, cell, n * (n * p).sum(1, keepdim=True)) + reg * cm
    xs = torch.linalg.solve(M, b)
    zc, rc = cd % (Z - 1), cd // (Z - 1)
    yc, xc = rc % (Y - 1), rc // (Y - 1)
    lo = torch.from_numpy(np.stack([coords(xc, X), coords(yc, Y), coords(zc, Z)], 1))
    hi = torch.from_numpy(np.stack([coords(xc + 1, X), coords(yc + 1, Y), coords(zc + 1, Z)], 1))
    xs = torch.minimum(torch.maximum(xs, lo), hi)
    return xs[torch.from_numpy(vcell.astype(np.int64))]


def test_dc_backward_matches_fp64_autodiff(gpu):
    import isoext_amd as iso
    from isoext_amd.autograd import dual_contouring
    from isoext_amd.dc import compute_normals, dc_raw, get_intersection
    shape = (22, 20, 24)
    vals = multi_sphere_field(shape, 5, seed=21)
    g = iso.UniformGrid(list(shape))
    g.set_values(torch.from_numpy(vals).cuda())
    its = get_intersection(g, 0.0)
    normals = compute_normals(g, its).requires_grad_(True)
    t = torch.from_numpy(vals).cuda().requires_grad_(True)
    v, f = dual_contouring(g, t, normals, 0.0, 1e-2, 1e-6)
    _, _, info = dc_raw(g, its, normals.detach(), 1e-2, 1e-6, want_vcell=True)
    vcell = info["vcell"].cpu().numpy()
    w = torch.from_numpy(np.random.default_rng(0).standard_normal(tuple(v.shape))).float().cuda()
    (v * w).sum().backward()

    vals64 = torch.from_numpy(vals.reshape(-1).astype(np.float64)).requires_grad_(True)
    n64 = normals.detach().double().cpu().requires_grad_(True)
    xs = reference(vals64, n64, its, shape, vcell, 1e-2)
    assert np.abs(xs.detach().numpy() - v.detach().double().cpu().numpy()).max() < 1e-5
    (xs * w.double().cpu()).sum().backward()
    gv = t.grad.double().cpu().numpy().reshape(-1)
    gr = vals64.grad.numpy()
    assert np.abs(gv - gr).max() <= 1e-4 * np.abs(gr).max(), np.abs(gv - gr).max()
    gn = normals.grad.double().cpu().numpy()
    grn = n64.grad.numpy()
    assert np.abs(gn - grn).max() <= 1e-4 * np.abs(grn).max()
    # FP64 central differences on values touched by intersections
    touched = np.nonzero(gr)[0]
    rng = np.random.default_rng(1)
    for i in rng.choice(touched, 20, replace=False):
        def loss(d):
            vv = vals64.detach().clone()
            vv[i] += d
            return float((reference(vv, n64.detach(), its, shape, vcell, 1e-2) * w.double().cpu()).sum())
        fd = (loss(1e-7) - loss(-1e-7)) / 2e-7
        assert abs(fd - gv[i]) <= 1e-3 * (abs(fd) + 1e-2 * np.abs(gr).max()), (i, fd, gv[i])


def test_dc_autograd_api(gpu):
    import isoext_amd as iso
    from isoext_amd.autograd import dual_contouring
    shape = (16, 16, 16)
    vals = multi_sphere_field(shape, 2, seed=4)
    g = iso.UniformGrid(list(shape))
    g.set_values(torch.from_numpy(vals).cuda())
    t = g.values.clone().requires_grad_(True)
    v, f = dual_contouring(g, t)
    v.sum().backward()
    assert t.grad is not None and torch.isfinite(t.grad).all() and t.grad.abs().sum() > 0
    v2, f2 = iso.dual_contouring(g)
    assert torch.equal(v.detach(), v2) and torch.equal(f, f2)
