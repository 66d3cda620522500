"""Backward of marching cubes vs FP64 references.

The hand-written backward kernel is checked against (a) torch autodiff of an
FP64 restatement of the forward on the same (frozen) topology, and (b) FP64
central finite differences of that forward.  For sparse grids every copy of
a perturbed corner moves together and the gradients of its copies are
summed."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

EA = np.array([0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 3, 2])
EB = np.array([1, 3, 3, 2, 5, 7, 7, 6, 4, 5, 7, 6])


def coords(i, n, a=-1.0, b=1.0):
    return a + (i / (n - 1)) * (b - a)


def dense_edges(keys, shape):
    """(value index a, value index b, position a, position b) of each vertex's edge."""
    X, Y, Z = shape
    axis = keys % 3
    p = keys // 3
    z, y, x = p % Z, (p // Z) % Y, p // (Y * Z)
    stride = np.where(axis == 0, Y * Z, np.where(axis == 1, Z, 1))
    q = p + stride
    pa = np.stack([coords(x, X), coords(y, Y), coords(z, Z)], 1)
    pb = pa.copy()
    pb[axis == 0, 0] = coords(x[axis == 0] + 1, X)
    pb[axis == 1, 1] = coords(y[axis == 1] + 1, Y)
    pb[axis == 2, 2] = coords(z[axis == 2] + 1, Z)
    return p, q, pa, pb


def forward64(vals64, ia, ib, pa, pb, level):
    va, vb = vals64[ia], vals64[ib]
    t = (level - va) / (vb - va)
    return (1 - t)[:, None] * pa + t[:, None] * pb


def canonical_weights(v, seed=0):
    order = np.lexsort((v[:, 2], v[:, 1], v[:, 0]))
    w = np.random.default_rng(seed).standard_normal((len(v), 3))
    out = np.empty_like(w)
    out[order] = w
    return out


def check(grad_gpu, vals, ia, ib, pa, pb, level, w, entries, groups=None, fd_step=1e-4):
    vals64 = torch.tensor(vals.reshape(-1), dtype=torch.float64, requires_grad=True)
    pos = forward64(vals64, torch.from_numpy(ia), torch.from_numpy(ib), torch.from_numpy(pa),
                    torch.from_numpy(pb), level)
    loss = (pos * torch.from_numpy(w)).sum()
    (g64,) = torch.autograd.grad(loss, vals64)
    g64 = g64.numpy()
    gg = grad_gpu.reshape(-1).astype(np.float64)
    scale = np.abs(g64).max() + 1e-30
    assert np.abs(gg - g64).max() <= 1e-4 * scale, "backward kernel != FP64 autodiff"
    # FP64 central differences on selected entries (all copies together)
    base = vals.reshape(-1).astype(np.float64)
    for e in entries:
        idx = groups[e] if groups is not None else [e]
        lp, lm = base.copy(), base.copy()
        lp[idx] += fd_step
        lm[idx] -= fd_step
        fp = (forward64(torch.from_numpy(lp), torch.from_numpy(ia), torch.from_numpy(ib), torch.from_numpy(pa),
                        torch.from_numpy(pb), level).numpy() * w).sum()
        fm = (forward64(torch.from_numpy(lm), torch.from_numpy(ia), torch.from_numpy(ib), torch.from_numpy(pa),
                        torch.from_numpy(pb), level).numpy() * w).sum()
        fd = (fp - fm) / (2 * fd_step)
        an = gg[idx].sum()
        assert abs(fd - an) <= 1e-3 * (abs(fd) + 1e-2 * scale), (e, fd, an)


def test_dense_backward(gpu):
    import isoext_amd as iso
    from conftest import multi_sphere_field
    from isoext_amd.autograd import marching_cubes
    from isoext_amd.mc import mc_dense_raw
    shape = (28, 30, 26)
    vals = multi_sphere_field(shape, 6, seed=11)
    g = iso.UniformGrid(list(shape))
    t = torch.from_numpy(vals).cuda().requires_grad_(True)
    v, f = marching_cubes(g, t, 0.0)
    w = canonical_weights(v.detach().cpu().numpy())
    (v * torch.from_numpy(w).float().cuda()).sum().backward()
    _, _, info = mc_dense_raw(t.detach(), shape, g.aabb_min, g.aabb_max, 0.0, 0, want_keys=True)
    keys = info["keys"].cpu().numpy()
    ia, ib, pa, pb = dense_edges(keys, shape)
    touched = np.unique(np.concatenate([ia, ib]))
    entries = np.random.default_rng(1).choice(touched, 200, replace=False)
    check(t.grad.cpu().numpy(), vals, ia, ib, pa, pb, 0.0, w, entries)


def test_sparse_backward_corner_copies(gpu):
    import isoext_amd as iso
    from isoext_amd.autograd import marching_cubes
    from isoext_amd.sparse import mc_sparse
    shape = (24, 24, 24)
    g = iso.SparseGrid(list(shape))
    ids = torch.arange(23 ** 3, device="cuda", dtype=torch.int64)
    sdf = lambda p: p.norm(dim=-1) - 0.6 + 0.03 * torch.sin(7 * p[..., 0]) * torch.cos(5 * p[..., 1])
    vals_all = sdf(g.get_points_by_cell_indices(ids))
    act = g.filter_cell_indices(ids, vals_all, 0.0)
    g.add_cells(act)
    g.set_values(sdf(g.get_points()))
    vals = g.values.detach().clone().requires_grad_(True)
    v, f = marching_cubes(g, vals, 0.0)
    w = canonical_weights(v.detach().cpu().numpy(), seed=3)
    (v * torch.from_numpy(w).float().cuda()).sum().backward()
    _, _, src = mc_sparse(g, 0.0, 0, want_src=True)
    src = src.cpu().numpy()
    slot, e = src // 12, src % 12
    cell = g.cell_ids.cpu().numpy()[slot]
    ia, ib = 8 * slot + EA[e], 8 * slot + EB[e]
    cz, r = cell % 23, cell // 23
    cy, cx = r % 23, r // 23
    def corner_pos(k):
        return np.stack([coords(cx + (k >> 2), 24), coords(cy + ((k >> 1) & 1), 24), coords(cz + (k & 1), 24)], 1)
    pa, pb = corner_pos(EA[e]), corner_pos(EB[e])
    # groups: all copies of one global grid point
    cells = g.cell_ids.cpu().numpy()
    gx, gy, gz = cells // (23 * 23), (cells // 23) % 23, cells % 23
    k = np.arange(8)
    pid = ((gx[:, None] + (k >> 2)) * 24 + gy[:, None] + ((k >> 1) & 1)) * 24 + gz[:, None] + (k & 1)
    pid = pid.reshape(-1)
    order = np.argsort(pid, kind="stable")
    uniq, start = np.unique(pid[order], return_index=True)
    groups = {}
    touched = np.unique(pid[np.concatenate([ia, ib])])
    pick = np.random.default_rng(2).choice(touched, 100, replace=False)
    for p in pick:
        j = np.searchsorted(uniq, p)
        s0 = start[j]
        s1 = start[j + 1] if j + 1 < len(start) else len(order)
        groups[p] = order[s0:s1]
    check(vals.grad.cpu().numpy(), vals.detach().cpu().numpy(), ia, ib, pa, pb, 0.0, w, list(pick), groups)
