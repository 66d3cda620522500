"""Parity at the sizes the benchmark workloads are quoted on.

* MC on the 1024^3 min-of-64-spheres field (seed 0): the mesh equals the
  oracle's bit for bit after canonical ordering.  More than 2^31 value bytes
  and 3N > 2^31 edge keys: the regime the reference's 32-bit `cell_idx * 8`
  overflows (include/utils.cuh:48).
* DC with more than 1,048,576 active cells (the chunked block-sum scans),
  against the FP64 oracle, through the lookup tables, the active-cell hash
  and the grid mode.
* Sparse MC and its backward on SparseGrid([2048]^3) with cell ids above
  2^32 (the reference cannot build it: src/grid/sparse.cu:13-18).
* Serving with grid sides log-uniform in [32, 512], 80% MC / 20% DC, every
  request on its own grid: each result equals the plain per-call path.
"""

import os

import numpy as np
import pytest

import oracle
from conftest import grid_points

pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_mc_1024_64_spheres_equals_oracle(gpu):
    import torch

    from isoext_amd import workloads
    from isoext_amd.mc import mc_dense_raw
    shape = (1024, 1024, 1024)
    print("mc 1024^3: extracting on the GPU", flush=True)
    vals = workloads.multi_sphere(shape, 64, seed=0)
    v, f, _ = mc_dense_raw(vals, shape, (-1, -1, -1), (1, 1, 1), 0.0, 0)
    vg, fg = v.cpu().numpy(), f.cpu().numpy()
    host = vals.cpu().numpy()
    del vals, v, f
    torch.cuda.empty_cache()
    assert (len(vg), len(fg)) == (7056716, 14113026)
    print("mc 1024^3: oracle", flush=True)
    vo, fo = oracle.mc_dense(host, 0.0, threads=_threads())
    del host
    assert (len(vo), len(fo)) == (7056716, 14113026)
    print("mc 1024^3: canonical comparison", flush=True)
    assert _same(oracle.canonical(vo, fo), oracle.canonical(vg, fg))


def _gyroid(n, scale):
    p = grid_points((n, n, n))
    s = np.float32(scale)
    x, y, z = p[..., 0] * s, p[..., 1] * s, p[..., 2] * s
    return (np.sin(x) * np.cos(y) + np.sin(y) * np.cos(z) + np.sin(z) * np.cos(x)).astype(np.float32)


def test_dc_over_a_million_active_cells(gpu):
    """1.22 M active cells, 3.7 M intersections: every block-sum scan takes
    the chunked three-kernel path.  Tables, hash and grid mode give the same
    mesh; the dual vertices are within 1e-4 h of the FP64 oracle."""
    import torch

    import isoext_amd as iso
    from isoext_amd.dc import dc_raw
    from test_dc_gpu import compare_dc, make_grid
    vals = _gyroid(256, 12)
    its_np = oracle.intersection(vals, 0.0, compute_normals=True)
    assert len(its_np["cell_indices"]) > 1048576
    err, flips, info = compare_dc(vals, its_np)
    print(f"dc gyroid 256^3: max error {err:.2e} h, {flips} diagonal flips near ties", flush=True)
    g = make_grid(vals)
    its = iso.get_intersection(g, 0.0, compute_normals=True)
    v1, f1, _ = dc_raw(g, its, its.normals, 1e-2, 1e-6)
    its._tables = None
    v0, f0, _ = dc_raw(g, its, its.normals, 1e-2, 1e-6)
    assert torch.equal(v1, v0) and torch.equal(f1, f0)
    vg, fg = iso.dual_contouring(g, 0.0)
    assert torch.equal(vg, v1) and torch.equal(fg, f1)


def _sparse_band(n, center, radius, band_h=2.0):
    import torch

    import isoext_amd as iso
    from isoext_amd import workloads
    c = torch.tensor(center, device="cuda")

    def sdf(p):
        return (p - c).norm(dim=-1) - radius
    ids = workloads.narrow_band_ids(n, band_h, sdf)
    g = iso.SparseGrid([n, n, n])
    g.add_cells(ids)
    vals = torch.cat([sdf(q).reshape(-1, 8) for q in g.get_points().reshape(-1, 8, 3).split(1 << 21)])
    g.set_values(vals.contiguous())
    return g


def test_sparse_2048_ids_above_2p32_equals_oracle(gpu):
    import isoext_amd as iso
    n = 2048
    g = _sparse_band(n, (0.55, 0.02, -0.03), 0.25)
    ids = g.cell_ids.cpu().numpy()
    assert ids.min() > 2 ** 32, ids.min()
    print(f"sparse 2048^3: {len(ids)} cells", flush=True)
    for method in ("nagae", "lorensen"):
        v, f = iso.marching_cubes(g, 0.0, method)
        vo, fo = oracle.mc_sparse(ids, g.values.cpu().numpy(), (n, n, n), 0.0, method)
        assert _same(oracle.canonical(vo, fo), oracle.canonical(v.cpu().numpy(), f.cpu().numpy())), method


def test_sparse_2048_backward_finite_differences(gpu):
    """dL/dvalues of L = sum(w * v) against FP64 central differences (step
    1e-6) on the frozen topology: every copy of a grid point perturbed
    together, the vertices on its edges re-interpolated in FP64."""
    import torch

    from isoext_amd.autograd import marching_cubes
    from isoext_amd.sparse import mc_sparse
    n = 2048
    g = _sparse_band(n, (0.55, 0.02, -0.03), 0.25)
    vals = g.values.detach().clone().requires_grad_(True)
    v, _ = marching_cubes(g, vals, 0.0)
    w = torch.from_numpy(np.random.default_rng(3).standard_normal((v.shape[0], 3))).float().cuda()
    (v * w).sum().backward()
    grad = vals.grad.double().cpu().numpy().reshape(-1)
    _, _, src = mc_sparse(g, 0.0, 0, want_src=True)
    src = src.cpu().numpy()
    slot, e = src // 12, src % 12
    EA = np.array([0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 3, 2])
    EB = np.array([1, 3, 3, 2, 5, 7, 7, 6, 4, 5, 7, 6])
    ia, ib = 8 * slot + EA[e], 8 * slot + EB[e]
    v0 = g.values.double().cpu().numpy().reshape(-1)
    cells = g.cell_ids.cpu().numpy()
    cz, r = cells % (n - 1), cells // (n - 1)
    cy, cx = r % (n - 1), r // (n - 1)
    k = np.arange(8)
    pid = (((cx[:, None] + (k >> 2)) * n + cy[:, None] + ((k >> 1) & 1)) * n + cz[:, None] + (k & 1)).reshape(-1)

    def pos(flat):
        p = pid[flat]
        zz, rr = p % n, p // n
        return np.stack([-1.0 + (rr // n) / (n - 1) * 2.0, -1.0 + (rr % n) / (n - 1) * 2.0,
                         -1.0 + zz / (n - 1) * 2.0], 1)
    pa, pb = pos(ia), pos(ib)
    wn = w.double().cpu().numpy()
    rng = np.random.default_rng(4)
    touched = np.unique(np.concatenate([pid[ia], pid[ib]]))
    pick = rng.choice(touched, 300, replace=False)
    order = np.argsort(pid, kind="stable")
    spid = pid[order]
    scale = np.abs(grad).max()
    worst = 0.0
    for p in pick:
        copies = order[np.searchsorted(spid, p):np.searchsorted(spid, p, side="right")]
        verts = np.nonzero(np.isin(ia, copies) | np.isin(ib, copies))[0]
        in_a, in_b = np.isin(ia[verts], copies), np.isin(ib[verts], copies)

        def loss(d):
            va = v0[ia[verts]] + np.where(in_a, d, 0.0)
            vb = v0[ib[verts]] + np.where(in_b, d, 0.0)
            t = (0.0 - va) / (vb - va)
            return float((((1 - t)[:, None] * pa[verts] + t[:, None] * pb[verts]) * wn[verts]).sum())
        fd = (loss(1e-6) - loss(-1e-6)) / 2e-6
        worst = max(worst, abs(fd - grad[copies].sum()) / (abs(fd) + 1e-2 * scale))
    assert worst < 1e-3, worst


def test_serving_sides_32_to_512_equal_alone(gpu):
    """100 requests, each on its own grid (side log-uniform in [32, 512],
    1-64 spheres), 80% MC / 20% DC, submitted as a Poisson stream: every
    result equals the plain path's."""
    import torch

    from isoext_amd import serving
    pool = serving.make_pool(100, seed=21, side_lo=32, side_hi=512)
    torch.cuda.synchronize()
    rng = np.random.default_rng(22)
    t = np.cumsum(rng.exponential(1.0 / 3000.0, size=len(pool)))
    reqs = []
    for i, grid in enumerate(pool):
        kind = "dc" if rng.random() < 0.2 else "mc"
        reqs.append(serving.Request(rid=i, values=grid, kind=kind,
                                    method=int(rng.integers(0, 2)) if kind == "mc" else 0, arrival=float(t[i])))
    srv = serving.Server(hbm_budget=200e9, workers=8)
    try:
        serving.drive(srv, reqs)
    finally:
        srv.close()
    assert srv.stats["dc_requests"] > 0 and srv.stats["batched_requests"] > 0
    assert all(r.error is None for r in reqs), [r.error for r in reqs if r.error][:3]
    for r in reqs:
        a, b = r.result, serving.run_plain(r)
        if b[0] is None:
            assert a[0] is None, r.rid
            continue
        assert _same(oracle.canonical(a[0].cpu().numpy(), a[1].cpu().numpy()),
                     oracle.canonical(b[0].cpu().numpy(), b[1].cpu().numpy())), (r.rid, r.kind, tuple(r.values.shape))
