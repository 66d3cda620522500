"""SparseGrid population, sparse marching cubes, sparse intersections and
sparse dual contouring on the GPU vs the CPU oracle."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def populate(grid, sdf, chunk=None, level=0.0):
    """The reference's documented population loop (tests/conftest.py:39-61)."""
    X, Y, Z = grid.shape
    chunks = grid.get_potential_cell_indices(chunk or X * Y * Z)
    for c in chunks:
        pts = grid.get_points_by_cell_indices(c)
        act = grid.filter_cell_indices(c, sdf(pts), level=level)
        if act is not None and act.numel() > 0:
            grid.add_cells(act)
    if grid.get_num_cells() > 0:
        grid.set_values(sdf(grid.get_points()))
    return len(chunks)


def sphere(r, c=(0.0, 0.0, 0.0)):
    import torch

    def f(p):
        return (p - torch.tensor(c, device=p.device)).norm(dim=-1) - r
    return f


def sparse_host(grid):
    return grid.cell_ids.cpu().numpy(), grid.values.cpu().numpy()


def assert_same(vo, fo, vg, fg):
    assert (vo is None) == (vg is None)
    if vo is None:
        return
    a, b = oracle.canonical(vo, fo), oracle.canonical(vg.cpu().numpy(), fg.cpu().numpy())
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("method", ["nagae", "lorensen"])
def test_sparse_mc_matches_oracle(gpu, method):
    import isoext_amd as iso
    g = iso.SparseGrid([32, 32, 32])
    populate(g, sphere(0.5, (0.02, -0.01, 0.03)))
    ids, vals = sparse_host(g)
    v, f = iso.marching_cubes(g, 0.0, method)
    vo, fo = oracle.mc_sparse(ids, vals, (32, 32, 32), 0.0, method)
    assert_same(vo, fo, v, f)


def test_sparse_mc_equals_dense(gpu):
    import isoext_amd as iso
    sdf = sphere(0.55)
    g = iso.SparseGrid([40, 40, 40])
    populate(g, sdf)
    d = iso.UniformGrid([40, 40, 40])
    d.set_values(sdf(d.get_points()))
    vs, fs = iso.marching_cubes(g)
    vd, fd = iso.marching_cubes(d)
    a = oracle.canonical(vs.cpu().numpy(), fs.cpu().numpy())
    b = oracle.canonical(vd.cpu().numpy(), fd.cpu().numpy())
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_sparse_inconsistent_values_weld_by_position(gpu):
    """Per-cell values disagree across shared corners: vertices are welded
    only where positions coincide, as in the reference."""
    import torch

    import isoext_amd as iso
    g = iso.SparseGrid([12, 12, 12])
    rng = np.random.default_rng(0)
    ids = np.unique(rng.integers(0, 11 ** 3, 600))
    g.add_cells(torch.from_numpy(ids.astype(np.int32)).cuda())
    vals = rng.standard_normal((len(ids), 8)).astype(np.float32)
    vals[::3] = np.round(vals[::3])     # exact ties at level 0 in some cells
    g.set_values(torch.from_numpy(vals).cuda())
    v, f = iso.marching_cubes(g, 0.0)
    vo, fo = oracle.mc_sparse(ids, vals, (12, 12, 12), 0.0)
    assert_same(vo, fo, v, f)


def test_sparse_intersection_and_dc(gpu):
    import isoext_amd as iso
    from isoext_amd.dc import dc_raw
    g = iso.SparseGrid([32, 32, 32])
    populate(g, sphere(0.5))
    ids, vals = sparse_host(g)
    its = iso.get_intersection(g, 0.0, compute_normals=True)
    ref = oracle.intersection(vals, 0.0, compute_normals=True, cell_ids=ids, shape=(32, 32, 32))
    assert np.array_equal(its.cell_indices.cpu().numpy(), ref["cell_indices"])
    assert np.array_equal(its.cell_offsets.cpu().numpy(), ref["cell_offsets"])
    assert np.array_equal(its.get_points().cpu().numpy(), ref["points"])
    assert np.array_equal(its.get_normals().cpu().numpy(), ref["normals"])
    v, f, info, dual = dc_raw(g, its, its.normals, 1e-2, 1e-6, want_dual=True)
    r = oracle.dual_contouring(ref, (32, 32, 32), slot_cell=ids)
    h = 2.0 / 31
    assert np.abs(dual.cpu().numpy() - r["dual_v64"]).max() <= 1e-4 * h
    assert f.shape[0] == len(r["F"])
    vv, ff = iso.dual_contouring(g, 0.0)
    assert ff.shape[0] == len(r["F"])


def test_sparse_api(gpu):
    import torch

    import isoext
    g = isoext.SparseGrid([16, 16, 16])
    assert g.get_num_cells() == 0 and g.get_points() is None
    g.add_cells(torch.arange(0, 1000, device="cuda", dtype=torch.int32))
    assert g.get_num_cells() == 1000 and g.get_num_points() == 8000
    g.add_cells(torch.arange(500, 1200, device="cuda", dtype=torch.int32))
    assert g.get_num_cells() == 1200
    assert bool((g.get_values() == torch.finfo(torch.float32).max).all())
    g.remove_cells(torch.arange(0, 100, device="cuda", dtype=torch.int32))
    assert g.get_num_cells() == 1100
    ci = g.get_cell_indices()
    assert ci.dtype == torch.int32 and int(ci[0]) == 100
    assert g.get_points().shape == (1100, 8, 3)
    assert g.get_cells().shape == (1100, 8)
    with pytest.raises(RuntimeError):
        g.set_values(torch.zeros((10, 8), device="cuda"))
    with pytest.raises(TypeError):
        g.set_values(torch.zeros((1100, 8), dtype=torch.float64, device="cuda"))
    assert len(g.get_potential_cell_indices(1000)) == 5       # ceil(16^3 / 1000)


def test_notebook_sparse_1024(gpu):
    """doc/grids.ipynb:307-309: 10,738 chunks; 2,416,778 active cells;
    2,416,776 vertices for |p| - 0.7 at [1024]^3 with chunk 100000."""
    import isoext_amd as iso
    g = iso.SparseGrid([1024, 1024, 1024])
    sdf = sphere(0.7)
    n = populate(g, sdf, chunk=100000)
    assert n == 10738
    assert g.get_num_cells() == 2416778
    v, f = iso.marching_cubes(g)
    assert v.shape[0] == 2416776
