"""SparseGrid population on the GPU: the sorted-key union / difference
kernels against set operations computed independently (numpy), the
reference's value reset, the keep_values option, and the cell enumeration
(src/grid/sparse.cu:71-142)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _grid(shape=(64, 64, 64)):
    import isoext_amd as iso
    return iso.SparseGrid(list(shape))


def _set(g, ids_np):
    g.cell_ids = torch.from_numpy(np.asarray(ids_np, np.int64)).cuda()
    g._reset_values()


@pytest.mark.parametrize("na,nb,hi", [(0, 0, 100), (0, 50, 100), (50, 0, 100), (1000, 3000, 5000),
                                      (200000, 500000, 1 << 40), (3_000_000, 2_000_000, 8_000_000_000)])
def test_union_and_difference(gpu, na, nb, hi):
    rng = np.random.default_rng(na + nb)
    a = np.unique(rng.integers(0, hi, na))
    b = rng.integers(0, hi, nb)
    if na and nb:
        b[: nb // 3] = rng.choice(a, nb // 3)      # overlap with the grid's cells
    g = _grid((2049, 2049, 2049))
    _set(g, a)
    g.add_cells(torch.from_numpy(b).cuda())
    assert np.array_equal(g.cell_ids.cpu().numpy(), np.union1d(a, b))
    assert g.values.shape == (len(np.union1d(a, b)), 8)
    _set(g, a)
    g.remove_cells(torch.from_numpy(b).cuda())
    assert np.array_equal(g.cell_ids.cpu().numpy(), np.setdiff1d(a, b))


def test_values_reset_and_keep(gpu):
    g = _grid()
    g.add_cells(torch.tensor([5, 1, 9], dtype=torch.int32, device="cuda"))
    assert torch.all(g.values == g.default_value)
    g.set_values(torch.arange(24, dtype=torch.float32, device="cuda").reshape(3, 8))
    g.add_cells(torch.tensor([7, 1, 7], dtype=torch.int32, device="cuda"), keep_values=True)
    ids = g.cell_ids.cpu().tolist()
    assert ids == [1, 5, 7, 9]
    v = g.values.cpu().numpy()
    old = np.arange(24, dtype=np.float32).reshape(3, 8)   # rows of cells 1, 5, 9
    assert np.array_equal(v[0], old[0]) and np.array_equal(v[1], old[1]) and np.array_equal(v[3], old[2])
    assert np.all(v[2] == g.default_value)
    g.remove_cells(torch.tensor([5], dtype=torch.int32, device="cuda"), keep_values=True)
    assert g.cell_ids.cpu().tolist() == [1, 7, 9]
    assert np.array_equal(g.values.cpu().numpy()[2], old[2])
    g.add_cells(torch.tensor([3], dtype=torch.int32, device="cuda"))   # reference behaviour: reset
    assert torch.all(g.values == g.default_value)


def test_potential_cell_indices(gpu):
    g = _grid((10, 11, 12))
    pts = torch.cat(g.get_potential_cell_indices(100))
    assert pts.numel() == 10 * 11 * 12 and torch.equal(pts.cpu(), torch.arange(10 * 11 * 12, dtype=pts.dtype))
    cells = torch.cat(g.get_potential_cell_indices(100, cells_only=True))
    assert cells.numel() == 9 * 10 * 11 and int(cells.max()) == 9 * 10 * 11 - 1


def test_population_flow_matches_dense(gpu):
    """The reference's notebook flow (enumerate, corner points, SDF, filter,
    add, set values) reproduces the dense grid's marching-cubes mesh."""
    import isoext_amd as iso
    import oracle
    n = 48
    g = iso.SparseGrid([n, n, n])
    sdf = lambda p: torch.linalg.norm(p, dim=-1) - 0.6
    for chunk in g.get_potential_cell_indices(20000, cells_only=True):
        pts = g.get_points_by_cell_indices(chunk)
        keep = g.filter_cell_indices(chunk, sdf(pts))
        if keep is not None and keep.numel():
            g.add_cells(keep)
    g.set_values(sdf(g.get_points()))
    v, f = iso.marching_cubes(g)
    d = iso.UniformGrid([n, n, n])
    d.set_values(sdf(d.get_points()))
    vd, fd = iso.marching_cubes(d)
    a = oracle.canonical(vd.cpu().numpy(), fd.cpu().numpy())
    b = oracle.canonical(v.cpu().numpy(), f.cpu().numpy())
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
