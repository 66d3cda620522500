"""Batched dual contouring (no intersection set stored: each cell's edges are
recomputed from its corner values): every grid's mesh in a batch equals dual
contouring of that grid alone through a stored intersection set
(get_intersection with normals, then the DC kernels), element for element
(same vertex and face order), including grids without cells or crossings."""

import numpy as np
import pytest

from conftest import multi_sphere_field

pytestmark = pytest.mark.gpu


def alone(job):
    from isoext_amd.dc import dc_raw, get_intersection
    from isoext_amd.grid import UniformGrid
    g = UniformGrid.wrap(job.values, job.aabb_min, job.aabb_max)
    its = get_intersection(g, job.level, True)
    if len(its) == 0:
        return None, None
    v, f, _ = dc_raw(g, its, its.normals, 1e-2, 1e-6)
    return (None, None) if f.shape[0] == 0 else (v, f)


def assert_same(got, ref, tag):
    if ref[0] is None:
        assert got[0] is None and got[1] is None, tag
        return
    assert got[0] is not None, tag
    assert got[0].shape == ref[0].shape and got[1].shape == ref[1].shape, tag
    assert np.array_equal(got[0].cpu().numpy(), ref[0].cpu().numpy(), equal_nan=True), tag
    assert np.array_equal(got[1].cpu().numpy(), ref[1].cpu().numpy()), tag


def make_jobs(seed, n, lo=2, hi=72):
    import torch

    from isoext_amd.batch import GridJob
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        shape = tuple(int(s) for s in rng.integers(lo, hi, size=3))
        if i % 5 == 0:
            vals = rng.standard_normal(shape).astype(np.float32)
        else:
            vals = multi_sphere_field(shape, int(rng.integers(1, 10)), seed=int(rng.integers(1 << 30)))
        level = float(rng.choice([0.0, 0.05, -0.1]))
        amin = tuple(float(x) for x in rng.uniform(-2.0, -0.5, size=3))
        amax = tuple(float(x) for x in rng.uniform(0.5, 2.0, size=3))
        jobs.append(GridJob(torch.from_numpy(vals).cuda(), level, 0, amin, amax))
    return jobs


@pytest.mark.parametrize("seed", [3, 11])
def test_dc_batch_equals_alone(gpu, seed):
    from isoext_amd.batch import dc_batch
    jobs = make_jobs(seed, 20)
    out = dc_batch(jobs)
    assert len(out) == len(jobs)
    for k, (j, got) in enumerate(zip(jobs, out)):
        assert_same(got, alone(j), (seed, k, tuple(j.values.shape)))


def test_dc_batch_empty_and_degenerate_items(gpu):
    import torch

    from isoext_amd.batch import GridJob, dc_batch
    sphere = torch.from_numpy(multi_sphere_field((40, 33, 27), 3, seed=5)).cuda()
    jobs = [
        GridJob(torch.ones((16, 16, 16), device="cuda")),                   # no crossing
        GridJob(sphere),
        GridJob(torch.zeros((1, 9, 9), device="cuda")),                     # no cells
        GridJob(torch.from_numpy(multi_sphere_field((2, 2, 2), 1, seed=1)).cuda()),
        GridJob(sphere, level=0.02),
        GridJob(torch.ones((5, 5, 5), device="cuda")),                      # no crossing, last item
    ]
    out = dc_batch(jobs)
    for k, (j, got) in enumerate(zip(jobs, out)):
        assert_same(got, alone(j), k)
    # a batch with no crossing anywhere
    none = dc_batch([GridJob(torch.ones((8, 8, 8), device="cuda"))] * 3)
    assert all(v is None and f is None for v, f in none)


def test_dc_batch_repeated_calls_and_one_grid(gpu):
    from isoext_amd.batch import dc_batch
    jobs = make_jobs(23, 6, lo=20, hi=64)
    first = dc_batch(jobs)
    second = dc_batch(jobs[2:3])   # smaller batch after a larger one reuses the workspaces
    third = dc_batch(jobs)
    for k, j in enumerate(jobs):
        ref = alone(j)
        assert_same(first[k], ref, ("first", k))
        assert_same(third[k], ref, ("third", k))
    assert_same(second[0], alone(jobs[2]), "second")


def test_dc_grid_path_special_values(gpu):
    """dual_contouring(grid) without an intersection (no stored set) equals
    the stored-set path on grids with NaN / inf values, values exactly at the
    level, a reversed box and a one-cell-thick slab."""
    import torch

    import isoext_amd as iso
    from isoext_amd.batch import GridJob
    rng = np.random.default_rng(41)
    base = multi_sphere_field((30, 31, 29), 4, seed=3)
    odd = base.copy()
    odd[rng.random(odd.shape) < 0.01] = np.nan
    odd[rng.random(odd.shape) < 0.005] = np.inf
    snapped = np.round(base * 8.0).astype(np.float32) / np.float32(8.0)   # many values exactly at 0
    cases = [
        GridJob(torch.from_numpy(odd).cuda()),
        GridJob(torch.from_numpy(snapped).cuda()),
        GridJob(torch.from_numpy(base).cuda(), 0.0, 0, (1.0, 1.0, 1.0), (-1.0, -1.0, -1.0)),
        GridJob(torch.from_numpy(multi_sphere_field((2, 40, 40), 3, seed=8)).cuda()),
    ]
    for k, j in enumerate(cases):
        g = iso.UniformGrid.wrap(j.values, j.aabb_min, j.aabb_max)
        got = iso.dual_contouring(g, level=j.level)
        assert_same(got, alone(j), k)
