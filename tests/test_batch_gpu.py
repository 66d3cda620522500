"""Ragged batched marching cubes: every grid's result in a batch equals its
result alone (bit-exact after canonical ordering) and the oracle's."""

import numpy as np
import pytest

import oracle
from conftest import multi_sphere_field

pytestmark = pytest.mark.gpu


def canon(v, f):
    return oracle.canonical(v.cpu().numpy(), f.cpu().numpy())


def alone(t, level, method):
    from isoext_amd.mc import mc_dense_raw
    v, f, _ = mc_dense_raw(t, tuple(t.shape), (-1, -1, -1), (1, 1, 1), level, method)
    return canon(v, f)


def make_jobs(seed, n):
    import torch

    from isoext_amd.batch import GridJob
    rng = np.random.default_rng(seed)
    jobs, fields = [], []
    for i in range(n):
        shape = tuple(int(s) for s in rng.integers(2, 90, size=3))
        if i % 4 == 0:
            vals = rng.standard_normal(shape).astype(np.float32)
        else:
            vals = multi_sphere_field(shape, int(rng.integers(1, 12)), seed=int(rng.integers(1 << 30)))
        level = float(rng.choice([0.0, 0.05, -0.1]))
        method = int(rng.integers(0, 2))
        fields.append(vals)
        jobs.append(GridJob(torch.from_numpy(vals).cuda(), level, method))
    return jobs, fields


@pytest.mark.parametrize("cfg", [0, 2, 3])
def test_batch_equals_alone_and_oracle(gpu, cfg):
    from isoext_amd.batch import mc_batch
    jobs, fields = make_jobs(7 + cfg, 24)
    out = mc_batch(jobs, cfg=cfg)
    assert len(out) == len(jobs)
    for j, vals, (v, f) in zip(jobs, fields, out):
        got = canon(v, f)
        ref = alone(j.values, j.level, j.method)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        vo, fo = oracle.mc_dense(vals, j.level, ["nagae", "lorensen"][j.method])
        ro = oracle.canonical(vo, fo)
        assert np.array_equal(got[0], ro[0]) and np.array_equal(got[1], ro[1])


def test_batch_degenerate_and_snapped_items(gpu):
    import torch

    from isoext_amd.batch import GridJob, mc_batch
    occ = (multi_sphere_field((40, 33, 70), 5, seed=3) < 0).astype(np.float32)   # snapped corners at level 0
    flat = np.ones((1, 20, 20), np.float32)                                          # no cells
    empty = np.ones((12, 12, 12), np.float32)                                        # no surface
    jobs = [GridJob(torch.from_numpy(x).cuda(), lvl) for x, lvl in ((occ, 0.0), (flat, 0.0), (empty, 0.0), (occ, 0.5))]
    out = mc_batch(jobs)
    assert out[1][0].shape[0] == 0 and out[2][0].shape[0] == 0
    for k in (0, 3):
        got = canon(*out[k])
        ref = alone(jobs[k].values, jobs[k].level, 0)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_batch_mixed_tile_configs(gpu):
    """Grids of very different sizes each keep their own tile configuration
    (separate emit launches per configuration) and still equal their alone
    results."""
    import torch

    from isoext_amd import _lib, workloads
    from isoext_amd.batch import GridJob, mc_batch
    shapes = [(20, 30, 40), (300, 290, 310), (40, 40, 40), (520, 520, 128), (7, 90, 65), (600, 600, 300),
              (200, 250, 256)]
    L = _lib.lib()
    cfgs = {int(L.isoext_mc_auto_config(_lib.i64x3(s))) for s in shapes}
    assert len(cfgs) >= 3
    jobs = []
    for k, s in enumerate(shapes):
        t = torch.empty(s, dtype=torch.float32, device="cuda")
        workloads.multi_sphere(s, 6, 100 + k, out=t)
        jobs.append(GridJob(t, 0.0, k % 2))
    out = mc_batch(jobs, cfg=None)   # each grid its own configuration
    for j, (v, f) in zip(jobs, out):
        got = canon(v, f)
        ref = alone(j.values, j.level, j.method)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
