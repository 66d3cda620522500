"""Captured marching-cubes plans (isoext_mc_plan_*): every replay of the
graph equals a plain marching_cubes call on the grid's current values, with
eager device work on the same tensor between the replays."""

from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu


def _plain(values, level, method, cfg):
    from isoext_amd.mc import mc_dense_raw
    v, f, _ = mc_dense_raw(values, values.shape, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), level, method, cfg)
    return v, f


@pytest.mark.parametrize("shape,cfg,method", [((64, 70, 90), 5, "nagae"), ((128, 128, 128), None, "lorensen"),
                                              ((48, 7, 300), 4, "nagae")])
def test_plan_replays_equal_plain_calls(gpu, shape, cfg, method):
    from isoext_amd.mc import METHODS, MarchingCubesPlan
    from isoext_amd.workloads import multi_sphere
    print(f"plan {shape} cfg {cfg} {method}", flush=True)
    values = multi_sphere(shape, 12, 3).contiguous()
    base = values.clone()
    plan = MarchingCubesPlan(values, 0.0, method, cfg=cfg)
    try:
        for it in range(5):
            # eager edits between replays: a shifted, scaled field
            values.copy_(base * (1.0 + 0.1 * it) + 0.02 * it)
            torch.cuda.synchronize()
            v, f = plan.run()
            v_ref, f_ref = _plain(values, 0.0, METHODS[method], plan.cfg)
            assert f is not None and f.shape[0] > 0
            assert torch.equal(v, v_ref) and torch.equal(f, f_ref), it
    finally:
        plan.close()


def test_plan_grows_its_buffers(gpu):
    """A field whose surface grows past the plan's buffers: the run rebuilds
    the plan and still equals the plain call."""
    from isoext_amd.mc import MarchingCubesPlan
    from isoext_amd.workloads import multi_sphere
    shape = (96, 96, 96)
    small = multi_sphere(shape, 2, 5)
    big = multi_sphere(shape, 64, 6)
    values = small.clone()
    plan = MarchingCubesPlan(values, 0.0, headroom=1.0)
    try:
        cap0 = plan.cap_f
        values.copy_(big)
        v, f = plan.run()
        v_ref, f_ref = _plain(values, 0.0, 0, plan.cfg)
        assert f_ref.shape[0] > cap0 and plan.cap_f > cap0
        assert torch.equal(v, v_ref) and torch.equal(f, f_ref)
        values.copy_(small)
        v, f = plan.run()
        v_ref, f_ref = _plain(values, 0.0, 0, plan.cfg)
        assert torch.equal(v, v_ref) and torch.equal(f, f_ref)
    finally:
        plan.close()


def test_plan_empty_surface(gpu):
    from isoext_amd.mc import MarchingCubesPlan
    values = torch.ones((40, 40, 40), device="cuda")
    plan = MarchingCubesPlan(values, 0.0)
    try:
        assert plan.run() == (None, None)
        values[10:20, 10:20, 10:20] = -1.0
        v, f = plan.run()
        v_ref, f_ref = _plain(values, 0.0, 0, plan.cfg)
        assert torch.equal(v, v_ref) and torch.equal(f, f_ref)
    finally:
        plan.close()


def test_plan_grows_its_special_lists(gpu):
    """A plan started with tiny special-vertex lists on a field whose values
    sit on a coarse lattice (many corners exactly at the level): the run
    reports the lists' need, the plan rebuilds with it, and the mesh equals
    the plain call's."""
    import numpy as np
    from isoext_amd.mc import MarchingCubesPlan
    n = 64
    ax = np.linspace(-1.0, 1.0, n, dtype=np.float32)
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    field = np.round((np.sqrt(x * x + y * y + z * z) - 0.6) * 8) / 8
    values = torch.from_numpy(field.astype(np.float32)).cuda()
    plan = MarchingCubesPlan(values, 0.0, special_cap=16)
    try:
        assert plan.special == 16
        v, f = plan.run()
        v_ref, f_ref = _plain(values, 0.0, 0, plan.cfg)
        assert plan.special > 16
        assert torch.equal(v, v_ref) and torch.equal(f, f_ref)
    finally:
        plan.close()
