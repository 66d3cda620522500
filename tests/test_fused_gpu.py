"""Marching cubes of the sphere-union field evaluated inside the kernels:
the mesh equals (element for element, same order) marching cubes on the
grid that the sphere-union generator writes for the same spheres, and the
oracle's on that grid."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def grid_result(shape, c, r, level, method, amin, amax):
    from isoext_amd.mc import mc_dense_raw
    from isoext_amd.workloads import spheres_field
    vals = spheres_field(shape, c, r, aabb_min=amin, aabb_max=amax)
    v, f, _ = mc_dense_raw(vals, shape, amin, amax, level, method)
    return vals, v, f


def fused_result(shape, c, r, level, method, amin, amax):
    from isoext_amd.fused import sphere_tensor
    from isoext_amd.mc import mc_dense_raw
    v, f, _ = mc_dense_raw(None, shape, amin, amax, level, method, spheres=sphere_tensor(c, r))
    return v, f


def spheres(seed, k, lo=-0.8, hi=0.8, rmin=0.05, rmax=0.25):
    rng = np.random.default_rng(seed)
    return (rng.uniform(lo, hi, size=(k, 3)).astype(np.float32),
            rng.uniform(rmin, rmax, size=k).astype(np.float32))


CASES = [
    ((48, 40, 33), 5, 1, 0.0, 0, (-1, -1, -1), (1, 1, 1)),
    ((64, 64, 64), 64, 2, 0.0, 1, (-1, -1, -1), (1, 1, 1)),
    ((37, 90, 129), 20, 3, 0.05, 0, (-1.2, -0.9, -1.1), (0.9, 1.3, 1.0)),
    ((100, 17, 70), 200, 4, -0.02, 1, (-1, -1, -1), (1, 1, 1)),
    ((2, 2, 2), 1, 5, 0.0, 0, (-1, -1, -1), (1, 1, 1)),
    ((130, 64, 128), 256, 6, 0.0, 0, (-1, -1, -1), (1, 1, 1)),
]


@pytest.mark.parametrize("case", CASES, ids=[str(c[0]) + f"-k{c[1]}" for c in CASES])
def test_fused_equals_grid(gpu, case):
    shape, k, seed, level, method, amin, amax = case
    c, r = spheres(seed, k)
    vals, vg, fg = grid_result(shape, c, r, level, method, amin, amax)
    vf, ff = fused_result(shape, c, r, level, method, amin, amax)
    assert np.array_equal(vf.cpu().numpy(), vg.cpu().numpy()), case
    assert np.array_equal(ff.cpu().numpy(), fg.cpu().numpy()), case
    if np.prod(shape) <= 400_000 and ff.shape[0] > 0:
        vo, fo = oracle.mc_dense(vals.cpu().numpy(), level, ["nagae", "lorensen"][method], aabb_min=amin,
                                 aabb_max=amax)
        got, ref = oracle.canonical(vf.cpu().numpy(), ff.cpu().numpy()), oracle.canonical(vo, fo)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_fused_edge_fields(gpu):
    """A sphere around the whole domain (all inside), spheres outside it (all
    outside), and a sphere tangent-ish to the box: equal to the grid path."""
    import isoext_amd.fused as fused
    shape = (33, 34, 35)
    for c, r in [(np.zeros((1, 3), np.float32), np.array([3.0], np.float32)),
                 (np.array([[5.0, 5.0, 5.0], [-4.0, 0.0, 0.0]], np.float32), np.array([0.5, 1.0], np.float32)),
                 (np.array([[1.5, 0.0, 0.0]], np.float32), np.array([0.5], np.float32))]:
        vals, vg, fg = grid_result(shape, c, r, 0.0, 0, (-1, -1, -1), (1, 1, 1))
        v, f = fused.marching_cubes_spheres(shape, c, r)
        if fg.shape[0] == 0:
            assert v is None and f is None
        else:
            assert np.array_equal(v.cpu().numpy(), vg.cpu().numpy())
            assert np.array_equal(f.cpu().numpy(), fg.cpu().numpy())
    with pytest.raises(ValueError):
        fused.marching_cubes_spheres(shape, np.zeros((257, 3)), np.ones(257))
    with pytest.raises(RuntimeError):
        fused.marching_cubes_spheres(shape, np.zeros((1, 3)), np.ones(1), method="x")


def test_fused_256_multi_sphere(gpu):
    """The benchmark field at 256^3 (64 seeded spheres)."""
    from isoext_amd.workloads import sphere_params
    shape = (256, 256, 256)
    c, r = sphere_params(64, 0)
    _, vg, fg = grid_result(shape, c, r, 0.0, 0, (-1, -1, -1), (1, 1, 1))
    vf, ff = fused_result(shape, c, r, 0.0, 0, (-1, -1, -1), (1, 1, 1))
    assert vf.shape[0] > 100_000
    assert np.array_equal(vf.cpu().numpy(), vg.cpu().numpy())
    assert np.array_equal(ff.cpu().numpy(), fg.cpu().numpy())
