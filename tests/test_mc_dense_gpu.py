"""Dense marching cubes on the GPU vs the CPU oracle (bit-exact after
canonical ordering: same vertex set, same faces with the same corner order)."""

import numpy as np
import pytest

import oracle
from conftest import multi_sphere_field, sphere_field

pytestmark = pytest.mark.gpu


def run_gpu(vals, level=0.0, method="nagae", cfg=None, aabb_min=(-1, -1, -1), aabb_max=(1, 1, 1)):
    import torch

    from isoext_amd.mc import METHODS, mc_dense_raw
    t = torch.from_numpy(np.ascontiguousarray(vals, np.float32)).cuda()
    v, f, info = mc_dense_raw(t, vals.shape, aabb_min, aabb_max, level, METHODS[method], cfg)
    return v.cpu().numpy(), f.cpu().numpy(), info


def assert_same_mesh(vals, level=0.0, method="nagae", cfg=None, **kw):
    vo, fo = oracle.mc_dense(vals, level, method, **kw)
    vg, fg, info = run_gpu(vals, level, method, cfg, **kw)
    if vo is None:
        assert len(fg) == 0
        return info
    assert len(vg) == len(vo), f"vertex count {len(vg)} != {len(vo)}"
    assert len(fg) == len(fo), f"face count {len(fg)} != {len(fo)}"
    assert fg.min() >= 0 and fg.max() < len(vg)
    a, b = oracle.canonical(vo, fo), oracle.canonical(vg, fg)
    assert np.array_equal(a[0], b[0]), "vertex sets differ"
    assert np.array_equal(a[1], b[1]), "face sets differ"
    return info


def test_sphere_64_known_answer(gpu):
    vals = sphere_field((64, 64, 64), 0.7)
    vg, fg, _ = run_gpu(vals)
    assert (len(vg), len(fg)) == (9168, 18332)   # doc/marching_cubes.ipynb:60-61
    assert_same_mesh(vals)


@pytest.mark.parametrize("method", ["nagae", "lorensen"])
@pytest.mark.parametrize("level", [-0.2, 0.0, 0.2])
def test_sphere_config1(gpu, method, level):
    assert_same_mesh(sphere_field((64, 64, 64), 0.8), level, method)


@pytest.mark.parametrize("cfg", range(6))
def test_every_tile_config(gpu, cfg):
    vals = multi_sphere_field((70, 45, 300), 12, seed=cfg)
    assert_same_mesh(vals, 0.0, "nagae", cfg)


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 3, 3), (16, 32, 48), (8, 64, 16), (5, 7, 300),
                                   (33, 17, 65), (65, 65, 129), (40, 3, 520)])
def test_shapes(gpu, shape):
    rng = np.random.default_rng(sum(shape))
    vals = rng.standard_normal(shape).astype(np.float32)
    assert_same_mesh(vals, 0.0)
    assert_same_mesh(vals, 0.3, "lorensen")


def test_multi_sphere_256(gpu):
    vals = multi_sphere_field((256, 256, 256), 64, seed=7)
    assert_same_mesh(vals)


def test_empty(gpu):
    import torch

    import isoext_amd as iso
    g = iso.UniformGrid([8, 8, 8])
    g.set_values(torch.ones((8, 8, 8), device="cuda"))
    v, f = iso.marching_cubes(g, 0.0)
    assert v is None and f is None


def test_occupancy_half_level(gpu):
    occ = (sphere_field((64, 64, 64), 0.7) < 0).astype(np.float32)
    assert_same_mesh(occ, 0.5)


@pytest.mark.parametrize("level", [0.0, 1.0])
def test_occupancy_snapped_level(gpu, level):
    """Every vertex lands exactly on a grid corner: massive position welding
    and degenerate-triangle drops."""
    occ = (sphere_field((32, 32, 32), 0.6) < 0).astype(np.float32)
    assert_same_mesh(occ, level)


def test_quantised_field_ties(gpu):
    """Values on a coarse lattice so many corners equal the level exactly."""
    vals = np.round(sphere_field((48, 48, 48), 0.6) * 8) / 8
    for lvl in (0.0, 0.125):
        assert_same_mesh(vals.astype(np.float32), lvl)
        assert_same_mesh(vals.astype(np.float32), lvl, "lorensen")


def test_near_corner_snaps(gpu):
    """Corners within a few ulps of the level: vertices coincide after rounding."""
    vals = sphere_field((40, 40, 40), 0.5)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, vals.size, 4000)
    flat = vals.reshape(-1)
    flat[idx] = rng.choice(np.array([0.0, 1e-8, -1e-8, 3e-7, -3e-7, 1e-6], np.float32), idx.size)
    assert_same_mesh(vals)


def test_non_unit_aabb(gpu):
    vals = sphere_field((30, 40, 50), 0.5)
    assert_same_mesh(vals, 0.0, aabb_min=(-3.0, 0.5, 10.0), aabb_max=(2.0, 4.0, 11.5))


def test_deterministic(gpu):
    vals = multi_sphere_field((96, 96, 96), 20, seed=1)
    v1, f1, _ = run_gpu(vals)
    v2, f2, _ = run_gpu(vals)
    assert np.array_equal(v1, v2) and np.array_equal(f1, f2)


def test_api_drop_in(gpu):
    import torch

    import isoext
    from isoext.sdf import SphereSDF
    g = isoext.UniformGrid([32, 32, 32], aabb_min=[-1, -1, -1], aabb_max=[1, 1, 1])
    g.set_values(SphereSDF(0.5)(g.get_points()))
    v, f = isoext.marching_cubes(g, level=0.0)
    assert v.dtype == torch.float32 and f.dtype == torch.int32 and v.shape[1] == 3 and f.shape[1] == 3
    with pytest.raises(RuntimeError, match="Unknown method"):
        isoext.marching_cubes(g, method="nope")
    with pytest.raises(TypeError):
        g.set_values(torch.zeros((32, 32, 32), dtype=torch.float64, device="cuda"))
    with pytest.raises(RuntimeError):
        g.set_values(torch.zeros((32, 32, 31), device="cuda"))


@pytest.mark.parametrize("cfg", range(6))
def test_sparse_surface_repeated_calls(gpu, cfg):
    """A small sphere leaves most emit tiles without a face (only the tiles
    the count pass lists run); moving it between calls on the same shape and
    workspace must not leave tiles of the previous call listed or flagged."""
    shape = (100, 90, 129)
    for center in [(-0.7, -0.6, -0.8), (0.6, 0.5, 0.7), (0.0, -0.7, 0.75)]:
        axes = [np.linspace(-1, 1, n, dtype=np.float32) for n in shape]
        x, y, z = np.meshgrid(*axes, indexing="ij")
        vals = (np.sqrt((x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2) - 0.2).astype(np.float32)
        assert_same_mesh(vals, 0.0, "nagae", cfg)


@pytest.mark.parametrize("method", ["nagae", "lorensen"])
def test_non_finite_values(gpu, method):
    """NaN, +inf and -inf corners.  The rule the oracle states: NaN is
    outside (v - level < 0 is false), a parameter t that comes out NaN
    (inf/inf, NaN operands) is 0, so the vertex sits on the edge's first
    endpoint and welds by position like any exact-level corner."""
    rng = np.random.default_rng(17)
    vals = multi_sphere_field((48, 40, 70), 10, seed=17)
    flat = vals.reshape(-1)
    idx = rng.integers(0, flat.size, 1500)
    flat[idx] = rng.choice(np.array([np.nan, np.inf, -np.inf], np.float32), idx.size)
    assert_same_mesh(vals, 0.0, method)


def test_notebook_gyroid_64(gpu):
    from conftest import grid_points
    p = grid_points((64, 64, 64))
    s = np.float32(6)
    x, y, z = p[..., 0] * s, p[..., 1] * s, p[..., 2] * s
    g = (np.sin(x) * np.cos(y) + np.sin(y) * np.cos(z) + np.sin(z) * np.cos(x)).astype(np.float32)
    vg, fg, _ = run_gpu(g)
    assert len(vg) == 38760                            # doc/grids.ipynb:172
    assert_same_mesh(g)


def test_notebook_occupancy_smoothed_64(gpu):
    """occupancy_grids notebook: gaussian_smooth(sigma=5) of the r = 0.7
    occupancy, level 0.5 -> 8,232 V / 16,460 F (doc/occupancy_grids.ipynb:129).
    The smoothing runs in torch on the host, so the field does not depend on
    the device's convolution."""
    import torch

    from isoext_amd.utils import gaussian_smooth
    occ = torch.from_numpy((sphere_field((64, 64, 64), 0.7) < 0).astype(np.float32))
    sm = gaussian_smooth(occ, sigma=5.0).numpy()
    vg, fg, _ = run_gpu(sm, 0.5)
    assert (len(vg), len(fg)) == (8232, 16460)
    assert_same_mesh(sm, 0.5)


def test_notebook_quickstart_256(gpu):
    import torch

    import isoext_amd as iso
    from isoext_amd.sdf import IntersectionOp, NegationOp, RotationOp, SphereSDF, TorusSDF, UnionOp
    g = iso.UniformGrid([256, 256, 256])
    torus = TorusSDF(R=0.75, r=0.15)
    shape = IntersectionOp([SphereSDF(0.75), NegationOp(UnionOp([
        torus, RotationOp(torus, axis=[1, 0, 0], angle=90), RotationOp(torus, axis=[0, 1, 0], angle=90)]))])
    pts = g.get_points().cpu()
    g.set_values(shape(pts).cuda())
    v, f = iso.marching_cubes(g, 0.0)
    assert (v.shape[0], f.shape[0]) == (196176, 392348)   # doc/quickstart.ipynb:21
    vals = g.values.cpu().numpy()
    vo, fo = oracle.mc_dense(vals, 0.0)
    a, b = oracle.canonical(vo, fo), oracle.canonical(v.cpu().numpy(), f.cpu().numpy())
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert torch.is_tensor(v)
