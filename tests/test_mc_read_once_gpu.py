"""The read-once marching-cubes pipeline (tile configuration 6: k_stream +
k_faces) against the oracle, bit-exact after canonical ordering.

Shapes cover regions cut by the grid's end on every axis (X not a multiple
of 16, Y not of 32, Z not of 64, Z < 64, a single word), fields whose waves
stage more vertices than LDS holds (random values: the re-read path),
corners exactly at the level (weld candidates found through the flagged
rows), both methods, and the keyed variant used by the backward pass."""

import numpy as np
import pytest

import oracle
from conftest import multi_sphere_field, sphere_field

pytestmark = pytest.mark.gpu

CFG = 6


def _run(vals, level=0.0, method="nagae", keyed=False):
    import torch

    from isoext_amd.mc import METHODS, mc_dense_raw
    t = torch.from_numpy(np.ascontiguousarray(vals, np.float32)).cuda()
    v, f, info = mc_dense_raw(t, vals.shape, (-1, -1, -1), (1, 1, 1), level, METHODS[method], CFG, want_keys=keyed)
    return v.cpu().numpy(), f.cpu().numpy(), info


def _same(vals, level=0.0, method="nagae"):
    print(f"read-once {vals.shape} level {level} {method}", flush=True)
    vo, fo = oracle.mc_dense(vals, level, method)
    vg, fg, _ = _run(vals, level, method)
    if vo is None:
        assert len(fg) == 0
        return 0
    assert (len(vg), len(fg)) == (len(vo), len(fo))
    assert fg.min() >= 0 and fg.max() < len(vg)
    a, b = oracle.canonical(vo, fo), oracle.canonical(vg, fg)
    assert np.array_equal(a[0], b[0]), "vertex sets differ"
    assert np.array_equal(a[1], b[1]), "face sets differ"
    return len(vo)


@pytest.mark.parametrize("shape", [(2, 2, 2), (17, 33, 65), (16, 32, 64), (40, 70, 130), (33, 31, 7),
                                   (97, 45, 200), (130, 129, 128), (5, 300, 64)])
def test_shapes_multi_sphere(gpu, shape):
    _same(multi_sphere_field(shape, 12, seed=sum(shape)))


@pytest.mark.parametrize("method", ["nagae", "lorensen"])
def test_random_values_overfull_waves(gpu, method):
    """Independent random values: about half of all edges cross, far more
    vertices than a wave stages, so every wave re-reads its rows."""
    rng = np.random.default_rng(5)
    _same(rng.standard_normal((48, 64, 150)).astype(np.float32), 0.0, method)
    _same(rng.standard_normal((33, 40, 64)).astype(np.float32), 0.3, method)


def test_exact_level_corners(gpu):
    vals = (np.round(sphere_field((70, 66, 140), 0.6) * 8) / 8).astype(np.float32)
    _same(vals, 0.0)
    _same(vals, 0.125, "lorensen")
    occ = (sphere_field((50, 47, 90), 0.7) < 0).astype(np.float32)
    _same(occ, 1.0)
    _same(occ, 0.5)


def test_multi_sphere_384(gpu):
    assert _same(multi_sphere_field((384, 384, 384), 40, seed=2)) > 100000


def test_keys_name_each_vertex_edge(gpu):
    """The keyed variant: each vertex's key (owner point * 3 + axis) names a
    crossing edge whose lerp gives the vertex."""
    import torch

    from isoext_amd.mc import mc_dense_raw
    vals = multi_sphere_field((60, 70, 90), 10, seed=9)
    t = torch.from_numpy(vals).cuda()
    v, f, info = mc_dense_raw(t, vals.shape, (-1, -1, -1), (1, 1, 1), 0.0, 0, CFG, want_keys=True)
    keys = info["keys"].cpu().numpy()
    v = v.cpu().numpy()
    X, Y, Z = vals.shape
    owner, axis = keys // 3, keys % 3
    x, r = owner // (Y * Z), owner % (Y * Z)
    y, z = r // Z, r % Z
    step = np.stack([(axis == 0), (axis == 1), (axis == 2)], 1).astype(np.int64)
    flat = vals.reshape(-1)
    va = flat[owner]
    vb = flat[((x + step[:, 0]) * Y + (y + step[:, 1])) * Z + z + step[:, 2]]
    assert np.all((va < 0) != (vb < 0))
    assert len(np.unique(keys)) == len(keys)
    # the vertex lies on its key's edge: off-axis coordinates are the owner's
    coord = [np.float32(-1) + (np.arange(n, dtype=np.float32) / np.float32(n - 1)) * np.float32(2)
             for n in (X, Y, Z)]
    own = np.stack([coord[0][x], coord[1][y], coord[2][z]], 1)
    off = np.abs(v - own) * (1 - step)
    assert off.max() < 1e-6


def test_repeated_calls_shrinking(gpu):
    """The same workspace serves a large, then a smaller grid: nothing of the
    first call's look-back, tickets or tile flags leaks into the second."""
    a = multi_sphere_field((200, 180, 260), 30, seed=11)
    b = multi_sphere_field((90, 70, 100), 8, seed=12)
    _same(a)
    _same(b)
    _same(a)
