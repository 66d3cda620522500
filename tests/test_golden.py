"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every stored vector, and the stored vectors
satisfy what the mathematics and the reference's notebooks fix (single-cube
closed forms, the 64^3 sphere's notebook counts, manifoldness).  GPU: the
HIP path reproduces them (canonical order, bit-exact)."""

from pathlib import Path

import numpy as np
import pytest

import oracle

G = Path(__file__).resolve().parent / "golden"
METHODS = ("nagae", "lorensen")
EA = np.array([0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 3, 2])
EB = np.array([1, 3, 3, 2, 5, 7, 7, 6, 4, 5, 7, 6])


def load(name):
    return np.load(G / name, allow_pickle=False)


def cube_values(case):
    k = np.arange(8)
    mag = (k + 1).astype(np.float32) / np.float32(16)
    return np.where((case >> k) & 1, -mag, mag).astype(np.float32).reshape(2, 2, 2)


def split(d, method, i):
    v, f, vo, fo = d[f"{method}_v"], d[f"{method}_f"], d[f"{method}_vo"], d[f"{method}_fo"]
    return v[vo[i]:vo[i + 1]], f[fo[i]:fo[i + 1]]


def random_fields(d):
    out, o = [], 0
    for s in d["shapes"]:
        n = int(np.prod(s))
        out.append(d["values"][o:o + n].reshape(tuple(int(x) for x in s)))
        o += n
    return out


def same(a, b):
    return a[0].shape == b[0].shape and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---- CPU: the oracle against the fixtures and the fixtures against closed forms ----

@pytest.mark.parametrize("method", METHODS)
def test_oracle_single_cube(method):
    d = load("mc_single_cube.npz")
    for c in range(256):
        want = split(d, method, c)
        got = oracle.canonical(*oracle.mc_dense(cube_values(c), 0.0, method))
        assert same(got, want), c
        # closed form: one vertex per crossing edge
        bits = (c >> np.arange(8)) & 1
        assert len(want[0]) == int((bits[EA] != bits[EB]).sum()), c


@pytest.mark.parametrize("method", METHODS)
def test_oracle_random_small(method):
    d = load("mc_random_small.npz")
    for i, vals in enumerate(random_fields(d)):
        assert same(oracle.canonical(*oracle.mc_dense(vals, 0.0, method)), split(d, method, i)), i


@pytest.mark.parametrize("method", METHODS)
def test_oracle_snap(method):
    d = load("mc_snap.npz")
    got = oracle.canonical(*oracle.mc_dense(d["values"], 0.0, method))
    assert same(got, (d[f"{method}_v"], d[f"{method}_f"]))
    assert (d["values"] == 0).any()


@pytest.mark.parametrize("method", METHODS)
def test_sphere64_fixture_matches_notebook(method):
    d = load("mc_sphere64.npz")
    v, f = d[f"r07_{method}_v"], d[f"r07_{method}_f"]
    assert (len(v), len(f)) == (9168, 18332)        # doc/marching_cubes.ipynb:60-61,113-114
    for tag, r in (("r07", 0.7), ("r08", 0.8)):
        got = oracle.canonical(*oracle.mc_dense(sphere64(r), 0.0, method))
        assert same(got, (d[f"{tag}_{method}_v"], d[f"{tag}_{method}_f"]))
    # closed surface: every edge used exactly twice, in opposite directions
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    fwd = set(map(tuple, e))
    assert len(fwd) == len(e) and all((b, a) in fwd for a, b in fwd)


def test_oracle_dc_small():
    d = load("dc_small.npz")
    i = 0
    while f"f{i}_values" in d:
        vals = d[f"f{i}_values"]
        its = oracle.intersection(vals, 0.0, compute_normals=True)
        ref = oracle.dual_contouring(its, vals.shape)
        assert np.array_equal(ref["dual_v64"], d[f"f{i}_dual_v64"])
        assert np.array_equal(np.asarray(ref["quads"], np.int64), d[f"f{i}_quads"])
        assert (len(ref["V"]), len(ref["F"])) == (int(d[f"f{i}_nv"]), int(d[f"f{i}_nf"]))
        i += 1
    assert i >= 4


# ---- GPU: the HIP path against the fixtures ---------------------------------------

def gpu_mesh(vals, method, cfg=None):
    import torch

    from isoext_amd.mc import METHODS as M, mc_dense_raw
    t = torch.from_numpy(np.ascontiguousarray(vals, np.float32)).cuda()
    v, f, _ = mc_dense_raw(t, vals.shape, (-1, -1, -1), (1, 1, 1), 0.0, M[method], cfg)
    return oracle.canonical(v.cpu().numpy(), f.cpu().numpy())


@pytest.mark.gpu
@pytest.mark.parametrize("method", METHODS)
def test_gpu_single_cube(gpu, method):
    d = load("mc_single_cube.npz")
    for c in range(256):
        assert same(gpu_mesh(cube_values(c), method), split(d, method, c)), c


@pytest.mark.gpu
@pytest.mark.parametrize("method", METHODS)
def test_gpu_random_small(gpu, method):
    d = load("mc_random_small.npz")
    for i, vals in enumerate(random_fields(d)):
        assert same(gpu_mesh(vals, method), split(d, method, i)), i


@pytest.mark.gpu
@pytest.mark.parametrize("method", METHODS)
def test_gpu_snap_and_sphere64(gpu, method):
    d = load("mc_snap.npz")
    assert same(gpu_mesh(d["values"], method), (d[f"{method}_v"], d[f"{method}_f"]))
    s = load("mc_sphere64.npz")
    for tag, r in (("r07", 0.7), ("r08", 0.8)):
        assert same(gpu_mesh(sphere64(r), method), (s[f"{tag}_{method}_v"], s[f"{tag}_{method}_f"]))


@pytest.mark.gpu
def test_gpu_dc_small(gpu):
    import torch

    import isoext_amd as iso
    from isoext_amd.dc import dc_raw
    d = load("dc_small.npz")
    i = 0
    while f"f{i}_values" in d:
        vals = d[f"f{i}_values"]
        g = iso.UniformGrid(list(vals.shape))
        g.set_values(torch.from_numpy(vals).cuda())
        its = iso.get_intersection(g, 0.0, compute_normals=True)
        v, f, info, dual = dc_raw(g, its, its.normals, 1e-2, 1e-6, want_dual=True)
        h = 2.0 / (max(vals.shape) - 1)
        err = np.abs(dual.cpu().numpy().astype(np.float64) - d[f"f{i}_dual_v64"]).max() / h
        assert err <= 1e-4, (i, err)
        assert len(f) == 2 * len(d[f"f{i}_quads"])
        i += 1


def sphere64(r):
    n = 64
    i = np.arange(n, dtype=np.float32)
    a = np.float32(-1) + (i / np.float32(n - 1)) * np.float32(2)
    p = np.stack(np.meshgrid(a, a, a, indexing="ij"), -1)
    return (np.sqrt((p * p).sum(-1)) - np.float32(r)).astype(np.float32)
