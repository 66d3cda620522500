"""Pins the CPU oracle against the reference's published known answers,
closed forms and mesh invariants (no GPU)."""

import numpy as np
import pytest

import oracle
from conftest import grid_points, sphere_field

pytestmark = []


def _edges(f):
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    return e


def assert_closed_oriented_manifold(v, f):
    d = _edges(f)
    # every directed edge once, and its reverse once: consistent winding
    key = d[:, 0].astype(np.int64) * len(v) + d[:, 1]
    rkey = d[:, 1].astype(np.int64) * len(v) + d[:, 0]
    assert len(np.unique(key)) == len(key), "directed edge repeated"
    assert np.array_equal(np.sort(key), np.sort(rkey)), "edge without its twin"
    return len(key) // 2


# ---- known answers from the reference's notebooks (doc/*.ipynb) -------------

def test_notebook_sphere_64():
    vals = sphere_field((64, 64, 64), 0.7)
    for method in ("nagae", "lorensen"):
        v, f = oracle.mc_dense(vals, 0.0, method)
        assert (len(v), len(f)) == (9168, 18332)     # doc/marching_cubes.ipynb:60-61,113-114
    assert len(oracle.mc_dense(vals, -0.2)[0]) == 4728   # doc/marching_cubes.ipynb:144
    assert len(oracle.mc_dense(vals, 0.2)[0]) == 15072   # doc/marching_cubes.ipynb:146


def test_notebook_occupancy_64():
    occ = (sphere_field((64, 64, 64), 0.7) < 0).astype(np.float32)
    v, f = oracle.mc_dense(occ, 0.5)
    assert (len(v), len(f)) == (9168, 18332)         # doc/occupancy_grids.ipynb:88


def test_notebook_gyroid_64():
    p = grid_points((64, 64, 64)).astype(np.float32)
    x, y, z = p[..., 0] * np.float32(6), p[..., 1] * np.float32(6), p[..., 2] * np.float32(6)
    g = (np.sin(x) * np.cos(y) + np.sin(y) * np.cos(z) + np.sin(z) * np.cos(x)).astype(np.float32)
    v, f = oracle.mc_dense(g, 0.0)
    assert len(v) == 38760                            # doc/grids.ipynb:172


def test_notebook_dc_cuboid_64():
    p = grid_points((64, 64, 64))
    q = np.abs(p) - np.float32(0.5)
    qm = q.max(-1)
    cube = (np.sqrt((np.maximum(q, 0) ** 2).sum(-1)) + np.minimum(qm, 0)).astype(np.float32)
    its = oracle.intersection(cube, 0.0, compute_normals=True)
    assert len(its["points"]) == 24576               # doc/dual_contouring.ipynb:104
    r = oracle.dual_contouring(its, (64, 64, 64))
    assert (len(r["V"]), len(r["F"])) == (6146, 12288)  # doc/dual_contouring.ipynb:51-52


@pytest.mark.slow
def test_notebook_quickstart_256():
    torch = pytest.importorskip("torch")
    from isoext_amd.sdf import IntersectionOp, NegationOp, RotationOp, SphereSDF, TorusSDF, UnionOp
    P = torch.from_numpy(grid_points((256, 256, 256)))
    torus = TorusSDF(R=0.75, r=0.15)
    shape = IntersectionOp([SphereSDF(0.75), NegationOp(UnionOp([
        torus, RotationOp(torus, axis=[1, 0, 0], angle=90), RotationOp(torus, axis=[0, 1, 0], angle=90)]))])
    v, f = oracle.mc_dense(shape(P).numpy(), 0.0)
    assert (len(v), len(f)) == (196176, 392348)       # doc/quickstart.ipynb:21


# ---- closed forms ------------------------------------------------------------

def test_all_256_single_cube_cases():
    """A 2x2x2 grid with corner values -1 (inside) / +1: one vertex per crossing
    edge at its midpoint and exactly the LUT's triangle count."""
    import json
    from pathlib import Path
    luts = json.loads((Path(__file__).resolve().parents[1] / "isoext_amd/luts/mc_luts.json").read_text())
    edges = luts["edge_corners"]
    for method in ("nagae", "lorensen"):
        for c in range(256):
            vals = np.ones((2, 2, 2), np.float32)
            for k in range(8):
                if (c >> k) & 1:
                    vals[k >> 2, (k >> 1) & 1, k & 1] = -1.0
            v, f = oracle.mc_dense(vals, 0.0, method)
            row = luts[method][c]
            ntri = sum(1 for k in range(0, len(row), 3) if row[k] != -1)
            cross = [e for e, (a, b) in enumerate(edges) if ((c >> a) & 1) != ((c >> b) & 1)]
            if ntri == 0:
                assert v is None
                continue
            assert len(f) == ntri and len(v) == len(cross)
            mids = sorted(tuple(float((np.array([a >> 2, (a >> 1) & 1, a & 1]) +
                                       np.array([b >> 2, (b >> 1) & 1, b & 1]))[i] - 1.0)
                                for i in range(3)) for a, b in (edges[e] for e in cross))
            assert sorted(map(tuple, v.tolist())) == mids


def test_vertices_lie_on_grid_edges():
    vals = sphere_field((33, 33, 33), 0.55, center=(0.03, -0.02, 0.05))
    v, f = oracle.mc_dense(vals, 0.0)
    g = (v + 1.0) * 16.0          # grid units (h = 1/16)
    frac = np.abs(g - np.round(g))
    on_line = (frac < 1e-4).sum(1)
    assert np.all(on_line >= 2)   # two coordinates on grid lines -> on an edge


def test_sphere_is_closed_manifold_with_euler_2():
    vals = sphere_field((40, 40, 40), 0.6, center=(0.01, 0.02, -0.03))
    v, f = oracle.mc_dense(vals, 0.0, "nagae")
    ne = assert_closed_oriented_manifold(v, f)
    assert len(v) - ne + len(f) == 2


def test_exact_level_corner_welds_and_drops():
    """A corner exactly at the level: vertices on its edges collapse onto it
    (welded) and triangles between them are dropped."""
    vals = np.ones((3, 3, 3), np.float32)
    vals[1, 1, 1] = 0.0
    vals[0, 0, 0] = -1.0
    v, f, ndeg = oracle.mc_dense(vals, 0.0, return_degenerate=True)
    assert f is not None
    # every face has 3 distinct vertices
    assert np.all((f[:, 0] != f[:, 1]) & (f[:, 1] != f[:, 2]) & (f[:, 0] != f[:, 2]))
    # positions are unique
    assert len(np.unique(v, axis=0)) == len(v)


def test_sparse_equals_dense_on_band():
    vals = sphere_field((24, 24, 24), 0.5)
    v0, f0 = oracle.mc_dense(vals, 0.0)
    n = 23
    ids = np.arange(n ** 3)
    x, y, z = ids // (n * n), (ids // n) % n, ids % n
    cv = np.stack([vals[x + (k >> 2), y + ((k >> 1) & 1), z + (k & 1)] for k in range(8)], 1)
    inside = (cv < 0)
    act = inside.any(1) & ~inside.all(1)
    v1, f1 = oracle.mc_sparse(ids[act], cv[act], (24, 24, 24), 0.0)
    a, b = oracle.canonical(v0, f0), oracle.canonical(v1, f1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---- DC solve ---------------------------------------------------------------

def test_dc_solve_matches_lstsq_and_recovers_corner():
    """Three orthogonal planes meeting at a point inside the cell: the QEF
    minimiser is the point (regularisation 0)."""
    rng = np.random.default_rng(0)
    corner = np.array([0.031, -0.012, 0.047])
    pts, nrm = [], []
    for axis in range(3):
        for _ in range(2):
            n = np.zeros(3)
            n[axis] = 1.0
            p = corner + rng.uniform(-0.05, 0.05, 3)
            p[axis] = corner[axis]
            pts.append(p)
            nrm.append(n)
    its = {"points": np.array(pts, np.float32), "normals": np.array(nrm, np.float32),
           "edges": np.zeros((6, 2), np.int64), "is_out": np.ones(6, bool),
           "cell_indices": np.array([0]), "cell_offsets": np.array([0, 6])}
    # one cell spanning [-1, 1]^3 (2x2x2 grid)
    r = oracle.dual_contouring(its, (2, 2, 2), reg=0.0, svd_tol=1e-6)
    assert np.allclose(r["dual_v64"][0], np.array(pts, np.float32).astype(np.float64)[[0, 2, 4], [0, 1, 2]],
                       atol=1e-7)


def test_notebook_occupancy_smoothed_64():
    torch = pytest.importorskip("torch")
    from isoext_amd.utils import gaussian_smooth
    occ = torch.from_numpy((sphere_field((64, 64, 64), 0.7) < 0).astype(np.float32))
    sm = gaussian_smooth(occ, sigma=5.0).numpy()
    v, f = oracle.mc_dense(sm, 0.5)
    assert (len(v), len(f)) == (8232, 16460)         # doc/occupancy_grids.ipynb:129
