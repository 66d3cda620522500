"""Independent pins for the oracle's trilinear normals and truncated solve.

The normal (src/its.cu:186-268) is a central difference of the cell's
trilinear interpolant at the point's cell coordinates clamped to
[0.01, 0.99].  A trilinear interpolant is affine along each axis, so the
difference quotient equals the partial derivative whatever the step: on a
linear field the normal is a/|a|; on the bilinear field (x - x0)(y - y0) it is
(y_c - y0, x_c - x0, 0) normalised, with x_c, y_c the CLAMPED coordinates,
which is what makes a wrong clamp, axis or cell size visible.  The step
(eps = 0.02) is the one constant mathematics cannot see: it cancels exactly
for an affine function and only moves the last bits.

The solve (src/batched_la.cu:151-178) is the pseudo-inverse that drops
singular values s_j <= tol * s_max: numpy.linalg.pinv(A, rcond=tol) states
the same rule, and plane configurations with reg = 0 have closed-form
minimum-norm solutions.
"""

import numpy as np
import pytest

import oracle
from conftest import grid_points

CORNERS = np.array([[k >> 2, (k >> 1) & 1, k & 1] for k in range(8)], np.float64)


def _cell(rng):
    lo = rng.uniform(-0.6, 0.6, 3)
    size = rng.uniform(0.4, 1.6, 3)          # anisotropic on purpose
    cp = (lo + CORNERS * size).astype(np.float32)
    return cp


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _points_in_cell(cp, rng, n):
    """Cell coordinates mixing interior points, near-boundary points (< 0.01
    from a face, so they clamp) and points on edges (two coordinates 0/1)."""
    lo, hi = cp[0].astype(np.float64), cp[7].astype(np.float64)
    ts = []
    for i in range(n):
        kind = i % 3
        t = rng.uniform(0.0, 1.0, 3)
        if kind == 1:
            ax = rng.integers(3)
            t[ax] = rng.choice([rng.uniform(0, 0.01), rng.uniform(0.99, 1.0)])
        elif kind == 2:
            ax = rng.permutation(3)[:2]
            t[ax] = rng.integers(0, 2, 2)
        ts.append(t)
    ts = np.array(ts)
    return (lo + ts * (hi - lo)).astype(np.float32)


def _clamped(cp, p):
    lo, hi = cp[0].astype(np.float64), cp[7].astype(np.float64)
    t = (p.astype(np.float64) - lo) / (hi - lo)
    return lo + np.clip(t, 0.01, 0.99) * (hi - lo)


def test_normal_of_linear_field_is_gradient_direction():
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(200):
        cp = _cell(rng)
        a = rng.normal(size=3)
        c = rng.uniform(-0.3, 0.3)
        cv = (cp.astype(np.float64) @ a + c).astype(np.float32)
        for p in _points_in_cell(cp, rng, 6):
            n = oracle.cell_normal(cp, cv, p)
            worst = max(worst, np.abs(n - _unit(a)).max())
    assert worst < 2e-5, worst


def test_normal_of_bilinear_field_uses_clamped_coordinates():
    rng = np.random.default_rng(12)
    worst = 0.0
    for _ in range(200):
        cp = _cell(rng)
        x0, y0 = rng.uniform(-2.5, -1.5), rng.uniform(1.5, 2.5)   # keep the gradient away from 0
        P = cp.astype(np.float64)
        cv = ((P[:, 0] - x0) * (P[:, 1] - y0)).astype(np.float32)
        for p in _points_in_cell(cp, rng, 6):
            n = oracle.cell_normal(cp, cv, p)
            pc = _clamped(cp, p)
            want = _unit([pc[1] - y0, pc[0] - x0, 0.0])
            worst = max(worst, np.abs(n - want).max())
    assert worst < 2e-5, worst


def test_bilinear_normal_distinguishes_the_clamp():
    """The same field evaluated at the unclamped edge coordinate is off by far
    more than the tolerance: the 0.01 / 0.99 clamp is observable."""
    cp = (CORNERS * np.array([1.0, 1.0, 1.0])).astype(np.float32)
    P = cp.astype(np.float64)
    x0, y0 = -2.0, 2.0
    cv = ((P[:, 0] - x0) * (P[:, 1] - y0)).astype(np.float32)
    p = np.array([0.25, 0.0, 0.0], np.float32)        # on an x-edge: ty = tz = 0
    n = oracle.cell_normal(cp, cv, p)
    clamped = _unit([0.01 - y0, 0.25 - x0, 0.0])
    unclamped = _unit([0.0 - y0, 0.25 - x0, 0.0])
    assert np.abs(n - clamped).max() < 1e-5
    assert np.abs(n - unclamped).max() > 1e-3


def test_normal_fallback_at_saddle_centre():
    """Dyadic cells, so the centre, its cell coordinates (0.5) and the corner
    values (+-q) are exact: every difference cancels and the normal is the
    (0, 0, 1) fallback."""
    rng = np.random.default_rng(13)
    for _ in range(20):
        lo = rng.integers(-8, 8, 3) / 16.0
        size = 2.0 ** rng.integers(-3, 1, 3)
        cp = (lo + CORNERS * size).astype(np.float32)
        P = cp.astype(np.float64)
        mid = (P[0] + P[7]) / 2
        cv = ((P[:, 0] - mid[0]) * (P[:, 1] - mid[1])).astype(np.float32)
        n = oracle.cell_normal(cp, cv, mid.astype(np.float32))
        assert n.tolist() == [0.0, 0.0, 1.0]


def test_grid_normals_of_bilinear_field():
    """oracle.intersection's normals on a grid: every crossing point's normal
    equals the closed form at its cell's clamped coordinates."""
    shape = (9, 7, 6)
    p = grid_points(shape).astype(np.float64)
    x0, y0 = -1.7, 1.9
    vals = ((p[..., 0] - x0) * (p[..., 1] - y0) - (0.1 - x0) * (0.2 - y0)).astype(np.float32)
    its = oracle.intersection(vals, 0.0, compute_normals=True)
    assert len(its["points"]) > 20
    pts = grid_points(shape)
    cz, cy = shape[2] - 1, shape[1] - 1
    off = its["cell_offsets"]
    for i, cid in enumerate(its["cell_indices"]):
        x, y, z = cid // (cy * cz), (cid // cz) % cy, cid % cz
        lo, hi = pts[x, y, z].astype(np.float64), pts[x + 1, y + 1, z + 1].astype(np.float64)
        for j in range(off[i], off[i + 1]):
            q = its["points"][j].astype(np.float64)
            t = np.clip((q - lo) / (hi - lo), 0.01, 0.99)
            qc = lo + t * (hi - lo)
            want = _unit([qc[1] - y0, qc[0] - x0, 0.0])
            assert np.abs(its["normals"][j] - want).max() < 2e-5


# ---- truncated pseudo-inverse -------------------------------------------------

def _sym(rng, eig):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    a = q @ np.diag(eig) @ q.T
    return (a + a.T) / 2


@pytest.mark.parametrize("eig,tol", [
    ([1.0, 0.3, 0.02], 1e-6),               # nothing truncated
    ([100.0, 3e-4, 5e-5], 1e-6),            # threshold 1e-4: drops only 5e-5 (an absolute 1e-6 would keep it)
    ([2.0, 1e-3, 1e-9], 1e-6),
    ([5.0, 1e-8, 2e-9], 1e-6),              # rank 1
    ([1.0, 0.5, 1e-3], 1e-2),               # a large tolerance drops 1e-3
    ([3.0, -2.0, 1e-9], 1e-6),              # indefinite: singular values are |eigenvalues|
])
def test_pinv_matches_numpy(eig, tol):
    rng = np.random.default_rng(int(abs(np.log10(eig[2])) * 7 + len(eig)))
    for _ in range(50):
        A = _sym(rng, np.array(eig) * rng.uniform(0.5, 2.0))
        b = rng.normal(size=3)
        x, cond = oracle.pinv_solve(A, b, tol)
        want = np.linalg.pinv(A, rcond=tol) @ b
        assert np.allclose(x, want, rtol=1e-8, atol=1e-10 * np.abs(want).max()), (x, want)
        s = np.linalg.svd(A, compute_uv=False)
        assert cond == pytest.approx(s.max() / s.min(), rel=1e-6)


def _its_dict(pts, nrm, cells, offsets):
    n = len(pts)
    return {"points": np.asarray(pts, np.float32), "normals": np.asarray(nrm, np.float32),
            "edges": np.zeros((n, 2), np.int64), "is_out": np.ones(n, bool),
            "cell_indices": np.asarray(cells, np.int64), "cell_offsets": np.asarray(offsets, np.int64)}


def test_plane_gives_origin_centred_minimum_norm_vertex():
    """reg = 0, one plane z = d: A = diag(0, 0, k) has two zero singular
    values, so x = (0, 0, mean p_z) -- the origin-centred minimum-norm
    solution -- and the clip then moves x, y to the cell box."""
    shape = (9, 9, 9)
    p = grid_points(shape)
    d = np.float32(0.1)
    vals = (p[..., 2] - d).astype(np.float32)
    its = oracle.intersection(vals, 0.0)
    its["normals"] = np.tile(np.array([0, 0, 1], np.float32), (len(its["points"]), 1))
    r = oracle.dual_contouring(its, shape, reg=0.0, svd_tol=1e-6)
    ax = p[:, 0, 0, 0]
    cz, cy = shape[2] - 1, shape[1] - 1
    for i, cid in enumerate(its["cell_indices"]):
        x, y = cid // (cy * cz), (cid // cz) % cy
        lo, hi = its["cell_offsets"][i], its["cell_offsets"][i + 1]
        want_x = np.clip(0.0, ax[x], ax[x + 1])
        want_y = np.clip(0.0, ax[y], ax[y + 1])
        want_z = its["points"][lo:hi, 2].astype(np.float64).mean()
        assert r["dual_v64"][i, 0] == want_x and r["dual_v64"][i, 1] == want_y
        assert abs(r["dual_v64"][i, 2] - want_z) < 1e-12
    assert np.all(r["cond"] == np.inf)


def test_two_planes_give_their_line_point_nearest_origin():
    rng = np.random.default_rng(5)
    a, b = 0.37, -0.21
    pts, nrm = [], []
    for _ in range(3):
        pts.append([a, rng.uniform(-0.9, 0.9), rng.uniform(-0.9, 0.9)])
        nrm.append([1, 0, 0])
        pts.append([rng.uniform(-0.9, 0.9), b, rng.uniform(-0.9, 0.9)])
        nrm.append([0, 1, 0])
    r = oracle.dual_contouring(_its_dict(pts, nrm, [0], [0, 6]), (2, 2, 2), reg=0.0, svd_tol=1e-6)
    want = np.array([np.float32(a), np.float32(b), 0.0])
    assert np.allclose(r["dual_v64"][0], want, atol=1e-12)


def test_regularised_qef_is_the_least_squares_minimiser():
    """With reg > 0 the vertex minimises sum (n.(x - p))^2 + reg |x - pbar|^2:
    numpy's lstsq of the stacked system gives the same point."""
    rng = np.random.default_rng(6)
    for _ in range(100):
        k = rng.integers(3, 8)
        pts = rng.uniform(-0.4, 0.4, (k, 3)).astype(np.float32)
        nrm = np.array([_unit(rng.normal(size=3)) for _ in range(k)], np.float32)
        reg = float(rng.choice([1e-2, 1e-1, 1.0]))
        r = oracle.dual_contouring(_its_dict(pts, nrm, [0], [0, k]), (2, 2, 2), reg=reg, svd_tol=1e-6)
        P, N = pts.astype(np.float64), nrm.astype(np.float64)
        M = np.vstack([N, np.sqrt(np.float32(reg)) * np.eye(3)])
        rhs = np.concatenate([(N * P).sum(1), np.sqrt(np.float32(reg)) * P.mean(0)])
        x = np.linalg.lstsq(M, rhs, rcond=None)[0]
        x = np.clip(x, -1.0, 1.0)
        assert np.allclose(r["dual_v64"][0], x, atol=1e-10)
