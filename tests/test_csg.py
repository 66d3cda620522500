"""SDF programs on the CPU: the numpy restatement of the reference's SDF
classes (oracle/sdf_eval.py) pinned by closed forms, by a torch evaluation
of the same trees and by finite differences; the program compiler
(isoext_amd/csg.py) checked by interpreting its output against the tree;
program validation through the C ABI (no device work)."""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import torch

from isoext_amd import csg
from isoext_amd import sdf as S
from isoext_amd import workloads
from oracle import sdf_eval as O

RNG = np.random.default_rng(7)


def rand_pts(n=4000, lo=-1.2, hi=1.2, seed=0):
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3)).astype(np.float32)


def random_tree(rng, depth=0, smooth=False):
    """A random tree of every primitive and operator."""
    if depth >= 3 or rng.random() < 0.3:
        k = rng.integers(4)
        if k == 0:
            return S.SphereSDF(float(rng.uniform(0.1, 0.6)))
        if k == 1:
            return S.TorusSDF(float(rng.uniform(0.2, 0.6)), float(rng.uniform(0.05, 0.2)))
        if k == 2:
            return S.CuboidSDF([float(x) for x in rng.uniform(0.2, 1.2, 3)])
        return S.CylinderSDF(float(rng.uniform(0.1, 0.5)))
    k = rng.integers(6 if smooth else 5)
    kids = [random_tree(rng, depth + 1, smooth) for _ in range(int(rng.integers(2, 4)))]
    if k == 0:
        return S.UnionOp(kids)
    if k == 1:
        return S.IntersectionOp(kids)
    if k == 2:
        return S.NegationOp(kids[0])
    if k == 3:
        return S.TranslationOp(kids[0], [float(x) for x in rng.uniform(-0.4, 0.4, 3)])
    if k == 4:
        return S.RotationOp(kids[0], [float(x) for x in rng.normal(size=3)], float(rng.uniform(0, 360)))
    return S.SmoothUnionOp(kids, float(rng.uniform(0.02, 0.2)))


# ---- closed forms -------------------------------------------------------------------

def test_primitives_closed_forms():
    p = np.array([[0.6, 0.0, 0.0], [0.0, 0.0, -0.25], [0.3, 0.4, 0.0]], np.float32)
    np.testing.assert_allclose(O.eval_tree(S.SphereSDF(0.5), p), [0.1, -0.25, 0.0], atol=1e-7)
    # torus R=0.5, r=0.1: distance to the tube circle minus r
    pt = np.array([[0.5, 0.0, 0.0], [0.8, 0.0, 0.0], [0.0, 0.5, 0.3], [0.0, 0.0, 0.0]], np.float32)
    np.testing.assert_allclose(O.eval_tree(S.TorusSDF(0.5, 0.1), pt), [-0.1, 0.2, 0.2, 0.4], atol=1e-7)
    # box of full size (1, 2, 4): inside the face distance, outside the Euclidean corner distance
    pb = np.array([[0.0, 0.0, 0.0], [0.6, 0.0, 0.0], [0.8, 1.3, 2.0], [0.4, 0.9, 1.5]], np.float32)
    np.testing.assert_allclose(O.eval_tree(S.CuboidSDF([1.0, 2.0, 4.0]), pb),
                               [-0.5, 0.1, math.sqrt(0.09 + 0.09), -0.1], atol=1e-6)
    pc = np.array([[0.3, 0.4, 7.0], [0.0, 0.0, -3.0]], np.float32)
    np.testing.assert_allclose(O.eval_tree(S.CylinderSDF(0.25), pc), [0.25, -0.25], atol=1e-7)


def test_operators_closed_forms():
    a, b = S.SphereSDF(0.5), S.TranslationOp(S.SphereSDF(0.25), [1.0, 0.0, 0.0])
    p = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 0.0, 0.0]], np.float32)
    da, db = O.eval_tree(a, p), O.eval_tree(b, p)
    np.testing.assert_allclose(db, [0.75, -0.25, 0.25], atol=1e-7)
    np.testing.assert_array_equal(O.eval_tree(S.UnionOp([a, b]), p), np.minimum(da, db))
    np.testing.assert_array_equal(O.eval_tree(S.IntersectionOp([a, b]), p), np.maximum(da, db))
    np.testing.assert_array_equal(O.eval_tree(S.NegationOp(a), p), -da)
    # smooth union of two equal values d: d - k log 2
    k = 0.1
    q = np.array([[0.2, 0.0, 0.0]], np.float32)
    su = O.eval_tree(S.SmoothUnionOp([a, a], k), q)
    np.testing.assert_allclose(su, O.eval_tree(a, q) - k * math.log(2.0), rtol=1e-6)
    # a 90-degree rotation about z: p @ R maps (0, 1, 0) onto (1, 0, 0)
    r = S.RotationOp(S.TranslationOp(S.SphereSDF(0.1), [1.0, 0.0, 0.0]), [0, 0, 1], 90.0)
    np.testing.assert_allclose(O.eval_tree(r, np.array([[0.0, 1.0, 0.0]], np.float32)), [-0.1], atol=1e-6)


def test_oracle_matches_torch_within_ulps():
    """The numpy restatement against a torch evaluation of the same trees
    (torch's norm reduction rounds its sum its own way)."""
    rng = np.random.default_rng(3)
    p = rand_pts(3000)
    for i in range(25):
        t = random_tree(rng, smooth=(i % 3 == 0))
        ref = t(torch.from_numpy(p)).numpy()
        got = O.eval_tree(t, p)
        tol = 4 * np.spacing(np.maximum(np.abs(ref), 1.0).astype(np.float32)) + 2e-7
        assert np.all(np.abs(ref - got) <= tol), (i, float(np.max(np.abs(ref - got))))


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = rand_pts(500, seed=9).astype(np.float64)
    h = 1e-6
    for i in range(12):
        t = random_tree(rng, smooth=(i % 2 == 0))
        _, g = O.eval_tree_grad(t, p)
        fd = np.stack([(O.eval_tree_grad(t, p + h * e)[0] - O.eval_tree_grad(t, p - h * e)[0]) / (2 * h)
                       for e in np.eye(3)], -1)
        # kinks (min/max switches, box edges) make a few points disagree
        ok = np.all(np.abs(fd - g) < 1e-5, axis=-1)
        assert ok.mean() > 0.97, (i, ok.mean())


def test_gradient_matches_torch_autograd():
    rng = np.random.default_rng(11)
    p = rand_pts(800, seed=4)
    for i in range(10):
        t = random_tree(rng, smooth=(i % 2 == 1))
        g_ref = S.get_sdf_grad(t, torch.from_numpy(p.copy())).numpy()
        _, g = O.eval_tree_grad(t, p)
        ok = np.all(np.abs(g - g_ref) < 1e-4, axis=-1)
        assert ok.mean() > 0.99, (i, ok.mean())


# ---- the compiler -------------------------------------------------------------------

def test_program_equals_tree_bit_for_bit():
    rng = np.random.default_rng(1)
    p = rand_pts(3000, seed=2)
    for i in range(40):
        t = random_tree(rng)
        ops, prm = csg.compile_sdf(t)
        np.testing.assert_array_equal(O.eval_program(ops, prm, p), O.eval_tree(t, p), err_msg=str(i))


def test_program_smooth_union_close():
    rng = np.random.default_rng(2)
    p = rand_pts(2000, seed=3)
    for i in range(20):
        t = random_tree(rng, smooth=True)
        ops, prm = csg.compile_sdf(t)
        got, ref = O.eval_program(ops, prm, p), O.eval_tree(t, p)
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6, err_msg=str(i))


def test_workload_tree_equals_torch_workload():
    """The dual-contouring workload's tree against its torch formula."""
    t = workloads.csg_box_cylinder_sdf(16, 0)
    p = rand_pts(5000, seed=6)
    ref = workloads.csg_eval(torch.from_numpy(p)).numpy()
    got = O.eval_tree(t, p)
    assert np.max(np.abs(ref - got)) <= 4e-7
    ops, prm = csg.compile_sdf(t)
    assert ops.shape == (20, 2)   # box, cylinder, negate, max, 16 spheres min-folded
    np.testing.assert_array_equal(O.eval_program(ops, prm, p), got)


def test_compile_limits():
    deep = S.SphereSDF(0.1)
    for _ in range(9):
        deep = S.UnionOp([S.SphereSDF(0.2), S.NegationOp(deep)])
    with pytest.raises(ValueError):
        csg.compile_sdf(deep)
    nested = S.SphereSDF(0.1)
    for _ in range(4):
        nested = S.TranslationOp(S.RotationOp(nested, [0, 0, 1], 10.0), [0.1, 0, 0])
    with pytest.raises(ValueError):
        csg.compile_sdf(nested)
    many = S.UnionOp([S.TorusSDF(0.3, 0.1)] * 300)
    with pytest.raises(ValueError):
        csg.compile_sdf(many)


# ---- validation through the C ABI (host only) ------------------------------------------

def _lib():
    from pathlib import Path
    L = ctypes.CDLL(str(Path(__file__).resolve().parents[1] / "isoext_amd/libisoext_amd.so"))
    L.isoext_sdf_upload.restype = ctypes.c_int
    L.isoext_sdf_upload.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                    ctypes.c_void_p, ctypes.c_void_p]
    L.isoext_last_error.restype = ctypes.c_char_p
    L.isoext_sdf_program_bytes.restype = ctypes.c_int64
    return L


@pytest.mark.parametrize("ops,prm,msg", [
    ([[99, 0]], [], b"unknown"),
    ([[1 | 9 << 8, 0]], [0.5], b"slots"),              # value slot 9
    ([[1 | 5 << 12, 0]], [0.5], b"slots"),             # point slot 5
    ([[2, 1]], [0.1, 0.2, 0.3, 0.4], b"offset"),       # 4 params from offset 1 of 4
    ([[9 | 1 << 8 | 1 << 16, 0]], [0, 0, 0], b"slots"),   # a transform cannot combine
    ([[8 | 1 << 16, 0]], [], b"combine"),
])
def test_upload_rejects_bad_programs(ops, prm, msg):
    L = _lib()
    o = np.asarray(ops, np.int32).reshape(-1, 2)
    p = np.asarray(prm if prm else [0.0], np.float32)
    # validation fails before any device work: a dummy aligned pointer is never touched
    rc = L.isoext_sdf_upload(o.ctypes.data, o.shape[0], p.ctypes.data, len(prm), 256, None)
    assert rc == -1
    assert msg in L.isoext_last_error()


def test_program_bytes():
    L = _lib()
    assert 4096 < L.isoext_sdf_program_bytes() < 16384
