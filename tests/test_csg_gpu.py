"""SDF programs on the GPU: the program's stored grid against the numpy
restatement of the reference's SDF classes (bit for bit without smooth
unions), analytic normals against the restatement's float64 gradients, and
the fused extractions (marching cubes, dual contouring, ragged batches)
against the same extractions on the stored grid, element for element."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import sdf_eval as O
from tests.conftest import grid_points

pytestmark = pytest.mark.gpu


def trees():
    from isoext_amd import sdf as S
    from isoext_amd import workloads
    torus = S.RotationOp(S.TorusSDF(0.55, 0.15), [1.0, 0.3, 0.0], 35.0)
    return {
        "csg_box_cylinder": workloads.csg_box_cylinder_sdf(16, 0),
        "torus_box": S.UnionOp([torus, S.TranslationOp(S.CuboidSDF([0.5, 0.3, 0.8]), [0.2, -0.4, 0.1])]),
        "carved": S.IntersectionOp([S.SphereSDF(0.8), S.NegationOp(S.UnionOp([
            S.CylinderSDF(0.2), S.RotationOp(S.CylinderSDF(0.2), [1, 0, 0], 90.0),
            S.RotationOp(S.CylinderSDF(0.2), [0, 1, 0], 90.0)]))]),
        "smooth": S.SmoothUnionOp([S.TranslationOp(S.SphereSDF(0.35), [-0.3, 0, 0]),
                                   S.TranslationOp(S.SphereSDF(0.3), [0.3, 0.1, 0]),
                                   S.TorusSDF(0.5, 0.08)], 0.08),
    }


def program(tree):
    from isoext_amd.csg import SdfProgram
    return SdfProgram(tree, "cuda")


@pytest.mark.parametrize("name", ["csg_box_cylinder", "torus_box", "carved", "smooth"])
def test_program_grid_matches_restatement(gpu, name):
    tree = trees()[name]
    shape = (33, 40, 70)
    amin, amax = (-1.1, -0.9, -1.0), (0.9, 1.2, 1.05)
    got = program(tree).grid(shape, amin, amax).cpu().numpy()
    ref = O.eval_tree(tree, grid_points(shape, amin, amax))
    if name == "smooth":   # exp / log of the device library vs numpy: a few ulps
        np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-6)
    else:
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("name", ["csg_box_cylinder", "torus_box", "carved", "smooth"])
def test_program_points_and_normals(gpu, name):
    tree = trees()[name]
    p = np.random.default_rng(1).uniform(-1, 1, (20000, 3)).astype(np.float32)
    vals, nrm = program(tree).points(torch.from_numpy(p).cuda(), normals=True)
    vals, nrm = vals.cpu().numpy(), nrm.cpu().numpy()
    ref = O.eval_tree(tree, p)
    if name == "smooth":
        np.testing.assert_allclose(vals, ref, rtol=2e-6, atol=2e-6)
    else:
        np.testing.assert_array_equal(vals, ref)
    _, g = O.eval_tree_grad(tree, p)
    n_ref = g / np.maximum(np.linalg.norm(g, axis=-1, keepdims=True), 1e-12)
    ok = np.all(np.abs(nrm - n_ref) < 1e-5, axis=-1)
    assert ok.mean() > 0.999, ok.mean()
    # against torch autograd's get_sdf_normal on the reference-style classes
    from isoext_amd import sdf as S
    n_t = S.get_sdf_normal(tree, torch.from_numpy(p.copy())).numpy()
    ok = np.all(np.abs(nrm - n_t) < 1e-4, axis=-1)
    assert ok.mean() > 0.999, ok.mean()


CASES = [
    ("csg_box_cylinder", (96, 96, 96), 0.0, "nagae"),
    ("csg_box_cylinder", (65, 70, 130), 0.03, "lorensen"),
    ("torus_box", (80, 81, 200), 0.0, "nagae"),
    ("carved", (64, 64, 64), -0.02, "nagae"),
    ("smooth", (50, 60, 70), 0.0, "nagae"),
    ("csg_box_cylinder", (256, 256, 256), 0.0, "nagae"),
]


@pytest.mark.parametrize("name,shape,level,method", CASES)
def test_mc_sdf_equals_stored_grid(gpu, name, shape, level, method):
    from isoext_amd import marching_cubes
    from isoext_amd.fused import marching_cubes_sdf
    from isoext_amd.grid import UniformGrid
    prog = program(trees()[name])
    print(f"mc_sdf {name} {shape} level {level} {method}", flush=True)
    v_ref, f_ref = marching_cubes(UniformGrid.wrap(prog.grid(shape)), level, method)
    v, f = marching_cubes_sdf(prog, shape, level, method)
    assert f is not None and f.shape[0] > 0
    assert torch.equal(v, v_ref) and torch.equal(f, f_ref)


def test_mc_sdf_sphere_union_equals_sphere_path(gpu):
    """A 64-sphere union compiled as a program: the sphere generator's
    formula, so the mesh equals the fused sphere path's."""
    from isoext_amd import sdf as S
    from isoext_amd.fused import marching_cubes_sdf, marching_cubes_spheres
    from isoext_amd.workloads import sphere_params
    c, r = sphere_params(64, 0)
    tree = S.UnionOp([S.TranslationOp(S.SphereSDF(float(r[k])), [float(x) for x in c[k]]) for k in range(64)])
    shape = (200, 200, 200)
    v1, f1 = marching_cubes_sdf(tree, shape)
    v2, f2 = marching_cubes_spheres(shape, c, r)
    assert torch.equal(v1, v2) and torch.equal(f1, f2)


@pytest.mark.parametrize("name,shape", [("csg_box_cylinder", (96, 96, 96)), ("torus_box", (70, 64, 90)),
                                        ("smooth", (48, 56, 64))])
def test_dc_sdf_trilinear_equals_stored_grid(gpu, name, shape):
    from isoext_amd import dual_contouring
    from isoext_amd.fused import dual_contouring_sdf
    from isoext_amd.grid import UniformGrid
    prog = program(trees()[name])
    print(f"dc_sdf trilinear {name} {shape}", flush=True)
    v_ref, f_ref = dual_contouring(UniformGrid.wrap(prog.grid(shape)), 0.0)
    v, f = dual_contouring_sdf(prog, shape, normals="trilinear")
    assert torch.equal(v, v_ref) and torch.equal(f, f_ref)


@pytest.mark.parametrize("name,shape", [("csg_box_cylinder", (96, 96, 96)), ("carved", (64, 72, 80)),
                                        ("csg_box_cylinder", (256, 256, 256))])
def test_dc_sdf_analytic_equals_intersection_path(gpu, name, shape):
    """Fused DC with analytic normals equals the reference's recipe on the
    stored grid: get_intersection, set_normals(get_sdf_normal(points)),
    dual_contouring -- with the program's normals."""
    from isoext_amd import dual_contouring, get_intersection
    from isoext_amd.fused import dual_contouring_sdf
    from isoext_amd.grid import UniformGrid
    prog = program(trees()[name])
    print(f"dc_sdf analytic {name} {shape}", flush=True)
    grid = UniformGrid.wrap(prog.grid(shape))
    its = get_intersection(grid, 0.0)
    _, n = prog.points(its.get_points(), normals=True)
    its.set_normals(n)
    v_ref, f_ref = dual_contouring(grid, 0.0, its)
    v, f = dual_contouring_sdf(prog, shape, normals="analytic")
    assert torch.equal(v, v_ref) and torch.equal(f, f_ref)


def test_batches_mix_grids_and_programs(gpu):
    from isoext_amd.batch import GridJob, dc_batch, mc_batch
    from isoext_amd.workloads import multi_sphere
    t = trees()
    progs = {k: program(v) for k, v in t.items()}
    jobs = [GridJob(multi_sphere((40, 50, 60), 8, 1)),
            GridJob(None, 0.0, 0, sdf=progs["csg_box_cylinder"], shape=(64, 64, 80)),
            GridJob(multi_sphere((70, 30, 90), 16, 2), 0.02, 1),
            GridJob(None, 0.01, 1, sdf=progs["torus_box"], shape=(50, 90, 33)),
            GridJob(None, 0.0, 0, sdf=progs["smooth"], shape=(40, 40, 40))]
    got = mc_batch(jobs)
    for j, (v, f) in zip(jobs, got):
        vals = j.values if j.sdf is None else j.sdf.grid(j.shape)
        ref = mc_batch([GridJob(vals, j.level, j.method)])[0]
        assert torch.equal(v, ref[0]) and torch.equal(f, ref[1])
    for j in jobs:
        j.method = 0
        j.sdf_normals = j.sdf is not None
    got = dc_batch(jobs)
    for j, (v, f) in zip(jobs, got):
        alone = dc_batch([GridJob(j.values, j.level, 0, sdf=j.sdf, shape=j.shape, sdf_normals=j.sdf_normals)])[0]
        assert torch.equal(v, alone[0]) and torch.equal(f, alone[1])


@pytest.mark.parametrize("shape,cfg", [((2, 2, 1200), 4), ((3, 5, 700), 4), ((9, 2, 300), 0)])
def test_mc_sdf_thin_grids(gpu, shape, cfg):
    """Grids too thin for the undecided-block list (fewer than 8 point rows
    along x or y with deep emit segments): the second classify phase walks
    every block and repeats the centre test."""
    from isoext_amd import sdf as S
    from isoext_amd.mc import mc_dense_raw
    tree = S.UnionOp([S.TorusSDF(0.5, 0.2), S.TranslationOp(S.SphereSDF(0.3), [0.0, 0.0, 0.5])])
    prog = program(tree)
    amin, amax = (-0.05, -0.05, -1.0), (0.05, 0.05, 1.0)
    v_ref, f_ref, _ = mc_dense_raw(prog.grid(shape, amin, amax), shape, amin, amax, 0.0, 0, cfg)
    v, f, _ = mc_dense_raw(None, shape, amin, amax, 0.0, 0, cfg, program=prog)
    assert f_ref.shape[0] > 0
    assert torch.equal(v, v_ref) and torch.equal(f, f_ref)


def test_sdf_single_primitives_and_empty(gpu):
    from isoext_amd import marching_cubes
    from isoext_amd import sdf as S
    from isoext_amd.fused import dual_contouring_sdf, marching_cubes_sdf
    from isoext_amd.grid import UniformGrid
    for tree in (S.SphereSDF(0.6), S.TorusSDF(0.5, 0.2), S.CuboidSDF([1.0, 0.6, 1.4]), S.CylinderSDF(0.4),
                 S.NegationOp(S.SphereSDF(0.6))):
        prog = program(tree)
        shape = (40, 44, 48)
        v_ref, f_ref = marching_cubes(UniformGrid.wrap(prog.grid(shape)), 0.0)
        v, f = marching_cubes_sdf(prog, shape)
        assert torch.equal(v, v_ref) and torch.equal(f, f_ref), type(tree).__name__
    # no crossing: every block is decided in the first phase
    far = program(S.TranslationOp(S.SphereSDF(0.1), [5.0, 5.0, 5.0]))
    assert marching_cubes_sdf(far, (64, 64, 64)) == (None, None)
    assert dual_contouring_sdf(far, (64, 64, 64)) == (None, None)
    inside_all = program(S.SphereSDF(10.0))
    assert marching_cubes_sdf(inside_all, (64, 64, 64), level=0.0) == (None, None)


@pytest.mark.parametrize("waves", ["one item", "several items"])
def test_batches_with_thin_program_grids(gpu, waves):
    """Program items too thin for the undecided-block list (every block is
    tested again in the second classify phase) batched alone and beside
    listed items: each mesh equals the stored-grid extraction."""
    from isoext_amd import sdf as S
    from isoext_amd.batch import GridJob, mc_batch
    tree = S.UnionOp([S.TorusSDF(0.5, 0.2), S.TranslationOp(S.SphereSDF(0.3), [0.0, 0.0, 0.5])])
    prog = program(tree)
    thin = [(2, 2, 1200), (3, 5, 700), (9, 2, 300)]
    box = dict(aabb_min=(-0.05, -0.05, -1.0), aabb_max=(0.05, 0.05, 1.0))   # along the z axis: crossings
    jobs = [GridJob(None, 0.0, 0, sdf=prog, shape=s, **box) for s in thin]
    if waves == "several items":
        jobs += [GridJob(None, 0.0, 0, sdf=program(trees()["torus_box"]), shape=(48, 40, 64)),
                 GridJob(None, 0.01, 1, sdf=program(trees()["carved"]), shape=(40, 44, 36))]
    else:
        jobs = jobs[:1]
    got = mc_batch(jobs)
    for j, (v, f) in zip(jobs, got):
        vals = j.sdf.grid(j.shape, j.aabb_min, j.aabb_max)
        ref = mc_batch([GridJob(vals, j.level, j.method, aabb_min=j.aabb_min, aabb_max=j.aabb_max)])[0]
        assert f is not None and f.shape[0] > 0, j.shape
        assert torch.equal(v, ref[0]) and torch.equal(f, ref[1]), j.shape
