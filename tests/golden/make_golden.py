"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):

* ``mc_single_cube.npz`` -- all 256 single-cube cases x {nagae, lorensen}:
  corner values -(k+1)/16 when corner k is inside, +(k+1)/16 otherwise on a
  2x2x2 grid over [-1, 1]^3; canonical vertices / faces concatenated with
  per-case offsets.
* ``mc_random_small.npz`` -- seeded random 4^3..8^3 grids (values
  U[-1, 1], with a few exact zeros), both methods, canonical meshes.
* ``mc_snap.npz`` -- a 16^3 sphere field with values snapped to the level
  at grid points near the surface (edges that merge at corners, dropped
  degenerate triangles), both methods.
* ``mc_sphere64.npz`` -- 64^3 spheres |p| - r at level 0, canonical meshes
  of both methods: r = 0.7 (the notebook's known answer, 9,168 / 18,332) and
  r = 0.8 (the reference's CPU-runnable benchmark configuration).
* ``dc_small.npz`` -- dual contouring of seeded 6^3..10^3 sphere / box
  fields: FP64 dual vertices, quads and diagonal choice from the oracle.

The oracle is itself pinned by tests/test_oracle.py (notebook counts,
closed forms, invariants); these files freeze its outputs so that the GPU
path and the oracle are both checked against the same stored vectors.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402

METHODS = ("nagae", "lorensen")


def grid_points(shape, amin=-1.0, amax=1.0):
    axes = []
    for n in shape:
        i = np.arange(n, dtype=np.float32)
        axes.append(np.float32(amin) + (i / np.float32(n - 1)) * (np.float32(amax) - np.float32(amin)))
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1)


def single_cube_values(case):
    k = np.arange(8)
    mag = (k + 1).astype(np.float32) / np.float32(16)
    v = np.where((case >> k) & 1, -mag, mag).astype(np.float32)
    return v.reshape(2, 2, 2)   # corner k at (k>>2, (k>>1)&1, k&1)


def pack(meshes):
    """List of (v, f) -> flat arrays + offsets."""
    vo, fo, vs, fs = [0], [0], [], []
    for v, f in meshes:
        vs.append(np.asarray(v, np.float32).reshape(-1, 3))
        fs.append(np.asarray(f, np.int64).reshape(-1, 3))
        vo.append(vo[-1] + len(vs[-1]))
        fo.append(fo[-1] + len(fs[-1]))
    return (np.concatenate(vs) if vs else np.zeros((0, 3), np.float32),
            np.concatenate(fs) if fs else np.zeros((0, 3), np.int64),
            np.asarray(vo, np.int64), np.asarray(fo, np.int64))


def mesh(vals, method, level=0.0):
    v, f = oracle.mc_dense(vals, level, method)
    return oracle.canonical(v, f)


def random_small(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in rng.integers(4, 9, size=3))
    vals = rng.uniform(-1, 1, size=shape).astype(np.float32)
    vals[rng.random(shape) < 0.05] = 0.0
    return vals


def snap_field():
    p = grid_points((16, 16, 16))
    vals = (np.sqrt((p * p).sum(-1)) - np.float32(0.55)).astype(np.float32)
    near = np.abs(vals) < 0.04
    vals[near & (np.arange(vals.size).reshape(vals.shape) % 3 == 0)] = 0.0
    return vals


def sphere64(r):
    p = grid_points((64, 64, 64))
    return (np.sqrt((p * p).sum(-1)) - np.float32(r)).astype(np.float32)


def dc_fields():
    out = []
    for seed in range(4):
        rng = np.random.default_rng(100 + seed)
        n = int(rng.integers(6, 11))
        p = grid_points((n, n, n))
        c = rng.uniform(-0.2, 0.2, 3).astype(np.float32)
        if seed % 2 == 0:
            vals = np.sqrt(((p - c) ** 2).sum(-1)) - np.float32(0.6)
        else:
            q = np.abs(p - c) - np.float32(0.5)
            vals = np.sqrt((np.maximum(q, 0) ** 2).sum(-1)) + np.minimum(q.max(-1), 0)
        out.append(vals.astype(np.float32))
    return out


def main():
    cube = {}
    for method in METHODS:
        meshes = [mesh(single_cube_values(c), method) for c in range(256)]
        v, f, vo, fo = pack(meshes)
        cube[f"{method}_v"], cube[f"{method}_f"], cube[f"{method}_vo"], cube[f"{method}_fo"] = v, f, vo, fo
    np.savez_compressed(HERE / "mc_single_cube.npz", **cube)

    rnd = {}
    fields = [random_small(s) for s in range(12)]
    rnd["shapes"] = np.asarray([f.shape for f in fields], np.int64)
    rnd["values"] = np.concatenate([f.reshape(-1) for f in fields])
    for method in METHODS:
        v, f, vo, fo = pack([mesh(x, method) for x in fields])
        rnd[f"{method}_v"], rnd[f"{method}_f"], rnd[f"{method}_vo"], rnd[f"{method}_fo"] = v, f, vo, fo
    np.savez_compressed(HERE / "mc_random_small.npz", **rnd)

    sv = snap_field()
    snap = {"values": sv}
    for method in METHODS:
        snap[f"{method}_v"], snap[f"{method}_f"] = mesh(sv, method)
    np.savez_compressed(HERE / "mc_snap.npz", **snap)

    sph = {}
    for tag, r in (("r07", 0.7), ("r08", 0.8)):
        s64 = sphere64(r)
        for method in METHODS:
            sph[f"{tag}_{method}_v"], sph[f"{tag}_{method}_f"] = mesh(s64, method)
    np.savez_compressed(HERE / "mc_sphere64.npz", **sph)

    dc = {}
    for i, vals in enumerate(dc_fields()):
        its = oracle.intersection(vals, 0.0, compute_normals=True)
        ref = oracle.dual_contouring(its, vals.shape)
        dc[f"f{i}_values"] = vals
        dc[f"f{i}_dual_v64"] = ref["dual_v64"]
        dc[f"f{i}_quads"] = np.asarray(ref["quads"], np.int64)
        dc[f"f{i}_split02"] = np.asarray(ref["split02"], bool)
        dc[f"f{i}_nv"] = np.int64(len(ref["V"]))
        dc[f"f{i}_nf"] = np.int64(len(ref["F"]))
    np.savez_compressed(HERE / "dc_small.npz", **dc)
    for p in sorted(HERE.glob("*.npz")):
        print(p.name, p.stat().st_size)


if __name__ == "__main__":
    main()
