"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


def grid_axis(n, amin=-1.0, amax=1.0):
    """Grid coordinates with get_vtx_pos_op's float op order."""
    i = np.arange(n, dtype=np.float32)
    return np.float32(amin) + (i / np.float32(n - 1)) * (np.float32(amax) - np.float32(amin))


def grid_points(shape, amin=(-1, -1, -1), amax=(1, 1, 1)):
    ax = [grid_axis(n, a, b) for n, a, b in zip(shape, amin, amax)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return np.stack([X, Y, Z], -1).astype(np.float32)


def sphere_field(shape, r, center=(0.0, 0.0, 0.0)):
    p = grid_points(shape) - np.asarray(center, np.float32)
    return (np.sqrt((p[..., 0] * p[..., 0] + p[..., 1] * p[..., 1]) + p[..., 2] * p[..., 2])
            - np.float32(r)).astype(np.float32)


def multi_sphere_field(shape, n_spheres, seed, rmin=0.05, rmax=0.25, lo=-0.8, hi=0.8):
    """min over seeded spheres (centres U[lo,hi]^3, radii U[rmin,rmax]), fp32."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, size=(n_spheres, 3)).astype(np.float32)
    r = rng.uniform(rmin, rmax, size=n_spheres).astype(np.float32)
    p = grid_points(shape)
    out = np.full(shape, np.inf, np.float32)
    for k in range(n_spheres):
        d = p - c[k]
        v = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]) - r[k]
        out = np.minimum(out, v)
    return out.astype(np.float32)


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)
