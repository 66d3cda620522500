"""Two host threads extracting different grids on the same (default) stream:
each result must equal the same extraction run alone.  A call runs in phases
with host synchronisations in between, so the threads' calls interleave on
the stream; their workspaces must not be shared."""

import threading

import numpy as np
import pytest

from conftest import multi_sphere_field

pytestmark = pytest.mark.gpu


def _mc(iso, grid):
    v, f = iso.marching_cubes(grid, 0.0)
    return v.cpu().numpy(), f.cpu().numpy()


def _dc_its(iso, grid):
    its = iso.get_intersection(grid, 0.0, compute_normals=True)
    v, f = iso.dual_contouring(grid, 0.0, its)
    return its.get_points().cpu().numpy(), v.cpu().numpy(), f.cpu().numpy()


def test_two_threads_default_stream_equal_alone(gpu):
    import torch

    import isoext_amd as iso
    shapes = [(96, 80, 130), (70, 110, 64)]
    grids = []
    for k, s in enumerate(shapes):
        g = iso.UniformGrid(list(s))
        g.set_values(torch.from_numpy(multi_sphere_field(s, 12, seed=40 + k)).cuda())
        grids.append(g)
    alone = [(_mc(iso, g), _dc_its(iso, g)) for g in grids]
    torch.cuda.synchronize()

    errors = []
    barrier = threading.Barrier(2)

    def work(k):
        try:
            barrier.wait()
            for it in range(12):
                mc = _mc(iso, grids[k])
                dc = _dc_its(iso, grids[k])
                for a, b in zip(mc + dc, alone[k][0] + alone[k][1]):
                    if not np.array_equal(a, b):
                        errors.append(f"thread {k} iteration {it}: result differs from the run alone")
                        return
        except Exception as e:          # surfaced below
            errors.append(f"thread {k}: {e!r}")

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
