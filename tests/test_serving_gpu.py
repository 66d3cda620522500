"""The continuous-batching server: every request's result equals the plain
per-call path on the same request (alone vs in the mix), under HBM
admission pressure too."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def canon(res):
    v, f = res
    if v is None:
        return None
    return oracle.canonical(v.cpu().numpy(), f.cpu().numpy())


def same(a, b):
    if a is None or b is None:
        return a is None and b is None
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def run(budget, workers, n=48, seed=4):
    import torch

    from isoext_amd import serving
    pool = serving.make_pool(10, seed=seed, side_lo=16, side_hi=96)
    torch.cuda.synchronize()
    reqs = serving.make_stream(pool, n, rate=2000.0, seed=seed, dc_frac=0.25)
    srv = serving.Server(hbm_budget=budget, max_batch=8, workers=workers)
    try:
        wall, lat = serving.drive(srv, reqs)
    finally:
        srv.close()
    assert all(r.error is None for r in reqs), [r.error for r in reqs if r.error][:3]
    for r in reqs:
        assert r.t_done >= r.t_admit >= 0
        assert same(canon(r.result), canon(serving.run_plain(r))), r.rid
    return srv, reqs, lat


def test_mix_equals_alone(gpu):
    srv, reqs, lat = run(200e9, workers=2)
    assert srv.stats["batches"] >= 1 and srv.stats["dc_requests"] >= 1
    assert srv.stats["batched_requests"] + srv.stats["dc_requests"] == len(reqs)
    assert np.all(lat >= 0)


def test_program_requests_mix_equals_alone(gpu):
    """Requests whose fields are SDF programs (never stored as grids) mixed
    with stored-grid requests: each result equals the per-call path, which
    writes the program's grid and extracts from it."""
    import torch

    from isoext_amd import serving
    grids = serving.make_pool(6, seed=5, side_lo=16, side_hi=96)
    progs = serving.make_sdf_pool(6, seed=6, side_lo=16, side_hi=120)
    torch.cuda.synchronize()
    reqs = serving.make_stream(grids + progs, 48, rate=3000.0, seed=8, dc_frac=0.25)
    assert any(r.sdf is not None for r in reqs) and any(r.sdf is None for r in reqs)
    assert any(r.sdf is not None and r.kind == "dc" for r in reqs)
    srv = serving.Server(hbm_budget=200e9, max_batch=8, workers=2)
    try:
        serving.drive(srv, reqs)
    finally:
        srv.close()
    assert all(r.error is None for r in reqs), [r.error for r in reqs if r.error][:3]
    for r in reqs:
        assert same(canon(r.result), canon(serving.run_plain(r))), (r.rid, r.kind, r.sdf is not None)


def test_admission_under_pressure(gpu):
    from isoext_amd import serving
    # a budget that holds about one large request at a time
    srv, reqs, lat = run(48e6, workers=2, n=24, seed=9)
    biggest = max(serving.footprint(r) for r in reqs)
    assert srv.stats["max_in_use"] <= max(48e6, biggest)


def test_failing_request_leaves_batch_mates_alone(gpu, monkeypatch):
    """A batch that fails (forced from the host: mc_batch raises whenever one
    marked grid is in it) is re-run request by request; only the marked
    request reports an error, every other result is bit-identical to the
    per-call path.  A request that fails validation (float64 values)
    completes at once with its error and never joins a batch."""
    import torch

    from isoext_amd import batch, serving
    pool = serving.make_pool(8, seed=7, side_lo=24, side_hi=80)
    torch.cuda.synchronize()
    marked = pool[3]
    real = batch.mc_batch

    def failing(jobs, *a, **k):
        if any(j.values is marked for j in jobs):
            raise RuntimeError("forced batch failure")
        return real(jobs, *a, **k)
    monkeypatch.setattr(batch, "mc_batch", failing)
    reqs = [serving.Request(rid=i, values=g, kind="mc", method=i % 2) for i, g in enumerate(pool)]
    reqs.append(serving.Request(rid=99, values=pool[0].double(), kind="mc"))
    srv = serving.Server(hbm_budget=200e9, max_batch=16, workers=1)
    try:
        srv.submit_many(reqs)
        for r in reqs:
            assert r.done.wait(120)
    finally:
        srv.close()
    assert reqs[-1].error and "float32" in reqs[-1].error
    for r in reqs[:-1]:
        if r.values is marked:
            assert r.error and "forced" in r.error
            continue
        assert r.error is None, r.error
        assert same(canon(r.result), canon(serving.run_plain(r))), r.rid
    assert srv.stats.get("isolated_batches", 0) >= 1
