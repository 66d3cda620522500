"""Benchmark: plain marching cubes on a dense 1024^3 fp32 UniformGrid.

Workload (the reference metric's configuration): SDF = min over 64 spheres
(centres U[-0.8, 0.8]^3, radii U[0.05, 0.25], seed 0) on [-1, 1]^3, iso 0,
method "nagae".  Values are generated on the GPU before timing; a step is one
isoext_amd.marching_cubes call (kernel + output sizing + host sync).

Prints one JSON line: Gvoxels/s (grid points / step time); the roofline of
the dominant kernel, classify (it streams every value once: algorithmic
bytes 4*N_points over its mean HIP-event duration, against 8 TB/s); the
whole extraction's algorithmic bytes 4*N_points + 12*V + 12*F over its device
time; and the CPU oracle timed on a bounded slab of the same field on this
host, on all cores and on one.

    python bench.py [--steps K] [--warmup W] [--size N] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

METRIC = "MC Gvoxels/s & % HBM peak at 1024³ fp32; serving p50/p99 vs per-call plain"
HBM_PEAK_GBS = 8000.0
READ_ONCE_CFG = 6   # tile configuration served by the read-once pipeline (k_stream + k_faces)
POOL_GRIDS = 320    # distinct serving grids: more than the 8 x 32 requests in flight


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--spheres", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--method", default="nagae")
    ap.add_argument("--cfg", type=int, default=None, help="tile configuration (default: autotuned)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fused", action="store_true", help="skip the fused sphere-field extraction")
    ap.add_argument("--cpu-planes", type=int, default=257, help="x-planes in the CPU sample")
    ap.add_argument("--no-dc", action="store_true", help="skip the 256^3 dual-contouring summary")
    ap.add_argument("--no-sparse", action="store_true", help="skip the 2048^3 sparse summary")
    ap.add_argument("--serving", type=int, default=10000,
                    help="requests in the serving comparison (0 = skip)")
    ap.add_argument("--serving-sdf", type=int, default=2000,
                    help="requests in the serving comparison with SDF-program fields (0 = skip)")
    ap.add_argument("--workload", default="mc1024", choices=["mc1024", "dc256", "sparse2048", "csg"],
                    help="mc1024: plain MC at 1024^3 (default); dc256: dual contouring of the 256^3 CSG "
                         "field; sparse2048: differentiable sparse MC on the 2048^3 narrow band; csg: the CSG "
                         "field's fused (never stored) extractions against generate-then-extract")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the timed steps, run a plain 4-byte elementwise copy (torch.neg) over MIB "
                         "MiB three times, so that one counter run also sees a kernel of known traffic")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU: summarise two rocprofv3 counter_collection.csv files (FETCH_SIZE run, "
                         "WRITE_SIZE run) of a --calibrate-copy run into HBM bytes per k_classify launch")
    return ap.parse_args()


def _pmc_per_kernel(path, counter):
    """Mean KiB per dispatch of `counter`, by kernel kind (classify, count,
    emit, sp_gather, copy)."""
    import csv
    per = {}
    with open(path, newline="") as fh:
        for row in csv.DictReader(fh):
            if row.get("Counter_Name") != counter:
                continue
            name = row["Kernel_Name"]
            kind = next((k for k in ("classify", "count", "emit", "sp_gather") if f"k_{k}" in name),
                        "copy" if "neg" in name else None)
            if kind is None:
                continue
            key = (kind, row["Dispatch_Id"])
            per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    out = {}
    for (kind, _), v in per.items():
        out.setdefault(kind, []).append(v)
    return {k: (sum(v) / len(v), len(v)) for k, v in out.items()}


def pmc_summary(fetch_csv, write_csv, copy_mib, npts):
    """HBM traffic of one classify launch from FETCH_SIZE / WRITE_SIZE
    (KiB per dispatch), each scaled by the factor that maps the same counter
    on the calibration copy (reads copy_mib MiB, writes copy_mib MiB) to its
    known bytes."""
    f = _pmc_per_kernel(fetch_csv, "FETCH_SIZE")
    w = _pmc_per_kernel(write_csv, "WRITE_SIZE")
    known = copy_mib * 2 ** 20
    kf = known / (f["copy"][0] * 1024)
    kw = known / (w["copy"][0] * 1024)
    read_b = f["classify"][0] * 1024 * kf
    write_b = w["classify"][0] * 1024 * kw
    alg = 4 * npts + npts // 8
    return {"kernel": "k_classify", "grid_points": npts, "alg_bytes": alg,
            "fetch_kib_raw": f["classify"][0], "write_kib_raw": w["classify"][0],
            "dispatches": {"fetch": f["classify"][1], "write": w["classify"][1]},
            "copy_calibration": {"bytes_each_way": known, "fetch_kib_raw": f["copy"][0],
                                 "write_kib_raw": w["copy"][0], "fetch_factor": kf, "write_factor": kw},
            "read_bytes": read_b, "write_bytes": write_b, "traffic_bytes": read_b + write_b,
            "traffic_over_alg": (read_b + write_b) / alg,
            "other_kernels_bytes": {k: {"read": f[k][0] * 1024 * kf, "write": w[k][0] * 1024 * kw}
                                    for k in ("count", "emit", "sp_gather") if k in f and k in w}}


def timed_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_dc256(args):
    """Dual contouring on a 256^3 CSG field (box half-extent 0.6 minus a
    z-cylinder r 0.3, union 16 seeded spheres r in [0.05, 0.15]) with
    analytic-gradient normals.  A step = get_intersection + set_normals-fed
    dual_contouring (the analytic normals are computed once, outside the
    timed region, as a user would supply them).  Reports the per-vertex error
    against the FP64 QEF oracle in units of h and the FP64 re-solve count."""
    import numpy as np
    import torch

    import isoext_amd as iso
    import oracle
    from isoext_amd import workloads
    from isoext_amd.dc import dc_raw, get_intersection
    n = 256
    shape = (n, n, n)
    vals = workloads.csg_box_cylinder(shape).contiguous()
    g = iso.UniformGrid.wrap(vals)
    its = get_intersection(g, 0.0)
    normals = workloads.csg_normals(its.points).contiguous()
    torch.cuda.synchronize()

    def step():
        it = get_intersection(g, 0.0)
        dc_raw(g, it, normals, 1e-2, 1e-6)
    ms = timed_steps(step, args.steps, args.warmup) * 1e3
    v, f, info, dual = dc_raw(g, its, normals, 1e-2, 1e-6, want_dual=True)
    fused = csg_dc_fused(shape, args.steps, args.warmup)
    # accuracy against the FP64 oracle on the same intersections and normals
    vals_np = vals.cpu().numpy()
    its_np = oracle.intersection(vals_np, 0.0)
    assert np.array_equal(its_np["points"], its.points.cpu().numpy())
    its_np["normals"] = normals.cpu().numpy()
    t0 = time.perf_counter()
    ref = oracle.dual_contouring(its_np, shape, reg=1e-2, svd_tol=1e-6)
    t_cpu = time.perf_counter() - t0
    h = 2.0 / (n - 1)
    err = np.abs(dual.cpu().numpy().astype(np.float64) - ref["dual_v64"]).max(axis=1) / h
    npts = n ** 3
    ni = its.points.shape[0]
    nv, nf = v.shape[0], f.shape[0]
    alg = 4 * npts + 24 * ni + 12 * nv + 12 * nf
    return {
        "metric": "DC Gvoxels/s at 256^3 fp32 (CSG, analytic normals), error vs FP64 oracle",
        "value": npts / (ms * 1e-3) / 1e9, "unit": "Gvoxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "dual contouring, 256^3 CSG box-cylinder + 16 spheres, reg 1e-2, svd_tol 1e-6",
                   "active_cells": int(its.cell_indices.numel()), "intersections": int(ni), "vertices": nv,
                   "faces": nf, "fp64_resolves": info["fp64_resolves"],
                   "dual_error_h": {"max": float(err.max()), "p99": float(np.percentile(err, 99)),
                                    "mean": float(err.mean())},
                   "achieved_GBs": alg / (ms * 1e-3) / 1e9, "fused": fused},
        "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": {"value": npts / t_cpu / 1e9, "unit": "Gvoxels/s", "cores": 1, "kind": "port",
                         "sample": f"oracle FP64 QEF dual contouring on the same intersections, {t_cpu:.2f} s"},
    }


def csg_dc_fused(shape, steps, warmup):
    """Dual contouring of the CSG workload end to end, field included: fused
    (the SDF program evaluated inside the kernels, analytic normals, no grid)
    against the stored-grid recipe (program grid, intersections, program
    normals at the points, DC) and a torch evaluation of the grid."""
    import torch

    import isoext_amd as iso
    from isoext_amd import workloads
    from isoext_amd.csg import SdfProgram
    from isoext_amd.dc import dc_raw, get_intersection
    from isoext_amd.fused import dual_contouring_sdf
    prog = SdfProgram(workloads.csg_box_cylinder_sdf(16, 0))
    vals = torch.empty(shape, dtype=torch.float32, device="cuda")
    g = iso.UniformGrid.wrap(vals)

    def stored():
        prog.grid(shape, out=vals)
        it = get_intersection(g, 0.0)
        _, nrm = prog.points(it.points, normals=True)
        dc_raw(g, it, nrm, 1e-2, 1e-6)
    ms_stored = timed_steps(stored, steps, warmup) * 1e3
    ms_fused = timed_steps(lambda: dual_contouring_sdf(prog, shape), steps, warmup) * 1e3
    ms_gen = timed_steps(lambda: prog.grid(shape, out=vals), steps, warmup) * 1e3
    ms_torch = timed_steps(lambda: workloads.csg_box_cylinder(shape, program=False), 2, 1) * 1e3
    v, f = dual_contouring_sdf(prog, shape)
    return {"fused_ms": ms_fused, "stored_recipe_ms": ms_stored, "program_grid_ms": ms_gen,
            "torch_grid_ms": ms_torch, "vertices": int(v.shape[0]), "faces": int(f.shape[0]),
            "note": "fused = dual_contouring_sdf (analytic normals, no grid); stored recipe = program grid + "
                    "get_intersection + program normals + DC; identical meshes (tests/test_csg_gpu.py)"}


def csg_mc_fused(sizes=(256, 1024), steps=5, warmup=2):
    """Marching cubes of the CSG workload's field, fused (never stored)
    against generate-then-extract, per grid size."""
    import torch

    import isoext_amd as iso
    from isoext_amd import workloads
    from isoext_amd.csg import SdfProgram
    from isoext_amd.fused import marching_cubes_sdf
    from isoext_amd.mc import mc_dense_raw
    prog = SdfProgram(workloads.csg_box_cylinder_sdf(16, 0))
    out = {}
    for n in sizes:
        shape = (n, n, n)
        ms_fused = timed_steps(lambda: marching_cubes_sdf(prog, shape), steps, warmup) * 1e3
        vals = torch.empty(shape, dtype=torch.float32, device="cuda")
        ms_gen = timed_steps(lambda: prog.grid(shape, out=vals), steps, warmup) * 1e3
        ms_mc = timed_steps(lambda: mc_dense_raw(vals, shape, (-1.0,) * 3, (1.0,) * 3, 0.0, 0), steps, warmup) * 1e3
        v, f = marching_cubes_sdf(prog, shape)
        out[f"{n}^3"] = {"fused_ms": ms_fused, "grid_ms": ms_gen, "mc_on_grid_ms": ms_mc,
                         "stored_total_ms": ms_gen + ms_mc, "vertices": int(v.shape[0]), "faces": int(f.shape[0])}
        del vals
        torch.cuda.empty_cache()
    return out


def bench_sparse2048(args):
    """Differentiable MC on a SparseGrid at 2048^3 effective resolution:
    narrow band |sdf| < 2h around a sphere r 0.7 plus 3-octave seeded
    Fourier noise (amp 0.02).  A step = forward sparse MC + backward of
    loss = sum(w * v) (w ~ N(0, 1) in canonical vertex order) to the (N, 8)
    values.  The gradient is checked against FP64 central differences (step
    1e-3, all copies of a corner perturbed together) on 1000 seeded grid
    points."""
    import numpy as np
    import torch

    import isoext_amd as iso
    from isoext_amd import workloads
    from isoext_amd.autograd import marching_cubes
    from isoext_amd.sparse import mc_sparse
    n = 2048
    sdf = lambda p: workloads.noisy_sphere_eval(p, r=0.7, amp=0.02, seed=0)
    t0 = time.perf_counter()
    ids = workloads.narrow_band_ids(n, 2.0, sdf)
    print(f"sparse2048: narrow band {ids.numel()} cells", file=sys.stderr, flush=True)
    g = iso.SparseGrid([n, n, n])
    g.add_cells(ids)
    pts = g.get_points()
    vals0 = torch.cat([sdf(c).reshape(-1, 8) for c in pts.reshape(-1, 8, 3).split(1 << 21)]).contiguous()
    del pts
    g.set_values(vals0)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    v, f = marching_cubes(g, vals0, 0.0)
    order = np.lexsort(v.detach().cpu().numpy().T[::-1])
    w_np = np.empty((v.shape[0], 3))
    w_np[order] = np.random.default_rng(0).standard_normal((v.shape[0], 3))
    w = torch.from_numpy(w_np).float().cuda()

    def fwd():
        mc_sparse(g, 0.0, 0, want_src=True)

    def fwd_bwd():
        vals = vals0.detach().requires_grad_(True)
        vv, _ = marching_cubes(g, vals, 0.0)
        (vv * w).sum().backward()
    print(f"sparse2048: {ids.numel()} cells, build {t_build:.1f} s", file=sys.stderr, flush=True)
    ms_f = timed_steps(fwd, args.steps, args.warmup) * 1e3
    ms = timed_steps(fwd_bwd, args.steps, args.warmup) * 1e3
    print(f"sparse2048: forward {ms_f:.2f} ms, forward+backward {ms:.2f} ms", file=sys.stderr, flush=True)
    # gradient check
    vals = vals0.detach().requires_grad_(True)
    vv, _ = marching_cubes(g, vals, 0.0)
    (vv * w).sum().backward()
    grad = vals.grad.double().cpu().numpy().reshape(-1)
    _, _, src = mc_sparse(g, 0.0, 0, want_src=True)
    check = {f"step_{st:g}": sparse_fd_check(g, vals0, src, w_np, grad, 1000, st) for st in (1e-3, 1e-6)}
    N = ids.numel()
    nv, nf = v.shape[0], f.shape[0]
    alg_f = 32 * N + 8 * N + 12 * nv + 12 * nf
    alg = alg_f + 12 * nv + 32 * N
    return {
        "metric": "sparse differentiable MC at 2048^3 narrow band: cells/s (forward + backward)",
        "value": N / (ms * 1e-3) / 1e9, "unit": "Gcells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "SparseGrid 2048^3, |sdf| < 2h band, sphere r 0.7 + 3-octave noise amp 0.02",
                   "active_cells": int(N), "vertices": nv, "faces": nf, "forward_ms": ms_f,
                   "build_s": t_build, "grad_check": check,
                   "roofline_forward": {"bound": "hbm", "achieved": alg_f / (ms_f * 1e-3) / 1e9,
                                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                        "frac": alg_f / (ms_f * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                        "alg_bytes": alg_f}},
        "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
    }


def sparse_fd_check(g, vals0, src, w, grad, n_check, step):
    """FP64 central differences of loss = sum(w * v) on the frozen topology,
    perturbing every copy of a grid point together, against the summed
    analytic gradient of those copies."""
    import numpy as np
    src = src.cpu().numpy()
    slot, e = src // 12, src % 12
    EA = np.array([0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 3, 2])
    EB = np.array([1, 3, 3, 2, 5, 7, 7, 6, 4, 5, 7, 6])
    ca, cb = EA[e], EB[e]
    vals = vals0.double().cpu().numpy().reshape(-1)
    ia, ib = 8 * slot + ca, 8 * slot + cb
    cells = g.cell_ids.cpu().numpy()
    X, Y, Z = g.shape
    cz_, r = cells % (Z - 1), cells // (Z - 1)
    cy_, cx_ = r % (Y - 1), r // (Y - 1)
    k = np.arange(8)
    pid = (((cx_[:, None] + (k >> 2)) * Y + cy_[:, None] + ((k >> 1) & 1)) * Z + cz_[:, None] + (k & 1)).reshape(-1)
    coord = lambda i, nn: -1.0 + (i / (nn - 1)) * 2.0
    def pos(flat):
        p = pid[flat]
        z, rr = p % Z, p // Z
        y, x = rr % Y, rr // Y
        return np.stack([coord(x, X), coord(y, Y), coord(z, Z)], 1)
    pa, pb = pos(ia), pos(ib)
    # value copies per grid point
    order = np.argsort(pid, kind="stable")
    spid = pid[order]
    uniq, start, cnt = np.unique(spid, return_index=True, return_counts=True)
    touched = np.unique(np.concatenate([pid[ia], pid[ib]]))
    pick = np.random.default_rng(1).choice(touched, min(n_check, len(touched)), replace=False)
    # vertices touching each point
    vp = np.concatenate([pid[ia], pid[ib]])
    vi = np.concatenate([np.arange(len(ia)), np.arange(len(ib))])
    vo = np.argsort(vp, kind="stable")
    vps, vis = vp[vo], vi[vo]
    errs = []
    near = 0
    scale = np.abs(grad).max() + 1e-30
    for q, p in enumerate(pick):
        if q % 200 == 0:
            print(f"fd check {q}/{len(pick)}", file=sys.stderr, flush=True)
        j = np.searchsorted(uniq, p)
        copies = order[start[j]:start[j] + cnt[j]]
        lo, hi = np.searchsorted(vps, p), np.searchsorted(vps, p, side="right")
        verts = np.unique(vis[lo:hi])
        ia_v, ib_v = ia[verts], ib[verts]
        in_a, in_b = np.isin(ia_v, copies), np.isin(ib_v, copies)

        def loss(delta):
            va = vals[ia_v] + np.where(in_a, delta, 0.0)
            vb = vals[ib_v] + np.where(in_b, delta, 0.0)
            t = (0.0 - va) / (vb - va)
            pos_ = (1 - t)[:, None] * pa[verts] + t[:, None] * pb[verts]
            return float((pos_ * w[verts]).sum())
        fd = (loss(step) - loss(-step)) / (2 * step)
        an = float(grad[copies].sum())
        errs.append(abs(fd - an) / (abs(fd) + 1e-2 * scale))
        # the step is not small against an edge's value difference: the
        # lerp is far from linear over [-step, step]
        near += int(step > 0.1 * np.abs(vals[ib_v] - vals[ia_v]).min())
    errs = np.array(errs)
    return {"entries": int(len(pick)), "fd_step": step, "max_rel_err": float(errs.max()),
            "p50_rel_err": float(np.median(errs)), "p99_rel_err": float(np.percentile(errs, 99)),
            "entries_step_over_0.1_dv": near}


def serving_bench(n_requests: int, seed: int = 0, sdf: bool = False):
    """Continuous-batching server vs the plain per-call path on the same
    seeded request stream: dense grids with cubic sides log-uniform in
    [32, 512], SDF = union of 1-64 random spheres, 80% marching cubes / 20%
    dual contouring, HBM admission budget 200 GB.  Inputs come from a pool
    of 320 distinct grids generated on the GPU before timing, taken in a
    fixed permutation, so that the 8 workers x 32 requests in flight never
    hold one grid twice (no batch sees a grid that another request of the
    batch already pulled into the caches).  sdf=True: the same distribution
    of fields as SDF programs (union-of-spheres trees) that the batched
    kernels evaluate without storing a grid; the plain path writes each
    request's grid (set_values) and then extracts."""
    import numpy as np
    import torch

    from isoext_amd import serving
    pool = serving.make_sdf_pool(POOL_GRIDS, seed=seed) if sdf else serving.make_pool(POOL_GRIDS, seed=seed)
    torch.cuda.synchronize()
    pts = float(np.mean([(e[1][0] ** 3 if isinstance(e, tuple) else e.numel()) for e in pool]))

    def stream(rate, s):
        return serving.make_stream(pool, n_requests, rate=rate, seed=s, dc_frac=0.2, cycle=True)

    # warm-up both paths (library load, workspaces grown to their steady-state
    # sizes, autotune tables): a full saturated stream of another seed through
    # the server, a fifth of one through the plain path
    def consume(r):   # the client takes each mesh as it completes (10k meshes would hold ~70 GB)
        r.result = None

    def drop(reqs):
        for r in reqs:
            r.result = None

    warm = stream(0.0, seed + 99)
    # defaults: 32 requests / 2^30 points per batch, 8 workers
    srv = serving.Server(hbm_budget=200e9, on_done=consume)
    serving.drive(srv, warm)
    drop(warm)
    for r in warm[: max(16, n_requests // 5)]:
        serving.run_plain(r)
    drop(warm)
    del warm
    torch.cuda.synchronize()
    # saturation throughput: every request available at t = 0; the median of
    # three passes over the same stream for each path
    walls_b, walls_p = [], []
    for _ in range(3):
        sat = stream(0.0, seed)
        w, lat_sat = serving.drive(srv, sat)
        walls_b.append(w)
        drop(sat)
    for _ in range(3):
        plain = stream(0.0, seed)
        w, _ = serving.drive_plain(plain, on_done=consume)
        walls_p.append(w)
        drop(plain)
    wall_b, wall_p = sorted(walls_b)[1], sorted(walls_p)[1]
    thr_b, thr_p = n_requests / wall_b, n_requests / wall_p
    # latency under a Poisson load at 80% of the plain path's capacity
    rate = 0.8 * thr_p
    lb = stream(rate, seed + 1)
    _, lat_b = serving.drive(srv, lb)
    drop(lb)
    lp = stream(rate, seed + 1)
    _, lat_p = serving.drive_plain(lp, on_done=consume)
    drop(lp)
    stats = dict(srv.stats)
    srv.close()
    q = lambda a, x: float(np.percentile(a, x) * 1e3)
    return {
        "requests": n_requests, "dc_fraction": 0.2, "mean_points": pts, "pool_grids": len(pool),
        "fields": "SDF programs evaluated in the kernels (no stored grid)" if sdf else "stored grids",
        "meshes_per_s_batched": thr_b, "meshes_per_s_plain": thr_p, "throughput_ratio": thr_b / thr_p,
        "saturation_passes_ms": {"batched": [w * 1e3 for w in walls_b], "plain": [w * 1e3 for w in walls_p]},
        "poisson_rate_per_s": rate,
        "latency_ms_batched": {"p50": q(lat_b, 50), "p99": q(lat_b, 99)},
        "latency_ms_plain": {"p50": q(lat_p, 50), "p99": q(lat_p, 99)},
        "saturation_latency_ms_batched": {"p50": q(lat_sat, 50), "p99": q(lat_sat, 99)},
        "server": {"batches": stats["batches"], "batched_requests": stats["batched_requests"],
                   "dc_requests": stats["dc_requests"], "max_hbm_in_use_GB": stats["max_in_use"] / 1e9},
    }


def fused_bench(args, shape, m, v_grid, f_grid, grid_ms_step):
    """The same extraction with the sphere-union field evaluated inside the
    kernels (isoext_mc_spheres): time per extraction, whether the mesh is
    identical to the grid path's, and the time the generator takes to write
    the grid that the grid path reads."""
    import torch

    from isoext_amd.fused import sphere_tensor
    from isoext_amd.mc import mc_dense_raw
    from isoext_amd.workloads import sphere_params, spheres_field
    c, r = sphere_params(args.spheres, args.seed)
    sp = sphere_tensor(c, r)
    amin, amax = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    for _ in range(2):
        v, f, _ = mc_dense_raw(None, shape, amin, amax, 0.0, m, args.cfg, spheres=sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = max(3, args.steps // 2)
    for _ in range(steps):
        v, f, _ = mc_dense_raw(None, shape, amin, amax, 0.0, m, args.cfg, spheres=sp)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    same = bool(v.shape == v_grid.shape and f.shape == f_grid.shape and torch.equal(v, v_grid)
                and torch.equal(f, f_grid))
    same_set = same
    if not same and v.shape == v_grid.shape and f.shape == f_grid.shape:
        # the grid path at this size runs the read-once pipeline, whose
        # vertex order differs: compare the canonical meshes
        import numpy as np

        import oracle
        a = oracle.canonical(v.cpu().numpy(), f.cpu().numpy())
        b = oracle.canonical(v_grid.cpu().numpy(), f_grid.cpu().numpy())
        same_set = bool(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]))
    del v, f
    buf = torch.empty(shape, dtype=torch.float32, device="cuda")
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    spheres_field(shape, c, r, out=buf)
    b.record()
    torch.cuda.synchronize()
    gen_ms = a.elapsed_time(b)
    del buf
    npts = shape[0] * shape[1] * shape[2]
    return {"workload": f"marching cubes of the {args.spheres}-sphere union evaluated in the kernels "
                        f"(isoext_mc_spheres), no grid stored",
            "ms_per_extraction": ms, "gvoxels_s": npts / (ms * 1e-3) / 1e9,
            "identical_to_grid_path": same, "same_mesh_as_grid_path": same_set,
            "grid_path_ms_per_step": grid_ms_step,
            "grid_generation_ms": gen_ms}


def host_info():
    """nproc, the affinity mask's size, the CPU model (lscpu) and the OpenMP
    thread setting of this host."""
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("Socket(s):"):
                info["sockets"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        info["model"] = None
    return info


def cpu_threads():
    """Threads the CPU baseline uses: the host's share for one GPU (the
    OpenMP setting, 16 on the GPU hosts), at most the affinity mask."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    share = int(omp) if omp and omp.isdigit() else 16
    return max(1, min(share, aff))


def cpu_baseline(values_gpu, planes: int, level: float, method: str):
    """The CPU oracle (C++ restatement of the reference, OpenMP, parallel
    weld sort) on the first `planes` x-planes of the same field; returns
    Gvoxels/s and details."""
    import numpy as np

    import oracle
    sample = values_gpu[:planes].contiguous().cpu().numpy()
    threads = cpu_threads()
    oracle.mc_dense(sample[:9], level, method, threads=threads)   # warm-up (load, page-in)

    def timed(arr, nthreads, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            oracle.mc_dense(arr, level, method, threads=nthreads)
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]
    t = timed(sample, threads, 5)
    pts = int(np.prod(sample.shape))
    one = sample[: max(2, planes // 8)]
    t1 = timed(one, 1, 3)
    return {"value": pts / t / 1e9, "unit": "Gvoxels/s", "cores": threads, "kind": "port",
            "sample": f"oracle marching cubes ({method}) on the first {planes} x-planes "
                      f"({sample.shape[0]}x{sample.shape[1]}x{sample.shape[2]} points) of the same "
                      f"field, {threads} threads, median of 5 runs, {t:.2f} s each",
            "host": host_info(),
            "single_core": {"value": int(np.prod(one.shape)) / t1 / 1e9, "unit": "Gvoxels/s", "cores": 1,
                            "sample": f"first {one.shape[0]} x-planes, median of 3 runs, {t1:.2f} s each"}}


def main():
    args = parse()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary, args.calibrate_copy or 1024, args.size ** 3), indent=1))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(every box has one MI355X).", file=sys.stderr)
        sys.exit(2)
    if args.workload != "mc1024":
        import torch
        torch.cuda.set_device(0)
        fn = {"dc256": bench_dc256, "sparse2048": bench_sparse2048,
              "csg": lambda a: {"csg_mc": csg_mc_fused(steps=a.steps, warmup=a.warmup),
                                "csg_dc256": csg_dc_fused((256, 256, 256), a.steps, a.warmup)}}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return
    import torch

    import isoext_amd as iso
    from isoext_amd import _lib, workloads
    from isoext_amd.mc import METHODS, mc_dense_raw

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    N = args.size
    shape = (N, N, N)
    grid = iso.UniformGrid(list(shape))
    workloads.multi_sphere(shape, args.spheres, args.seed, out=grid.values)
    torch.cuda.synchronize()
    L = _lib.lib()
    m = METHODS[args.method]

    import ctypes
    kernel_ms = []
    stages = []
    st = (ctypes.c_float * 5)()
    nv = nf = 0
    cfg = None
    for _ in range(args.warmup):
        v, f, info = mc_dense_raw(grid.values, shape, grid.aabb_min, grid.aabb_max, 0.0, m, args.cfg)
    # the timed steps run the path as a caller does: no stage events
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        v, f, info = mc_dense_raw(grid.values, shape, grid.aabb_min, grid.aabb_max, 0.0, m, args.cfg)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # stage durations from HIP events on the launch stream, in steps of their own
    L.isoext_set_timing(1)
    for _ in range(max(5, args.steps)):
        v, f, info = mc_dense_raw(grid.values, shape, grid.aabb_min, grid.aabb_max, 0.0, m, args.cfg)
        kernel_ms.append(L.isoext_last_kernel_ms())
        L.isoext_mc_last_stage_ms(st)
        stages.append(list(st))
    torch.cuda.synchronize()
    L.isoext_set_timing(0)
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy * 2 ** 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    nv, nf, cfg = v.shape[0], f.shape[0], info["cfg"]
    ms_step = (t1 - t0) * 1e3 / args.steps
    npts = N * N * N
    kms = sorted(kernel_ms)[len(kernel_ms) // 2]
    kmean = sum(kernel_ms) / len(kernel_ms)
    read_once = cfg == READ_ONCE_CFG
    names = (["reset", "stream", "scan_flush", "faces", "total"] if read_once
             else ["classify", "count", "scan", "emit", "total"])
    stage_mean = {k: sum(s[i] for s in stages) / len(stages) for i, k in enumerate(names)}
    alg_bytes = 4 * npts + 12 * nv + 12 * nf
    pipe_gbs = alg_bytes / (stage_mean["total"] * 1e-3) / 1e9
    if read_once:
        # dominant kernel: k_stream reads every value once and writes the vertices
        dom = {"kernel": "k_stream (4 B per grid point read + 12 B per vertex written)",
               "bytes": 4 * npts + 12 * nv, "ms": stage_mean["stream"]}
        tname = "mc_read_once_1024_traffic.json"
    else:
        dom = {"kernel": "k_classify (4 B per grid point read)", "bytes": 4 * npts, "ms": stage_mean["classify"]}
        tname = "mc_dense_1024_traffic.json"
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    traffic, traffic_source = None, None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", tname)
    if os.path.exists(tfile) and N == 1024:
        try:
            traffic = json.load(open(tfile)).get("traffic_bytes")
            traffic_source = f"profiles/{tname}: rocprofv3 FETCH_SIZE + WRITE_SIZE runs of this kernel at 1024^3"
        except (OSError, ValueError):
            traffic = None
    out = {
        "metric": METRIC,
        "value": npts / (ms_step * 1e-3) / 1e9,
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain marching cubes ({args.method}) on a dense {N}^3 fp32 UniformGrid over "
                        f"[-1,1]^3, SDF = min of {args.spheres} seeded spheres (seed {args.seed}), iso 0",
            "grid_points": [N, N, N],
            "vertices": nv,
            "faces": nf,
            "tile_config": cfg,
            "kernel_ms_median": kms,
            "kernel_ms_mean": kmean,
            "stage_ms_mean": stage_mean,
            "pipeline": {"alg_bytes": alg_bytes, "device_ms": stage_mean["total"], "achieved_GBs": pipe_gbs,
                         "frac": pipe_gbs / HBM_PEAK_GBS},
            "cell_rate_gcells_s": (N - 1) ** 3 / (ms_step * 1e-3) / 1e9,
        },
        "roofline": {
            "kernel": dom["kernel"],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_source,
        },
        "cpu_baseline": None,
    }
    if not args.no_fused:
        out["config"]["fused_spheres"] = fused_bench(args, shape, m, v, f, ms_step)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(grid.values, min(args.cpu_planes, N), 0.0, args.method)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": repr(e)}
    del grid, v, f
    torch.cuda.empty_cache()
    sub = argparse.Namespace(steps=10, warmup=2)
    if not args.no_dc:
        r = bench_dc256(sub)
        out["config"]["dc256"] = {"workload": r["config"]["workload"], "ms_per_step": r["ms_per_step"],
                                  "gvoxels_s": r["value"], "dual_error_h": r["config"]["dual_error_h"],
                                  "fp64_resolves": r["config"]["fp64_resolves"],
                                  "active_cells": r["config"]["active_cells"],
                                  "intersections": r["config"]["intersections"], "roofline": r["roofline"],
                                  "cpu_baseline": r["cpu_baseline"], "fused": r["config"]["fused"]}
        torch.cuda.empty_cache()
        out["config"]["csg_mc"] = csg_mc_fused()
        torch.cuda.empty_cache()
    if not args.no_sparse:
        r = bench_sparse2048(sub)
        out["config"]["sparse2048"] = {"workload": r["config"]["workload"], "ms_per_step": r["ms_per_step"],
                                       "forward_ms": r["config"]["forward_ms"], "gcells_s": r["value"],
                                       "active_cells": r["config"]["active_cells"],
                                       "vertices": r["config"]["vertices"], "faces": r["config"]["faces"],
                                       "grad_check": r["config"]["grad_check"], "roofline": r["roofline"],
                                       "roofline_forward": r["config"].get("roofline_forward")}
        torch.cuda.empty_cache()
    if args.serving > 0:
        out["config"]["serving"] = serving_bench(args.serving, args.seed)
        torch.cuda.empty_cache()
    if args.serving_sdf > 0:
        out["config"]["serving_sdf"] = serving_bench(args.serving_sdf, args.seed, sdf=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
