"""Continuous-batching extraction server on one GPU.

Requests (marching cubes or dual contouring on a dense grid already in HBM,
or on the field of an SDF program -- csg.SdfProgram -- that the kernels
evaluate where they need it, so the request's grid is never stored) arrive
at any time.  A request is *admitted* when its estimated device
footprint (mask workspace, count arrays, outputs) fits in the HBM budget
that in-flight requests leave free; admitted marching-cubes requests are
grouped by whichever worker becomes free next into one ragged batch
(``isoext_amd.batch.mc_batch``: one launch per stage for the whole batch),
and so are dual-contouring requests (``isoext_amd.batch.dc_batch``).
Workers own separate HIP streams,
so batches overlap on the device; a request that arrives while others run is
picked up by the next free worker (continuous batching, no fixed batch
boundaries).  Each result is identical to the plain per-call path on the
same request (tests/test_serving_gpu.py).
"""

from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class Request:
    rid: int
    values: torch.Tensor | None     # (X, Y, Z) float32 on the server's device (None with sdf)
    kind: str = "mc"                # "mc" or "dc"
    level: float = 0.0
    method: int = 0                 # marching cubes: 0 nagae, 1 lorensen
    arrival: float = 0.0            # seconds after the stream start
    # filled in by the server
    t_admit: float = -1.0
    t_done: float = -1.0
    result: tuple | None = None
    error: str | None = None
    done: threading.Event = field(default_factory=threading.Event)
    sdf: object = None              # csg.SdfProgram on the server's device (values None)
    shape: tuple | None = None      # grid points per axis of an sdf request

    @property
    def grid_shape(self) -> tuple:
        return tuple(int(s) for s in (self.values.shape if self.sdf is None else self.shape))

    @property
    def points(self) -> int:
        X, Y, Z = self.grid_shape
        return X * Y * Z


def work_bytes(req: Request) -> int:
    """Device bytes of a request's intermediate state: the inside masks and
    count arrays (about N/8 + N/16 bytes) plus a fixed 1 MiB."""
    n = req.points
    return n // 8 + n // 16 + (1 << 20)


def out_estimate(req: Request) -> int:
    """Output bytes assumed at admission, before the request's counts are
    known: a surface crossing every x-plane (12 B per vertex + 24 B per face,
    at most 3 vertices per cross-section cell; twice that for DC)."""
    X, Y, Z = req.grid_shape
    out = 3 * max(X * Y, Y * Z, X * Z) * (12 + 24)
    return 2 * out if req.kind == "dc" else out


def footprint(req: Request) -> int:
    """Device bytes a request is admitted with (work + output estimate);
    after its batch's first phase the reservation becomes work + the exact
    output bytes."""
    return work_bytes(req) + out_estimate(req)


def validate(req: Request) -> str | None:
    """Why a request cannot run (None when it can): the checks of the
    per-call API, done at submission so that a bad request never joins a
    batch."""
    if req.sdf is not None:
        if req.values is not None:
            return "a request holds values or an SDF program, not both"
        if req.shape is None or len(req.shape) != 3 or any(int(s) < 1 for s in req.shape):
            return "an SDF request needs a 3-entry positive shape"
        if not hasattr(req.sdf, "ptr") or getattr(req.sdf, "device", None) is None:
            return "sdf must be an isoext_amd.csg.SdfProgram"
    else:
        v = req.values
        if not isinstance(v, torch.Tensor) or v.dtype != torch.float32 or v.dim() != 3 or not v.is_cuda:
            return "values must be a 3-D float32 CUDA tensor"
        if not v.is_contiguous():
            return "values must be C-contiguous"
    if req.kind not in ("mc", "dc"):
        return f"unknown request kind {req.kind!r}"
    if req.kind == "mc" and req.method not in (0, 1):
        return f"Unknown method: {req.method}"
    return None


class Server:
    def __init__(self, device=0, hbm_budget: float = 200e9, max_batch: int = 32,
                 max_batch_points: int = 1 << 30, workers: int = 8, cfg: int | None = None, on_done=None):
        """on_done(request), if given, runs on the worker thread when a
        request completes (a client that consumes each mesh as it arrives,
        e.g. copies it out and drops it)."""
        self.on_done = on_done
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        self.max_batch_points = max_batch_points
        self.cfg = cfg
        self._lock = threading.Condition()
        self._pending: collections.deque[Request] = collections.deque()   # arrived, not admitted
        self._mc: collections.deque[Request] = collections.deque()         # admitted, not started
        self._dc: collections.deque[Request] = collections.deque()
        self._in_use = 0
        self._seq = 0
        self._stop = False
        self._t0 = time.perf_counter()
        self.stats = {"batches": 0, "batched_requests": 0, "dc_requests": 0, "max_in_use": 0}
        self._workers = [threading.Thread(target=self._work, args=(i,), daemon=True) for i in range(workers)]
        for w in self._workers:
            w.start()

    # ---- client side -----------------------------------------------------------
    def now(self) -> float:
        return time.perf_counter() - self._t0

    def reset_clock(self):
        self._t0 = time.perf_counter()

    def submit(self, req: Request):
        self.submit_many([req])

    def submit_many(self, reqs):
        """Enqueue requests that arrived together (one admission pass).  A
        request that fails validation completes at once with its error."""
        ok = []
        for r in reqs:
            err = validate(r)
            if err is None:
                ok.append(r)
                continue
            r.error = err
            r.t_admit = r.t_done = self.now()
            if self.on_done is not None:
                self.on_done(r)
            r.done.set()
        with self._lock:
            self._pending.extend(ok)
            self._admit_locked()
            self._lock.notify_all()

    def close(self):
        with self._lock:
            self._stop = True
            self._lock.notify_all()
        for w in self._workers:
            w.join()

    # ---- admission ---------------------------------------------------------------
    def _admit_locked(self):
        while self._pending:
            r = self._pending[0]
            need = footprint(r)
            if self._in_use + need > self.budget and self._in_use > 0:
                break   # wait for in-flight requests to free memory (FIFO, no overtaking)
            self._pending.popleft()
            r._bytes = need
            self._in_use += need
            self.stats["max_in_use"] = max(self.stats["max_in_use"], self._in_use)
            r.t_admit = self.now()
            r._seq = self._seq
            self._seq += 1
            (self._dc if r.kind == "dc" else self._mc).append(r)

    def _release(self, reqs):
        with self._lock:
            for r in reqs:
                self._in_use -= r._bytes
            self._admit_locked()
            self._lock.notify_all()

    def _sized(self, reqs, out_bytes):
        """A batch's first phase has counted its meshes: each reservation
        becomes work + exact output bytes (frees budget when the estimate was
        high, takes more when it was low)."""
        with self._lock:
            for r, ob in zip(reqs, out_bytes):
                exact = work_bytes(r) + int(ob)
                self._in_use += exact - r._bytes
                r._bytes = exact
            self.stats["max_in_use"] = max(self.stats["max_in_use"], self._in_use)
            self._admit_locked()
            self._lock.notify_all()

    # ---- workers -----------------------------------------------------------------
    def _take_locked(self):
        """Next unit of work: the oldest admitted request decides its kind,
        and the batch takes every admitted request of that kind (in
        admission order, up to the size limits)."""
        if not self._mc and not self._dc:
            return None
        q = self._dc if self._dc and (not self._mc or self._dc[0]._seq < self._mc[0]._seq) else self._mc
        batch, pts = [], 0
        while q and len(batch) < self.max_batch:
            r = q[0]
            if batch and pts + r.points > self.max_batch_points:
                break
            batch.append(q.popleft())
            pts += r.points
        return batch

    def _work(self, wid: int):
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)
        while True:
            with self._lock:
                while not self._mc and not self._dc and not self._stop:
                    self._lock.wait()
                if self._stop and not self._mc and not self._dc:
                    return
                work = self._take_locked()
            if not work:
                continue
            try:
                with torch.cuda.stream(stream):
                    if work[0].kind == "dc":
                        self._run_dc(work)
                    else:
                        self._run_mc(work)
                stream.synchronize()
            except Exception as e:  # the batch failed: run its requests one by one
                if len(work) == 1:
                    work[0].error = repr(e)
                else:
                    self._run_alone(work, stream)
            t = self.now()
            for r in work:
                r.t_done = t
                if self.on_done is not None:
                    self.on_done(r)
                r.done.set()
            self._release(work)

    def _run_alone(self, reqs, stream):
        """After a batch failure: each request by itself on the worker's
        stream, so one failing request leaves its batch-mates' results
        untouched (each equals its per-call result)."""
        for r in reqs:
            try:
                with torch.cuda.stream(stream):
                    if r.kind == "dc":
                        self._run_dc([r])
                    else:
                        self._run_mc([r])
                stream.synchronize()
                r.error = None
            except Exception as e:
                r.result = None
                r.error = repr(e)
        with self._lock:
            self.stats["isolated_batches"] = self.stats.get("isolated_batches", 0) + 1

    def _run_mc(self, reqs):
        from .batch import BATCH_CFG, GridJob, mc_batch
        jobs = [GridJob(r.values, r.level, r.method, sdf=r.sdf, shape=r.shape) for r in reqs]
        out = mc_batch(jobs, cfg=self.cfg if self.cfg is not None else BATCH_CFG, ws_kind="serving",
                       on_sizes=lambda b: self._sized(reqs, b))
        for r, (v, f) in zip(reqs, out):
            r.result = (v, f) if f.shape[0] > 0 else (None, None)
        with self._lock:
            self.stats["batches"] += 1
            self.stats["batched_requests"] += len(reqs)

    def _run_dc(self, reqs):
        from .batch import GridJob, dc_batch
        # program requests take trilinear normals, as dual_contouring on their stored grid
        out = dc_batch([GridJob(r.values, r.level, sdf=r.sdf, shape=r.shape) for r in reqs], ws_kind="serving_dc",
                       on_sizes=lambda b: self._sized(reqs, b))
        for r, res in zip(reqs, out):
            r.result = res
        with self._lock:
            self.stats["dc_requests"] += len(reqs)
            self.stats["dc_batches"] = self.stats.get("dc_batches", 0) + 1


# ---- plain per-call path (the baseline) --------------------------------------------

def run_plain(r: Request):
    """One request through the plain API (one extraction per call)."""
    from .dc import dual_contouring
    from .grid import UniformGrid
    from .mc import marching_cubes
    # a program request: set_values(sdf(points)) first, as the reference's users do
    g = UniformGrid.wrap(r.values if r.sdf is None else r.sdf.grid(r.shape))
    if r.kind == "dc":
        return dual_contouring(g, level=r.level)
    return marching_cubes(g, r.level, ["nagae", "lorensen"][r.method])


# ---- request streams ---------------------------------------------------------------

def make_pool(n_grids: int, seed: int = 0, side_lo: int = 32, side_hi: int = 512, device="cuda",
              max_points: int | None = None):
    """Distinct synthetic SDF grids: cubic sides log-uniform in [side_lo,
    side_hi], each the union (min) of 1-64 random spheres (centres
    U[-0.8, 0.8]^3, radii U[0.05, 0.25], as in the 1024^3 benchmark field)."""
    from .workloads import spheres_field
    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(n_grids):
        side = int(round(np.exp(rng.uniform(np.log(side_lo), np.log(side_hi)))))
        if max_points is not None:
            side = min(side, int(max_points ** (1 / 3)))
        k = int(rng.integers(1, 65))
        c = rng.uniform(-0.8, 0.8, size=(k, 3)).astype(np.float32)
        r = rng.uniform(0.05, 0.25, size=k).astype(np.float32)
        pool.append(spheres_field((side, side, side), c, r, device=device))
    return pool


def make_sdf_pool(n_programs: int, seed: int = 0, side_lo: int = 32, side_hi: int = 512, device="cuda"):
    """Request fields as SDF programs, never stored: (program, shape) pairs
    with the make_pool distribution (cubic sides log-uniform in [side_lo,
    side_hi], unions of 1-64 random spheres) compiled from isoext.sdf trees."""
    from . import sdf as S
    from .csg import SdfProgram
    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(n_programs):
        side = int(round(np.exp(rng.uniform(np.log(side_lo), np.log(side_hi)))))
        k = int(rng.integers(1, 65))
        c = rng.uniform(-0.8, 0.8, size=(k, 3)).astype(np.float32)
        r = rng.uniform(0.05, 0.25, size=k).astype(np.float32)
        tree = S.UnionOp([S.TranslationOp(S.SphereSDF(float(r[j])), [float(x) for x in c[j]]) for j in range(k)])
        pool.append((SdfProgram(tree, device), (side, side, side)))
    return pool


def make_stream(pool, n: int, rate: float, seed: int = 0, dc_frac: float = 0.2, cycle: bool = False):
    """n requests, Poisson arrivals at `rate` per second (rate <= 0: all at
    t = 0), dc_frac of them dual contouring.  Grids are drawn from the pool
    at random, or with cycle=True in a fixed seeded permutation repeated, so
    that any len(pool) consecutive requests use distinct grids.  Pool entries
    are grids or (SdfProgram, shape) pairs (make_sdf_pool)."""
    rng = np.random.default_rng(seed + 1)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n)) if rate > 0 else np.zeros(n)
    perm = np.random.default_rng(seed + 2).permutation(len(pool))
    reqs = []
    for i in range(n):
        kind = "dc" if rng.random() < dc_frac else "mc"
        g = int(perm[i % len(pool)]) if cycle else int(rng.integers(len(pool)))
        method = int(rng.integers(0, 2)) if kind == "mc" else 0
        if isinstance(pool[g], tuple):
            reqs.append(Request(rid=i, values=None, kind=kind, method=method, arrival=float(t[i]),
                                sdf=pool[g][0], shape=pool[g][1]))
        else:
            reqs.append(Request(rid=i, values=pool[g], kind=kind, method=method, arrival=float(t[i])))
    return reqs


def drive(server: Server, reqs):
    """Submit requests at their arrival times, wait for all; returns
    (wall seconds, latencies in seconds)."""
    server.reset_clock()
    t0 = time.perf_counter()
    i = 0
    while i < len(reqs):
        dt = reqs[i].arrival - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        # every request that has arrived by now goes in one submission
        now = time.perf_counter() - t0
        j = i + 1
        while j < len(reqs) and reqs[j].arrival <= now:
            j += 1
        server.submit_many(reqs[i:j])
        i = j
    for r in reqs:
        r.done.wait()
    wall = time.perf_counter() - t0
    lat = np.array([r.t_done - r.arrival for r in reqs])
    return wall, lat


def drive_plain(reqs, on_done=None):
    """The per-call baseline on the same stream: one request at a time, in
    arrival order, each started no earlier than its arrival; on_done as for
    Server."""
    t0 = time.perf_counter()
    lat = []
    for r in reqs:
        dt = r.arrival - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        r.result = run_plain(r)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0 - r.arrival)
        if on_done is not None:
            on_done(r)
    return time.perf_counter() - t0, np.array(lat)
