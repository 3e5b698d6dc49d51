"""Serving: a request loop that admits simulation jobs into device-resident job pools by
their HBM footprint and advances every admitted job together (continuous batching).

Jobs arrive on a seeded Poisson process. Each job kind has its own pool and its own loop,
the loops running side by side: at every turn a loop admits, in arrival order, each job of
its kind that has arrived and fits in the pool's memory (its upload runs on the pool's upload
stream beside the steps, and it joins the first step after the upload); then it advances all admitted jobs
of the pool together (one persistent launch per step for a hydro pool, the five 3D MHD
kernels each over all jobs for an MHD pool, replayed from a captured graph); finished jobs
are collected and retired, which frees memory for the queue. Time-to-solution is measured
from a job's arrival to its completion; throughput over the whole stream.

A job's result does not depend on what else runs with it: the pool steps each job with
the per-cell code of its single-grid run (see tests/test_serving_gpu.py).
"""
from __future__ import annotations

import math
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import native, problems
from .params import make_params
from .pool import MT19937_64, Pool, PoolFull


@dataclass
class Job:
    id: int
    kind: str          # "hydro_kh" (3D hydro Kelvin-Helmholtz shear layer) or "mhd_blast"
    n: int             # grid edge (n^3 cells)
    steps: int
    arrival: float     # seconds after the stream starts
    seed: int
    params: object = None
    U0: np.ndarray = None
    B0: float = 0.0    # field strength of an MHD blast
    admit_t: float = -1.0
    done_t: float = -1.0
    result: np.ndarray = field(default=None, repr=False)
    error: str = None  # why the job failed (it was rejected or its context could not be made)

    @property
    def cells(self) -> int:
        return self.n ** 3


def kh3d(p, seed: int) -> np.ndarray:
    """3D Kelvin-Helmholtz shear layer on the unit cube: rho = 2, vx = +0.5 where
    |z - 1/2| < 1/4, rho = 1, vx = -0.5 elsewhere; vy = 0.01 sin(2 pi x) plus uniform
    noise of amplitude 1e-3 (numpy PCG64 seeded with `seed`); vz = 0; p = 2.5."""
    x = p.xmin + p.dx / 2 + np.arange(p.nx) * p.dx
    z = p.zmin + p.dz / 2 + np.arange(p.nz) * p.dz
    inner = (np.abs(z - 0.5) < 0.25)[:, None, None]
    rng = np.random.default_rng(seed)
    noise = rng.uniform(-1e-3, 1e-3, size=(p.nz, p.ny, p.nx))
    rho = np.where(inner, 2.0, 1.0) * np.ones((p.nz, p.ny, p.nx))
    vx = np.where(inner, 0.5, -0.5) * np.ones((p.nz, p.ny, p.nx))
    vy = 0.01 * np.sin(2 * np.pi * x)[None, None, :] + noise
    pr = 2.5
    U = np.zeros((5, p.nz, p.ny, p.nx))
    U[0] = rho
    U[2] = rho * vx
    U[3] = rho * vy
    U[1] = pr / (p.gamma0 - 1.0) + 0.5 * rho * (vx * vx + vy * vy)
    return U


def job_params(kind: str, n: int, precision: int):
    if kind == "hydro_kh":
        return make_params("Hydro_Muscl_3D", n, n, n, gamma0="1.4", cfl="0.5", slope_type="2",
                           riemann="hllc", bc=(3,) * 6, precision=precision)
    if kind == "mhd_blast":
        return make_params("MHD_Muscl_3D", n, n, n, gamma0="1.6666667", cfl="0.8", slope_type="2",
                           riemann="hlld", bc=(3,) * 6, precision=precision)
    raise ValueError(f"unknown job kind {kind}")


def job_state(j: "Job") -> np.ndarray:
    if j.kind == "hydro_kh":
        return kh3d(j.params, j.seed)
    return problems.mhd_blast3d(j.params, B0=j.B0)


def poisson_stream(n_jobs: int, rate: float, seed: int, kinds=("hydro_kh", "mhd_blast"), edge=(64, 192), steps=50,
                   precision: int = native.EK_FP32, build_state: bool = True, pinned: bool = True):
    """Seeded job stream: exponential inter-arrival times of mean 1/rate seconds, kind
    uniform over `kinds`, edge uniform over the integers [edge[0], edge[1]]. With
    build_state, each job's initial state is built now and, with `pinned`, held in
    page-locked host memory, as a request's input buffer would be, so that its upload at
    admission runs at the link's full speed."""
    rng = MT19937_64(seed)
    t = 0.0
    jobs = []
    for i in range(n_jobs):
        t += -math.log(1.0 - rng.uniform()) / rate
        kind = kinds[min(len(kinds) - 1, int(rng.uniform() * len(kinds)))]
        n = edge[0] + min(edge[1] - edge[0], int(rng.uniform() * (edge[1] - edge[0] + 1)))
        js = rng.next_u64() & 0x7FFFFFFF
        j = Job(i, kind, n, steps, t, js)
        j.B0 = 0.1 + 0.9 * rng.uniform()
        j.params = job_params(kind, n, precision)
        if build_state:
            j.U0 = native.pinned_copy(job_state(j)) if pinned else job_state(j)
        jobs.append(j)
    return jobs


class Server:
    """Job pools per kind sharing an HBM budget; continuous batching over arrivals. Each
    kind's pool has its own loop (admission in arrival order as memory allows, then all of
    its admitted jobs advanced together), and the loops run side by side."""

    def __init__(self, hbm_bytes: int, precision: int = native.EK_FP32, kinds=("hydro_kh", "mhd_blast"),
                 max_jobs=512, max_chunk=6):
        self.pools = {}
        share = hbm_bytes // len(kinds)
        for k in kinds:
            self.pools[k] = Pool(job_params(k, 64, precision), share, max_jobs)
        self.precision = precision
        self.max_chunk = max_chunk  # most steps advanced per loop turn (between admissions)
        self._ex = ThreadPoolExecutor(max_workers=len(self.pools))

    def close(self):
        for p in self.pools.values():
            p.close()
        self._ex.shutdown()

    def serve(self, jobs, keep_results: bool = False, clock=time.perf_counter, sleep=time.sleep):
        """Serve a job stream: one admission-and-stepping loop per job kind (each kind has its
        own pool and its own share of HBM), the loops running side by side."""
        t0 = clock()
        by_kind = {k: [j for j in jobs if j.kind == k] for k in self.pools}
        futs = [self._ex.submit(self._serve_kind, k, js, t0, keep_results, clock, sleep)
                for k, js in by_kind.items() if js]
        launches = sum(f.result() for f in futs)
        return summarize(jobs, launches)

    def _serve_kind(self, kind, jobs, t0, keep_results, clock, sleep) -> int:
        pool = self.pools[kind]
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        active = {}  # jid -> Job
        launches = 0
        step_s = 0.0  # wall time of one step of this pool, last measured
        while pending or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                j = pending[0]
                try:
                    # the upload runs beside the steps of the jobs already admitted; the job
                    # joins the first step after it has completed
                    jid = pool.admit(j.params, j.U0, j.steps, wait=False)
                except PoolFull:
                    break  # first come, first served: wait for memory
                except native.EkError as e:  # rejected (e.g. invalid parameters): fail this job only
                    pending.popleft()
                    j.error = str(e)
                    j.done_t = clock() - t0
                    continue
                j.admit_t = clock() - t0
                active[jid] = j
                pending.popleft()
            if not active:
                if pending:
                    sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            # advance to the earliest completion (at most max_chunk steps), but do not run
            # past the next arrival by more than one step
            nstep = max(1, min(self.max_chunk, min(j.steps - pool.status(jid)[1] for jid, j in active.items())))
            if pending and step_s > 0.0:
                gap = pending[0].arrival - (clock() - t0)
                nstep = max(1, min(nstep, int(gap / step_s)))
            ts = clock()
            pool.step(nstep)
            launches += nstep
            step_s = (clock() - ts) / nstep
            done_t = clock() - t0
            for jid in list(active):
                _, it, fin = pool.status(jid)
                if fin:
                    j = active.pop(jid)
                    err = pool.failed(jid) if hasattr(pool, "failed") else None
                    if err is not None:
                        j.error = str(err)
                    elif keep_results:
                        j.result = pool.download(jid)
                    pool.retire(jid)
                    j.done_t = done_t
        return launches


def summarize(jobs, launches=0):
    """Throughput and time-to-solution percentiles over the jobs that completed; failed jobs
    are counted apart."""
    ok = [j for j in jobs if j.error is None]
    out = {"jobs": len(ok), "failed": len(jobs) - len(ok), "pool_launches": launches}
    if not ok:
        return out
    tts = np.array([j.done_t - j.arrival for j in ok])
    first = min(j.arrival for j in ok)
    last = max(j.done_t for j in ok)
    cells = sum(j.cells * j.steps for j in ok)
    out.update({"jobs_per_s": len(ok) / (last - first), "cell_updates_per_s": cells / (last - first),
                "tts_p50_s": float(np.percentile(tts, 50)), "tts_p99_s": float(np.percentile(tts, 99)),
                "tts_max_s": float(tts.max()), "span_s": last - first})
    return out


def serve_one_at_a_time(jobs, precision=native.EK_FP64, clock=time.perf_counter, sleep=time.sleep):
    """The plain path: one job at a time in arrival order, each in its own context (ek_run)."""
    t0 = clock()
    for j in sorted(jobs, key=lambda j: j.arrival):
        now = clock() - t0
        if j.arrival > now:
            sleep(j.arrival - now)
        p = job_params(j.kind, j.n, precision)
        ctx = native.Context(p)
        try:
            ctx.upload(j.U0, native.EK_LAYOUT_SOA)
            ctx.run(j.steps, j.steps, 1e30)
            ctx.sync()
        finally:
            ctx.close()
        j.done_t = clock() - t0
    return summarize(jobs)
