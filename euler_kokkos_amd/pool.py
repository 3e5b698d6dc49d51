"""Job pool (include/ek_pool.h) from Python, and the batched-run generator.

A :class:`Pool` keeps many independent grids (hydro 2D/3D, or 3D MHD) resident in one
device arena and advances all unfinished jobs by one step of their own dt per pool step (a
persistent kernel over a (job, tile) work list for hydro, the five 3D MHD pipeline kernels
each over all jobs' chunks for MHD, replayed from a captured HIP graph). A job's result
equals its single-grid :class:`native.Context` run.

:func:`random_blast_batch` draws the batched workload (256 independent 2D blasts) from
a seeded std::mt19937_64 stream with a hand-written uniform mapping, so the same seed
gives the same runs in any language; every run is expressible as a reference
``Hydro_Muscl_2D`` blast .ini (each real goes through the .ini float parser).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native, problems
from .native import EK_LAYOUT_SOA, EkError, EkParams
from .params import getfloat, make_params

EK_EFULL = 4

POOL_EXPORTS = ("ek_pool_create", "ek_pool_destroy", "ek_pool_last_error", "ek_pool_job_bytes",
                "ek_pool_free_bytes", "ek_pool_active", "ek_pool_admit", "ek_pool_step",
                "ek_pool_status", "ek_pool_download", "ek_pool_budgets", "ek_pool_retire",
                "ek_pool_autotune", "ek_pool_set_tile", "ek_pool_last_step_time", "ek_pool_test_fail_admit",
                "ek_pool_admit_async", "ek_pool_test_defer_join")

_BOUND = False


def _lib() -> C.CDLL:
    global _BOUND
    L = native.lib()
    if not _BOUND:
        P, V = C.POINTER, C.c_void_p
        D = P(C.c_double)
        L.ek_pool_create.argtypes = [P(EkParams), C.c_longlong, C.c_int, P(V)]
        L.ek_pool_destroy.argtypes = [V]
        L.ek_pool_destroy.restype = None
        L.ek_pool_last_error.argtypes = [V]
        L.ek_pool_last_error.restype = C.c_char_p
        L.ek_pool_job_bytes.argtypes = [V, P(EkParams)]
        L.ek_pool_job_bytes.restype = C.c_longlong
        L.ek_pool_free_bytes.argtypes = [V]
        L.ek_pool_free_bytes.restype = C.c_longlong
        L.ek_pool_active.argtypes = [V]
        L.ek_pool_admit.argtypes = [V, P(EkParams), D, C.c_int, C.c_int, C.c_double, P(C.c_int)]
        L.ek_pool_admit_async.argtypes = [V, P(EkParams), D, C.c_int, C.c_double, P(C.c_int)]
        L.ek_pool_step.argtypes = [V, C.c_int]
        L.ek_pool_status.argtypes = [V, C.c_int, D, P(C.c_int), P(C.c_int)]
        L.ek_pool_download.argtypes = [V, C.c_int, D, C.c_int]
        L.ek_pool_budgets.argtypes = [V, C.c_int, D]
        L.ek_pool_retire.argtypes = [V, C.c_int]
        L.ek_pool_autotune.argtypes = [V, C.c_int, P(C.c_int), D, C.c_int]
        L.ek_pool_set_tile.argtypes = [V, C.c_int]
        L.ek_pool_last_step_time.argtypes = [V, D, P(C.c_int)]
        L.ek_pool_test_fail_admit.argtypes = [V, C.c_int]
        L.ek_pool_test_defer_join.argtypes = [V, C.c_int]
        _BOUND = True
    return L


class PoolFull(EkError):
    """The job does not fit in the arena (or no job slot is free) right now."""


class Pool:
    def __init__(self, kind: EkParams, arena_bytes: int, max_jobs: int = 1024):
        self._L = _lib()
        h = C.c_void_p()
        rc = self._L.ek_pool_create(C.byref(kind), int(arena_bytes), int(max_jobs), C.byref(h))
        if rc != 0:
            raise EkError(f"ek_pool_create failed ({rc}): {self._L.ek_pool_last_error(None).decode()}")
        self.h = h
        self.kind = kind
        self.params = {}
        self._inputs = {}  # host arrays of asynchronously admitted jobs, held until retire

    def _check(self, rc, what):
        if rc == EK_EFULL:
            raise PoolFull(f"{what}: {self._L.ek_pool_last_error(self.h).decode()}")
        if rc != 0:
            raise EkError(f"{what} failed ({rc}): {self._L.ek_pool_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self._L.ek_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def job_bytes(self, p: EkParams) -> int:
        return int(self._L.ek_pool_job_bytes(self.h, C.byref(p)))

    def free_bytes(self) -> int:
        return int(self._L.ek_pool_free_bytes(self.h))

    def active(self) -> int:
        return int(self._L.ek_pool_active(self.h))

    def admit(self, p: EkParams, U: np.ndarray, nStepmax: int, tEnd: float = 1e30,
              layout: int = EK_LAYOUT_SOA, wait: bool = True) -> int:
        """Admit a job. With wait=False (SoA layout only) the upload runs on the pool's upload
        stream while later steps advance the other jobs, and the job joins the first step after
        it has completed; the host array is held until the job is retired (page-locked memory,
        e.g. native.pinned_copy, lets the copy overlap the steps)."""
        a = np.ascontiguousarray(U, dtype=np.float64)
        jid = C.c_int(-1)
        ptr = a.ctypes.data_as(C.POINTER(C.c_double))
        if wait:
            self._check(self._L.ek_pool_admit(self.h, C.byref(p), ptr, layout, int(nStepmax), float(tEnd),
                                              C.byref(jid)), "ek_pool_admit")
        else:
            if layout != EK_LAYOUT_SOA:
                raise ValueError("asynchronous admission takes the SoA layout only")
            self._check(self._L.ek_pool_admit_async(self.h, C.byref(p), ptr, int(nStepmax), float(tEnd),
                                                    C.byref(jid)), "ek_pool_admit_async")
            self._inputs[jid.value] = a
        self.params[jid.value] = p
        return jid.value

    def step(self, nsteps: int = 1) -> None:
        self._check(self._L.ek_pool_step(self.h, int(nsteps)), "ek_pool_step")

    def status(self, jid: int):
        t, it, fin = C.c_double(), C.c_int(), C.c_int()
        self._check(self._L.ek_pool_status(self.h, jid, C.byref(t), C.byref(it), C.byref(fin)),
                    "ek_pool_status")
        return t.value, it.value, bool(fin.value)

    def download(self, jid: int) -> np.ndarray:
        p = self.params[jid]
        shape = (p.nbvar, p.ny, p.nx) if p.dim == 2 else (p.nbvar, p.nz, p.ny, p.nx)
        a = np.empty(shape, dtype=np.float64)
        self._check(self._L.ek_pool_download(self.h, jid, a.ctypes.data_as(C.POINTER(C.c_double)),
                                             EK_LAYOUT_SOA), "ek_pool_download")
        return a

    def budgets(self, jid: int) -> np.ndarray:
        out = np.zeros(8)
        self._check(self._L.ek_pool_budgets(self.h, jid, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "ek_pool_budgets")
        return out[: self.params[jid].nbvar]

    def retire(self, jid: int) -> None:
        self._check(self._L.ek_pool_retire(self.h, jid), "ek_pool_retire")
        self.params.pop(jid, None)
        self._inputs.pop(jid, None)

    def autotune(self, steps: int = 3):
        ch = C.c_int(-1)
        ms = np.zeros(8)
        self._check(self._L.ek_pool_autotune(self.h, int(steps), C.byref(ch),
                                             ms.ctypes.data_as(C.POINTER(C.c_double)), 8), "ek_pool_autotune")
        n = 4
        return ch.value, ms[:n].tolist()

    def set_tile(self, choice: int) -> None:
        self._check(self._L.ek_pool_set_tile(self.h, int(choice)), "ek_pool_set_tile")

    def last_step_time(self):
        ms, n = C.c_double(), C.c_int()
        self._check(self._L.ek_pool_last_step_time(self.h, C.byref(ms), C.byref(n)), "ek_pool_last_step_time")
        return ms.value, n.value

    def test_fail_admit(self, stage: int) -> None:
        """Make every following admission fail at `stage` (1 before the upload, 2 after it, 3
        after the job's device record is written; 0 turns it off) -- for rollback tests."""
        self._check(self._L.ek_pool_test_fail_admit(self.h, int(stage)), "ek_pool_test_fail_admit")

    def test_defer_join(self, calls: int) -> None:
        """Asynchronously admitted jobs join no earlier than the `calls`-th step() call after
        their admission while other jobs are stepping (0 turns it off) -- for tests."""
        self._check(self._L.ek_pool_test_defer_join(self.h, int(calls)), "ek_pool_test_defer_join")


# ------------------------------------------------------------------ seeded generator


class MT19937_64:
    """std::mt19937_64 (the 64-bit Mersenne Twister of the C++ standard library)."""

    _M = 0xFFFFFFFFFFFFFFFF

    def __init__(self, seed: int):
        self.mt = [0] * 312
        self.mt[0] = seed & self._M
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & self._M
        self.idx = 312

    def _twist(self):
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def next_u64(self) -> int:
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self._M

    def uniform(self) -> float:
        """[0, 1) from the top 53 bits: (x >> 11) * 2^-53."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def random_blast_batch(n_runs: int = 256, seed: int = 2024, precision: int = native.EK_FP32):
    """The batched 2D hydro workload: n_runs independent blasts.

    Per run: edge N = 64 + 32k with k uniform in {0..14}; centre uniform in [0.2,0.8]^2;
    radius uniform in [0.05,0.15]; rho = 1 inside and out; p_out = 0.1 and
    p_in = 0.1 * 10^(1+2u) (inner/outer ratio log-uniform in [10,1000]); gamma 1.4,
    HLLC, MC slopes (slope_type 2), CFL 0.5, reflecting walls. Returns a list of
    (params, initial SoA state, description) tuples."""
    rng = MT19937_64(seed)
    runs = []
    for r in range(n_runs):
        k = min(14, int(rng.uniform() * 15))
        n = 64 + 32 * k
        cx = repr(0.2 + 0.6 * rng.uniform())
        cy = repr(0.2 + 0.6 * rng.uniform())
        rad = repr(0.05 + 0.1 * rng.uniform())
        pin = repr(0.1 * 10.0 ** (1.0 + 2.0 * rng.uniform()))
        p = make_params("Hydro_Muscl_2D", n, n, gamma0="1.4", cfl="0.5", slope_type="2",
                        riemann="hllc", bc=(1, 1, 1, 1, 1, 1), precision=precision)
        U = problems.blast(p, radius=rad, center_x=cx, center_y=cy, density_in="1",
                           density_out="1", pressure_in=pin, pressure_out="0.1")
        runs.append((p, U, dict(n=n, center=(getfloat(cx), getfloat(cy)), radius=getfloat(rad),
                                pressure_in=getfloat(pin))))
    return runs
