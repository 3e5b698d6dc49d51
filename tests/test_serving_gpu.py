"""Continuous batching on the GPU: each job's result is the same alone and in the mix
(bitwise: the pool steps every job with its single-grid per-tile code), jobs wait for HBM
when the arena is full, and time-to-solution is measured from arrival."""
import numpy as np
import pytest

from euler_kokkos_amd import native, serving
from euler_kokkos_amd.pool import Pool

pytestmark = pytest.mark.gpu


def small_stream(seed=5, edge=(20, 40)):
    return serving.poisson_stream(8, rate=400.0, seed=seed, edge=edge, steps=7)


def alone(j):
    P = Pool(j.params, (64 if j.kind == "hydro_kh" else 512) << 20, 4)
    try:
        jid = P.admit(j.params, j.U0, j.steps)
        P.step(j.steps + 1)
        assert P.status(jid)[2]
        return P.download(jid)
    finally:
        P.close()


def test_each_result_same_alone_and_in_the_mix():
    jobs = small_stream()
    S = serving.Server(256 << 20)
    try:
        stats = S.serve(jobs, keep_results=True)
    finally:
        S.close()
    print(f"case=serving mix: {stats}", flush=True)
    assert stats["jobs"] == 8 and {j.kind for j in jobs} == {"hydro_kh", "mhd_blast"}
    for j in jobs:
        assert j.done_t >= j.arrival and j.result is not None
        ref = alone(j)
        assert np.array_equal(j.result, ref), j.id


def test_admission_waits_for_memory():
    jobs = [j for j in small_stream(seed=9, edge=(32, 32))]
    for j in jobs:  # one kind, so that the budget is one pool's
        j.kind = "hydro_kh"
        j.params = serving.job_params("hydro_kh", j.n, native.EK_FP32)
        j.U0 = serving.job_state(j)
    biggest = max(2 * j.cells * 5 * 4 for j in jobs)
    S = serving.Server(biggest + (1 << 16), kinds=("hydro_kh",))  # room for one job at a time
    try:
        stats = S.serve(jobs, keep_results=True)
    finally:
        S.close()
    spans = sorted((j.admit_t, j.done_t) for j in jobs)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0 - 1e-9  # never two jobs resident together
    for j in jobs:
        assert np.array_equal(j.result, alone(j))


def test_bad_jobs_fail_alone_in_the_mix():
    # one hydro and one MHD job that their pools refuse at admission (slope_type 0 is not
    # the pools' kind, and not an MHD trace option at all): both end with an error, every
    # other job's result is still bitwise its result alone
    jobs = small_stream(seed=11)
    bad = {next(j.id for j in jobs if j.kind == "hydro_kh"), next(j.id for j in jobs if j.kind == "mhd_blast")}
    for j in jobs:
        if j.id in bad:
            j.params = serving.job_params(j.kind, j.n, native.EK_FP32)
            j.params.slope_type = 0.0
    S = serving.Server(256 << 20)
    try:
        stats = S.serve(jobs, keep_results=True)
    finally:
        S.close()
    print(f"case=serving mix with 2 bad jobs: {stats}", flush=True)
    assert stats["failed"] == 2 and stats["jobs"] == len(jobs) - 2
    for j in jobs:
        if j.id in bad:
            assert j.error and j.result is None
        else:
            assert j.error is None and np.array_equal(j.result, alone(j)), j.id
