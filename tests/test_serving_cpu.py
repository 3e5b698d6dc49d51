"""Host side of serving: the seeded Poisson stream, the 3D Kelvin-Helmholtz state, and the
serving loop's handling of jobs that fail."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from euler_kokkos_amd import native, serving


def test_poisson_stream_is_seeded_and_in_range():
    a = serving.poisson_stream(200, rate=20.0, seed=3, edge=(64, 192), build_state=False)
    b = serving.poisson_stream(200, rate=20.0, seed=3, edge=(64, 192), build_state=False)
    assert [(j.arrival, j.n, j.seed) for j in a] == [(j.arrival, j.n, j.seed) for j in b]
    t = np.array([j.arrival for j in a])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert 0.035 < gaps.mean() < 0.065  # mean inter-arrival 1/20 s
    assert min(j.n for j in a) >= 64 and max(j.n for j in a) <= 192
    assert len({j.n for j in a}) > 50


def test_kh3d_state():
    j = serving.poisson_stream(1, rate=20.0, seed=1, edge=(16, 16), build_state=False)[0]
    U = serving.kh3d(j.params, 7)
    p = j.params
    z = p.zmin + p.dz / 2 + np.arange(p.nz) * p.dz
    inner = np.abs(z - 0.5) < 0.25
    assert np.all(U[0][inner] == 2.0) and np.all(U[0][~inner] == 1.0)
    vx = U[2] / U[0]
    assert np.allclose(vx[inner], 0.5) and np.allclose(vx[~inner], -0.5)
    vy = U[3] / U[0]
    x = p.xmin + p.dx / 2 + np.arange(p.nx) * p.dx
    assert np.abs(vy - 0.01 * np.sin(2 * np.pi * x)).max() <= 1e-3 + 1e-12
    pr = (p.gamma0 - 1) * (U[1] - 0.5 * (U[2] ** 2 + U[3] ** 2) / U[0])
    assert np.allclose(pr, 2.5, rtol=1e-12)
    assert np.array_equal(U, serving.kh3d(j.params, 7))


class _FakePool:
    """Stands in for a job pool in the serving loop: a job's params say how it behaves
    ("ok", "reject": refused at admission, "fail": its context fails after admission)."""

    def __init__(self, async_failures):
        self.jobs = {}
        self.next = 0
        self.async_failures = async_failures

    def admit(self, params, U, steps, wait=True):
        if params == "reject":
            raise native.EkError("ek_create failed (2): invalid parameters")
        jid, self.next = self.next, self.next + 1
        self.jobs[jid] = dict(params=params, it=0, steps=steps)
        return jid

    def step(self, n):
        for j in self.jobs.values():
            if j["params"] == "ok":
                j["it"] = min(j["steps"], j["it"] + n)

    def status(self, jid):
        j = self.jobs[jid]
        return 0.0, j["it"], j["it"] >= j["steps"] or j["params"] == "fail"

    def download(self, jid):
        return np.full(3, float(jid))

    def retire(self, jid):
        del self.jobs[jid]


class _FakeGroup(_FakePool):
    def failed(self, jid):
        return RuntimeError("context creation failed") if self.jobs[jid]["params"] == "fail" else None


def test_serving_loop_fails_bad_jobs_alone():
    # a job refused at admission and a job whose context fails later end with an error; the
    # others complete with their results, and the summary counts the failures apart
    S = serving.Server.__new__(serving.Server)
    S.pools = {"hydro_kh": _FakePool(False), "mhd_blast": _FakeGroup(True)}
    S.max_chunk = 4
    S.precision = native.EK_FP32
    S._ex = ThreadPoolExecutor(max_workers=2)
    kinds = ["hydro_kh", "mhd_blast"]
    behaviour = ["ok", "reject", "ok", "fail", "ok", "ok"]
    jobs = [serving.Job(i, kinds[i % 2], 8, 5, 0.001 * i, i, params=b) for i, b in enumerate(behaviour)]
    try:
        stats = S.serve(jobs, keep_results=True)
    finally:
        S._ex.shutdown()
    assert stats["jobs"] == 4 and stats["failed"] == 2
    for j, b in zip(jobs, behaviour):
        assert j.done_t >= 0
        if b == "ok":
            assert j.error is None and j.result is not None
        else:
            assert j.error and j.result is None
    assert not S.pools["hydro_kh"].jobs and not S.pools["mhd_blast"].jobs
