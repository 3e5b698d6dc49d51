"""The five benchmark configurations at their full sizes, checked against the oracle.

C1 (2D implosion 512^2, 200 steps) is in test_parity_gpu.py. Here:
  C2  3D blast 256^3 periodic, HLLC, CFL 0.5, 100 steps: FP64 bitwise with the oracle, FP32
      bulk density/energy L1 against the FP64 run;
  C3  2D Orszag-Tang 1024^2, gamma 1.6666667 (as the .ini parser rounds it), HLLD + CT,
      500 steps: FP64 bitwise with the oracle, divergence of B and conservation;
  C4  256 random 2D blasts (seed 2024) for 100 steps in the FP32 and FP64 job pools: FP64
      jobs (smallest and largest edges among them) bitwise with the oracle, every FP32 job's
      density L1 against its FP64 twin;
  C5  the serving stream (32 Poisson jobs, 64^3-192^3, 50 steps, seed 77) through the
      server: every served result bitwise its run alone, FP32 L1 against FP64 for every job,
      and two jobs of each kind in FP64 bitwise with the oracle.
The oracle runs with OpenMP on the host; each comparison takes seconds to a minute.
"""
import numpy as np
import pytest

from euler_kokkos_amd import native, pool, problems, serving
from euler_kokkos_amd.params import make_params

pytestmark = pytest.mark.gpu
ID, IP, IU, IV, IW, IA, IB, IC = range(8)


def gpu_run(ekp, U0, nsteps, layout=native.EK_LAYOUT_REFERENCE):
    ctx = native.Context(ekp)
    try:
        ctx.upload(U0, layout)
        t, it, _ = ctx.run(nsteps, nsteps, 1e30)
        return ctx.download(layout), t, it
    finally:
        ctx.close()


@pytest.mark.timeout(600)
def test_c2_blast_256_fp64_bitwise_and_fp32_l1(oracle):
    kw = dict(gamma0="1.4", cfl="0.5", slope_type="2", riemann="hllc", bc=(3,) * 6)
    ekp = make_params("Hydro_Muscl_3D", 256, 256, 256, **kw)
    op = oracle.from_ek(ekp)
    # the [blast] section of the configuration, each real through the .ini float parser
    U0 = problems.reference_from_soa(ekp, problems.blast(ekp, radius="0.1", density_in="1", density_out="1",
                                                         pressure_in="10", pressure_out="0.1"))
    print("case=C2 256^3 blast FP64, 100 steps: GPU", flush=True)
    G, t, it = gpu_run(ekp, U0, 100)
    print("case=C2 256^3 blast FP64, 100 steps: oracle", flush=True)
    R = U0.copy()
    t_o, it_o, _ = oracle.run(op, R, 100, 1e30)
    assert it == it_o == 100 and t == t_o
    g, o = oracle.interior(op, G), oracle.interior(op, R)
    print(f"case=C2 FP64: max|gpu-oracle| = {np.abs(g - o).max():.3e}", flush=True)
    assert np.array_equal(g, o)
    del R, o
    G32, _, it32 = gpu_run(make_params("Hydro_Muscl_3D", 256, 256, 256, precision=native.EK_FP32, **kw), U0, 100)
    g32 = oracle.interior(op, G32)
    l1_rho = np.abs(g32[..., ID] - g[..., ID]).mean() / np.abs(g[..., ID]).mean()
    l1_e = np.abs(g32[..., IP] - g[..., IP]).mean() / np.abs(g[..., IP]).mean()
    print(f"case=C2 FP32 vs FP64: rel L1 rho {l1_rho:.3e}, E {l1_e:.3e}", flush=True)
    assert it32 == 100 and l1_rho <= 1e-4 and l1_e <= 1e-4


def divb_dx(U):
    """Face-field imbalance |dBx + dBy| of every cell of a periodic SoA state (= |div B| dx)."""
    A, B = U[IA], U[IB]
    return np.abs((np.roll(A, -1, axis=1) - A) + (np.roll(B, -1, axis=0) - B)).max()


@pytest.mark.timeout(600)
def test_c3_orszag_tang_1024_fp64_bitwise(oracle):
    ekp = make_params("MHD_Muscl_2D", 1024, 1024, gamma0="1.6666667", cfl="0.8", slope_type="2",
                      riemann="hlld", smallr="1e-7", smallc="1e-7", bc=(3,) * 6)
    assert ekp.gamma0 == float(np.float32(1.6666667))
    op = oracle.from_ek(ekp)
    U0 = oracle.orszag_tang(op)
    # the product's initial condition is the oracle's, bit for bit
    assert np.array_equal(problems.orszag_tang(ekp), problems.soa_from_reference(ekp, U0))
    print("case=C3 OT 1024^2 FP64, 500 steps: GPU", flush=True)
    ctx = native.Context(ekp)
    try:
        ctx.upload(U0)
        b0 = ctx.budgets()
        t, it, _ = ctx.run(500, 500, 1e30)
        b1 = ctx.budgets()
        G = ctx.download()
    finally:
        ctx.close()
    print("case=C3 OT 1024^2 FP64, 500 steps: oracle", flush=True)
    R = U0.copy()
    t_o, it_o, _ = oracle.mhd_run(op, R, 500, 1e30)
    assert it == it_o == 500 and t == t_o
    g, o = oracle.interior(op, G), oracle.interior(op, R)
    print(f"case=C3: max|gpu-oracle| = {np.abs(g - o).max():.3e}", flush=True)
    assert np.array_equal(g, o)
    S0 = problems.soa_from_reference(ekp, U0)
    S = problems.soa_from_reference(ekp, G)
    # div B: the initial field has none; after 500 steps the imbalance stays at round-off
    # (|div B| dx <= 1e-12 max|B|); the periodic grid conserves every budget to round-off
    assert divb_dx(S0) == 0.0
    bmax = np.abs(S[IA:IC]).max()
    print(f"case=C3: max|div B| dx = {divb_dx(S):.3e}, max|B| = {bmax:.3f}", flush=True)
    assert divb_dx(S) <= 1e-12 * bmax
    for v in (ID, IP, IU, IV, IA, IB):
        assert abs(b1[v] - b0[v]) <= 1e-12 * max(abs(b0[v]), np.abs(S0[v]).sum()), v


@pytest.mark.timeout(600)
def test_c4_batched_blasts_against_oracle(oracle):
    steps = 100
    runs = {prec: pool.random_blast_batch(256, 2024, prec) for prec in (native.EK_FP64, native.EK_FP32)}
    final = {}
    for prec, rl in runs.items():
        elem = 8 if prec == native.EK_FP64 else 4
        P = pool.Pool(rl[0][0], sum(2 * p.nx * p.ny * 4 * elem + 4096 for p, _, _ in rl) + (64 << 20), 256)
        try:
            ids = [P.admit(p, U, steps) for p, U, _ in rl]
            P.autotune(2)
            P.step(steps)
            for j in ids:
                t, it, fin = P.status(j)
                assert fin and it == steps
            final[prec] = [P.download(j) for j in ids]
        finally:
            P.close()
        print(f"case=C4 pool prec={prec}: 256 runs x {steps} steps done", flush=True)
    rl = runs[native.EK_FP64]
    edges = [p.nx for p, _, _ in rl]
    assert min(edges) == 64 and max(edges) == 512
    # FP64 pool jobs against the oracle: the smallest and largest edges and six more spread
    # over the batch
    pick = sorted({edges.index(64), edges.index(512), *range(7, 256, 41)})
    assert len(pick) >= 8
    for k in pick:
        p, U, desc = rl[k]
        op = oracle.from_ek(p)
        R = problems.reference_from_soa(p, U)
        _, it_o, _ = oracle.run(op, R, steps, 1e30)
        o = problems.soa_from_reference(p, R)
        print(f"case=C4 run {k} edge {p.nx} p_in {desc['pressure_in']:.4g}: "
              f"max|pool-oracle| = {np.abs(final[native.EK_FP64][k] - o).max():.3e}", flush=True)
        assert it_o == steps and np.array_equal(final[native.EK_FP64][k], o), k
    # every FP32 job against its FP64 twin
    l1 = [np.abs(a[ID] - b[ID]).mean() / np.abs(b[ID]).mean()
          for a, b in zip(final[native.EK_FP32], final[native.EK_FP64])]
    print(f"case=C4 FP32 vs FP64: max rel L1 rho {max(l1):.3e}", flush=True)
    assert max(l1) <= 1e-4


@pytest.mark.timeout(600)
def test_c5_serving_stream_against_alone_and_oracle(oracle):
    jobs = serving.poisson_stream(32, 20.0, 77, edge=(64, 192), steps=50)
    S = serving.Server(48 << 30, native.EK_FP32)
    try:
        stats = S.serve(jobs, keep_results=True)
    finally:
        S.close()
    print(f"case=C5 served: {stats}", flush=True)
    assert stats["jobs"] == 32 and stats["failed"] == 0
    assert {j.kind for j in jobs} == {"hydro_kh", "mhd_blast"}
    for j in jobs:
        assert j.done_t >= j.arrival and j.admit_t >= j.arrival
        # the served result is the job's FP32 run alone, bit for bit
        alone, _, it = gpu_run(j.params, j.U0, j.steps, native.EK_LAYOUT_SOA)
        assert it == j.steps and np.array_equal(j.result, alone), j.id
        # and within the FP32 tolerance of its FP64 run
        p64 = serving.job_params(j.kind, j.n, native.EK_FP64)
        d64, _, _ = gpu_run(p64, j.U0, j.steps, native.EK_LAYOUT_SOA)
        l1 = np.abs(j.result[ID] - d64[ID]).mean() / np.abs(d64[ID]).mean()
        assert l1 <= 1e-4, (j.id, l1)
        j.fp64 = d64
    # two jobs of each kind in FP64 (the two smallest) bitwise with the oracle
    for kind, orun in (("hydro_kh", oracle.run), ("mhd_blast", oracle.mhd3d_run)):
        for j in sorted((j for j in jobs if j.kind == kind), key=lambda j: j.n)[:2]:
            p64 = serving.job_params(j.kind, j.n, native.EK_FP64)
            op = oracle.from_ek(p64)
            R = problems.reference_from_soa(p64, np.asarray(j.U0))
            _, it_o, _ = orun(op, R, j.steps, 1e30)
            o = problems.soa_from_reference(p64, R)
            print(f"case=C5 job {j.id} {kind} {j.n}^3 FP64: max|gpu-oracle| = {np.abs(j.fp64 - o).max():.3e}",
                  flush=True)
            assert it_o == j.steps and np.array_equal(j.fp64, o), j.id
