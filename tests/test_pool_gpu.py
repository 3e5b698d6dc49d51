"""The job pool on the GPU: every job's result equals its single-grid run (FP64 and FP32,
bitwise), whether it runs alone or in a mix, is admitted at once or while others are
already stepping; arena accounting; the autotuner leaves the jobs' states unchanged."""
import numpy as np
import pytest

from euler_kokkos_amd import native, pool, problems
from euler_kokkos_amd.params import make_params

pytestmark = pytest.mark.gpu


def single_run(p, U0, nsteps, tEnd=1e30):
    ctx = native.Context(p)
    try:
        ctx.upload(U0, native.EK_LAYOUT_SOA)
        t, it, _ = ctx.run(nsteps, nsteps, tEnd)
        return ctx.download(native.EK_LAYOUT_SOA), t, it
    finally:
        ctx.close()


def jobs2d(precision):
    out = []
    for n, (nx, ny, bc, rad) in enumerate([(64, 64, (1,) * 4, "0.1"), (96, 40, (3, 3, 1, 2), "0.2"),
                                           (33, 77, (2, 1, 3, 3), "0.15"), (128, 128, (1,) * 4, "0.05")]):
        p = make_params("Hydro_Muscl_2D", nx, ny, gamma0="1.4", cfl="0.5", slope_type="2",
                        riemann="hllc", bc=bc + (1, 1), precision=precision)
        U = problems.blast(p, radius=rad, center_x="0.4", center_y="0.55", density_out="1",
                           pressure_in=str(10 ** (1 + n)), pressure_out="0.1")
        out.append((p, U))
    return out


@pytest.mark.parametrize("precision", [native.EK_FP64, native.EK_FP32])
def test_pool_2d_matches_single_grid_with_staggered_admission(precision):
    jobs = jobs2d(precision)
    steps = [20, 13, 27, 9]
    ref = [single_run(p, U, s) for (p, U), s in zip(jobs, steps)]
    P = pool.Pool(jobs[0][0], 256 << 20, 16)
    try:
        ids = [P.admit(jobs[0][0], jobs[0][1], steps[0]), P.admit(jobs[1][0], jobs[1][1], steps[1])]
        P.step(5)       # the first two run alone for a while
        ids.append(P.admit(jobs[2][0], jobs[2][1], steps[2]))
        P.step(7)
        ids.append(P.admit(jobs[3][0], jobs[3][1], steps[3]))
        P.step(40)      # everybody reaches nStepmax
        for k, jid in enumerate(ids):
            t, it, fin = P.status(jid)
            G = P.download(jid)
            print(f"case=pool2d prec={precision} job={k}: it={it} max|diff|={np.abs(G - ref[k][0]).max():.3e}",
                  flush=True)
            assert fin and it == steps[k] and it == ref[k][2] and t == ref[k][1]
            assert np.array_equal(G, ref[k][0])
    finally:
        P.close()


def test_pool_3d_matches_single_grid():
    specs = [(32, 24, 20, (3,) * 6), (40, 32, 48, (1, 2, 3, 3, 1, 1))]
    jobs = []
    for nx, ny, nz, bc in specs:
        p = make_params("Hydro_Muscl_3D", nx, ny, nz, gamma0="1.4", cfl="0.5", slope_type="2",
                        riemann="hllc", bc=bc)
        jobs.append((p, problems.blast(p, radius="0.2", density_out="1", pressure_out="0.1")))
    ref = [single_run(p, U, 8) for p, U in jobs]
    P = pool.Pool(jobs[0][0], 256 << 20, 8)
    try:
        ids = [P.admit(p, U, 8) for p, U in jobs]
        P.step(8)
        for k, jid in enumerate(ids):
            assert np.array_equal(P.download(jid), ref[k][0])
    finally:
        P.close()


def test_pool_arena_full_and_retire():
    p = make_params("Hydro_Muscl_2D", 128, 128, gamma0="1.4", cfl="0.5", slope_type="2", riemann="hllc",
                    precision=native.EK_FP32)
    U = problems.blast(p, density_out="1")
    P = pool.Pool(p, 3 * (2 * 128 * 128 * 4 * 4) + 4096, 8)
    try:
        need = P.job_bytes(p)
        a = [P.admit(p, U, 4) for _ in range(3)]
        with pytest.raises(pool.PoolFull):
            P.admit(p, U, 4)
        P.retire(a[1])
        assert P.free_bytes() >= need
        a[1] = P.admit(p, U, 4)
        P.step(4)
        ref = single_run(p, U, 4)[0]
        for jid in a:
            assert np.array_equal(P.download(jid), ref)
    finally:
        P.close()


def test_autotune_restores_state_and_keeps_results():
    jobs = jobs2d(native.EK_FP64)
    ref = [single_run(p, U, 12) for p, U in jobs]
    P = pool.Pool(jobs[0][0], 256 << 20, 8)
    try:
        ids = [P.admit(p, U, 12) for p, U in jobs]
        P.step(3)
        choice, ms = P.autotune(2)
        print(f"case=autotune choice={choice} ms={ms}", flush=True)
        assert 0 <= choice < 4 and all(m > 0 for m in ms)
        P.step(9)
        for k, jid in enumerate(ids):
            assert np.array_equal(P.download(jid), ref[k][0])
            b = P.budgets(jid)
            assert abs(b[0] - ref[k][0][0].sum()) <= 1e-10 * ref[k][0][0].sum()
    finally:
        P.close()


def test_finished_job_stays_frozen_while_others_step():
    # a job that reached its nStepmax keeps its final time, iteration and state while the
    # pool goes on stepping the others over several calls
    jobs = jobs2d(native.EK_FP64)[:2]
    ref = [single_run(jobs[0][0], jobs[0][1], 5), single_run(jobs[1][0], jobs[1][1], 20)]
    P = pool.Pool(jobs[0][0], 256 << 20, 8)
    try:
        a = P.admit(jobs[0][0], jobs[0][1], 5)
        b = P.admit(jobs[1][0], jobs[1][1], 20)
        for n in (6, 1, 2):
            P.step(n)
            t, it, fin = P.status(a)
            print(f"case=frozen job after step({n}): it={it} fin={fin}", flush=True)
            assert fin and it == 5 and t == ref[0][1]
            assert np.array_equal(P.download(a), ref[0][0])
        P.step(12)
        t, it, fin = P.status(b)
        assert fin and it == 20 and t == ref[1][1]
        assert np.array_equal(P.download(b), ref[1][0])
        assert np.array_equal(P.download(a), ref[0][0])
    finally:
        P.close()


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_failed_admission_leaves_the_pool_as_it_was(stage):
    # an admission that fails after its arena range was reserved gives the range and the slot
    # back; the jobs already admitted keep stepping to their single-grid results and the
    # next admission succeeds
    jobs = jobs2d(native.EK_FP64)
    ref = [single_run(p, U, 12) for p, U in jobs]
    P = pool.Pool(jobs[0][0], 256 << 20, 8)
    try:
        ids = [P.admit(p, U, 12) for p, U in jobs[:2]]
        P.step(3)
        free0, act0 = P.free_bytes(), P.active()
        P.test_fail_admit(stage)
        with pytest.raises(native.EkError):
            P.admit(jobs[2][0], jobs[2][1], 12)
        print(f"case=admission rollback stage {stage}: free {P.free_bytes()} / {free0}", flush=True)
        assert P.free_bytes() == free0 and P.active() == act0
        P.step(2)
        P.test_fail_admit(0)
        ids.append(P.admit(jobs[3][0], jobs[3][1], 12))
        P.step(20)
        for k, jid in zip((0, 1, 3), ids):
            t, it, fin = P.status(jid)
            assert fin and it == 12 and t == ref[k][1]
            assert np.array_equal(P.download(jid), ref[k][0]), k
    finally:
        P.close()


def jobs_mhd3d(precision):
    out = []
    for nx, ny, nz, bc, B0 in [(24, 20, 16, (3,) * 6, 0.5), (16, 28, 20, (1, 1, 3, 3, 2, 2), 1.0),
                               (32, 32, 32, (3,) * 6, 0.2)]:
        p = make_params("MHD_Muscl_3D", nx, ny, nz, gamma0="1.6666667", cfl="0.8", slope_type="2",
                        riemann="hlld", bc=bc, precision=precision)
        out.append((p, problems.mhd_blast3d(p, B0=B0, radius="0.2", center=("0.45", "0.5", "0.55"))))
    return out


@pytest.mark.parametrize("precision", [native.EK_FP64, native.EK_FP32])
def test_pool_mhd3d_matches_single_grid(precision):
    # 3D MHD jobs of different shapes, boundaries and fields, admitted while others step and
    # stopping at different nStepmax: each ends bitwise at its single-grid ek_run result
    jobs = jobs_mhd3d(precision)
    steps = [9, 14, 7]
    ref = [single_run(p, U, s) for (p, U), s in zip(jobs, steps)]
    P = pool.Pool(jobs[0][0], 512 << 20, 8)
    try:
        free0 = P.free_bytes()
        ids = [P.admit(jobs[0][0], jobs[0][1], steps[0]), P.admit(jobs[1][0], jobs[1][1], steps[1])]
        assert P.free_bytes() == free0 - P.job_bytes(jobs[0][0]) - P.job_bytes(jobs[1][0])
        P.step(4)
        ids.append(P.admit(jobs[2][0], jobs[2][1], steps[2]))
        P.step(20)
        for k, jid in enumerate(ids):
            t, it, fin = P.status(jid)
            G = P.download(jid)
            print(f"case=pool mhd3d prec={precision} job={k}: it={it} max|diff|={np.abs(G - ref[k][0]).max():.3e}",
                  flush=True)
            assert fin and it == steps[k] == ref[k][2] and t == ref[k][1]
            assert np.array_equal(G, ref[k][0])
        for jid in ids:
            P.retire(jid)
        assert P.free_bytes() == free0 and P.active() == 0
    finally:
        P.close()


def jobs3d(precision):
    out = []
    for nx, ny, nz, bc in [(32, 24, 20, (3,) * 6), (40, 32, 48, (1, 2, 3, 3, 1, 1)), (24, 40, 28, (2,) * 6)]:
        p = make_params("Hydro_Muscl_3D", nx, ny, nz, gamma0="1.4", cfl="0.5", slope_type="2",
                        riemann="hllc", bc=bc, precision=precision)
        out.append((p, problems.blast(p, radius="0.2", density_out="1", pressure_out="0.1")))
    return out


@pytest.mark.parametrize("kind,precision", [("hydro2d", native.EK_FP64), ("hydro3d", native.EK_FP32),
                                            ("hydro3d", native.EK_FP64), ("mhd3d", native.EK_FP32)])
def test_async_admission_matches_single_grid(kind, precision):
    # jobs admitted without waiting for their upload join while another job steps: one after
    # an odd number of steps (its state is moved to the other arena buffer), one after an even
    # number, one at the next step; each ends bitwise at its single-grid result. A job retired
    # while still uploading, and one downloaded before it joined, leave the pool consistent.
    jobs = {"hydro2d": jobs2d, "hydro3d": jobs3d, "mhd3d": jobs_mhd3d}[kind](precision)[:3]
    steps = [12, 9, 7]
    ref = [single_run(p, U, s) for (p, U), s in zip(jobs, steps)]
    P = pool.Pool(jobs[0][0], 512 << 20, 8)
    try:
        free0 = P.free_bytes()
        inputs = [native.pinned_copy(U) for _, U in jobs]
        ids = [P.admit(jobs[0][0], inputs[0], steps[0], wait=False)]
        P.step(2)                   # nothing else to step: the first step waits for its upload
        assert P.status(ids[0])[1] == 2
        P.test_defer_join(2)        # joins at the second step() call: after one more step
        ids.append(P.admit(jobs[1][0], inputs[1], steps[1], wait=False))
        P.step(1)
        assert P.status(ids[1])[1] == 0
        P.test_defer_join(3)        # after two more steps
        ids.append(P.admit(jobs[2][0], inputs[2], steps[2], wait=False))
        P.step(1)
        P.step(1)
        # a job that is retired before it joins, and one downloaded before it joins
        P.test_defer_join(0)
        gone = P.admit(jobs[0][0], inputs[0], 5, wait=False)
        P.retire(gone)
        early = P.admit(jobs[1][0], inputs[1], 3, wait=False)
        D = P.download(early)       # waits for the upload: the initial state as the pool holds it
        E = single_run(jobs[1][0], jobs[1][1], 0)[0]
        assert np.array_equal(D, E)
        P.retire(early)
        P.step(30)
        for k, jid in enumerate(ids):
            t, it, fin = P.status(jid)
            G = P.download(jid)
            print(f"case=pool async {kind} prec={precision} job={k}: it={it} "
                  f"max|diff|={np.abs(G - ref[k][0]).max():.3e}", flush=True)
            assert fin and it == steps[k] == ref[k][2] and t == ref[k][1]
            assert np.array_equal(G, ref[k][0])
        for jid in ids:
            P.retire(jid)
        assert P.free_bytes() == free0 and P.active() == 0
    finally:
        P.close()
