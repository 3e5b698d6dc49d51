"""Parity of the 2D MHD HIP path (MHD_Muscl_2D through libek_hip.so) with the oracle.

FP64 must match the oracle cell by cell (the build targets bit-exact agreement; the
test bound is 1e-12 relative). The full-size Orszag-Tang configuration (1024^2, 500 steps)
is in test_configs_gpu.py.
"""
import numpy as np
import pytest

from euler_kokkos_amd import native, problems
from euler_kokkos_amd.params import make_params

pytestmark = pytest.mark.gpu

RTOL = 1e-12
ID, IP, IU, IV, IW, IA, IB, IC = range(8)


def rel_err(a, b):
    scale = np.maximum(np.abs(b), 1e-12 * np.max(np.abs(b)))
    return float(np.max(np.abs(a - b) / scale))


def mparams(nx, ny, **kw):
    kw.setdefault("gamma0", "1.666")
    kw.setdefault("cfl", "0.8")
    kw.setdefault("slope_type", "2")
    kw.setdefault("riemann", "hlld")
    kw.setdefault("smallr", "1e-7")
    kw.setdefault("smallc", "1e-7")
    kw.setdefault("bc", (3,) * 6)
    return make_params("MHD_Muscl_2D", nx, ny, **kw)


def run_pair(oracle, ekp, R0, nsteps, tEnd=1e30, use_step=False):
    op = oracle.from_ek(ekp)
    R = R0.copy()
    t_o, it_o, _ = oracle.mhd_run(op, R, nsteps, tEnd)
    ctx = native.Context(ekp)
    try:
        ctx.upload(R0)
        if use_step:
            t, it = 0.0, 0
            for _ in range(nsteps):
                dt = ekp.cfl / ctx.inv_dt()
                if t + dt > tEnd:
                    dt = tEnd - t
                ctx.step(dt)
                t += dt
                it += 1
        else:
            t, it, _ = ctx.run(nsteps, nsteps, tEnd)
        G = ctx.download()
    finally:
        ctx.close()
    assert it == it_o
    assert t == t_o
    return oracle.interior(op, G), oracle.interior(op, R)


def test_oblique_alfven_wave_bitwise(oracle):
    # the oracle's second-order oblique Alfven wave (test_oracle_mhd.py) to the time it returns
    # to its initial state: the GPU result is bitwise the oracle's
    from test_oracle_mhd import oblique_alfven_2d
    ekp = mparams(64, 64)
    op, R0, T = oblique_alfven_2d(oracle, 64)
    assert (op.nx, op.ny, op.gamma0) == (ekp.nx, ekp.ny, oracle.from_ek(ekp).gamma0)
    g, o = run_pair(oracle, ekp, R0, 100000, tEnd=T)
    print(f"case=mhd2d oblique Alfven 64^2 to T={T:.6f}", flush=True)
    assert np.array_equal(g, o), rel_err(g, o)


@pytest.mark.parametrize("rs,st", [("hlld", "2"), ("hlld", "1"), ("hll", "2"), ("llf", "1")])
def test_orszag_tang_fp64_parity(oracle, rs, st):
    ekp = mparams(64, 48, riemann=rs, slope_type=st)  # ragged against the 14x14 tiles
    R0 = oracle.orszag_tang(oracle.from_ek(ekp))
    g, o = run_pair(oracle, ekp, R0, 24)
    print(f"case=mhd2d orszag-tang 64x48 rs={rs} st={st}: max|diff|={np.max(np.abs(g - o)):.3e}",
          flush=True)
    assert rel_err(g, o) <= RTOL


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (2, 2, 2, 2), (1, 2, 3, 3), (3, 3, 1, 2)])
def test_mixed_boundaries_fp64_parity(oracle, bc):
    ekp = mparams(45, 30, bc=bc + (3, 3))
    R0 = oracle.orszag_tang(oracle.from_ek(ekp))
    g, o = run_pair(oracle, ekp, R0, 16)
    print(f"case=mhd2d bc={bc}: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert rel_err(g, o) <= RTOL


def test_host_driven_steps_and_tend_clamp(oracle):
    ekp = mparams(40, 40)
    R0 = oracle.orszag_tang(oracle.from_ek(ekp))
    g, o = run_pair(oracle, ekp, R0, 12, use_step=True)
    assert rel_err(g, o) <= RTOL
    # a tEnd reached mid-run: the last dt is clamped exactly as compute_dt does
    g, o = run_pair(oracle, ekp, R0, 1000, tEnd=0.01)
    assert rel_err(g, o) <= RTOL


def test_inv_dt_and_ghost_download(oracle):
    ekp = mparams(30, 26, bc=(1, 2, 3, 3, 3, 3))
    op = oracle.from_ek(ekp)
    R0 = oracle.orszag_tang(op)
    ctx = native.Context(ekp)
    try:
        ctx.upload(R0)
        oracle.lib().eko_make_boundaries(__import__("ctypes").byref(op), oracle.dp(R0))
        assert ctx.inv_dt() == oracle.mhd_inv_dt(op, R0)
        G = ctx.download()
    finally:
        ctx.close()
    # every ghost cell (corners included) as make_boundaries fills it
    assert np.array_equal(G, R0)


def test_product_initial_condition_matches_oracle(oracle):
    ekp = mparams(50, 38)
    op = oracle.from_ek(ekp)
    ctx = native.Context(ekp)
    try:
        ctx.upload(problems.orszag_tang(ekp), native.EK_LAYOUT_SOA)
        A = ctx.download(native.EK_LAYOUT_SOA)
    finally:
        ctx.close()
    assert np.array_equal(A, problems.soa_from_reference(ekp, oracle.orszag_tang(op)))


def test_orszag_tang_fp32_l1():
    ekp64 = mparams(256, 256)
    ekp32 = mparams(256, 256, precision=native.EK_FP32)
    U0 = problems.orszag_tang(ekp64)
    out = []
    for ekp in (ekp64, ekp32):
        ctx = native.Context(ekp)
        try:
            ctx.upload(U0, native.EK_LAYOUT_SOA)
            ctx.run(100, 100, 1e30)
            out.append(ctx.download(native.EK_LAYOUT_SOA))
        finally:
            ctx.close()
    d64, d32 = out
    l1 = np.abs(d32[ID] - d64[ID]).mean() / np.abs(d64[ID]).mean()
    print(f"case=OT 256^2 100 steps FP32 vs FP64: rel L1(rho)={l1:.3e}", flush=True)
    assert np.isfinite(d32).all()
    assert l1 <= 1e-4


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (2, 2, 2, 2), (1, 2, 3, 3), (3, 3, 2, 1)])
def test_mixed_boundaries_fp32_l1(bc):
    # the FP32 step's 30x14 tiles on a grid that leaves partial tiles on both axes, with
    # reflecting / outflow / periodic sides: relative L1 of every variable against the
    # FP64 step (itself bitwise the oracle's on these boundaries) within 1e-4
    ekp64 = mparams(97, 61, bc=bc + (3, 3))
    ekp32 = mparams(97, 61, bc=bc + (3, 3), precision=native.EK_FP32)
    U0 = problems.orszag_tang(ekp64)
    out = []
    for ekp in (ekp64, ekp32):
        ctx = native.Context(ekp)
        try:
            ctx.upload(U0, native.EK_LAYOUT_SOA)
            t, it, _ = ctx.run(40, 40, 1e30)
            out.append(ctx.download(native.EK_LAYOUT_SOA))
        finally:
            ctx.close()
        assert it == 40
    d64, d32 = out
    assert np.isfinite(d32).all()
    for v in (ID, IP, IU, IV, IA, IB):
        l1 = np.abs(d32[v] - d64[v]).mean() / np.abs(d64[v]).mean()
        print(f"case=mhd2d FP32 bc={bc} var={v}: rel L1 {l1:.3e}", flush=True)
        assert l1 <= 1e-4
