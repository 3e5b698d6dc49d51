"""Parity of the 3D MHD HIP pipeline (MHD_Muscl_3D through libek_hip.so) with the oracle."""
import ctypes as C

import numpy as np
import pytest

from euler_kokkos_amd import native, problems
from euler_kokkos_amd.params import make_params

pytestmark = pytest.mark.gpu
ID, IP, IU, IV, IW, IA, IB, IC = range(8)


def m3(nx, ny, nz, **kw):
    kw.setdefault("gamma0", "1.666")
    kw.setdefault("cfl", "0.8")
    kw.setdefault("slope_type", "2")
    kw.setdefault("riemann", "hlld")
    kw.setdefault("smallr", "1e-7")
    kw.setdefault("smallc", "1e-7")
    kw.setdefault("bc", (3,) * 6)
    return make_params("MHD_Muscl_3D", nx, ny, nz, **kw)


def smooth_state(oracle, ekp):
    """A generic smooth 3D state (every variable varies; each face field is independent
    of its own coordinate, so div B = 0) with a pressure blast."""
    op = oracle.from_ek(ekp)
    U = np.zeros(oracle.ref_shape(op))
    g = 3
    xs = [op.xmin + op.dx / 2 + (np.arange(op.isize) - g) * op.dx,
          op.ymin + op.dy / 2 + (np.arange(op.jsize) - g) * op.dy,
          op.zmin + op.dz / 2 + (np.arange(op.ksize) - g) * op.dz]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    tp = 2 * np.pi
    rho = 1.0 + 0.3 * np.sin(tp * X + 1.0) * np.cos(tp * Y)
    u, v, w = 0.3 * np.sin(tp * Y), 0.2 * np.sin(tp * Z), 0.25 * np.sin(tp * X)
    A = 0.3 + 0.1 * np.sin(tp * (Y + Z))
    B = 0.2 + 0.1 * np.sin(tp * (X + Z))
    Cf = 0.25 + 0.1 * np.cos(tp * (X + Y))
    pr = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2 < 0.04, 5.0, 1.0)
    U[..., ID] = rho
    U[..., IU], U[..., IV], U[..., IW] = rho * u, rho * v, rho * w
    U[..., IA], U[..., IB], U[..., IC] = A, B, Cf
    U[..., IP] = pr / (op.gamma0 - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (A * A + B * B + Cf * Cf)
    return op, U


def run_pair(oracle, ekp, R0, nsteps, use_step=False):
    op = oracle.from_ek(ekp)
    R = R0.copy()
    t_o, it_o, _ = oracle.mhd3d_run(op, R, nsteps, 1e30)
    ctx = native.Context(ekp)
    try:
        ctx.upload(R0)
        if use_step:
            t = 0.0
            for _ in range(nsteps):
                dt = ekp.cfl / ctx.inv_dt()
                ctx.step(dt)
                t += dt
            it = nsteps
        else:
            t, it, _ = ctx.run(nsteps, nsteps, 1e30)
        G = ctx.download()
    finally:
        ctx.close()
    assert it == it_o and t == t_o
    return G[3:-3, 3:-3, 3:-3], R[3:-3, 3:-3, 3:-3]


def rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * np.abs(b).max())))


@pytest.mark.parametrize("rs,st,bc", [("hlld", "2", (3,) * 6), ("hlld", "1", (1, 2, 3, 3, 2, 1)),
                                      ("hll", "2", (2, 2, 1, 1, 3, 3)), ("llf", "2", (3, 3, 3, 3, 1, 1))])
def test_mhd3d_fp64_parity(oracle, rs, st, bc):
    ekp = m3(18, 14, 22, riemann=rs, slope_type=st, bc=bc)
    _, R0 = smooth_state(oracle, ekp)
    g, o = run_pair(oracle, ekp, R0, 6)
    print(f"case=mhd3d 18x14x22 rs={rs} st={st} bc={bc}: max|diff|={np.abs(g - o).max():.3e}", flush=True)
    # FP64 parity is bitwise: same IEEE operations in the same order as the oracle
    assert np.array_equal(g, o), rel(g, o)


def test_mhd3d_oblique_alfven_wave_bitwise(oracle):
    # the oracle's second-order oblique Alfven wave (test_oracle_mhd3d.py), run to the time it
    # returns to its initial state, on the GPU: bitwise the oracle's result
    from test_oracle_mhd3d import oblique_alfven
    ekp = m3(24, 24, 24)
    op, R0, T = oblique_alfven(oracle, 24)
    assert (op.nx, op.gamma0, op.cfl) == (oracle.from_ek(ekp).nx, oracle.from_ek(ekp).gamma0, ekp.cfl)
    R = R0.copy()
    t_o, it_o, _ = oracle.mhd3d_run(op, R, 100000, T)
    ctx = native.Context(ekp)
    try:
        ctx.upload(R0)
        t, it, _ = ctx.run(100000, 100000, T)
        G = ctx.download()
    finally:
        ctx.close()
    assert (t, it) == (t_o, it_o)
    print(f"case=mhd3d oblique Alfven 24^3 to T={T:.6f} ({it} steps)", flush=True)
    assert np.array_equal(G[3:-3, 3:-3, 3:-3], R[3:-3, 3:-3, 3:-3])


def test_mhd3d_blast_host_driven(oracle):
    ekp = m3(20, 20, 20)
    U = problems.mhd_blast3d(ekp, B0=0.6)
    ctx = native.Context(ekp)
    try:
        ctx.upload(U, native.EK_LAYOUT_SOA)
        R0 = ctx.download()
    finally:
        ctx.close()
    g, o = run_pair(oracle, ekp, R0, 5, use_step=True)
    assert np.array_equal(g, o), rel(g, o)


def test_mhd3d_divb_conservation_64():
    ekp = m3(64, 64, 64)
    U0 = problems.mhd_blast3d(ekp, B0=0.8)
    ctx = native.Context(ekp)
    try:
        ctx.upload(U0, native.EK_LAYOUT_SOA)
        b0 = ctx.budgets()
        ctx.run(40, 40, 1e30)
        b1 = ctx.budgets()
        U = ctx.download(native.EK_LAYOUT_SOA)  # [v][k][j][i]
    finally:
        ctx.close()
    A, B, Cf = U[IA], U[IB], U[IC]
    d = ((np.roll(A, -1, 2) - A) / ekp.dx + (np.roll(B, -1, 1) - B) / ekp.dy + (np.roll(Cf, -1, 0) - Cf) / ekp.dz)
    print(f"case=mhd3d blast 64^3 40 steps: max|divB|*dx={np.abs(d).max() * ekp.dx:.3e}", flush=True)
    assert np.abs(d).max() <= 1e-12 * np.abs(U[IA:]).max() / ekp.dx
    for v in (ID, IP, IU, IV, IW, IA, IB, IC):
        assert abs(b1[v] - b0[v]) <= 1e-12 * max(abs(b0[v]), np.abs(U0[v]).sum(), 1.0), v


def test_mhd3d_fp32_l1():
    out = []
    for prec in (native.EK_FP64, native.EK_FP32):
        ekp = m3(48, 48, 48, precision=prec)
        ctx = native.Context(ekp)
        try:
            ctx.upload(problems.mhd_blast3d(ekp, B0=0.5), native.EK_LAYOUT_SOA)
            ctx.run(20, 20, 1e30)
            out.append(ctx.download(native.EK_LAYOUT_SOA))
        finally:
            ctx.close()
    l1 = np.abs(out[1][ID] - out[0][ID]).mean() / np.abs(out[0][ID]).mean()
    print(f"case=mhd3d FP32 vs FP64 48^3: rel L1(rho)={l1:.3e}", flush=True)
    assert np.isfinite(out[1]).all() and l1 <= 1e-4


@pytest.mark.parametrize("bc", [(1, 2, 3, 3, 2, 1), (2, 2, 1, 1, 3, 3)])
def test_mhd3d_fp32_ragged_mixed_bc_l1(bc):
    # the FP32 pipeline on a ragged grid with reflecting / outflow / periodic sides: relative
    # L1 of every variable against the FP64 pipeline (bitwise the oracle's) within 1e-4
    out = []
    for prec in (native.EK_FP64, native.EK_FP32):
        ekp = m3(29, 21, 17, bc=bc, precision=prec)
        ctx = native.Context(ekp)
        try:
            ctx.upload(problems.mhd_blast3d(ekp, B0=0.5), native.EK_LAYOUT_SOA)
            t, it, _ = ctx.run(20, 20, 1e30)
            out.append(ctx.download(native.EK_LAYOUT_SOA))
        finally:
            ctx.close()
        assert it == 20
    assert np.isfinite(out[1]).all()
    for v in range(8):
        l1 = np.abs(out[1][v] - out[0][v]).mean() / np.abs(out[0][v]).mean()
        print(f"case=mhd3d FP32 bc={bc} var={v}: rel L1 {l1:.3e}", flush=True)
        assert l1 <= 1e-4
