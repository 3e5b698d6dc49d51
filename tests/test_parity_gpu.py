"""Parity of the HIP path (through the C ABI of libek_hip.so) with the oracle.

FP64 is required to match the oracle cell by cell to 1e-12 relative (the build targets
bit-exact agreement: same IEEE operations, no contraction); FP32 is checked for its L1
error against FP64. Sizes are chosen so the oracle finishes in seconds; the full-size
configurations are checked through size-independent properties (conservation).
"""
import numpy as np
import pytest

from euler_kokkos_amd import native
from euler_kokkos_amd.params import make_params

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def rel_err(a, b):
    scale = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / np.maximum(scale, 1e-12 * np.max(np.abs(b)))))


def run_pair(oracle, ekp, U0, nsteps, tEnd=1e30, use_step=False):
    """Advance the same initial state on the GPU and with the oracle; return both interiors."""
    op = oracle.from_ek(ekp)
    U = U0.copy()
    t_o, it_o, dt_o = oracle.run(op, U, nsteps, tEnd)
    ctx = native.Context(ekp)
    try:
        ctx.upload(U0)
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
    return oracle.interior(op, G), oracle.interior(op, U)


@pytest.mark.parametrize("rs", ["hllc", "approx", "llf", "hll"])
@pytest.mark.parametrize("st", ["1", "2"])
def test_hydro2d_implode_fp64_parity(oracle, rs, st):
    ekp = make_params("Hydro_Muscl_2D", 64, 64, riemann=rs, slope_type=st, cfl="0.8")
    U0 = oracle.implode(oracle.from_ek(ekp))
    g, o = run_pair(oracle, ekp, U0, 20)
    print(f"case=hydro2d implode 64x64 rs={rs} st={st}: max|diff|={np.max(np.abs(g - o)):.3e}",
          flush=True)
    assert rel_err(g, o) <= RTOL


def test_hydro2d_c1_full_fp64_parity(oracle):
    # 2D implosion, 512x512, HLLC, gamma 1.4, CFL 0.8, reflecting walls, 200 steps
    ekp = make_params("Hydro_Muscl_2D", 512, 512, riemann="hllc", slope_type="2", cfl="0.8",
                      gamma0="1.4")
    U0 = oracle.implode(oracle.from_ek(ekp))
    g, o = run_pair(oracle, ekp, U0, 200)
    print(f"case=C1 implode 512^2 200 steps: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert rel_err(g, o) <= RTOL


@pytest.mark.parametrize("dim,rs", [(2, "hllc"), (2, "hll"), (3, "hllc"), (3, "approx")])
def test_first_order_slope_type_0_fp64_parity(oracle, dim, rs):
    # slope_type 0: every limited slope is zero (the reference reaches it through 0 * (q - q-));
    # the kernels skip the limiter. A state at rest outside a blast keeps exact zeros in it.
    if dim == 2:
        ekp = make_params("Hydro_Muscl_2D", 72, 56, riemann=rs, slope_type="0", cfl="0.8", bc=(1, 2, 3, 3))
    else:
        ekp = make_params("Hydro_Muscl_3D", 40, 36, 28, riemann=rs, slope_type="0", cfl="0.5",
                          bc=(3, 3, 1, 2, 2, 1))
    U0 = oracle.blast(oracle.from_ek(ekp), radius=0.2, center_x=0.45, center_y=0.55, center_z=0.5,
                      density_in=1.0, density_out=1.0, pressure_in=10.0, pressure_out=0.1)
    g, o = run_pair(oracle, ekp, U0, 12)
    print(f"case=slope_type 0 dim={dim} rs={rs}: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("rs", ["hllc", "approx", "llf"])
def test_hydro3d_blast_fp64_parity(oracle, rs):
    ekp = make_params("Hydro_Muscl_3D", 32, 32, 32, riemann=rs, slope_type="2", cfl="0.5",
                      bc=(3,) * 6)
    U0 = oracle.blast(oracle.from_ek(ekp), radius=0.2, center_x=0.5, center_y=0.5, center_z=0.5,
                      density_in=1.0, density_out=1.0, pressure_in=10.0, pressure_out=0.1)
    g, o = run_pair(oracle, ekp, U0, 10)
    print(f"case=hydro3d blast 32^3 rs={rs}: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert rel_err(g, o) <= RTOL


@pytest.mark.parametrize("shape,cz", [((256, 128, 60), 8), ((192, 192, 100), 12), ((256, 256, 90), 23)])
def test_hydro3d_z_chunk_lengths_fp64_parity(oracle, shape, cz):
    # grids whose z-chunk length (chosen from the grid size) is 8 or 12 with a partial last
    # chunk, or 23 (24 spread evenly over 4 chunks): the chunking changes where planes are
    # traced, never a result
    ekp = make_params("Hydro_Muscl_3D", *shape, riemann="hllc", slope_type="2", cfl="0.5", bc=(3, 3, 1, 1, 2, 2))
    U0 = oracle.blast(oracle.from_ek(ekp), radius=0.2, center_x=0.45, center_y=0.55, center_z=0.5,
                      density_in=1.0, density_out=1.0, pressure_in=10.0, pressure_out=0.1)
    g, o = run_pair(oracle, ekp, U0, 3)
    print(f"case=hydro3d {shape} (z-chunk {cz}): max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)


def test_hydro3d_ragged_mixed_bc_fp64_parity(oracle):
    # sizes that are not multiples of the tile, every boundary type, non-cubic cells
    ekp = make_params("Hydro_Muscl_3D", 37, 21, 19, xmax="1.3", riemann="hllc", slope_type="2",
                      cfl="0.6", bc=(1, 2, 3, 3, 2, 1))
    U0 = oracle.blast(oracle.from_ek(ekp), radius=0.25, center_x=0.3, center_y=0.6, center_z=0.4,
                      density_in=2.0, density_out=1.0, pressure_in=5.0, pressure_out=0.2)
    g, o = run_pair(oracle, ekp, U0, 12)
    print(f"case=hydro3d ragged 37x21x19: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert rel_err(g, o) <= RTOL


def test_hydro2d_ragged_mixed_bc_fp64_parity(oracle):
    ekp = make_params("Hydro_Muscl_2D", 45, 70, ymax="0.7", riemann="hllc", slope_type="1",
                      cfl="0.7", bc=(2, 1, 3, 3, 1, 1))
    U0 = oracle.blast(oracle.from_ek(ekp), radius=0.2, center_x=0.4, center_y=0.35,
                      density_in=1.0, density_out=1.2, pressure_in=10.0, pressure_out=0.1)
    g, o = run_pair(oracle, ekp, U0, 25)
    assert rel_err(g, o) <= RTOL


def test_host_driven_step_matches_device_loop(oracle):
    ekp = make_params("Hydro_Muscl_2D", 48, 48, riemann="hllc", slope_type="2", cfl="0.8")
    U0 = oracle.implode(oracle.from_ek(ekp))
    g, o = run_pair(oracle, ekp, U0, 9, use_step=True)
    assert rel_err(g, o) <= RTOL


@pytest.mark.parametrize("dim", [2, 3])
def test_run_in_pieces_and_capture_only_calls(oracle, dim):
    # runs of odd and even lengths alternate the start buffer (one cached graph per buffer);
    # nsteps = 0 only captures and changes nothing; the pieces equal one run and the oracle
    if dim == 2:
        ekp = make_params("Hydro_Muscl_2D", 40, 28, riemann="hllc", slope_type="2", cfl="0.8")
        U0 = oracle.implode(oracle.from_ek(ekp))
    else:
        ekp = make_params("Hydro_Muscl_3D", 20, 18, 16, riemann="hllc", slope_type="2", cfl="0.5",
                          bc=(3,) * 6)
        U0 = oracle.blast(oracle.from_ek(ekp), radius=0.2, center_x=0.5, center_y=0.5, center_z=0.5,
                          density_in=1.0, density_out=1.0, pressure_in=10.0, pressure_out=0.1)
    g, o = run_pair(oracle, ekp, U0, 31)
    ctx = native.Context(ekp)
    try:
        ctx.upload(U0)
        t, it = 0.0, 0
        for n in (0, 7, 0, 6, 12, 0, 1, 5):
            t2, it2, _ = ctx.run(n, 31, 1e30, t, it)
            if n == 0:
                assert (t2, it2) == (t, it)
            t, it = t2, it2
        assert it == 31
        G = ctx.download()
    finally:
        ctx.close()
    assert np.array_equal(oracle.interior(oracle.from_ek(ekp), G), g)
    assert rel_err(g, o) <= RTOL


def test_tend_clamp_and_finish(oracle):
    # the device loop clamps the last dt to tEnd and stops (SolverBase::compute_dt/finished)
    ekp = make_params("Hydro_Muscl_2D", 32, 32, riemann="hllc", slope_type="2", cfl="0.8")
    U0 = oracle.implode(oracle.from_ek(ekp))
    g, o = run_pair(oracle, ekp, U0, 40, tEnd=0.013)
    assert rel_err(g, o) <= RTOL


@pytest.mark.parametrize("dim", [2, 3])
def test_inv_dt_and_ghost_download_bitwise(oracle, dim):
    solver = "Hydro_Muscl_2D" if dim == 2 else "Hydro_Muscl_3D"
    ekp = make_params(solver, 20, 18, 17 if dim == 3 else 1, bc=(1, 2, 3, 3, 1, 2))
    op = oracle.from_ek(ekp)
    rng = np.random.default_rng(7)
    U = np.zeros(oracle.ref_shape(op))
    U[..., 0] = rng.uniform(0.5, 2, U.shape[:-1])
    U[..., 2:] = rng.uniform(-1, 1, U.shape[:-1] + (dim,))
    U[..., 1] = rng.uniform(1, 3, U.shape[:-1]) + 0.5 * (U[..., 2:] ** 2).sum(-1) / U[..., 0]
    ctx = native.Context(ekp)
    ctx.upload(U)
    assert ctx.inv_dt() == oracle.inv_dt(op, U)
    G = ctx.download()
    import ctypes as C
    oracle.lib().eko_make_boundaries(C.byref(op), oracle.dp(U))
    np.testing.assert_array_equal(G, U)
    ctx.close()


def test_c2_full_size_fp64_conservation():
    # 3D blast 256^3 periodic (the bench workload): mass, momentum and energy conserved to
    # round-off over 10 steps (periodic box: every flux enters one cell and leaves another)
    ekp = make_params("Hydro_Muscl_3D", 256, 256, 256, riemann="hllc", slope_type="2", cfl="0.5",
                      bc=(3,) * 6)
    ctx = native.Context(ekp)
    n = 256
    x = (np.arange(n) + 0.5) / n
    r2 = (x[:, None, None] - 0.5) ** 2 + (x[None, :, None] - 0.5) ** 2 + (x[None, None, :] - 0.5) ** 2
    U = np.zeros((5, n, n, n))
    U[0] = 1.0
    U[1] = np.where(r2 < 0.01, 10.0, 0.1) / (ekp.gamma0 - 1)
    ctx.upload(U, native.EK_LAYOUT_SOA)
    b0 = ctx.budgets()
    t, it, dt = ctx.run(10, 10, 1e30)
    b1 = ctx.budgets()
    ctx.close()
    assert it == 10 and dt > 0
    scale = np.array([n ** 3, b0[1], n ** 3, n ** 3, n ** 3], float)
    assert np.all(np.abs(b1 - b0) <= 1e-12 * scale), (b0, b1)


def test_fp32_bulk_l1_vs_fp64(oracle):
    ekp64 = make_params("Hydro_Muscl_3D", 48, 48, 48, riemann="hllc", slope_type="2", cfl="0.5",
                        bc=(3,) * 6)
    ekp32 = make_params("Hydro_Muscl_3D", 48, 48, 48, riemann="hllc", slope_type="2", cfl="0.5",
                        bc=(3,) * 6, precision=native.EK_FP32)
    U0 = oracle.blast(oracle.from_ek(ekp64), radius=0.1, center_x=0.5, center_y=0.5, center_z=0.5,
                      density_in=1.0, density_out=1.0, pressure_in=10.0, pressure_out=0.1)
    out = []
    for ekp in (ekp64, ekp32):
        ctx = native.Context(ekp)
        ctx.upload(U0)
        ctx.run(20, 20, 1e30)
        out.append(ctx.download(native.EK_LAYOUT_SOA))
        ctx.close()
    l1_rho = np.abs(out[1][0] - out[0][0]).mean()
    l1_e = np.abs(out[1][1] - out[0][1]).mean() / np.abs(out[0][1]).mean()
    print(f"case=fp32 vs fp64 48^3: L1 rho={l1_rho:.3e} L1 E(rel)={l1_e:.3e}", flush=True)
    assert l1_rho <= 1e-4 and l1_e <= 1e-4


@pytest.mark.parametrize("bc", [(1, 2, 3, 3, 2, 1), (2, 2, 1, 1, 3, 3), (1,) * 6])
def test_fp32_bulk_ragged_mixed_bc_l1(oracle, bc):
    # FP32 bulk on a grid that is not a multiple of the tile or of the z chunk, with
    # reflecting / outflow / periodic sides: relative L1 of every variable against the FP64
    # step (itself bitwise the oracle's on these boundaries) within 1e-4
    kw = dict(riemann="hllc", slope_type="2", cfl="0.6", bc=bc, xmax="1.3")
    ekp64 = make_params("Hydro_Muscl_3D", 37, 21, 19, **kw)
    ekp32 = make_params("Hydro_Muscl_3D", 37, 21, 19, precision=native.EK_FP32, **kw)
    U0 = oracle.blast(oracle.from_ek(ekp64), radius=0.25, center_x=0.3, center_y=0.6, center_z=0.4,
                      density_in=2.0, density_out=1.0, pressure_in=5.0, pressure_out=0.2)
    out = []
    for ekp in (ekp64, ekp32):
        ctx = native.Context(ekp)
        try:
            ctx.upload(U0)
            t, it, _ = ctx.run(25, 25, 1e30)
            out.append(ctx.download(native.EK_LAYOUT_SOA))
        finally:
            ctx.close()
        assert it == 25
    assert np.isfinite(out[1]).all()
    for v in range(5):
        l1 = np.abs(out[1][v] - out[0][v]).mean() / np.abs(out[0][v]).mean()
        print(f"case=hydro3d FP32 bc={bc} var={v}: rel L1 {l1:.3e}", flush=True)
        assert l1 <= 1e-4
