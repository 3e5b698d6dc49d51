"""ek_run_noalloc (the time loop on caller-owned device buffers, scratch in a fixed device
array) against ek_create/ek_upload/ek_run/ek_download on the same state: same time, same
iteration count, same final state bit for bit. Each case makes two calls in a row, the
second on a smaller grid, so the second call starts from scratch the first one left.

The caller's device buffers come from hipMalloc of the HIP runtime the library itself is
linked against (the test is the caller that can allocate; the entry point allocates nothing)."""
import ctypes as C

import numpy as np
import pytest

from euler_kokkos_amd import native, problems
from euler_kokkos_amd.params import make_params

pytestmark = pytest.mark.gpu


def initial_state(ekp):
    if ekp.mhd and ekp.dim == 3:
        return problems.mhd_blast3d(ekp, B0=0.5, radius="0.2", center=("0.45", "0.5", "0.55"))
    if ekp.mhd:
        return problems.orszag_tang(ekp)
    return problems.blast(ekp, radius="0.2", center_x="0.45", center_y="0.5", center_z="0.55")


def with_context(ekp, U0, nsteps, nStepmax, tEnd):
    ctx = native.Context(ekp)
    try:
        ctx.upload(U0, native.EK_LAYOUT_SOA)
        t, it, dt = ctx.run(nsteps, nStepmax, tEnd)
        return ctx.download(native.EK_LAYOUT_SOA), t, it, dt
    finally:
        ctx.close()


def hip_runtime():
    native.lib()  # loads the library and, with it, its HIP runtime
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


H2D, D2H = 1, 2


def without_alloc(ekp, U0, nsteps, nStepmax, tEnd):
    hip = hip_runtime()
    host = np.ascontiguousarray(U0 if ekp.precision == native.EK_FP64 else U0.astype(np.float32))
    nbytes = host.nbytes
    bufs = [C.c_void_p(), C.c_void_p()]
    try:
        for b in bufs:
            assert hip.hipMalloc(C.byref(b), nbytes) == 0
        assert hip.hipMemcpy(bufs[0], host.ctypes.data, nbytes, H2D) == 0
        assert hip.hipMemset(bufs[1], 0xFF, nbytes) == 0  # NaN pattern: the unused buffer must not leak in
        t, it, dt, in2 = native.run_noalloc(ekp, bufs[0].value, bufs[1].value, nsteps, nStepmax, tEnd)
        out = np.empty_like(host)
        assert hip.hipMemcpy(out.ctypes.data, bufs[in2], nbytes, D2H) == 0
    finally:
        for b in bufs:
            if b.value:
                hip.hipFree(b)
    assert in2 == it % 2
    return out.astype(np.float64), t, it, dt


CASES = {
    "hydro3d_fp64": ("Hydro_Muscl_3D", [(40, 36, 28), (24, 20, 16)], native.EK_FP64, "hllc", (3, 3, 1, 1, 2, 2)),
    "hydro3d_fp32": ("Hydro_Muscl_3D", [(36, 32, 30), (20, 16, 12)], native.EK_FP32, "hllc", (3,) * 6),
    "hydro2d_fp64": ("Hydro_Muscl_2D", [(96, 80, 1), (40, 56, 1)], native.EK_FP64, "hllc", (1, 2, 3, 3, 3, 3)),
    "mhd2d_fp64": ("MHD_Muscl_2D", [(64, 64, 1), (32, 24, 1)], native.EK_FP64, "hlld", (3,) * 6),
    "mhd3d_fp64": ("MHD_Muscl_3D", [(20, 18, 16), (12, 10, 14)], native.EK_FP64, "hlld", (3, 3, 1, 1, 3, 3)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_matches_ek_run_on_two_calls(case):
    solver, shapes, prec, rs, bc = CASES[case]
    for call, (nx, ny, nz) in enumerate(shapes):
        gamma = "1.6666667" if "MHD" in solver else "1.4"
        ekp = make_params(solver, nx, ny, nz, gamma0=gamma, cfl="0.6", slope_type="2", riemann=rs, bc=bc,
                          precision=prec)
        U0 = initial_state(ekp)
        # the second call runs into nStepmax (an odd count), so its state ends in U2
        nsteps, nStepmax = (8, 8) if call == 0 else (10, 7)
        print(f"case={case} call={call} grid={nx}x{ny}x{nz} nsteps={nsteps} nStepmax={nStepmax}", flush=True)
        G, t, it, dt = with_context(ekp, U0, nsteps, nStepmax, 1e30)
        N, tn, itn, dtn = without_alloc(ekp, U0, nsteps, nStepmax, 1e30)
        assert it == itn == min(nsteps, nStepmax)
        assert t == tn and dt == dtn
        assert np.array_equal(G, N), f"max diff {np.nanmax(np.abs(G - N))}"


def test_noalloc_rejects_oversized_3d_mhd_before_launching():
    ekp = make_params("MHD_Muscl_3D", 64, 64, 64, gamma0="1.6666667", riemann="hlld", bc=(3,) * 6)
    with pytest.raises(native.EkError, match="too large for ek_run_noalloc"):
        native.run_noalloc(ekp, 1, 2, 4, 4, 1e30)
