"""CPU-side checks of the native library: it loads, exports every entry point declared in
include/ek_hip.h, and rejects invalid configurations before touching a device."""
import ctypes as C
import os
import re

import pytest

from euler_kokkos_amd import native
from euler_kokkos_amd.params import make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?\s\**)(ek\w+)\s*\(", src, re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = native.lib()
    names = declared_symbols()
    assert {"ek_create", "ek_run", "ek_step", "ek_inv_dt"} <= names
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(native.EXPORTS) <= names


@pytest.mark.parametrize("bad", [
    dict(riemann="hll", solver="Hydro_Muscl_3D"),   # reference 3D dispatch has no HLL
    dict(slope_type="1.5"),                          # reference leaves slopes unset
    dict(riemann="hlld"),                            # MHD-only solver
])
def test_create_rejects_unsupported(bad):
    solver = bad.pop("solver", "Hydro_Muscl_2D")
    p = make_params(solver, 16, 16, 16 if solver.endswith("3D") else 1, **bad)
    h = C.c_void_p()
    rc = native.lib().ek_create(C.byref(p), C.byref(h))
    assert rc == 1
    assert native.lib().ek_last_error(None)


@pytest.mark.parametrize("case", ["mhd3d_too_large", "cIso", "mhd_slope", "bc", "tiny", "precision", "dim"])
def test_create_rejects_invalid_before_allocating(case):
    # every check runs on the host, before any device call (so it runs here without a GPU)
    if case in ("mhd3d_too_large", "cIso", "mhd_slope"):
        n = 1200 if case == "mhd3d_too_large" else 16
        p = make_params("MHD_Muscl_3D", n, n, n, riemann="hlld", slope_type="2")
    else:
        p = make_params("Hydro_Muscl_3D", 16, 16, 16)
    want = {"mhd3d_too_large": "4 GiB", "cIso": "isothermal", "mhd_slope": "slope_type",
            "bc": "boundary", "tiny": "ghostWidth", "precision": "precision", "dim": "dim"}[case]
    if case == "cIso":
        p.cIso = 0.5
    elif case == "mhd_slope":
        p.slope_type = 0.0
    elif case == "bc":
        p.bc[3] = 4
    elif case == "tiny":
        p.ny = 1
    elif case == "precision":
        p.precision = 7
    elif case == "dim":
        p.dim = 4
    h = C.c_void_p()
    rc = native.lib().ek_create(C.byref(p), C.byref(h))
    assert rc == 1 and not h.value
    msg = native.lib().ek_last_error(None).decode()
    assert want in msg, msg



def test_noalloc_rejects_on_the_host():
    """ek_run_noalloc refuses invalid calls before touching a device: null buffers, an invalid
    configuration, and a 3D MHD grid whose work area does not fit its fixed scratch (78 reals
    per ghost-padded cell: a 23^3 FP64 grid fits, 24^3 does not)."""
    ok = make_params("Hydro_Muscl_3D", 16, 16, 16, riemann="hllc")
    with pytest.raises(native.EkError, match="null argument"):
        native.run_noalloc(ok, 0, 0, 1, 1, 1e30)
    with pytest.raises(native.EkError, match="nsteps"):
        native.run_noalloc(ok, 1, 2, -1, 1, 1e30)
    bad = make_params("Hydro_Muscl_3D", 16, 16, 16, riemann="hll")
    with pytest.raises(native.EkError, match="riemann"):
        native.run_noalloc(bad, 1, 2, 1, 1, 1e30)
    for n, fits in ((23, True), (24, False)):
        mhd = make_params("MHD_Muscl_3D", n, n, n, gamma0="1.6666667", riemann="hlld", bc=(3,) * 6)
        assert ((n + 6) ** 3 * 78 * 8 <= native.NOALLOC_SCRATCH_BYTES - 49408) == fits
        if not fits:
            with pytest.raises(native.EkError, match="too large for ek_run_noalloc"):
                native.run_noalloc(mhd, 1, 2, 1, 1, 1e30)
