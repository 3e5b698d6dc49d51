"""The .ini-driven host layer (ConfigMap -> HydroParams -> SolverFactory -> HIP solver) on
the GPU, against the oracle run of the same parameters; and the executable's VTI output."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from euler_kokkos_amd import native
from euler_kokkos_amd.params import make_params

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

INI_2D = """[run]
solver_name=Hydro_Muscl_2D
tEnd=0.02
nStepmax=1000
nOutput=0
nlog=7
[mesh]
nx=96
ny=64
xmax=1.5
boundary_type_xmin=1
boundary_type_xmax=2
boundary_type_ymin=3
boundary_type_ymax=3
[hydro]
gamma0=1.666
cfl=0.8
slope_type=2
problem=blast
riemann=hllc
[blast]
radius=0.2
center_x=0.6
center_y=0.4
density_in=1.0
density_out=1.2
"""


def ini_run(path, n):
    L = native.lib()
    L.ek_ini_run.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_longlong, C.POINTER(C.c_double),
                             C.POINTER(C.c_int)]
    L.ek_host_last_error.restype = C.c_char_p
    out = np.zeros(n)
    t, it = C.c_double(), C.c_int()
    rc = L.ek_ini_run(str(path).encode(), out.ctypes.data_as(C.POINTER(C.c_double)), n, C.byref(t),
                      C.byref(it))
    assert rc == 0, L.ek_host_last_error()
    return out, t.value, it.value


def test_ini_blast_2d_matches_oracle(tmp_path, oracle):
    ini = tmp_path / "b.ini"
    ini.write_text(INI_2D)
    ekp = make_params("Hydro_Muscl_2D", 96, 64, xmax="1.5", gamma0="1.666", cfl="0.8",
                      slope_type="2", riemann="hllc", bc=(1, 2, 3, 3, 1, 1))
    op = oracle.from_ek(ekp)
    U = oracle.blast(op, radius=float(np.float32(0.2)), center_x=float(np.float32(0.6)),
                     center_y=float(np.float32(0.4)), center_z=0.5, density_in=1.0,
                     density_out=float(np.float32(1.2)), pressure_in=10.0,
                     pressure_out=float(np.float32(0.1)))
    t_o, it_o, _ = oracle.run(op, U, 1000, float(np.float32(0.02)))
    G, t, it = ini_run(ini, 4 * 96 * 64)
    assert it == it_o and t == t_o
    want = U[2:-2, 2:-2, :].transpose(2, 1, 0).reshape(-1)
    np.testing.assert_array_equal(G, want)


def test_executable_writes_reference_vti(tmp_path, oracle):
    exe = os.path.join(ROOT, "euler_kokkos_amd", "bin", "euler_kokkos_amd")
    ini = tmp_path / "o.ini"
    ini.write_text(INI_2D.replace("nOutput=0", "nOutput=-1").replace("nStepmax=1000", "nStepmax=3")
                   + f"[output]\noutputDir={tmp_path}\noutputPrefix=blast\n")
    r = subprocess.run([exe, str(ini)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Perf (interior)" in r.stdout
    files = sorted(p for p in os.listdir(tmp_path) if p.endswith(".vti"))
    assert len(files) >= 2
    raw = open(tmp_path / files[-1], "rb").read()
    cut = raw.index(b"<AppendedData encoding=\"raw\">")
    head, data = raw[:cut], raw[cut:]
    assert b'WholeExtent="0 96 0 64 0 0"' in head
    blob = data[data.index(b"_") + 1:]
    arrays = []
    off = 0
    for _ in range(4):
        nbytes = int(np.frombuffer(blob[off:off + 8], dtype=np.uint64)[0])
        arrays.append(np.frombuffer(blob[off + 8:off + 8 + nbytes], dtype=np.float64))
        off += 8 + nbytes
    rho = arrays[0].reshape(64, 96)
    assert rho.min() > 0 and np.isfinite(rho).all()


INI_MHD = """[run]
solver_name=MHD_Muscl_2D
tEnd=1e30
nStepmax=25
nOutput=0
nlog=10
[mesh]
nx=60
ny=44
boundary_type_xmin=3
boundary_type_xmax=3
boundary_type_ymin=3
boundary_type_ymax=3
[hydro]
gamma0=1.666
cfl=0.8
slope_type=2
problem=orszag_tang
riemann=hlld
smallr=1e-7
smallc=1e-7
"""


def test_ini_orszag_tang_mhd2d_matches_oracle(tmp_path, oracle):
    ini = tmp_path / "ot.ini"
    ini.write_text(INI_MHD)
    ekp = make_params("MHD_Muscl_2D", 60, 44, gamma0="1.666", cfl="0.8", slope_type="2",
                      riemann="hlld", smallr="1e-7", smallc="1e-7", bc=(3,) * 6)
    op = oracle.from_ek(ekp)
    U = oracle.orszag_tang(op)
    t_o, it_o, _ = oracle.mhd_run(op, U, 25, 1e30)
    G, t, it = ini_run(ini, 8 * 60 * 44)
    assert it == it_o == 25 and t == t_o
    want = U[3:-3, 3:-3, :].transpose(2, 1, 0).reshape(-1)
    np.testing.assert_array_equal(G, want)


INI_MHD3D = """[run]
solver_name=MHD_Muscl_3D
tEnd=1e30
nStepmax=6
nOutput=0
[mesh]
nx=14
ny=12
nz=16
boundary_type_xmin=3
boundary_type_xmax=3
boundary_type_ymin=1
boundary_type_ymax=2
boundary_type_zmin=3
boundary_type_zmax=3
[hydro]
gamma0=1.666
cfl=0.8
slope_type=2
problem=blast
riemann=hlld
[blast]
radius=0.25
bx_in=0.4
bz_in=0.7
bx_out=0.3
bz_out=0.2
"""


def test_ini_mhd3d_blast_matches_oracle(tmp_path, oracle):
    ini = tmp_path / "m3.ini"
    ini.write_text(INI_MHD3D)
    ekp = make_params("MHD_Muscl_3D", 14, 12, 16, gamma0="1.666", cfl="0.8", slope_type="2",
                      riemann="hlld", bc=(3, 3, 1, 2, 3, 3))
    op = oracle.from_ek(ekp)
    f = lambda s: float(np.float32(s))  # ConfigMap::getFloat
    U = np.zeros(oracle.ref_shape(op))
    g = 3
    ax = [lo + d / 2 + (np.arange(n) - g) * d for n, d, lo in ((op.isize, op.dx, op.xmin),
                                                                 (op.jsize, op.dy, op.ymin),
                                                                 (op.ksize, op.dz, op.zmin))]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    inside = (X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2 < f("0.25") * f("0.25")
    a = np.where(inside, f("0.4"), f("0.3"))
    b = np.where(inside, f("0.7"), f("0.2"))
    U[..., 0] = np.where(inside, 1.0, f("1.2"))
    U[..., 5], U[..., 6] = a, b
    U[..., 1] = np.where(inside, 10.0, f("0.1")) / (op.gamma0 - 1.0) + 0.5 * (a * a + b * b + 0.0)
    t_o, it_o, _ = oracle.mhd3d_run(op, U, 6, 1e30)
    G, t, it = ini_run(ini, 8 * 14 * 12 * 16)
    assert it == it_o == 6 and t == t_o
    want = U[3:-3, 3:-3, 3:-3, :].transpose(3, 2, 1, 0).reshape(-1)
    np.testing.assert_array_equal(G, want)
