"""The .ini-driven host layer (ConfigMap -> HydroParams -> SolverFactory -> HIP solver) on
the GPU, against the oracle run of the same parameters; and the executable's VTI output."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from euler_kokkos_amd import native, problems
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


def mhd3d_blast_oracle_ic(oracle, op):
    """The state InitBlastFunctor3D_MHD makes from INI_MHD3D's [blast] section (BlastParams
    reads the bz keys into by and leaves bz at zero), every real through strtof."""
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
    return U


MHD3D_EKP = dict(gamma0="1.666", cfl="0.8", slope_type="2", riemann="hlld", bc=(3, 3, 1, 2, 3, 3))


def test_ini_mhd3d_blast_matches_oracle(tmp_path, oracle):
    ini = tmp_path / "m3.ini"
    ini.write_text(INI_MHD3D)
    ekp = make_params("MHD_Muscl_3D", 14, 12, 16, **MHD3D_EKP)
    op = oracle.from_ek(ekp)
    U = mhd3d_blast_oracle_ic(oracle, op)
    t_o, it_o, _ = oracle.mhd3d_run(op, U, 6, 1e30)
    G, t, it = ini_run(ini, 8 * 14 * 12 * 16)
    assert it == it_o == 6 and t == t_o
    want = U[3:-3, 3:-3, 3:-3, :].transpose(3, 2, 1, 0).reshape(-1)
    np.testing.assert_array_equal(G, want)


# ---------------------------------------------------------------- VTI output content
INI_3D = """[run]
solver_name=Hydro_Muscl_3D
tEnd=1e30
nStepmax=4
nOutput=0
[mesh]
nx=20
ny=16
nz=12
boundary_type_xmin=1
boundary_type_xmax=2
boundary_type_ymin=3
boundary_type_ymax=3
boundary_type_zmin=2
boundary_type_zmax=1
[hydro]
gamma0=1.4
cfl=0.5
slope_type=2
problem=blast
riemann=hllc
[blast]
radius=0.3
center_x=0.4
center_y=0.5
center_z=0.55
density_in=1.0
density_out=1.0
pressure_in=10.0
pressure_out=0.1
"""

NAMES = ["rho", "energy", "rho_vx", "rho_vy", "rho_vz", "bx", "by", "bz"]


def parse_vti(path):
    """(extent, [(name, values)]) of a VTI file the reference's save_VTK_2D/3D layout:
    ascii DataArrays, or raw appended data (a UInt64 byte count before each array) whose
    offsets the DataArray elements give."""
    import re
    raw = open(path, "rb").read()
    ascii_ = b'format="ascii"' in raw
    cut = len(raw) if ascii_ else raw.index(b"<AppendedData")
    head = raw[:cut].decode()
    extent = [int(v) for v in re.search(r'WholeExtent="([^"]*)"', head).group(1).split()]
    out = []
    if ascii_:
        for m in re.finditer(r'<DataArray type="Float64" Name="([^"]+)" format="ascii" >\n(.*?)\n    </DataArray>',
                             head, re.S):
            out.append((m.group(1), m.group(2).split()))
        return extent, out
    blob = raw[raw.index(b"_", cut) + 1:]
    for m in re.finditer(r'<DataArray type="Float64" Name="([^"]+)" format="appended" offset="(\d+)" />', head):
        off = int(m.group(2))
        nbytes = int(np.frombuffer(blob[off:off + 8], dtype=np.uint64)[0])
        out.append((m.group(1), np.frombuffer(blob[off + 8:off + 8 + nbytes], dtype=np.float64)))
    return extent, out


def check_vti(path, want_soa, ascii_):
    """The file holds the interior cells of every variable, x fastest, in the reference's
    variable order: bitwise in binary, as C++'s default 6-digit formatting in ascii."""
    nv = want_soa.shape[0]
    extent, arrays = parse_vti(path)
    dims = want_soa.shape[1:][::-1]  # nx, ny[, nz]
    assert extent == [0, dims[0], 0, dims[1], 0, dims[2] if len(dims) == 3 else 0]
    assert [a[0] for a in arrays] == NAMES[:nv]
    for v, (name, vals) in enumerate(arrays):
        flat = want_soa[v].reshape(-1)
        if ascii_:
            assert vals == ["%g" % x for x in flat], name
        else:
            assert np.array_equal(vals, flat), name


def write_outputs(tmp_path, ini_text, tag, ascii_):
    d = tmp_path / tag
    d.mkdir()
    ini = tmp_path / f"{tag}.ini"
    ini.write_text(ini_text.replace("nOutput=0", "nOutput=-1") + f"[output]\noutputDir={d}\noutputPrefix=o\n"
                   + ("outputVtkAscii=true\n" if ascii_ else ""))
    run_exe(ini, tmp_path)
    files = sorted(p for p in os.listdir(d) if p.endswith(".vti"))
    return d / files[0], d / files[-1]


@pytest.mark.parametrize("ascii_", [False, True])
def test_vti_hydro_2d_content(tmp_path, oracle, ascii_):
    ini = INI_2D.replace("nStepmax=1000", "nStepmax=3")
    first, last = write_outputs(tmp_path, ini, "h2", ascii_)
    ekp = make_params("Hydro_Muscl_2D", 96, 64, xmax="1.5", gamma0="1.666", cfl="0.8",
                      slope_type="2", riemann="hllc", bc=(1, 2, 3, 3, 1, 1))
    op = oracle.from_ek(ekp)
    f = lambda s: float(np.float32(s))
    U = oracle.blast(op, radius=f("0.2"), center_x=f("0.6"), center_y=f("0.4"), center_z=0.5, density_in=1.0,
                     density_out=f("1.2"), pressure_in=10.0, pressure_out=f("0.1"))
    check_vti(first, problems.soa_from_reference(ekp, U), ascii_)
    _, it, _ = oracle.run(op, U, 3, f("0.02"))
    assert it == 3
    check_vti(last, problems.soa_from_reference(ekp, U), ascii_)


@pytest.mark.parametrize("ascii_", [False, True])
def test_vti_hydro_3d_content(tmp_path, oracle, ascii_):
    first, last = write_outputs(tmp_path, INI_3D, "h3", ascii_)
    ekp = make_params("Hydro_Muscl_3D", 20, 16, 12, gamma0="1.4", cfl="0.5", slope_type="2", riemann="hllc",
                      bc=(1, 2, 3, 3, 2, 1))
    op = oracle.from_ek(ekp)
    f = lambda s: float(np.float32(s))
    U = oracle.blast(op, radius=f("0.3"), center_x=f("0.4"), center_y=0.5, center_z=f("0.55"), density_in=1.0,
                     density_out=1.0, pressure_in=10.0, pressure_out=f("0.1"))
    check_vti(first, problems.soa_from_reference(ekp, U), ascii_)
    _, it, _ = oracle.run(op, U, 4, 1e30)
    assert it == 4
    check_vti(last, problems.soa_from_reference(ekp, U), ascii_)


def test_vti_mhd_2d_content(tmp_path, oracle):
    first, last = write_outputs(tmp_path, INI_MHD.replace("nStepmax=25", "nStepmax=5"), "m2", False)
    ekp = make_params("MHD_Muscl_2D", 60, 44, gamma0="1.666", cfl="0.8", slope_type="2",
                      riemann="hlld", smallr="1e-7", smallc="1e-7", bc=(3,) * 6)
    op = oracle.from_ek(ekp)
    U = oracle.orszag_tang(op)
    check_vti(first, problems.soa_from_reference(ekp, U), False)
    _, it, _ = oracle.mhd_run(op, U, 5, 1e30)
    assert it == 5
    check_vti(last, problems.soa_from_reference(ekp, U), False)


def test_vti_mhd_3d_content(tmp_path, oracle):
    first, last = write_outputs(tmp_path, INI_MHD3D, "m3", False)
    ekp = make_params("MHD_Muscl_3D", 14, 12, 16, **MHD3D_EKP)
    op = oracle.from_ek(ekp)
    U = mhd3d_blast_oracle_ic(oracle, op)
    check_vti(first, problems.soa_from_reference(ekp, U), False)
    _, it, _ = oracle.mhd3d_run(op, U, 6, 1e30)
    assert it == 6
    check_vti(last, problems.soa_from_reference(ekp, U), False)


# ---------------------------------------------------------------- HDF5 output and restart
EXE = os.path.join(ROOT, "euler_kokkos_amd", "bin", "euler_kokkos_amd")
H5LIB = os.path.join(ROOT, "euler_kokkos_amd", "libek_h5.so")
needs_h5 = pytest.mark.skipif(not os.path.exists(H5LIB), reason="libek_h5.so not built (no HDF5 here)")


def h5_read(path, shape, names):
    L = C.CDLL(H5LIB)
    rank = len(shape)
    dims = (C.c_ulonglong * rank)(*shape)
    nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
    outs = [np.zeros(shape) for _ in names]
    ptrs = (C.POINTER(C.c_double) * len(outs))(*[o.ctypes.data_as(C.POINTER(C.c_double)) for o in outs])
    step, t = C.c_int(), C.c_double()
    err = C.create_string_buffer(512)
    rc = L.ek_h5_read(str(path).encode(), rank, dims, len(names), nm, ptrs, C.byref(step), C.byref(t), err, 512)
    assert rc == 0, err.value
    return outs, step.value, t.value


def run_exe(ini, cwd):
    print(f"case=exe {ini.name}", flush=True)
    r = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=300, cwd=str(cwd))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def h5_files(d):
    return sorted(p for p in os.listdir(d) if p.endswith(".h5"))


HYDRO_NAMES = ["rho", "energy", "rho_vx", "rho_vy"]


@needs_h5
def test_hdf5_output_ghosts_and_restart_continue_bitwise(tmp_path):
    """Save_HDF5 / Load_HDF5 through the executable: interior files, ghost-included files
    whose interior is the same and whose ghosts follow the boundary rules, and a restart from
    an intermediate file (with or without ghosts) that ends bitwise on the uninterrupted run."""
    base = INI_2D.replace("nOutput=0", "nOutput=4")
    out = {}
    for name, extra in (("a", ""), ("g", "ghostIncluded=true\n")):
        d = tmp_path / name
        d.mkdir()
        ini = tmp_path / f"{name}.ini"
        ini.write_text(base + f"[output]\noutputDir={d}\noutputPrefix=blast\nvtk_enabled=false\nhdf5_enabled=true\n"
                       + extra)
        run_exe(ini, tmp_path)
        out[name] = h5_files(d)
    assert out["a"] == out["g"] and len(out["a"]) >= 5
    xmf = (tmp_path / "blast.xmf").read_text()
    assert xmf.count('GridType="Uniform"') == len(out["a"]) and "blast_0000002.h5:/rho_vx" in xmf
    for f in out["a"]:
        A, sa, ta = h5_read(tmp_path / "a" / f, (64, 96), HYDRO_NAMES)
        G, sg, tg = h5_read(tmp_path / "g" / f, (68, 100), HYDRO_NAMES)
        assert sa == sg and ta == tg
        for a, g in zip(A, G):
            assert np.array_equal(g[2:-2, 2:-2], a)
        # x min is reflecting: ghost column i = 1 mirrors i = 2, normal momentum negated
        assert np.array_equal(G[0][:, 1], G[0][:, 2]) and np.array_equal(G[2][:, 1], -G[2][:, 2])
        # y min/max periodic
        assert np.array_equal(G[0][0:2, 2:-2], G[0][64:66, 2:-2])
    last = out["a"][-1]
    A_end, s_end, t_end = h5_read(tmp_path / "a" / last, (64, 96), HYDRO_NAMES)
    for src in ("a", "g"):
        d = tmp_path / f"r{src}"
        d.mkdir()
        ini = tmp_path / f"r{src}.ini"
        ghost = "ghostIncluded=true\n" if src == "g" else ""
        ini.write_text(base.replace("[run]\n", f"[run]\nrestart_enabled=1\nrestart_filename={tmp_path / src / out[src][2]}\n")
                       + f"[output]\noutputDir={d}\noutputPrefix=blast\nvtk_enabled=false\nhdf5_enabled=true\n" + ghost)
        log = run_exe(ini, tmp_path)
        assert "This is a restarted run" in log
        files = h5_files(d)
        assert files[0] == out[src][2] and files[-1] == last
        shape = (68, 100) if src == "g" else (64, 96)
        R, s, t = h5_read(d / last, shape, HYDRO_NAMES)
        assert s == s_end and t == t_end
        for r, a in zip(R, A_end):
            assert np.array_equal(r[2:-2, 2:-2] if src == "g" else r, a)


@needs_h5
def test_hdf5_restart_upscale_mhd(tmp_path):
    """[run] restart_upscale: a 30x22 Orszag-Tang file (ghosts included) restarts a 60x44 run;
    every fine cell takes its coarse cell, and a fine face between two coarse faces takes
    their mean (Load_HDF5::copy_buffer, IO_HDF5.h:1593-1650)."""
    names = ["rho", "energy", "rho_vx", "rho_vy", "rho_vz", "bx", "by", "bz"]
    coarse = INI_MHD.replace("nx=60", "nx=30").replace("ny=44", "ny=22").replace("nStepmax=25", "nStepmax=3") \
                    .replace("nOutput=0", "nOutput=1")
    dc = tmp_path / "c"
    dc.mkdir()
    ini = tmp_path / "c.ini"
    ini.write_text(coarse + f"[output]\noutputDir={dc}\noutputPrefix=ot\nvtk_enabled=false\nhdf5_enabled=true\n"
                   "ghostIncluded=true\n")
    run_exe(ini, tmp_path)
    cf = h5_files(dc)[-1]
    Cv, cstep, ct = h5_read(dc / cf, (28, 36), names)
    df = tmp_path / "f"
    df.mkdir()
    ini = tmp_path / "f.ini"
    ini.write_text(INI_MHD.replace("nStepmax=25", "nStepmax=0").replace("nOutput=0", "nOutput=1")
                   .replace("[run]\n", f"[run]\nrestart_enabled=1\nrestart_filename={dc / cf}\nrestart_upscale=true\n")
                   + f"[output]\noutputDir={df}\noutputPrefix=ot\nvtk_enabled=false\nhdf5_enabled=true\n"
                   "ghostIncluded=true\n")
    run_exe(ini, tmp_path)
    F, fstep, ft = h5_read(df / h5_files(df)[0], (50, 66), names)
    assert fstep == cstep and ft == ct
    gw = 3
    J, I = np.meshgrid(np.arange(gw, 44 + gw), np.arange(gw, 60 + gw), indexing="ij")
    jl, il = (J + gw) // 2, (I + gw) // 2
    for v, (c, f) in enumerate(zip(Cv, F)):
        want = c[jl, il]
        if names[v] == "bx":
            odd = (I + gw - 2 * il) != 0
            want = np.where(odd, (c[jl, il] + c[jl, np.minimum(il + 1, 35)]) / 2, want)
        if names[v] == "by":
            odd = (J + gw - 2 * jl) != 0
            want = np.where(odd, (c[jl, il] + c[np.minimum(jl + 1, 27), il]) / 2, want)
        np.testing.assert_array_equal(f[gw:-gw, gw:-gw], want, err_msg=names[v])
