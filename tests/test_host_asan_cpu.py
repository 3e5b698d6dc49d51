"""Host layer under AddressSanitizer and UndefinedBehaviorSanitizer (g++, no device code):
the .ini parser (src/utils/config/ConfigMap.cpp), HydroParams (src/shared/HydroParams.cpp:37-180),
the initial conditions (HydroInitFunctors2D/3D.h, MHDInitFunctors2D/3D.h), the VTI writer
(src/utils/io/IO_VTK.cpp:25-410) and the HDF5 writer / restart reader
(src/utils/io/IO_HDF5.h:76-525, 1527-2143) driven by tests/host_io_driver.cpp on small
grids of every solver kind, ascii and binary VTK. Any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "euler_kokkos_amd", "host")
H5LIB = os.path.join(ROOT, "euler_kokkos_amd", "libek_h5.so")
UNITS = ["ConfigMap.cpp", "HydroParams.cpp", "Problems.cpp", "IO_VTK.cpp", "IO_HDF5.cpp"]

INIS = {
    "implode2d": """[run]
solver_name=Hydro_Muscl_2D
[mesh]
nx=37
ny=21
xmin=0.0
xmax=1.0
ymin=0.0
ymax=1.0
boundary_type_xmin=1
boundary_type_xmax=2
boundary_type_ymin=3
boundary_type_ymax=3
[hydro]
gamma0=1.4
problem=implode
[output]
outputPrefix=imp
outputVtkAscii={ascii}
hdf5_enabled=true
""",
    "blast3d": """[run]
solver_name=Hydro_Muscl_3D
[mesh]
nx=12
ny=9
nz=7
boundary_type_xmin=3
boundary_type_xmax=3
boundary_type_ymin=3
boundary_type_ymax=3
boundary_type_zmin=3
boundary_type_zmax=3
[hydro]
gamma0=1.4
problem=blast
[blast]
radius=0.2
[output]
outputPrefix=b3
outputVtkAscii={ascii}
""",
    "orszag_tang": """[run]
solver_name=MHD_Muscl_2D
[mesh]
nx=16
ny=11
boundary_type_xmin=3
boundary_type_xmax=3
boundary_type_ymin=3
boundary_type_ymax=3
[hydro]
gamma0=1.666
problem=orszag_tang
riemann=hlld
[output]
outputPrefix=ot
outputVtkAscii={ascii}
""",
    "mhd_blast3d": """[run]
solver_name=MHD_Muscl_3D
[mesh]
nx=9
ny=8
nz=6
boundary_type_xmin=3
boundary_type_xmax=3
boundary_type_ymin=3
boundary_type_ymax=3
boundary_type_zmin=3
boundary_type_zmax=3
[hydro]
gamma0=1.666
problem=blast
riemann=hlld
[blast]
radius=0.25
[output]
outputPrefix=m3
outputVtkAscii={ascii}
outputHdf5CompressionLevel=4
""",
}


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("asan")
    exe = str(d / "host_io_driver")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", HOST, "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "host_io_driver.cpp")] + [os.path.join(HOST, u) for u in UNITS] + \
          ["-o", exe, "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr or "").lower() and "cannot find" in r.stderr:
        pytest.skip("the sanitizer runtime is not installed")
    assert r.returncode == 0, r.stderr[-3000:]
    # the HDF5 reader/writer library is looked up next to the code that loads it
    if os.path.exists(H5LIB):
        os.symlink(H5LIB, str(d / "libek_h5.so"))
    return exe


@pytest.mark.parametrize("ascii", ["true", "false"])
@pytest.mark.parametrize("case", sorted(INIS))
def test_host_io_clean_under_sanitizers(driver, tmp_path, case, ascii):
    ini = tmp_path / f"{case}.ini"
    ini.write_text(INIS[case].format(ascii=ascii) + f"outputDir={tmp_path}\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")
    r = subprocess.run([driver, str(ini)], capture_output=True, text=True, cwd=str(tmp_path), env=env, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.stdout.startswith("ok "), r.stdout
    vti = [f for f in os.listdir(tmp_path) if f.endswith(".vti")]
    assert len(vti) == 1 and os.path.getsize(tmp_path / vti[0]) > 0
    if os.path.exists(H5LIB):
        assert r.stdout.strip().endswith("h5")
        assert any(f.endswith(".h5") for f in os.listdir(tmp_path))
        assert any(f.endswith(".xmf") for f in os.listdir(tmp_path))
