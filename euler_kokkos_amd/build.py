"""Builds the native parts of euler_kokkos_amd in-tree.

* ``euler_kokkos_amd/libek_hip.so`` -- HIP kernels for gfx950 plus the C ABI of
  ``include/ek_hip.h`` and the C++ host layer (ConfigMap / HydroParams / solvers).
* ``euler_kokkos_amd/bin/euler_kokkos_amd`` -- the .ini-driven executable (the
  reference's ``src/main.cpp`` equivalent) linked against that library.

Usage: ``python -m euler_kokkos_amd.build [-j N]``. Every call recompiles.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libek_hip.so")
BIN_DIR = os.path.join(PKG, "bin")
EXE = os.path.join(BIN_DIR, "euler_kokkos_amd")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code; FP64 translation units never
# contract a*b+c into an FMA (bit-exact parity with the reference's CPU build). The FP32
# bulk path allows contraction and the hardware's approximate FP32 divide/sqrt.
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", HOST,
    "-Wno-unused-command-line-argument",
]
# FP32 bulk path: contraction, the hardware's approximate divide/sqrt, and FP32 denormals
# flushed (so a divide is a reciprocal and a multiply, without the denormal rescaling).
FP32_FLAGS = ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
              "-fgpu-flush-denormals-to-zero"]
UNITS = [
    ("ek_fp64.hip", ["-ffp-contract=off"]),
    ("ek_fp32.hip", FP32_FLAGS),
    ("ek_capi.hip", ["-ffp-contract=off"]),
    ("ek_mhd_fp64.hip", ["-ffp-contract=off"]),
    ("ek_mhd_fp32.hip", FP32_FLAGS),
    ("ek_pool_fp64.hip", ["-ffp-contract=off"]),
    ("ek_pool_fp32.hip", FP32_FLAGS),
    ("ek_pool.hip", ["-ffp-contract=off"]),
    ("ek_mhd3d_fp64.hip", ["-ffp-contract=off"]),
    ("ek_mhd3d_fp32.hip", FP32_FLAGS),
]
HOST_UNITS = ["ConfigMap.cpp", "HydroParams.cpp", "SolverBase.cpp", "SolverHydroMuscl.cpp", "SolverMHDMuscl.cpp",
              "SolverFactory.cpp", "Problems.cpp", "IO_VTK.cpp", "IO_HDF5.cpp", "ek_host_capi.cpp"]

# Optional HDF5 output / restart: libek_h5.so (host/ek_h5.cpp) is built against the first
# HDF5 installation found and opened by the host layer on first use.
H5_LIB = os.path.join(PKG, "libek_h5.so")
HDF5_DIRS = [d for d in (os.environ.get("EK_HDF5_DIR"), "/opt/conda", "/usr") if d]


def find_hdf5():
    for d in HDF5_DIRS:
        for inc in (os.path.join(d, "include"), os.path.join(d, "include", "hdf5", "serial")):
            if os.path.exists(os.path.join(inc, "hdf5.h")):
                for lib in (os.path.join(d, "lib"), os.path.join(d, "lib", "x86_64-linux-gnu", "hdf5", "serial"),
                            os.path.join(d, "lib", "x86_64-linux-gnu")):
                    if any(os.path.exists(os.path.join(lib, n)) for n in ("libhdf5.so", "libhdf5.so.103")):
                        return inc, lib
    return None
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
              "-I", os.path.join(ROOT, "include"), "-I", HOST]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN_DIR, exist_ok=True)
    cmds = []
    objs = []
    for src, extra in UNITS:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        cmds.append([HIPCC, *COMMON, *extra, "-c", "-o", o, os.path.join(CSRC, src)])
        objs.append(o)
    host_srcs = [h for h in HOST_UNITS if os.path.exists(os.path.join(HOST, h))]
    for src in host_srcs:
        o = os.path.join(OBJ, "host_" + src.replace(".cpp", ".o"))
        cmds.append(["g++", *HOST_FLAGS, "-c", "-o", o, os.path.join(HOST, src)])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB, *objs, "-ldl"])
    h5 = find_hdf5()
    if h5:
        inc, lib = h5
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared", "-I", inc, "-o", H5_LIB,
              os.path.join(HOST, "ek_h5.cpp"), f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}"])
    elif os.path.exists(H5_LIB):
        os.remove(H5_LIB)
    main_cpp = os.path.join(HOST, "main.cpp")
    if os.path.exists(main_cpp):
        _run(["g++", *HOST_FLAGS, "-o", EXE, main_cpp, LIB,
              f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=4)
    a = ap.parse_args()
    build(a.j, verbose=True)
