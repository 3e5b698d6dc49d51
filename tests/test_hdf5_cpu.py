"""HDF5 output/restart files (libek_h5.so) checked with the HDF5 tools' own reader (h5dump):
the structure Save_HDF5 writes (src/utils/io/IO_HDF5.h:270-500) -- one chunked, shuffled,
deflated FP64 dataset per variable in C order, scalar int / double root attributes and a
variable-length creation date -- and the reader's round trip and shape check."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "euler_kokkos_amd")
H5LIB = os.path.join(PKG, "libek_h5.so")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(not os.path.exists(H5LIB), reason="libek_h5.so not built (no HDF5 here)")


def lib():
    L = C.CDLL(H5LIB)
    L.ek_h5_write.restype = C.c_int
    L.ek_h5_read.restype = C.c_int
    return L


def write(path, arrays, names, chunk, level=0, step=3, t=0.25, nxyz=(4, 3, 2), ghost=0):
    L = lib()
    rank = arrays[0].ndim
    dims = (C.c_ulonglong * rank)(*arrays[0].shape)
    ch = (C.c_ulonglong * rank)(*chunk)
    nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    ptrs = (C.POINTER(C.c_double) * len(keep))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in keep])
    err = C.create_string_buffer(512)
    rc = L.ek_h5_write(str(path).encode(), rank, dims, ch, level, len(names), nm, ptrs, step, C.c_double(t),
                       nxyz[0], nxyz[1], nxyz[2], ghost, err, 512)
    assert rc == 0, err.value


def read(path, shape, names):
    L = lib()
    rank = len(shape)
    dims = (C.c_ulonglong * rank)(*shape)
    nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
    outs = [np.zeros(shape) for _ in names]
    ptrs = (C.POINTER(C.c_double) * len(outs))(*[o.ctypes.data_as(C.POINTER(C.c_double)) for o in outs])
    step, t = C.c_int(), C.c_double()
    err = C.create_string_buffer(512)
    rc = L.ek_h5_read(str(path).encode(), rank, dims, len(names), nm, ptrs, C.byref(step), C.byref(t), err, 512)
    return rc, outs, step.value, t.value, err.value.decode()


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not found")
def test_writer_structure_read_by_h5dump(tmp_path):
    nz, ny, nx = 2, 3, 4
    rng = np.random.default_rng(5)
    arrs = [rng.standard_normal((nz, ny, nx)) for _ in range(5)]
    names = ["rho", "energy", "rho_vx", "rho_vy", "rho_vz"]
    f = tmp_path / "out_0000003.h5"
    write(f, arrs, names, chunk=(nz, ny, nx), level=6, step=3, t=0.25, nxyz=(nx, ny, nz))
    txt = subprocess.run([H5DUMP, "-p", "-H", str(f)], capture_output=True, text=True, check=True).stdout
    for n in names:
        blk = txt[txt.index(f'DATASET "{n}"'):]
        blk = blk[:blk.index("}\n   }") + 1] if "}\n   }" in blk else blk
        assert "H5T_IEEE_F64LE" in blk
        assert f"SIMPLE {{ ( {nz}, {ny}, {nx} ) / ( {nz}, {ny}, {nx} ) }}" in blk
        assert f"CHUNKED ( {nz}, {ny}, {nx} )" in blk
        assert "PREPROCESSING SHUFFLE" in blk and "COMPRESSION DEFLATE { LEVEL 6 }" in blk
    attrs = subprocess.run([H5DUMP, "-A", "-g", "/", str(f)], capture_output=True, text=True, check=True).stdout

    def attr(name):
        i = attrs.index(f'ATTRIBUTE "{name}"')
        return attrs[i:attrs.index("}\n   }", i)]
    for name, typ, val in (("time step", "H5T_STD_I32LE", "3"), ("total time", "H5T_IEEE_F64LE", "0.25"),
                           ("nx", "H5T_STD_I32LE", str(nx)), ("ny", "H5T_STD_I32LE", str(ny)),
                           ("nz", "H5T_STD_I32LE", str(nz)), ("ghost zone included", "H5T_STD_I32LE", "0")):
        a = attr(name)
        assert typ in a and "SCALAR" in a and re.search(r"\(0\): " + re.escape(val) + r"\b", a), (name, a)
    d = attr("creation date")
    assert "H5T_VARIABLE" in d and "SIMPLE { ( 1 ) / ( 1 ) }" in d
    assert re.search(r'"\d{4}-\d\d-\d\d \d\d:\d\d:\d\d', d)
    # values, in C order (z, y, x), through h5dump's own reader
    vals = subprocess.run([H5DUMP, "-d", "/rho_vy", "-m", "%.17g", "-y", "-w", "0", str(f)], capture_output=True,
                          text=True, check=True).stdout
    body = vals[vals.index("DATA {") + 6:vals.index("}", vals.index("DATA {"))]
    got = np.array([float(x) for x in body.replace(",", " ").split()])
    assert np.array_equal(got, arrs[3].ravel())


def test_read_roundtrip_and_shape_check(tmp_path):
    ny, nx = 5, 7
    a, b = np.arange(35.0).reshape(ny, nx), np.linspace(-1, 1, 35).reshape(ny, nx)
    f = tmp_path / "r.h5"
    write(f, [a, b], ["rho", "energy"], chunk=(ny, nx), step=12, t=1.5, nxyz=(nx, ny, 1))
    rc, (ra, rb), step, t, _ = read(f, (ny, nx), ["rho", "energy"])
    assert rc == 0 and step == 12 and t == 1.5
    assert np.array_equal(ra, a) and np.array_equal(rb, b)
    rc, _, _, _, msg = read(f, (ny + 1, nx), ["rho"])
    assert rc != 0 and "shape" in msg
    rc, _, _, _, msg = read(f, (ny, nx), ["bx"])
    assert rc != 0 and "bx" in msg
    rc, _, _, _, msg = read(tmp_path / "missing.h5", (ny, nx), ["rho"])
    assert rc != 0 and "cannot open" in msg
