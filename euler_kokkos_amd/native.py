"""ctypes binding of the C ABI in ``include/ek_hip.h`` (``libek_hip.so``).

The library is loaded from the package directory (built in-tree by
``python -m euler_kokkos_amd.build``). There is no Python or CPU fallback: if the
library is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libek_hip.so")

EK_FP64, EK_FP32 = 0, 1
EK_LAYOUT_REFERENCE, EK_LAYOUT_SOA = 0, 1
RIEMANN = {"approx": 0, "llf": 1, "hll": 2, "hllc": 3, "hlld": 4}

# Exported symbols declared by include/ek_hip.h (checked by the CPU test-suite).
EXPORTS = ("ek_create", "ek_destroy", "ek_last_error", "ek_upload", "ek_download", "ek_inv_dt",
           "ek_step", "ek_run", "ek_budgets", "ek_sync", "ek_device_bytes", "ek_last_run_time",
           "ek_host_alloc", "ek_host_free", "ek_device_info", "ek_run_noalloc")
# size of the fixed device array behind ek_run_noalloc (EK_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 16 << 20


class EkParams(C.Structure):
    """Mirror of ``ek_params`` (include/ek_hip.h)."""

    _fields_ = [
        ("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("gw", C.c_int), ("nbvar", C.c_int), ("mhd", C.c_int), ("riemann", C.c_int),
        ("bc", C.c_int * 6), ("precision", C.c_int), ("reserved", C.c_int),
        ("xmin", C.c_double), ("ymin", C.c_double), ("zmin", C.c_double),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("gamma0", C.c_double), ("gamma6", C.c_double), ("cfl", C.c_double),
        ("slope_type", C.c_double), ("smallr", C.c_double), ("smallc", C.c_double),
        ("smallp", C.c_double), ("smallpp", C.c_double), ("cIso", C.c_double),
    ]


_LIB = None


def lib() -> C.CDLL:
    """Load libek_hip.so once; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run `python -m euler_kokkos_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.ek_create.argtypes = [P(EkParams), P(C.c_void_p)]
    L.ek_create.restype = C.c_int
    L.ek_destroy.argtypes = [C.c_void_p]
    L.ek_destroy.restype = None
    L.ek_last_error.argtypes = [C.c_void_p]
    L.ek_last_error.restype = C.c_char_p
    L.ek_upload.argtypes = [C.c_void_p, P(C.c_double), C.c_int]
    L.ek_download.argtypes = [C.c_void_p, P(C.c_double), C.c_int]
    L.ek_inv_dt.argtypes = [C.c_void_p, P(C.c_double)]
    L.ek_step.argtypes = [C.c_void_p, C.c_double]
    L.ek_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, P(C.c_double), P(C.c_int),
                         P(C.c_double)]
    L.ek_budgets.argtypes = [C.c_void_p, P(C.c_double)]
    L.ek_sync.argtypes = [C.c_void_p]
    L.ek_last_run_time.argtypes = [C.c_void_p, P(C.c_double), P(C.c_int)]
    L.ek_last_run_time.restype = C.c_int
    L.ek_device_bytes.argtypes = [C.c_void_p]
    L.ek_device_bytes.restype = C.c_longlong
    L.ek_host_alloc.argtypes = [C.c_size_t, P(C.c_void_p)]
    L.ek_host_free.argtypes = [C.c_void_p]
    L.ek_device_info.argtypes = [P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), C.c_char_p, C.c_int]
    L.ek_run_noalloc.argtypes = [P(EkParams), C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double,
                                 P(C.c_double), P(C.c_int), P(C.c_double), P(C.c_int), C.c_void_p]
    for n in ("ek_upload", "ek_download", "ek_inv_dt", "ek_step", "ek_run", "ek_budgets", "ek_sync",
              "ek_host_alloc", "ek_host_free", "ek_device_info", "ek_run_noalloc"):
        getattr(L, n).restype = C.c_int
    _LIB = L
    return L


def device_info() -> dict:
    """Compute units, memory clock and bus width, engine clock and name of the current device
    (ek_device_info)."""
    cus, mclk, bw, sclk = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    name = C.create_string_buffer(256)
    rc = lib().ek_device_info(C.byref(cus), C.byref(mclk), C.byref(bw), C.byref(sclk), name, 256)
    if rc != 0:
        raise EkError(f"ek_device_info failed ({rc})")
    return {"cus": cus.value, "mem_clock_khz": mclk.value, "bus_width_bits": bw.value,
            "sclk_khz": sclk.value, "name": name.value.decode(errors="replace")}


def run_noalloc(params: EkParams, U: int, U2: int, nsteps: int, nStepmax: int, tEnd: float,
                t: float = 0.0, it: int = 0, stream: int = 0):
    """ek_run_noalloc on caller-owned device buffers (device addresses U, U2 holding SoA state
    in the params' precision): returns (t, it, dt_last, in_U2), in_U2 saying which buffer
    holds the final state."""
    t_c, it_c, dt_c, in2 = C.c_double(t), C.c_int(it), C.c_double(0.0), C.c_int(0)
    rc = lib().ek_run_noalloc(C.byref(params), C.c_void_p(U), C.c_void_p(U2), nsteps, nStepmax, tEnd,
                              C.byref(t_c), C.byref(it_c), C.byref(dt_c), C.byref(in2), C.c_void_p(stream))
    if rc != 0:
        raise EkError(f"ek_run_noalloc failed ({rc}): {lib().ek_last_error(None).decode()}")
    return t_c.value, it_c.value, dt_c.value, in2.value


class _Pinned:
    """Owner of one page-locked host block (ek_host_alloc); freed with its last array."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        rc = lib().ek_host_alloc(C.c_size_t(nbytes), C.byref(self.ptr))
        if rc != 0:
            raise MemoryError(f"ek_host_alloc({nbytes}) failed ({rc})")

    def __del__(self):
        ptr, self.ptr = self.ptr, C.c_void_p()
        if ptr and _LIB is not None:  # at interpreter exit the binding may already be gone
            _LIB.ek_host_free(ptr)


def pinned_empty(shape, dtype=np.float64) -> np.ndarray:
    """An uninitialised numpy array in page-locked host memory: uploads from it and
    downloads into it run at the link's full speed. The memory is freed when the array and
    every view of it are gone."""
    dt = np.dtype(dtype)
    count = int(np.prod(shape, dtype=np.int64))
    nbytes = max(count * dt.itemsize, 1)
    owner = _Pinned(nbytes)
    buf = (C.c_char * nbytes).from_address(owner.ptr.value)
    buf._owner = owner  # the ctypes buffer keeps the block alive; numpy keeps the buffer
    return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)


def pinned_copy(a: np.ndarray) -> np.ndarray:
    """A C-contiguous copy of `a` in page-locked host memory."""
    out = pinned_empty(a.shape, a.dtype)
    out[...] = a
    return out


class EkError(RuntimeError):
    pass


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One device state (``ek_ctx``): a hydro grid resident in HBM."""

    def __init__(self, params: EkParams):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.ek_create(C.byref(params), C.byref(h))
        if rc != 0:
            raise EkError(f"ek_create failed ({rc}): {self._L.ek_last_error(None).decode()}")
        self.h = h
        self.params = params

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise EkError(f"{what} failed ({rc}): {self._L.ek_last_error(self.h).decode()}")

    def close(self) -> None:
        if self.h:
            self._L.ek_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # shapes --------------------------------------------------------------------------
    def ref_shape(self):
        p = self.params
        s = [p.nx + 2 * p.gw, p.ny + 2 * p.gw]
        if p.dim == 3:
            s.append(p.nz + 2 * p.gw)
        return tuple(s) + (p.nbvar,)

    def soa_shape(self):
        p = self.params
        if p.dim == 2:
            return (p.nbvar, p.ny, p.nx)
        return (p.nbvar, p.nz, p.ny, p.nx)

    # transfers -----------------------------------------------------------------------
    def upload(self, arr: np.ndarray, layout: int = EK_LAYOUT_REFERENCE) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        want = self.ref_shape() if layout == EK_LAYOUT_REFERENCE else self.soa_shape()
        if a.shape != want:
            raise ValueError(f"shape {a.shape} != {want}")
        self._check(self._L.ek_upload(self.h, _dptr(a), layout), "ek_upload")

    def download(self, layout: int = EK_LAYOUT_REFERENCE) -> np.ndarray:
        shape = self.ref_shape() if layout == EK_LAYOUT_REFERENCE else self.soa_shape()
        a = np.empty(shape, dtype=np.float64)
        self._check(self._L.ek_download(self.h, _dptr(a), layout), "ek_download")
        return a

    # compute -------------------------------------------------------------------------
    def inv_dt(self) -> float:
        v = C.c_double()
        self._check(self._L.ek_inv_dt(self.h, C.byref(v)), "ek_inv_dt")
        return v.value

    def step(self, dt: float) -> None:
        self._check(self._L.ek_step(self.h, dt), "ek_step")

    def run(self, nsteps: int, nStepmax: int, tEnd: float, t: float = 0.0, it: int = 0):
        tt, ii, dd = C.c_double(t), C.c_int(it), C.c_double(0.0)
        self._check(self._L.ek_run(self.h, nsteps, nStepmax, tEnd, C.byref(tt), C.byref(ii),
                                   C.byref(dd)), "ek_run")
        return tt.value, ii.value, dd.value

    def budgets(self) -> np.ndarray:
        out = np.zeros(8, dtype=np.float64)
        self._check(self._L.ek_budgets(self.h, _dptr(out)), "ek_budgets")
        return out[: self.params.nbvar]

    def sync(self) -> None:
        self._check(self._L.ek_sync(self.h), "ek_sync")

    def last_run_time(self):
        """(ms, launches) of the step launches of the last run(), from HIP events."""
        ms, n = C.c_double(), C.c_int()
        self._check(self._L.ek_last_run_time(self.h, C.byref(ms), C.byref(n)), "ek_last_run_time")
        return ms.value, n.value

    def device_bytes(self) -> int:
        return int(self._L.ek_device_bytes(self.h))
