"""ctypes binding of the oracle (oracle/libek_oracle.so).

TEST INFRASTRUCTURE: the CPU restatement of the reference step that the HIP path is
checked against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libek_oracle.so")


class EkoParams(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("gw", C.c_int),
        ("nbvar", C.c_int), ("isize", C.c_int), ("jsize", C.c_int), ("ksize", C.c_int),
        ("imin", C.c_int), ("imax", C.c_int), ("jmin", C.c_int), ("jmax", C.c_int),
        ("kmin", C.c_int), ("kmax", C.c_int), ("bc", C.c_int * 6), ("riemann", C.c_int),
        ("mhd", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
        ("zmin", C.c_double), ("zmax", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("dz", C.c_double), ("gamma0", C.c_double), ("gamma6", C.c_double), ("cfl", C.c_double),
        ("slope_type", C.c_double), ("smallr", C.c_double), ("smallc", C.c_double),
        ("smallp", C.c_double), ("smallpp", C.c_double), ("cIso", C.c_double),
    ]


class EkoImplode(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rho_out", "p_out", "u_out", "v_out", "w_out",
                                          "rho_in", "p_in", "u_in", "v_in", "w_in")] + \
               [("shape", C.c_int)]


class EkoBlast(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("radius", "center_x", "center_y", "center_z",
                                          "density_in", "density_out", "pressure_in",
                                          "pressure_out")]


_L = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        _L = _bind(C.CDLL(LIB_PATH))
    return _L


def load(path: str):
    """A separately built copy of the oracle (e.g. one with a slip planted by a test), bound
    like the oracle itself; make the module's functions use it with `using`."""
    return _bind(C.CDLL(path))


@contextlib.contextmanager
def using(L):
    """Within the block, every function of this module calls library L instead of the oracle."""
    global _L
    old = _L
    _L = L
    try:
        yield L
    finally:
        _L = old


def _bind(L):
    """Sets the argument and result types of every oracle entry point on L."""
    P = C.POINTER
    D = P(C.c_double)
    L.eko_params_init.argtypes = [P(EkoParams)]
    L.eko_compute_primitives.argtypes = [P(EkoParams), D, D, D]
    L.eko_slope_scalar.argtypes = [P(EkoParams), C.c_double, C.c_double, C.c_double, D]
    L.eko_trace_along_dir.argtypes = [P(EkoParams), D, D, D, D, C.c_double, C.c_double,
                                      C.c_double, C.c_int, D]
    L.eko_riemann_hydro.argtypes = [P(EkoParams), D, D, D]
    L.eko_init_implode.argtypes = [P(EkoParams), P(EkoImplode), D]
    L.eko_init_blast.argtypes = [P(EkoParams), P(EkoBlast), D]
    L.eko_make_boundaries.argtypes = [P(EkoParams), D]
    L.eko_compute_inv_dt.argtypes = [P(EkoParams), D]
    L.eko_compute_inv_dt.restype = C.c_double
    L.eko_godunov_unsplit_impl.argtypes = [P(EkoParams), D, D, D, D, D, D, C.c_double]
    L.eko_run.argtypes = [P(EkoParams), D, C.c_int, C.c_double, D, P(C.c_int), D]
    L.eko_run.restype = C.c_int
    L.eko_num_threads.restype = C.c_int
    A8 = C.POINTER(C.c_double * 8)
    L.eko_mhd_constoprim.argtypes = [P(EkoParams), D, D, D, D]
    L.eko_mhd_speed_info.argtypes = [P(EkoParams), D, D]
    L.eko_riemann_mhd.argtypes = [P(EkoParams), D, D, D]
    L.eko_compute_emfz.argtypes = [P(EkoParams), A8]
    L.eko_compute_emfz.restype = C.c_double
    L.eko_mhd_trace_2d.argtypes = [P(EkoParams), D, D, C.c_double, C.c_double, D, D, D]
    L.eko_mhd_init_orszag_tang.argtypes = [P(EkoParams), D]
    L.eko_mhd_convert_to_primitives.argtypes = [P(EkoParams), D, D]
    L.eko_mhd_inv_dt.argtypes = [P(EkoParams), D]
    L.eko_mhd_inv_dt.restype = C.c_double
    L.eko_mhd_run.argtypes = [P(EkoParams), D, C.c_int, C.c_double, D, P(C.c_int), D]
    L.eko_mhd_run.restype = C.c_int
    L.eko_mhd3d_run.argtypes = [P(EkoParams), D, C.c_int, C.c_double, D, P(C.c_int), D]
    L.eko_mhd3d_run.restype = C.c_int
    L.eko_mhd_run_dt.argtypes = [P(EkoParams), D, C.c_int, C.c_double]
    L.eko_mhd3d_run_dt.argtypes = [P(EkoParams), D, C.c_int, C.c_double]
    L.eko_mhd3d_convert_to_primitives.argtypes = [P(EkoParams), D, D]
    L.eko_mhd3d_inv_dt.argtypes = [P(EkoParams), D]
    L.eko_mhd3d_inv_dt.restype = C.c_double
    L.eko_compute_emf_dir.argtypes = [P(EkoParams), A8, C.c_int]
    L.eko_compute_emf_dir.restype = C.c_double
    return L


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def from_ek(ekp) -> EkoParams:
    """Oracle params from the product's ek_params (same field values)."""
    p = EkoParams()
    for f in ("dim", "nx", "ny", "nz", "gw", "nbvar", "riemann", "mhd", "xmin", "xmax", "ymin",
              "ymax", "zmin", "zmax", "gamma0", "cfl", "slope_type", "smallr", "smallc", "cIso"):
        setattr(p, f, getattr(ekp, f))
    for i in range(6):
        p.bc[i] = ekp.bc[i]
    lib().eko_params_init(C.byref(p))
    return p


def ref_shape(p: EkoParams):
    s = (p.isize, p.jsize) + ((p.ksize,) if p.dim == 3 else ())
    return s + (p.nbvar,)


def implode(p: EkoParams, **kw) -> np.ndarray:
    ip = EkoImplode(rho_out=1.0, p_out=1.0, rho_in=0.125, p_in=float(np.float32(0.14)))
    for k, v in kw.items():
        setattr(ip, k, v)
    U = np.zeros(ref_shape(p))
    lib().eko_init_implode(C.byref(p), C.byref(ip), dp(U))
    return U


def blast(p: EkoParams, **kw) -> np.ndarray:
    bp = EkoBlast()
    for k, v in kw.items():
        setattr(bp, k, v)
    U = np.zeros(ref_shape(p))
    lib().eko_init_blast(C.byref(p), C.byref(bp), dp(U))
    return U


def run(p: EkoParams, U: np.ndarray, nStepmax: int, tEnd: float, t: float = 0.0, it: int = 0):
    """Reference time loop; U is advanced in place. Returns (t, it, last dt)."""
    tt, ii, dd = C.c_double(t), C.c_int(it), C.c_double()
    rc = lib().eko_run(C.byref(p), dp(U), nStepmax, tEnd, C.byref(tt), C.byref(ii), C.byref(dd))
    if rc:
        raise MemoryError("oracle allocation failed")
    return tt.value, ii.value, dd.value


def inv_dt(p: EkoParams, U: np.ndarray) -> float:
    return lib().eko_compute_inv_dt(C.byref(p), dp(U))


def interior(p: EkoParams, U: np.ndarray) -> np.ndarray:
    g = p.gw
    if p.dim == 2:
        return U[g:-g, g:-g, :]
    return U[g:-g, g:-g, g:-g, :]


# ---------------------------------------------------------------- 2D MHD


def orszag_tang(p: EkoParams) -> np.ndarray:
    """InitOrszagTangFunctor2D (src/muscl/MHDInitFunctors2D.h:345-479) on the reference layout."""
    U = np.zeros(ref_shape(p))
    lib().eko_mhd_init_orszag_tang(C.byref(p), dp(U))
    return U


def mhd_run(p: EkoParams, U: np.ndarray, nStepmax: int, tEnd: float, t: float = 0.0, it: int = 0):
    """SolverMHDMuscl<2> time loop; U is advanced in place. Returns (t, it, last dt)."""
    tt, ii, dd = C.c_double(t), C.c_int(it), C.c_double()
    rc = lib().eko_mhd_run(C.byref(p), dp(U), nStepmax, tEnd, C.byref(tt), C.byref(ii),
                           C.byref(dd))
    if rc:
        raise MemoryError("oracle allocation failed")
    return tt.value, ii.value, dd.value


def mhd_primitives(p: EkoParams, U: np.ndarray) -> np.ndarray:
    Q = np.zeros_like(U)
    lib().eko_mhd_convert_to_primitives(C.byref(p), dp(U), dp(Q))
    return Q


def mhd_inv_dt(p: EkoParams, U: np.ndarray) -> float:
    """Max over the interior of the fast speeds of the primitives of U (ghosts as given)."""
    return lib().eko_mhd_inv_dt(C.byref(p), dp(mhd_primitives(p, U)))


def mhd_constoprim(p: EkoParams, u, mag):
    u = np.ascontiguousarray(u, dtype=np.float64)
    mag = np.ascontiguousarray(mag, dtype=np.float64)
    q = np.zeros(8)
    c = C.c_double()
    lib().eko_mhd_constoprim(C.byref(p), dp(u), dp(mag), C.byref(c), dp(q))
    return q, c.value


def mhd_speed_info(p: EkoParams, q) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros(3)
    lib().eko_mhd_speed_info(C.byref(p), dp(q), dp(out))
    return out


def riemann_mhd(p: EkoParams, ql, qr) -> np.ndarray:
    a = np.array(ql, dtype=np.float64)
    b = np.array(qr, dtype=np.float64)
    f = np.zeros(8)
    lib().eko_riemann_mhd(C.byref(p), dp(a), dp(b), dp(f))
    return f


def compute_emfz(p: EkoParams, qedge) -> float:
    """compute_emf<EMFZ>; qedge rows are the RT, RB, LT, LB edge states."""
    q = np.ascontiguousarray(qedge, dtype=np.float64).reshape(4, 8)
    return lib().eko_compute_emfz(C.byref(p), q.ctypes.data_as(C.POINTER(C.c_double * 8)))


def mhd_trace_2d(p: EkoParams, qnb, bfnb, dtdx: float, dtdy: float):
    """trace_unsplit_mhd_2d: qnb [3][3][8], bfnb [4][4][3] -> (qm[2][8], qp[2][8], qEdge[4][8])."""
    qnb = np.ascontiguousarray(qnb, dtype=np.float64).reshape(3, 3, 8)
    bfnb = np.ascontiguousarray(bfnb, dtype=np.float64).reshape(4, 4, 3)
    qm = np.zeros((2, 8))
    qp = np.zeros((2, 8))
    qe = np.zeros((4, 8))
    lib().eko_mhd_trace_2d(C.byref(p), dp(qnb), dp(bfnb), dtdx, dtdy, dp(qm), dp(qp), dp(qe))
    return qm, qp, qe


# ---------------------------------------------------------------- 3D MHD


def mhd3d_run(p: EkoParams, U: np.ndarray, nStepmax: int, tEnd: float, t: float = 0.0, it: int = 0):
    """SolverMHDMuscl<3> time loop; U (reference layout) is advanced in place."""
    tt, ii, dd = C.c_double(t), C.c_int(it), C.c_double()
    rc = lib().eko_mhd3d_run(C.byref(p), dp(U), nStepmax, tEnd, C.byref(tt), C.byref(ii), C.byref(dd))
    if rc:
        raise MemoryError("oracle allocation failed")
    return tt.value, ii.value, dd.value


def mhd3d_inv_dt(p: EkoParams, U: np.ndarray) -> float:
    Q = np.zeros_like(U)
    lib().eko_mhd3d_convert_to_primitives(C.byref(p), dp(U), dp(Q))
    return lib().eko_mhd3d_inv_dt(C.byref(p), dp(Q))


def compute_emf_dir(p: EkoParams, qedge, direction: int) -> float:
    q = np.ascontiguousarray(qedge, dtype=np.float64).reshape(4, 8)
    return lib().eko_compute_emf_dir(C.byref(p), q.ctypes.data_as(C.POINTER(C.c_double * 8)), direction)


def mhd_run_dt(p: EkoParams, U: np.ndarray, nsteps: int, dt: float) -> None:
    """nsteps steps with a fixed dt (2D or 3D by p.dim), U advanced in place."""
    f = lib().eko_mhd3d_run_dt if p.dim == 3 else lib().eko_mhd_run_dt
    if f(C.byref(p), dp(U), nsteps, dt):
        raise MemoryError("oracle allocation failed")
