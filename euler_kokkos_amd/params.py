"""HydroParams mirror for Python callers (tests, bench, serving).

Builds the ``ek_params`` POD the C ABI takes, with the reference's conventions:
every real goes through ``ConfigMap::getFloat`` -- parsed by ``strtof`` and widened
to double (src/utils/config/ConfigMap.cpp:37-46) -- and the derived quantities follow
``HydroParams::init`` (src/shared/HydroParams.cpp:326-359).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util

from .native import EK_FP64, EkParams, RIEMANN

_libc = C.CDLL(ctypes.util.find_library("c"))
_libc.strtof.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]
_libc.strtof.restype = C.c_float


def getfloat(x) -> float:
    """Value the reference's ConfigMap::getFloat returns for `x` (a string or a number)."""
    s = x if isinstance(x, str) else repr(float(x))
    return float(_libc.strtof(s.encode(), None))


SOLVERS = {
    "Hydro_Muscl_2D": dict(dim=2, nbvar=4, gw=2, mhd=0),
    "Hydro_Muscl_3D": dict(dim=3, nbvar=5, gw=2, mhd=0),
    "MHD_Muscl_2D": dict(dim=2, nbvar=8, gw=3, mhd=1),
    "MHD_Muscl_3D": dict(dim=3, nbvar=8, gw=3, mhd=1),
}


def make_params(solver_name: str, nx: int, ny: int, nz: int = 1, *, xmin="0", xmax="1", ymin="0",
                ymax="1", zmin="0", zmax="1", gamma0="1.4", cfl="0.5", slope_type="1.0",
                smallr="1e-10", smallc="1e-10", riemann: str = "approx", bc=(1, 1, 1, 1, 1, 1),
                precision: int = EK_FP64) -> EkParams:
    """ek_params for `solver_name`; defaults are the reference's HydroParams defaults."""
    s = SOLVERS[solver_name]
    p = EkParams()
    p.dim, p.nbvar, p.gw, p.mhd = s["dim"], s["nbvar"], s["gw"], s["mhd"]
    p.nx, p.ny, p.nz = nx, ny, nz
    p.riemann = RIEMANN[riemann]
    for i, b in enumerate(bc):
        p.bc[i] = int(b)
    p.precision = precision
    p.xmin, p.xmax = getfloat(xmin), getfloat(xmax)
    p.ymin, p.ymax = getfloat(ymin), getfloat(ymax)
    p.zmin, p.zmax = getfloat(zmin), getfloat(zmax)
    p.dx = (p.xmax - p.xmin) / nx
    p.dy = (p.ymax - p.ymin) / ny
    p.dz = (p.zmax - p.zmin) / nz
    p.gamma0 = getfloat(gamma0)
    p.cfl = getfloat(cfl)
    p.slope_type = getfloat(slope_type)
    p.smallr = getfloat(smallr)
    p.smallc = getfloat(smallc)
    p.smallp = p.smallc * p.smallc / p.gamma0
    p.smallpp = p.smallr * p.smallp
    p.gamma6 = (p.gamma0 + 1.0) / (2.0 * p.gamma0)
    p.cIso = 0.0
    return p
