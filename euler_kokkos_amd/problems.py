"""Initial conditions for Python callers (bench, batching, serving), in the device SoA layout.

Each builder evaluates the reference's initialisation functor at the interior cell centres
with the same floating-point expressions, so the state is identical to what the
reference's InitXxxFunctor writes (ghost cells are not stored on the device).
Parameters are given as the reference's problem sections would hold them (strings or
numbers, float-parsed like ConfigMap::getFloat).
"""
from __future__ import annotations

import numpy as np

from .params import getfloat


def _centres(p):
    """Cell-centre coordinates x = xmin + dx/2 + (i-gw+gw)*dx (HydroInitFunctors2D.h:79-80)."""
    xs = [p.xmin + p.dx / 2 + np.arange(p.nx, dtype=np.float64) * p.dx,
          p.ymin + p.dy / 2 + np.arange(p.ny, dtype=np.float64) * p.dy]
    if p.dim == 3:
        xs.append(p.zmin + p.dz / 2 + np.arange(p.nz, dtype=np.float64) * p.dz)
    return xs


def _grid(p):
    xs = _centres(p)
    if p.dim == 2:
        y, x = np.meshgrid(xs[1], xs[0], indexing="ij")
        return x, y, None
    z, y, x = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    return x, y, z


def blast(p, radius="0.1", center_x="0.5", center_y="0.5", center_z="0.5", density_in="1.0",
          density_out="1.2", pressure_in="10.0", pressure_out="0.1") -> np.ndarray:
    """InitBlastFunctor2D/3D (src/muscl/HydroInitFunctors2D.h:127-209, 3D.h:173-265)."""
    r = getfloat(radius)
    cx, cy, cz = getfloat(center_x), getfloat(center_y), getfloat(center_z)
    din, dout = getfloat(density_in), getfloat(density_out)
    pin, pout = getfloat(pressure_in), getfloat(pressure_out)
    x, y, z = _grid(p)
    d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy)
    if p.dim == 3:
        d2 = d2 + (z - cz) * (z - cz)
    inside = d2 < r * r
    U = np.zeros((p.nbvar,) + x.shape)
    U[0] = np.where(inside, din, dout)
    U[1] = np.where(inside, pin / (p.gamma0 - 1.0), pout / (p.gamma0 - 1.0))
    return U


def implode(p, density_outer="1.0", pressure_outer="1.0", density_inner="0.125",
            pressure_inner="0.14", shape_region=0) -> np.ndarray:
    """InitImplodeFunctor2D/3D with zero velocities (src/muscl/HydroInitFunctors2D.h:33-122,
    3D.h:68-168)."""
    ro, po = getfloat(density_outer), getfloat(pressure_outer)
    ri, pi = getfloat(density_inner), getfloat(pressure_inner)
    x, y, z = _grid(p)
    if p.dim == 2:
        if shape_region == 1:
            out = (x + y * y > 0.5) & (x + y * y < 1.5)
        else:
            out = x + y > (p.xmin + p.xmax) / 2.0 + p.ymin
    else:
        if shape_region == 1:
            out = (x + y + z > 0.5) & (x + y + z < 2.5)
        else:
            out = x + y + z > (p.xmin + p.xmax) / 2.0 + p.ymin + p.zmin
    U = np.zeros((p.nbvar,) + x.shape)
    U[0] = np.where(out, ro, ri)
    U[1] = np.where(out, po / (p.gamma0 - 1.0) + 0.5 * ro * 0.0, pi / (p.gamma0 - 1.0) + 0.5 * ri * 0.0)
    return U


def orszag_tang(p) -> np.ndarray:
    """InitOrszagTangFunctor2D (src/muscl/MHDInitFunctors2D.h:345-479), 2D MHD.

    Evaluated on the full ghost-extended grid exactly as the functor does (the energy pass
    reads the face fields of the right and top neighbours, ghosts included), with libm's
    sin through ``math.sin``; the interior is returned in the device SoA layout."""
    import math
    g = p.gw
    isz, jsz = p.nx + 2 * g, p.ny + 2 * g
    twopi = 2 * math.pi
    gamma0 = p.gamma0
    B0 = 1.0 / math.sqrt(2 * twopi)
    p0 = gamma0 / (2 * twopi)
    d0 = gamma0 * p0
    xs = [p.xmin + p.dx / 2 + (i - g) * p.dx for i in range(isz)]
    ys = [p.ymin + p.dy / 2 + (j - g) * p.dy for j in range(jsz)]
    sy = np.array([math.sin(y * twopi) for y in ys])
    sx = np.array([math.sin(x * twopi) for x in xs])
    s2x = np.array([math.sin(2.0 * x * twopi) for x in xs])
    U = np.zeros((8, jsz, isz))
    U[0] = d0
    U[2] = (-d0 * 1.0 * sy)[:, None]
    U[3] = (d0 * 1.0 * sx)[None, :]
    U[5] = (-B0 * sy)[:, None]
    U[6] = (B0 * s2x)[None, :]
    two_pi_e = 4.0 * math.asin(1.0)
    p0e = gamma0 / (2.0 * two_pi_e)
    d, u, v, a, b = U[0, :-1, :-1], U[2, :-1, :-1], U[3, :-1, :-1], U[5], U[6]
    bx = (a[:-1, :-1] + a[:-1, 1:]) / 2
    by = (b[:-1, :-1] + b[1:, :-1]) / 2
    U[1, :-1, :-1] = p0e / (gamma0 - 1.0) + 0.5 * ((u * u) / d + (v * v) / d) + 0.5 * (bx * bx + by * by)
    return np.ascontiguousarray(U[:, g:-g, g:-g])


def mhd_blast3d(p, B0=1.0, radius="0.1", center=("0.5", "0.5", "0.5"), density="1.0",
                pressure_in="10.0", pressure_out="0.1") -> np.ndarray:
    """3D MHD blast with a uniform field B = B0 (1,1,1)/sqrt(3) on every face: rho = 1, v = 0,
    p = pressure_in inside the sphere, pressure_out outside; E = p/(gamma-1) + B^2/2.
    (The reference's InitBlastFunctor3D_MHD reads its field from BlastParams, which leaves
    bz undefined, src/shared/problems/BlastParams.h:47-52; this builder states B fully.)"""
    r = getfloat(radius)
    cx, cy, cz = (getfloat(c) for c in center)
    x, y, z = _grid(p)
    d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)
    inside = d2 < r * r
    b = float(B0) / np.sqrt(3.0)
    U = np.zeros((8,) + x.shape)
    U[0] = getfloat(density)
    U[5] = U[6] = U[7] = b
    pin, pout = getfloat(pressure_in), getfloat(pressure_out)
    U[1] = np.where(inside, pin, pout) / (p.gamma0 - 1.0) + 0.5 * (b * b + b * b + b * b)
    return U


def soa_from_reference(p, R: np.ndarray) -> np.ndarray:
    """Interior of a reference-layout array U(i,j[,k],var) as the device SoA U[var][k][j][i]."""
    g = p.gw
    if p.dim == 2:
        return np.ascontiguousarray(R[g:-g, g:-g, :].transpose(2, 1, 0))
    return np.ascontiguousarray(R[g:-g, g:-g, g:-g, :].transpose(3, 2, 1, 0))


def reference_from_soa(p, S: np.ndarray) -> np.ndarray:
    """Device SoA U[var][k][j][i] as a reference-layout array U(i,j[,k],var) with zero ghost
    cells (a solver fills them before it reads them)."""
    g = p.gw
    if p.dim == 2:
        R = np.zeros((p.nx + 2 * g, p.ny + 2 * g, p.nbvar))
        R[g:-g, g:-g, :] = S.transpose(2, 1, 0)
    else:
        R = np.zeros((p.nx + 2 * g, p.ny + 2 * g, p.nz + 2 * g, p.nbvar))
        R[g:-g, g:-g, g:-g, :] = S.transpose(3, 2, 1, 0)
    return R
