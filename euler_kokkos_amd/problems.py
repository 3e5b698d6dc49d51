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


def soa_from_reference(p, R: np.ndarray) -> np.ndarray:
    """Interior of a reference-layout array U(i,j[,k],var) as the device SoA U[var][k][j][i]."""
    g = p.gw
    if p.dim == 2:
        return np.ascontiguousarray(R[g:-g, g:-g, :].transpose(2, 1, 0))
    return np.ascontiguousarray(R[g:-g, g:-g, g:-g, :].transpose(3, 2, 1, 0))
