// ek_geom_host.h -- host-side construction of the kernels' geometry and gas constants from
// an ek_params (shared by the single-grid context and the job pool).
#pragma once

#include <cmath>

#include "ek_hip.h"
#include "ek_hydro.h"

namespace ek {

inline bool is_pow2(double x)
{
  if (!(x > 0.0) || !std::isfinite(x))
    return false;
  int e;
  const double m = std::frexp(x, &e);
  return m == 0.5;
}

template <class T>
inline void fill_geom(const ek_params& p, ek::Geom<T>& g, ek::Gas<T>& gas)
{
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = p.dim == 3 ? p.nz : 1;
  for (int i = 0; i < 6; i++)
    g.bc[i] = p.bc[i];
  g.plane = (long long)p.nx * p.ny;
  g.vstride = g.plane * g.nz;
  g.dx = p.dx;
  g.dy = p.dy;
  g.dz = p.dz;
  g.tdx = (T)p.dx;
  g.tdy = (T)p.dy;
  g.tdz = (T)p.dz;
  g.pow2 = is_pow2(p.dx) && is_pow2(p.dy) && (p.dim == 2 || is_pow2(p.dz));
  g.idx = (T)(1.0 / p.dx);
  g.idy = (T)(1.0 / p.dy);
  g.idz = (T)(1.0 / p.dz);
  gas.gamma0 = (T)p.gamma0;
  gas.gamma6 = (T)p.gamma6;
  gas.gm1 = (T)(p.gamma0 - 1.0);
  gas.smallr = (T)p.smallr;
  gas.smallc = (T)p.smallc;
  gas.smallp = (T)p.smallp;
  gas.smallpp = (T)p.smallpp;
  gas.smallc2 = (T)(p.smallc * p.smallc);
  gas.entho = (T)(1.0 / (p.gamma0 - 1.0));
}

} // namespace ek
