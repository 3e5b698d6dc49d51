// Problems.cpp -- see Problems.h.
#include "Problems.h"

#include <cmath>

namespace euler_kokkos {

namespace {
struct Shape {
  int nx, ny, nz, nv;
  long long cells;
  long long at(int v, int i, int j, int k) const { return ((long long)v * nz + k) * ny * (long long)nx + (long long)j * nx + i; }
};
Shape shape_of(const HydroParams& p)
{
  Shape s;
  s.nx = p.nx;
  s.ny = p.ny;
  s.nz = p.dimType == THREE_D ? p.nz : 1;
  s.nv = p.nbvar;
  s.cells = (long long)s.nx * s.ny * s.nz;
  return s;
}
} // namespace

void init_implode(const HydroParams& p, const ConfigMap& c, std::vector<double>& U)
{
  const double rho_out = c.getFloat("implode", "density_outer", 1.0);
  const double p_out = c.getFloat("implode", "pressure_outer", 1.0);
  const double u_out = c.getFloat("implode", "vx_outer", 0.0);
  const double v_out = c.getFloat("implode", "vy_outer", 0.0);
  const double w_out = c.getFloat("implode", "vz_outer", 0.0);
  const double rho_in = c.getFloat("implode", "density_inner", 0.125);
  const double p_in = c.getFloat("implode", "pressure_inner", 0.14);
  const double u_in = c.getFloat("implode", "vx_inner", 0.0);
  const double v_in = c.getFloat("implode", "vy_inner", 0.0);
  const double w_in = c.getFloat("implode", "vz_inner", 0.0);
  const int shape = (int)c.getInteger("implode", "shape_region", 0);
  const Shape s = shape_of(p);
  U.assign((size_t)(s.cells * s.nv), 0.0);
  const int gw = p.ghostWidth;
  const double gamma0 = p.settings.gamma0;
  for (int k = 0; k < s.nz; k++)
    for (int j = 0; j < s.ny; j++)
      for (int i = 0; i < s.nx; i++) {
        // x = xmin + dx/2 + (i_padded - ghostWidth)*dx with i_padded = i + gw
        const double x = p.xmin + p.dx / 2 + ((i + gw) - gw) * p.dx;
        const double y = p.ymin + p.dy / 2 + ((j + gw) - gw) * p.dy;
        bool out;
        if (p.dimType == TWO_D) {
          if (shape == 1)
            out = x + y * y > 0.5 && x + y * y < 1.5;
          else
            out = x + y > (p.xmin + p.xmax) / 2. + p.ymin;
          U[s.at(ID, i, j, k)] = out ? rho_out : rho_in;
          U[s.at(IP, i, j, k)] = out ? p_out / (gamma0 - 1.0) + 0.5 * rho_out * (u_out * u_out + v_out * v_out)
                                     : p_in / (gamma0 - 1.0) + 0.5 * rho_in * (u_in * u_in + v_in * v_in);
          U[s.at(IU, i, j, k)] = out ? u_out : u_in;
          U[s.at(IV, i, j, k)] = out ? v_out : v_in;
        } else {
          const double z = p.zmin + p.dz / 2 + ((k + gw) - gw) * p.dz;
          if (shape == 1)
            out = x + y + z > 0.5 && x + y + z < 2.5;
          else
            out = x + y + z > (p.xmin + p.xmax) / 2. + p.ymin + p.zmin;
          U[s.at(ID, i, j, k)] = out ? rho_out : rho_in;
          U[s.at(IP, i, j, k)] =
            out ? p_out / (gamma0 - 1.0) + 0.5 * rho_out * (u_out * u_out + v_out * v_out + w_out * w_out)
                : p_in / (gamma0 - 1.0) + 0.5 * rho_in * (u_in * u_in + v_in * v_in + w_in * w_in);
          // the reference stores the velocity in the momentum slots (harmless at rest)
          U[s.at(IU, i, j, k)] = out ? u_out : u_in;
          U[s.at(IV, i, j, k)] = out ? v_out : v_in;
          U[s.at(IW, i, j, k)] = out ? w_out : w_in;
        }
      }
}

void init_blast(const HydroParams& p, const ConfigMap& c, std::vector<double>& U)
{
  const double xmin = c.getFloat("mesh", "xmin", 0.0), ymin = c.getFloat("mesh", "ymin", 0.0);
  const double zmin = c.getFloat("mesh", "zmin", 0.0), xmax = c.getFloat("mesh", "xmax", 1.0);
  const double ymax = c.getFloat("mesh", "ymax", 1.0), zmax = c.getFloat("mesh", "zmax", 1.0);
  const double radius = c.getFloat("blast", "radius", (xmin + xmax) / 2.0 / 10);
  const double cx = c.getFloat("blast", "center_x", (xmin + xmax) / 2);
  const double cy = c.getFloat("blast", "center_y", (ymin + ymax) / 2);
  const double cz = c.getFloat("blast", "center_z", (zmin + zmax) / 2);
  const double din = c.getFloat("blast", "density_in", 1.0);
  const double dout = c.getFloat("blast", "density_out", 1.2);
  const double pin = c.getFloat("blast", "pressure_in", 10.0);
  const double pout = c.getFloat("blast", "pressure_out", 0.1);
  const Shape s = shape_of(p);
  U.assign((size_t)(s.cells * s.nv), 0.0);
  const int gw = p.ghostWidth;
  const double gamma0 = p.settings.gamma0;
  const double radius2 = radius * radius;
  for (int k = 0; k < s.nz; k++)
    for (int j = 0; j < s.ny; j++)
      for (int i = 0; i < s.nx; i++) {
        const double x = p.xmin + p.dx / 2 + ((i + gw) - gw) * p.dx;
        const double y = p.ymin + p.dy / 2 + ((j + gw) - gw) * p.dy;
        double d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
        if (p.dimType == THREE_D) {
          const double z = p.zmin + p.dz / 2 + ((k + gw) - gw) * p.dz;
          d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
        }
        const bool in = d2 < radius2;
        U[s.at(ID, i, j, k)] = in ? din : dout;
        U[s.at(IP, i, j, k)] = in ? pin / (gamma0 - 1.0) : pout / (gamma0 - 1.0);
      }
}

// InitOrszagTangFunctor2D (src/muscl/MHDInitFunctors2D.h:345-479): both passes run over the
// ghost-extended grid as in the reference (the energy pass reads the face fields of the
// right and top neighbours, ghosts included); the interior is kept.
void init_orszag_tang(const HydroParams& p, const ConfigMap& c, std::vector<double>& U)
{
  (void)c;
  const int gw = p.ghostWidth;
  const int isz = p.nx + 2 * gw, jsz = p.ny + 2 * gw;
  const double TWOPI = 2 * 3.141592653589793238462643383279502884;
  const double gamma0 = p.settings.gamma0;
  const double B0 = 1.0 / std::sqrt(2 * TWOPI);
  const double p0 = gamma0 / (2 * TWOPI);
  const double d0 = gamma0 * p0;
  const double v0 = 1.0;
  std::vector<double> R((size_t)isz * jsz * 8, 0.0);
  auto at = [&](int i, int j, int v) -> double& { return R[((size_t)i * jsz + j) * 8 + v]; };
  for (int i = 0; i < isz; i++)
    for (int j = 0; j < jsz; j++) {
      const double xPos = p.xmin + p.dx / 2 + (i - gw) * p.dx;
      const double yPos = p.ymin + p.dy / 2 + (j - gw) * p.dy;
      at(i, j, ID) = d0;
      at(i, j, IU) = -d0 * v0 * std::sin(yPos * TWOPI);
      at(i, j, IV) = d0 * v0 * std::sin(xPos * TWOPI);
      at(i, j, IW) = 0.0;
      at(i, j, IA) = -B0 * std::sin(yPos * TWOPI);
      at(i, j, IB) = B0 * std::sin(2.0 * xPos * TWOPI);
      at(i, j, IC) = 0.0;
    }
  const double TwoPi = 4.0 * std::asin(1.0);
  const double p0e = gamma0 / (2.0 * TwoPi);
  for (int i = 0; i < isz - 1; i++)
    for (int j = 0; j < jsz - 1; j++) {
      const double bx = (at(i, j, IA) + at(i + 1, j, IA)) / 2;
      const double by = (at(i, j, IB) + at(i, j + 1, IB)) / 2;
      at(i, j, IP) = p0e / (gamma0 - 1.0) +
                     0.5 * (at(i, j, IU) * at(i, j, IU) / at(i, j, ID) + at(i, j, IV) * at(i, j, IV) / at(i, j, ID)) +
                     0.5 * (bx * bx + by * by);
    }
  const Shape s = shape_of(p);
  U.assign((size_t)(s.cells * 8), 0.0);
  for (int j = 0; j < s.ny; j++)
    for (int i = 0; i < s.nx; i++)
      for (int v = 0; v < 8; v++)
        U[s.at(v, i, j, 0)] = at(i + gw, j + gw, v);
}

// InitBlastFunctor3D_MHD (src/muscl/MHDInitFunctors3D.h:267-371) with BlastParams
// (src/shared/problems/BlastParams.h:28-53). BlastParams reads the by keys and then
// overwrites by_in / by_out with the bz_in / bz_out keys, leaving bz_in / bz_out unset; this
// build keeps that key mapping (By takes the bz keys) and sets Bz to 0.
void init_blast_mhd3d(const HydroParams& p, const ConfigMap& c, std::vector<double>& U)
{
  const double xmin = c.getFloat("mesh", "xmin", 0.0), ymin = c.getFloat("mesh", "ymin", 0.0);
  const double zmin = c.getFloat("mesh", "zmin", 0.0), xmax = c.getFloat("mesh", "xmax", 1.0);
  const double ymax = c.getFloat("mesh", "ymax", 1.0), zmax = c.getFloat("mesh", "zmax", 1.0);
  const double radius = c.getFloat("blast", "radius", (xmin + xmax) / 2.0 / 10);
  const double cx = c.getFloat("blast", "center_x", (xmin + xmax) / 2);
  const double cy = c.getFloat("blast", "center_y", (ymin + ymax) / 2);
  const double cz = c.getFloat("blast", "center_z", (zmin + zmax) / 2);
  const double din = c.getFloat("blast", "density_in", 1.0);
  const double dout = c.getFloat("blast", "density_out", 1.2);
  const double pin = c.getFloat("blast", "pressure_in", 10.0);
  const double pout = c.getFloat("blast", "pressure_out", 0.1);
  const double bxi = c.getFloat("blast", "bx_in", 0.5), byi = c.getFloat("blast", "bz_in", 0.5);
  const double bxo = c.getFloat("blast", "bx_out", 0.5), byo = c.getFloat("blast", "bz_out", 0.5);
  const double bzi = 0.0, bzo = 0.0;
  const Shape s = shape_of(p);
  U.assign((size_t)(s.cells * 8), 0.0);
  const double gamma0 = p.settings.gamma0;
  const double radius2 = radius * radius;
  for (int k = 0; k < s.nz; k++)
    for (int j = 0; j < s.ny; j++)
      for (int i = 0; i < s.nx; i++) {
        const double x = p.xmin + p.dx / 2 + i * p.dx;
        const double y = p.ymin + p.dy / 2 + j * p.dy;
        const double z = p.zmin + p.dz / 2 + k * p.dz;
        const double d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
        const bool in = d2 < radius2;
        const double a = in ? bxi : bxo, b = in ? byi : byo, cc = in ? bzi : bzo;
        U[s.at(ID, i, j, k)] = in ? din : dout;
        U[s.at(IA, i, j, k)] = a;
        U[s.at(IB, i, j, k)] = b;
        U[s.at(IC, i, j, k)] = cc;
        U[s.at(IP, i, j, k)] = (in ? pin : pout) / (gamma0 - 1.0) + 0.5 * (a * a + b * b + cc * cc);
      }
}

bool init_problem_mhd(const std::string& problem, const HydroParams& params, const ConfigMap& configMap,
                      std::vector<double>& U)
{
  if (problem == "blast" && params.dimType == THREE_D) {
    init_blast_mhd3d(params, configMap, U);
    return true;
  }
  if (problem == "orszag_tang" && params.dimType == TWO_D) {
    init_orszag_tang(params, configMap, U);
    return true;
  }
  return false;
}

bool init_problem(const std::string& problem, const HydroParams& params, const ConfigMap& configMap,
                  std::vector<double>& U)
{
  if (problem == "implode") {
    init_implode(params, configMap, U);
    return true;
  }
  if (problem == "blast") {
    init_blast(params, configMap, U);
    return true;
  }
  return false;
}

} // namespace euler_kokkos
