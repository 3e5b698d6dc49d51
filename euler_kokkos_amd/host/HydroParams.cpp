// HydroParams.cpp -- see HydroParams.h.
#include "HydroParams.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>

namespace euler_kokkos {

// src/shared/HydroParams.cpp:15-33
DimensionType get_dim(const ConfigMap& configMap)
{
  const std::string s = configMap.getString("run", "solver_name", "unknown");
  if (s == "Hydro_Muscl_2D" || s == "MHD_Muscl_2D")
    return TWO_D;
  if (s == "Hydro_Muscl_3D" || s == "MHD_Muscl_3D")
    return THREE_D;
  std::cerr << "Solver name not valid : " << s << "\n";
  return TWO_D;
}

// src/shared/HydroParams.cpp:37-180
HydroParams::HydroParams(const ConfigMap& c)
  : nStepmax((int)c.getInteger("run", "nstepmax", 1000))
  , tEnd(c.getFloat("run", "tend", 0.0))
  , nOutput((int)c.getInteger("run", "noutput", 100))
  , enableOutput(nOutput != 0)
  , mhdEnabled(false)
  , nlog((int)c.getInteger("run", "nlog", 10))
  , nx((int)c.getInteger("mesh", "nx", 1))
  , ny((int)c.getInteger("mesh", "ny", 1))
  , nz((int)c.getInteger("mesh", "nz", 1))
  , ghostWidth(2)
  , nbvar(4)
  , dimType(TWO_D)
  , imin(0), imax(0), jmin(0), jmax(0), kmin(0), kmax(0)
  , isize(0), jsize(0), ksize(0)
  , xmin(c.getFloat("mesh", "xmin", 0.0))
  , xmax(c.getFloat("mesh", "xmax", 1.0))
  , ymin(c.getFloat("mesh", "ymin", 0.0))
  , ymax(c.getFloat("mesh", "ymax", 1.0))
  , zmin(c.getFloat("mesh", "zmin", 0.0))
  , zmax(c.getFloat("mesh", "zmax", 1.0))
  , dx(0.0), dy(0.0), dz(0.0)
  , boundary_type_xmin((BoundaryConditionType)c.getInteger("mesh", "boundary_type_xmin", BC_DIRICHLET))
  , boundary_type_xmax((BoundaryConditionType)c.getInteger("mesh", "boundary_type_xmax", BC_DIRICHLET))
  , boundary_type_ymin((BoundaryConditionType)c.getInteger("mesh", "boundary_type_ymin", BC_DIRICHLET))
  , boundary_type_ymax((BoundaryConditionType)c.getInteger("mesh", "boundary_type_ymax", BC_DIRICHLET))
  , boundary_type_zmin((BoundaryConditionType)c.getInteger("mesh", "boundary_type_zmin", BC_DIRICHLET))
  , boundary_type_zmax((BoundaryConditionType)c.getInteger("mesh", "boundary_type_zmax", BC_DIRICHLET))
  , ioVTK(true)
  , ioHDF5(false)
  , settings()
  , niter_riemann((int)c.getInteger("hydro", "niter_riemann", 10))
  , riemannSolverType(RIEMANN_APPROX)
{
  dimType = get_dim(c);
  const std::string solver_name = c.getString("run", "solver_name", "unknown");
  if (solver_name == "Hydro_Muscl_2D") {
    nbvar = 4;
    ghostWidth = 2;
  } else if (solver_name == "Hydro_Muscl_3D") {
    nbvar = 5;
    ghostWidth = 2;
  } else if (solver_name == "MHD_Muscl_2D" || solver_name == "MHD_Muscl_3D") {
    nbvar = 8;
    ghostWidth = 3;
    mhdEnabled = true;
  } else {
    std::cerr << "Solver name not valid : " << solver_name << "\n";
  }

  const std::string rs = c.getString("hydro", "riemann", "approx");
  if (rs == "approx")
    riemannSolverType = RIEMANN_APPROX;
  else if (rs == "llf")
    riemannSolverType = RIEMANN_LLF;
  else if (rs == "hll")
    riemannSolverType = RIEMANN_HLL;
  else if (rs == "hllc")
    riemannSolverType = RIEMANN_HLLC;
  else if (rs == "hlld")
    riemannSolverType = RIEMANN_HLLD;
  else {
    std::cout << "Riemann Solver specified in parameter file is invalid\n";
    std::cout << "Use the default one : approx\n";
    riemannSolverType = RIEMANN_APPROX;
  }

  implementationVersion = (int)c.getFloat("OTHER", "implementationVersion", 0);
  if (implementationVersion != 0 && implementationVersion != 1 && implementationVersion != 2) {
    std::cout << "Implementation version is invalid (must be 0, 1 or 2)\n";
    std::cout << "Use the default : 0\n";
    implementationVersion = 0;
  }

  settings.gamma0 = c.getFloat("hydro", "gamma0", 1.4);
  settings.cfl = c.getFloat("hydro", "cfl", 0.5);
  settings.iorder = (int)c.getInteger("hydro", "iorder", 2);
  settings.slope_type = c.getFloat("hydro", "slope_type", 1.0);
  settings.smallc = c.getFloat("hydro", "smallc", 1e-10);
  settings.smallr = c.getFloat("hydro", "smallr", 1e-10);
  settings.cp = c.getFloat("hydro", "cp", 0.0);
  settings.mu = c.getFloat("hydro", "mu", 0.0);
  settings.kappa = c.getFloat("hydro", "kappa", 0.0);
  const std::string prec = c.getString("hydro", "precision", "fp64");
  precision = (prec == "fp32" || prec == "float" || prec == "single") ? EK_FP32 : EK_FP64;
  init();
}

// src/shared/HydroParams.cpp:326-359
void HydroParams::init()
{
  imin = jmin = kmin = 0;
  imax = nx - 1 + 2 * ghostWidth;
  jmax = ny - 1 + 2 * ghostWidth;
  kmax = nz - 1 + 2 * ghostWidth;
  isize = imax - imin + 1;
  jsize = jmax - jmin + 1;
  ksize = kmax - kmin + 1;
  dx = (xmax - xmin) / nx;
  dy = (ymax - ymin) / ny;
  dz = (zmax - zmin) / nz;
  settings.smallp = settings.smallc * settings.smallc / settings.gamma0;
  settings.smallpp = settings.smallr * settings.smallp;
  settings.gamma6 = (settings.gamma0 + 1.0) / (2.0 * settings.gamma0);
}

ek_params HydroParams::to_ek() const
{
  ek_params p{};
  p.dim = dimType;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.gw = ghostWidth;
  p.nbvar = nbvar;
  p.mhd = mhdEnabled ? 1 : 0;
  p.riemann = riemannSolverType;
  p.bc[0] = boundary_type_xmin;
  p.bc[1] = boundary_type_xmax;
  p.bc[2] = boundary_type_ymin;
  p.bc[3] = boundary_type_ymax;
  p.bc[4] = boundary_type_zmin;
  p.bc[5] = boundary_type_zmax;
  p.precision = precision;
  p.xmin = xmin;
  p.ymin = ymin;
  p.zmin = zmin;
  p.xmax = xmax;
  p.ymax = ymax;
  p.zmax = zmax;
  p.dx = dx;
  p.dy = dy;
  p.dz = dz;
  p.gamma0 = settings.gamma0;
  p.gamma6 = settings.gamma6;
  p.cfl = settings.cfl;
  p.slope_type = settings.slope_type;
  p.smallr = settings.smallr;
  p.smallc = settings.smallc;
  p.smallp = settings.smallp;
  p.smallpp = settings.smallpp;
  p.cIso = settings.cIso;
  return p;
}

// src/shared/HydroParams.cpp:364-412
void HydroParams::print() const
{
  printf("##########################\n");
  printf("Simulation run parameters:\n");
  printf("##########################\n");
  printf("nx         : %d\n", nx);
  printf("ny         : %d\n", ny);
  printf("nz         : %d\n", nz);
  printf("dx         : %f\n", dx);
  printf("dy         : %f\n", dy);
  printf("dz         : %f\n", dz);
  printf("imin       : %d\n", imin);
  printf("imax       : %d\n", imax);
  printf("jmin       : %d\n", jmin);
  printf("jmax       : %d\n", jmax);
  printf("kmin       : %d\n", kmin);
  printf("kmax       : %d\n", kmax);
  printf("ghostWidth : %d\n", ghostWidth);
  printf("nbvar      : %d\n", nbvar);
  printf("nStepmax   : %d\n", nStepmax);
  printf("tEnd       : %f\n", tEnd);
  printf("nOutput    : %d\n", nOutput);
  printf("gamma0     : %f\n", settings.gamma0);
  printf("gamma6     : %f\n", settings.gamma6);
  printf("cfl        : %f\n", settings.cfl);
  printf("smallr     : %12.10f\n", settings.smallr);
  printf("smallc     : %12.10f\n", settings.smallc);
  printf("smallp     : %12.10f\n", settings.smallp);
  printf("smallpp    : %g\n", settings.smallpp);
  printf("iorder     : %d\n", settings.iorder);
  printf("slope_type : %f\n", settings.slope_type);
  printf("riemann    : %d\n", riemannSolverType);
  printf("implementation version : %d\n", implementationVersion);
  printf("precision  : %s\n", precision == EK_FP64 ? "fp64" : "fp32");
  printf("##########################\n");
}

} // namespace euler_kokkos
