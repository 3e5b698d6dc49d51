// HydroParams.h -- run / mesh / scheme parameters read from a ConfigMap.
//
// Field names, .ini keys, defaults and derived quantities follow euler_kokkos'
// HydroParams / HydroSettings (src/shared/HydroParams.h:32-197, HydroParams.cpp:37-180,326-359),
// so an existing .ini file produces the same parameter values. Reals are float-parsed
// (ConfigMap::getFloat) exactly as the reference does.
#pragma once

#include <string>

#include "ConfigMap.h"
#include "ek_hip.h"

namespace euler_kokkos {

enum DimensionType { TWO_D = 2, THREE_D = 3 };
enum RiemannSolverType { RIEMANN_APPROX, RIEMANN_LLF, RIEMANN_HLL, RIEMANN_HLLC, RIEMANN_HLLD };
enum BoundaryConditionType { BC_UNDEFINED, BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC, BC_COPY };
enum VarIndex { ID = 0, IP = 1, IE = 1, IU = 2, IV = 3, IW = 4, IA = 5, IB = 6, IC = 7 };

struct HydroSettings {
  double gamma0 = 1.4;
  double gamma6 = 1.0;
  double cfl = 1.0;
  double slope_type = 2.0;
  int iorder = 1;
  double smallr = 1e-8;
  double smallc = 1e-8;
  double smallp = 1e-6;
  double smallpp = 1e-6;
  double cIso = 0;
  double Omega0 = 0.0;
  double cp = 0.0;
  double mu = 0.0;
  double kappa = 0.0;
};

DimensionType get_dim(const ConfigMap& configMap);

class HydroParams {
public:
  explicit HydroParams(const ConfigMap& configMap);
  virtual ~HydroParams() = default;

  int nStepmax;
  double tEnd;
  int nOutput;
  bool enableOutput;
  bool mhdEnabled;
  int nlog;

  int nx, ny, nz;
  int ghostWidth;
  int nbvar;
  DimensionType dimType;
  int imin, imax, jmin, jmax, kmin, kmax;
  int isize, jsize, ksize;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double dx, dy, dz;
  BoundaryConditionType boundary_type_xmin, boundary_type_xmax, boundary_type_ymin,
    boundary_type_ymax, boundary_type_zmin, boundary_type_zmax;
  bool ioVTK, ioHDF5;
  HydroSettings settings;
  int niter_riemann;
  int riemannSolverType;
  int implementationVersion = 0;
  // precision of the device state: [hydro] precision = fp64 (default) | fp32
  int precision = EK_FP64;

  void print() const;
  // the POD the device library takes (include/ek_hip.h)
  ek_params to_ek() const;

private:
  void init();
};

} // namespace euler_kokkos
