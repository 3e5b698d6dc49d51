// Problems.h -- initial conditions, evaluated on the host at interior cell centres and
// written in the device SoA layout U[var][k][j][i].
//
// Each builder reads its .ini section like the reference's *Params struct and evaluates
// the reference's init functor expression for every interior cell (ghost cells are not
// stored; the device fills them through the boundary map):
//   implode: ImplodeParams (src/shared/problems/ImplodeParams.h),
//            InitImplodeFunctor2D/3D (src/muscl/HydroInitFunctors2D.h:33-122, 3D.h:68-168)
//   blast:   BlastParams (src/shared/problems/BlastParams.h:28-53),
//            InitBlastFunctor2D/3D (src/muscl/HydroInitFunctors2D.h:127-209, 3D.h:173-265)
//   orszag_tang (MHD 2D): InitOrszagTangFunctor2D (src/muscl/MHDInitFunctors2D.h:345-479)
//   blast (MHD 3D): InitBlastFunctor3D_MHD (src/muscl/MHDInitFunctors3D.h:267-371)
#pragma once

#include <string>
#include <vector>

#include "ConfigMap.h"
#include "HydroParams.h"

namespace euler_kokkos {

// false when `problem` has no builder here
bool init_problem(const std::string& problem, const HydroParams& params, const ConfigMap& configMap,
                  std::vector<double>& U);

void init_implode(const HydroParams& params, const ConfigMap& configMap, std::vector<double>& U);
void init_blast(const HydroParams& params, const ConfigMap& configMap, std::vector<double>& U);

// MHD problems (SolverMHDMuscl<dim>::init, src/muscl/SolverMHDMuscl.h:645-705)
bool init_problem_mhd(const std::string& problem, const HydroParams& params, const ConfigMap& configMap,
                      std::vector<double>& U);
void init_orszag_tang(const HydroParams& params, const ConfigMap& configMap, std::vector<double>& U);
void init_blast_mhd3d(const HydroParams& params, const ConfigMap& configMap, std::vector<double>& U);

} // namespace euler_kokkos
