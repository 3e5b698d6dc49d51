// SolverFactory.cpp -- see SolverFactory.h.
#include "SolverFactory.h"

#include <cstdio>
#include <cstdlib>

#include "SolverHydroMuscl.h"
#include "SolverMHDMuscl.h"

namespace euler_kokkos {

SolverFactory::SolverFactory()
{
  registerSolver("Hydro_Muscl_2D", &muscl::SolverHydroMuscl<2>::create);
  registerSolver("Hydro_Muscl_3D", &muscl::SolverHydroMuscl<3>::create);
  registerSolver("MHD_Muscl_2D", &muscl::SolverMHDMuscl<2>::create);
  registerSolver("MHD_Muscl_3D", &muscl::SolverMHDMuscl<3>::create);
}

SolverFactory& SolverFactory::Instance()
{
  static SolverFactory instance;
  return instance;
}

SolverBase* SolverFactory::create(const std::string& solver_name, HydroParams& params, ConfigMap& configMap)
{
  auto it = m_solverCreateMap.find(solver_name);
  if (it != m_solverCreateMap.end()) {
    SolverBase* solver = it->second(params, configMap);
    solver->init_io();
    return solver;
  }
  printf("############ WARNING: ############\n");
  printf("%s: is not recognized as a valid application name key.\n", solver_name.c_str());
  printf("Valid solver names are:\n");
  for (const auto& kv : m_solverCreateMap)
    printf("%s\n", kv.first.c_str());
  printf("############ WARNING: ############\n");
  printf("Solver application name not found\n");
  std::abort();
  return nullptr;
}

} // namespace euler_kokkos
