// SolverFactory.h -- name -> solver registry (the reference's plugin point).
//
// Same interface as euler_kokkos' SolverFactory (src/shared/SolverFactory.h:43-116,
// SolverFactory.cpp:12-24): registerSolver(name, create function), create(name, ...) that
// calls init_io() and aborts on an unknown name. The registered names are the reference's
// ("Hydro_Muscl_2D", "Hydro_Muscl_3D"); here they create the HIP-backed solvers.
#pragma once

#include <map>
#include <string>

#include "SolverBase.h"

namespace euler_kokkos {

class SolverFactory {
public:
  using SolverCreateFn = SolverBase* (*)(HydroParams& params, ConfigMap& configMap);

  static SolverFactory& Instance();
  void registerSolver(const std::string& key, SolverCreateFn cfn) { m_solverCreateMap[key] = cfn; }
  SolverBase* create(const std::string& solver_name, HydroParams& params, ConfigMap& configMap);
  bool has(const std::string& name) const { return m_solverCreateMap.count(name) != 0; }

  SolverFactory(const SolverFactory&) = delete;
  SolverFactory& operator=(const SolverFactory&) = delete;

private:
  SolverFactory();
  std::map<std::string, SolverCreateFn> m_solverCreateMap;
};

} // namespace euler_kokkos
