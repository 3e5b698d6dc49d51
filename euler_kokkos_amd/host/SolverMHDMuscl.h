// SolverMHDMuscl.h -- MHD MUSCL-Hancock solver (2D/3D, HLLD + constrained transport) backed
// by the HIP library.
//
// Drop-in for euler_kokkos' muscl::SolverMHDMuscl<dim> (src/muscl/SolverMHDMuscl.h:44-903,
// SolverMHDMuscl.cpp:369-512): registered as "MHD_Muscl_2D" / "MHD_Muscl_3D", same .ini keys, time loop and
// output. As in the reference the time step of an iteration is computed inside
// godunov_unsplit, from the primitives of the state it advances.
#pragma once

#include <vector>

#include "SolverBase.h"
#include "ek_hip.h"

namespace euler_kokkos {
namespace muscl {

template <int dim>
class SolverMHDMuscl : public SolverBase {
public:
  SolverMHDMuscl(HydroParams& params, ConfigMap& configMap);
  ~SolverMHDMuscl() override;

  static SolverBase* create(HydroParams& params, ConfigMap& configMap)
  {
    return new SolverMHDMuscl<dim>(params, configMap);
  }

  double compute_dt_local() override;
  void next_iteration_impl() override;
  void save_solution_impl() override;
  void run() override;

  void godunov_unsplit();

  // current state, SoA interior [var][k][j][i] (FP64)
  std::vector<double> state() const;
  // current state in the reference layout U(i,j[,k],var) with the ghosts filled
  std::vector<double> state_reference() const;
  ek_ctx* context() { return m_ctx; }

  int isize, jsize, ksize;

private:
  void check(int rc, const char* what) const;
  ek_ctx* m_ctx = nullptr;
};

} // namespace muscl
} // namespace euler_kokkos
