// SolverHydroMuscl.h -- hydro MUSCL-Hancock solver (2D/3D) backed by the HIP library.
//
// Drop-in for euler_kokkos' muscl::SolverHydroMuscl<dim> (src/muscl/SolverHydroMuscl.h:
// 163-762, SolverHydroMuscl.cpp:19-390): same registration names, .ini keys, time loop and
// output. The state lives on the device in an ek_ctx (include/ek_hip.h); compute_dt_local
// is ek_inv_dt and godunov_unsplit is ek_step. run() advances many iterations per host
// round trip with the device-resident time loop (ek_run) whenever no log line or output
// file falls between them; the sequence of time steps is identical either way.
#pragma once

#include <vector>

#include "SolverBase.h"
#include "ek_hip.h"

namespace euler_kokkos {
namespace muscl {

template <int dim>
class SolverHydroMuscl : public SolverBase {
public:
  SolverHydroMuscl(HydroParams& params, ConfigMap& configMap);
  ~SolverHydroMuscl() override;

  static SolverBase* create(HydroParams& params, ConfigMap& configMap)
  {
    return new SolverHydroMuscl<dim>(params, configMap);
  }

  double compute_dt_local() override;
  void next_iteration_impl() override;
  void save_solution_impl() override;
  void run() override;

  void godunov_unsplit(double dt);

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
