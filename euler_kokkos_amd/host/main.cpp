// main.cpp -- .ini-driven run, the counterpart of euler_kokkos' src/main.cpp:53-159.
//
//   euler_kokkos_amd settings.ini
//
// Reads the parameter file, builds the solver named by [run] solver_name through the
// SolverFactory, writes the initial and final outputs when nOutput != 0 (VTK and/or HDF5,
// plus the Xdmf wrapper of the HDF5 files), runs the time loop and prints the timing summary (including the reference's Mcell-updates/s figure,
// which counts ghost cells, and the interior-cell rate).
#include <cstdio>
#include <iostream>
#include <string>

#include "ConfigMap.h"
#include "HydroParams.h"
#include "IO_HDF5.h"
#include "SolverFactory.h"

int main(int argc, char* argv[])
{
  namespace ek = ::euler_kokkos;
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <parameter file .ini>\n", argv[0]);
    return 1;
  }
  ek::ConfigMap configMap = ek::broadcast_parameters(std::string(argv[1]));
  if (configMap.ParseError() < 0)
    std::fprintf(stderr, "warning: cannot read %s, using defaults\n", argv[1]);
  ek::HydroParams params(configMap);
  const std::string solver_name = configMap.getString("run", "solver_name", "Unknown");
  ek::SolverBase* solver = nullptr;
  try {
    solver = ek::SolverFactory::Instance().create(solver_name, params, configMap);
    if (params.nOutput != 0)
      solver->save_solution();
    std::cout << "Start computation....\n";
    solver->timers[TIMER_TOTAL]->start();
    solver->run();
    solver->timers[TIMER_TOTAL]->stop();
    if (params.nOutput != 0)
      solver->save_solution();
    // Xdmf wrapper of the HDF5 outputs (main.cpp:140-147)
    if (configMap.getBool("output", "hdf5_enabled", false))
      ek::writeXdmfForHdf5Wrapper(params, configMap, solver->m_variables_names, solver->m_times_saved - 1, false);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    delete solver;
    return 2;
  }
  std::printf("final time is %f\n", solver->m_t);
  ek::print_solver_monitoring_info(solver);
  delete solver;
  return 0;
}
