// SolverBase.h -- abstract solver with the reference's virtual interface and time bookkeeping.
//
// Mirrors euler_kokkos' SolverBase (src/shared/SolverBase.h:53-259, SolverBase.cpp:22-265):
// the same public state (m_t, m_dt, m_iteration, m_tEnd, m_nCells, timers,
// m_variables_names, m_times_saved) and the same virtuals, so a driver written against the
// reference (src/main.cpp) runs unchanged. The device work of a concrete solver goes through
// the C ABI of include/ek_hip.h.
#pragma once

#include <chrono>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ConfigMap.h"
#include "HydroParams.h"

enum TimerIds {
  TIMER_TOTAL = 0,
  TIMER_IO = 1,
  TIMER_DT = 2,
  TIMER_COMPUTE_PRIMITIVES = 3,
  TIMER_COMPUTE_FLUXES = 4,
  TIMER_HYDRO_UPDATE = 5,
  TIMER_BOUNDARIES = 6,
  TIMER_NUM_SCHEME = 7
};

namespace euler_kokkos {

// Wall-clock timer with the start/stop/elapsed interface of the reference's OpenMPTimer.
class Timer {
public:
  void start() { m_t0 = std::chrono::steady_clock::now(); }
  void stop()
  {
    m_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - m_t0).count();
  }
  double elapsed() const { return m_total; }

private:
  std::chrono::steady_clock::time_point m_t0;
  double m_total = 0.0;
};

class SolverBase {
public:
  SolverBase(HydroParams& params, ConfigMap& configMap);
  virtual ~SolverBase() = default;

  HydroParams& params;
  ConfigMap& configMap;

  int m_restart_run_enabled;
  std::string m_restart_run_filename;
  double m_t;
  double m_dt;
  int m_iteration;
  double m_tEnd;
  double m_cfl;
  int m_nlog;
  long long m_nCells;       // cells including ghosts, as the reference counts them
  long long m_nDofsPerCell;
  std::string m_problem_name;
  std::string m_solver_name;

  virtual void compute_dt();
  virtual double compute_dt_local();
  virtual int finished();
  virtual void next_iteration();
  virtual void next_iteration_impl();
  virtual int should_save_solution();
  virtual void save_solution();
  virtual void save_solution_impl();
  virtual void read_restart_file();
  virtual void init_io();

  // Advance the time loop to completion (main.cpp:126-131). Solvers that keep the time step
  // on the device override it to run many iterations per host round trip when no host-side
  // event (log line, output file) falls between them.
  virtual void run();

  // IO_ReadWrite::save_data (src/utils/io/IO_ReadWrite.cpp:74-168): VTK when [output]
  // vtk_enabled (default true), HDF5 when [output] hdf5_enabled (default false).
  // Usoa: interior SoA state; Uref: reference layout with ghosts, needed only when
  // hdf5_needs_ghosts().
  void save_data(const std::vector<double>& Usoa, const std::vector<double>& Uref);
  bool hdf5_needs_ghosts() const;

  // init_restart (src/muscl/SolverHydroMuscl.h:510-580, IO_ReadWrite::load_data_impl): with
  // [run] restart_enabled, the state, m_times_saved and m_t come from the HDF5 file
  // [run] restart_filename (which needs [output] hdf5_enabled and the .h5 suffix, as there);
  // [run] restart_reset_totaltime sets m_t = 0. Returns false when no restart is asked for.
  bool init_restart(std::vector<double>& Usoa);

  int m_times_saved;
  std::map<int, std::string> m_variables_names;
  using TimerMap = std::map<int, std::shared_ptr<Timer>>;
  TimerMap timers;
};

// print monitoring information (src/shared/solver_utils.h:14-55), plus the interior-cell rate
void print_solver_monitoring_info(SolverBase* solver);

} // namespace euler_kokkos
