// SolverBase.cpp -- see SolverBase.h.
#include "SolverBase.h"

#include <cmath>
#include <cstdio>
#include <iostream>
#include <stdexcept>

#include "IO_HDF5.h"
#include "IO_VTK.h"

namespace euler_kokkos {

// src/shared/SolverBase.cpp:22-110
SolverBase::SolverBase(HydroParams& params, ConfigMap& configMap)
  : params(params)
  , configMap(configMap)
  , m_restart_run_enabled((int)configMap.getInteger("run", "restart_enabled", 0))
  , m_restart_run_filename(configMap.getString("run", "restart_filename", ""))
  , m_t(configMap.getFloat("run", "tCurrent", 0.0))
  , m_dt(0.0)
  , m_iteration(0)
  , m_tEnd(configMap.getFloat("run", "tEnd", 0.0))
  , m_cfl(configMap.getFloat("hydro", "cfl", 1.0))
  , m_nlog((int)configMap.getInteger("run", "nlog", 10))
  , m_nCells(-1)
  , m_nDofsPerCell(-1)
  , m_problem_name(configMap.getString("hydro", "problem", "unknown"))
  , m_solver_name(configMap.getString("run", "solver_name", "unknown"))
  , m_times_saved(0)
{
  for (int id : { TIMER_TOTAL, TIMER_IO, TIMER_DT, TIMER_COMPUTE_PRIMITIVES, TIMER_COMPUTE_FLUXES,
                  TIMER_HYDRO_UPDATE, TIMER_BOUNDARIES, TIMER_NUM_SCHEME })
    timers[id] = std::make_shared<Timer>();
  m_variables_names[ID] = "rho";
  m_variables_names[IP] = "energy";
  m_variables_names[IU] = "rho_vx";
  m_variables_names[IV] = "rho_vy";
  m_variables_names[IW] = "rho_vz";
  m_variables_names[IA] = "bx";
  m_variables_names[IB] = "by";
  m_variables_names[IC] = "bz";
}

// src/shared/SolverBase.cpp:124-154 (single subdomain: no allreduce)
void SolverBase::compute_dt()
{
  m_dt = compute_dt_local();
  if (m_t + m_dt > m_tEnd)
    m_dt = m_tEnd - m_t;
}

double SolverBase::compute_dt_local()
{
  return m_tEnd;
}

// src/shared/SolverBase.cpp:170-176
int SolverBase::finished()
{
  return m_t >= (m_tEnd - 1e-14) || m_iteration >= params.nStepmax;
}

// src/shared/SolverBase.cpp:180-195
void SolverBase::next_iteration()
{
  next_iteration_impl();
  ++m_iteration;
  m_t += m_dt;
}

void SolverBase::next_iteration_impl() {}

// src/shared/SolverBase.cpp:239-265
int SolverBase::should_save_solution()
{
  const double interval = m_tEnd / params.nOutput;
  if (params.nOutput < 0)
    return 1;
  if ((m_t - (m_times_saved - 1) * interval) > interval)
    return 1;
  if (std::fabs(m_t - m_tEnd) < 1e-12) // ISFUZZYNULL (src/shared/utils.h:20-22)
    return 1;
  return 0;
}

void SolverBase::save_solution()
{
  save_solution_impl();
  ++m_times_saved;
}

void SolverBase::save_solution_impl() {}
void SolverBase::read_restart_file() {}
void SolverBase::init_io() {}

void SolverBase::run()
{
  while (!finished())
    next_iteration();
}

void SolverBase::save_data(const std::vector<double>& Usoa, const std::vector<double>& Uref)
{
  if (configMap.getBool("output", "vtk_enabled", true))
    save_VTK(Usoa.data(), params, configMap, params.nbvar, m_variables_names, m_times_saved);
  if (configMap.getBool("output", "hdf5_enabled", false))
    save_HDF5(Usoa, Uref, params, configMap, m_variables_names, m_times_saved, m_t);
}

bool SolverBase::hdf5_needs_ghosts() const
{
  return configMap.getBool("output", "hdf5_enabled", false) && configMap.getBool("output", "ghostIncluded", false);
}

bool SolverBase::init_restart(std::vector<double>& Usoa)
{
  if (!configMap.getBool("run", "restart_enabled", false))
    return false;
  const std::string f = configMap.getString("run", "restart_filename", "");
  const bool h5 = f.size() >= 3 && f.compare(f.size() - 3, 3, ".h5") == 0;
  if (!configMap.getBool("output", "hdf5_enabled", false) || !h5)
    throw std::runtime_error("restart needs [output] hdf5_enabled=true and an .h5 [run] restart_filename (got '" +
                             f + "')");
  int iStep = 0;
  double t = 0.0;
  load_HDF5(f, params, configMap, m_variables_names, Usoa, iStep, t);
  m_times_saved = iStep;
  m_t = t;
  if (configMap.getBool("run", "restart_reset_totaltime", false))
    m_t = 0;
  std::cout << "### This is a restarted run ! Current time is " << m_t << " ###\n";
  return true;
}

// src/shared/solver_utils.h:14-55
void print_solver_monitoring_info(SolverBase* solver)
{
  const double t_tot = solver->timers[TIMER_TOTAL]->elapsed();
  const double t_comp = solver->timers[TIMER_NUM_SCHEME]->elapsed();
  const double t_dt = solver->timers[TIMER_DT]->elapsed();
  const double t_bound = solver->timers[TIMER_BOUNDARIES]->elapsed();
  const double t_io = solver->timers[TIMER_IO]->elapsed();
  printf("total        time : %5.3f secondes\n", t_tot);
  printf("godunov      time : %5.3f secondes %5.2f%%\n", t_comp, 100 * t_comp / t_tot);
  printf("compute dt   time : %5.3f secondes %5.2f%%\n", t_dt, 100 * t_dt / t_tot);
  printf("boundaries   time : %5.3f secondes %5.2f%%\n", t_bound, 100 * t_bound / t_tot);
  printf("io           time : %5.3f secondes %5.2f%%\n", t_io, 100 * t_io / t_tot);
  printf("Perf              : %5.3f number of Mcell-updates/s\n",
         solver->m_iteration * solver->m_nCells / t_tot * 1e-6);
  const auto& p = solver->params;
  const double interior = (double)p.nx * p.ny * (p.dimType == THREE_D ? p.nz : 1);
  printf("Perf (interior)   : %5.3f number of Mcell-updates/s\n",
         solver->m_iteration * interior / t_tot * 1e-6);
}

} // namespace euler_kokkos
