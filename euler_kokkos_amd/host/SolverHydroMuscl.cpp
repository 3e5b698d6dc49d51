// SolverHydroMuscl.cpp -- see SolverHydroMuscl.h.
#include "SolverHydroMuscl.h"

#include <cstdio>
#include <iostream>
#include <stdexcept>

#include "Problems.h"

namespace euler_kokkos {
namespace muscl {

template <int dim>
void SolverHydroMuscl<dim>::check(int rc, const char* what) const
{
  if (rc != EK_OK)
    throw std::runtime_error(std::string(what) + ": " + ek_last_error(m_ctx));
}

// src/muscl/SolverHydroMuscl.h:163-307
template <int dim>
SolverHydroMuscl<dim>::SolverHydroMuscl(HydroParams& params, ConfigMap& configMap)
  : SolverBase(params, configMap), isize(params.isize), jsize(params.jsize), ksize(params.ksize)
{
  m_nCells = dim == 2 ? (long long)params.isize * params.jsize
                      : (long long)params.isize * params.jsize * params.ksize;
  m_nDofsPerCell = 1;
  const ek_params ekp = params.to_ek();
  if (ek_create(&ekp, &m_ctx) != EK_OK)
    throw std::runtime_error(std::string("ek_create: ") + ek_last_error(nullptr));

  // initial condition (SolverHydroMuscl.cpp:78-207); unknown problems fall back to implode
  std::vector<double> U;
  if (init_restart(U)) {
    // state, output counter and time from the restart file
  } else if (!init_problem(m_problem_name, params, configMap, U)) {
    std::cout << "Problem : " << m_problem_name << " is not recognized / implemented." << std::endl;
    std::cout << "Use default - implode" << std::endl;
    init_implode(params, configMap, U);
  }
  check(ek_upload(m_ctx, U.data(), EK_LAYOUT_SOA), "ek_upload");
  compute_dt();

  std::cout << "##########################" << "\n";
  std::cout << "Solver is " << m_solver_name << "\n";
  std::cout << "Problem (init condition) is " << m_problem_name << "\n";
  std::cout << "##########################" << "\n";
  params.print();
  std::cout << "##########################" << "\n";
  std::cout << "Memory requested : " << (ek_device_bytes(m_ctx) / 1e6) << " MBytes\n";
  std::cout << "##########################" << "\n";
}

template <int dim>
SolverHydroMuscl<dim>::~SolverHydroMuscl()
{
  ek_destroy(m_ctx);
}

// src/muscl/SolverHydroMuscl.h:589-631
template <int dim>
double SolverHydroMuscl<dim>::compute_dt_local()
{
  double invDt = 0.0;
  check(ek_inv_dt(m_ctx, &invDt), "ek_inv_dt");
  return params.settings.cfl / invDt;
}

// src/muscl/SolverHydroMuscl.h:635-679
template <int dim>
void SolverHydroMuscl<dim>::next_iteration_impl()
{
  if (m_iteration % m_nlog == 0)
    printf("time step=%7d (dt=% 10.8f t=% 10.8f)\n", m_iteration, m_dt, m_t);
  if (params.enableOutput && should_save_solution()) {
    std::cout << "Output results at time t=" << m_t << " step " << m_iteration << " dt=" << m_dt << std::endl;
    save_solution();
  }
  timers[TIMER_DT]->start();
  compute_dt();
  timers[TIMER_DT]->stop();
  godunov_unsplit(m_dt);
}

template <int dim>
void SolverHydroMuscl<dim>::godunov_unsplit(double dt)
{
  timers[TIMER_NUM_SCHEME]->start();
  check(ek_step(m_ctx, dt), "ek_step");
  check(ek_sync(m_ctx), "ek_sync");
  timers[TIMER_NUM_SCHEME]->stop();
}

template <int dim>
void SolverHydroMuscl<dim>::run()
{
  if (params.enableOutput) {
    // output times depend on t, which only the per-iteration path sees in time
    SolverBase::run();
    return;
  }
  while (!finished()) {
    if (m_iteration % m_nlog == 0)
      printf("time step=%7d (dt=% 10.8f t=% 10.8f)\n", m_iteration, m_dt, m_t);
    // iterations until the next log line (or the end)
    int n = m_nlog - m_iteration % m_nlog;
    if (m_iteration + n > params.nStepmax)
      n = params.nStepmax - m_iteration;
    timers[TIMER_NUM_SCHEME]->start();
    double t = m_t, dt = m_dt;
    int it = m_iteration;
    check(ek_run(m_ctx, n, params.nStepmax, m_tEnd, &t, &it, &dt), "ek_run");
    timers[TIMER_NUM_SCHEME]->stop();
    if (it == m_iteration)
      break;
    m_t = t;
    m_dt = dt;
    m_iteration = it;
  }
}

template <int dim>
std::vector<double> SolverHydroMuscl<dim>::state_reference() const
{
  const long long n = (long long)params.isize * params.jsize * (dim == 3 ? params.ksize : 1) * params.nbvar;
  std::vector<double> U((size_t)n);
  check(ek_download(m_ctx, U.data(), EK_LAYOUT_REFERENCE), "ek_download");
  return U;
}

template <int dim>
std::vector<double> SolverHydroMuscl<dim>::state() const
{
  const long long n = (long long)params.nx * params.ny * (dim == 3 ? params.nz : 1) * params.nbvar;
  std::vector<double> U((size_t)n);
  check(ek_download(m_ctx, U.data(), EK_LAYOUT_SOA), "ek_download");
  return U;
}

// src/muscl/SolverHydroMuscl.h:749-762 (VTK output)
template <int dim>
void SolverHydroMuscl<dim>::save_solution_impl()
{
  timers[TIMER_IO]->start();
  const std::vector<double> U = state();
  save_data(U, hdf5_needs_ghosts() ? state_reference() : std::vector<double>());
  timers[TIMER_IO]->stop();
}

template class SolverHydroMuscl<2>;
template class SolverHydroMuscl<3>;

} // namespace muscl
} // namespace euler_kokkos
