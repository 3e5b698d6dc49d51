// ek_pool_mhd3d.h -- 3D MHD jobs of the job pool: one pool step runs the five kernels of
// the 3D MHD pipeline (ek_mhd3d.h) once each over every active job.
//
// Every job keeps its state in the pool's arena (same offset in both ping-pong buffers) and
// its own padded work area in the pool's work buffer. Two work lists describe the active
// jobs: (job, 4096-cell chunk of the padded grid) items for the four padded-grid kernels,
// which run one block per 256 cells of them (four per 256 cells for the flux/EMF kernel, one
// per task), and (job, 4096-cell chunk of the interior) items for the update, a persistent
// grid whose blocks walk contiguous ranges of them; a job's first interior chunk keeps the
// job's time bookkeeping. Each job
// reads its own dt from its own time slots. The per-cell code is the single-grid pipeline's
// (padprims_cell ... update_cell), so a job's result is its single-grid ek_run result.
#pragma once

#include "ek_mhd3d.h"
#include "ek_pool_kernels.h"

namespace ek {

template <class T>
struct PoolJobM {
  long long off;   // element offset of the job's state in each arena buffer
  long long woff;  // element offset of the job's work area in the work buffer
  int slot0;       // index of the job's first TimeSlot (3 per job)
  int nStepmax;
  double cfl, tEnd;
  mhd3::Work3 w;
  Geom<T> geo;
  Gas<T> gas;
};

// cells per work item: padded-grid kernels, update. A block walks its range in units of 256
// cells (one per thread), reading the item and its job record again for every unit: kept
// across units, the record's fields and the array bases derived from them would take the
// scalar registers that the scalar-base addressing of the per-cell code needs (measured: the
// flux/EMF code went from 66 to 123 VGPRs).
constexpr int MP_CHUNK = 4096;
constexpr int MU_CHUNK = 4096;
constexpr int UNITS_P = MP_CHUNK / 256, UNITS_U = MU_CHUNK / 256;

template <class T>
__device__ __forceinline__ StepCtl pool_ctl(const PoolJobM<T>& J, TimeSlot* ts, int slot)
{
  StepCtl ctl;
  ctl.ts = ts + J.slot0;
  ctl.spread = nullptr;
  ctl.slot = slot;
  ctl.mode = 0;
  ctl.dt_given = 0.0;
  ctl.cfl = J.cfl;
  ctl.tEnd = J.tEnd;
  ctl.nStepmax = J.nStepmax;
  return ctl;
}

// One unit per block (TASKS units per 256 cells: the flux/EMF kernel runs its four tasks as
// units of their own, as the single-grid kernel runs them in blocks of their own): the block
// reads its item, the item's job record and the job's dt, then `body` runs for this thread's
// padded cell c and the unit's task. One unit per block lets the hardware overlap those
// dependent reads across the many resident blocks; persistent blocks walking ranges of units
// serialized them with every unit's cell loads (0.78 against 0.60 ms per 128^3 FP32 job).
template <class T, int TASKS, class Body>
__device__ __forceinline__ void pool_padded_unit(const PoolJobM<T>* __restrict__ jobs,
                                                 const PoolWork* __restrict__ work, TimeSlot* ts, int slot,
                                                 Body body)
{
  constexpr int PER = UNITS_P * TASKS;
  const int u = blockIdx.x;
  const int q = u / PER, r = u - q * PER;
  const PoolWork it = work[q];
  const PoolJobM<T> J = jobs[it.job];
  double dt;
  if (!mhd3::step_dt(pool_ctl(J, ts, slot), dt))
    return;
  const unsigned c = ((unsigned)it.tx * (unsigned)UNITS_P + (unsigned)(r / TASKS)) * 256u + threadIdx.x;
  if ((long long)c < J.w.PN)
    body(J, dt, c, r % TASKS);
}

template <class T>
__global__ __launch_bounds__(256) void mhd3_pool_padprims(const T* __restrict__ Ua, T* __restrict__ W,
                                                          const PoolJobM<T>* __restrict__ jobs,
                                                          const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                          int slot)
{
  pool_padded_unit<T, 1>(jobs, work, ts, slot, [&](const PoolJobM<T>& J, double, unsigned c, int) {
    int i, j, k;
    mhd3::split3(c, (unsigned)J.w.PX, (unsigned)J.w.PY, i, j, k);
    mhd3::padprims_cell<T>(Ua + J.off, W + J.woff, J.w, J.geo, J.gas, i, j, k);
  });
}

template <class T, int ST>
__global__ __launch_bounds__(256) void mhd3_pool_efield(T* __restrict__ W, const PoolJobM<T>* __restrict__ jobs,
                                                        const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                        int slot)
{
  pool_padded_unit<T, 1>(jobs, work, ts, slot, [&](const PoolJobM<T>& J, double, unsigned c, int) {
    int i, j, k;
    mhd3::split3(c, (unsigned)J.w.PX, (unsigned)J.w.PY, i, j, k);
    mhd3::efield_cell<T, ST>(W + J.woff, J.w, i, j, k);
  });
}

template <class T, int ST>
__global__ __launch_bounds__(256) void mhd3_pool_trace(T* __restrict__ W, const PoolJobM<T>* __restrict__ jobs,
                                                       const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                       int slot)
{
  pool_padded_unit<T, 1>(jobs, work, ts, slot, [&](const PoolJobM<T>& J, double dt, unsigned c, int) {
    int i, j, k;
    mhd3::split3(c, (unsigned)J.w.PX, (unsigned)J.w.PY, i, j, k);
    mhd3::trace_cell<T, ST>(W + J.woff, J.w, J.geo, J.gas, dt, i, j, k);
  });
}

template <class T, int RS>
__global__ __launch_bounds__(256) void mhd3_pool_fluxemf(T* __restrict__ W, const PoolJobM<T>* __restrict__ jobs,
                                                         const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                         int slot)
{
  pool_padded_unit<T, 4>(jobs, work, ts, slot, [&](const PoolJobM<T>& J, double, unsigned c, int task) {
    mhd3::fluxemf_cell<T, RS>(W + J.woff, J.w, J.gas, c, task);
  });
}

// Update. A block's range holds consecutive units of few jobs: the block keeps a running CFL
// maximum and adds it to a job's next slot once, when its units of that job end. The job's
// first unit keeps the job's time bookkeeping. (Contiguous ranges measured faster than
// grid-strided units here: 0.78 against 0.88 ms per 128^3 FP32 job in a pool of 8.)
template <class T>
__global__ __launch_bounds__(256) void mhd3_pool_update(const T* __restrict__ W, T* __restrict__ Ub,
                                                        const PoolJobM<T>* __restrict__ jobs,
                                                        const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                        int slot)
{
  __shared__ double sred[4];
  int u0, u1;
  pool_range(UNITS_U * nwork, u0, u1);
  double local = 0.0;
  int pending = -1;  // TimeSlot word index of the job whose maximum `local` holds
  int pjob = -1;
  for (int u = u0; u < u1; u++) {
    const int q = u / UNITS_U, r = u - q * UNITS_U;
    const PoolWork it = work[q];
    if (pending >= 0 && it.job != pjob) {
      block_max_to<256>(&ts[pending].invdt_bits, local, sred);
      local = 0.0;
      pending = -1;
    }
    const PoolJobM<T> J = jobs[it.job];
    const StepCtl ctl = pool_ctl(J, ts, slot);
    const bool first = (it.tz_first & 1) != 0 && r == 0;
    double dt;
    TimeSlot cur;
    if (!step_begin(ctl, first, dt, cur))
      continue;
    pending = J.slot0 + (slot + 1) % 3;
    pjob = it.job;
    const T dtdx = T(dt / J.geo.dx), dtdy = T(dt / J.geo.dy), dtdz = T(dt / J.geo.dz);
    const long long c = ((long long)it.tx * UNITS_U + r) * 256 + threadIdx.x;
    if (c < J.geo.vstride) {
      const double sv = mhd3::update_cell<T>(W + J.woff, Ub + J.off, J.w, J.geo, J.gas, dtdx, dtdy, dtdz, c);
      local = (sv > local || !(sv == sv)) ? sv : local;
    }
    if (first && threadIdx.x < 64)
      step_end_bookkeeping(ctl, cur, dt);
  }
  if (pending >= 0)
    block_max_to<256>(&ts[pending].invdt_bits, local, sred);
}

namespace pool {

// persistent grid of each of the five kernels: CUs x blocks that fit on one CU at once
template <class T, int RS, int ST>
inline void mhd3_grids(int cus, int* grid)
{
  const void* f[5] = { reinterpret_cast<const void*>(mhd3_pool_padprims<T>),
                       reinterpret_cast<const void*>(mhd3_pool_efield<T, ST>),
                       reinterpret_cast<const void*>(mhd3_pool_trace<T, ST>),
                       reinterpret_cast<const void*>(mhd3_pool_fluxemf<T, RS>),
                       reinterpret_cast<const void*>(mhd3_pool_update<T>) };
  for (int i = 0; i < 5; i++) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f[i], 256, 0) != hipSuccess || n < 1)
      n = 1;
    grid[i] = cus * n;
  }
}

template <class T, int RS, int ST>
inline hipError_t mhd3_launch(const T* Ua, T* Ub, T* W, const PoolJobM<T>* jobs, const PoolWork* wp, int np,
                              const PoolWork* wu, int nu, TimeSlot* ts, int slot, const int* grid, hipStream_t s)
{
  if (np <= 0 || nu <= 0)
    return hipSuccess;
  auto g = [&](int i, int n) { return dim3((unsigned)(grid[i] < n ? grid[i] : n)); };
  hipLaunchKernelGGL(mhd3_pool_padprims<T>, dim3(UNITS_P * np), dim3(256), 0, s, Ua, W, jobs, wp, np, ts, slot);
  hipLaunchKernelGGL((mhd3_pool_efield<T, ST>), dim3(UNITS_P * np), dim3(256), 0, s, W, jobs, wp, np, ts, slot);
  hipLaunchKernelGGL((mhd3_pool_trace<T, ST>), dim3(UNITS_P * np), dim3(256), 0, s, W, jobs, wp, np, ts, slot);
  hipLaunchKernelGGL((mhd3_pool_fluxemf<T, RS>), dim3(4 * UNITS_P * np), dim3(256), 0, s, W, jobs, wp, np, ts, slot);
  hipLaunchKernelGGL(mhd3_pool_update<T>, g(4, UNITS_U * nu), dim3(256), 0, s, W, Ub, jobs, wu, nu, ts, slot);
  return hipGetLastError();
}

template <class T>
inline hipError_t mhd3_step(int rs, int st, const T* Ua, T* Ub, T* W, const PoolJobM<T>* jobs, const PoolWork* wp,
                            int np, const PoolWork* wu, int nu, TimeSlot* ts, int slot, const int* grid, hipStream_t s)
{
  const bool st1 = st == 1;
  switch (rs) {
    case RS_HLL: return st1 ? mhd3_launch<T, RS_HLL, 1>(Ua, Ub, W, jobs, wp, np, wu, nu, ts, slot, grid, s)
                            : mhd3_launch<T, RS_HLL, 2>(Ua, Ub, W, jobs, wp, np, wu, nu, ts, slot, grid, s);
    case RS_LLF: return st1 ? mhd3_launch<T, RS_LLF, 1>(Ua, Ub, W, jobs, wp, np, wu, nu, ts, slot, grid, s)
                            : mhd3_launch<T, RS_LLF, 2>(Ua, Ub, W, jobs, wp, np, wu, nu, ts, slot, grid, s);
    default: return st1 ? mhd3_launch<T, RS_HLLD, 1>(Ua, Ub, W, jobs, wp, np, wu, nu, ts, slot, grid, s)
                        : mhd3_launch<T, RS_HLLD, 2>(Ua, Ub, W, jobs, wp, np, wu, nu, ts, slot, grid, s);
  }
}

template <class T>
inline void mhd3_grids_for(int rs, int st, int cus, int* grid)
{
  const bool st1 = st == 1;
  switch (rs) {
    case RS_HLL: return st1 ? mhd3_grids<T, RS_HLL, 1>(cus, grid) : mhd3_grids<T, RS_HLL, 2>(cus, grid);
    case RS_LLF: return st1 ? mhd3_grids<T, RS_LLF, 1>(cus, grid) : mhd3_grids<T, RS_LLF, 2>(cus, grid);
    default: return st1 ? mhd3_grids<T, RS_HLLD, 1>(cus, grid) : mhd3_grids<T, RS_HLLD, 2>(cus, grid);
  }
}

} // namespace pool
} // namespace ek
