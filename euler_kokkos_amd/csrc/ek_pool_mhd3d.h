// ek_pool_mhd3d.h -- 3D MHD jobs of the job pool: one pool step runs the five kernels of
// the 3D MHD pipeline (ek_mhd3d.h) once each over every active job.
//
// Every job keeps its state in the pool's arena (same offset in both ping-pong buffers) and
// its own padded work area in the pool's work buffer. Each kernel is a persistent grid whose
// blocks walk a contiguous range of a work list: the four padded-grid kernels one list of
// (job, 256-cell chunk of the padded grid) items, the update another of (job, 4096-cell
// chunk of the interior) items whose first chunk keeps the job's time bookkeeping. Each job
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

// cells per work item: padded-grid kernels, update
constexpr int MP_CHUNK = 256;
constexpr int MU_CHUNK = 4096;

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

// padded cell of item `it` for this thread, or false past the job's padded grid
template <class T>
__device__ __forceinline__ bool pool_pcell(const PoolJobM<T>& J, const PoolWork& it, unsigned& c)
{
  c = (unsigned)it.tx * (unsigned)MP_CHUNK + threadIdx.x;
  return (long long)c < J.w.PN;
}

template <class T>
__global__ __launch_bounds__(256) void mhd3_pool_padprims(const T* __restrict__ Ua, T* __restrict__ W,
                                                          const PoolJobM<T>* __restrict__ jobs,
                                                          const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                          int slot)
{
  int w0, w1;
  pool_range(nwork, w0, w1);
  for (int q = w0; q < w1; q++) {
    const PoolWork it = work[q];
    const PoolJobM<T> J = jobs[it.job];
    double dt;
    unsigned c;
    if (!mhd3::step_dt(pool_ctl(J, ts, slot), dt) || !pool_pcell(J, it, c))
      continue;
    int i, j, k;
    mhd3::split3(c, (unsigned)J.w.PX, (unsigned)J.w.PY, i, j, k);
    mhd3::padprims_cell<T>(Ua + J.off, W + J.woff, J.w, J.geo, J.gas, i, j, k);
  }
}

template <class T, int ST>
__global__ __launch_bounds__(256) void mhd3_pool_efield(T* __restrict__ W, const PoolJobM<T>* __restrict__ jobs,
                                                        const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                        int slot)
{
  int w0, w1;
  pool_range(nwork, w0, w1);
  for (int q = w0; q < w1; q++) {
    const PoolWork it = work[q];
    const PoolJobM<T> J = jobs[it.job];
    double dt;
    unsigned c;
    if (!mhd3::step_dt(pool_ctl(J, ts, slot), dt) || !pool_pcell(J, it, c))
      continue;
    int i, j, k;
    mhd3::split3(c, (unsigned)J.w.PX, (unsigned)J.w.PY, i, j, k);
    mhd3::efield_cell<T, ST>(W + J.woff, J.w, i, j, k);
  }
}

template <class T, int ST>
__global__ __launch_bounds__(256) void mhd3_pool_trace(T* __restrict__ W, const PoolJobM<T>* __restrict__ jobs,
                                                       const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                       int slot)
{
  int w0, w1;
  pool_range(nwork, w0, w1);
  for (int q = w0; q < w1; q++) {
    const PoolWork it = work[q];
    const PoolJobM<T> J = jobs[it.job];
    double dt;
    unsigned c;
    if (!mhd3::step_dt(pool_ctl(J, ts, slot), dt) || !pool_pcell(J, it, c))
      continue;
    int i, j, k;
    mhd3::split3(c, (unsigned)J.w.PX, (unsigned)J.w.PY, i, j, k);
    mhd3::trace_cell<T, ST>(W + J.woff, J.w, J.geo, J.gas, dt, i, j, k);
  }
}

// the four tasks of a chunk (three Riemann problems; EMF z, y, x) one after the other in the
// same block, so they share its CU's cache lines of the chunk's records
template <class T, int RS>
__global__ __launch_bounds__(256) void mhd3_pool_fluxemf(T* __restrict__ W, const PoolJobM<T>* __restrict__ jobs,
                                                         const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                         int slot)
{
  int w0, w1;
  pool_range(nwork, w0, w1);
  for (int q = w0; q < w1; q++) {
    const PoolWork it = work[q];
    const PoolJobM<T> J = jobs[it.job];
    double dt;
    unsigned c;
    if (!mhd3::step_dt(pool_ctl(J, ts, slot), dt) || !pool_pcell(J, it, c))
      continue;
    for (int task = 0; task < 4; task++)
      mhd3::fluxemf_cell<T, RS>(W + J.woff, J.w, J.gas, c, task);
  }
}

// Update. A block's range holds consecutive chunks of few jobs: the block keeps a running CFL
// maximum and adds it to a job's next slot once, when its items of that job end.
template <class T>
__global__ __launch_bounds__(256) void mhd3_pool_update(const T* __restrict__ W, T* __restrict__ Ub,
                                                        const PoolJobM<T>* __restrict__ jobs,
                                                        const PoolWork* __restrict__ work, int nwork, TimeSlot* ts,
                                                        int slot)
{
  __shared__ double sred[4];
  int w0, w1;
  pool_range(nwork, w0, w1);
  double local = 0.0;
  int pending = -1;  // job whose maximum `local` holds
  for (int q = w0; q < w1; q++) {
    const PoolWork it = work[q];
    if (pending >= 0 && it.job != pending) {
      block_max_to<256>(&ts[jobs[pending].slot0 + (slot + 1) % 3].invdt_bits, local, sred);
      local = 0.0;
      pending = -1;
    }
    const PoolJobM<T> J = jobs[it.job];
    const StepCtl ctl = pool_ctl(J, ts, slot);
    const bool first = (it.tz_first & 1) != 0;
    double dt;
    TimeSlot cur;
    if (!step_begin(ctl, first, dt, cur))
      continue;
    pending = it.job;
    const T dtdx = T(dt / J.geo.dx), dtdy = T(dt / J.geo.dy), dtdz = T(dt / J.geo.dz);
    const long long n = J.geo.vstride;
    for (int s = 0; s < MU_CHUNK / 256; s++) {
      const long long c = (long long)it.tx * MU_CHUNK + s * 256 + threadIdx.x;
      if (c < n) {
        const double sv = mhd3::update_cell<T>(W + J.woff, Ub + J.off, J.w, J.geo, J.gas, dtdx, dtdy, dtdz, c);
        local = (sv > local || !(sv == sv)) ? sv : local;
      }
    }
    if (first && threadIdx.x < 64)
      step_end_bookkeeping(ctl, cur, dt);
  }
  if (pending >= 0)
    block_max_to<256>(&ts[jobs[pending].slot0 + (slot + 1) % 3].invdt_bits, local, sred);
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
  hipLaunchKernelGGL(mhd3_pool_padprims<T>, g(0, np), dim3(256), 0, s, Ua, W, jobs, wp, np, ts, slot);
  hipLaunchKernelGGL((mhd3_pool_efield<T, ST>), g(1, np), dim3(256), 0, s, W, jobs, wp, np, ts, slot);
  hipLaunchKernelGGL((mhd3_pool_trace<T, ST>), g(2, np), dim3(256), 0, s, W, jobs, wp, np, ts, slot);
  hipLaunchKernelGGL((mhd3_pool_fluxemf<T, RS>), g(3, np), dim3(256), 0, s, W, jobs, wp, np, ts, slot);
  hipLaunchKernelGGL(mhd3_pool_update<T>, g(4, nu), dim3(256), 0, s, W, Ub, jobs, wu, nu, ts, slot);
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
