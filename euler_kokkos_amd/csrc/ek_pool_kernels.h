// ek_pool_kernels.h -- batched steps of many independent grids (a job pool) in one persistent launch.
//
// Every job keeps its own state in a shared device arena (two ping-pong buffers, the job at
// the same element offset in both), its own geometry, gas constants and time slots. One
// launch advances every active job by one step of its own dt: the blocks walk a work list
// of (job, tile) items and run the same per-tile code as the single-grid kernels
// (hydro2d_tile / hydro3d_column), so a job's results are those of its single-grid run.
// Each job's first tile keeps that job's time bookkeeping; every tile reduces its CFL term
// into the job's next slot.
#pragma once

#include "ek_hydro.h"
#include "ek_hydro3d.h"

namespace ek {

template <class T>
struct PoolJob {
  long long off;   // element offset of the job's state in each arena buffer
  int slot0;       // index of the job's first TimeSlot (3 per job)
  int nStepmax;
  int cz;          // z planes per column chunk (3D)
  int ntx, nty;    // tiles per axis
  double cfl, tEnd;
  Geom<T> geo;
  Gas<T> gas;
};

struct PoolWork {
  int job;
  int tx, ty;
  int tz_first; // z chunk * 2 + (1 if this item does the job's bookkeeping)
};

// Items of block b: a contiguous range of the work list, so neighbouring tiles of a job run
// one after the other on the same XCD (their halo lines meet in that XCD's L2).
__device__ __forceinline__ void pool_range(int nwork, int& w0, int& w1)
{
  const int per = nwork / (int)gridDim.x, rem = nwork % (int)gridDim.x;
  const int b = blockIdx.x;
  w0 = b * per + (b < rem ? b : rem);
  w1 = w0 + per + (b < rem ? 1 : 0);
}

template <class T, int TX, int TY, int NT, int RS, int ST>
__global__ __launch_bounds__(NT) void hydro2d_pool_step(const T* __restrict__ Ua, T* __restrict__ Ub,
                                                        const PoolJob<T>* __restrict__ jobs,
                                                        const PoolWork* __restrict__ work, int nwork,
                                                        TimeSlot* ts, int slot)
{
  __shared__ double sred[NT / 64];
  int w0, w1;
  pool_range(nwork, w0, w1);
  for (int w = w0; w < w1; w++) {
    const PoolWork it = work[w];
    const PoolJob<T> J = jobs[it.job];
    StepCtl ctl;
    ctl.ts = ts + J.slot0;
    ctl.spread = nullptr; // jobs' atomics already land on different slots
    ctl.slot = slot;
    ctl.mode = 0;
    ctl.dt_given = 0.0;
    ctl.cfl = J.cfl;
    ctl.tEnd = J.tEnd;
    ctl.nStepmax = J.nStepmax;
    const bool first = (it.tz_first & 1) != 0;
    double dt;
    TimeSlot cur;
    if (!step_begin(ctl, first, dt, cur))
      continue;
    __syncthreads(); // the previous item's shared-memory reads are done
    const double local =
        hydro2d_tile<T, TX, TY, NT, RS, ST>(Ua + J.off, Ub + J.off, J.geo, J.gas, dt, it.tx * TX, it.ty * TY);
    block_max_to<NT>(&ctl.ts[(slot + 1) % 3].invdt_bits, local, sred);
    if (first && threadIdx.x < 64)
      step_end_bookkeeping(ctl, cur, dt);
  }
}

// Occupancy targets: FP64 3 waves per SIMD (what the shared memory allows anyway); FP32 5,
// which its shared memory also allows: the compiler would otherwise keep 119 registers (4
// waves); at 96 it spills 12 and runs a serving burst of 3D jobs 7 % faster.
template <class T, int TX, int TY, int NT, int RS, int ST>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 5))) void hydro3d_pool_step(const T* __restrict__ Ua, T* __restrict__ Ub,
                                                        const PoolJob<T>* __restrict__ jobs,
                                                        const PoolWork* __restrict__ work, int nwork,
                                                        TimeSlot* ts, int slot)
{
  __shared__ double sred[NT / 64];
  int w0, w1;
  pool_range(nwork, w0, w1);
  for (int w = w0; w < w1; w++) {
    const PoolWork it = work[w];
    const PoolJob<T> J = jobs[it.job];
    StepCtl ctl;
    ctl.ts = ts + J.slot0;
    ctl.spread = nullptr; // jobs' atomics already land on different slots
    ctl.slot = slot;
    ctl.mode = 0;
    ctl.dt_given = 0.0;
    ctl.cfl = J.cfl;
    ctl.tEnd = J.tEnd;
    ctl.nStepmax = J.nStepmax;
    const bool first = (it.tz_first & 1) != 0;
    double dt;
    TimeSlot cur;
    if (!step_begin(ctl, first, dt, cur))
      continue;
    __syncthreads();
    const double local = hydro3d_column<T, TX, TY, NT, RS, ST, false>(Ua + J.off, Ub + J.off, J.geo, J.gas, dt,
                                                               it.tx * TX, it.ty * TY, (it.tz_first >> 1) * J.cz,
                                                               J.cz);
    block_max_to<NT>(&ctl.ts[(slot + 1) % 3].invdt_bits, local, sred);
    if (first && threadIdx.x < 64)
      step_end_bookkeeping(ctl, cur, dt);
  }
}

} // namespace ek
