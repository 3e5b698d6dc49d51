// ek_pool_launch.h -- instantiation and launch of the job-pool kernels for one real type.
// Included by exactly one translation unit per precision (ek_pool_fp64.hip, ek_pool_fp32.hip).
#pragma once

#include "ek_pool_kernels.h"

namespace ek {
namespace pool {

// 2D tile shapes the autotuner chooses from: (TX, TY, threads)
constexpr int N2VARIANTS = 4;
constexpr int V2[N2VARIANTS][3] = { { 16, 16, 256 }, { 32, 8, 256 }, { 16, 8, 128 }, { 8, 8, 64 } };
// 3D column tile (the z chunk depth is the tuned parameter)
constexpr int P3X = 16, P3Y = 8, P3T = 256;

template <class T>
using Pool2Fn = void (*)(const T*, T*, const PoolJob<T>*, const PoolWork*, int, TimeSlot*, int);

template <class T, int V, int ST>
inline Pool2Fn<T> pick2v()
{
  return hydro2d_pool_step<T, V2[V][0], V2[V][1], V2[V][2], RS_HLLC, ST>;
}

template <class T>
inline Pool2Fn<T> pick2(int variant, int st)
{
  switch (variant * 2 + (st == 1 ? 0 : 1)) {
    case 0: return pick2v<T, 0, 1>();
    case 1: return pick2v<T, 0, 2>();
    case 2: return pick2v<T, 1, 1>();
    case 3: return pick2v<T, 1, 2>();
    case 4: return pick2v<T, 2, 1>();
    case 5: return pick2v<T, 2, 2>();
    case 6: return pick2v<T, 3, 1>();
    default: return pick2v<T, 3, 2>();
  }
}

template <class T>
inline Pool2Fn<T> pick3(int st)
{
  if (st == 1)
    return hydro3d_pool_step<T, P3X, P3Y, P3T, RS_HLLC, 1>;
  return hydro3d_pool_step<T, P3X, P3Y, P3T, RS_HLLC, 2>;
}

// blocks of `variant` that fit on one CU at once
template <class T>
inline int occupancy(int dim, int variant, int st)
{
  int n = 0;
  if (dim == 2) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(pick2<T>(variant, st)),
                                                     V2[variant][2], 0) != hipSuccess)
      return 1;
  } else {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(pick3<T>(st)), P3T, 0) !=
        hipSuccess)
      return 1;
  }
  return n > 0 ? n : 1;
}

template <class T>
inline hipError_t launch_step(int dim, int variant, int st, const T* Ua, T* Ub, const PoolJob<T>* jobs,
                              const PoolWork* work, int nwork, TimeSlot* ts, int slot, int grid, hipStream_t s)
{
  if (nwork <= 0)
    return hipSuccess;
  if (grid > nwork)
    grid = nwork;
  if (dim == 2)
    hipLaunchKernelGGL(pick2<T>(variant, st), dim3(grid), dim3(V2[variant][2]), 0, s, Ua, Ub, jobs, work, nwork,
                       ts, slot);
  else
    hipLaunchKernelGGL(pick3<T>(st), dim3(grid), dim3(P3T), 0, s, Ua, Ub, jobs, work, nwork, ts, slot);
  return hipGetLastError();
}

} // namespace pool
} // namespace ek
