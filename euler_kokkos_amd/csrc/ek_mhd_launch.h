// ek_mhd_launch.h -- instantiation and launch of the 2D MHD kernels for one real type.
// Included by exactly one translation unit per precision (ek_mhd_fp64.hip, ek_mhd_fp32.hip).
#pragma once

#include "ek_mhd.h"

namespace ek {
namespace mhd {

// 14x14 cells per block: the 16x16 trace region gives one record per thread (256).
constexpr int M2X = 14, M2Y = 14, M2T = 256;

template <class T>
using MStepFn = void (*)(const T*, T*, Geom<T>, Gas<T>, StepCtl);

template <class T, int RS>
inline MStepFn<T> pick(int st)
{
  if (st == 1)
    return mhd2d_step<T, M2X, M2Y, M2T, RS, 1>;
  return mhd2d_step<T, M2X, M2Y, M2T, RS, 2>;
}

template <class T>
inline hipError_t launch_step(int rs, int st, const T* Uin, T* Uout, const Geom<T>& geo, const Gas<T>& gas,
                              const StepCtl& ctl, hipStream_t s)
{
  MStepFn<T> f;
  switch (rs) {
    case RS_HLL: f = pick<T, RS_HLL>(st); break;
    case RS_LLF: f = pick<T, RS_LLF>(st); break;
    default: f = pick<T, RS_HLLD>(st); break;
  }
  dim3 grid((geo.nx + M2X - 1) / M2X, (geo.ny + M2Y - 1) / M2Y, 1);
  hipLaunchKernelGGL(f, grid, dim3(M2T), 0, s, Uin, Uout, geo, gas, ctl);
  return hipGetLastError();
}

template <class T>
inline hipError_t launch_inv_dt(const T* U, const Geom<T>& geo, const Gas<T>& gas, unsigned long long* dst,
                                unsigned long long* spread, hipStream_t s)
{
  long long nb = (geo.vstride + 255) / 256;
  if (nb > 4096)
    nb = 4096;
  hipLaunchKernelGGL((mhd2d_inv_dt_kernel<T, 256>), dim3((unsigned)nb), dim3(256), 0, s, U, geo, gas, dst, spread);
  return hipGetLastError();
}

} // namespace mhd
} // namespace ek
