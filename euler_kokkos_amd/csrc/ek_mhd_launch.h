// ek_mhd_launch.h -- instantiation and launch of the 2D MHD kernels for one real type.
// Included by exactly one translation unit per precision (ek_mhd_fp64.hip, ek_mhd_fp32.hip).
#pragma once

#include "ek_mhd.h"

namespace ek {
namespace mhd {

// Tile of the 2D MHD step: the (TX+2) x (TY+2) trace region gives one record per thread.
// FP64: 14x14 cells, 256 threads (81 KB of LDS, two blocks per CU). FP32: 30x14, 512
// threads, compact LDS layout (53.7 KB, three blocks per CU, 71 VGPRs, six waves per SIMD).
// Per 1024^2 Orszag-Tang step: FP32 0.139 ms at 14x14, 0.097 at 30x14, 0.088 with the
// compact layout; FP64 0.246 ms at 14x14, 0.260 at 30x14, 0.249 with the compact layout at
// three waves per SIMD (every variant bitwise the same).
template <class T>
struct M2Tile {
  static constexpr bool F64 = sizeof(T) == 8;
  static constexpr int X = F64 ? 14 : 30, Y = 14, NT = F64 ? 256 : 512;
  static constexpr bool CMP = !F64;
  static constexpr int WPE = F64 ? 2 : 6;
};

template <class T>
using MStepFn = void (*)(const T*, T*, Geom<T>, Gas<T>, StepCtl);

template <class T, int RS>
inline MStepFn<T> pick(int st)
{
  using M = M2Tile<T>;
  if (st == 1)
    return mhd2d_step<T, M::X, M::Y, M::NT, RS, 1, M::CMP, M::WPE>;
  return mhd2d_step<T, M::X, M::Y, M::NT, RS, 2, M::CMP, M::WPE>;
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
  using M = M2Tile<T>;
  dim3 grid((geo.nx + M::X - 1) / M::X, (geo.ny + M::Y - 1) / M::Y, 1);
  hipLaunchKernelGGL(f, grid, dim3(M::NT), 0, s, Uin, Uout, geo, gas, ctl);
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
