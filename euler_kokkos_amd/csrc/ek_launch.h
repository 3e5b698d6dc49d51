// ek_launch.h -- kernel instantiation and launch for one real type. Included by exactly one
// translation unit per precision: ek_fp64.hip (built with -ffp-contract=off, the bit-exact
// path) and ek_fp32.hip (FP32 bulk path, contraction allowed).
#pragma once

#include "ek_hydro.h"
#include "ek_hydro3d.h"

namespace ek {

// Tile shapes. 2D: 16x16 cells per block (256 threads). 3D: T3X x T3Y column tiles marching
// over z-chunks (ek_hydro3d.h; N3T threads: one per primitive column of the tile and its
// 2-cell halo).
constexpr int T2X = 16, T2Y = 16, N2T = 256;

template <class T>
using Step2Fn = void (*)(const T*, T*, Geom<T>, Gas<T>, StepCtl);
template <class T>
using Step3Fn = void (*)(const T*, T*, Geom<T>, Gas<T>, StepCtl, int, int, int, int);

template <class T, int RS>
inline Step2Fn<T> pick2(int st)
{
  if (st == 0)
    return hydro2d_step<T, T2X, T2Y, N2T, RS, 0>;
  if (st == 1)
    return hydro2d_step<T, T2X, T2Y, N2T, RS, 1>;
  return hydro2d_step<T, T2X, T2Y, N2T, RS, 2>;
}
template <class T, int RS>
inline Step3Fn<T> pick3(int st)
{
  if (st == 0)
    return hydro3d_step<T, T3X, T3Y, N3T, RS, 0>;
  if (st == 1)
    return hydro3d_step<T, T3X, T3Y, N3T, RS, 1>;
  return hydro3d_step<T, T3X, T3Y, N3T, RS, 2>;
}

template <class T>
inline hipError_t launch_step(int dim, int rs, int st, const T* Uin, T* Uout, const Geom<T>& geo,
                              const Gas<T>& gas, const StepCtl& ctl, int cz, hipStream_t s)
{
  if (dim == 2) {
    Step2Fn<T> f;
    switch (rs) {
      case RS_APPROX: f = pick2<T, RS_APPROX>(st); break;
      case RS_LLF: f = pick2<T, RS_LLF>(st); break;
      case RS_HLL: f = pick2<T, RS_HLL>(st); break;
      default: f = pick2<T, RS_HLLC>(st); break;
    }
    dim3 grid((geo.nx + T2X - 1) / T2X, (geo.ny + T2Y - 1) / T2Y, 1);
    hipLaunchKernelGGL(f, grid, dim3(N2T), 0, s, Uin, Uout, geo, gas, ctl);
  } else {
    Step3Fn<T> f;
    switch (rs) {
      case RS_APPROX: f = pick3<T, RS_APPROX>(st); break;
      case RS_LLF: f = pick3<T, RS_LLF>(st); break;
      default: f = pick3<T, RS_HLLC>(st); break;
    }
    const int ntx = (geo.nx + T3X - 1) / T3X, nty = (geo.ny + T3Y - 1) / T3Y;
    const int nch = (geo.nz + cz - 1) / cz;
    hipLaunchKernelGGL(f, dim3(ntx * nty * nch), dim3(N3T), 0, s, Uin, Uout, geo, gas, ctl, cz, ntx,
                       nty, nch);
  }
  return hipGetLastError();
}

template <class T>
inline hipError_t launch_inv_dt(int dim, const T* U, const Geom<T>& geo, const Gas<T>& gas,
                                unsigned long long* dst, unsigned long long* spread, hipStream_t s)
{
  const long long n = geo.vstride;
  long long nb = (n + 255) / 256;
  if (nb > 4096)
    nb = 4096;
  if (dim == 2)
    hipLaunchKernelGGL((inv_dt_kernel<T, 2, 256>), dim3((unsigned)nb), dim3(256), 0, s, U, geo, gas, dst, spread, n);
  else
    hipLaunchKernelGGL((inv_dt_kernel<T, 3, 256>), dim3((unsigned)nb), dim3(256), 0, s, U, geo, gas, dst, spread, n);
  return hipGetLastError();
}

template <class T>
inline hipError_t launch_budget(int nv, const T* U, long long ncell, long long vstride, double* partial,
                                int nblocks, hipStream_t s)
{
  if (nv == 4)
    hipLaunchKernelGGL((budget_kernel<T, 4, 256>), dim3(nblocks), dim3(256), 0, s, U, ncell, vstride, partial);
  else if (nv == 5)
    hipLaunchKernelGGL((budget_kernel<T, 5, 256>), dim3(nblocks), dim3(256), 0, s, U, ncell, vstride, partial);
  else
    hipLaunchKernelGGL((budget_kernel<T, 8, 256>), dim3(nblocks), dim3(256), 0, s, U, ncell, vstride, partial);
  return hipGetLastError();
}

template <class T>
inline hipError_t launch_ref_to_soa(int dim, const double* R, T* U, const Geom<T>& geo, int gw, int nbvar,
                                    int isize, int jsize, int ksize, hipStream_t s)
{
  long long nb = (geo.vstride + 255) / 256;
  if (nb > 65535)
    nb = 65535;
  if (dim == 2)
    hipLaunchKernelGGL((ref_to_soa<T, 2>), dim3((unsigned)nb), dim3(256), 0, s, R, U, geo, gw, nbvar, isize, jsize, ksize);
  else
    hipLaunchKernelGGL((ref_to_soa<T, 3>), dim3((unsigned)nb), dim3(256), 0, s, R, U, geo, gw, nbvar, isize, jsize, ksize);
  return hipGetLastError();
}

template <class T>
inline hipError_t launch_soa_to_ref(int dim, const T* U, double* R, const Geom<T>& geo, int gw, int nbvar,
                                    int isize, int jsize, int ksize, int mhd, hipStream_t s)
{
  const long long n = (long long)isize * jsize * (dim == 3 ? ksize : 1);
  long long nb = (n + 255) / 256;
  if (nb > 65535)
    nb = 65535;
  if (dim == 2)
    hipLaunchKernelGGL((soa_to_ref<T, 2>), dim3((unsigned)nb), dim3(256), 0, s, U, R, geo, gw, nbvar, isize, jsize, ksize, mhd);
  else
    hipLaunchKernelGGL((soa_to_ref<T, 3>), dim3((unsigned)nb), dim3(256), 0, s, U, R, geo, gw, nbvar, isize, jsize, ksize, mhd);
  return hipGetLastError();
}

} // namespace ek
