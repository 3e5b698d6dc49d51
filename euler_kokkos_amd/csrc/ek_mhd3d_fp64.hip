// ek_mhd3d_fp64.hip -- 3D MHD step pipeline, double instantiation.
#include "ek_internal.h"
#include "ek_mhd3d.h"
#include "ek_pool_mhd3d.h"

namespace ek {
namespace fp64 {

template <int RS, int ST>
static hipError_t launch3(const double* Uin, double* Uout, double* S, const mhd3::Work3& w, const Geom<double>& geo,
                          const Gas<double>& gas, const StepCtl& ctl, hipStream_t s)
{
  const unsigned nb = (unsigned)((w.PN + 255) / 256);
  const long long nu = (geo.vstride + 255) / 256;
  const unsigned ni = (unsigned)(nu < mhd3::MHD3_UPDATE_BLOCKS ? nu : mhd3::MHD3_UPDATE_BLOCKS);
  hipLaunchKernelGGL(mhd3::padprims_kernel<double>, dim3(nb), dim3(256), 0, s, Uin, S, w, geo, gas, ctl);
  hipLaunchKernelGGL((mhd3::efield_kernel<double, ST>), dim3(nb), dim3(256), 0, s, S, w, ctl);
  hipLaunchKernelGGL((mhd3::trace_kernel<double, ST>), dim3(nb), dim3(256), 0, s, S, w, geo, gas, ctl);
  const unsigned nfe = 32u * ((nb + 7u) / 8u); // 4 tasks per chunk, chunks in groups of 8
  hipLaunchKernelGGL((mhd3::fluxemf_kernel<double, RS>), dim3(nfe), dim3(256), 0, s, S, w, gas, ctl);
  hipLaunchKernelGGL(mhd3::update_kernel<double>, dim3(ni), dim3(256), 0, s, S, Uout, w, geo, gas, ctl);
  return hipGetLastError();
}

hipError_t mhd3_step(int rs, int st, const double* Uin, double* Uout, double* scratch, const Geom<double>& geo,
                     const Gas<double>& gas, const StepCtl& ctl, hipStream_t s)
{
  mhd3::Work3 w;
  w.nx = geo.nx;
  w.ny = geo.ny;
  w.nz = geo.nz;
  w.PX = geo.nx + 6;
  w.PY = geo.ny + 6;
  w.PZ = geo.nz + 6;
  w.PN = (long long)w.PX * w.PY * w.PZ;
  const bool st1 = st == 1;
  switch (rs) {
    case RS_HLL: return st1 ? launch3<RS_HLL, 1>(Uin, Uout, scratch, w, geo, gas, ctl, s)
                            : launch3<RS_HLL, 2>(Uin, Uout, scratch, w, geo, gas, ctl, s);
    case RS_LLF: return st1 ? launch3<RS_LLF, 1>(Uin, Uout, scratch, w, geo, gas, ctl, s)
                            : launch3<RS_LLF, 2>(Uin, Uout, scratch, w, geo, gas, ctl, s);
    default: return st1 ? launch3<RS_HLLD, 1>(Uin, Uout, scratch, w, geo, gas, ctl, s)
                        : launch3<RS_HLLD, 2>(Uin, Uout, scratch, w, geo, gas, ctl, s);
  }
}

hipError_t mhd3_inv_dt(const double* U, const Geom<double>& geo, const Gas<double>& gas, unsigned long long* dst,
                       unsigned long long* spread, hipStream_t s)
{
  long long nb = (geo.vstride + 255) / 256;
  if (nb > 4096)
    nb = 4096;
  hipLaunchKernelGGL(mhd3::inv_dt3_kernel<double>, dim3((unsigned)nb), dim3(256), 0, s, U, geo, gas, dst, spread);
  return hipGetLastError();
}

long long mhd3_scratch_elems(int nx, int ny, int nz)
{
  return (long long)mhd3::A_TOTAL * (nx + 6) * (ny + 6) * (long long)(nz + 6);
}

hipError_t pool_mhd3_step(int rs, int st, const double* Ua, double* Ub, double* W, const PoolJobM<double>* jobs,
                          const PoolWork* wp, int np, const PoolWork* wu, int nu, TimeSlot* ts, int slot,
                          const int* grid, hipStream_t s)
{
  return pool::mhd3_step<double>(rs, st, Ua, Ub, W, jobs, wp, np, wu, nu, ts, slot, grid, s);
}

void pool_mhd3_grids(int rs, int st, int cus, int* grid)
{
  pool::mhd3_grids_for<double>(rs, st, cus, grid);
}

} // namespace fp64
} // namespace ek
