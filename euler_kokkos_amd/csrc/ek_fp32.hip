// ek_fp32.hip -- FP32 bulk path: FP32 state and fluxes, contraction allowed; dt and budgets stay FP64.
#include "ek_internal.h"
#include "ek_launch.h"

namespace ek {
namespace fp32 {
hipError_t step(int dim, int rs, int st, const float* Uin, float* Uout, const Geom<float>& geo,
                const Gas<float>& gas, const StepCtl& ctl, int cz, hipStream_t s)
{
  return launch_step<float>(dim, rs, st, Uin, Uout, geo, gas, ctl, cz, s);
}
hipError_t inv_dt(int dim, const float* U, const Geom<float>& geo, const Gas<float>& gas,
                  unsigned long long* dst, unsigned long long* spread, hipStream_t s)
{
  return launch_inv_dt<float>(dim, U, geo, gas, dst, spread, s);
}
hipError_t budget(int nv, const float* U, long long ncell, long long vstride, double* partial,
                  int nblocks, hipStream_t s)
{
  return launch_budget<float>(nv, U, ncell, vstride, partial, nblocks, s);
}
hipError_t ref_to_soa(int dim, const double* R, float* U, const Geom<float>& geo, int gw, int nbvar,
                      int isize, int jsize, int ksize, hipStream_t s)
{
  return launch_ref_to_soa<float>(dim, R, U, geo, gw, nbvar, isize, jsize, ksize, s);
}
hipError_t soa_to_ref(int dim, const float* U, double* R, const Geom<float>& geo, int gw, int nbvar,
                      int isize, int jsize, int ksize, int mhd, hipStream_t s)
{
  return launch_soa_to_ref<float>(dim, U, R, geo, gw, nbvar, isize, jsize, ksize, mhd, s);
}
} // namespace fp32
} // namespace ek
