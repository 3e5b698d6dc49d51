// ek_mhd_fp64.hip -- 2D MHD kernels, double instantiation. Built with -ffp-contract=off (bit-exact path).
#include "ek_internal.h"
#include "ek_mhd_launch.h"

namespace ek {
namespace fp64 {
hipError_t mhd_step(int rs, int st, const double* Uin, double* Uout, const Geom<double>& geo, const Gas<double>& gas,
                    const StepCtl& ctl, hipStream_t s)
{
  return mhd::launch_step<double>(rs, st, Uin, Uout, geo, gas, ctl, s);
}
hipError_t mhd_inv_dt(const double* U, const Geom<double>& geo, const Gas<double>& gas, unsigned long long* dst,
                      unsigned long long* spread, hipStream_t s)
{
  return mhd::launch_inv_dt<double>(U, geo, gas, dst, spread, s);
}
} // namespace fp64
} // namespace ek
