// ek_mhd_fp32.hip -- 2D MHD kernels, float instantiation. FP32 bulk path.
#include "ek_internal.h"
#include "ek_mhd_launch.h"

namespace ek {
namespace fp32 {
hipError_t mhd_step(int rs, int st, const float* Uin, float* Uout, const Geom<float>& geo, const Gas<float>& gas,
                    const StepCtl& ctl, hipStream_t s)
{
  return mhd::launch_step<float>(rs, st, Uin, Uout, geo, gas, ctl, s);
}
hipError_t mhd_inv_dt(const float* U, const Geom<float>& geo, const Gas<float>& gas, unsigned long long* dst,
                      unsigned long long* spread, hipStream_t s)
{
  return mhd::launch_inv_dt<float>(U, geo, gas, dst, spread, s);
}
} // namespace fp32
} // namespace ek
