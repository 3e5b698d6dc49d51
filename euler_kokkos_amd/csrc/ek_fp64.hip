// ek_fp64.hip -- FP64 path: built with -ffp-contract=off so that every multiply and add rounds separately, as in the reference's CPU build.
#include "ek_internal.h"
#include "ek_launch.h"

namespace ek {
namespace fp64 {
hipError_t step(int dim, int rs, int st, const double* Uin, double* Uout, const Geom<double>& geo,
                const Gas<double>& gas, const StepCtl& ctl, int cz, hipStream_t s)
{
  return launch_step<double>(dim, rs, st, Uin, Uout, geo, gas, ctl, cz, s);
}
hipError_t inv_dt(int dim, const double* U, const Geom<double>& geo, const Gas<double>& gas,
                  unsigned long long* dst, unsigned long long* spread, hipStream_t s)
{
  return launch_inv_dt<double>(dim, U, geo, gas, dst, spread, s);
}
hipError_t budget(int nv, const double* U, long long ncell, long long vstride, double* partial,
                  int nblocks, hipStream_t s)
{
  return launch_budget<double>(nv, U, ncell, vstride, partial, nblocks, s);
}
hipError_t ref_to_soa(int dim, const double* R, double* U, const Geom<double>& geo, int gw, int nbvar,
                      int isize, int jsize, int ksize, hipStream_t s)
{
  return launch_ref_to_soa<double>(dim, R, U, geo, gw, nbvar, isize, jsize, ksize, s);
}
hipError_t soa_to_ref(int dim, const double* U, double* R, const Geom<double>& geo, int gw, int nbvar,
                      int isize, int jsize, int ksize, int mhd, hipStream_t s)
{
  return launch_soa_to_ref<double>(dim, U, R, geo, gw, nbvar, isize, jsize, ksize, mhd, s);
}
} // namespace fp64
} // namespace ek
