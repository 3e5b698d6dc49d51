// ek_pool_fp32.hip -- job-pool kernels, float instantiation.
#include "ek_internal.h"
#include "ek_pool_launch.h"

namespace ek {
namespace fp32 {
hipError_t pool_step(int dim, int variant, int st, const float* Ua, float* Ub, const PoolJob<float>* jobs,
                     const PoolWork* work, int nwork, TimeSlot* ts, int slot, int grid, hipStream_t s)
{
  return pool::launch_step<float>(dim, variant, st, Ua, Ub, jobs, work, nwork, ts, slot, grid, s);
}
int pool_occupancy(int dim, int variant, int st)
{
  return pool::occupancy<float>(dim, variant, st);
}
} // namespace fp32
} // namespace ek
