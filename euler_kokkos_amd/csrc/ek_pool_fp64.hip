// ek_pool_fp64.hip -- job-pool kernels, double instantiation.
#include "ek_internal.h"
#include "ek_pool_launch.h"

namespace ek {
namespace fp64 {
hipError_t pool_step(int dim, int variant, int st, const double* Ua, double* Ub, const PoolJob<double>* jobs,
                     const PoolWork* work, int nwork, TimeSlot* ts, int slot, int grid, hipStream_t s)
{
  return pool::launch_step<double>(dim, variant, st, Ua, Ub, jobs, work, nwork, ts, slot, grid, s);
}
int pool_occupancy(int dim, int variant, int st)
{
  return pool::occupancy<double>(dim, variant, st);
}
} // namespace fp64
} // namespace ek
