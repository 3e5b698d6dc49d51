// ek_internal.h -- per-precision launch entry points, implemented in ek_fp64.hip
// (bit-exact FP64) and ek_fp32.hip (FP32 bulk), the 2D MHD ones in ek_mhd_fp64.hip /
// ek_mhd_fp32.hip; called by the C ABI in ek_capi.hip.
#pragma once

#include "ek_hydro.h"
#include "ek_pool_kernels.h"

#include <string>

#include "ek_hip.h"

namespace ek {
template <class T>
struct PoolJobM;  // 3D MHD job record of the job pool (ek_pool_mhd3d.h)

// Checks an ek_params against what this library provides (ek_capi.hip); 0 when valid.
int validate_params(const ek_params* p, std::string& why);
// FP32 copy of n FP64 values already on the device (round to nearest, denormals kept, as
// the host's (float) conversion); ek_capi.hip.
hipError_t cast_f64_f32(const double* src, float* dst, long long n, hipStream_t s);
// n FP64 host values into the FP32 device array dst: chunks through the FP64 device buffer
// stage of stage_elems values, rounded on the device; ek_capi.hip.
hipError_t upload_f64_as_f32(const double* host, float* dst, long long n, double* stage, long long stage_elems,
                             hipStream_t s);

namespace fp64 {
hipError_t step(int dim, int rs, int st, const double* Uin, double* Uout, const Geom<double>& geo,
                const Gas<double>& gas, const StepCtl& ctl, int cz, hipStream_t s);
hipError_t inv_dt(int dim, const double* U, const Geom<double>& geo, const Gas<double>& gas,
                  unsigned long long* dst, unsigned long long* spread, hipStream_t s);
hipError_t budget(int nv, const double* U, long long ncell, long long vstride, double* partial,
                  int nblocks, hipStream_t s);
hipError_t ref_to_soa(int dim, const double* R, double* U, const Geom<double>& geo, int gw, int nbvar,
                      int isize, int jsize, int ksize, hipStream_t s);
hipError_t soa_to_ref(int dim, const double* U, double* R, const Geom<double>& geo, int gw, int nbvar,
                      int isize, int jsize, int ksize, int mhd, hipStream_t s);
// 2D MHD (ek_mhd_fp64.hip)
hipError_t mhd_step(int rs, int st, const double* Uin, double* Uout, const Geom<double>& geo, const Gas<double>& gas,
                    const StepCtl& ctl, hipStream_t s);
hipError_t mhd_inv_dt(const double* U, const Geom<double>& geo, const Gas<double>& gas, unsigned long long* dst,
                      unsigned long long* spread, hipStream_t s);
// job pool (ek_pool_fp64.hip)
hipError_t pool_step(int dim, int variant, int st, const double* Ua, double* Ub, const PoolJob<double>* jobs,
                     const PoolWork* work, int nwork, TimeSlot* ts, int slot, int grid, hipStream_t s);
int pool_occupancy(int dim, int variant, int st);
// 3D MHD (ek_mhd3d_fp64.hip): a pipeline over a ghost-padded work area `scratch` of
// mhd3_scratch_elems(nx, ny, nz) reals
hipError_t mhd3_step(int rs, int st, const double* Uin, double* Uout, double* scratch, const Geom<double>& geo,
                     const Gas<double>& gas, const StepCtl& ctl, hipStream_t s);
hipError_t mhd3_inv_dt(const double* U, const Geom<double>& geo, const Gas<double>& gas, unsigned long long* dst,
                       unsigned long long* spread, hipStream_t s);
long long mhd3_scratch_elems(int nx, int ny, int nz);
// 3D MHD jobs of the job pool (ek_mhd3d_fp64.hip): the five pipeline kernels over the work
// lists wp (padded-grid chunks) and wu (interior chunks); grid[5] from pool_mhd3_grids
hipError_t pool_mhd3_step(int rs, int st, const double* Ua, double* Ub, double* W, const PoolJobM<double>* jobs,
                          const PoolWork* wp, int np, const PoolWork* wu, int nu, TimeSlot* ts, int slot,
                          const int* grid, hipStream_t s);
void pool_mhd3_grids(int rs, int st, int cus, int* grid);
} // namespace fp64
namespace fp32 {
hipError_t step(int dim, int rs, int st, const float* Uin, float* Uout, const Geom<float>& geo,
                const Gas<float>& gas, const StepCtl& ctl, int cz, hipStream_t s);
hipError_t inv_dt(int dim, const float* U, const Geom<float>& geo, const Gas<float>& gas,
                  unsigned long long* dst, unsigned long long* spread, hipStream_t s);
hipError_t budget(int nv, const float* U, long long ncell, long long vstride, double* partial,
                  int nblocks, hipStream_t s);
hipError_t ref_to_soa(int dim, const double* R, float* U, const Geom<float>& geo, int gw, int nbvar,
                      int isize, int jsize, int ksize, hipStream_t s);
hipError_t soa_to_ref(int dim, const float* U, double* R, const Geom<float>& geo, int gw, int nbvar,
                      int isize, int jsize, int ksize, int mhd, hipStream_t s);
// 2D MHD (ek_mhd_fp32.hip)
hipError_t mhd_step(int rs, int st, const float* Uin, float* Uout, const Geom<float>& geo, const Gas<float>& gas,
                    const StepCtl& ctl, hipStream_t s);
hipError_t mhd_inv_dt(const float* U, const Geom<float>& geo, const Gas<float>& gas, unsigned long long* dst,
                      unsigned long long* spread, hipStream_t s);
// job pool (ek_pool_fp32.hip)
hipError_t pool_step(int dim, int variant, int st, const float* Ua, float* Ub, const PoolJob<float>* jobs,
                     const PoolWork* work, int nwork, TimeSlot* ts, int slot, int grid, hipStream_t s);
int pool_occupancy(int dim, int variant, int st);
// 3D MHD (ek_mhd3d_fp32.hip): a pipeline over a ghost-padded work area `scratch` of
// mhd3_scratch_elems(nx, ny, nz) reals
hipError_t mhd3_step(int rs, int st, const float* Uin, float* Uout, float* scratch, const Geom<float>& geo,
                     const Gas<float>& gas, const StepCtl& ctl, hipStream_t s);
hipError_t mhd3_inv_dt(const float* U, const Geom<float>& geo, const Gas<float>& gas, unsigned long long* dst,
                       unsigned long long* spread, hipStream_t s);
long long mhd3_scratch_elems(int nx, int ny, int nz);
// 3D MHD jobs of the job pool (ek_mhd3d_fp32.hip): the five pipeline kernels over the work
// lists wp (padded-grid chunks) and wu (interior chunks); grid[5] from pool_mhd3_grids
hipError_t pool_mhd3_step(int rs, int st, const float* Ua, float* Ub, float* W, const PoolJobM<float>* jobs,
                          const PoolWork* wp, int np, const PoolWork* wu, int nu, TimeSlot* ts, int slot,
                          const int* grid, hipStream_t s);
void pool_mhd3_grids(int rs, int st, int cus, int* grid);
} // namespace fp32
} // namespace ek
