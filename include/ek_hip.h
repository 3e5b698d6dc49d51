/*
 * ek_hip.h -- C ABI of the MI355X-native MUSCL-Hancock Godunov step (libek_hip.so).
 *
 * The library replaces the device side of euler_kokkos' MUSCL solvers -- hydro
 * (SolverHydroMuscl<2|3>, src/muscl/SolverHydroMuscl.{h,cpp}) and MHD
 * (SolverMHDMuscl<2|3>, src/muscl/SolverMHDMuscl.{h,cpp}): the Kokkos functors of one
 * impl-0 godunov step, the ghost fill and the CFL reduction. A host solver keeps the
 * reference's Solver/ConfigMap interface and calls these entry points (see the C++ host
 * in euler_kokkos_amd/host/). Paths cited are relative to the euler_kokkos source tree.
 *
 * Conventions: every int-returning call returns 0 on success and a nonzero EK_E* code on
 * failure, with a message available from ek_last_error(ctx). No C++ exception crosses
 * this boundary. One context per host thread; contexts are independent of each other.
 */
#ifndef EK_HIP_H_
#define EK_HIP_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EK_OK 0
#define EK_EINVAL 1   /* invalid argument or unsupported configuration */
#define EK_EHIP 2     /* HIP runtime error */
#define EK_ENOMEM 3   /* device allocation failed */

/* precision of the device state */
#define EK_FP64 0 /* bit-exact with the reference's CPU results */
#define EK_FP32 1 /* FP32 state and fluxes, FP64 dt and budgets */

/* host array layouts for ek_upload / ek_download */
#define EK_LAYOUT_REFERENCE 0 /* Kokkos OpenMP LayoutRight U(i,j[,k],var), ghosts included, FP64 */
#define EK_LAYOUT_SOA 1       /* U[var][k][j][i], interior cells only, FP64 */

typedef struct ek_ctx ek_ctx;

/*
 * POD mirror of HydroParams / HydroSettings (src/shared/HydroParams.h:32-69,81-197) after
 * HydroParams::init (src/shared/HydroParams.cpp:326-359). Reals carry the values the
 * reference's ConfigMap::getFloat produces (float-parsed, then widened to double).
 */
typedef struct {
  int dim;       /* 2 or 3 */
  int nx, ny, nz;
  int gw;        /* ghostWidth: 2 (hydro), 3 (MHD) */
  int nbvar;     /* 4 (hydro 2D), 5 (hydro 3D), 8 (MHD) */
  int mhd;       /* 0 hydro, 1 MHD (2D: HLLD/HLL/LLF fluxes + constrained transport) */
  int riemann;   /* 0 approx, 1 llf, 2 hll, 3 hllc, 4 hlld (src/shared/enums.h:142-149) */
  int bc[6];     /* xmin,xmax,ymin,ymax,zmin,zmax: 1 reflecting, 2 outflow, 3 periodic */
  int precision; /* EK_FP64 or EK_FP32 */
  int reserved;
  double xmin, ymin, zmin, xmax, ymax, zmax;
  double dx, dy, dz;
  double gamma0, gamma6, cfl, slope_type, smallr, smallc, smallp, smallpp, cIso;
} ek_params;

/* Allocates the device state (two state buffers, time slots) on the current HIP device.
 * Replaces the View allocations of the SolverHydroMuscl constructor
 * (src/muscl/SolverHydroMuscl.h:163-274). */
int ek_create(const ek_params* params, ek_ctx** out);

/* Frees everything ek_create allocated. Safe on NULL. */
void ek_destroy(ek_ctx* ctx);

/* Message of the last failure on this context (or of ek_create when ctx is NULL). */
const char* ek_last_error(const ek_ctx* ctx);

/* Host -> device copy of the current state (replaces Kokkos::deep_copy(U, Uhost)). */
int ek_upload(ek_ctx* ctx, const double* host, int layout);

/* Device -> host copy of the current state. In EK_LAYOUT_REFERENCE the ghost cells are
 * filled as SolverBase::make_boundaries would fill them from the current interior
 * (src/shared/SolverBase.cpp:489-512, src/shared/BoundariesFunctors.h). */
int ek_download(ek_ctx* ctx, double* host, int layout);

/* CFL reduction on the current state: *inv_dt = max over interior cells of
 * (c+|u|)/dx + (c+|v|)/dy [+ (c+|w|)/dz]. Replaces ComputeDtFunctor2D/3D::apply
 * (src/muscl/HydroRunFunctors2D.h:18-82, HydroRunFunctors3D.h:16-87). For MHD the speeds
 * are the fast magnetosonic ones of the primitives (convertToPrimitives +
 * ComputeDtFunctor2D_MHD, src/muscl/MHDRunFunctors2D.h:16-159). Synchronizes. */
int ek_inv_dt(ek_ctx* ctx, double* inv_dt);

/* One godunov_unsplit step with a host-chosen dt (SolverHydroMuscl::godunov_unsplit,
 * src/muscl/SolverHydroMuscl.h:686-700 and .cpp:214-390): ghost fill, primitives,
 * slopes/trace, Riemann fluxes and update, fused in one launch. For MHD the launch is
 * SolverMHDMuscl<2>::godunov_unsplit_impl (src/muscl/SolverMHDMuscl.cpp:369-454) after
 * its compute_dt: trace, face fluxes, edge EMFs, flux and EMF updates. Asynchronous. */
int ek_step(ek_ctx* ctx, double dt);

/* Up to nsteps iterations of the reference time loop (src/main.cpp:126-131 with
 * SolverBase::finished/next_iteration/compute_dt, src/shared/SolverBase.cpp:124-195) with
 * the time step kept on the device: each launch computes dt = cfl/invDt (clamped to
 * tEnd), advances the state and reduces the next invDt. On entry *t_io / *it_io hold the
 * current time and iteration; on exit they hold the values after the last iteration run,
 * and *dt_last the last dt. Stops early (without error) when t >= tEnd-1e-14 or
 * it >= nStepmax. Launches are replayed from a captured HIP graph (one per state buffer the
 * run starts from, kept across calls). nsteps = 0 launches nothing and leaves *t_io, *it_io
 * and *dt_last as they are; it only captures the graph for this start buffer, nStepmax and
 * tEnd if it is not there yet, so that the capture stays out of a following timed run.
 * Synchronizes (nsteps > 0). */
int ek_run(ek_ctx* ctx, int nsteps, int nStepmax, double tEnd, double* t_io, int* it_io,
           double* dt_last);

/* Size of the fixed device array behind ek_run_noalloc (bytes, per device). */
#define EK_NOALLOC_SCRATCH_BYTES (16u << 20)

/* The time loop of ek_run on caller-owned device state, allocating nothing. No reference
 * counterpart: it is for callers that cannot allocate device memory (code recorded under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * U holds the state on entry in EK_LAYOUT_SOA order (interior cells only,
 * U[var][k][j][i]) in the params' precision (double for EK_FP64, float for EK_FP32); U2 is
 * a second device buffer of the same size. Steps alternate between the two; on return
 * *in_U2 is 1 when the final state is in U2, else 0. t_io / it_io / dt_last / nStepmax /
 * tEnd mean what they mean for ek_run. Kernels are launched on `stream` (a hipStream_t, NULL
 * for the default stream), one launch per step (no graph), and the call synchronizes with
 * that stream to read the final time. The time slots, CFL words and, for 3D MHD, the padded
 * work area live in a fixed device array of EK_NOALLOC_SCRATCH_BYTES per device, held from
 * the library's first use on that device until the process ends: a 3D MHD grid whose work
 * area (78 reals per ghost-padded cell) does not fit is rejected with EK_EINVAL. Every call
 * on one device uses the same array, so calls that overlap on one device race; each call
 * clears what it needs first. Errors are reported through ek_last_error(NULL). */
int ek_run_noalloc(const ek_params* params, void* U, void* U2, int nsteps, int nStepmax, double tEnd,
                   double* t_io, int* it_io, double* dt_last, int* in_U2, void* stream);

/* Conservation budgets: Neumaier-compensated FP64 sums over the interior of every
 * conserved variable (rho, E, rho*u, rho*v[, rho*w, Bx, By, Bz]) into sums[0..nbvar-1].
 * Synchronizes. (No reference counterpart; used for round-off conservation checks.) */
int ek_budgets(ek_ctx* ctx, double* sums);

/* Blocks until all work queued on the context's stream has finished. */
int ek_sync(ek_ctx* ctx);

/* Device time of the step launches of the last ek_run, from HIP events recorded on the
 * context's stream around them (in ms), and the number of launches it covered.
 * (No reference counterpart: the reference times whole phases on the host,
 * src/shared/SolverBase.h:35-45.) */
int ek_last_run_time(const ek_ctx* ctx, double* ms, int* launches);

/* Bytes of device memory held by the context. */
long long ek_device_bytes(const ek_ctx* ctx);

/* Page-locked host memory (hipHostMalloc) for states that are uploaded or downloaded:
 * ek_upload / ek_download / ek_pool_admit / ek_pool_download copy from and to it at the
 * link's full speed instead of through the runtime's bounce buffers. Returns EK_OK or
 * EK_ENOMEM; free with ek_host_free. (No reference counterpart: the reference's host
 * mirrors are Kokkos HostSpace views.) */
int ek_host_alloc(size_t bytes, void** ptr);
int ek_host_free(void* ptr);

/* Properties of the current HIP device that measurements report: compute units, memory
 * clock (kHz) and bus width (bits) as the runtime gives them, peak engine clock (kHz), and
 * the device name with its architecture (name_len bytes, NUL-terminated). (No reference
 * counterpart: measurement support.) */
int ek_device_info(int* cus, int* mem_clock_khz, int* bus_width_bits, int* sclk_khz, char* name, int name_len);

#ifdef __cplusplus
}
#endif

#endif /* EK_HIP_H_ */
