// ek_capi.hip -- the C ABI of include/ek_hip.h: device state, uploads/downloads, the
// device-resident time loop (captured into a HIP graph) and the budget reductions.
#include "ek_hip.h"
#include "ek_internal.h"
#include "ek_geom_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace ek {

// FP64 -> FP32, round to nearest; this unit flushes no denormals, so the result is the
// host's (float) conversion
__global__ __launch_bounds__(256) void cast_f64_f32_kernel(const double* __restrict__ src, float* __restrict__ dst,
                                                           long long n)
{
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    dst[i] = (float)src[i];
}

hipError_t cast_f64_f32(const double* src, float* dst, long long n, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  const long long nb = std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(cast_f64_f32_kernel, dim3((unsigned)nb), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t upload_f64_as_f32(const double* host, float* dst, long long n, double* stage, long long stage_elems,
                             hipStream_t s)
{
  // one staging buffer: each chunk's copy is queued behind the previous chunk's rounding
  for (long long off = 0; off < n; off += stage_elems) {
    const long long len = std::min(stage_elems, n - off);
    hipError_t e = hipMemcpyAsync(stage, host + off, (size_t)len * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = cast_f64_f32(stage, dst + off, len, s);
    if (e != hipSuccess)
      return e;
  }
  return hipSuccess;
}

} // namespace ek

// staging chunk of FP32 uploads (FP64 values)
static constexpr long long kUploadChunk = 1LL << 22;

struct ek_ctx {
  ek_params p{};
  std::string err;
  hipStream_t stream = nullptr;
  int device = 0;
  void* U[2] = { nullptr, nullptr };
  int cur = 0;
  ek::TimeSlot* ts = nullptr;
  unsigned long long* scratch_bits = nullptr;
  unsigned long long* spread = nullptr; // extra CFL words of the three time slots
  double* partial = nullptr;
  double* stage = nullptr;   // FP64 staging buffer for reference-layout transfers
  void* work = nullptr;      // 3D MHD ghost-padded work area
  long long work_bytes = 0;
  long long stage_elems = 0;
  int nblocks_budget = 0;
  size_t elem = 8;
  long long ncell = 0;
  ek::Geom<double> g64{};
  ek::Gas<double> gas64{};
  ek::Geom<float> g32{};
  ek::Gas<float> gas32{};
  int st = 2;
  int cz = 32;
  long long bytes = 0;
  // captured 6-step graphs (6 = lcm of the 2 state buffers and the 3 time slots), one per
  // physical buffer the run starts from: a run of an odd number of steps leaves the state in
  // the other buffer, and the next run then replays the other graph instead of capturing anew
  struct Graph {
    hipGraphExec_t exec = nullptr;
    int nStepmax = -1;
    double tEnd = 0.0;
    void* U0 = nullptr;
  } graphs[2];
  hipEvent_t ev0 = nullptr, ev1 = nullptr; // bracket the step launches of the last ek_run
  double last_run_ms = 0.0;
  int last_run_launches = 0;
};

// message of the last failed ek_create on this thread (contexts may be created concurrently)
static thread_local std::string g_create_err;

static int fail(ek_ctx* c, int code, const std::string& msg)
{
  if (c)
    c->err = msg;
  else
    g_create_err = msg;
  return code;
}

static int hip_fail(ek_ctx* c, hipError_t e, const char* what)
{
  return fail(c, EK_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define EK_HIP(c, call)                      \
  do {                                       \
    hipError_t e_ = (call);                  \
    if (e_ != hipSuccess)                    \
      return hip_fail((c), e_, #call);       \
  } while (0)

static int validate_common(const ek_params* p, std::string& why)
{
  if (p->precision != EK_FP64 && p->precision != EK_FP32)
    return why = "precision must be EK_FP64 or EK_FP32", 1;
  for (int i = 0; i < 2 * p->dim; i++)
    if (p->bc[i] < 1 || p->bc[i] > 3)
      return why = "boundary types must be 1 (reflecting), 2 (outflow) or 3 (periodic)", 1;
  if (!(p->dx > 0) || !(p->dy > 0) || (p->dim == 3 && !(p->dz > 0)))
    return why = "cell sizes must be positive", 1;
  if (p->nx < p->gw || p->ny < p->gw || (p->dim == 3 && p->nz < p->gw))
    return why = "every axis needs at least ghostWidth interior cells", 1;
  return 0;
}

int ek::validate_params(const ek_params* p, std::string& why)
{
  if (p->dim != 2 && p->dim != 3)
    return why = "dim must be 2 or 3", 1;
  if (p->mhd) {
    // SolverMHDMuscl<2|3>, impl 0 (src/muscl/SolverMHDMuscl.cpp:369-512)
    if (p->gw != 3)
      return why = "MHD MUSCL needs ghostWidth 3", 1;
    if (p->nbvar != 8)
      return why = "nbvar must be 8 for MHD", 1;
    if (p->riemann != ek::RS_HLLD && p->riemann != ek::RS_HLL && p->riemann != ek::RS_LLF)
      return why = "MHD riemann solver must be hlld, hll or llf (riemann_mhd dispatch)", 1;
    if (!(p->slope_type == 1.0 || p->slope_type == 2.0))
      return why = "the MHD trace supports slope_type 1 and 2 only", 1;
    if (p->cIso > 0)
      return why = "isothermal MHD (cIso > 0) is not provided", 1;
    // the 3D pipeline addresses each work-area array with 32-bit byte offsets
    if (p->dim == 3 &&
        (double)(p->nx + 6) * (double)(p->ny + 6) * (double)(p->nz + 6) * (p->precision == EK_FP64 ? 8.0 : 4.0) >=
            4294967296.0)
      return why = "3D MHD grid too large: each padded array must stay below 4 GiB", 1;
    return validate_common(p, why);
  }
  if (p->gw != 2)
    return why = "hydro MUSCL needs ghostWidth 2", 1;
  if (p->nbvar != (p->dim == 2 ? 4 : 5))
    return why = "nbvar must be 4 (2D) or 5 (3D) for hydro", 1;
  if (p->riemann != ek::RS_APPROX && p->riemann != ek::RS_LLF && p->riemann != ek::RS_HLLC &&
      !(p->riemann == ek::RS_HLL && p->dim == 2))
    return why = "riemann solver not available for hydro in this dimension (3D has no HLL in the "
                 "reference dispatch; HLLD is MHD-only)",
           1;
  if (!(p->slope_type == 0.0 || p->slope_type == 1.0 || p->slope_type == 2.0))
    return why = "slope_type must be 0, 1 or 2", 1;
  return validate_common(p, why);
}

extern "C" int ek_create(const ek_params* params, ek_ctx** out)
{
  if (!params || !out)
    return fail(nullptr, EK_EINVAL, "null argument");
  *out = nullptr;
  std::string why;
  if (ek::validate_params(params, why))
    return fail(nullptr, EK_EINVAL, why);
  ek_ctx* c = new ek_ctx();
  c->p = *params;
  c->st = (int)params->slope_type;
  c->elem = params->precision == EK_FP64 ? 8 : 4;
  ek::fill_geom<double>(c->p, c->g64, c->gas64);
  ek::fill_geom<float>(c->p, c->g32, c->gas32);
  c->ncell = c->g64.vstride;
  c->cz = params->dim == 3 ? ek::hydro3d_z_chunk(params->nx, params->ny, params->nz) : 1;
  hipError_t e;
  if ((e = hipGetDevice(&c->device)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
    std::string m = std::string("HIP init: ") + hipGetErrorString(e);
    delete c;
    return fail(nullptr, EK_EHIP, m);
  }
  const size_t state_bytes = (size_t)c->ncell * c->p.nbvar * c->elem;
  c->nblocks_budget = (int)std::min<long long>(1024, (c->ncell + 255) / 256);
  if (hipMalloc(&c->U[0], state_bytes) != hipSuccess || hipMalloc(&c->U[1], state_bytes) != hipSuccess ||
      hipMalloc(&c->ts, 4 * sizeof(ek::TimeSlot)) != hipSuccess ||
      hipMalloc(&c->scratch_bits, (size_t)ek::NSPREAD * ek::SPREAD_STRIDE * 8) != hipSuccess ||
      hipMalloc(&c->spread, ek::SPREAD_ELEMS * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->partial, (size_t)c->nblocks_budget * 16 * sizeof(double)) != hipSuccess) {
    ek_destroy(c);
    return fail(nullptr, EK_ENOMEM, "device allocation failed");
  }
  c->bytes = (long long)(2 * state_bytes + 4 * sizeof(ek::TimeSlot) + ek::NSPREAD * ek::SPREAD_STRIDE * 8 +
                         ek::SPREAD_ELEMS * 8 +
                         (size_t)c->nblocks_budget * 16 * sizeof(double));
  if (c->p.mhd && c->p.dim == 3) {
    const size_t wb = (size_t)ek::fp64::mhd3_scratch_elems(c->p.nx, c->p.ny, c->p.nz) * c->elem;
    if (hipMalloc(&c->work, wb) != hipSuccess) {
      ek_destroy(c);
      return fail(nullptr, EK_ENOMEM, "device allocation of the 3D MHD work area failed");
    }
    c->bytes += (long long)wb;
    c->work_bytes = (long long)wb;
  }
  hipMemsetAsync(c->U[0], 0, state_bytes, c->stream);
  hipMemsetAsync(c->U[1], 0, state_bytes, c->stream);
  hipMemsetAsync(c->ts, 0, 4 * sizeof(ek::TimeSlot), c->stream);
  hipMemsetAsync(c->spread, 0, ek::SPREAD_ELEMS * sizeof(unsigned long long), c->stream);
  if ((e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
    ek_destroy(c);
    return fail(nullptr, EK_EHIP, std::string("HIP init: ") + hipGetErrorString(e));
  }
  *out = c;
  return EK_OK;
}

extern "C" void ek_destroy(ek_ctx* c)
{
  if (!c)
    return;
  if (c->stream)
    hipStreamSynchronize(c->stream);
  for (auto& g : c->graphs)
    if (g.exec)
      hipGraphExecDestroy(g.exec);
  if (c->ev0)
    hipEventDestroy(c->ev0);
  if (c->ev1)
    hipEventDestroy(c->ev1);
  hipFree(c->U[0]);
  hipFree(c->U[1]);
  hipFree(c->ts);
  hipFree(c->scratch_bits);
  hipFree(c->spread);
  hipFree(c->partial);
  hipFree(c->stage);
  hipFree(c->work);
  if (c->stream)
    hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* ek_last_error(const ek_ctx* c)
{
  return c ? c->err.c_str() : g_create_err.c_str();
}

static long long ref_elems(const ek_params& p)
{
  const long long is = p.nx + 2LL * p.gw, js = p.ny + 2LL * p.gw, ks = p.nz + 2LL * p.gw;
  return is * js * (p.dim == 3 ? ks : 1) * p.nbvar;
}

static int ensure_stage(ek_ctx* c, long long n)
{
  if (c->stage_elems >= n)
    return EK_OK;
  hipFree(c->stage);
  c->stage = nullptr;
  c->stage_elems = 0;
  if (hipMalloc(&c->stage, (size_t)n * sizeof(double)) != hipSuccess)
    return fail(c, EK_ENOMEM, "staging allocation failed");
  c->stage_elems = n;
  return EK_OK;
}

extern "C" int ek_upload(ek_ctx* c, const double* host, int layout)
{
  if (!c || !host)
    return fail(c, EK_EINVAL, "null argument");
  const ek_params& p = c->p;
  const int is = p.nx + 2 * p.gw, js = p.ny + 2 * p.gw, ks = p.nz + 2 * p.gw;
  if (layout == EK_LAYOUT_SOA) {
    const long long n = c->ncell * p.nbvar;
    if (p.precision == EK_FP64) {
      EK_HIP(c, hipMemcpyAsync(c->U[c->cur], host, n * 8, hipMemcpyHostToDevice, c->stream));
    } else {
      // FP64 values through a device staging buffer, rounded to FP32 on the device; a 3D
      // MHD context stages in its work area, which holds nothing between steps
      double* stage = (double*)c->work;
      long long se = c->work_bytes / (long long)sizeof(double);
      if (!stage) {
        if (int r = ensure_stage(c, std::min(n, kUploadChunk)))
          return r;
        stage = c->stage;
        se = c->stage_elems;
      }
      EK_HIP(c, ek::upload_f64_as_f32(host, (float*)c->U[c->cur], n, stage, se, c->stream));
    }
  } else if (layout == EK_LAYOUT_REFERENCE) {
    const long long n = ref_elems(p);
    if (int r = ensure_stage(c, n))
      return r;
    EK_HIP(c, hipMemcpyAsync(c->stage, host, n * 8, hipMemcpyHostToDevice, c->stream));
    if (p.precision == EK_FP64)
      EK_HIP(c, ek::fp64::ref_to_soa(p.dim, c->stage, (double*)c->U[c->cur], c->g64, p.gw, p.nbvar, is, js,
                                    ks, c->stream));
    else
      EK_HIP(c, ek::fp32::ref_to_soa(p.dim, c->stage, (float*)c->U[c->cur], c->g32, p.gw, p.nbvar, is, js,
                                    ks, c->stream));
  } else {
    return fail(c, EK_EINVAL, "unknown layout");
  }
  EK_HIP(c, hipStreamSynchronize(c->stream));
  return EK_OK;
}

extern "C" int ek_download(ek_ctx* c, double* host, int layout)
{
  if (!c || !host)
    return fail(c, EK_EINVAL, "null argument");
  const ek_params& p = c->p;
  const int is = p.nx + 2 * p.gw, js = p.ny + 2 * p.gw, ks = p.nz + 2 * p.gw;
  if (layout == EK_LAYOUT_SOA) {
    const long long n = c->ncell * p.nbvar;
    if (p.precision == EK_FP64) {
      EK_HIP(c, hipMemcpyAsync(host, c->U[c->cur], n * 8, hipMemcpyDeviceToHost, c->stream));
      EK_HIP(c, hipStreamSynchronize(c->stream));
    } else {
      std::vector<float> tmp(n);
      EK_HIP(c, hipMemcpyAsync(tmp.data(), c->U[c->cur], n * 4, hipMemcpyDeviceToHost, c->stream));
      EK_HIP(c, hipStreamSynchronize(c->stream));
      for (long long i = 0; i < n; i++)
        host[i] = (double)tmp[i];
    }
  } else if (layout == EK_LAYOUT_REFERENCE) {
    const long long n = ref_elems(p);
    if (int r = ensure_stage(c, n))
      return r;
    if (p.precision == EK_FP64)
      EK_HIP(c, ek::fp64::soa_to_ref(p.dim, (const double*)c->U[c->cur], c->stage, c->g64, p.gw, p.nbvar, is,
                                    js, ks, p.mhd, c->stream));
    else
      EK_HIP(c, ek::fp32::soa_to_ref(p.dim, (const float*)c->U[c->cur], c->stage, c->g32, p.gw, p.nbvar, is,
                                    js, ks, p.mhd, c->stream));
    EK_HIP(c, hipMemcpyAsync(host, c->stage, n * 8, hipMemcpyDeviceToHost, c->stream));
    EK_HIP(c, hipStreamSynchronize(c->stream));
  } else {
    return fail(c, EK_EINVAL, "unknown layout");
  }
  return EK_OK;
}

// CFL reduction of the current state into dst; with spread set, blocks also use the words
// spread + k*SPREAD_STRIDE (k = 1..63) and the result is the max of dst and those words.
static hipError_t launch_inv_dt_into(ek_ctx* c, unsigned long long* dst, unsigned long long* spread)
{
  if (c->p.mhd && c->p.dim == 3) {
    if (c->p.precision == EK_FP64)
      return ek::fp64::mhd3_inv_dt((const double*)c->U[c->cur], c->g64, c->gas64, dst, spread, c->stream);
    return ek::fp32::mhd3_inv_dt((const float*)c->U[c->cur], c->g32, c->gas32, dst, spread, c->stream);
  }
  if (c->p.mhd) {
    if (c->p.precision == EK_FP64)
      return ek::fp64::mhd_inv_dt((const double*)c->U[c->cur], c->g64, c->gas64, dst, spread, c->stream);
    return ek::fp32::mhd_inv_dt((const float*)c->U[c->cur], c->g32, c->gas32, dst, spread, c->stream);
  }
  if (c->p.precision == EK_FP64)
    return ek::fp64::inv_dt(c->p.dim, (const double*)c->U[c->cur], c->g64, c->gas64, dst, spread, c->stream);
  return ek::fp32::inv_dt(c->p.dim, (const float*)c->U[c->cur], c->g32, c->gas32, dst, spread, c->stream);
}

extern "C" int ek_inv_dt(ek_ctx* c, double* inv_dt)
{
  if (!c || !inv_dt)
    return fail(c, EK_EINVAL, "null argument");
  const size_t nb = (size_t)ek::NSPREAD * ek::SPREAD_STRIDE * sizeof(unsigned long long);
  EK_HIP(c, hipMemsetAsync(c->scratch_bits, 0, nb, c->stream));
  EK_HIP(c, launch_inv_dt_into(c, c->scratch_bits, c->scratch_bits));
  std::vector<unsigned long long> w((size_t)ek::NSPREAD * ek::SPREAD_STRIDE);
  EK_HIP(c, hipMemcpyAsync(w.data(), c->scratch_bits, nb, hipMemcpyDeviceToHost, c->stream));
  EK_HIP(c, hipStreamSynchronize(c->stream));
  unsigned long long b = 0;
  for (int k = 0; k < ek::NSPREAD; k++)
    b = std::max(b, w[(size_t)k * ek::SPREAD_STRIDE]);
  std::memcpy(inv_dt, &b, 8);
  return EK_OK;
}

static hipError_t launch_one(ek_ctx* c, const ek::StepCtl& ctl, void* in, void* out)
{
  const ek_params& p = c->p;
  if (p.mhd && p.dim == 3) {
    if (p.precision == EK_FP64)
      return ek::fp64::mhd3_step(p.riemann, c->st, (const double*)in, (double*)out, (double*)c->work, c->g64,
                                 c->gas64, ctl, c->stream);
    return ek::fp32::mhd3_step(p.riemann, c->st, (const float*)in, (float*)out, (float*)c->work, c->g32, c->gas32,
                               ctl, c->stream);
  }
  if (p.mhd) {
    if (p.precision == EK_FP64)
      return ek::fp64::mhd_step(p.riemann, c->st, (const double*)in, (double*)out, c->g64, c->gas64, ctl,
                                c->stream);
    return ek::fp32::mhd_step(p.riemann, c->st, (const float*)in, (float*)out, c->g32, c->gas32, ctl, c->stream);
  }
  if (p.precision == EK_FP64)
    return ek::fp64::step(p.dim, p.riemann, c->st, (const double*)in, (double*)out, c->g64, c->gas64, ctl,
                          c->cz, c->stream);
  return ek::fp32::step(p.dim, p.riemann, c->st, (const float*)in, (float*)out, c->g32, c->gas32, ctl, c->cz,
                        c->stream);
}

extern "C" int ek_step(ek_ctx* c, double dt)
{
  if (!c)
    return fail(c, EK_EINVAL, "null argument");
  EK_HIP(c, hipMemsetAsync(c->ts, 0, 3 * sizeof(ek::TimeSlot), c->stream));
  ek::StepCtl ctl{};
  ctl.ts = c->ts;
  ctl.spread = c->spread;
  ctl.slot = 0;
  ctl.mode = 1;
  ctl.dt_given = dt;
  ctl.cfl = c->p.cfl;
  ctl.tEnd = 0.0;
  ctl.nStepmax = 0;
  EK_HIP(c, launch_one(c, ctl, c->U[c->cur], c->U[1 - c->cur]));
  c->cur = 1 - c->cur;
  return EK_OK;
}

// (Re)captures graphs[s] for a run that starts from the current U[0].
static int build_graph(ek_ctx* c, int s, int nStepmax, double tEnd)
{
  ek_ctx::Graph& G = c->graphs[s];
  if (G.exec) {
    hipGraphExecDestroy(G.exec);
    G.exec = nullptr;
    G.U0 = nullptr;
  }
  hipGraph_t g = nullptr;
  EK_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  for (int i = 0; i < 6; i++) {
    ek::StepCtl ctl{};
    ctl.ts = c->ts;
    ctl.spread = c->spread;
    ctl.slot = i % 3;
    ctl.mode = 0;
    ctl.cfl = c->p.cfl;
    ctl.tEnd = tEnd;
    ctl.nStepmax = nStepmax;
    hipError_t e = launch_one(c, ctl, c->U[i % 2], c->U[(i + 1) % 2]);
    if (e != hipSuccess) {
      hipStreamEndCapture(c->stream, &g);
      if (g)
        hipGraphDestroy(g);
      return hip_fail(c, e, "graph capture launch");
    }
  }
  EK_HIP(c, hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    G.exec = nullptr;
    return hip_fail(c, e, "hipGraphInstantiate");
  }
  G.nStepmax = nStepmax;
  G.tEnd = tEnd;
  G.U0 = c->U[0];
  return EK_OK;
}

extern "C" int ek_run(ek_ctx* c, int nsteps, int nStepmax, double tEnd, double* t_io, int* it_io,
                      double* dt_last)
{
  if (!c || !t_io || !it_io)
    return fail(c, EK_EINVAL, "null argument");
  if (nsteps < 0)
    return fail(c, EK_EINVAL, "nsteps must be >= 0");
  // the current state always starts in U[0] so that the captured graph's buffer order holds
  if (c->cur == 1) {
    std::swap(c->U[0], c->U[1]);
    c->cur = 0;
  }
  // the graph of this start buffer: the one captured from it, else the slot not holding the
  // other buffer's graph
  const int gs = c->graphs[0].U0 == c->U[0] ? 0 : (c->graphs[1].U0 == c->U[0] ? 1 : (c->graphs[0].U0 ? 1 : 0));
  const ek_ctx::Graph& G = c->graphs[gs];
  const bool stale = !G.exec || G.nStepmax != nStepmax || G.tEnd != tEnd || G.U0 != c->U[0];
  if (nsteps == 0) // warm-up call: capture only
    return stale ? build_graph(c, gs, nStepmax, tEnd) : EK_OK;
  ek::TimeSlot init[3];
  std::memset(init, 0, sizeof(init));
  init[0].t = *t_io;
  init[0].it = *it_io;
  EK_HIP(c, hipMemcpyAsync(c->ts, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  EK_HIP(c, hipMemsetAsync(c->spread, 0, ek::SPREAD_ELEMS * sizeof(unsigned long long), c->stream));
  EK_HIP(c, launch_inv_dt_into(c, &c->ts[0].invdt_bits, c->spread)); // slot 0's words
  const int nfull = nsteps / 6, rem = nsteps % 6;
  if (nfull > 0 && stale)
    if (int r = build_graph(c, gs, nStepmax, tEnd))
      return r;
  EK_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (nfull > 0) {
    for (int g = 0; g < nfull; g++)
      EK_HIP(c, hipGraphLaunch(G.exec, c->stream));
  }
  for (int r = 0; r < rem; r++) {
    ek::StepCtl ctl{};
    ctl.ts = c->ts;
    ctl.spread = c->spread;
    ctl.slot = r % 3;
    ctl.mode = 0;
    ctl.cfl = c->p.cfl;
    ctl.tEnd = tEnd;
    ctl.nStepmax = nStepmax;
    EK_HIP(c, launch_one(c, ctl, c->U[r % 2], c->U[(r + 1) % 2]));
  }
  EK_HIP(c, hipEventRecord(c->ev1, c->stream));
  ek::TimeSlot fin;
  EK_HIP(c, hipMemcpyAsync(&fin, &c->ts[nsteps % 3], sizeof(fin), hipMemcpyDeviceToHost, c->stream));
  EK_HIP(c, hipStreamSynchronize(c->stream));
  const int ran = fin.it - *it_io;
  c->cur = ran % 2;
  float ms = 0.f;
  EK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_run_ms = ms;
  c->last_run_launches = nsteps;
  *t_io = fin.t;
  *it_io = fin.it;
  if (dt_last)
    *dt_last = fin.dt;
  return EK_OK;
}

extern "C" int ek_budgets(ek_ctx* c, double* sums)
{
  if (!c || !sums)
    return fail(c, EK_EINVAL, "null argument");
  const int nv = c->p.nbvar;
  const int nb = c->nblocks_budget;
  if (c->p.precision == EK_FP64)
    EK_HIP(c, ek::fp64::budget(nv, (const double*)c->U[c->cur], c->ncell, c->ncell, c->partial, nb, c->stream));
  else
    EK_HIP(c, ek::fp32::budget(nv, (const float*)c->U[c->cur], c->ncell, c->ncell, c->partial, nb, c->stream));
  std::vector<double> h((size_t)nb * nv * 2);
  EK_HIP(c, hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  EK_HIP(c, hipStreamSynchronize(c->stream));
  for (int v = 0; v < nv; v++) {
    double S = 0.0, C = 0.0;
    for (int b = 0; b < nb; b++)
      for (int k = 0; k < 2; k++) {
        const double x = h[((size_t)b * nv + v) * 2 + k];
        const double t = S + x;
        if (std::fabs(S) >= std::fabs(x))
          C += (S - t) + x;
        else
          C += (x - t) + S;
        S = t;
      }
    sums[v] = S + C;
  }
  return EK_OK;
}

extern "C" int ek_sync(ek_ctx* c)
{
  if (!c)
    return fail(c, EK_EINVAL, "null argument");
  EK_HIP(c, hipStreamSynchronize(c->stream));
  return EK_OK;
}

extern "C" int ek_last_run_time(const ek_ctx* c, double* ms, int* launches)
{
  if (!c || !ms || !launches)
    return EK_EINVAL;
  *ms = c->last_run_ms;
  *launches = c->last_run_launches;
  return EK_OK;
}

extern "C" long long ek_device_bytes(const ek_ctx* c)
{
  return c ? c->bytes + c->stage_elems * 8 : 0;
}

extern "C" int ek_host_alloc(size_t bytes, void** ptr)
{
  if (!ptr)
    return EK_EINVAL;
  *ptr = nullptr;
  if (bytes == 0)
    return EK_OK;
  return hipHostMalloc(ptr, bytes, hipHostMallocDefault) == hipSuccess ? EK_OK : EK_ENOMEM;
}

extern "C" int ek_host_free(void* ptr)
{
  if (!ptr)
    return EK_OK;
  return hipHostFree(ptr) == hipSuccess ? EK_OK : EK_EHIP;
}

extern "C" int ek_device_info(int* cus, int* mem_clock_khz, int* bus_width_bits, int* sclk_khz, char* name,
                              int name_len)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return EK_EHIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return EK_EHIP;
  if (cus)
    *cus = prop.multiProcessorCount;
  if (mem_clock_khz)
    *mem_clock_khz = prop.memoryClockRate;
  if (bus_width_bits)
    *bus_width_bits = prop.memoryBusWidth;
  if (sclk_khz)
    *sclk_khz = prop.clockRate;
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  return EK_OK;
}

// ---------------------------------------------------------------------------------------
// ek_run_noalloc: the time loop of ek_run on caller-owned device state, for callers that
// cannot allocate device memory. Everything the kernels need beyond the two state buffers
// (the three time slots, the spread CFL words and, for 3D MHD, the padded work area) lives
// in `scratch`, a fixed array of every device's code object. Its device address is taken
// once per device with hipGetSymbolAddress; host code only ever uses that address.
// Calls that overlap on one device share the array and race.

namespace {
constexpr size_t kScratchBytes = EK_NOALLOC_SCRATCH_BYTES;
constexpr size_t kScratchSlots = 0;                  // 3 TimeSlots (96 B), padded to 256
constexpr size_t kScratchSpread = 256;               // SPREAD_ELEMS words (48 KiB)
constexpr size_t kScratchWork = kScratchSpread + ((size_t)ek::SPREAD_ELEMS * 8 + 255) / 256 * 256;
static_assert(3 * sizeof(ek::TimeSlot) <= kScratchSpread - kScratchSlots, "time slots fit");
static_assert(kScratchWork < kScratchBytes, "room for a work area");
constexpr int kMaxDevices = 64;
} // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device

// Device address of this device's `scratch` (looked up on the first call per device).
static int scratch_address(int dev, unsigned char** out)
{
  if (dev < 0 || dev >= kMaxDevices)
    return fail(nullptr, EK_EINVAL, "device ordinal out of range for ek_run_noalloc");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(nullptr, EK_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return EK_OK;
}

namespace ek {
// Clears the three time slots and their spread words and puts (t0, it0) in slot 0: a kernel,
// not hipMemsetAsync, so that the call can also be recorded into a graph.
__global__ __launch_bounds__(256) void noalloc_init_kernel(TimeSlot* ts, unsigned long long* spread, double t0,
                                                           int it0)
{
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < SPREAD_ELEMS; i += (long long)gridDim.x * 256)
    spread[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x < 3) {
    TimeSlot s{};
    if (threadIdx.x == 0) {
      s.t = t0;
      s.it = it0;
    }
    ts[threadIdx.x] = s;
  }
}
} // namespace ek

extern "C" int ek_run_noalloc(const ek_params* params, void* U, void* U2, int nsteps, int nStepmax, double tEnd,
                              double* t_io, int* it_io, double* dt_last, int* in_U2, void* stream)
{
  if (!params || !U || !U2 || !t_io || !it_io || !in_U2)
    return fail(nullptr, EK_EINVAL, "null argument");
  if (nsteps < 0)
    return fail(nullptr, EK_EINVAL, "nsteps must be >= 0");
  std::string why;
  if (ek::validate_params(params, why))
    return fail(nullptr, EK_EINVAL, why);
  // a context on the host stack only: its device pointers are the caller's buffers and
  // `scratch`, and nothing in it is freed
  ek_ctx c;
  c.p = *params;
  c.st = (int)params->slope_type;
  c.elem = params->precision == EK_FP64 ? 8 : 4;
  ek::fill_geom<double>(c.p, c.g64, c.gas64);
  ek::fill_geom<float>(c.p, c.g32, c.gas32);
  c.ncell = c.g64.vstride;
  c.cz = params->dim == 3 ? ek::hydro3d_z_chunk(params->nx, params->ny, params->nz) : 1;
  c.stream = (hipStream_t)stream;
  c.U[0] = U;
  c.U[1] = U2;
  if (c.p.mhd && c.p.dim == 3) {
    const long long wb = ek::fp64::mhd3_scratch_elems(c.p.nx, c.p.ny, c.p.nz) * (long long)c.elem;
    if ((double)wb > (double)(kScratchBytes - kScratchWork))
      return fail(nullptr, EK_EINVAL,
                  "grid too large for ek_run_noalloc: the 3D MHD work area needs " + std::to_string(wb) +
                      " bytes and the fixed scratch holds " + std::to_string(kScratchBytes - kScratchWork) +
                      " (use ek_create/ek_run)");
  }
  if (hipGetDevice(&c.device) != hipSuccess)
    return fail(nullptr, EK_EHIP, "hipGetDevice failed");
  unsigned char* base = nullptr;
  if (int r = scratch_address(c.device, &base))
    return r;
  c.ts = (ek::TimeSlot*)(base + kScratchSlots);
  c.spread = (unsigned long long*)(base + kScratchSpread);
  c.work = (c.p.mhd && c.p.dim == 3) ? (void*)(base + kScratchWork) : nullptr;
  auto hip_err = [](hipError_t e, const char* what) {
    return fail(nullptr, EK_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  *in_U2 = 0;
  if (nsteps == 0)
    return EK_OK;
  hipLaunchKernelGGL(ek::noalloc_init_kernel, dim3(24), dim3(256), 0, c.stream, c.ts, c.spread, *t_io, *it_io);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return hip_err(e, "time-slot initialisation");
  if ((e = launch_inv_dt_into(&c, &c.ts[0].invdt_bits, c.spread)) != hipSuccess)
    return hip_err(e, "initial CFL reduction");
  for (int r = 0; r < nsteps; r++) {
    ek::StepCtl ctl{};
    ctl.ts = c.ts;
    ctl.spread = c.spread;
    ctl.slot = r % 3;
    ctl.mode = 0;
    ctl.cfl = c.p.cfl;
    ctl.tEnd = tEnd;
    ctl.nStepmax = nStepmax;
    if ((e = launch_one(&c, ctl, c.U[r % 2], c.U[(r + 1) % 2])) != hipSuccess)
      return hip_err(e, "step launch");
  }
  ek::TimeSlot fin;
  if ((e = hipMemcpyAsync(&fin, &c.ts[nsteps % 3], sizeof(fin), hipMemcpyDeviceToHost, c.stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c.stream)) != hipSuccess)
    return hip_err(e, "reading the final time slot");
  *in_U2 = (fin.it - *it_io) % 2;
  *t_io = fin.t;
  *it_io = fin.it;
  if (dt_last)
    *dt_last = fin.dt;
  return EK_OK;
}
