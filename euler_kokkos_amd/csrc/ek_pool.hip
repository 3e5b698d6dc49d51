// ek_pool.hip -- the job pool of include/ek_pool.h: arena allocation, admission and
// retirement, the work list, graph-captured pool steps and the tile autotuner.
#include "ek_pool.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ek_geom_host.h"
#include "ek_internal.h"
#include "ek_pool_launch.h"
#include "ek_pool_mhd3d.h"

namespace {

struct HostJob {
  bool used = false;
  // admitted by ek_pool_admit_async and still uploading on the upload stream: not in the work
  // lists until activate_pending sees its event; upar = the arena buffer it is uploaded into
  bool pending = false;
  int upar = 0;
  long long seq = 0; // admission order
  long long join_after = 0; // no earlier than this ek_pool_step call (ek_pool_test_defer_join)
  ek_params p{};
  long long off = 0, nelem = 0, alloc = 0;
  long long woff = 0, walloc = 0; // 3D MHD: the job's padded work area in the work buffer
  long long adm_step = 0;
  int nStepmax = 0;
  double tEnd = 0.0;
  double t = 0.0;
  int it = 0;
  int finished = 0;
  long long ncell = 0;
};

constexpr long long kAlign = 64; // elements
constexpr int kGraphSteps = 6;   // lcm of the 2 arena buffers and the 3 time slots
constexpr int kCz[4] = { 8, 16, 32, 64 };
// 3D MHD pools: share of the arena for the two state buffers; the rest is the work buffer
// (78 reals per ghost-padded cell against 2 x 8 per cell: 13-16 % for edges 64-192)
constexpr double kMhdStateShare = 0.16;

long long round_up(long long n)
{
  return (n + kAlign - 1) / kAlign * kAlign;
}

} // namespace

struct ek_pool {
  ek_params kind{};
  std::string err;
  hipStream_t stream = nullptr;
  bool fp64 = true;
  size_t elem = 8;
  long long cap = 0; // elements per arena buffer
  void* A[2] = { nullptr, nullptr };
  ek::TimeSlot* ts = nullptr;
  int max_jobs = 0;
  void* djobs = nullptr;
  ek::PoolWork* dwork = nullptr;
  int work_cap = 0;
  int nwork = 0;
  std::vector<HostJob> jobs;
  std::map<long long, long long> freel; // offset -> length (elements)
  long long nstep = 0;
  int variant = 0;
  int cus = 1;
  int grid = 1;
  hipGraphExec_t graph = nullptr;
  bool graph_ok = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int last_launches = 0;
  double* stage = nullptr;
  long long stage_elems = 0;
  double* partial = nullptr;
  // 3D MHD kind: the jobs' work areas (first-fit like the arena) and the update's work list
  bool mhd3 = false;
  void* W = nullptr;
  long long wcap = 0;
  std::map<long long, long long> wfree;
  ek::PoolWork* dworkU = nullptr;
  int nworkU = 0;
  int workU_cap = 0;
  int mgrid[5] = { 1, 1, 1, 1, 1 };
  int test_fail_admit = 0; // ek_pool_test_fail_admit: 0 off, else the admission step that fails
  // asynchronous admission: uploads (and the FP32 rounding) run on their own stream, with
  // their own staging buffer, one completion event per job slot
  hipStream_t ustream = nullptr;
  double* ustage = nullptr;
  long long ustage_elems = 0;
  std::vector<hipEvent_t> jev;
  long long admissions = 0;
  long long step_calls = 0;
  int test_defer_join = 0; // ek_pool_test_defer_join
  ek::TimeSlot zslots[3] = {}; // t = 0, it = 0: the source of every job's time-slot reset
};

static int pfail(ek_pool* p, int code, const std::string& m)
{
  if (p)
    p->err = m;
  return code;
}

#define PHIP(p, call)                                                  \
  do {                                                                 \
    hipError_t e_ = (call);                                            \
    if (e_ != hipSuccess)                                              \
      return pfail((p), EK_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// message of the last failed ek_pool_create on this thread
static thread_local std::string g_pool_create_err;

static int kind_dim(const ek_pool* p)
{
  return p->kind.dim;
}

// z-chunk length of a 3D job: the tile choice's depth, with the job's planes then spread
// evenly over that many chunks (a chunk's block lasts as long as its chunk is deep)
static int cz_of(const ek_pool* p, const ek_params& j)
{
  const int c = kCz[p->variant & 3];
  if (c >= j.nz)
    return j.nz;
  const int nch = (j.nz + c - 1) / c;
  return (j.nz + nch - 1) / nch;
}

// PoolJob record of job `id` for the current tile choice
template <class T>
static ek::PoolJob<T> make_record(const ek_pool* p, int id)
{
  const HostJob& h = p->jobs[id];
  ek::PoolJob<T> r{};
  r.off = h.off;
  r.slot0 = 3 * id;
  r.nStepmax = h.nStepmax;
  r.cz = h.p.dim == 3 ? cz_of(p, h.p) : 1;
  r.cfl = h.p.cfl;
  r.tEnd = h.tEnd;
  ek::fill_geom<T>(h.p, r.geo, r.gas);
  return r;
}

// 3D MHD record of job `id`
template <class T>
static ek::PoolJobM<T> make_record_m(const ek_pool* p, int id)
{
  const HostJob& h = p->jobs[id];
  ek::PoolJobM<T> r{};
  r.off = h.off;
  r.woff = h.woff;
  r.slot0 = 3 * id;
  r.nStepmax = h.nStepmax;
  r.cfl = h.p.cfl;
  r.tEnd = h.tEnd;
  r.w.nx = h.p.nx;
  r.w.ny = h.p.ny;
  r.w.nz = h.p.nz;
  r.w.PX = h.p.nx + 6;
  r.w.PY = h.p.ny + 6;
  r.w.PZ = h.p.nz + 6;
  r.w.PN = (long long)r.w.PX * r.w.PY * r.w.PZ;
  ek::fill_geom<T>(h.p, r.geo, r.gas);
  return r;
}

static size_t record_bytes(const ek_pool* p)
{
  if (p->mhd3)
    return p->fp64 ? sizeof(ek::PoolJobM<double>) : sizeof(ek::PoolJobM<float>);
  return p->fp64 ? sizeof(ek::PoolJob<double>) : sizeof(ek::PoolJob<float>);
}

static int upload_record(ek_pool* p, int id)
{
  if (p->mhd3) {
    if (p->fp64) {
      const ek::PoolJobM<double> r = make_record_m<double>(p, id);
      PHIP(p, hipMemcpyAsync((ek::PoolJobM<double>*)p->djobs + id, &r, sizeof(r), hipMemcpyHostToDevice, p->stream));
    } else {
      const ek::PoolJobM<float> r = make_record_m<float>(p, id);
      PHIP(p, hipMemcpyAsync((ek::PoolJobM<float>*)p->djobs + id, &r, sizeof(r), hipMemcpyHostToDevice, p->stream));
    }
    PHIP(p, hipStreamSynchronize(p->stream));
    return EK_OK;
  }
  if (p->fp64) {
    const ek::PoolJob<double> r = make_record<double>(p, id);
    PHIP(p, hipMemcpyAsync((ek::PoolJob<double>*)p->djobs + id, &r, sizeof(r), hipMemcpyHostToDevice, p->stream));
  } else {
    const ek::PoolJob<float> r = make_record<float>(p, id);
    PHIP(p, hipMemcpyAsync((ek::PoolJob<float>*)p->djobs + id, &r, sizeof(r), hipMemcpyHostToDevice, p->stream));
  }
  PHIP(p, hipStreamSynchronize(p->stream));
  return EK_OK;
}

static void tiles_of(const ek_pool* p, const ek_params& j, int& ntx, int& nty, int& ntz)
{
  if (j.dim == 2) {
    ntx = (j.nx + ek::pool::V2[p->variant][0] - 1) / ek::pool::V2[p->variant][0];
    nty = (j.ny + ek::pool::V2[p->variant][1] - 1) / ek::pool::V2[p->variant][1];
    ntz = 1;
  } else {
    ntx = (j.nx + ek::pool::P3X - 1) / ek::pool::P3X;
    nty = (j.ny + ek::pool::P3Y - 1) / ek::pool::P3Y;
    const int cz = cz_of(p, j);
    ntz = (j.nz + cz - 1) / cz;
  }
}

// Device copy of a host work list into *buf (capacity *cap items, grown as needed); on
// failure the previous list stays valid.
static int put_list(ek_pool* p, const std::vector<ek::PoolWork>& w, ek::PoolWork** buf, int* cap)
{
  if ((int)w.size() > *cap) {
    const int nc = (int)std::max<size_t>(w.size() * 2, 1024);
    ek::PoolWork* nw = nullptr;
    PHIP(p, hipMalloc(&nw, (size_t)nc * sizeof(ek::PoolWork)));
    PHIP(p, hipStreamSynchronize(p->stream)); // no launch still reads the old list
    hipFree(*buf);
    *buf = nw;
    *cap = nc;
  }
  if (!w.empty())
    PHIP(p, hipMemcpyAsync(*buf, w.data(), w.size() * sizeof(ek::PoolWork), hipMemcpyHostToDevice, p->stream));
  PHIP(p, hipStreamSynchronize(p->stream));
  return EK_OK;
}

// 3D MHD work lists: (job, 256-cell chunk of its padded grid) for the four padded-grid
// kernels, (job, 4096-cell chunk of its interior) for the update, the job's first chunk
// flagged for its time bookkeeping
static int rebuild_work_mhd3(ek_pool* p)
{
  std::vector<ek::PoolWork> wp, wu;
  for (int id = 0; id < (int)p->jobs.size(); id++) {
    const HostJob& h = p->jobs[id];
    if (!h.used || h.finished || h.pending)
      continue;
    const long long pn = (long long)(h.p.nx + 6) * (h.p.ny + 6) * (long long)(h.p.nz + 6);
    const long long np = (pn + ek::MP_CHUNK - 1) / ek::MP_CHUNK, nu = (h.ncell + ek::MU_CHUNK - 1) / ek::MU_CHUNK;
    for (long long c = 0; c < np; c++)
      wp.push_back(ek::PoolWork{ id, (int)c, 0, 0 });
    for (long long c = 0; c < nu; c++)
      wu.push_back(ek::PoolWork{ id, (int)c, 0, c == 0 ? 1 : 0 });
  }
  if (int rc = put_list(p, wp, &p->dwork, &p->work_cap))
    return rc;
  if (int rc = put_list(p, wu, &p->dworkU, &p->workU_cap))
    return rc;
  p->nwork = (int)wp.size();
  p->nworkU = (int)wu.size();
  p->graph_ok = false;
  return EK_OK;
}

static int rebuild_work(ek_pool* p)
{
  if (p->mhd3)
    return rebuild_work_mhd3(p);
  std::vector<ek::PoolWork> w;
  for (int id = 0; id < (int)p->jobs.size(); id++) {
    const HostJob& h = p->jobs[id];
    if (!h.used || h.finished || h.pending)
      continue;
    int ntx, nty, ntz;
    tiles_of(p, h.p, ntx, nty, ntz);
    for (int tz = 0; tz < ntz; tz++)
      for (int ty = 0; ty < nty; ty++)
        for (int tx = 0; tx < ntx; tx++)
          w.push_back(ek::PoolWork{ id, tx, ty, tz * 2 + ((tx | ty | tz) == 0 ? 1 : 0) });
  }
  if ((int)w.size() > p->work_cap) {
    // the new list buffer first: on failure the previous list (and nwork) stay valid
    const int cap = (int)std::max<size_t>(w.size() * 2, 1024);
    ek::PoolWork* nw = nullptr;
    PHIP(p, hipMalloc(&nw, (size_t)cap * sizeof(ek::PoolWork)));
    PHIP(p, hipStreamSynchronize(p->stream)); // no launch still reads the old list
    hipFree(p->dwork);
    p->dwork = nw;
    p->work_cap = cap;
  }
  if (!w.empty())
    PHIP(p, hipMemcpyAsync(p->dwork, w.data(), w.size() * sizeof(ek::PoolWork), hipMemcpyHostToDevice, p->stream));
  PHIP(p, hipStreamSynchronize(p->stream));
  p->nwork = (int)w.size();
  p->graph_ok = false;
  return EK_OK;
}

static void set_grid(ek_pool* p)
{
  if (p->mhd3) {
    const int rs = p->kind.riemann, st = (int)p->kind.slope_type;
    if (p->fp64)
      ek::fp64::pool_mhd3_grids(rs, st, p->cus, p->mgrid);
    else
      ek::fp32::pool_mhd3_grids(rs, st, p->cus, p->mgrid);
    p->grid = p->mgrid[0];
    return;
  }
  const int occ = p->fp64 ? ek::fp64::pool_occupancy(kind_dim(p), p->variant, (int)p->kind.slope_type)
                          : ek::fp32::pool_occupancy(kind_dim(p), p->variant, (int)p->kind.slope_type);
  p->grid = p->cus * occ;
}

static hipError_t launch_one(ek_pool* p, long long n)
{
  const int par = (int)(n % 2), slot = (int)(n % 3);
  const int st = (int)p->kind.slope_type;
  if (p->mhd3) {
    const int rs = p->kind.riemann;
    if (p->fp64)
      return ek::fp64::pool_mhd3_step(rs, st, (const double*)p->A[par], (double*)p->A[1 - par], (double*)p->W,
                                      (const ek::PoolJobM<double>*)p->djobs, p->dwork, p->nwork, p->dworkU,
                                      p->nworkU, p->ts, slot, p->mgrid, p->stream);
    return ek::fp32::pool_mhd3_step(rs, st, (const float*)p->A[par], (float*)p->A[1 - par], (float*)p->W,
                                    (const ek::PoolJobM<float>*)p->djobs, p->dwork, p->nwork, p->dworkU, p->nworkU,
                                    p->ts, slot, p->mgrid, p->stream);
  }
  if (p->fp64)
    return ek::fp64::pool_step(kind_dim(p), p->variant, st, (const double*)p->A[par], (double*)p->A[1 - par],
                               (const ek::PoolJob<double>*)p->djobs, p->dwork, p->nwork, p->ts, slot, p->grid,
                               p->stream);
  return ek::fp32::pool_step(kind_dim(p), p->variant, st, (const float*)p->A[par], (float*)p->A[1 - par],
                             (const ek::PoolJob<float>*)p->djobs, p->dwork, p->nwork, p->ts, slot, p->grid,
                             p->stream);
}

static int capture_graph(ek_pool* p)
{
  if (p->graph) {
    hipGraphExecDestroy(p->graph);
    p->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  PHIP(p, hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed));
  for (int i = 0; i < kGraphSteps; i++) {
    hipError_t e = launch_one(p, i);
    if (e != hipSuccess) {
      hipStreamEndCapture(p->stream, &g);
      if (g)
        hipGraphDestroy(g);
      return pfail(p, EK_EHIP, std::string("pool graph capture: ") + hipGetErrorString(e));
    }
  }
  PHIP(p, hipStreamEndCapture(p->stream, &g));
  hipError_t e = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess)
    return pfail(p, EK_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  p->graph_ok = true;
  return EK_OK;
}

static int refresh_status(ek_pool* p, bool& changed)
{
  changed = false;
  std::vector<ek::TimeSlot> h((size_t)p->max_jobs * 3);
  PHIP(p, hipMemcpyAsync(h.data(), p->ts, h.size() * sizeof(ek::TimeSlot), hipMemcpyDeviceToHost, p->stream));
  PHIP(p, hipStreamSynchronize(p->stream));
  const int slot = (int)(p->nstep % 3);
  for (int id = 0; id < (int)p->jobs.size(); id++) {
    HostJob& j = p->jobs[id];
    // a finished job has left the work list: its slots are no longer forwarded, so its
    // t / it / finished stay as the refresh that saw it finish left them
    if (!j.used || j.finished || j.pending)
      continue;
    const ek::TimeSlot& s = h[(size_t)id * 3 + slot];
    j.t = s.t;
    j.it = s.it;
    const int fin = (s.finished || s.t >= (j.tEnd - 1e-14) || s.it >= j.nStepmax) ? 1 : 0;
    if (fin && !j.finished)
      changed = true;
    j.finished = fin;
  }
  return EK_OK;
}

// arena ------------------------------------------------------------------------------

// first fit in a free list (offset -> length, elements)
static long long list_alloc(std::map<long long, long long>& fl, long long n)
{
  for (auto it = fl.begin(); it != fl.end(); ++it) {
    if (it->second >= n) {
      const long long off = it->first, len = it->second;
      fl.erase(it);
      if (len > n)
        fl[off + n] = len - n;
      return off;
    }
  }
  return -1;
}

static void list_free(std::map<long long, long long>& fl, long long off, long long n)
{
  auto it = fl.emplace(off, n).first;
  auto nx = std::next(it);
  if (nx != fl.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    fl.erase(nx);
  }
  if (it != fl.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      fl.erase(it);
    }
  }
}

static long long arena_alloc(ek_pool* p, long long n)
{
  return list_alloc(p->freel, n);
}

static void arena_free(ek_pool* p, long long off, long long n)
{
  list_free(p->freel, off, n);
}

// elements of a 3D MHD job's padded work area
static long long work_elems(const ek_params& j)
{
  return ek::fp64::mhd3_scratch_elems(j.nx, j.ny, j.nz);
}

static int ensure_stage(ek_pool* p, long long n)
{
  if (p->stage_elems >= n)
    return EK_OK;
  hipFree(p->stage);
  p->stage = nullptr;
  p->stage_elems = 0;
  PHIP(p, hipMalloc(&p->stage, (size_t)n * sizeof(double)));
  p->stage_elems = n;
  return EK_OK;
}

static int job_parity(const ek_pool* p, const HostJob& j)
{
  return (int)((j.adm_step + j.it) % 2);
}

// C ABI ------------------------------------------------------------------------------

extern "C" int ek_pool_create(const ek_params* kind, long long arena_bytes, int max_jobs, ek_pool** out)
{
  if (!kind || !out || arena_bytes <= 0 || max_jobs <= 0) {
    g_pool_create_err = "invalid argument";
    return EK_EINVAL;
  }
  *out = nullptr;
  std::string why;
  if (ek::validate_params(kind, why)) {
    g_pool_create_err = why;
    return EK_EINVAL;
  }
  const bool mhd3 = kind->mhd && kind->dim == 3;
  if ((kind->mhd && !mhd3) || (!kind->mhd && kind->riemann != ek::RS_HLLC) ||
      !(kind->slope_type == 1.0 || kind->slope_type == 2.0)) {
    g_pool_create_err = "the job pool runs hydro HLLC jobs (2D or 3D) and 3D MHD jobs, with slope_type 1 or 2";
    return EK_EINVAL;
  }
  ek_pool* p = new ek_pool();
  p->kind = *kind;
  p->mhd3 = mhd3;
  p->fp64 = kind->precision == EK_FP64;
  p->elem = p->fp64 ? 8 : 4;
  const double state_share = mhd3 ? kMhdStateShare : 1.0;
  p->cap = (long long)(arena_bytes * state_share) / 2 / (long long)p->elem / kAlign * kAlign;
  if (mhd3) {
    p->wcap = (long long)(arena_bytes * (1.0 - state_share)) / (long long)p->elem / kAlign * kAlign;
    p->wfree[0] = p->wcap;
  }
  p->max_jobs = max_jobs;
  p->jobs.resize(max_jobs);
  p->jev.assign(max_jobs, nullptr);
  p->freel[0] = p->cap;
  p->variant = (kind->dim == 2 || mhd3) ? 0 : 2; // 16x16 tiles (2D), 32-plane chunks (3D) until tuned
  int dev = 0;
  hipError_t e;
  const size_t recb = record_bytes(p);
  if ((e = hipGetDevice(&dev)) != hipSuccess ||
      (e = hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&p->ustream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&p->A[0], (size_t)p->cap * p->elem)) != hipSuccess ||
      (e = hipMalloc(&p->A[1], (size_t)p->cap * p->elem)) != hipSuccess ||
      (mhd3 && (e = hipMalloc(&p->W, (size_t)p->wcap * p->elem)) != hipSuccess) ||
      (e = hipMalloc(&p->ts, (size_t)max_jobs * 3 * sizeof(ek::TimeSlot))) != hipSuccess ||
      (e = hipMalloc(&p->djobs, (size_t)max_jobs * recb)) != hipSuccess ||
      (e = hipMalloc(&p->partial, (size_t)1024 * 16 * sizeof(double))) != hipSuccess ||
      (e = hipEventCreate(&p->ev0)) != hipSuccess || (e = hipEventCreate(&p->ev1)) != hipSuccess ||
      (e = hipMemsetAsync(p->ts, 0, (size_t)max_jobs * 3 * sizeof(ek::TimeSlot), p->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(p->stream)) != hipSuccess) {
    g_pool_create_err = std::string("pool allocation: ") + hipGetErrorString(e);
    ek_pool_destroy(p);
    return e == hipErrorOutOfMemory ? EK_ENOMEM : EK_EHIP;
  }
  set_grid(p);
  *out = p;
  return EK_OK;
}

extern "C" void ek_pool_destroy(ek_pool* p)
{
  if (!p)
    return;
  if (p->ustream)
    hipStreamSynchronize(p->ustream);
  if (p->stream)
    hipStreamSynchronize(p->stream);
  for (hipEvent_t ev : p->jev)
    if (ev)
      hipEventDestroy(ev);
  if (p->graph)
    hipGraphExecDestroy(p->graph);
  if (p->ev0)
    hipEventDestroy(p->ev0);
  if (p->ev1)
    hipEventDestroy(p->ev1);
  hipFree(p->A[0]);
  hipFree(p->A[1]);
  hipFree(p->ts);
  hipFree(p->djobs);
  hipFree(p->dwork);
  hipFree(p->dworkU);
  hipFree(p->W);
  hipFree(p->stage);
  hipFree(p->ustage);
  hipFree(p->partial);
  if (p->stream)
    hipStreamDestroy(p->stream);
  if (p->ustream)
    hipStreamDestroy(p->ustream);
  delete p;
}

extern "C" const char* ek_pool_last_error(const ek_pool* p)
{
  return p ? p->err.c_str() : g_pool_create_err.c_str();
}

static long long job_cells(const ek_params& j)
{
  return (long long)j.nx * j.ny * (j.dim == 3 ? j.nz : 1);
}

extern "C" long long ek_pool_job_bytes(const ek_pool* p, const ek_params* j)
{
  if (!p || !j)
    return -1;
  long long n = 2 * round_up(job_cells(*j) * j->nbvar);
  if (p->mhd3)
    n += round_up(work_elems(*j));
  return n * (long long)p->elem;
}

extern "C" long long ek_pool_free_bytes(const ek_pool* p)
{
  if (!p)
    return -1;
  long long f = 0;
  for (const auto& kv : p->freel)
    f += kv.second;
  f *= 2;
  for (const auto& kv : p->wfree)
    f += kv.second;
  return f * (long long)p->elem;
}

extern "C" int ek_pool_active(const ek_pool* p)
{
  if (!p)
    return 0;
  int n = 0;
  for (const HostJob& j : p->jobs)
    n += (j.used && !j.finished) ? 1 : 0;
  return n;
}

// Uploads the admitted job's initial state into arena buffer `par` at its offset, resets its
// time slots and reduces its first invDt. Touches nothing outside the job's own arena
// range, slots and the staging buffer.
static int init_slots(ek_pool* p, const ek_params* jp, int id, const char* dst);

static int admit_upload(ek_pool* p, const ek_params* jp, const double* host, int layout, int id, long long off,
                        long long nelem)
{
  const int par = (int)(p->nstep % 2);
  char* dst = (char*)p->A[par] + (size_t)off * p->elem;
  if (p->test_fail_admit == 1)
    return pfail(p, EK_EHIP, "injected admission failure (upload)");
  if (layout == EK_LAYOUT_SOA) {
    if (p->fp64) {
      PHIP(p, hipMemcpyAsync(dst, host, (size_t)nelem * 8, hipMemcpyHostToDevice, p->stream));
    } else {
      // FP64 values through the device staging buffer, rounded to FP32 on the device
      if (int rc = ensure_stage(p, std::min<long long>(nelem, 1LL << 22)))
        return rc;
      PHIP(p, ek::upload_f64_as_f32(host, (float*)dst, nelem, p->stage, p->stage_elems, p->stream));
    }
  } else if (layout == EK_LAYOUT_REFERENCE) {
    const int is = jp->nx + 2 * jp->gw, js = jp->ny + 2 * jp->gw, ks = jp->nz + 2 * jp->gw;
    const long long nr = (long long)is * js * (jp->dim == 3 ? ks : 1) * jp->nbvar;
    if (int rc = ensure_stage(p, nr))
      return rc;
    PHIP(p, hipMemcpyAsync(p->stage, host, (size_t)nr * 8, hipMemcpyHostToDevice, p->stream));
    if (p->fp64) {
      ek::Geom<double> g;
      ek::Gas<double> gs;
      ek::fill_geom<double>(*jp, g, gs);
      PHIP(p, ek::fp64::ref_to_soa(jp->dim, p->stage, (double*)dst, g, jp->gw, jp->nbvar, is, js, ks, p->stream));
    } else {
      ek::Geom<float> g;
      ek::Gas<float> gs;
      ek::fill_geom<float>(*jp, g, gs);
      PHIP(p, ek::fp32::ref_to_soa(jp->dim, p->stage, (float*)dst, g, jp->gw, jp->nbvar, is, js, ks, p->stream));
    }
  } else {
    return pfail(p, EK_EINVAL, "unknown layout");
  }
  PHIP(p, hipStreamSynchronize(p->stream));
  if (int rc = init_slots(p, jp, id, dst))
    return rc;
  PHIP(p, hipStreamSynchronize(p->stream));
  if (p->test_fail_admit == 2)
    return pfail(p, EK_EHIP, "injected admission failure (after the upload)");
  return EK_OK;
}

// Time slots of job id on the pool stream: t = 0, it = 0, then the invDt of its state at dst
// in the slot the next pool step reads. Does not synchronize.
static int init_slots(ek_pool* p, const ek_params* jp, int id, const char* dst)
{
  PHIP(p, hipMemcpyAsync(p->ts + 3 * id, p->zslots, sizeof(p->zslots), hipMemcpyHostToDevice, p->stream));
  unsigned long long* bits = &p->ts[3 * id + (int)(p->nstep % 3)].invdt_bits;
  if (p->fp64) {
    ek::Geom<double> g;
    ek::Gas<double> gs;
    ek::fill_geom<double>(*jp, g, gs);
    if (p->mhd3)
      PHIP(p, ek::fp64::mhd3_inv_dt((const double*)dst, g, gs, bits, nullptr, p->stream));
    else
      PHIP(p, ek::fp64::inv_dt(jp->dim, (const double*)dst, g, gs, bits, nullptr, p->stream));
  } else {
    ek::Geom<float> g;
    ek::Gas<float> gs;
    ek::fill_geom<float>(*jp, g, gs);
    if (p->mhd3)
      PHIP(p, ek::fp32::mhd3_inv_dt((const float*)dst, g, gs, bits, nullptr, p->stream));
    else
      PHIP(p, ek::fp32::inv_dt(jp->dim, (const float*)dst, g, gs, bits, nullptr, p->stream));
  }
  return EK_OK;
}

// Checks a job against the pool's kind and reserves a free job slot with its arena range (and,
// for 3D MHD, its work-area range). The slot is filled but stays unused (out of the work list,
// status and downloads) until the admission has succeeded; release_job gives everything back.
// Returns EK_OK and the slot in *id_out.
static int reserve_job(ek_pool* p, const ek_params* jp, int layout, int nStepmax, double tEnd, int* id_out)
{
  std::string why;
  if (ek::validate_params(jp, why))
    return pfail(p, EK_EINVAL, why);
  if (jp->dim != p->kind.dim || jp->mhd != p->kind.mhd || jp->riemann != p->kind.riemann ||
      jp->slope_type != p->kind.slope_type || jp->precision != p->kind.precision)
    return pfail(p, EK_EINVAL, "job does not belong to the pool's kind (dim, mhd, riemann, slope_type, precision)");
  if (layout != EK_LAYOUT_SOA && layout != EK_LAYOUT_REFERENCE)
    return pfail(p, EK_EINVAL, "unknown layout");
  int id = -1;
  for (int i = 0; i < p->max_jobs; i++)
    if (!p->jobs[i].used) {
      id = i;
      break;
    }
  if (id < 0)
    return pfail(p, EK_EFULL, "no free job slot");
  const long long ncell = job_cells(*jp), nelem = ncell * jp->nbvar, need = round_up(nelem);
  const long long wneed = p->mhd3 ? round_up(work_elems(*jp)) : 0;
  if (p->mhd3 && (double)(jp->nx + 6) * (jp->ny + 6) * (jp->nz + 6) * (double)p->elem >= 4294967296.0)
    return pfail(p, EK_EINVAL, "3D MHD job too large: each padded work array must stay below 4 GiB");
  const long long off = arena_alloc(p, need);
  if (off < 0)
    return pfail(p, EK_EFULL, "arena full");
  const long long woff = p->mhd3 ? list_alloc(p->wfree, wneed) : 0;
  if (woff < 0) {
    arena_free(p, off, need);
    return pfail(p, EK_EFULL, "work buffer full");
  }
  HostJob& h = p->jobs[id];
  h = HostJob{};
  h.p = *jp;
  h.off = off;
  h.nelem = nelem;
  h.alloc = need;
  h.woff = woff;
  h.walloc = wneed;
  h.ncell = ncell;
  h.adm_step = p->nstep;
  h.nStepmax = nStepmax;
  h.tEnd = tEnd;
  h.seq = p->admissions++;
  *id_out = id;
  return EK_OK;
}

// Gives a reserved or admitted job's slot and memory back.
static void release_job(ek_pool* p, int id)
{
  HostJob& h = p->jobs[id];
  arena_free(p, h.off, h.alloc);
  if (h.walloc > 0)
    list_free(p->wfree, h.woff, h.walloc);
  h = HostJob{};
}

extern "C" int ek_pool_admit(ek_pool* p, const ek_params* jp, const double* host, int layout, int nStepmax,
                             double tEnd, int* job_id)
{
  if (!p || !jp || !host || !job_id)
    return pfail(p, EK_EINVAL, "null argument");
  int id = -1;
  if (int rc = reserve_job(p, jp, layout, nStepmax, tEnd, &id))
    return rc;
  // any failure gives the arena range and the slot back, so the pool is as it was before
  auto rollback = [&](int rc) {
    release_job(p, id);
    return rc;
  };
  HostJob& h = p->jobs[id];
  if (int rc = admit_upload(p, jp, host, layout, id, h.off, h.nelem))
    return rollback(rc);
  if (int rc = upload_record(p, id))
    return rollback(rc);
  h.used = true;
  if (p->test_fail_admit == 3) {
    pfail(p, EK_EHIP, "injected admission failure (work list)");
    return rollback(EK_EHIP);
  }
  if (int rc = rebuild_work(p))
    return rollback(rc); // rebuild_work leaves the previous work list in place on failure
  *job_id = id;
  return EK_OK;
}

extern "C" int ek_pool_admit_async(ek_pool* p, const ek_params* jp, const double* host, int nStepmax, double tEnd,
                                   int* job_id)
{
  if (!p || !jp || !host || !job_id)
    return pfail(p, EK_EINVAL, "null argument");
  int id = -1;
  if (int rc = reserve_job(p, jp, EK_LAYOUT_SOA, nStepmax, tEnd, &id))
    return rc;
  HostJob& h = p->jobs[id];
  // once an upload is queued, a failure waits for the upload stream before the range is
  // given back
  auto rollback = [&](int rc) {
    hipStreamSynchronize(p->ustream);
    release_job(p, id);
    return rc;
  };
  // into the buffer the next pool step reads; activate_pending moves it if the job joins
  // after an odd number of steps
  const int par = (int)(p->nstep % 2);
  char* dst = (char*)p->A[par] + (size_t)h.off * p->elem;
  hipError_t e;
  if (p->fp64) {
    e = hipMemcpyAsync(dst, host, (size_t)h.nelem * 8, hipMemcpyHostToDevice, p->ustream);
  } else {
    if (!p->ustage) {
      const long long n = 1LL << 22;
      if (hipMalloc(&p->ustage, (size_t)n * sizeof(double)) != hipSuccess) {
        p->ustage = nullptr;
        return rollback(pfail(p, EK_ENOMEM, "upload staging allocation failed"));
      }
      p->ustage_elems = n;
    }
    e = ek::upload_f64_as_f32(host, (float*)dst, h.nelem, p->ustage, p->ustage_elems, p->ustream);
  }
  if (e == hipSuccess && !p->jev[id])
    e = hipEventCreateWithFlags(&p->jev[id], hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipEventRecord(p->jev[id], p->ustream);
  if (e != hipSuccess)
    return rollback(pfail(p, EK_EHIP, std::string("asynchronous admission: ") + hipGetErrorString(e)));
  if (int rc = upload_record(p, id))
    return rollback(rc);
  h.used = true;
  h.pending = true;
  h.upar = par;
  h.join_after = p->step_calls + p->test_defer_join;
  *job_id = id;
  return EK_OK;
}

// Lets uploaded jobs join the pool: each pending job whose upload event has completed is
// moved into the buffer the next step reads (when it was uploaded into the other one), gets
// its time slots and first invDt, and enters the work lists. mode 0: only jobs already
// uploaded; 1: waits for the earliest admitted pending job first; 2: waits for all of them.
static int activate_pending(ek_pool* p, int mode)
{
  int first = -1;
  for (int id = 0; id < p->max_jobs; id++) {
    const HostJob& h = p->jobs[id];
    if (h.used && h.pending && (first < 0 || h.seq < p->jobs[first].seq))
      first = id;
  }
  if (first < 0)
    return EK_OK;
  if (mode == 1)
    PHIP(p, hipEventSynchronize(p->jev[first]));
  else if (mode == 2)
    PHIP(p, hipStreamSynchronize(p->ustream));
  bool any = false;
  for (int id = 0; id < p->max_jobs; id++) {
    HostJob& h = p->jobs[id];
    if (!h.used || !h.pending || (mode == 0 && p->step_calls < h.join_after))
      continue;
    const hipError_t q = hipEventQuery(p->jev[id]);
    if (q == hipErrorNotReady)
      continue;
    if (q != hipSuccess)
      return pfail(p, EK_EHIP, std::string("asynchronous admission: ") + hipGetErrorString(q));
    const int par = (int)(p->nstep % 2);
    char* dst = (char*)p->A[par] + (size_t)h.off * p->elem;
    if (par != h.upar)
      PHIP(p, hipMemcpyAsync(dst, (const char*)p->A[h.upar] + (size_t)h.off * p->elem, (size_t)h.nelem * p->elem,
                             hipMemcpyDeviceToDevice, p->stream));
    h.adm_step = p->nstep;
    if (int rc = init_slots(p, &h.p, id, dst))
      return rc;
    h.pending = false;
    any = true;
  }
  if (!any)
    return EK_OK;
  PHIP(p, hipStreamSynchronize(p->stream));
  return rebuild_work(p);
}

extern "C" int ek_pool_test_defer_join(ek_pool* p, int calls)
{
  if (!p || calls < 0)
    return EK_EINVAL;
  p->test_defer_join = calls;
  return EK_OK;
}

extern "C" int ek_pool_test_fail_admit(ek_pool* p, int stage)
{
  if (!p || stage < 0 || stage > 3)
    return EK_EINVAL;
  p->test_fail_admit = stage;
  return EK_OK;
}

extern "C" int ek_pool_step(ek_pool* p, int nsteps)
{
  if (!p || nsteps < 0)
    return pfail(p, EK_EINVAL, "invalid argument");
  // jobs whose upload has completed join now; with none stepping, wait for the earliest
  p->step_calls++;
  {
    bool stepping = false;
    for (const HostJob& h : p->jobs)
      stepping = stepping || (h.used && !h.pending && !h.finished);
    if (int rc = activate_pending(p, stepping ? 0 : 1))
      return rc;
  }
  int launches = 0;
  PHIP(p, hipEventRecord(p->ev0, p->stream));
  int left = nsteps;
  while (left > 0) {
    if (p->nstep % kGraphSteps == 0 && left >= kGraphSteps && p->nwork > 0) {
      if (!p->graph_ok)
        if (int rc = capture_graph(p))
          return rc;
      PHIP(p, hipGraphLaunch(p->graph, p->stream));
      p->nstep += kGraphSteps;
      left -= kGraphSteps;
      launches += kGraphSteps;
    } else {
      PHIP(p, launch_one(p, p->nstep));
      p->nstep += 1;
      left -= 1;
      launches += 1;
    }
  }
  PHIP(p, hipEventRecord(p->ev1, p->stream));
  PHIP(p, hipStreamSynchronize(p->stream));
  float ms = 0.f;
  PHIP(p, hipEventElapsedTime(&ms, p->ev0, p->ev1));
  p->last_ms = ms;
  p->last_launches = launches;
  bool changed = false;
  if (int rc = refresh_status(p, changed))
    return rc;
  if (changed)
    return rebuild_work(p); // drop the tiles of jobs that finished
  return EK_OK;
}

extern "C" int ek_pool_status(const ek_pool* p, int id, double* t, int* it, int* finished)
{
  if (!p || id < 0 || id >= p->max_jobs || !p->jobs[id].used)
    return EK_EINVAL;
  if (t)
    *t = p->jobs[id].t;
  if (it)
    *it = p->jobs[id].it;
  if (finished)
    *finished = p->jobs[id].finished;
  return EK_OK;
}

extern "C" int ek_pool_download(ek_pool* p, int id, double* host, int layout)
{
  if (!p || !host || id < 0 || id >= p->max_jobs || !p->jobs[id].used)
    return pfail(p, EK_EINVAL, "invalid job");
  if (p->jobs[id].pending)
    if (int rc = activate_pending(p, 2))
      return rc;
  const HostJob& h = p->jobs[id];
  const char* src = (const char*)p->A[job_parity(p, h)] + (size_t)h.off * p->elem;
  if (layout == EK_LAYOUT_SOA) {
    if (p->fp64) {
      PHIP(p, hipMemcpyAsync(host, src, (size_t)h.nelem * 8, hipMemcpyDeviceToHost, p->stream));
      PHIP(p, hipStreamSynchronize(p->stream));
    } else {
      std::vector<float> tmp((size_t)h.nelem);
      PHIP(p, hipMemcpyAsync(tmp.data(), src, (size_t)h.nelem * 4, hipMemcpyDeviceToHost, p->stream));
      PHIP(p, hipStreamSynchronize(p->stream));
      for (long long i = 0; i < h.nelem; i++)
        host[i] = tmp[i];
    }
    return EK_OK;
  }
  if (layout != EK_LAYOUT_REFERENCE)
    return pfail(p, EK_EINVAL, "unknown layout");
  const ek_params& jp = h.p;
  const int is = jp.nx + 2 * jp.gw, js = jp.ny + 2 * jp.gw, ks = jp.nz + 2 * jp.gw;
  const long long nr = (long long)is * js * (jp.dim == 3 ? ks : 1) * jp.nbvar;
  if (int rc = ensure_stage(p, nr))
    return rc;
  if (p->fp64) {
    ek::Geom<double> g;
    ek::Gas<double> gs;
    ek::fill_geom<double>(jp, g, gs);
    PHIP(p, ek::fp64::soa_to_ref(jp.dim, (const double*)src, p->stage, g, jp.gw, jp.nbvar, is, js, ks, jp.mhd,
                                 p->stream));
  } else {
    ek::Geom<float> g;
    ek::Gas<float> gs;
    ek::fill_geom<float>(jp, g, gs);
    PHIP(p, ek::fp32::soa_to_ref(jp.dim, (const float*)src, p->stage, g, jp.gw, jp.nbvar, is, js, ks, jp.mhd,
                                 p->stream));
  }
  PHIP(p, hipMemcpyAsync(host, p->stage, (size_t)nr * 8, hipMemcpyDeviceToHost, p->stream));
  PHIP(p, hipStreamSynchronize(p->stream));
  return EK_OK;
}

extern "C" int ek_pool_budgets(ek_pool* p, int id, double* sums)
{
  if (!p || !sums || id < 0 || id >= p->max_jobs || !p->jobs[id].used)
    return pfail(p, EK_EINVAL, "invalid job");
  if (p->jobs[id].pending)
    if (int rc = activate_pending(p, 2))
      return rc;
  const HostJob& h = p->jobs[id];
  const int nv = h.p.nbvar;
  const int nb = (int)std::min<long long>(1024, (h.ncell + 255) / 256);
  const char* src = (const char*)p->A[job_parity(p, h)] + (size_t)h.off * p->elem;
  if (p->fp64)
    PHIP(p, ek::fp64::budget(nv, (const double*)src, h.ncell, h.ncell, p->partial, nb, p->stream));
  else
    PHIP(p, ek::fp32::budget(nv, (const float*)src, h.ncell, h.ncell, p->partial, nb, p->stream));
  std::vector<double> part((size_t)nb * nv * 2);
  PHIP(p, hipMemcpyAsync(part.data(), p->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  PHIP(p, hipStreamSynchronize(p->stream));
  for (int v = 0; v < nv; v++) {
    double S = 0.0, C = 0.0;
    for (int b = 0; b < nb; b++)
      for (int k = 0; k < 2; k++) {
        const double x = part[((size_t)b * nv + v) * 2 + k];
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

extern "C" int ek_pool_retire(ek_pool* p, int id)
{
  if (!p || id < 0 || id >= p->max_jobs || !p->jobs[id].used)
    return pfail(p, EK_EINVAL, "invalid job");
  const bool pending = p->jobs[id].pending;
  if (pending)
    PHIP(p, hipEventSynchronize(p->jev[id])); // its upload still writes the arena range
  release_job(p, id);
  return pending ? EK_OK : rebuild_work(p);
}

static int apply_variant(ek_pool* p, int v)
{
  const int nvar = p->mhd3 ? 1 : (p->kind.dim == 2 ? ek::pool::N2VARIANTS : 4);
  if (v < 0 || v >= nvar)
    return pfail(p, EK_EINVAL, "tile choice out of range");
  p->variant = v;
  set_grid(p);
  for (int id = 0; id < p->max_jobs; id++)
    if (p->jobs[id].used)
      if (int rc = upload_record(p, id))
        return rc;
  return rebuild_work(p);
}

extern "C" int ek_pool_set_tile(ek_pool* p, int choice)
{
  if (!p)
    return EK_EINVAL;
  return apply_variant(p, choice);
}

extern "C" int ek_pool_autotune(ek_pool* p, int steps, int* choice, double* ms_out, int n_ms)
{
  if (!p || steps <= 0)
    return pfail(p, EK_EINVAL, "invalid argument");
  if (int rc = activate_pending(p, 2)) // the snapshot below holds every job's state
    return rc;
  const int nvar = p->mhd3 ? 1 : (p->kind.dim == 2 ? ek::pool::N2VARIANTS : 4);
  // snapshot of the current buffer and the time slots, restored after every trial
  long long hi = 0;
  for (const HostJob& h : p->jobs)
    if (h.used)
      hi = std::max(hi, h.off + h.alloc);
  void* snap = nullptr;
  ek::TimeSlot* tsnap = nullptr;
  const size_t sbytes = (size_t)hi * p->elem, tbytes = (size_t)p->max_jobs * 3 * sizeof(ek::TimeSlot);
  if (hipMalloc(&snap, sbytes ? sbytes : 8) != hipSuccess || hipMalloc(&tsnap, tbytes) != hipSuccess) {
    hipFree(snap);
    return pfail(p, EK_ENOMEM, "autotune snapshot allocation failed");
  }
  const int par = (int)(p->nstep % 2);
  int rc = EK_OK;
  double best = 1e300;
  int bestv = p->variant;
  auto restore = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(p->A[par], snap, sbytes, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(p->ts, tsnap, tbytes, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(p->stream);
    return e;
  };
  hipError_t e = hipMemcpyAsync(snap, p->A[par], sbytes, hipMemcpyDeviceToDevice, p->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(tsnap, p->ts, tbytes, hipMemcpyDeviceToDevice, p->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(p->stream);
  for (int v = 0; v < nvar && e == hipSuccess && rc == EK_OK; v++) {
    if ((rc = apply_variant(p, v)))
      break;
    // one untimed step (first use of the kernel), restore, then the timed steps
    if ((e = launch_one(p, p->nstep)) != hipSuccess || (e = restore()) != hipSuccess)
      break;
    if ((e = hipEventRecord(p->ev0, p->stream)) != hipSuccess)
      break;
    for (int s = 0; s < steps && e == hipSuccess; s++)
      e = launch_one(p, p->nstep + s);
    if (e != hipSuccess || (e = hipEventRecord(p->ev1, p->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(p->stream)) != hipSuccess)
      break;
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, p->ev0, p->ev1)) != hipSuccess)
      break;
    const double per = ms / steps;
    if (ms_out && v < n_ms)
      ms_out[v] = per;
    if (per < best) {
      best = per;
      bestv = v;
    }
    if ((e = restore()) != hipSuccess)
      break;
  }
  hipFree(snap);
  hipFree(tsnap);
  if (e != hipSuccess)
    return pfail(p, EK_EHIP, std::string("autotune: ") + hipGetErrorString(e));
  if (rc)
    return rc;
  if ((rc = apply_variant(p, bestv)))
    return rc;
  if (choice)
    *choice = bestv;
  return EK_OK;
}

extern "C" int ek_pool_last_step_time(const ek_pool* p, double* ms, int* launches)
{
  if (!p || !ms || !launches)
    return EK_EINVAL;
  *ms = p->last_ms;
  *launches = p->last_launches;
  return EK_OK;
}
