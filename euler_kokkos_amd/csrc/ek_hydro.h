// ek_hydro.h -- fused MUSCL-Hancock hydro step kernels for gfx950 (templated on the real type).
//
// Device state is SoA per variable, interior cells only, x fastest:
//   U[v][k][j][i]  (2D: U[v][j][i]), plane stride = nx*ny, var stride = nx*ny*nz.
// No ghost cells are stored: the boundary conditions of SolverBase::make_boundaries
// (src/shared/BoundariesFunctors.h:17-523) are applied when a tile loads its halo, by
// mapping every out-of-domain index back to the interior cell the reference's X, then Y,
// then Z ghost passes would copy from (the composition of the per-axis maps, with the
// normal momentum negated once per reflecting axis). That reproduces the reference's
// ghost values exactly, corners included.
//
// One launch performs one complete godunov_unsplit_impl step (src/muscl/SolverHydroMuscl.cpp:
// 214-390): primitives, slopes, trace, Riemann fluxes, the conservative update and the
// CFL reduction for the next step, with the time-step bookkeeping of SolverBase::compute_dt /
// next_iteration (src/shared/SolverBase.cpp:124-195) kept on the device.
#pragma once

#include "ek_cellmath.h"

namespace ek {

// Device-resident time bookkeeping. Step n reads slot n%3, writes slot (n+1)%3 and
// clears the CFL accumulator of slot (n+2)%3, so no launch both reads and writes a slot.
struct TimeSlot {
  double t;
  double dt;
  unsigned long long invdt_bits; // max-reduced as the bit pattern of a non-negative double
  int it;
  int finished;
};

struct StepCtl {
  TimeSlot* ts;
  unsigned long long* spread; // extra CFL words of the slots (see invdt_word), or null
  int slot;        // slot this launch reads
  int mode;        // 0: dt = cfl/invdt (clamped to tEnd), 1: dt = dt_given
  double dt_given;
  double cfl;
  double tEnd;
  int nStepmax;
};

template <class T>
struct Geom {
  int nx, ny, nz;
  int bc[6];
  long long plane; // nx*ny
  long long vstride;
  double dx, dy, dz;
  T tdx, tdy, tdz;    // dx.. in T
  T idx, idy, idz;    // 1/dx.. (exact when pow2)
  int pow2;
};

// Source index and reflection flag of position g on an axis of n cells.
__device__ __forceinline__ int bc_map(int g, int n, int bc_lo, int bc_hi, bool& neg)
{
  neg = false;
  int s = g;
  if (g < 0) {
    if (bc_lo == BC_DIRICHLET) {
      neg = true;
      s = -1 - g;
    } else if (bc_lo == BC_NEUMANN) {
      s = 0;
    } else {
      s = n + g;
    }
  } else if (g >= n) {
    if (bc_hi == BC_DIRICHLET) {
      neg = true;
      s = 2 * n - 1 - g;
    } else if (bc_hi == BC_NEUMANN) {
      s = n - 1;
    } else {
      s = g - n;
    }
  }
  s = s < 0 ? 0 : (s >= n ? n - 1 : s);
  return s;
}

__device__ __forceinline__ double bits_to_double(unsigned long long b)
{
  return __longlong_as_double((long long)b);
}
__device__ __forceinline__ unsigned long long double_to_bits(double d)
{
  return (unsigned long long)__double_as_longlong(d);
}

// Wave-wide maxima with DPP row shifts and row broadcasts (GFX9 DPP), result valid in
// lane 63 and returned via readlane. dpp_shr_bits moves a 64-bit value by one step of the
// pattern; lanes with no source read 0, the identity of both maxima below.
__device__ __forceinline__ unsigned long long dpp_shr_bits(unsigned long long v, int ctrl_sel)
{
  const int lo = (int)(unsigned)v;
  const int hi = (int)(unsigned)(v >> 32);
  int lo2, hi2;
  switch (ctrl_sel) {
    case 1:
      lo2 = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xf, 0xf, false);
      hi2 = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xf, 0xf, false);
      break;
    case 2:
      lo2 = __builtin_amdgcn_update_dpp(0, lo, 0x112, 0xf, 0xf, false);
      hi2 = __builtin_amdgcn_update_dpp(0, hi, 0x112, 0xf, 0xf, false);
      break;
    case 4:
      lo2 = __builtin_amdgcn_update_dpp(0, lo, 0x114, 0xf, 0xf, false);
      hi2 = __builtin_amdgcn_update_dpp(0, hi, 0x114, 0xf, 0xf, false);
      break;
    case 8:
      lo2 = __builtin_amdgcn_update_dpp(0, lo, 0x118, 0xf, 0xf, false);
      hi2 = __builtin_amdgcn_update_dpp(0, hi, 0x118, 0xf, 0xf, false);
      break;
    case 15:
      lo2 = __builtin_amdgcn_update_dpp(0, lo, 0x142, 0xa, 0xf, false);
      hi2 = __builtin_amdgcn_update_dpp(0, hi, 0x142, 0xa, 0xf, false);
      break;
    default:
      lo2 = __builtin_amdgcn_update_dpp(0, lo, 0x143, 0xc, 0xf, false);
      hi2 = __builtin_amdgcn_update_dpp(0, hi, 0x143, 0xc, 0xf, false);
      break;
  }
  return ((unsigned long long)(unsigned)hi2 << 32) | (unsigned)lo2;
}

__device__ __forceinline__ double dpp_max_shr(double v, int ctrl_sel)
{
  const double o = bits_to_double(dpp_shr_bits(double_to_bits(v), ctrl_sel));
  // values are >= 0
  return v > o ? v : (o == o ? o : v);
}

__device__ __forceinline__ unsigned long long wave_umax(unsigned long long v)
{
  unsigned long long o;
  o = dpp_shr_bits(v, 1);
  v = o > v ? o : v;
  o = dpp_shr_bits(v, 2);
  v = o > v ? o : v;
  o = dpp_shr_bits(v, 4);
  v = o > v ? o : v;
  o = dpp_shr_bits(v, 8);
  v = o > v ? o : v;
  o = dpp_shr_bits(v, 15);
  v = o > v ? o : v;
  o = dpp_shr_bits(v, 31);
  v = o > v ? o : v;
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ double wave_max(double v)
{
  v = dpp_max_shr(v, 1);
  v = dpp_max_shr(v, 2);
  v = dpp_max_shr(v, 4);
  v = dpp_max_shr(v, 8);
  v = dpp_max_shr(v, 15);
  v = dpp_max_shr(v, 31);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// CFL words of a time slot. Every block of a step ends with one device-scope atomic max,
// and such atomics on one address serialize: with thousands of blocks per launch that cost
// up to three quarters of a kernel's time. With StepCtl::spread set, the maximum of slot s
// lives in NSPREAD words 256 bytes apart -- ts[s].invdt_bits (key 0) and
// spread[(s*NSPREAD + k)*SPREAD_STRIDE] for k = 1..NSPREAD-1 -- block b updates word
// b % NSPREAD and readers take the max of all of them. Max is exact, so dt is unchanged.
constexpr int NSPREAD = 64;
constexpr int SPREAD_STRIDE = 32;
constexpr long long SPREAD_ELEMS = 3LL * NSPREAD * SPREAD_STRIDE;

__device__ __forceinline__ unsigned long long* invdt_word(const StepCtl& c, int slot, unsigned key)
{
  const unsigned k = key & (NSPREAD - 1);
  if (!c.spread || k == 0)
    return &c.ts[slot].invdt_bits;
  return c.spread + ((long long)slot * NSPREAD + k) * SPREAD_STRIDE;
}

// Max over the words of `slot`; `own` is ts[slot].invdt_bits as already read. Call with
// the whole wave active (lane k reads word k).
__device__ __forceinline__ unsigned long long invdt_gather(const StepCtl& c, int slot, unsigned long long own)
{
  static_assert(NSPREAD == 64, "one word per lane");
  if (!c.spread)
    return own;
  const int lane = threadIdx.x & 63;
  const unsigned long long v =
      lane == 0 ? own : c.spread[((long long)slot * NSPREAD + lane) * SPREAD_STRIDE];
  return wave_umax(v);
}

// Word of block blockIdx.x in a standalone CFL reduction: dst itself, or, with spread
// words (spread + k*SPREAD_STRIDE for k = 1..NSPREAD-1, k = 0 being dst), word b % NSPREAD.
__device__ __forceinline__ unsigned long long* reduce_word(unsigned long long* dst, unsigned long long* spread)
{
  const unsigned k = blockIdx.x & (NSPREAD - 1);
  return (spread && k) ? spread + (long long)k * SPREAD_STRIDE : dst;
}

// Block max then one atomic per block. NaN (a failed cell) wins: its bit pattern is above
// every finite double's, so a blown-up state yields dt = NaN instead of being skipped.
template <int NT>
__device__ __forceinline__ void block_max_to(unsigned long long* dst, double v, double* sred)
{
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const bool bad = !(v == v);
  const unsigned long long anybad = __ballot(bad);
  double m = wave_max(bad ? 0.0 : v);
  if (anybad)
    m = __longlong_as_double(0x7ff8000000000000LL);
  if (lane == 0)
    sred[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long best = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
      const unsigned long long b = double_to_bits(sred[w]);
      best = b > best ? b : best;
    }
    atomicMax(dst, best);
  }
}

// Per-launch time step (SolverBase::compute_dt + SolverHydroMuscl::compute_dt_local:
// dt = cfl/invDt, clamped so that t+dt does not pass tEnd). Returns false when the run
// has finished (SolverBase::finished), in which case block 0 forwards the slot.
__device__ __forceinline__ bool step_begin(const StepCtl& c, bool first_block, double& dt,
                                           TimeSlot& cur)
{
  cur = c.ts[c.slot];
  const int nxt = (c.slot + 1) % 3;
  if (c.mode == 0) {
    if (cur.finished || cur.t >= (c.tEnd - 1e-14) || cur.it >= c.nStepmax) {
      if (first_block && threadIdx.x == 0) {
        TimeSlot f = cur;
        f.finished = 1;
        c.ts[nxt] = f;
      }
      return false;
    }
    dt = c.cfl / bits_to_double(invdt_gather(c, c.slot, cur.invdt_bits));
    if (cur.t + dt > c.tEnd)
      dt = c.tEnd - cur.t;
  } else {
    dt = c.dt_given;
  }
  return true;
}

// Called by the first wave of the step's first block: lane 0 writes the next slot's time
// and clears the slot after it; the other lanes clear that slot's spread words.
__device__ __forceinline__ void step_end_bookkeeping(const StepCtl& c, const TimeSlot& cur, double dt)
{
  const int nxt = (c.slot + 1) % 3;
  const int nn = (c.slot + 2) % 3;
  const int lane = threadIdx.x & 63;
  if (lane == 0) {
    c.ts[nxt].t = cur.t + dt;
    c.ts[nxt].dt = dt;
    c.ts[nxt].it = cur.it + 1;
    c.ts[nxt].finished = 0;
    c.ts[nn].invdt_bits = 0ull;
  } else if (c.spread) {
    c.spread[((long long)nn * NSPREAD + lane) * SPREAD_STRIDE] = 0ull;
  }
}

// CFL term of one cell of the new state. The FP64 state uses the reference's expression
// as is; an FP32 state is widened so that the dt reduction always runs in FP64, with one
// FP64 division (1/rho) where the reference's expression has five (FP32 3D step 0.553 ->
// 0.538 ms at 256^3; the divisions by the cell size stay: a hoisted 1/dx costs registers).
template <class T, int DIM>
__device__ __forceinline__ double cfl_term(const Gas<T>& gas, const T* u, const Geom<T>& geo)
{
  if constexpr (sizeof(T) == 8) {
    return (double)cell_inv_dt<T, DIM>(gas, u, geo.tdx, geo.tdy, geo.tdz, geo.idx, geo.idy, geo.idz,
                                       geo.pow2 != 0);
  } else {
    const double d = fmax((double)u[ID], (double)gas.smallr);
    const double id = 1.0 / d;
    const double ux = (double)u[IU] * id, uy = (double)u[IV] * id;
    double ek = ux * ux + uy * uy;
    double uz = 0.0;
    if constexpr (DIM == 3) {
      uz = (double)u[IW] * id;
      ek += uz * uz;
    }
    const double e = (double)u[IP] * id - 0.5 * ek;
    const double p = fmax((double)gas.gm1 * d * e, d * (double)gas.smallp);
    const double c = sqrt((double)gas.gamma0 * p * id);
    double s = (c + fabs(ux)) / geo.dx + (c + fabs(uy)) / geo.dy;
    if constexpr (DIM == 3)
      s += (c + fabs(uz)) / geo.dz;
    return s;
  }
}

// ------------------------------------------------------------------------------------
// 2D: one block per TX x TY tile. Work items per tile: primitives on (TX+4)x(TY+4)
// positions, slopes+trace on the tile and its 1-wide halo cross, (TX+1)*TY x-faces and
// TX*(TY+1) y-faces, then the update of the TX*TY cells. Every face flux is solved once
// per tile; slopes are computed once per position (the reference recomputes them up to 3x).
// ------------------------------------------------------------------------------------
// One TX x TY tile of the 2D step, origin (i0, j0), with time step dt. Returns this
// thread's contribution to the CFL maximum of the new state. Shared by the single-grid
// kernel below and the batched (job pool) kernel.
template <class T, int TX, int TY, int NT, int RS, int ST>
__device__ __forceinline__ double hydro2d_tile(const T* __restrict__ Uin, T* __restrict__ Uout,
                                               const Geom<T>& geo, const Gas<T>& gas, double dt, int i0, int j0)
{
  constexpr int NV = 4;
  constexpr int QX = TX + 4, QY = TY + 4, QN = QX * QY;
  __shared__ T sQ[NV][QN];
  __shared__ T sX[2][NV][TY][TX + 1]; // [0]: left state at face f (cell f-1), [1]: right state (cell f)
  __shared__ T sY[2][NV][TY + 1][TX];

  const int tid = threadIdx.x;
  const T dtdx[2] = { T(dt / geo.dx), T(dt / geo.dy) };
  const int nx = geo.nx, ny = geo.ny;

  // primitives of the tile plus a 2-cell halo, ghosts resolved through the boundary map
  for (int w = tid; w < QN; w += NT) {
    const int qx = w % QX, qy = w / QX;
    bool nx_, ny_;
    const int si = bc_map(i0 - 2 + qx, nx, geo.bc[0], geo.bc[1], nx_);
    const int sj = bc_map(j0 - 2 + qy, ny, geo.bc[2], geo.bc[3], ny_);
    const long long o = (long long)sj * nx + si;
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
      u[v] = Uin[v * geo.vstride + o];
    if (nx_)
      u[IU] = u[IU] * T(-1);
    if (ny_)
      u[IV] = u[IV] * T(-1);
    T q[NV];
    prims<T, 2>(gas, u, q);
#pragma unroll
    for (int v = 0; v < NV; v++)
      sQ[v][w] = q[v];
  }
  __syncthreads();

  // slopes + face states: tile cells give 4 states, halo cells the one state a tile face needs
  constexpr int PT = TX * TY, PH = 2 * TY + 2 * TX;
  for (int w = tid; w < PT + PH; w += NT) {
    int cx, cy;
    unsigned want;
    if (w < PT) {
      cx = w % TX;
      cy = w / TX;
      want = 0xF;
    } else {
      const int h = w - PT;
      if (h < TY) {
        cx = -1; cy = h; want = 0x2;          // left halo: its max-x face
      } else if (h < 2 * TY) {
        cx = TX; cy = h - TY; want = 0x1;     // right halo: its min-x face
      } else if (h < 2 * TY + TX) {
        cx = h - 2 * TY; cy = -1; want = 0x8; // bottom halo: its max-y face
      } else {
        cx = h - 2 * TY - TX; cy = TY; want = 0x4;
      }
    }
    const int c = (cy + 2) * QX + (cx + 2);
    T q[NV], dq[3][5];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      q[v] = sQ[v][c];
      if constexpr (ST == 0) {
        dq[0][v] = T(0);
        dq[1][v] = T(0);
      } else {
        dq[0][v] = slope1<T, ST>(q[v], sQ[v][c + 1], sQ[v][c - 1]);
        dq[1][v] = slope1<T, ST>(q[v], sQ[v][c + QX], sQ[v][c - QX]);
      }
    }
    T fm[3][5], fp[3][5];
    trace_cell<T, 2>(gas, q, dq, dtdx, want, fm, fp);
    if (want & 0x2) {
#pragma unroll
      for (int v = 0; v < NV; v++)
        sX[0][v][cy][cx + 1] = fm[0][v];
    }
    if (want & 0x1) {
#pragma unroll
      for (int v = 0; v < NV; v++)
        sX[1][v][cy][cx] = fp[0][v];
    }
    if (want & 0x8) {
#pragma unroll
      for (int v = 0; v < NV; v++)
        sY[0][v][cy + 1][cx] = fm[1][v];
    }
    if (want & 0x4) {
#pragma unroll
      for (int v = 0; v < NV; v++)
        sY[1][v][cy][cx] = fp[1][v];
    }
  }
  __syncthreads();

  // Riemann problems: x faces (no swap), y faces (IU<->IV swapped around the solve)
  constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1);
  for (int w = tid; w < NFX + NFY; w += NT) {
    T ql[NV], qr[NV], f[NV];
    if (w < NFX) {
      const int fx = w % (TX + 1), fy = w / (TX + 1);
#pragma unroll
      for (int v = 0; v < NV; v++) {
        ql[v] = sX[0][v][fy][fx];
        qr[v] = sX[1][v][fy][fx];
      }
      riemann<T, NV, RS>(gas, ql, qr, f);
#pragma unroll
      for (int v = 0; v < NV; v++)
        sX[0][v][fy][fx] = f[v] * dtdx[0];
    } else {
      const int h = w - NFX;
      const int fx = h % TX, fy = h / TX;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        ql[v] = sY[0][v][fy][fx];
        qr[v] = sY[1][v][fy][fx];
      }
      T t;
      t = ql[IU]; ql[IU] = ql[IV]; ql[IV] = t;
      t = qr[IU]; qr[IU] = qr[IV]; qr[IV] = t;
      riemann<T, NV, RS>(gas, ql, qr, f);
      sY[0][ID][fy][fx] = f[ID] * dtdx[1];
      sY[0][IP][fy][fx] = f[IP] * dtdx[1];
      sY[0][IU][fy][fx] = f[IV] * dtdx[1];
      sY[0][IV][fy][fx] = f[IU] * dtdx[1];
    }
  }
  __syncthreads();

  // update (UpdateFunctor2D order: +Fx(i) -Fx(i+1) +Fy(j) -Fy(j+1)) and CFL of the new state
  double local = 0.0;
  for (int w = tid; w < PT; w += NT) {
    const int cx = w % TX, cy = w / TX;
    const int gi = i0 + cx, gj = j0 + cy;
    if (gi < nx && gj < ny) {
      const long long o = (long long)gj * nx + gi;
      T u[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        T a = Uin[v * geo.vstride + o];
        a += sX[0][v][cy][cx];
        a -= sX[0][v][cy][cx + 1];
        a += sY[0][v][cy][cx];
        a -= sY[0][v][cy + 1][cx];
        u[v] = a;
        Uout[v * geo.vstride + o] = a;
      }
      const double s = cfl_term<T, 2>(gas, u, geo);
      local = (s > local || !(s == s)) ? s : local;
    }
  }
  return local;
}

template <class T, int TX, int TY, int NT, int RS, int ST>
__global__ __launch_bounds__(NT) void hydro2d_step(const T* __restrict__ Uin, T* __restrict__ Uout,
                                                   Geom<T> geo, Gas<T> gas, StepCtl ctl)
{
  __shared__ double sred[NT / 64];
  const bool first_block = blockIdx.x == 0 && blockIdx.y == 0;
  double dt;
  TimeSlot cur;
  if (!step_begin(ctl, first_block, dt, cur))
    return;
  const double local =
      hydro2d_tile<T, TX, TY, NT, RS, ST>(Uin, Uout, geo, gas, dt, blockIdx.x * TX, blockIdx.y * TY);
  block_max_to<NT>(invdt_word(ctl, (ctl.slot + 1) % 3, blockIdx.x + blockIdx.y * gridDim.x), local, sred);
  if (first_block && threadIdx.x < 64)
    step_end_bookkeeping(ctl, cur, dt);
}

// ------------------------------------------------------------------------------------
// CFL reduction over the current state (ComputeDtFunctor2D/3D), into slot `slot`.
// ------------------------------------------------------------------------------------
template <class T, int DIM, int NT>
__global__ __launch_bounds__(NT) void inv_dt_kernel(const T* __restrict__ U, Geom<T> geo, Gas<T> gas,
                                                    unsigned long long* dst, unsigned long long* spread,
                                                    long long ncell)
{
  __shared__ double sred[NT / 64];
  constexpr int NV = DIM == 3 ? 5 : 4;
  double local = 0.0;
  for (long long c = (long long)blockIdx.x * NT + threadIdx.x; c < ncell; c += (long long)gridDim.x * NT) {
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
      u[v] = U[v * geo.vstride + c];
    const double s = cfl_term<T, DIM>(gas, u, geo);
    local = (s > local || !(s == s)) ? s : local;
  }
  block_max_to<NT>(reduce_word(dst, spread), local, sred);
}

// Compensated (Neumaier) FP64 sums of each conserved variable over the interior:
// per-block partials, summed again on the host in block order (deterministic).
template <class T, int NV, int NT>
__global__ __launch_bounds__(NT) void budget_kernel(const T* __restrict__ U, long long ncell,
                                                    long long vstride, double* partial)
{
  __shared__ double ss[NV][NT];
  __shared__ double sc[NV][NT];
  double s[NV], cc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    s[v] = 0.0;
    cc[v] = 0.0;
  }
  for (long long c = (long long)blockIdx.x * NT + threadIdx.x; c < ncell; c += (long long)gridDim.x * NT) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const double x = (double)U[v * vstride + c];
      const double t = s[v] + x;
      if (fabs(s[v]) >= fabs(x))
        cc[v] += (s[v] - t) + x;
      else
        cc[v] += (x - t) + s[v];
      s[v] = t;
    }
  }
#pragma unroll
  for (int v = 0; v < NV; v++) {
    ss[v][threadIdx.x] = s[v];
    sc[v][threadIdx.x] = cc[v];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int v = threadIdx.x;
    double S = 0.0, C = 0.0;
    for (int i = 0; i < NT; i++) {
      const double parts[2] = { ss[v][i], sc[v][i] };
      for (int k = 0; k < 2; k++) {
        const double x = parts[k];
        const double t = S + x;
        if (fabs(S) >= fabs(x))
          C += (S - t) + x;
        else
          C += (x - t) + S;
        S = t;
      }
    }
    partial[(long long)blockIdx.x * NV * 2 + 2 * v] = S;
    partial[(long long)blockIdx.x * NV * 2 + 2 * v + 1] = C;
  }
}

// Layout conversion between the reference layout (LayoutRight (i,j[,k],var) with ghosts,
// FP64) and the device SoA interior layout, in either direction. On the way out, ghost
// cells are filled through the same boundary map the step kernels use.
template <class T, int DIM>
__global__ void ref_to_soa(const double* __restrict__ R, T* __restrict__ U, Geom<T> geo, int gw,
                           int nbvar, int isize, int jsize, int ksize)
{
  const long long n = (long long)geo.vstride;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % geo.nx);
    const int j = (int)((c / geo.nx) % geo.ny);
    const int k = DIM == 3 ? (int)(c / geo.plane) : 0;
    long long ro;
    if (DIM == 2)
      ro = ((long long)(i + gw) * jsize + (j + gw)) * nbvar;
    else
      ro = (((long long)(i + gw) * jsize + (j + gw)) * ksize + (k + gw)) * nbvar;
    for (int v = 0; v < nbvar; v++)
      U[v * geo.vstride + c] = (T)R[ro + v];
  }
}

template <class T, int DIM>
__global__ void soa_to_ref(const T* __restrict__ U, double* __restrict__ R, Geom<T> geo, int gw,
                           int nbvar, int isize, int jsize, int ksize, int mhd)
{
  const long long n = (long long)isize * jsize * (DIM == 3 ? ksize : 1);
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    if (DIM == 2) {
      i = (int)(c / jsize);
      j = (int)(c % jsize);
      k = 0;
    } else {
      i = (int)(c / ((long long)jsize * ksize));
      j = (int)((c / ksize) % jsize);
      k = (int)(c % ksize);
    }
    bool ni, nj, nk = false;
    const int si = bc_map(i - gw, geo.nx, geo.bc[0], geo.bc[1], ni);
    const int sj = bc_map(j - gw, geo.ny, geo.bc[2], geo.bc[3], nj);
    int sk = 0;
    if (DIM == 3)
      sk = bc_map(k - gw, geo.nz, geo.bc[4], geo.bc[5], nk);
    const long long o = (long long)sk * geo.plane + (long long)sj * geo.nx + si;
    for (int v = 0; v < nbvar; v++) {
      double x = (double)U[v * geo.vstride + o];
      if ((ni && (v == IU || (mhd && v == IA))) || (nj && (v == IV || (mhd && v == IB))) ||
          (nk && (v == IW || (mhd && v == IC))))
        x = x * -1.0;
      R[c * nbvar + v] = x;
    }
  }
}

} // namespace ek
