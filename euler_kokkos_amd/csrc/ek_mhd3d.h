// ek_mhd3d.h -- 3D MHD MUSCL-Hancock step (HLLD fluxes, 3D constrained transport) for gfx950.
//
// One step is SolverMHDMuscl<3>::godunov_unsplit_impl, impl 0 (src/muscl/SolverMHDMuscl.cpp:
// 462-512), as a pipeline of kernels over a ghost-padded SoA work area (x fastest, 3 ghost
// layers each side, so padded index (i,j,k) is the reference's U(i,j,k)):
//   padprims: interior state -> padded state, ghosts through the boundary map, and the
//             primitives with the right/top/front face fields (ConvertToPrimitivesFunctor3D_MHD)
//   efield  : edge electric fields and transverse face-field slopes (ComputeElecFieldFunctor3D,
//             ComputeMagSlopesFunctor3D)
//   trace   : the half-step record of every cell (trace_unsplit_mhd_3d_simpler): centred
//             state, face fields and half slopes, 35 values, from which each of the 6 face
//             states and 12 edge states is the centre plus one or two half slopes
//   fluxemf : rebuilds the face and edge states from the records with the reference's own
//             additions and floors, then 3 Riemann problems and 3 edge EMFs per cell
//             (ComputeFluxesAndStoreFunctor3D_MHD, ComputeEmfAndStoreFunctor3D)
//   update  : flux update, CT update and the CFL term of the new state (UpdateFunctor3D_MHD,
//             UpdateEmfFunctor3D, ComputeDtFunctor3D_MHD), with the device time bookkeeping
// Every kernel reads the step's dt from the device time slot.
#pragma once

#include "ek_mhd.h"

namespace ek {
namespace mhd3 {

using namespace ek::mhd;

// Work-area layout. All arrays have PN = PX*PY*PZ elements.
struct Work3 {
  int PX, PY, PZ;
  long long PN;
  int nx, ny, nz;
};

// offsets (in arrays of PN elements) inside the scratch block
enum {
  A_UP = 0,           // padded state, 8
  A_Q = 8,            // primitives, 8
  A_E = 16,           // Ex, Ey, Ez
  A_DB = 19,          // dA/dy, dA/dz, dB/dx, dB/dz, dC/dx, dC/dy (full, not halved)
  A_REC = 25,         // half-step record, K_N values (below)
  A_FLUX = 60,        // Fx, Fy, Fz: 3 x 5 (ID IP IU IV IW of the update, unrotated)
  A_EMF = 75,         // EMF x, y, z
  A_TOTAL = 78
};

// Half-step record of a cell: K_C + v the centred state (component order of the state),
// K_AL..K_CR the six face fields, then the half slopes along x, y and z, each group
// holding ID IP IU IV IW at 0..4 and the two transverse field components at 5, 6
// (x: IB IC; y: IA IC; z: IA IB).
enum { K_C = 0, K_AL = 8, K_AR, K_BL, K_BR, K_CL, K_CR, K_X = 14, K_Y = 21, K_Z = 28, K_N = 35 };

// padded linear index (below 2^30: the host keeps every array below 4 GiB)
__device__ __forceinline__ unsigned pidx(const Work3& w, int i, int j, int k)
{
  return ((unsigned)k * (unsigned)w.PY + (unsigned)j) * (unsigned)w.PX + (unsigned)i;
}

// Element at a 32-bit byte offset from a wave-uniform array base: the access addresses as a
// scalar base + one 32-bit lane offset per cell, instead of a 64-bit address per (array,
// cell), which takes two registers and 64-bit adds each. Each array of one grid is below
// 4 GiB (checked by the host). The base is the same in every lane; passing it through
// readfirstlane keeps the compiler from folding the array offset into a 64-bit per-lane
// address. Used where it measured faster: the record reads of the flux/EMF kernel (fewer
// instructions and registers) and the prims and update kernels; the pad, E-field and trace
// kernels keep plain indexing, with which the compiler spaces their loads better.
typedef char __attribute__((address_space(1))) * gbytes_t;

template <class T>
__device__ __forceinline__ gbytes_t sbase(const T* base)
{
  const unsigned long long a = reinterpret_cast<unsigned long long>(base);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  return (gbytes_t)(((unsigned long long)hi << 32) | lo);
}

template <class T>
__device__ __forceinline__ T ldb(const T* base, unsigned boff)
{
  return *(const T __attribute__((address_space(1)))*)(sbase(base) + boff);
}

template <class T>
__device__ __forceinline__ void stb(T* base, unsigned boff, T x)
{
  *(T __attribute__((address_space(1)))*)(sbase(base) + boff) = x;
}

// (i, j, k) of linear index c in a box of X x Y x (any) cells. 32-bit division: the host
// keeps every work-area array below 4 GiB, so every index here is below 2^30 (a 64-bit
// division would cost each thread a long software sequence).
__device__ __forceinline__ void split3(unsigned c, unsigned X, unsigned Y, int& i, int& j, int& k)
{
  const unsigned r = c / X;
  const unsigned z = r / Y;
  i = (int)(c - r * X);
  j = (int)(r - z * Y);
  k = (int)z;
}

template <class T>
__device__ __forceinline__ bool cell_of(const Work3& w, int& i, int& j, int& k)
{
  const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (unsigned long long)w.PN)
    return false;
  split3(c, (unsigned)w.PX, (unsigned)w.PY, i, j, k);
  return true;
}

// step_begin without the finished-slot forwarding (the update kernel owns it)
__device__ __forceinline__ bool step_dt(const StepCtl& c, double& dt)
{
  const TimeSlot cur = c.ts[c.slot];
  if (c.mode == 0) {
    if (cur.finished || cur.t >= (c.tEnd - 1e-14) || cur.it >= c.nStepmax)
      return false;
    dt = c.cfl / bits_to_double(invdt_gather(c, c.slot, cur.invdt_bits));
    if (cur.t + dt > c.tEnd)
      dt = c.tEnd - cur.t;
  } else {
    dt = c.dt_given;
  }
  return true;
}

// pad + prims in one pass. The padded state: ghost values as SolverBase::make_boundaries
// fills them (X, then Y, then Z passes; reflecting axes negate the normal momentum and the
// normal face field). On i < PX-1, j < PY-1, k < PZ-1 also the primitives
// (ConvertToPrimitivesFunctor3D_MHD), whose right / top / front face fields are the padded
// values of those neighbours, read from the interior state through the same map.
template <class T>
__device__ __forceinline__ void padprims_cell(const T* __restrict__ U, T* __restrict__ S, const Work3& w,
                                              const Geom<T>& geo, const Gas<T>& gas, int i, int j, int k)
{
  bool fx, fy, fz;
  const int si = bc_map(i - 3, w.nx, geo.bc[0], geo.bc[1], fx);
  const int sj = bc_map(j - 3, w.ny, geo.bc[2], geo.bc[3], fy);
  const int sk = bc_map(k - 3, w.nz, geo.bc[4], geo.bc[5], fz);
  const long long vs = geo.vstride;
  const long long o = ((long long)sk * w.ny + sj) * w.nx + si;
  const long long po = pidx(w, i, j, k), PN = w.PN;
  T u[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    T x = U[v * vs + o];
    if ((fx && (v == IU || v == IA)) || (fy && (v == IV || v == IB)) || (fz && (v == IW || v == IC)))
      x = x * T(-1);
    S[(A_UP + v) * PN + po] = x;
    u[v] = x;
  }
  if (i >= w.PX - 1 || j >= w.PY - 1 || k >= w.PZ - 1)
    return;
  // the padded IA of (i+1,j,k), IB of (i,j+1,k), IC of (i,j,k+1): only the reflection
  // along the field's own axis negates it
  bool gx, gy, gz;
  const int si1 = bc_map(i - 2, w.nx, geo.bc[0], geo.bc[1], gx);
  const int sj1 = bc_map(j - 2, w.ny, geo.bc[2], geo.bc[3], gy);
  const int sk1 = bc_map(k - 2, w.nz, geo.bc[4], geo.bc[5], gz);
  T mx = U[IA * vs + ((long long)sk * w.ny + sj) * w.nx + si1];
  T my = U[IB * vs + ((long long)sk * w.ny + sj1) * w.nx + si];
  T mz = U[IC * vs + ((long long)sk1 * w.ny + sj) * w.nx + si];
  if (gx)
    mx = mx * T(-1);
  if (gy)
    my = my * T(-1);
  if (gz)
    mz = mz * T(-1);
  T q[NV];
  constoprim<T>(gas, u, mx, my, mz, q);
#pragma unroll
  for (int v = 0; v < NV; v++)
    S[(A_Q + v) * PN + po] = q[v];
}

template <class T>
__global__ __launch_bounds__(256) void padprims_kernel(const T* __restrict__ U, T* __restrict__ S, Work3 w,
                                                       Geom<T> geo, Gas<T> gas, StepCtl ctl)
{
  double dt;
  if (!step_dt(ctl, dt))
    return;
  int i, j, k;
  if (!cell_of<T>(w, i, j, k))
    return;
  padprims_cell<T>(U, S, w, geo, gas, i, j, k);
}

// edge electric fields and transverse face-field slopes on 1 <= i,j,k < size-1
template <class T, int ST>
__device__ __forceinline__ void efield_cell(T* __restrict__ S, const Work3& w, int i, int j, int k)
{
  if (i < 1 || j < 1 || k < 1 || i >= w.PX - 1 || j >= w.PY - 1 || k >= w.PZ - 1)
    return;
  const long long PN = w.PN, o = pidx(w, i, j, k);
  const long long sx = 1, sy = w.PX, sz = (long long)w.PX * w.PY;
  const T* Q = S + A_Q * PN;
  const T* Up = S + A_UP * PN;
  auto q = [&](int v, long long off) -> T { return Q[v * PN + o + off]; };
  auto u = [&](int v, long long off) -> T { return Up[v * PN + o + off]; };
  T vv, ww, uu, A, B, C;
  vv = T(0.25) * (q(IV, -sy - sz) + q(IV, -sy) + q(IV, -sz) + q(IV, 0));
  ww = T(0.25) * (q(IW, -sy - sz) + q(IW, -sy) + q(IW, -sz) + q(IW, 0));
  B = T(0.5) * (u(IB, -sz) + u(IB, 0));
  C = T(0.5) * (u(IC, -sy) + u(IC, 0));
  S[(A_E + 0) * PN + o] = vv * C - ww * B;
  uu = T(0.25) * (q(IU, -sx - sz) + q(IU, -sx) + q(IU, -sz) + q(IU, 0));
  ww = T(0.25) * (q(IW, -sx - sz) + q(IW, -sx) + q(IW, -sz) + q(IW, 0));
  A = T(0.5) * (u(IA, -sz) + u(IA, 0));
  C = T(0.5) * (u(IC, -sx) + u(IC, 0));
  S[(A_E + 1) * PN + o] = ww * A - uu * C;
  uu = T(0.25) * (q(IU, -sx - sy) + q(IU, -sx) + q(IU, -sy) + q(IU, 0));
  vv = T(0.25) * (q(IV, -sx - sy) + q(IV, -sx) + q(IV, -sy) + q(IV, 0));
  A = T(0.5) * (u(IA, -sy) + u(IA, 0));
  B = T(0.5) * (u(IB, -sx) + u(IB, 0));
  S[(A_E + 2) * PN + o] = uu * B - vv * A;
  // slope_unsplit_mhd_3d with xslope_type = min(slope_type, 2) (MHDBaseFunctor3D.h:563-668)
  S[(A_DB + 0) * PN + o] = slope1<T, ST>(u(IA, 0), u(IA, sy), u(IA, -sy));
  S[(A_DB + 1) * PN + o] = slope1<T, ST>(u(IA, 0), u(IA, sz), u(IA, -sz));
  S[(A_DB + 2) * PN + o] = slope1<T, ST>(u(IB, 0), u(IB, sx), u(IB, -sx));
  S[(A_DB + 3) * PN + o] = slope1<T, ST>(u(IB, 0), u(IB, sz), u(IB, -sz));
  S[(A_DB + 4) * PN + o] = slope1<T, ST>(u(IC, 0), u(IC, sx), u(IC, -sx));
  S[(A_DB + 5) * PN + o] = slope1<T, ST>(u(IC, 0), u(IC, sy), u(IC, -sy));
}

template <class T, int ST>
__global__ __launch_bounds__(256) void efield_kernel(T* __restrict__ S, Work3 w, StepCtl ctl)
{
  double dt;
  if (!step_dt(ctl, dt))
    return;
  int i, j, k;
  if (!cell_of<T>(w, i, j, k))
    return;
  efield_cell<T, ST>(S, w, i, j, k);
}

// Inputs of the 3D trace of one cell, gathered from wherever the caller keeps them.
template <class T>
struct TraceIn {
  T qc[NV];                  // primitives of the cell
  T dqx[NV], dqy[NV], dqz[NV];  // their slopes (slope1 of the neighbours along x, y, z)
  T ELL, ELR, ERL, ERR;      // Ex at (j, k), (j, k+1), (j+1, k), (j+1, k+1)
  T FLL, FLR, FRL, FRR;      // Ey at (i, k), (i, k+1), (i+1, k), (i+1, k+1)
  T GLL, GLR, GRL, GRR;      // Ez at (i, j), (i, j+1), (i+1, j), (i+1, j+1)
  T AL, AR, BL, BR, CL, CR;  // the cell's own face fields (left/right)
};

// The half-step record of one cell (trace_unsplit_mhd_3d_simpler): `put(k, value)` receives
// the 35 values K_C .. K_Z+6.
template <class T, int ST, class Put>
__device__ __forceinline__ void trace_record(const TraceIn<T>& in, const Gas<T>& gas, T dtdx, T dtdy, T dtdz,
                                             Put put)
{
  const T* qc = in.qc;
  const T* dqx = in.dqx;
  const T* dqy = in.dqy;
  const T* dqz = in.dqz;
  const T ELL = in.ELL, ELR = in.ELR, ERL = in.ERL, ERR = in.ERR;
  const T FLL = in.FLL, FLR = in.FLR, FRL = in.FRL, FRR = in.FRR;
  const T GLL = in.GLL, GLR = in.GLR, GRL = in.GRL, GRR = in.GRR;
  T r = qc[ID], p = qc[IP], u = qc[IU], v = qc[IV], ww = qc[IW], A = qc[IA], B = qc[IB], C = qc[IC];
  T AL = in.AL, AR = in.AR, BL = in.BL, BR = in.BR, CL = in.CL, CR = in.CR;
  const T h = T(0.5);
  const T drx = dqx[ID] * h, dpx = dqx[IP] * h, dux = dqx[IU] * h, dvx = dqx[IV] * h, dwx = dqx[IW] * h;
  const T dCx = dqx[IC] * h, dBx = dqx[IB] * h;
  const T dry = dqy[ID] * h, dpy = dqy[IP] * h, duy = dqy[IU] * h, dvy = dqy[IV] * h, dwy = dqy[IW] * h;
  const T dCy = dqy[IC] * h, dAy = dqy[IA] * h;
  const T drz = dqz[ID] * h, dpz = dqz[IP] * h, duz = dqz[IU] * h, dvz = dqz[IV] * h, dwz = dqz[IW] * h;
  const T dAz = dqz[IA] * h, dBz = dqz[IB] * h;
  const T dAx = h * (AR - AL), dBy = h * (BR - BL), dCz = h * (CR - CL);
  const T gamma = gas.gamma0;
  const T sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy + (-ww * drz - dwz * r) * dtdz;
  const T su0 = (-u * dux - (dpx + B * dBx + C * dCx) / r) * dtdx + (-v * duy + B * dAy / r) * dtdy +
                (-ww * duz + C * dAz / r) * dtdz;
  const T sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - (dpy + A * dAy + C * dCy) / r) * dtdy +
                (-ww * dvz + C * dBz / r) * dtdz;
  const T sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy +
                (-ww * dwz - (dpz + A * dAz + B * dBz) / r) * dtdz;
  const T sp0 = (-u * dpx - dux * gamma * p) * dtdx + (-v * dpy - dvy * gamma * p) * dtdy +
                (-ww * dpz - dwz * gamma * p) * dtdz;
  const T sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy + (u * dCz + C * duz - ww * dAz - A * dwz) * dtdz;
  const T sB0 = (v * dAx + A * dvx - u * dBx - B * dux) * dtdx + (v * dCz + C * dvz - ww * dBz - B * dwz) * dtdz;
  const T sC0 = (ww * dAx + A * dwx - u * dCx - C * dux) * dtdx + (ww * dBy + B * dwy - v * dCy - C * dvy) * dtdy;
  const T sAL0 = +(GLR - GLL) * dtdy * h - (FLR - FLL) * dtdz * h;
  const T sAR0 = +(GRR - GRL) * dtdy * h - (FRR - FRL) * dtdz * h;
  const T sBL0 = -(GRL - GLL) * dtdx * h + (ELR - ELL) * dtdz * h;
  const T sBR0 = -(GRR - GLR) * dtdx * h + (ERR - ERL) * dtdz * h;
  const T sCL0 = +(FRL - FLL) * dtdx * h - (ERL - ELL) * dtdy * h;
  const T sCR0 = +(FRR - FLR) * dtdx * h - (ERR - ELR) * dtdy * h;
  r = r + sr0;
  u = u + su0;
  v = v + sv0;
  ww = ww + sw0;
  p = p + sp0;
  A = A + sA0;
  B = B + sB0;
  C = C + sC0;
  AL = AL + sAL0;
  AR = AR + sAR0;
  BL = BL + sBL0;
  BR = BR + sBR0;
  CL = CL + sCL0;
  CR = CR + sCR0;
  put(K_C + ID, r);
  put(K_C + IP, p);
  put(K_C + IU, u);
  put(K_C + IV, v);
  put(K_C + IW, ww);
  put(K_C + IA, A);
  put(K_C + IB, B);
  put(K_C + IC, C);
  put(K_AL, AL);
  put(K_AR, AR);
  put(K_BL, BL);
  put(K_BR, BR);
  put(K_CL, CL);
  put(K_CR, CR);
  put(K_X + 0, drx);
  put(K_X + 1, dpx);
  put(K_X + 2, dux);
  put(K_X + 3, dvx);
  put(K_X + 4, dwx);
  put(K_X + 5, dBx);
  put(K_X + 6, dCx);
  put(K_Y + 0, dry);
  put(K_Y + 1, dpy);
  put(K_Y + 2, duy);
  put(K_Y + 3, dvy);
  put(K_Y + 4, dwy);
  put(K_Y + 5, dAy);
  put(K_Y + 6, dCy);
  put(K_Z + 0, drz);
  put(K_Z + 1, dpz);
  put(K_Z + 2, duz);
  put(K_Z + 3, dvz);
  put(K_Z + 4, dwz);
  put(K_Z + 5, dAz);
  put(K_Z + 6, dBz);
}

// the 3D trace (trace_unsplit_mhd_3d_simpler) on 1 <= i,j,k < size-2
template <class T, int ST>
__device__ __forceinline__ void trace_cell(T* __restrict__ S, const Work3& w, const Geom<T>& geo, const Gas<T>& gas,
                                           double dt, int i, int j, int k)
{
  if (i < 1 || j < 1 || k < 1 || i >= w.PX - 2 || j >= w.PY - 2 || k >= w.PZ - 2)
    return;
  const T dtdx = T(dt / geo.dx), dtdy = T(dt / geo.dy), dtdz = T(dt / geo.dz);
  const long long PN = w.PN, o = pidx(w, i, j, k);
  const long long sx = 1, sy = w.PX, sz = (long long)w.PX * w.PY;
  const T* Q = S + A_Q * PN;
  const T* Up = S + A_UP * PN;
  const T* E = S + A_E * PN;
  TraceIn<T> in;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const T q = Q[v * PN + o];
    in.qc[v] = q;
    in.dqx[v] = slope1<T, ST>(q, Q[v * PN + o + sx], Q[v * PN + o - sx]);
    in.dqy[v] = slope1<T, ST>(q, Q[v * PN + o + sy], Q[v * PN + o - sy]);
    in.dqz[v] = slope1<T, ST>(q, Q[v * PN + o + sz], Q[v * PN + o - sz]);
  }
  in.ELL = E[o];
  in.ELR = E[o + sz];
  in.ERL = E[o + sy];
  in.ERR = E[o + sy + sz];
  in.FLL = E[PN + o];
  in.FLR = E[PN + o + sz];
  in.FRL = E[PN + o + sx];
  in.FRR = E[PN + o + sx + sz];
  in.GLL = E[2 * PN + o];
  in.GLR = E[2 * PN + o + sy];
  in.GRL = E[2 * PN + o + sx];
  in.GRR = E[2 * PN + o + sx + sy];
  in.AL = Up[IA * PN + o];
  in.AR = Up[IA * PN + o + sx];
  in.BL = Up[IB * PN + o];
  in.BR = Up[IB * PN + o + sy];
  in.CL = Up[IC * PN + o];
  in.CR = Up[IC * PN + o + sz];
  T* Rw = S + (long long)A_REC * PN + o;
  trace_record<T, ST>(in, gas, dtdx, dtdy, dtdz, [&](int kk, T x) { Rw[(long long)kk * PN] = x; });
}

template <class T, int ST>
__global__ __launch_bounds__(256) void trace_kernel(T* __restrict__ S, Work3 w, Geom<T> geo, Gas<T> gas, StepCtl ctl)
{
  double dt;
  if (!step_dt(ctl, dt))
    return;
  int i, j, k;
  if (!cell_of<T>(w, i, j, k))
    return;
  trace_cell<T, ST>(S, w, geo, gas, dt, i, j, k);
}

template <class T>
struct RecView {
  const T* S;
  long long PN;
  __device__ __forceinline__ T operator()(int kk, unsigned boff) const
  {
    return ldb(S + (long long)(A_REC + kk) * PN, boff);
  }
  __device__ __forceinline__ T db(int m, unsigned boff) const
  {
    return ldb(S + (long long)(A_DB + m) * PN, boff);
  }
};

// Face state of the cell at byte offset cb along dir (0..2): maxface -> qm (centre + half
// slope, FACE_*MAX), else qp (centre - half slope); the normal field is the face's own.
// Floors as the 3D trace (trace_unsplit_mhd_3d_simpler): rho by smallr, p by smallp.
template <class T>
__device__ __forceinline__ void face_state(const RecView<T>& R, unsigned cb, int dir, bool maxface,
                                           const Gas<T>& gas, T* q)
{
  const int K = K_X + 7 * dir;
#pragma unroll
  for (int v = 0; v < 5; v++)
    q[v] = maxface ? R(K_C + v, cb) + R(K + v, cb) : R(K_C + v, cb) - R(K + v, cb);
  const int m1 = dir == 0 ? IB : IA, m2 = dir == 2 ? IB : IC, mn = dir == 0 ? IA : (dir == 1 ? IB : IC);
  q[m1] = maxface ? R(K_C + m1, cb) + R(K + 5, cb) : R(K_C + m1, cb) - R(K + 5, cb);
  q[m2] = maxface ? R(K_C + m2, cb) + R(K + 6, cb) : R(K_C + m2, cb) - R(K + 6, cb);
  q[mn] = R((maxface ? K_AR : K_AL) + 2 * dir, cb);
  q[ID] = emax(gas.smallr, q[ID]);
  q[IP] = emax(gas.smallp, q[IP]);
}

// Edge state e (0..3 = RT, RB, LT, LB) along d (0..2 = X, Y, Z) of the cell at byte offset
// cb, as the edge solver reads it. Over the two transverse directions (X: y, z; Y: x, z;
// Z: x, y) a state is centre + (+-s1 +- s2) (RT ++, RB +-, LT -+, LB --); the two
// transverse face fields are the right or left face value plus or minus its own
// transverse half slope, h*DB at that face (ComputeMagSlopesFunctor3D's slopes, halved as
// the trace does). bx, by, bz: byte strides of the padded grid.
template <class T>
__device__ __forceinline__ Edge<T> edge_state(const RecView<T>& R, unsigned cb, int e, int d, unsigned bx,
                                              unsigned by, unsigned bz, const Gas<T>& gas)
{
  const T h = T(0.5);
  const int K1 = d == 0 ? K_Y : K_X, K2 = d == 2 ? K_Y : K_Z;
  const bool p1 = e < 2, p2 = (e & 1) == 0;
  auto comb = [&](T base, T a, T b) {
    const T t = p1 ? a : -a;
    return base + (p2 ? t + b : t - b);
  };
  T q[NV];
#pragma unroll
  for (int v = 0; v < 5; v++)
    q[v] = comb(R(K_C + v, cb), R(K1 + v, cb), R(K2 + v, cb));
  if (d == 0) {
    q[IA] = comb(R(K_C + IA, cb), R(K_Y + 5, cb), R(K_Z + 5, cb));
    const T sb = p1 ? h * R.db(3, cb + by) : h * R.db(3, cb);   // dBRz / dBLz
    q[IB] = (p1 ? R(K_BR, cb) : R(K_BL, cb)) + (p2 ? sb : -sb);
    const T sc = p2 ? h * R.db(5, cb + bz) : h * R.db(5, cb);   // dCRy / dCLy
    q[IC] = (p2 ? R(K_CR, cb) : R(K_CL, cb)) + (p1 ? sc : -sc);
  } else if (d == 1) {
    q[IB] = comb(R(K_C + IB, cb), R(K_X + 5, cb), R(K_Z + 6, cb));
    const T sa = p1 ? h * R.db(1, cb + bx) : h * R.db(1, cb);   // dARz / dALz
    q[IA] = (p1 ? R(K_AR, cb) : R(K_AL, cb)) + (p2 ? sa : -sa);
    const T sc = p2 ? h * R.db(4, cb + bz) : h * R.db(4, cb);   // dCRx / dCLx
    q[IC] = (p2 ? R(K_CR, cb) : R(K_CL, cb)) + (p1 ? sc : -sc);
  } else {
    q[IC] = comb(R(K_C + IC, cb), R(K_X + 6, cb), R(K_Y + 6, cb));
    const T sa = p1 ? h * R.db(0, cb + bx) : h * R.db(0, cb);   // dARy / dALy
    q[IA] = (p1 ? R(K_AR, cb) : R(K_AL, cb)) + (p2 ? sa : -sa);
    const T sb = p2 ? h * R.db(2, cb + by) : h * R.db(2, cb);   // dBRx / dBLx
    q[IB] = (p2 ? R(K_BR, cb) : R(K_BL, cb)) + (p1 ? sb : -sb);
  }
  Edge<T> s;
  s.r = emax(gas.smallr, q[ID]);
  s.p = emax(gas.smallp, q[IP]);
  if (d == 0) {
    s.u = q[IV]; s.v = q[IW]; s.a = q[IB]; s.b = q[IC]; s.c = q[IA];
  } else if (d == 1) {
    s.u = q[IW]; s.v = q[IU]; s.a = q[IC]; s.b = q[IA]; s.c = q[IB];
  } else {
    s.u = q[IU]; s.v = q[IV]; s.a = q[IA]; s.b = q[IB]; s.c = q[IC];
  }
  return s;
}

// Fluxes and EMFs on 3 <= i,j,k < size-2. Four tasks per chunk of 256 cells: the three
// Riemann problems of each cell, or one of the three edge EMFs. One thread doing all six
// holds too many rebuilt states at once (over 200 VGPRs); one task each stays near 128.
// Launch 32 * ceil(nchunk / 8) blocks: block b takes task (b / 8) % 4 of chunk
// (b / 32) * 8 + b % 8, so the four tasks of a chunk run on one XCD (blocks are dealt to
// the XCDs round robin) and share its L2 copy of the chunk's records.
// One task of padded cell c: 0 its three Riemann problems, 1..3 its EMF z, y or x.
template <class T, int RS>
__device__ __forceinline__ void fluxemf_cell(T* __restrict__ S, const Work3& w, const Gas<T>& gas, unsigned c,
                                             int task)
{
  int i, j, k;
  split3(c, (unsigned)w.PX, (unsigned)w.PY, i, j, k);
  if (i < 3 || j < 3 || k < 3 || i >= w.PX - 2 || j >= w.PY - 2 || k >= w.PZ - 2)
    return;
  const long long PN = w.PN;
  const unsigned ob = c * (unsigned)sizeof(T), bx = sizeof(T), by = (unsigned)w.PX * (unsigned)sizeof(T);
  const unsigned bz = (unsigned)(w.PX * w.PY) * (unsigned)sizeof(T);
  const RecView<T> R{ S, PN };
  if (task != 0) {
    auto E = [&](unsigned cb, int e, int d) { return edge_state<T>(R, cb, e, d, bx, by, bz, gas); };
    if (task == 1) {
      // EMF z: RT(i-1,j-1,k), RB(i-1,j,k), LT(i,j-1,k), LB(i,j,k) of the Z edges
      S[(A_EMF + 2) * PN + c] = emf_z<T>(gas, E(ob - bx - by, 0, 2), E(ob - bx, 1, 2), E(ob - by, 2, 2), E(ob, 3, 2));
    } else if (task == 2) {
      // EMF y: RT(i-1,j,k-1), LT(i,j,k-1) in the RB place, RB(i-1,j,k) in the LT place, LB(i,j,k)
      S[(A_EMF + 1) * PN + c] = emf_z<T>(gas, E(ob - bx - bz, 0, 1), E(ob - bz, 2, 1), E(ob - bx, 1, 1), E(ob, 3, 1));
    } else {
      // EMF x: RT(i,j-1,k-1), RB(i,j-1,k), LT(i,j,k-1), LB(i,j,k)
      S[(A_EMF + 0) * PN + c] = emf_z<T>(gas, E(ob - by - bz, 0, 0), E(ob - by, 1, 0), E(ob - bz, 2, 0), E(ob, 3, 0));
    }
    return;
  }
  T ql[NV], qr[NV], f[NV];
  // x: qm_x(i-1) | qp_x(i)
  face_state<T>(R, ob - bx, 0, true, gas, ql);
  face_state<T>(R, ob, 0, false, gas, qr);
  riemann<T, RS>(gas, ql, qr, f);
  {
    T* F = S + (A_FLUX + 0) * PN + c;
    F[0] = f[ID]; F[PN] = f[IP]; F[2 * PN] = f[IU]; F[3 * PN] = f[IV]; F[4 * PN] = f[IW];
  }
  // y: qm_y(j-1) | qp_y(j), IU<->IV and IA<->IB around the solve
  face_state<T>(R, ob - by, 1, true, gas, ql);
  face_state<T>(R, ob, 1, false, gas, qr);
  {
    T t;
    t = ql[IU]; ql[IU] = ql[IV]; ql[IV] = t;
    t = ql[IA]; ql[IA] = ql[IB]; ql[IB] = t;
    t = qr[IU]; qr[IU] = qr[IV]; qr[IV] = t;
    t = qr[IA]; qr[IA] = qr[IB]; qr[IB] = t;
  }
  riemann<T, RS>(gas, ql, qr, f);
  {
    T* F = S + (A_FLUX + 5) * PN + c;
    F[0] = f[ID]; F[PN] = f[IP]; F[2 * PN] = f[IV]; F[3 * PN] = f[IU]; F[4 * PN] = f[IW];
  }
  // z: qm_z(k-1) | qp_z(k), IU<->IW and IA<->IC around the solve
  face_state<T>(R, ob - bz, 2, true, gas, ql);
  face_state<T>(R, ob, 2, false, gas, qr);
  {
    T t;
    t = ql[IU]; ql[IU] = ql[IW]; ql[IW] = t;
    t = ql[IA]; ql[IA] = ql[IC]; ql[IC] = t;
    t = qr[IU]; qr[IU] = qr[IW]; qr[IW] = t;
    t = qr[IA]; qr[IA] = qr[IC]; qr[IC] = t;
  }
  riemann<T, RS>(gas, ql, qr, f);
  {
    T* F = S + (A_FLUX + 10) * PN + c;
    F[0] = f[ID]; F[PN] = f[IP]; F[2 * PN] = f[IW]; F[3 * PN] = f[IV]; F[4 * PN] = f[IU];
  }
}

template <class T, int RS>
__global__ __launch_bounds__(256) void fluxemf_kernel(T* __restrict__ S, Work3 w, Gas<T> gas, StepCtl ctl)
{
  double dt;
  if (!step_dt(ctl, dt))
    return;
  const unsigned b = blockIdx.x;
  const int task = (b >> 3) & 3;
  const unsigned c = (((b >> 5) << 3) | (b & 7)) * 256u + threadIdx.x;
  if (c >= (unsigned long long)w.PN)
    return;
  fluxemf_cell<T, RS>(S, w, gas, c, task);
}

// CFL term of a 3D cell from its new state and new right/top/front face fields
template <class T>
__device__ __forceinline__ double cfl3(const Gas<T>& gas, const T* u, T mx, T my, T mz, const Geom<T>& geo)
{
  if constexpr (sizeof(T) == 8) {
    T q[NV];
    constoprim<T>(gas, u, mx, my, mz, q);
    const T d = q[ID], p = q[IP], a = q[IA], b = q[IB], c = q[IC];
    const T b2 = a * a + b * b + c * c;
    const T c2 = gas.gamma0 * p / d;
    const T d2 = T(0.5) * (b2 / d + c2);
    const T vx = esqrt(d2 + esqrt(d2 * d2 - c2 * a * a / d)) + eabs(q[IU]);
    const T vy = esqrt(d2 + esqrt(d2 * d2 - c2 * b * b / d)) + eabs(q[IV]);
    const T vz = esqrt(d2 + esqrt(d2 * d2 - c2 * c * c / d)) + eabs(q[IW]);
    if (geo.pow2)
      return vx * geo.idx + vy * geo.idy + vz * geo.idz;
    return vx / geo.dx + vy / geo.dy + vz / geo.dz;
  } else {
    return mhd_cfl_widened<3>(gas, u, (double)mx, (double)my, (double)mz, geo.dx, geo.dy, geo.dz);
  }
}

// Flux update, CT update and the CFL term of the new state of interior cell c
// (UpdateFunctor3D_MHD + UpdateEmfFunctor3D + ComputeDtFunctor3D_MHD on the result).
template <class T>
__device__ __forceinline__ double update_cell(const T* __restrict__ S, T* __restrict__ Uout, const Work3& w,
                                              const Geom<T>& geo, const Gas<T>& gas, T dtdx, T dtdy, T dtdz,
                                              long long c)
{
  const long long n = geo.vstride;
  {
    int ii, jj, kk;
    split3((unsigned)c, (unsigned)w.nx, (unsigned)w.ny, ii, jj, kk);
    const long long PN = w.PN;
    // byte offsets: this cell of the padded grid, its strides, this cell of the output
    const unsigned o = pidx(w, ii + 3, jj + 3, kk + 3) * (unsigned)sizeof(T);
    const unsigned sx = sizeof(T), sy = (unsigned)w.PX * (unsigned)sizeof(T);
    const unsigned sz = (unsigned)(w.PX * w.PY) * (unsigned)sizeof(T);
    const unsigned ob = (unsigned)c * (unsigned)sizeof(T);
    auto at = [&](int a, unsigned off) -> T { return ldb(S + a * PN, off); };
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
      u[v] = at(A_UP + v, o);
    constexpr int hv[5] = { ID, IP, IU, IV, IW };
#pragma unroll
    for (int m = 0; m < 5; m++) {
      T a = u[hv[m]];
      a += at(A_FLUX + m, o) * dtdx;
      a -= at(A_FLUX + m, o + sx) * dtdx;
      a += at(A_FLUX + 5 + m, o) * dtdy;
      a -= at(A_FLUX + 5 + m, o + sy) * dtdy;
      a += at(A_FLUX + 10 + m, o) * dtdz;
      a -= at(A_FLUX + 10 + m, o + sz) * dtdz;
      u[hv[m]] = a;
    }
    auto EX = [&](unsigned off) -> T { return at(A_EMF + 0, off); };
    auto EY = [&](unsigned off) -> T { return at(A_EMF + 1, off); };
    auto EZ = [&](unsigned off) -> T { return at(A_EMF + 2, off); };
    u[IA] += (EZ(o + sy) - EZ(o)) * dtdy;
    u[IB] -= (EZ(o + sx) - EZ(o)) * dtdx;
    u[IA] -= (EY(o + sz) - EY(o)) * dtdz;
    u[IB] += (EX(o + sz) - EX(o)) * dtdz;
    u[IC] += (EY(o + sx) - EY(o)) * dtdx;
    u[IC] -= (EX(o + sy) - EX(o)) * dtdy;
#pragma unroll
    for (int v = 0; v < NV; v++)
      stb(Uout + v * n, ob, u[v]);
    // new face fields of the right / top / front neighbours, as the next ghost fill sees them
    T ar, bt, cf;
    if (ii + 1 < w.nx || geo.bc[1] == BC_PERIODIC) {
      const unsigned q = o + sx;
      ar = at(A_UP + IA, q);
      ar += (EZ(q + sy) - EZ(q)) * dtdy;
      ar -= (EY(q + sz) - EY(q)) * dtdz;
    } else {
      ar = geo.bc[1] == BC_DIRICHLET ? u[IA] * T(-1) : u[IA];
    }
    if (jj + 1 < w.ny || geo.bc[3] == BC_PERIODIC) {
      const unsigned q = o + sy;
      bt = at(A_UP + IB, q);
      bt -= (EZ(q + sx) - EZ(q)) * dtdx;
      bt += (EX(q + sz) - EX(q)) * dtdz;
    } else {
      bt = geo.bc[3] == BC_DIRICHLET ? u[IB] * T(-1) : u[IB];
    }
    if (kk + 1 < w.nz || geo.bc[5] == BC_PERIODIC) {
      const unsigned q = o + sz;
      cf = at(A_UP + IC, q);
      cf += (EY(q + sx) - EY(q)) * dtdx;
      cf -= (EX(q + sy) - EX(q)) * dtdy;
    } else {
      cf = geo.bc[5] == BC_DIRICHLET ? u[IC] * T(-1) : u[IC];
    }
    return cfl3<T>(gas, u, ar, bt, cf, geo);
  }
}

// update on the interior (UpdateFunctor3D_MHD + UpdateEmfFunctor3D) into the SoA output.
// A grid-stride loop over a bounded grid (MHD3_UPDATE_BLOCKS): each block ends with one
// atomic max on the time slot, and one short block per 256 cells (27k blocks at 192^3)
// made those atomics cost three quarters of the kernel's time.
constexpr unsigned MHD3_UPDATE_BLOCKS = 2048;

template <class T>
__global__ __launch_bounds__(256) void update_kernel(const T* __restrict__ S, T* __restrict__ Uout, Work3 w,
                                                     Geom<T> geo, Gas<T> gas, StepCtl ctl)
{
  __shared__ double sred[4];
  double dt;
  TimeSlot cur;
  const bool first_block = blockIdx.x == 0;
  if (!step_begin(ctl, first_block, dt, cur))
    return;
  const T dtdx = T(dt / geo.dx), dtdy = T(dt / geo.dy), dtdz = T(dt / geo.dz);
  const long long n = geo.vstride;
  double local = 0.0;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const double sv = update_cell<T>(S, Uout, w, geo, gas, dtdx, dtdy, dtdz, c);
    local = (sv > local || !(sv == sv)) ? sv : local;
  }
  block_max_to<256>(invdt_word(ctl, (ctl.slot + 1) % 3, blockIdx.x), local, sred);
  if (first_block && threadIdx.x < 64)
    step_end_bookkeeping(ctl, cur, dt);
}

// CFL reduction over a 3D MHD state (interior SoA), neighbours' face fields through the
// boundary map
template <class T>
__global__ __launch_bounds__(256) void inv_dt3_kernel(const T* __restrict__ U, Geom<T> geo, Gas<T> gas,
                                                      unsigned long long* dst, unsigned long long* spread)
{
  __shared__ double sred[4];
  const long long n = geo.vstride;
  const int nx = geo.nx, ny = geo.ny, nz = geo.nz;
  double local = 0.0;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
    int i, j, k;
    split3((unsigned)c, (unsigned)nx, (unsigned)ny, i, j, k);
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
      u[v] = U[v * n + c];
    bool fx, fy, fz;
    const int si = bc_map(i + 1, nx, geo.bc[0], geo.bc[1], fx);
    const int sj = bc_map(j + 1, ny, geo.bc[2], geo.bc[3], fy);
    const int sk = bc_map(k + 1, nz, geo.bc[4], geo.bc[5], fz);
    T mx = U[IA * n + ((long long)k * ny + j) * nx + si];
    T my = U[IB * n + ((long long)k * ny + sj) * nx + i];
    T mz = U[IC * n + ((long long)sk * ny + j) * nx + i];
    if (fx)
      mx = mx * T(-1);
    if (fy)
      my = my * T(-1);
    if (fz)
      mz = mz * T(-1);
    const double s = cfl3<T>(gas, u, mx, my, mz, geo);
    local = (s > local || !(s == s)) ? s : local;
  }
  block_max_to<256>(reduce_word(dst, spread), local, sred);
}

} // namespace mhd3
} // namespace ek
