// ek_mhd.h -- 2D MHD MUSCL-Hancock step with HLLD fluxes and constrained transport, for gfx950.
//
// One launch performs one complete SolverMHDMuscl<2>::godunov_unsplit_impl step, impl 0
// (src/muscl/SolverMHDMuscl.cpp:369-454): ghost values, primitives, the 2D MHD trace,
// HLLD (or HLL / LLF) face fluxes, the 2D HLLD edge EMF, the flux and EMF updates, and the
// CFL reduction of the new state for the next step (ComputeDtFunctor2D_MHD), with the
// time-step bookkeeping kept on the device as for the hydro step.
//
// State layout as the hydro path: SoA, interior only, U[v][j][i] with 8 variables
// (ID IP IU IV IW IA IB IC); IA/IB are the face fields on the cell's left/bottom face, IC is
// cell centred. Ghost values come from the per-axis boundary map; reflecting axes negate
// the normal momentum and the normal face field (src/shared/BoundariesFunctors.h:534-748).
//
// Per block: a TX x TY tile of cells. Quantities that the reference computes per cell but
// that are functions of one face or one edge only are computed once per face/edge here,
// which leaves every value bitwise unchanged:
//   * the edge electric field u*B - v*A of trace_unsplit_mhd_2d (the 4 cells around an edge
//     form it from the same operands in the same order),
//   * the half-step face fields AL/AR/BL/BR (AR of cell i is AL of cell i+1),
//   * the transverse face-field slopes dALy/dARy/dBLx/dBRx.
// Every cell keeps a 22-value trace record (the half-step centred state and the half
// slopes); face and edge states are rebuilt from it with the reference's additions and
// floors when a Riemann problem needs them.
#pragma once

#include "ek_hydro.h"

namespace ek {
namespace mhd {

constexpr int NV = 8;

// constoprim_mhd (src/muscl/MHDBaseFunctor2D.h:192-242), cIso = 0.
template <class T>
__host__ __device__ inline void constoprim(const Gas<T>& g, const T* u, T mx, T my, T mz, T* q)
{
  q[ID] = emax(u[ID], g.smallr);
  q[IU] = u[IU] / q[ID];
  q[IV] = u[IV] / q[ID];
  q[IW] = u[IW] / q[ID];
  q[IA] = T(0.5) * (u[IA] + mx);
  q[IB] = T(0.5) * (u[IB] + my);
  q[IC] = T(0.5) * (u[IC] + mz);
  const T eken = T(0.5) * (q[IU] * q[IU] + q[IV] * q[IV] + q[IW] * q[IW]);
  const T emag = T(0.5) * (q[IA] * q[IA] + q[IB] * q[IB] + q[IC] * q[IC]);
  T eint = (u[IP] - emag) / q[ID] - eken;
  if constexpr (sizeof(T) == 4) {
    // FP32 state: p = (gamma-1)(E - rho v^2/2 - B^2/2) recovered in FP64 where more than 7
    // of the 24 mantissa bits cancel
    if (eint < T(0.0078125) * (u[IP] / q[ID])) {
      const double d = q[ID], ux = (double)u[IU] / d, uy = (double)u[IV] / d, uz = (double)u[IW] / d;
      const double bx = 0.5 * ((double)u[IA] + (double)mx), by = 0.5 * ((double)u[IB] + (double)my),
                   bz = 0.5 * ((double)u[IC] + (double)mz);
      const double em = 0.5 * (bx * bx + by * by + bz * bz);
      eint = (T)(((double)u[IP] - em) / d - 0.5 * (ux * ux + uy * uy + uz * uz));
    }
  }
  q[IP] = emax(g.gm1 * q[ID] * eint, q[ID] * g.smallp);
}

// find_speed_fast<dir> (src/shared/mhd_utils.h:89-117)
template <class T, int DIR>
__host__ __device__ inline T speed_fast(const Gas<T>& g, const T* q)
{
  const T d = q[ID], p = q[IP], a = q[IA], b = q[IB], c = q[IC];
  const T b2 = a * a + b * b + c * c;
  const T c2 = g.gamma0 * p / d;
  const T d2 = T(0.5) * (b2 / d + c2);
  const T n = DIR == 0 ? a : (DIR == 1 ? b : c);
  return esqrt(d2 + esqrt(d2 * d2 - c2 * n * n / d));
}

// find_speed_info<TWO_D> (mhd_utils.h:319-366) folded into ComputeDtFunctor2D_MHD
// (src/muscl/MHDRunFunctors2D.h:36-74): vx/dx + vy/dy.
template <class T>
__host__ __device__ inline T cell_inv_dt(const Gas<T>& g, const T* q, T dx, T dy, T idx, T idy, bool pow2)
{
  const T d = q[ID], p = q[IP], a = q[IA], b = q[IB], c = q[IC];
  const T b2 = a * a + b * b + c * c;
  const T c2 = g.gamma0 * p / d;
  const T d2 = T(0.5) * (b2 / d + c2);
  const T cfx = esqrt(d2 + esqrt(d2 * d2 - c2 * a * a / d));
  const T vx = cfx + eabs(q[IU]);
  const T cfy = esqrt(d2 + esqrt(d2 * d2 - c2 * b * b / d));
  const T vy = cfy + eabs(q[IV]);
  if (pow2)
    return vx * idx + vy * idy;
  return vx / dx + vy / dy;
}

// CFL term of an FP32 MHD cell, widened to FP64 so that the dt reduction always runs in
// FP64: the primitives with the cell-centred field from the face fields (mx, my, mz: the
// right / top / front ones; mz = 0 in 2D), then the fast speeds, with one division (1/rho)
// where the reference's expressions have eight (2D) or ten (3D).
template <int DIM>
__device__ __forceinline__ double mhd_cfl_widened(const Gas<float>& gas, const float* u, double mx, double my,
                                                  double mz, double dx, double dy, double dz)
{
  const double d = fmax((double)u[ID], (double)gas.smallr);
  const double id = 1.0 / d;
  const double ux = (double)u[IU] * id, uy = (double)u[IV] * id, uz = (double)u[IW] * id;
  const double a = 0.5 * ((double)u[IA] + mx), b = 0.5 * ((double)u[IB] + my), c = 0.5 * ((double)u[IC] + mz);
  const double eken = 0.5 * (ux * ux + uy * uy + uz * uz);
  const double b2 = a * a + b * b + c * c;
  const double eint = ((double)u[IP] - 0.5 * b2) * id - eken;
  const double p = fmax((double)gas.gm1 * d * eint, d * (double)gas.smallp);
  const double c2 = (double)gas.gamma0 * p * id;
  const double d2 = 0.5 * (b2 * id + c2);
  const double vx = sqrt(d2 + sqrt(d2 * d2 - c2 * a * a * id)) + fabs(ux);
  const double vy = sqrt(d2 + sqrt(d2 * d2 - c2 * b * b * id)) + fabs(uy);
  double s = vx / dx + vy / dy;
  if constexpr (DIM == 3) {
    const double vz = sqrt(d2 + sqrt(d2 * d2 - c2 * c * c * id)) + fabs(uz);
    s += vz / dz;
  }
  return s;
}

// CFL term of one cell: primitives (with the right/top face fields) then fast speeds. The
// FP64 state uses the reference's expressions; an FP32 state is widened (mhd_cfl_widened).
template <class T>
__device__ __forceinline__ double mhd_cfl_term(const Gas<T>& gas, const T* u, T mx, T my, const Geom<T>& geo)
{
  if constexpr (sizeof(T) == 8) {
    T q[NV];
    constoprim<T>(gas, u, mx, my, T(0), q);
    return cell_inv_dt<T>(gas, q, geo.tdx, geo.tdy, geo.idx, geo.idy, geo.pow2 != 0);
  } else {
    return mhd_cfl_widened<2>(gas, u, (double)mx, (double)my, 0.0, geo.dx, geo.dy, 1.0);
  }
}

// find_mhd_flux (mhd_utils.h:173-231)
template <class T>
__host__ __device__ inline void mhd_flux(const Gas<T>& g, const T* q, T* cv, T* ff)
{
  const T d = q[ID], p = q[IP], u = q[IU], v = q[IV], w = q[IW], a = q[IA], b = q[IB], c = q[IC];
  const T ecin = T(0.5) * (u * u + v * v + w * w) * d;
  const T emag = T(0.5) * (a * a + b * b + c * c);
  const T etot = p * g.entho + ecin + emag;
  const T ptot = p + emag;
  cv[ID] = d;
  cv[IP] = etot;
  cv[IU] = d * u;
  cv[IV] = d * v;
  cv[IW] = d * w;
  cv[IA] = a;
  cv[IB] = b;
  cv[IC] = c;
  ff[ID] = d * u;
  ff[IP] = (etot + ptot) * u - a * (a * u + b * v + c * w);
  ff[IU] = d * u * u - a * a + ptot;
  ff[IV] = d * u * v - a * b;
  ff[IW] = d * u * w - a * c;
  ff[IA] = T(0);
  ff[IB] = b * u - a * v;
  ff[IC] = c * u - a * w;
}

// riemann_hll, MHD (src/shared/RiemannSolvers_MHD.h:24-64)
template <class T>
__host__ __device__ inline void hll(const Gas<T>& g, T* ql, T* qr, T* flux)
{
  const T bx = T(0.5) * (ql[IA] + qr[IA]);
  ql[IA] = bx;
  qr[IA] = bx;
  T ul[NV], fl[NV], ur[NV], fr[NV];
  mhd_flux(g, ql, ul, fl);
  mhd_flux(g, qr, ur, fr);
  const T cfl = speed_fast<T, 0>(g, ql), cfr = speed_fast<T, 0>(g, qr);
  const T sl = emin(emin(ql[IU], qr[IU]) - emax(cfl, cfr), T(0));
  const T sr = emax(emax(ql[IU], qr[IU]) + emax(cfl, cfr), T(0));
#pragma unroll
  for (int n = 0; n < NV; n++)
    flux[n] = (sr * fl[n] - sl * fr[n] + sr * sl * (ur[n] - ul[n])) / (sr - sl);
}

// riemann_llf, MHD (RiemannSolvers_MHD.h:78-110)
template <class T>
__host__ __device__ inline void llf(const Gas<T>& g, T* ql, T* qr, T* flux)
{
  const T bx = T(0.5) * (ql[IA] + qr[IA]);
  ql[IA] = bx;
  qr[IA] = bx;
  T ul[NV], fl[NV], ur[NV], fr[NV];
  mhd_flux(g, ql, ul, fl);
  mhd_flux(g, qr, ur, fr);
#pragma unroll
  for (int n = 0; n < NV; n++)
    flux[n] = (fl[n] + fr[n]) / T(2);
  const T cl = speed_fast<T, 0>(g, ql) + eabs(ql[IU]);
  const T cr = speed_fast<T, 0>(g, qr) + eabs(qr[IU]);
  const T vel = emax(cl, cr);
#pragma unroll
  for (int n = 0; n < NV; n++)
    flux[n] -= vel * (ur[n] - ul[n]) / T(2);
}

// riemann_hlld (RiemannSolvers_MHD.h:130-364)
template <class T>
__host__ __device__ inline void hlld(const Gas<T>& g, const T* qleft, const T* qright, T* flux)
{
  const T a = T(0.5) * (qleft[IA] + qright[IA]);
  const T sgnm = (a >= T(0)) ? T(1) : T(-1);
  const T entho = g.entho;
  const T rl = qleft[ID], pl = qleft[IP], ul = qleft[IU], vl = qleft[IV], wl = qleft[IW];
  const T bl = qleft[IB], cl = qleft[IC];
  const T ecinl = T(0.5) * (ul * ul + vl * vl + wl * wl) * rl;
  const T emagl = T(0.5) * (a * a + bl * bl + cl * cl);
  const T etotl = pl * entho + ecinl + emagl;
  const T ptotl = pl + emagl;
  const T vdotbl = ul * a + vl * bl + wl * cl;
  const T rr = qright[ID], pr = qright[IP], ur = qright[IU], vr = qright[IV], wr = qright[IW];
  const T br = qright[IB], cr = qright[IC];
  const T ecinr = T(0.5) * (ur * ur + vr * vr + wr * wr) * rr;
  const T emagr = T(0.5) * (a * a + br * br + cr * cr);
  const T etotr = pr * entho + ecinr + emagr;
  const T ptotr = pr + emagr;
  const T vdotbr = ur * a + vr * br + wr * cr;
  T qtmp[NV];
  qtmp[ID] = rl; qtmp[IP] = pl; qtmp[IA] = a; qtmp[IB] = bl; qtmp[IC] = cl;
  const T cfastl = speed_fast<T, 0>(g, qtmp);
  qtmp[ID] = rr; qtmp[IP] = pr; qtmp[IB] = br; qtmp[IC] = cr;
  const T cfastr = speed_fast<T, 0>(g, qtmp);
  const T sl = emin(ul, ur) - emax(cfastl, cfastr);
  const T sr = emax(ul, ur) + emax(cfastl, cfastr);
  const T rcl = rl * (ul - sl);
  const T rcr = rr * (sr - ur);
  const T ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  const T ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);
  // left star region
  const T rstarl = rl * (sl - ul) / (sl - ustar);
  T estar = rl * (sl - ul) * (sl - ustar) - a * a;
  const T el = rl * (sl - ul) * (sl - ul) - a * a;
  T vstarl, wstarl, bstarl, cstarl;
  if (a * a > T(0) && eabs(estar / (a * a) - T(1)) <= T(1e-8)) {
    vstarl = vl; bstarl = bl; wstarl = wl; cstarl = cl;
  } else {
    vstarl = vl - a * bl * (ustar - ul) / estar;
    bstarl = bl * el / estar;
    wstarl = wl - a * cl * (ustar - ul) / estar;
    cstarl = cl * el / estar;
  }
  const T vdotbstarl = ustar * a + vstarl * bstarl + wstarl * cstarl;
  const T etotstarl = ((sl - ul) * etotl - ptotl * ul + ptotstar * ustar + a * (vdotbl - vdotbstarl)) / (sl - ustar);
  const T sqrrstarl = esqrt(rstarl);
  const T calfvenl = eabs(a) / sqrrstarl;
  const T sal = ustar - calfvenl;
  // right star region
  const T rstarr = rr * (sr - ur) / (sr - ustar);
  estar = rr * (sr - ur) * (sr - ustar) - a * a;
  const T er = rr * (sr - ur) * (sr - ur) - a * a;
  T vstarr, wstarr, bstarr, cstarr;
  if (a * a > T(0) && eabs(estar / (a * a) - T(1)) <= T(1e-8)) {
    vstarr = vr; bstarr = br; wstarr = wr; cstarr = cr;
  } else {
    vstarr = vr - a * br * (ustar - ur) / estar;
    bstarr = br * er / estar;
    wstarr = wr - a * cr * (ustar - ur) / estar;
    cstarr = cr * er / estar;
  }
  const T vdotbstarr = ustar * a + vstarr * bstarr + wstarr * cstarr;
  const T etotstarr = ((sr - ur) * etotr - ptotr * ur + ptotstar * ustar + a * (vdotbr - vdotbstarr)) / (sr - ustar);
  const T sqrrstarr = esqrt(rstarr);
  const T calfvenr = eabs(a) / sqrrstarr;
  const T sar = ustar + calfvenr;
  // double star region
  const T den = sqrrstarl + sqrrstarr;
  const T vstarstar = (sqrrstarl * vstarl + sqrrstarr * vstarr + sgnm * (bstarr - bstarl)) / den;
  const T wstarstar = (sqrrstarl * wstarl + sqrrstarr * wstarr + sgnm * (cstarr - cstarl)) / den;
  const T bstarstar = (sqrrstarl * bstarr + sqrrstarr * bstarl + sgnm * sqrrstarl * sqrrstarr * (vstarr - vstarl)) / den;
  const T cstarstar = (sqrrstarl * cstarr + sqrrstarr * cstarl + sgnm * sqrrstarl * sqrrstarr * (wstarr - wstarl)) / den;
  const T vdotbstarstar = ustar * a + vstarstar * bstarstar + wstarstar * cstarstar;
  const T etotstarstarl = etotstarl - sgnm * sqrrstarl * (vdotbstarl - vdotbstarstar);
  const T etotstarstarr = etotstarr + sgnm * sqrrstarr * (vdotbstarr - vdotbstarstar);
  T ro, uo, vo, wo, bo, co, ptoto, etoto, vdotbo;
  if (sl > T(0)) {
    ro = rl; uo = ul; vo = vl; wo = wl; bo = bl; co = cl; ptoto = ptotl; etoto = etotl; vdotbo = vdotbl;
  } else if (sal > T(0)) {
    ro = rstarl; uo = ustar; vo = vstarl; wo = wstarl; bo = bstarl; co = cstarl; ptoto = ptotstar;
    etoto = etotstarl; vdotbo = vdotbstarl;
  } else if (ustar > T(0)) {
    ro = rstarl; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarl; vdotbo = vdotbstarstar;
  } else if (sar > T(0)) {
    ro = rstarr; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarr; vdotbo = vdotbstarstar;
  } else if (sr > T(0)) {
    ro = rstarr; uo = ustar; vo = vstarr; wo = wstarr; bo = bstarr; co = cstarr; ptoto = ptotstar;
    etoto = etotstarr; vdotbo = vdotbstarr;
  } else {
    ro = rr; uo = ur; vo = vr; wo = wr; bo = br; co = cr; ptoto = ptotr; etoto = etotr; vdotbo = vdotbr;
  }
  flux[ID] = ro * uo;
  flux[IP] = (etoto + ptoto) * uo - a * vdotbo;
  flux[IU] = ro * uo * uo - a * a + ptoto;
  flux[IV] = ro * uo * vo - a * bo;
  flux[IW] = ro * uo * wo - a * co;
  flux[IA] = T(0);
  flux[IB] = bo * uo - a * vo;
  flux[IC] = co * uo - a * wo;
}

// riemann_mhd dispatch (RiemannSolvers_MHD.h:369-389)
template <class T, int RS>
__host__ __device__ inline void riemann(const Gas<T>& g, T* ql, T* qr, T* f)
{
  if constexpr (RS == RS_HLLD)
    hlld<T>(g, ql, qr, f);
  else if constexpr (RS == RS_HLL)
    hll<T>(g, ql, qr, f);
  else
    llf<T>(g, ql, qr, f);
}

// FMAX4 / FMIN4 / FMAX5 (src/shared/mhd_utils.h:17-60): sequential strict compares
template <class T>
__host__ __device__ inline T fmax4(T a0, T a1, T a2, T a3)
{
  T r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  return r;
}
template <class T>
__host__ __device__ inline T fmin4(T a0, T a1, T a2, T a3)
{
  T r = a0;
  r = (a1 < r) ? a1 : r;
  r = (a2 < r) ? a2 : r;
  r = (a3 < r) ? a3 : r;
  return r;
}
template <class T>
__host__ __device__ inline T fmax5(T a0, T a1, T a2, T a3, T a4)
{
  T r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  r = (a4 > r) ? a4 : r;
  return r;
}

// Edge state in the order the 2D magnetic solver reads it: r p u v a b c (w is not read).
template <class T>
struct Edge {
  T r, p, u, v, a, b, c;
};

template <class T>
__host__ __device__ inline T fast_x(const Gas<T>& g, const Edge<T>& s)
{
  const T b2 = s.a * s.a + s.b * s.b + s.c * s.c;
  const T c2 = g.gamma0 * s.p / s.r;
  const T d2 = T(0.5) * (b2 / s.r + c2);
  return esqrt(d2 + esqrt(d2 * d2 - c2 * s.a * s.a / s.r));
}
template <class T>
__host__ __device__ inline T fast_y(const Gas<T>& g, const Edge<T>& s)
{
  const T b2 = s.a * s.a + s.b * s.b + s.c * s.c;
  const T c2 = g.gamma0 * s.p / s.r;
  const T d2 = T(0.5) * (b2 / s.r + c2);
  return esqrt(d2 + esqrt(d2 * d2 - c2 * s.b * s.b / s.r));
}

// compute_emf<EMFZ> + mag_riemann2d_hlld (RiemannSolvers_MHD.h:395-627, 648-871). The
// edge states arrive as the trace produced them (qRT, qRB, qLT, qLB); the solver's
// LL/RL/LR/RR states are RT/LT/RB/LB with the transverse face fields averaged.
template <class T>
__host__ __device__ inline T emf_z(const Gas<T>& g, const Edge<T>& RT, const Edge<T>& RB, const Edge<T>& LT,
                                   const Edge<T>& LB)
{
  Edge<T> LL = RT, RL = LT, LR = RB, RR = LB;
  const T aT = T(0.5) * (RT.a + LT.a);
  const T aB = T(0.5) * (RB.a + LB.a);
  const T bL = T(0.5) * (RT.b + RB.b);
  const T bR = T(0.5) * (LT.b + LB.b);
  LL.a = aT; RL.a = aT; LR.a = aB; RR.a = aB;
  LL.b = bL; RL.b = bR; LR.b = bL; RR.b = bR;
  const T ELL = LL.u * LL.b - LL.v * LL.a;
  const T ERL = RL.u * RL.b - RL.v * RL.a;
  const T ELR = LR.u * LR.b - LR.v * LR.a;
  const T ERR = RR.u * RR.b - RR.v * RR.a;
  const T rLL = LL.r, uLL = LL.u, vLL = LL.v, aLL = LL.a, bLL = LL.b, cLL = LL.c, pLL = LL.p;
  const T rLR = LR.r, uLR = LR.u, vLR = LR.v, aLR = LR.a, bLR = LR.b, cLR = LR.c, pLR = LR.p;
  const T rRL = RL.r, uRL = RL.u, vRL = RL.v, aRL = RL.a, bRL = RL.b, cRL = RL.c, pRL = RL.p;
  const T rRR = RR.r, uRR = RR.u, vRR = RR.v, aRR = RR.a, bRR = RR.b, cRR = RR.c, pRR = RR.p;
  const T SL = fmin4(uLL, uLR, uRL, uRR) - fmax4(fast_x(g, LL), fast_x(g, LR), fast_x(g, RL), fast_x(g, RR));
  const T SR = fmax4(uLL, uLR, uRL, uRR) + fmax4(fast_x(g, LL), fast_x(g, LR), fast_x(g, RL), fast_x(g, RR));
  const T SB = fmin4(vLL, vLR, vRL, vRR) - fmax4(fast_y(g, LL), fast_y(g, LR), fast_y(g, RL), fast_y(g, RR));
  const T ST = fmax4(vLL, vLR, vRL, vRR) + fmax4(fast_y(g, LL), fast_y(g, LR), fast_y(g, RL), fast_y(g, RR));
  const T PtotLL = pLL + T(0.5) * (aLL * aLL + bLL * bLL + cLL * cLL);
  const T PtotLR = pLR + T(0.5) * (aLR * aLR + bLR * bLR + cLR * cLR);
  const T PtotRL = pRL + T(0.5) * (aRL * aRL + bRL * bRL + cRL * cRL);
  const T PtotRR = pRR + T(0.5) * (aRR * aRR + bRR * bRR + cRR * cRR);
  const T rcLLx = rLL * (uLL - SL), rcRLx = rRL * (SR - uRL), rcLRx = rLR * (uLR - SL), rcRRx = rRR * (SR - uRR);
  const T rcLLy = rLL * (vLL - SB), rcLRy = rLR * (ST - vLR), rcRLy = rRL * (vRL - SB), rcRRy = rRR * (ST - vRR);
  const T ustar = (rcLLx * uLL + rcLRx * uLR + rcRLx * uRL + rcRRx * uRR + (PtotLL - PtotRL + PtotLR - PtotRR)) /
                  (rcLLx + rcLRx + rcRLx + rcRRx);
  const T vstar = (rcLLy * vLL + rcLRy * vLR + rcRLy * vRL + rcRRy * vRR + (PtotLL - PtotLR + PtotRL - PtotRR)) /
                  (rcLLy + rcLRy + rcRLy + rcRRy);
  const T rstarLLx = rLL * (SL - uLL) / (SL - ustar);
  const T BstarLL = bLL * (SL - uLL) / (SL - ustar);
  const T rstarLLy = rLL * (SB - vLL) / (SB - vstar);
  const T AstarLL = aLL * (SB - vLL) / (SB - vstar);
  const T rstarLL = rLL * (SL - uLL) / (SL - ustar) * (SB - vLL) / (SB - vstar);
  const T EstarLLx = ustar * BstarLL - vLL * aLL;
  const T EstarLLy = uLL * bLL - vstar * AstarLL;
  const T EstarLL = ustar * BstarLL - vstar * AstarLL;
  const T rstarLRx = rLR * (SL - uLR) / (SL - ustar);
  const T BstarLR = bLR * (SL - uLR) / (SL - ustar);
  const T rstarLRy = rLR * (ST - vLR) / (ST - vstar);
  const T AstarLR = aLR * (ST - vLR) / (ST - vstar);
  const T rstarLR = rLR * (SL - uLR) / (SL - ustar) * (ST - vLR) / (ST - vstar);
  const T EstarLRx = ustar * BstarLR - vLR * aLR;
  const T EstarLRy = uLR * bLR - vstar * AstarLR;
  const T EstarLR = ustar * BstarLR - vstar * AstarLR;
  const T rstarRLx = rRL * (SR - uRL) / (SR - ustar);
  const T BstarRL = bRL * (SR - uRL) / (SR - ustar);
  const T rstarRLy = rRL * (SB - vRL) / (SB - vstar);
  const T AstarRL = aRL * (SB - vRL) / (SB - vstar);
  const T rstarRL = rRL * (SR - uRL) / (SR - ustar) * (SB - vRL) / (SB - vstar);
  const T EstarRLx = ustar * BstarRL - vRL * aRL;
  const T EstarRLy = uRL * bRL - vstar * AstarRL;
  const T EstarRL = ustar * BstarRL - vstar * AstarRL;
  const T rstarRRx = rRR * (SR - uRR) / (SR - ustar);
  const T BstarRR = bRR * (SR - uRR) / (SR - ustar);
  const T rstarRRy = rRR * (ST - vRR) / (ST - vstar);
  const T AstarRR = aRR * (ST - vRR) / (ST - vstar);
  const T rstarRR = rRR * (SR - uRR) / (SR - ustar) * (ST - vRR) / (ST - vstar);
  const T EstarRRx = ustar * BstarRR - vRR * aRR;
  const T EstarRRy = uRR * bRR - vstar * AstarRR;
  const T EstarRR = ustar * BstarRR - vstar * AstarRR;
  const T calfvenL = fmax5(eabs(aLR) / esqrt(rstarLRx), eabs(AstarLR) / esqrt(rstarLR), eabs(aLL) / esqrt(rstarLLx),
                           eabs(AstarLL) / esqrt(rstarLL), g.smallc);
  const T calfvenR = fmax5(eabs(aRR) / esqrt(rstarRRx), eabs(AstarRR) / esqrt(rstarRR), eabs(aRL) / esqrt(rstarRLx),
                           eabs(AstarRL) / esqrt(rstarRL), g.smallc);
  const T calfvenB = fmax5(eabs(bLL) / esqrt(rstarLLy), eabs(BstarLL) / esqrt(rstarLL), eabs(bRL) / esqrt(rstarRLy),
                           eabs(BstarRL) / esqrt(rstarRL), g.smallc);
  const T calfvenT = fmax5(eabs(bLR) / esqrt(rstarLRy), eabs(BstarLR) / esqrt(rstarLR), eabs(bRR) / esqrt(rstarRRy),
                           eabs(BstarRR) / esqrt(rstarRR), g.smallc);
  const T SAL = emin(ustar - calfvenL, T(0));
  const T SAR = emax(ustar + calfvenR, T(0));
  const T SAB = emin(vstar - calfvenB, T(0));
  const T SAT = emax(vstar + calfvenT, T(0));
  const T AstarT = (SAR * AstarRR - SAL * AstarLR) / (SAR - SAL);
  const T AstarB = (SAR * AstarRL - SAL * AstarLL) / (SAR - SAL);
  const T BstarR = (SAT * BstarRR - SAB * BstarRL) / (SAT - SAB);
  const T BstarL = (SAT * BstarLR - SAB * BstarLL) / (SAT - SAB);
  // branch selectors as 0/1 integers; the reference blends all five candidates with them,
  // so a non-finite unused candidate still reaches the result (0 * inf = NaN) -- kept.
  const int SB_pos = (int)(T(1) + ecopysign(T(1), SB)) / 2, SB_neg = 1 - SB_pos;
  const int ST_pos = (int)(T(1) + ecopysign(T(1), ST)) / 2, ST_neg = 1 - ST_pos;
  const int SL_pos = (int)(T(1) + ecopysign(T(1), SL)) / 2, SL_neg = 1 - SL_pos;
  const int SR_pos = (int)(T(1) + ecopysign(T(1), SR)) / 2, SR_neg = 1 - SR_pos;
  T E = T(0), tmpE;
  tmpE = (SAL * SAB * EstarRR - SAL * SAT * EstarRL - SAR * SAB * EstarLR + SAR * SAT * EstarLL) / (SAR - SAL) /
             (SAT - SAB) -
         SAT * SAB / (SAT - SAB) * (AstarT - AstarB) + SAR * SAL / (SAR - SAL) * (BstarR - BstarL);
  E += T(SB_neg * ST_pos * SL_neg * SR_pos) * tmpE;
  tmpE = (SAR * EstarLLx - SAL * EstarRLx + SAR * SAL * (bRL - bLL)) / (SAR - SAL);
  tmpE = T(SL_pos) * ELL + T(SL_neg * SR_neg) * ERL + T(SL_neg * SR_pos) * tmpE;
  E += T(SB_pos) * tmpE;
  tmpE = (SAR * EstarLRx - SAL * EstarRRx + SAR * SAL * (bRR - bLR)) / (SAR - SAL);
  tmpE = T(SL_pos) * ELR + T(SL_neg * SR_neg) * ERR + T(SL_neg * SR_pos) * tmpE;
  E += T(SB_neg * ST_neg) * tmpE;
  tmpE = (SAT * EstarLLy - SAB * EstarLRy - SAT * SAB * (aLR - aLL)) / (SAT - SAB);
  E += T(SB_neg * ST_pos * SL_pos) * tmpE;
  tmpE = (SAT * EstarRLy - SAB * EstarRRy - SAT * SAB * (aRR - aRL)) / (SAT - SAB);
  E += T(SB_neg * ST_pos * SL_neg * SR_neg) * tmpE;
  return E;
}

// Trace record of one cell: the centred state after the half-step source terms and the
// half slopes of trace_unsplit_mhd_2d (src/muscl/MHDBaseFunctor2D.h:774-1109).
enum Rec {
  R_R = 0, R_U, R_V, R_W, R_P, R_A, R_B, R_C,
  R_DRX, R_DUX, R_DVX, R_DWX, R_DPX, R_DBX, R_DCX,
  R_DRY, R_DUY, R_DVY, R_DWY, R_DPY, R_DAY, R_DCY,
  NREC
};

// ------------------------------------------------------------------------------------
// The step kernel. TX x TY cells per block of NT threads; (TX+2)*(TY+2) must equal NT
// (one trace record per thread).
// ------------------------------------------------------------------------------------
template <int TX, int TY>
struct Tile {
  static constexpr int QX = TX + 4, QY = TY + 4, QN = QX * QY;      // primitives / raw faces
  static constexpr int EX = TX + 3, EY = TY + 3, EN = EX * EY;      // edges -1..TX+1
  static constexpr int AXX = TX + 3, AXY = TY + 2, AXN = AXX * AXY; // x faces -1..TX+1, rows -1..TY
  static constexpr int BYX = TX + 2, BYY = TY + 3, BYN = BYX * BYY; // y faces: cols -1..TX, -1..TY+1
  static constexpr int RX = TX + 2, RN = (TX + 2) * (TY + 2);        // trace cells -1..TX
  static constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1), NE = (TX + 1) * (TY + 1);
  // shared memory regions (in reals)
  static constexpr int R1 = (NV * QN > 6 * NFX + 6 * NFY + NE) ? NV * QN : 6 * NFX + 6 * NFY + NE;
  static constexpr int R2 = (NREC * RN > EN) ? NREC * RN : EN;
  static constexpr int R3 = 2 * QN + 2 * AXN + 2 * BYN;
  static constexpr int TOTAL = R1 + R2 + R3;
  // compact layout: region CX holds, in turn, the primitives + edge fields + raw face fields
  // (phases A-B1), the trace records (C) and the fluxes and EMFs (D); region CY the half-step
  // face fields and transverse slopes (B0-C). The records and the Riemann results are
  // produced in registers and written after a barrier, so no two of them share the LDS.
  static constexpr int XA = NV * QN + EN + 2 * QN, XB = NREC * RN, XC = 6 * NFX + 6 * NFY + NE;
  static constexpr int CX = XA > XB ? (XA > XC ? XA : XC) : (XB > XC ? XB : XC);
  static constexpr int CTOTAL = CX + 2 * AXN + 2 * BYN;
};

template <class T>
__device__ __forceinline__ void floor_dp(const Gas<T>& g, T& r, T& p)
{
  r = emax(g.smallr, r);
  p = emax(g.smallp * r, p);
}

// WPE: waves per SIMD the registers are sized for. CMP: the compact LDS layout of Tile
// (two more barriers per step): the FP32 30x14 tile (512 threads) takes 53.7 KB with it,
// three blocks per CU, against 82 KB and two without. FP64 14x14 (256 threads) keeps the
// plain layout (81 KB, two blocks per CU): at three blocks the two extra barriers and the
// held results cost more than the third wave gains.
template <class T, int TX, int TY, int NT, int RS, int ST, bool CMP, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, WPE))) void mhd2d_step(const T* __restrict__ Uin, T* __restrict__ Uout, Geom<T> geo,
                                                 Gas<T> gas, StepCtl ctl)
{
  using L = Tile<TX, TY>;
  static_assert(L::RN == NT, "one trace record per thread");
  static_assert(ST == 1 || ST == 2, "MHD trace supports slope_type 1 and 2");
  constexpr int QX = L::QX, QN = L::QN;
  __shared__ T smem[CMP ? L::CTOTAL : L::TOTAL];
  __shared__ double sred[NT / 64];
  T* const sQ = smem;                                             // [NV][QN] primitives
  T* const sF = smem;                                             // after the trace: fluxes and EMFs
  T* const sEz = smem + (CMP ? NV * QN : L::R1);                  // [EN] edge fields u*B - v*A
  T* const sRec = CMP ? smem : smem + L::R1;                      // after the face evolution: [NREC][RN]
  T* const sBf = smem + (CMP ? NV * QN + L::EN : L::R1 + L::R2);  // [2][QN] raw face fields (IA, IB)
  T* const sAev = CMP ? smem + L::CX : sBf + 2 * QN;              // [AXN] half-step x-face field
  T* const sDAy = sAev + L::AXN;         // [AXN] transverse slope dA/dy at x faces (halved)
  T* const sBev = sDAy + L::AXN;         // [BYN] half-step y-face field
  T* const sDBx = sBev + L::BYN;         // [BYN] transverse slope dB/dx at y faces (halved)

  const int tid = threadIdx.x;
  const bool first_block = blockIdx.x == 0 && blockIdx.y == 0;
  double dt;
  TimeSlot cur;
  if (!step_begin(ctl, first_block, dt, cur))
    return;
  const T dtdx = T(dt / geo.dx), dtdy = T(dt / geo.dy);
  const int nx = geo.nx, ny = geo.ny;
  const long long vs = geo.vstride;
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;

  // A: primitives and raw face fields on cells -2..TX+1 (ghosts through the boundary map;
  // a reflecting x pass negates IU and IA, a reflecting y pass IV and IB)
  for (int w = tid; w < QN; w += NT) {
    const int lx = w % QX - 2, ly = w / QX - 2;
    bool fx, fy, fx1, fy1;
    const int si = bc_map(i0 + lx, nx, geo.bc[0], geo.bc[1], fx);
    const int sj = bc_map(j0 + ly, ny, geo.bc[2], geo.bc[3], fy);
    const int si1 = bc_map(i0 + lx + 1, nx, geo.bc[0], geo.bc[1], fx1);
    const int sj1 = bc_map(j0 + ly + 1, ny, geo.bc[2], geo.bc[3], fy1);
    const long long o = (long long)sj * nx + si;
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
      u[v] = Uin[v * vs + o];
    if (fx) {
      u[IU] = u[IU] * T(-1);
      u[IA] = u[IA] * T(-1);
    }
    if (fy) {
      u[IV] = u[IV] * T(-1);
      u[IB] = u[IB] * T(-1);
    }
    T mx = Uin[IA * vs + (long long)sj * nx + si1];
    if (fx1)
      mx = mx * T(-1);
    T my = Uin[IB * vs + (long long)sj1 * nx + si];
    if (fy1)
      my = my * T(-1);
    T q[NV];
    constoprim<T>(gas, u, mx, my, T(0), q);
#pragma unroll
    for (int v = 0; v < NV; v++)
      sQ[v * QN + w] = q[v];
    sBf[w] = u[IA];
    sBf[QN + w] = u[IB];
  }
  __syncthreads();

  // B0: edge fields (edges -1..TX+1), transverse face-field slopes
  {
    constexpr int N0 = L::EN, N1 = L::EN + L::AXN, N2 = N1 + L::BYN;
    for (int w = tid; w < N2; w += NT) {
      if (w < N0) {
        const int cx = w % L::EX - 1, cy = w / L::EX - 1;
        const int q00 = (cy + 1) * QX + (cx + 1); // cell (cx-1, cy-1)
        const int q01 = q00 + QX;                 // (cx-1, cy)
        const int q10 = q00 + 1;                  // (cx, cy-1)
        const int q11 = q01 + 1;                  // (cx, cy)
        const T u = T(0.25) * (sQ[IU * QN + q00] + sQ[IU * QN + q01] + sQ[IU * QN + q10] + sQ[IU * QN + q11]);
        const T v = T(0.25) * (sQ[IV * QN + q00] + sQ[IV * QN + q01] + sQ[IV * QN + q10] + sQ[IV * QN + q11]);
        const T A = T(0.5) * (sBf[q10] + sBf[q11]);
        const T B = T(0.5) * (sBf[QN + q01] + sBf[QN + q11]);
        sEz[w] = u * B - v * A;
      } else if (w < N1) {
        const int h = w - N0;
        const int fxl = h % L::AXX - 1, y = h / L::AXX - 1;
        const int c = (y + 2) * QX + (fxl + 2);
        sDAy[h] = T(0.5) * slope1<T, ST>(sBf[c], sBf[c + QX], sBf[c - QX]);
      } else {
        const int h = w - N1;
        const int x = h % L::BYX - 1, fyl = h / L::BYX - 1;
        const int c = (fyl + 2) * QX + (x + 2);
        sDBx[h] = T(0.5) * slope1<T, ST>(sBf[QN + c], sBf[QN + c + 1], sBf[QN + c - 1]);
      }
    }
  }
  __syncthreads();

  // B1: half-step face fields: AL += (ELR - ELL)*0.5*dtdy, BL += -(ERL - ELL)*0.5*dtdx
  for (int w = tid; w < L::AXN + L::BYN; w += NT) {
    if (w < L::AXN) {
      const int fxl = w % L::AXX - 1, y = w / L::AXX - 1;
      const int e0 = (y + 1) * L::EX + (fxl + 1); // edge (fxl, y)
      const T s = (sEz[e0 + L::EX] - sEz[e0]) * T(0.5) * dtdy;
      sAev[w] = sBf[(y + 2) * QX + (fxl + 2)] + s;
    } else {
      const int h = w - L::AXN;
      const int x = h % L::BYX - 1, fyl = h / L::BYX - 1;
      const int e0 = (fyl + 1) * L::EX + (x + 1); // edge (x, fyl)
      const T s = -(sEz[e0 + 1] - sEz[e0]) * T(0.5) * dtdx;
      sBev[h] = sBf[QN + (fyl + 2) * QX + (x + 2)] + s;
    }
  }
  __syncthreads();

  // B: trace record of cell (tid%RX - 1, tid/RX - 1)
  {
    const int lx = tid % L::RX - 1, ly = tid / L::RX - 1;
    const int c = (ly + 2) * QX + (lx + 2);
    T q[NV], dx_[NV], dy_[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      q[v] = sQ[v * QN + c];
      dx_[v] = slope1<T, ST>(q[v], sQ[v * QN + c + 1], sQ[v * QN + c - 1]);
      dy_[v] = slope1<T, ST>(q[v], sQ[v * QN + c + QX], sQ[v * QN + c - QX]);
    }
    const T r = q[ID], p = q[IP], u = q[IU], v = q[IV], w = q[IW], A = q[IA], B = q[IB], C = q[IC];
    const T AL = sBf[c], AR = sBf[c + 1], BL = sBf[QN + c], BR = sBf[QN + c + QX];
    T drx = dx_[ID]; drx *= T(0.5);
    T dpx = dx_[IP]; dpx *= T(0.5);
    T dux = dx_[IU]; dux *= T(0.5);
    T dvx = dx_[IV]; dvx *= T(0.5);
    T dwx = dx_[IW]; dwx *= T(0.5);
    T dCx = dx_[IC]; dCx *= T(0.5);
    T dBx = dx_[IB]; dBx *= T(0.5);
    T dry = dy_[ID]; dry *= T(0.5);
    T dpy = dy_[IP]; dpy *= T(0.5);
    T duy = dy_[IU]; duy *= T(0.5);
    T dvy = dy_[IV]; dvy *= T(0.5);
    T dwy = dy_[IW]; dwy *= T(0.5);
    T dCy = dy_[IC]; dCy *= T(0.5);
    T dAy = dy_[IA]; dAy *= T(0.5);
    const T dAx = T(0.5) * (AR - AL);
    const T dBy = T(0.5) * (BR - BL);
    const T gamma = gas.gamma0;
    const T sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy;
    const T su0 = (-u * dux - dpx / r - B * dBx / r - C * dCx / r) * dtdx + (-v * duy + B * dAy / r) * dtdy;
    const T sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - dpy / r - A * dAy / r - C * dCy / r) * dtdy;
    const T sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy;
    const T sp0 = (-u * dpx - dux * gamma * p) * dtdx + (-v * dpy - dvy * gamma * p) * dtdy;
    const T sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy;
    const T sB0 = (-u * dBx - B * dux + v * dAx + A * dvx) * dtdx;
    const T sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (-v * dCy - C * dvy + w * dBy + B * dwy) * dtdy;
    T rv[NREC];
    rv[R_R] = r + sr0;
    rv[R_U] = u + su0;
    rv[R_V] = v + sv0;
    rv[R_W] = w + sw0;
    rv[R_P] = p + sp0;
    rv[R_A] = A + sA0;
    rv[R_B] = B + sB0;
    rv[R_C] = C + sC0;
    rv[R_DRX] = drx;
    rv[R_DUX] = dux;
    rv[R_DVX] = dvx;
    rv[R_DWX] = dwx;
    rv[R_DPX] = dpx;
    rv[R_DBX] = dBx;
    rv[R_DCX] = dCx;
    rv[R_DRY] = dry;
    rv[R_DUY] = duy;
    rv[R_DVY] = dvy;
    rv[R_DWY] = dwy;
    rv[R_DPY] = dpy;
    rv[R_DAY] = dAy;
    rv[R_DCY] = dCy;
    if constexpr (CMP)
      __syncthreads(); // every read of the primitives and raw face fields is done
#pragma unroll
    for (int k = 0; k < NREC; k++)
      sRec[k * NT + tid] = rv[k];
  }
  __syncthreads();

  // C: Riemann problems (x faces then y faces, one list) and edge EMFs, in whole-wave tasks
  T* const sFx = sF;                       // [6][NFX] raw x fluxes (ID IP IU IV IW IC)
  T* const sFy = sF + 6 * L::NFX;          // [6][NFY] y fluxes back in x/y order
  T* const sE = sF + 6 * (L::NFX + L::NFY); // [NE]
  {
    constexpr int NF = L::NFX + L::NFY;
    constexpr int TF = (NF + 63) / 64, TE = (L::NE + 63) / 64;
    const int wid = tid >> 6, lane = tid & 63;
    auto rec = [&](int cell, int k) -> T { return sRec[k * NT + cell]; };
    // flux of face k (x faces first) as its solve returns it: ID IP IU IV IW IC, where a y
    // face's IU is the y momentum flux (rotated solve)
    auto face = [&](int k, T* out) {
      T ql[NV], qr[NV], f[NV];
      if (k < L::NFX) {
        const int fxl = k % (TX + 1), y = k / (TX + 1);
        const int cl = (y + 1) * L::RX + fxl, cr = cl + 1; // cells (fxl-1, y), (fxl, y)
        const T a = sAev[(y + 1) * L::AXX + fxl + 1];
        ql[ID] = rec(cl, R_R) + rec(cl, R_DRX);
        ql[IU] = rec(cl, R_U) + rec(cl, R_DUX);
        ql[IV] = rec(cl, R_V) + rec(cl, R_DVX);
        ql[IW] = rec(cl, R_W) + rec(cl, R_DWX);
        ql[IP] = rec(cl, R_P) + rec(cl, R_DPX);
        ql[IA] = a;
        ql[IB] = rec(cl, R_B) + rec(cl, R_DBX);
        ql[IC] = rec(cl, R_C) + rec(cl, R_DCX);
        floor_dp(gas, ql[ID], ql[IP]);
        qr[ID] = rec(cr, R_R) - rec(cr, R_DRX);
        qr[IU] = rec(cr, R_U) - rec(cr, R_DUX);
        qr[IV] = rec(cr, R_V) - rec(cr, R_DVX);
        qr[IW] = rec(cr, R_W) - rec(cr, R_DWX);
        qr[IP] = rec(cr, R_P) - rec(cr, R_DPX);
        qr[IA] = a;
        qr[IB] = rec(cr, R_B) - rec(cr, R_DBX);
        qr[IC] = rec(cr, R_C) - rec(cr, R_DCX);
        floor_dp(gas, qr[ID], qr[IP]);
      } else {
        const int h = k - L::NFX;
        const int x = h % TX, fyl = h / TX;
        const int cb = fyl * L::RX + x + 1, ct = cb + L::RX; // cells (x, fyl-1), (x, fyl)
        const T b = sBev[(fyl + 1) * L::BYX + x + 1];
        // qm_y of the lower cell and qp_y of the upper one, then IU<->IV, IA<->IB
        T r_ = rec(cb, R_R) + rec(cb, R_DRY);
        T p_ = rec(cb, R_P) + rec(cb, R_DPY);
        floor_dp(gas, r_, p_);
        ql[ID] = r_;
        ql[IP] = p_;
        ql[IV] = rec(cb, R_U) + rec(cb, R_DUY);
        ql[IU] = rec(cb, R_V) + rec(cb, R_DVY);
        ql[IW] = rec(cb, R_W) + rec(cb, R_DWY);
        ql[IB] = rec(cb, R_A) + rec(cb, R_DAY);
        ql[IA] = b;
        ql[IC] = rec(cb, R_C) + rec(cb, R_DCY);
        r_ = rec(ct, R_R) - rec(ct, R_DRY);
        p_ = rec(ct, R_P) - rec(ct, R_DPY);
        floor_dp(gas, r_, p_);
        qr[ID] = r_;
        qr[IP] = p_;
        qr[IV] = rec(ct, R_U) - rec(ct, R_DUY);
        qr[IU] = rec(ct, R_V) - rec(ct, R_DVY);
        qr[IW] = rec(ct, R_W) - rec(ct, R_DWY);
        qr[IB] = rec(ct, R_A) - rec(ct, R_DAY);
        qr[IA] = b;
        qr[IC] = rec(ct, R_C) - rec(ct, R_DCY);
      }
      riemann<T, RS>(gas, ql, qr, f);
      out[0] = f[ID];
      out[1] = f[IP];
      out[2] = f[IU];
      out[3] = f[IV];
      out[4] = f[IW];
      out[5] = f[IC];
    };
    // stores a face's fluxes; a y face's are turned back to x/y order
    auto store_face = [&](int k, const T* o) {
      if (k < L::NFX) {
#pragma unroll
        for (int n = 0; n < 6; n++)
          sFx[n * L::NFX + k] = o[n];
      } else {
        const int h = k - L::NFX;
        sFy[0 * L::NFY + h] = o[0];
        sFy[1 * L::NFY + h] = o[1];
        sFy[2 * L::NFY + h] = o[3];
        sFy[3 * L::NFY + h] = o[2];
        sFy[4 * L::NFY + h] = o[4];
        sFy[5 * L::NFY + h] = o[5];
      }
    };
    // EMF of edge k
    auto edge = [&](int k) -> T {
      const int cx = k % (TX + 1), cy = k / (TX + 1);
      const int cLB = (cy + 1) * L::RX + cx + 1; // cell (cx, cy)
      const int cRB = cLB - 1;                   // (cx-1, cy)
      const int cLT = cLB - L::RX;               // (cx, cy-1)
      const int cRT = cLT - 1;                   // (cx-1, cy-1)
      const int aT = cy * L::AXX + cx + 1;       // x face cx, row cy-1
      const int aB = aT + L::AXX;                // x face cx, row cy
      const int bL = (cy + 1) * L::BYX + cx;     // y face cy, column cx-1
      const int bR = bL + 1;                     // y face cy, column cx
      Edge<T> RT, RB, LT, LB;
      RT.r = rec(cRT, R_R) + (rec(cRT, R_DRX) + rec(cRT, R_DRY));
      RT.u = rec(cRT, R_U) + (rec(cRT, R_DUX) + rec(cRT, R_DUY));
      RT.v = rec(cRT, R_V) + (rec(cRT, R_DVX) + rec(cRT, R_DVY));
      RT.p = rec(cRT, R_P) + (rec(cRT, R_DPX) + rec(cRT, R_DPY));
      RT.a = sAev[aT] + sDAy[aT];
      RT.b = sBev[bL] + sDBx[bL];
      RT.c = rec(cRT, R_C) + (rec(cRT, R_DCX) + rec(cRT, R_DCY));
      floor_dp(gas, RT.r, RT.p);
      RB.r = rec(cRB, R_R) + (rec(cRB, R_DRX) - rec(cRB, R_DRY));
      RB.u = rec(cRB, R_U) + (rec(cRB, R_DUX) - rec(cRB, R_DUY));
      RB.v = rec(cRB, R_V) + (rec(cRB, R_DVX) - rec(cRB, R_DVY));
      RB.p = rec(cRB, R_P) + (rec(cRB, R_DPX) - rec(cRB, R_DPY));
      RB.a = sAev[aB] + (-sDAy[aB]);
      RB.b = sBev[bL] + sDBx[bL];
      RB.c = rec(cRB, R_C) + (rec(cRB, R_DCX) - rec(cRB, R_DCY));
      floor_dp(gas, RB.r, RB.p);
      LT.r = rec(cLT, R_R) + (-rec(cLT, R_DRX) + rec(cLT, R_DRY));
      LT.u = rec(cLT, R_U) + (-rec(cLT, R_DUX) + rec(cLT, R_DUY));
      LT.v = rec(cLT, R_V) + (-rec(cLT, R_DVX) + rec(cLT, R_DVY));
      LT.p = rec(cLT, R_P) + (-rec(cLT, R_DPX) + rec(cLT, R_DPY));
      LT.a = sAev[aT] + sDAy[aT];
      LT.b = sBev[bR] + (-sDBx[bR]);
      LT.c = rec(cLT, R_C) + (-rec(cLT, R_DCX) + rec(cLT, R_DCY));
      floor_dp(gas, LT.r, LT.p);
      LB.r = rec(cLB, R_R) + (-rec(cLB, R_DRX) - rec(cLB, R_DRY));
      LB.u = rec(cLB, R_U) + (-rec(cLB, R_DUX) - rec(cLB, R_DUY));
      LB.v = rec(cLB, R_V) + (-rec(cLB, R_DVX) - rec(cLB, R_DVY));
      LB.p = rec(cLB, R_P) + (-rec(cLB, R_DPX) - rec(cLB, R_DPY));
      LB.a = sAev[aB] + (-sDAy[aB]);
      LB.b = sBev[bR] + (-sDBx[bR]);
      LB.c = rec(cLB, R_C) + (-rec(cLB, R_DCX) - rec(cLB, R_DCY));
      floor_dp(gas, LB.r, LB.p);
      return emf_z<T>(gas, RT, RB, LT, LB);
    };
    if constexpr (!CMP) {
      // the same solves written out in place (the FP64 step built from the lambdas above
      // measured 1.2 % slower: same instructions, scheduled differently)
      for (int task = wid; task < TF + TE; task += NT / 64) {
        if (task < TF) {
          const int k = task * 64 + lane;
          if (k < NF) {
            T ql[NV], qr[NV], f[NV];
            if (k < L::NFX) {
              const int fxl = k % (TX + 1), y = k / (TX + 1);
              const int cl = (y + 1) * L::RX + fxl, cr = cl + 1; // cells (fxl-1, y), (fxl, y)
              const T a = sAev[(y + 1) * L::AXX + fxl + 1];
              ql[ID] = rec(cl, R_R) + rec(cl, R_DRX);
              ql[IU] = rec(cl, R_U) + rec(cl, R_DUX);
              ql[IV] = rec(cl, R_V) + rec(cl, R_DVX);
              ql[IW] = rec(cl, R_W) + rec(cl, R_DWX);
              ql[IP] = rec(cl, R_P) + rec(cl, R_DPX);
              ql[IA] = a;
              ql[IB] = rec(cl, R_B) + rec(cl, R_DBX);
              ql[IC] = rec(cl, R_C) + rec(cl, R_DCX);
              floor_dp(gas, ql[ID], ql[IP]);
              qr[ID] = rec(cr, R_R) - rec(cr, R_DRX);
              qr[IU] = rec(cr, R_U) - rec(cr, R_DUX);
              qr[IV] = rec(cr, R_V) - rec(cr, R_DVX);
              qr[IW] = rec(cr, R_W) - rec(cr, R_DWX);
              qr[IP] = rec(cr, R_P) - rec(cr, R_DPX);
              qr[IA] = a;
              qr[IB] = rec(cr, R_B) - rec(cr, R_DBX);
              qr[IC] = rec(cr, R_C) - rec(cr, R_DCX);
              floor_dp(gas, qr[ID], qr[IP]);
            } else {
              const int h = k - L::NFX;
              const int x = h % TX, fyl = h / TX;
              const int cb = fyl * L::RX + x + 1, ct = cb + L::RX; // cells (x, fyl-1), (x, fyl)
              const T b = sBev[(fyl + 1) * L::BYX + x + 1];
              // qm_y of the lower cell and qp_y of the upper one, then IU<->IV, IA<->IB
              T r_ = rec(cb, R_R) + rec(cb, R_DRY);
              T p_ = rec(cb, R_P) + rec(cb, R_DPY);
              floor_dp(gas, r_, p_);
              ql[ID] = r_;
              ql[IP] = p_;
              ql[IV] = rec(cb, R_U) + rec(cb, R_DUY);
              ql[IU] = rec(cb, R_V) + rec(cb, R_DVY);
              ql[IW] = rec(cb, R_W) + rec(cb, R_DWY);
              ql[IB] = rec(cb, R_A) + rec(cb, R_DAY);
              ql[IA] = b;
              ql[IC] = rec(cb, R_C) + rec(cb, R_DCY);
              r_ = rec(ct, R_R) - rec(ct, R_DRY);
              p_ = rec(ct, R_P) - rec(ct, R_DPY);
              floor_dp(gas, r_, p_);
              qr[ID] = r_;
              qr[IP] = p_;
              qr[IV] = rec(ct, R_U) - rec(ct, R_DUY);
              qr[IU] = rec(ct, R_V) - rec(ct, R_DVY);
              qr[IW] = rec(ct, R_W) - rec(ct, R_DWY);
              qr[IB] = rec(ct, R_A) - rec(ct, R_DAY);
              qr[IA] = b;
              qr[IC] = rec(ct, R_C) - rec(ct, R_DCY);
            }
            riemann<T, RS>(gas, ql, qr, f);
            if (k < L::NFX) {
              sFx[0 * L::NFX + k] = f[ID];
              sFx[1 * L::NFX + k] = f[IP];
              sFx[2 * L::NFX + k] = f[IU];
              sFx[3 * L::NFX + k] = f[IV];
              sFx[4 * L::NFX + k] = f[IW];
              sFx[5 * L::NFX + k] = f[IC];
            } else {
              const int h = k - L::NFX;
              sFy[0 * L::NFY + h] = f[ID];
              sFy[1 * L::NFY + h] = f[IP];
              sFy[2 * L::NFY + h] = f[IV];
              sFy[3 * L::NFY + h] = f[IU];
              sFy[4 * L::NFY + h] = f[IW];
              sFy[5 * L::NFY + h] = f[IC];
            }
          }
        } else {
          const int k = (task - TF) * 64 + lane;
          if (k < L::NE) {
            const int cx = k % (TX + 1), cy = k / (TX + 1);
            const int cLB = (cy + 1) * L::RX + cx + 1; // cell (cx, cy)
            const int cRB = cLB - 1;                   // (cx-1, cy)
            const int cLT = cLB - L::RX;               // (cx, cy-1)
            const int cRT = cLT - 1;                   // (cx-1, cy-1)
            const int aT = cy * L::AXX + cx + 1;       // x face cx, row cy-1
            const int aB = aT + L::AXX;                // x face cx, row cy
            const int bL = (cy + 1) * L::BYX + cx;     // y face cy, column cx-1
            const int bR = bL + 1;                     // y face cy, column cx
            Edge<T> RT, RB, LT, LB;
            RT.r = rec(cRT, R_R) + (rec(cRT, R_DRX) + rec(cRT, R_DRY));
            RT.u = rec(cRT, R_U) + (rec(cRT, R_DUX) + rec(cRT, R_DUY));
            RT.v = rec(cRT, R_V) + (rec(cRT, R_DVX) + rec(cRT, R_DVY));
            RT.p = rec(cRT, R_P) + (rec(cRT, R_DPX) + rec(cRT, R_DPY));
            RT.a = sAev[aT] + sDAy[aT];
            RT.b = sBev[bL] + sDBx[bL];
            RT.c = rec(cRT, R_C) + (rec(cRT, R_DCX) + rec(cRT, R_DCY));
            floor_dp(gas, RT.r, RT.p);
            RB.r = rec(cRB, R_R) + (rec(cRB, R_DRX) - rec(cRB, R_DRY));
            RB.u = rec(cRB, R_U) + (rec(cRB, R_DUX) - rec(cRB, R_DUY));
            RB.v = rec(cRB, R_V) + (rec(cRB, R_DVX) - rec(cRB, R_DVY));
            RB.p = rec(cRB, R_P) + (rec(cRB, R_DPX) - rec(cRB, R_DPY));
            RB.a = sAev[aB] + (-sDAy[aB]);
            RB.b = sBev[bL] + sDBx[bL];
            RB.c = rec(cRB, R_C) + (rec(cRB, R_DCX) - rec(cRB, R_DCY));
            floor_dp(gas, RB.r, RB.p);
            LT.r = rec(cLT, R_R) + (-rec(cLT, R_DRX) + rec(cLT, R_DRY));
            LT.u = rec(cLT, R_U) + (-rec(cLT, R_DUX) + rec(cLT, R_DUY));
            LT.v = rec(cLT, R_V) + (-rec(cLT, R_DVX) + rec(cLT, R_DVY));
            LT.p = rec(cLT, R_P) + (-rec(cLT, R_DPX) + rec(cLT, R_DPY));
            LT.a = sAev[aT] + sDAy[aT];
            LT.b = sBev[bR] + (-sDBx[bR]);
            LT.c = rec(cLT, R_C) + (-rec(cLT, R_DCX) + rec(cLT, R_DCY));
            floor_dp(gas, LT.r, LT.p);
            LB.r = rec(cLB, R_R) + (-rec(cLB, R_DRX) - rec(cLB, R_DRY));
            LB.u = rec(cLB, R_U) + (-rec(cLB, R_DUX) - rec(cLB, R_DUY));
            LB.v = rec(cLB, R_V) + (-rec(cLB, R_DVX) - rec(cLB, R_DVY));
            LB.p = rec(cLB, R_P) + (-rec(cLB, R_DPX) - rec(cLB, R_DPY));
            LB.a = sAev[aB] + (-sDAy[aB]);
            LB.b = sBev[bR] + (-sDBx[bR]);
            LB.c = rec(cLB, R_C) + (-rec(cLB, R_DCX) - rec(cLB, R_DCY));
            floor_dp(gas, LB.r, LB.p);
            sE[k] = emf_z<T>(gas, RT, RB, LT, LB);
          }
        }
      }
    } else {
      // the fluxes and EMFs go where the records are: every result is held until all
      // records have been read (EMF tasks first, so that at most the EMF and one face's
      // fluxes are held through a solve)
      constexpr int NW = NT / 64, MAXT = (TF + TE + NW - 1) / NW;
      T hold[MAXT][6];
#pragma unroll
      for (int j = 0; j < MAXT; j++) {
        const int task = wid + j * NW;
        if (task < TE) {
          const int k = task * 64 + lane;
          if (k < L::NE)
            hold[j][0] = edge(k);
        } else if (task < TE + TF) {
          const int k = (task - TE) * 64 + lane;
          if (k < NF)
            face(k, hold[j]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < MAXT; j++) {
        const int task = wid + j * NW;
        if (task < TE) {
          const int k = task * 64 + lane;
          if (k < L::NE)
            sE[k] = hold[j][0];
        } else if (task < TE + TF) {
          const int k = (task - TE) * 64 + lane;
          if (k < NF)
            store_face(k, hold[j]);
        }
      }
    }
  }
  __syncthreads();

  // D: flux update (UpdateFunctor2D_MHD), EMF update (UpdateEmfFunctor2D), then the CFL
  // term of the new state, whose primitives need the new face fields on the right and top
  double local = 0.0;
  for (int w = tid; w < TX * TY; w += NT) {
    const int lx = w % TX, ly = w / TX;
    const int gi = i0 + lx, gj = j0 + ly;
    if (gi < nx && gj < ny) {
      const long long o = (long long)gj * nx + gi;
      T u[NV];
#pragma unroll
      for (int v = 0; v < NV; v++)
        u[v] = Uin[v * vs + o];
      const int kx = ly * (TX + 1) + lx; // x face lx, row ly
      const int ky = ly * TX + lx;       // y face ly, column lx
      constexpr int vmap[6] = { ID, IP, IU, IV, IW, IC };
#pragma unroll
      for (int n = 0; n < 6; n++) {
        T a = u[vmap[n]];
        a += sFx[n * L::NFX + kx] * dtdx;
        a -= sFx[n * L::NFX + kx + 1] * dtdx;
        a += sFy[n * L::NFY + ky] * dtdy;
        a -= sFy[n * L::NFY + ky + TX] * dtdy;
        u[vmap[n]] = a;
      }
      const int e = ly * (TX + 1) + lx; // edge (lx, ly)
      const T E00 = sE[e], E01 = sE[e + TX + 1], E10 = sE[e + 1], E11 = sE[e + TX + 2];
      u[IA] += (E01 - E00) * dtdy;
      u[IB] -= (E10 - E00) * dtdx;
#pragma unroll
      for (int v = 0; v < NV; v++)
        Uout[v * vs + o] = u[v];
      // new IA on the right face and IB on the top face, as the next step's ghost fill
      // will see them
      T ar, bt;
      if (gi + 1 < nx) {
        ar = Uin[IA * vs + o + 1] + (E11 - E10) * dtdy;
      } else if (geo.bc[1] == BC_PERIODIC) {
        // edges on x = nx and x = 0 carry identical values on a periodic axis
        ar = Uin[IA * vs + (long long)gj * nx] + (E11 - E10) * dtdy;
      } else if (geo.bc[1] == BC_DIRICHLET) {
        ar = u[IA] * T(-1);
      } else {
        ar = u[IA];
      }
      if (gj + 1 < ny) {
        bt = Uin[IB * vs + o + nx] - (E11 - E01) * dtdx;
      } else if (geo.bc[3] == BC_PERIODIC) {
        bt = Uin[IB * vs + gi] - (E11 - E01) * dtdx;
      } else if (geo.bc[3] == BC_DIRICHLET) {
        bt = u[IB] * T(-1);
      } else {
        bt = u[IB];
      }
      const double s = mhd_cfl_term<T>(gas, u, ar, bt, geo);
      local = (s > local || !(s == s)) ? s : local;
    }
  }
  block_max_to<NT>(invdt_word(ctl, (ctl.slot + 1) % 3, blockIdx.x + blockIdx.y * gridDim.x), local, sred);
  if (first_block && tid < 64)
    step_end_bookkeeping(ctl, cur, dt);
}

// CFL reduction over the current MHD state (convertToPrimitives + ComputeDtFunctor2D_MHD):
// the primitives of a cell need IA of the right neighbour and IB of the top one, through
// the boundary map when they are ghosts.
template <class T, int NT>
__global__ __launch_bounds__(NT) void mhd2d_inv_dt_kernel(const T* __restrict__ U, Geom<T> geo, Gas<T> gas,
                                                           unsigned long long* dst, unsigned long long* spread)
{
  __shared__ double sred[NT / 64];
  const long long ncell = geo.vstride;
  const long long vs = geo.vstride;
  const int nx = geo.nx, ny = geo.ny;
  double local = 0.0;
  for (long long c = (long long)blockIdx.x * NT + threadIdx.x; c < ncell; c += (long long)gridDim.x * NT) {
    const int i = (int)(c % nx), j = (int)(c / nx);
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
      u[v] = U[v * vs + c];
    bool fx, fy;
    const int si = bc_map(i + 1, nx, geo.bc[0], geo.bc[1], fx);
    const int sj = bc_map(j + 1, ny, geo.bc[2], geo.bc[3], fy);
    T mx = U[IA * vs + (long long)j * nx + si];
    if (fx)
      mx = mx * T(-1);
    T my = U[IB * vs + (long long)sj * nx + i];
    if (fy)
      my = my * T(-1);
    const double s = mhd_cfl_term<T>(gas, u, mx, my, geo);
    local = (s > local || !(s == s)) ? s : local;
  }
  block_max_to<NT>(reduce_word(dst, spread), local, sred);
}

} // namespace mhd
} // namespace ek
