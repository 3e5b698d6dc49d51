// ek_cellmath.h -- per-cell arithmetic of the MUSCL-Hancock step (device + host inline).
//
// Every function keeps the reference's floating-point operation order exactly (no
// reassociation, no reciprocal-for-division), so that the FP64 build, compiled with
// -ffp-contract=off, reproduces the reference's CPU results bit for bit. Shared
// subexpressions are computed once only where they are literally the same IEEE
// expression in the reference (e.g. HALF*drx used by both faces of a cell).
// Paths cited are relative to the euler_kokkos source tree.
#pragma once

#include <hip/hip_runtime.h>

namespace ek {

enum Var { ID = 0, IP = 1, IU = 2, IV = 3, IW = 4, IA = 5, IB = 6, IC = 7 };
enum Riemann { RS_APPROX = 0, RS_LLF = 1, RS_HLL = 2, RS_HLLC = 3, RS_HLLD = 4 };
enum Bc { BC_UNDEFINED = 0, BC_DIRICHLET = 1, BC_NEUMANN = 2, BC_PERIODIC = 3 };

// Gas / scheme constants (HydroSettings, src/shared/HydroParams.h:32-69). smallc2 and entho
// are the literal products/quotients the reference recomputes at each use.
template <class T>
struct Gas {
  T gamma0, gamma6, gm1, smallr, smallc, smallp, smallpp, smallc2, entho;
};

template <class T>
__host__ __device__ inline T kHalf()
{
  return T(0.5);
}

__host__ __device__ inline double emax(double a, double b) { return fmax(a, b); }
__host__ __device__ inline float emax(float a, float b) { return fmaxf(a, b); }
__host__ __device__ inline double emin(double a, double b) { return fmin(a, b); }
__host__ __device__ inline float emin(float a, float b) { return fminf(a, b); }
__host__ __device__ inline double eabs(double a) { return fabs(a); }
__host__ __device__ inline float eabs(float a) { return fabsf(a); }
__host__ __device__ inline double esqrt(double a) { return sqrt(a); }
__host__ __device__ inline float esqrt(float a) { return sqrtf(a); }
__host__ __device__ inline double ecopysign(double a, double b) { return copysign(a, b); }
__host__ __device__ inline float ecopysign(float a, float b) { return copysignf(a, b); }

// computePrimitives (src/muscl/HydroBaseFunctor2D.h:79-105, HydroBaseFunctor3D.h:79-107).
// u/q are NV-long (4 in 2D, 5 in 3D); c is not produced here (only the dt needs it).
template <class T, int DIM>
__host__ __device__ inline void prims(const Gas<T>& g, const T* u, T* q)
{
  const T d = emax(u[ID], g.smallr);
  const T ux = u[IU] / d;
  const T uy = u[IV] / d;
  T eken;
  T uz = T(0);
  if constexpr (DIM == 3) {
    uz = u[IW] / d;
    eken = kHalf<T>() * (ux * ux + uy * uy + uz * uz);
  } else {
    eken = kHalf<T>() * (ux * ux + uy * uy);
  }
  T e = u[IP] / d - eken;
  if constexpr (sizeof(T) == 4) {
    // FP32 state: when the internal energy is a small remainder of E/rho (more than 7 of
    // the 24 mantissa bits cancelled), recover it in FP64 from the same conserved values.
    if (e < T(0.0078125) * (u[IP] / d)) {
      const double dd = d, ix = (double)u[IU] / dd, iy = (double)u[IV] / dd;
      double ek = ix * ix + iy * iy;
      if constexpr (DIM == 3) {
        const double iz = (double)u[IW] / dd;
        ek += iz * iz;
      }
      e = (T)((double)u[IP] / dd - 0.5 * ek);
    }
  }
  q[ID] = d;
  q[IP] = emax(g.gm1 * d * e, d * g.smallp);
  q[IU] = ux;
  q[IV] = uy;
  if constexpr (DIM == 3)
    q[IW] = uz;
}

// Per-cell contribution to the CFL reduction (ComputeDtFunctor2D/3D,
// src/muscl/HydroRunFunctors2D.h:48-78, HydroRunFunctors3D.h:46-82). Divisions by the
// cell size become multiplications only when the host proved 1/dx exact (power of two),
// which gives the identical IEEE result.
template <class T, int DIM>
__host__ __device__ inline T cell_inv_dt(const Gas<T>& g, const T* u, T dx, T dy, T dz, T idx,
                                         T idy, T idz, bool pow2)
{
  T q[5];
  prims<T, DIM>(g, u, q);
  const T c = esqrt(g.gamma0 * (q[IP]) / q[ID]);
  const T vx = c + eabs(q[IU]);
  const T vy = c + eabs(q[IV]);
  if constexpr (DIM == 3) {
    const T vz = c + eabs(q[IW]);
    if (pow2)
      return vx * idx + vy * idy + vz * idz;
    return vx / dx + vy / dy + vz / dz;
  } else {
    if (pow2)
      return vx * idx + vy * idy;
    return vx / dx + vy / dy;
  }
}

// One direction of slope_unsplit_hydro_{2d,3d}_scalar (src/muscl/HydroBaseFunctor2D.h:423-457).
// ST is slope_type (1 minmod, 2 MC); ST*x is written as the identical exact product.
template <class T, int ST>
__host__ __device__ inline T slope1(T q, T qp, T qm)
{
  T dlft, drgt;
  if constexpr (ST == 1) {
    dlft = q - qm;
    drgt = qp - q;
  } else {
    dlft = T(ST) * (q - qm);
    drgt = T(ST) * (qp - q);
  }
  const T dcen = kHalf<T>() * (qp - qm);
  const T slop = emin(eabs(dlft), eabs(drgt));
  const T dlim = ((dlft * drgt) <= T(0)) ? T(0) : slop;
  const T m = emin(dlim, eabs(dcen));
  return (dcen >= T(0)) ? m : -m;
}

// Source terms of the Hancock half step (the s vector of trace_unsplit_{2d,3d}_along_dir,
// src/muscl/HydroBaseFunctor2D.h:247-251, HydroBaseFunctor3D.h:317-321).
template <class T, int DIM>
__host__ __device__ inline void trace_source(const Gas<T>& g, const T* q, const T (*dq)[5], T* s)
{
  const T r = q[ID], p = q[IP], u = q[IU], v = q[IV];
  if constexpr (DIM == 2) {
    const T drx = dq[0][ID], dpx = dq[0][IP], dux = dq[0][IU], dvx = dq[0][IV];
    const T dry = dq[1][ID], dpy = dq[1][IP], duy = dq[1][IU], dvy = dq[1][IV];
    s[ID] = -u * drx - v * dry - (dux + dvy) * r;
    s[IP] = -u * dpx - v * dpy - (dux + dvy) * g.gamma0 * p;
    s[IU] = -u * dux - v * duy - (dpx) / r;
    s[IV] = -u * dvx - v * dvy - (dpy) / r;
  } else {
    const T w = q[IW];
    const T drx = dq[0][ID], dpx = dq[0][IP], dux = dq[0][IU], dvx = dq[0][IV], dwx = dq[0][IW];
    const T dry = dq[1][ID], dpy = dq[1][IP], duy = dq[1][IU], dvy = dq[1][IV], dwy = dq[1][IW];
    const T drz = dq[2][ID], dpz = dq[2][IP], duz = dq[2][IU], dvz = dq[2][IV], dwz = dq[2][IW];
    s[ID] = -u * drx - v * dry - w * drz - (dux + dvy + dwz) * r;
    s[IP] = -u * dpx - v * dpy - w * dpz - (dux + dvy + dwz) * g.gamma0 * p;
    s[IU] = -u * dux - v * duy - w * duz - (dpx) / r;
    s[IV] = -u * dvx - v * dvy - w * dvz - (dpy) / r;
    s[IW] = -u * dwx - v * dwy - w * dwz - (dpz) / r;
  }
}

// One face state of a cell: maxface -> q + dq/2 + s dt/(2dx) (FACE_XMAX/YMAX/ZMAX), else
// q - dq/2 + s dt/(2dx) (FACE_*MIN); density floored at smallr.
template <class T, int NV>
__host__ __device__ inline void trace_face(const Gas<T>& g, const T* q, const T* dqd, const T* s, T dtdxd,
                                           bool maxface, T* out)
{
  const T h = kHalf<T>();
#pragma unroll
  for (int vv = 0; vv < NV; vv++) {
    const T hd = h * dqd[vv];
    const T st = s[vv] * dtdxd * h;
    out[vv] = maxface ? q[vv] + hd + st : q[vv] - hd + st;
  }
  out[ID] = emax(g.smallr, out[ID]);
}

// The 2*DIM face states of one cell (trace_unsplit_{2d,3d}_along_dir,
// src/muscl/HydroBaseFunctor2D.h:218-293, HydroBaseFunctor3D.h:276-389).
// fm[d] = state at the cell's max face along d (FACE_XMAX/YMAX/ZMAX),
// fp[d] = state at its min face (FACE_XMIN/YMIN/ZMIN). dq[d][v] = slope along d.
// Only the faces selected by `want` (bit 2d = min face, bit 2d+1 = max face) are produced.
template <class T, int DIM>
__host__ __device__ inline void trace_cell(const Gas<T>& g, const T* q, const T (*dq)[5],
                                           const T* dtdx, unsigned want, T (*fm)[5], T (*fp)[5])
{
  constexpr int NV = DIM == 3 ? 5 : 4;
  T s[5];
  trace_source<T, DIM>(g, q, dq, s);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    if (want & (1u << (2 * d)))
      trace_face<T, NV>(g, q, dq[d], s, dtdx[d], false, fp[d]);
    if (want & (2u << (2 * d)))
      trace_face<T, NV>(g, q, dq[d], s, dtdx[d], true, fm[d]);
  }
}

// HLLC (riemann_hllc, src/shared/RiemannSolvers.h:445-568). The reference forms the
// left- and right-star states and then picks one of four; here only the star state the
// sampling will pick is formed (same expressions, same order), which halves the divides.
template <class T, int NV>
__host__ __device__ inline void hllc(const Gas<T>& g, const T* ql, const T* qr, T* f)
{
  const T h = kHalf<T>();
  const T rl = emax(ql[ID], g.smallr);
  const T pl = emax(ql[IP], rl * g.smallp);
  const T ul = ql[IU];
  T ecinl = h * rl * ul * ul;
  ecinl += h * rl * ql[IV] * ql[IV];
  if constexpr (NV == 5)
    ecinl += h * rl * ql[IW] * ql[IW];
  const T etotl = pl * g.entho + ecinl;
  const T rr = emax(qr[ID], g.smallr);
  const T pr = emax(qr[IP], rr * g.smallp);
  const T ur = qr[IU];
  T ecinr = h * rr * ur * ur;
  ecinr += h * rr * qr[IV] * qr[IV];
  if constexpr (NV == 5)
    ecinr += h * rr * qr[IW] * qr[IW];
  const T etotr = pr * g.entho + ecinr;
  const T cfastl = esqrt(emax(g.gamma0 * pl / rl, g.smallc2));
  const T cfastr = esqrt(emax(g.gamma0 * pr / rr, g.smallc2));
  const T SL = emin(ul, ur) - emax(cfastl, cfastr);
  const T SR = emax(ul, ur) + emax(cfastl, cfastr);
  const T rcl = rl * (ul - SL);
  const T rcr = rr * (SR - ur);
  const T ustar = (rcr * ur + rcl * ul + (pl - pr)) / (rcr + rcl);
  const T ptotstar = (rcr * pl + rcl * pr + rcl * rcr * (ul - ur)) / (rcr + rcl);
  // star state on the side of the contact that sampling at x/t=0 selects
  const bool lstar = ustar > T(0);
  const T rx = lstar ? rl : rr;
  const T Sx = lstar ? SL : SR;
  const T ux = lstar ? ul : ur;
  const T px = lstar ? pl : pr;
  const T ex = lstar ? etotl : etotr;
  const T smu = Sx - ux;
  const T rstar = rx * smu / (Sx - ustar);
  const T estar = (smu * ex - px * ux + ptotstar * ustar) / (Sx - ustar);
  T ro, uo, po, eo;
  if (SL > T(0)) {
    ro = rl;
    uo = ul;
    po = pl;
    eo = etotl;
  } else if (lstar || SR > T(0)) {
    ro = rstar;
    uo = ustar;
    po = ptotstar;
    eo = estar;
  } else {
    ro = rr;
    uo = ur;
    po = pr;
    eo = etotr;
  }
  const T fd = ro * uo;
  f[ID] = fd;
  f[IU] = fd * uo + po;
  f[IP] = (eo + po) * uo;
  const bool up = fd > T(0);
  f[IV] = fd * (up ? ql[IV] : qr[IV]);
  if constexpr (NV == 5)
    f[IW] = fd * (up ? ql[IW] : qr[IW]);
}

// cmpflx (src/shared/RiemannSolvers.h:19-48)
template <class T, int NV>
__host__ __device__ inline void cmpflx(const Gas<T>& g, const T* qg, T* f)
{
  const T h = kHalf<T>();
  f[ID] = qg[ID] * qg[IU];
  f[IU] = f[ID] * qg[IU] + qg[IP];
  f[IV] = f[ID] * qg[IV];
  if constexpr (NV == 5)
    f[IW] = f[ID] * qg[IW];
  T ekin = h * qg[ID] * (qg[IU] * qg[IU] + qg[IV] * qg[IV]);
  if constexpr (NV == 5)
    ekin += h * qg[ID] * (qg[IW] * qg[IW]);
  const T etot = qg[IP] * g.entho + ekin;
  f[IP] = qg[IU] * (etot + qg[IP]);
}

// Two-shock approximate solver (riemann_approx, src/shared/RiemannSolvers.h:59-199).
template <class T, int NV>
__host__ __device__ inline void approx(const Gas<T>& g, const T* ql, const T* qr, T* f)
{
  const T one = T(1), zero = T(0), h = kHalf<T>();
  const T rl = emax(ql[ID], g.smallr), ul = ql[IU], pl = emax(ql[IP], rl * g.smallp);
  const T rr = emax(qr[ID], g.smallr), ur = qr[IU], pr = emax(qr[IP], rr * g.smallp);
  const T cl = g.gamma0 * pl * rl;
  const T cr = g.gamma0 * pr * rr;
  T wl = esqrt(cl);
  T wr = esqrt(cr);
  T pstar = emax(((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr), zero);
  T pold = pstar;
  T conv = one;
  for (int iter = 0; (iter < 10) && (conv > T(1e-6)); ++iter) {
    const T wwl = esqrt(cl * (one + g.gamma6 * (pold - pl) / pl));
    const T wwr = esqrt(cr * (one + g.gamma6 * (pold - pr) / pr));
    const T qL = T(2) * wwl * wwl * wwl / (wwl * wwl + cl);
    const T qR = T(2) * wwr * wwr * wwr / (wwr * wwr + cr);
    const T usl = ul - (pold - pl) / wwl;
    const T usr = ur + (pold - pr) / wwr;
    const T delp = emax(qR * qL / (qR + qL) * (usl - usr), -pold);
    pold = pold + delp;
    conv = eabs(delp / (pold + g.smallpp));
  }
  pstar = pold;
  wl = esqrt(cl * (one + g.gamma6 * (pstar - pl) / pl));
  wr = esqrt(cr * (one + g.gamma6 * (pstar - pr) / pr));
  const T ustar = h * (ul + (pl - pstar) / wl + ur - (pr - pstar) / wr);
  const T sgnm = ecopysign(one, ustar);
  const bool left = sgnm > zero;
  const T ro = left ? rl : rr, uo = left ? ul : ur, po = left ? pl : pr, wo = left ? wl : wr;
  const T co = emax(g.smallc, esqrt(eabs(g.gamma0 * po / ro)));
  const T rstar = emax(ro / (one + ro * (po - pstar) / (wo * wo)), g.smallr);
  const T cstar = emax(g.smallc, esqrt(eabs(g.gamma0 * pstar / rstar)));
  T spout = co - sgnm * uo;
  T spin = cstar - sgnm * ustar;
  const T ushock = wo / ro - sgnm * uo;
  if (pstar >= po) {
    spin = ushock;
    spout = ushock;
  }
  const T scr = emax(spout - spin, g.smallc + eabs(spout + spin));
  T frac = h * (one + (spout + spin) / scr);
  if (frac != frac)
    frac = zero;
  else
    frac = frac >= one ? one : frac <= zero ? zero : frac;
  T qg[5];
  qg[ID] = frac * rstar + (one - frac) * ro;
  qg[IU] = frac * ustar + (one - frac) * uo;
  qg[IP] = frac * pstar + (one - frac) * po;
  if (spout < zero) {
    qg[ID] = ro;
    qg[IU] = uo;
    qg[IP] = po;
  }
  if (spin > zero) {
    qg[ID] = rstar;
    qg[IU] = ustar;
    qg[IP] = pstar;
  }
  qg[IV] = left ? ql[IV] : qr[IV];
  if constexpr (NV == 5)
    qg[IW] = left ? ql[IW] : qr[IW];
  cmpflx<T, NV>(g, qg, f);
}

// Conservative state and physical flux of a primitive state, as riemann_llf/riemann_hll
// build them (src/shared/RiemannSolvers.h:257-313, 382-422).
template <class T, int NV>
__host__ __device__ inline void cons_and_flux(const Gas<T>& g, const T* q, T* u, T* fl, bool hll)
{
  const T h = kHalf<T>();
  u[ID] = q[ID];
  u[IP] = q[IP] * g.entho + h * q[ID] * q[IU] * q[IU];
  u[IP] += h * q[ID] * q[IV] * q[IV];
  if constexpr (NV == 5)
    u[IP] += h * q[ID] * q[IW] * q[IW];
  u[IU] = q[ID] * q[IU];
  u[IV] = q[ID] * q[IV];
  if constexpr (NV == 5)
    u[IW] = q[ID] * q[IW];
  fl[ID] = hll ? u[IU] : u[ID] * q[IU];
  fl[IP] = q[IU] * (u[IP] + q[IP]);
  fl[IU] = q[IP] + u[IU] * q[IU];
  fl[IV] = fl[ID] * q[IV];
  if constexpr (NV == 5)
    fl[IW] = fl[ID] * q[IW];
}

// riemann_llf (src/shared/RiemannSolvers.h:217-327)
template <class T, int NV>
__host__ __device__ inline void llf(const Gas<T>& g, const T* ql, const T* qr, T* f)
{
  const T h = kHalf<T>();
  const T rl = emax(ql[ID], g.smallr), ul = ql[IU], pl = emax(ql[IP], rl * g.smallp);
  const T rr = emax(qr[ID], g.smallr), ur = qr[IU], pr = emax(qr[IP], rr * g.smallp);
  const T cl = esqrt(g.gamma0 * pl / rl);
  const T cr = esqrt(g.gamma0 * pr / rr);
  const T cmax = emax(eabs(ul) + cl, eabs(ur) + cr);
  T ul_[5], ur_[5], fl[5], fr[5];
  cons_and_flux<T, NV>(g, ql, ul_, fl, false);
  cons_and_flux<T, NV>(g, qr, ur_, fr, false);
#pragma unroll
  for (int v = 0; v < NV; v++)
    f[v] = h * (fl[v] + fr[v] - cmax * (ur_[v] - ul_[v]));
}

// riemann_hll (src/shared/RiemannSolvers.h:343-435); the reference's 3D dispatch has no
// HLL branch (RiemannSolvers.h:608-633), so the host rejects HLL in 3D.
template <class T, int NV>
__host__ __device__ inline void hll(const Gas<T>& g, const T* ql, const T* qr, T* f)
{
  const T zero = T(0);
  const T rl = emax(ql[ID], g.smallr), ul = ql[IU], pl = emax(ql[IP], rl * g.smallp);
  const T rr = emax(qr[ID], g.smallr), ur = qr[IU], pr = emax(qr[IP], rr * g.smallp);
  const T cl = esqrt(g.gamma0 * pl / rl);
  const T cr = esqrt(g.gamma0 * pr / rr);
  const T SL = emin(emin(ul, ur) - emax(cl, cr), zero);
  const T SR = emax(emax(ul, ur) + emax(cl, cr), zero);
  T ul_[5], ur_[5], fl[5], fr[5];
  cons_and_flux<T, NV>(g, ql, ul_, fl, true);
  cons_and_flux<T, NV>(g, qr, ur_, fr, true);
#pragma unroll
  for (int v = 0; v < NV; v++)
    f[v] = (SR * fl[v] - SL * fr[v] + SR * SL * (ur_[v] - ul_[v])) / (SR - SL);
}

template <class T, int NV, int RS>
__host__ __device__ inline void riemann(const Gas<T>& g, const T* ql, const T* qr, T* f)
{
  if constexpr (RS == RS_HLLC)
    hllc<T, NV>(g, ql, qr, f);
  else if constexpr (RS == RS_APPROX)
    approx<T, NV>(g, ql, qr, f);
  else if constexpr (RS == RS_LLF)
    llf<T, NV>(g, ql, qr, f);
  else
    hll<T, NV>(g, ql, qr, f);
}

} // namespace ek
