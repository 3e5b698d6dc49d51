/*
 * ek_oracle_mhd.c -- CPU restatement of euler_kokkos' 2D MHD MUSCL-Hancock step (impl 0).
 *
 * TEST INFRASTRUCTURE: the oracle the HIP MHD path is checked against, never the product.
 * Follows SolverMHDMuscl<2>::godunov_unsplit_impl (src/muscl/SolverMHDMuscl.cpp:369-454) on
 * the reference layout (LayoutRight U(i,j,var), ghosts included, gw = 3), in the
 * reference's operation order. Build with -ffp-contract=off. Paths are relative to the
 * euler_kokkos source tree.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ek_oracle.h"

#define ID 0
#define IP 1
#define IU 2
#define IV 3
#define IW 4
#define IA 5
#define IB 6
#define IC 7
#define NV 8

#define HALF 0.5
#define ONE 1.0
#define ZERO 0.0

/* edge indices (src/shared/enums.h:193-207) */
#define IRT 0
#define IRB 1
#define ILT 2
#define ILB 3
#define ILL 0
#define IRL 1
#define ILR 2
#define IRR 3

#define IX2(p, i, j, v) ((((long)(i)) * (p)->jsize + (j)) * NV + (v))

static double fmax4(double a0, double a1, double a2, double a3)
{
  /* FMAX4 (src/shared/mhd_utils.h:38-46) */
  double r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  return r;
}
static double fmin4(double a0, double a1, double a2, double a3)
{
  double r = a0;
  r = (a1 < r) ? a1 : r;
  r = (a2 < r) ? a2 : r;
  r = (a3 < r) ? a3 : r;
  return r;
}
static double fmax5(double a0, double a1, double a2, double a3, double a4)
{
  double r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  r = (a4 > r) ? a4 : r;
  return r;
}

/* constoprim_mhd (src/muscl/MHDBaseFunctor2D.h:192-242), cIso = 0 */
void eko_mhd_constoprim(const eko_params* p, const double* u, const double* mag, double* c, double* q)
{
  q[ID] = fmax(u[ID], p->smallr);
  q[IU] = u[IU] / q[ID];
  q[IV] = u[IV] / q[ID];
  q[IW] = u[IW] / q[ID];
  q[IA] = 0.5 * (u[IA] + mag[0]);
  q[IB] = 0.5 * (u[IB] + mag[1]);
  q[IC] = 0.5 * (u[IC] + mag[2]);
  double eken = 0.5 * (q[IU] * q[IU] + q[IV] * q[IV] + q[IW] * q[IW]);
  double emag = 0.5 * (q[IA] * q[IA] + q[IB] * q[IB] + q[IC] * q[IC]);
  double eint = (u[IP] - emag) / q[ID] - eken;
  q[IP] = fmax((p->gamma0 - 1.0) * q[ID] * eint, q[ID] * p->smallp);
  *c = sqrt(p->gamma0 * q[IP] / q[ID]);
}

/* slope_unsplit_hydro_2d_scalar, 3-argument form (MHDBaseFunctor2D.h:300-330) */
static double slope_mhd(const eko_params* p, double q, double qPlus, double qMinus)
{
  double dlft, drgt, dcen, dsgn, slop, dlim;
  dlft = p->slope_type * (q - qMinus);
  drgt = p->slope_type * (qPlus - q);
  dcen = 0.5 * (qPlus - qMinus);
  dsgn = (dcen >= ZERO) ? ONE : -ONE;
  slop = fmin(fabs(dlft), fabs(drgt));
  dlim = slop;
  if ((dlft * drgt) <= ZERO)
    dlim = ZERO;
  return dsgn * fmin(dlim, fabs(dcen));
}

/* slope_unsplit_mhd_2d (MHDBaseFunctor2D.h:456-509) */
static void slope_bf_2d(const eko_params* p, const double bfNeighbors[6], double dbf[2][3])
{
  const double bfx = bfNeighbors[0], bfx_yplus = bfNeighbors[1], bfx_yminus = bfNeighbors[2];
  const double bfy = bfNeighbors[3], bfy_xplus = bfNeighbors[4], bfy_xminus = bfNeighbors[5];
  for (int n = 0; n < 3; ++n) {
    dbf[0][n] = ZERO;
    dbf[1][n] = ZERO;
  }
  double dlft, drgt, dcen, dsgn, slop, dlim;
  dlft = p->slope_type * (bfx - bfx_yminus);
  drgt = p->slope_type * (bfx_yplus - bfx);
  dcen = HALF * (bfx_yplus - bfx_yminus);
  dsgn = (dcen >= ZERO) ? ONE : -ONE;
  slop = fmin(fabs(dlft), fabs(drgt));
  dlim = slop;
  if ((dlft * drgt) <= ZERO)
    dlim = ZERO;
  dbf[1][0] = dsgn * fmin(dlim, fabs(dcen));
  dlft = p->slope_type * (bfy - bfy_xminus);
  drgt = p->slope_type * (bfy_xplus - bfy);
  dcen = HALF * (bfy_xplus - bfy_xminus);
  dsgn = (dcen >= ZERO) ? ONE : -ONE;
  slop = fmin(fabs(dlft), fabs(drgt));
  dlim = slop;
  if ((dlft * drgt) <= ZERO)
    dlim = ZERO;
  dbf[0][1] = dsgn * fmin(dlim, fabs(dcen));
}

/* trace_unsplit_mhd_2d (MHDBaseFunctor2D.h:774-1109). qNb[3][3][8], bfNb[4][4][3];
 * outputs qm[2][8], qp[2][8], qEdge[4][8] (RT, RB, LT, LB). */
void eko_mhd_trace_2d(const eko_params* p, const double qNb[3][3][NV], const double bfNb[4][4][3],
                      double dtdx, double dtdy, double qm[2][NV], double qp[2][NV], double qEdge[4][NV])
{
  const double smallR = p->smallr, smallp = p->smallp, gamma = p->gamma0;
  const double* q = qNb[1][1];
  double Ez[2][2];
  for (int di = 0; di < 2; di++)
    for (int dj = 0; dj < 2; dj++) {
      int cX = 1 + di, cY = 1 + dj;
      double u = 0.25f * (qNb[cX - 1][cY - 1][IU] + qNb[cX - 1][cY][IU] + qNb[cX][cY - 1][IU] + qNb[cX][cY][IU]);
      double v = 0.25f * (qNb[cX - 1][cY - 1][IV] + qNb[cX - 1][cY][IV] + qNb[cX][cY - 1][IV] + qNb[cX][cY][IV]);
      double A = 0.5f * (bfNb[cX][cY - 1][0] + bfNb[cX][cY][0]);
      double B = 0.5f * (bfNb[cX - 1][cY][1] + bfNb[cX][cY][1]);
      Ez[di][dj] = u * B - v * A;
    }
  const double ELL = Ez[0][0], ELR = Ez[0][1], ERL = Ez[1][0], ERR = Ez[1][1];
  double r = q[ID], pr = q[IP], u = q[IU], v = q[IV], w = q[IW], A = q[IA], B = q[IB], C = q[IC];
  double AL = bfNb[1][1][0], AR = bfNb[2][1][0], BL = bfNb[1][1][1], BR = bfNb[1][2][1];
  double dq[2][NV];
  /* slope_unsplit_hydro_2d(qNb, dq): slope_type 1 or 2 (MHDBaseFunctor2D.h:345-390) */
  for (int n = 0; n < NV; n++) {
    dq[0][n] = slope_mhd(p, qNb[1][1][n], qNb[2][1][n], qNb[0][1][n]);
    dq[1][n] = slope_mhd(p, qNb[1][1][n], qNb[1][2][n], qNb[1][0][n]);
  }
  double drx = dq[0][ID]; drx *= 0.5;
  double dpx = dq[0][IP]; dpx *= 0.5;
  double dux = dq[0][IU]; dux *= 0.5;
  double dvx = dq[0][IV]; dvx *= 0.5;
  double dwx = dq[0][IW]; dwx *= 0.5;
  double dCx = dq[0][IC]; dCx *= 0.5;
  double dBx = dq[0][IB]; dBx *= 0.5;
  double dry = dq[1][ID]; dry *= 0.5;
  double dpy = dq[1][IP]; dpy *= 0.5;
  double duy = dq[1][IU]; duy *= 0.5;
  double dvy = dq[1][IV]; dvy *= 0.5;
  double dwy = dq[1][IW]; dwy *= 0.5;
  double dCy = dq[1][IC]; dCy *= 0.5;
  double dAy = dq[1][IA]; dAy *= 0.5;
  double bfn[6], dbf[2][3];
  bfn[0] = bfNb[1][1][0]; bfn[1] = bfNb[1][2][0]; bfn[2] = bfNb[1][0][0];
  bfn[3] = bfNb[1][1][1]; bfn[4] = bfNb[2][1][1]; bfn[5] = bfNb[0][1][1];
  slope_bf_2d(p, bfn, dbf);
  double dALy = 0.5 * dbf[1][0];
  double dBLx = 0.5 * dbf[0][1];
  bfn[0] = bfNb[2][1][0]; bfn[1] = bfNb[2][2][0]; bfn[2] = bfNb[2][0][0];
  bfn[3] = bfNb[2][1][1]; bfn[4] = bfNb[3][1][1]; bfn[5] = bfNb[1][1][1];
  slope_bf_2d(p, bfn, dbf);
  double dARy = 0.5 * dbf[1][0];
  bfn[0] = bfNb[1][2][0]; bfn[1] = bfNb[1][3][0]; bfn[2] = bfNb[1][1][0];
  bfn[3] = bfNb[1][2][1]; bfn[4] = bfNb[2][2][1]; bfn[5] = bfNb[0][2][1];
  slope_bf_2d(p, bfn, dbf);
  double dBRx = 0.5 * dbf[0][1];
  double dAx = 0.5 * (AR - AL);
  double dBy = 0.5 * (BR - BL);
  double sr0, su0, sv0, sw0, sp0, sA0, sB0, sC0, sAL0, sAR0, sBL0, sBR0;
  sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy;
  su0 = (-u * dux - dpx / r - B * dBx / r - C * dCx / r) * dtdx + (-v * duy + B * dAy / r) * dtdy;
  sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - dpy / r - A * dAy / r - C * dCy / r) * dtdy;
  sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy;
  sp0 = (-u * dpx - dux * gamma * pr) * dtdx + (-v * dpy - dvy * gamma * pr) * dtdy;
  sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy;
  sB0 = (-u * dBx - B * dux + v * dAx + A * dvx) * dtdx;
  sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (-v * dCy - C * dvy + w * dBy + B * dwy) * dtdy;
  sAL0 = +(ELR - ELL) * 0.5 * dtdy;
  sAR0 = +(ERR - ERL) * 0.5 * dtdy;
  sBL0 = -(ERL - ELL) * 0.5 * dtdx;
  sBR0 = -(ERR - ELR) * 0.5 * dtdx;
  r = r + sr0;
  u = u + su0;
  v = v + sv0;
  w = w + sw0;
  pr = pr + sp0;
  A = A + sA0;
  B = B + sB0;
  C = C + sC0;
  AL = AL + sAL0;
  AR = AR + sAR0;
  BL = BL + sBL0;
  BR = BR + sBR0;
  qp[0][ID] = r - drx; qp[0][IU] = u - dux; qp[0][IV] = v - dvx; qp[0][IW] = w - dwx;
  qp[0][IP] = pr - dpx; qp[0][IA] = AL; qp[0][IB] = B - dBx; qp[0][IC] = C - dCx;
  qp[0][ID] = fmax(smallR, qp[0][ID]);
  qp[0][IP] = fmax(smallp * qp[0][ID], qp[0][IP]);
  qm[0][ID] = r + drx; qm[0][IU] = u + dux; qm[0][IV] = v + dvx; qm[0][IW] = w + dwx;
  qm[0][IP] = pr + dpx; qm[0][IA] = AR; qm[0][IB] = B + dBx; qm[0][IC] = C + dCx;
  qm[0][ID] = fmax(smallR, qm[0][ID]);
  qm[0][IP] = fmax(smallp * qm[0][ID], qm[0][IP]);
  qp[1][ID] = r - dry; qp[1][IU] = u - duy; qp[1][IV] = v - dvy; qp[1][IW] = w - dwy;
  qp[1][IP] = pr - dpy; qp[1][IA] = A - dAy; qp[1][IB] = BL; qp[1][IC] = C - dCy;
  qp[1][ID] = fmax(smallR, qp[1][ID]);
  qp[1][IP] = fmax(smallp * qp[1][ID], qp[1][IP]);
  qm[1][ID] = r + dry; qm[1][IU] = u + duy; qm[1][IV] = v + dvy; qm[1][IW] = w + dwy;
  qm[1][IP] = pr + dpy; qm[1][IA] = A + dAy; qm[1][IB] = BR; qm[1][IC] = C + dCy;
  qm[1][ID] = fmax(smallR, qm[1][ID]);
  qm[1][IP] = fmax(smallp * qm[1][ID], qm[1][IP]);
  double* qRT = qEdge[IRT];
  double* qRB = qEdge[IRB];
  double* qLT = qEdge[ILT];
  double* qLB = qEdge[ILB];
  qRT[ID] = r + (+drx + dry); qRT[IU] = u + (+dux + duy); qRT[IV] = v + (+dvx + dvy);
  qRT[IW] = w + (+dwx + dwy); qRT[IP] = pr + (+dpx + dpy); qRT[IA] = AR + (+dARy);
  qRT[IB] = BR + (+dBRx); qRT[IC] = C + (+dCx + dCy);
  qRT[ID] = fmax(smallR, qRT[ID]);
  qRT[IP] = fmax(smallp * qRT[ID], qRT[IP]);
  qRB[ID] = r + (+drx - dry); qRB[IU] = u + (+dux - duy); qRB[IV] = v + (+dvx - dvy);
  qRB[IW] = w + (+dwx - dwy); qRB[IP] = pr + (+dpx - dpy); qRB[IA] = AR + (-dARy);
  qRB[IB] = BL + (+dBLx); qRB[IC] = C + (+dCx - dCy);
  qRB[ID] = fmax(smallR, qRB[ID]);
  qRB[IP] = fmax(smallp * qRB[ID], qRB[IP]);
  qLB[ID] = r + (-drx - dry); qLB[IU] = u + (-dux - duy); qLB[IV] = v + (-dvx - dvy);
  qLB[IW] = w + (-dwx - dwy); qLB[IP] = pr + (-dpx - dpy); qLB[IA] = AL + (-dALy);
  qLB[IB] = BL + (-dBLx); qLB[IC] = C + (-dCx - dCy);
  qLB[ID] = fmax(smallR, qLB[ID]);
  qLB[IP] = fmax(smallp * qLB[ID], qLB[IP]);
  qLT[ID] = r + (-drx + dry); qLT[IU] = u + (-dux + duy); qLT[IV] = v + (-dvx + dvy);
  qLT[IW] = w + (-dwx + dwy); qLT[IP] = pr + (-dpx + dpy); qLT[IA] = AL + (+dALy);
  qLT[IB] = BR + (-dBRx); qLT[IC] = C + (-dCx + dCy);
  qLT[ID] = fmax(smallR, qLT[ID]);
  qLT[IP] = fmax(smallp * qLT[ID], qLT[IP]);
}

/* find_speed_fast<dir> (src/shared/mhd_utils.h:89-117) */
static double speed_fast(const eko_params* p, const double* q, int dir)
{
  double d = q[ID], pr = q[IP], a = q[IA], b = q[IB], c = q[IC], b2, c2, d2, cf = 0;
  b2 = a * a + b * b + c * c;
  c2 = p->gamma0 * pr / d;
  d2 = 0.5 * (b2 / d + c2);
  if (dir == 0)
    cf = sqrt(d2 + sqrt(d2 * d2 - c2 * a * a / d));
  if (dir == 1)
    cf = sqrt(d2 + sqrt(d2 * d2 - c2 * b * b / d));
  if (dir == 2)
    cf = sqrt(d2 + sqrt(d2 * d2 - c2 * c * c / d));
  return cf;
}

/* find_speed_info<NDIM> (mhd_utils.h:319-366) */
void eko_mhd_speed_info(const eko_params* p, const double* q, double* out)
{
  double d = q[ID], pr = q[IP], a = q[IA], b = q[IB], c = q[IC], b2, c2, d2, cf;
  b2 = a * a + b * b + c * c;
  c2 = p->gamma0 * pr / d;
  d2 = 0.5 * (b2 / d + c2);
  cf = sqrt(d2 + sqrt(d2 * d2 - c2 * a * a / d));
  out[0] = cf + fabs(q[IU]);
  cf = sqrt(d2 + sqrt(d2 * d2 - c2 * b * b / d));
  out[1] = cf + fabs(q[IV]);
  if (p->dim == 3) {
    cf = sqrt(d2 + sqrt(d2 * d2 - c2 * c * c / d));
    out[2] = cf + fabs(q[IW]);
  }
}

/* find_mhd_flux (mhd_utils.h:173-231) */
static void mhd_flux(const eko_params* p, const double* q, double* cv, double* ff)
{
  const double pr = q[IP];
  const double entho = ONE / (p->gamma0 - ONE);
  double d = q[ID], u = q[IU], v = q[IV], w = q[IW], a = q[IA], b = q[IB], c = q[IC];
  double ecin = 0.5 * (u * u + v * v + w * w) * d;
  double emag = 0.5 * (a * a + b * b + c * c);
  double etot = pr * entho + ecin + emag;
  double ptot = pr + emag;
  cv[ID] = d; cv[IP] = etot; cv[IU] = d * u; cv[IV] = d * v; cv[IW] = d * w;
  cv[IA] = a; cv[IB] = b; cv[IC] = c;
  ff[ID] = d * u;
  ff[IP] = (etot + ptot) * u - a * (a * u + b * v + c * w);
  ff[IU] = d * u * u - a * a + ptot;
  ff[IV] = d * u * v - a * b;
  ff[IW] = d * u * w - a * c;
  ff[IA] = 0.0;
  ff[IB] = b * u - a * v;
  ff[IC] = c * u - a * w;
}

/* riemann_hll MHD (src/shared/RiemannSolvers_MHD.h:24-67) */
static void mhd_hll(const eko_params* p, double* ql, double* qr, double* flux)
{
  double bx_mean = 0.5 * (ql[IA] + qr[IA]);
  ql[IA] = bx_mean;
  qr[IA] = bx_mean;
  double ul[NV], fl[NV], ur[NV], fr[NV];
  mhd_flux(p, ql, ul, fl);
  mhd_flux(p, qr, ur, fr);
  double cfl = speed_fast(p, ql, 0), cfr = speed_fast(p, qr, 0);
  double vl = ql[IU], vr = qr[IU];
  double sl = fmin(fmin(vl, vr) - fmax(cfl, cfr), 0.0);
  double sr = fmax(fmax(vl, vr) + fmax(cfl, cfr), 0.0);
  for (int n = 0; n < NV; n++)
    flux[n] = (sr * fl[n] - sl * fr[n] + sr * sl * (ur[n] - ul[n])) / (sr - sl);
}

/* riemann_llf MHD (RiemannSolvers_MHD.h:80-110) */
static void mhd_llf(const eko_params* p, double* ql, double* qr, double* flux)
{
  double bx_mean = HALF * (ql[IA] + qr[IA]);
  ql[IA] = bx_mean;
  qr[IA] = bx_mean;
  double ul[NV], fl[NV], ur[NV], fr[NV];
  mhd_flux(p, ql, ul, fl);
  mhd_flux(p, qr, ur, fr);
  for (int n = 0; n < NV; n++)
    flux[n] = (fl[n] + fr[n]) / 2;
  double sil[3], sir[3];
  eko_params p1 = *p;
  p1.dim = 2;
  eko_mhd_speed_info(&p1, ql, sil);
  eko_mhd_speed_info(&p1, qr, sir);
  double vel_info = fmax(sil[0], sir[0]);
  for (int n = 0; n < NV; n++)
    flux[n] -= vel_info * (ur[n] - ul[n]) / 2;
}

/* riemann_hlld (RiemannSolvers_MHD.h:130-364) */
static void mhd_hlld(const eko_params* p, double* qleft, double* qright, double* flux)
{
  const double gamma0 = p->gamma0;
  const double entho = 1.0 / (gamma0 - 1.0);
  double a = 0.5 * (qleft[IA] + qright[IA]);
  double sgnm = (a >= 0) ? ONE : -ONE;
  qleft[IA] = a;
  qright[IA] = a;
  double rl = qleft[ID], pl = qleft[IP], ul = qleft[IU], vl = qleft[IV], wl = qleft[IW];
  double bl = qleft[IB], cl = qleft[IC];
  double ecinl = 0.5 * (ul * ul + vl * vl + wl * wl) * rl;
  double emagl = 0.5 * (a * a + bl * bl + cl * cl);
  double etotl = pl * entho + ecinl + emagl;
  double ptotl = pl + emagl;
  double vdotbl = ul * a + vl * bl + wl * cl;
  double rr = qright[ID], pr = qright[IP], ur = qright[IU], vr = qright[IV], wr = qright[IW];
  double br = qright[IB], cr = qright[IC];
  double ecinr = 0.5 * (ur * ur + vr * vr + wr * wr) * rr;
  double emagr = 0.5 * (a * a + br * br + cr * cr);
  double etotr = pr * entho + ecinr + emagr;
  double ptotr = pr + emagr;
  double vdotbr = ur * a + vr * br + wr * cr;
  double cfastl = speed_fast(p, qleft, 0);
  double cfastr = speed_fast(p, qright, 0);
  double sl = fmin(ul, ur) - fmax(cfastl, cfastr);
  double sr = fmax(ul, ur) + fmax(cfastl, cfastr);
  double rcl = rl * (ul - sl);
  double rcr = rr * (sr - ur);
  double ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  double ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);
  double estar, rstarl, el;
  rstarl = rl * (sl - ul) / (sl - ustar);
  estar = rl * (sl - ul) * (sl - ustar) - a * a;
  el = rl * (sl - ul) * (sl - ul) - a * a;
  double vstarl, wstarl, bstarl, cstarl;
  if (a * a > 0 && fabs(estar / (a * a) - ONE) <= 1e-8) {
    vstarl = vl; bstarl = bl; wstarl = wl; cstarl = cl;
  } else {
    vstarl = vl - a * bl * (ustar - ul) / estar;
    bstarl = bl * el / estar;
    wstarl = wl - a * cl * (ustar - ul) / estar;
    cstarl = cl * el / estar;
  }
  double vdotbstarl = ustar * a + vstarl * bstarl + wstarl * cstarl;
  double etotstarl = ((sl - ul) * etotl - ptotl * ul + ptotstar * ustar + a * (vdotbl - vdotbstarl)) / (sl - ustar);
  double sqrrstarl = sqrt(rstarl);
  double calfvenl = fabs(a) / sqrrstarl;
  double sal = ustar - calfvenl;
  double rstarr, er;
  rstarr = rr * (sr - ur) / (sr - ustar);
  estar = rr * (sr - ur) * (sr - ustar) - a * a;
  er = rr * (sr - ur) * (sr - ur) - a * a;
  double vstarr, wstarr, bstarr, cstarr;
  if (a * a > 0 && fabs(estar / (a * a) - ONE) <= 1e-8) {
    vstarr = vr; bstarr = br; wstarr = wr; cstarr = cr;
  } else {
    vstarr = vr - a * br * (ustar - ur) / estar;
    bstarr = br * er / estar;
    wstarr = wr - a * cr * (ustar - ur) / estar;
    cstarr = cr * er / estar;
  }
  double vdotbstarr = ustar * a + vstarr * bstarr + wstarr * cstarr;
  double etotstarr = ((sr - ur) * etotr - ptotr * ur + ptotstar * ustar + a * (vdotbr - vdotbstarr)) / (sr - ustar);
  double sqrrstarr = sqrt(rstarr);
  double calfvenr = fabs(a) / sqrrstarr;
  double sar = ustar + calfvenr;
  double vstarstar = (sqrrstarl * vstarl + sqrrstarr * vstarr + sgnm * (bstarr - bstarl)) / (sqrrstarl + sqrrstarr);
  double wstarstar = (sqrrstarl * wstarl + sqrrstarr * wstarr + sgnm * (cstarr - cstarl)) / (sqrrstarl + sqrrstarr);
  double bstarstar = (sqrrstarl * bstarr + sqrrstarr * bstarl + sgnm * sqrrstarl * sqrrstarr * (vstarr - vstarl)) / (sqrrstarl + sqrrstarr);
  double cstarstar = (sqrrstarl * cstarr + sqrrstarr * cstarl + sgnm * sqrrstarl * sqrrstarr * (wstarr - wstarl)) / (sqrrstarl + sqrrstarr);
  double vdotbstarstar = ustar * a + vstarstar * bstarstar + wstarstar * cstarstar;
  double etotstarstarl = etotstarl - sgnm * sqrrstarl * (vdotbstarl - vdotbstarstar);
  double etotstarstarr = etotstarr + sgnm * sqrrstarr * (vdotbstarr - vdotbstarstar);
  double ro, uo, vo, wo, bo, co, ptoto, etoto, vdotbo;
  if (sl > 0) {
    ro = rl; uo = ul; vo = vl; wo = wl; bo = bl; co = cl; ptoto = ptotl; etoto = etotl; vdotbo = vdotbl;
  } else if (sal > 0) {
    ro = rstarl; uo = ustar; vo = vstarl; wo = wstarl; bo = bstarl; co = cstarl; ptoto = ptotstar;
    etoto = etotstarl; vdotbo = vdotbstarl;
  } else if (ustar > 0) {
    ro = rstarl; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarl; vdotbo = vdotbstarstar;
  } else if (sar > 0) {
    ro = rstarr; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarr; vdotbo = vdotbstarstar;
  } else if (sr > 0) {
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
  flux[IA] = 0.0;
  flux[IB] = bo * uo - a * vo;
  flux[IC] = co * uo - a * wo;
}

/* riemann_mhd dispatch (RiemannSolvers_MHD.h:369-389); arguments are modified (Bn average) */
void eko_riemann_mhd(const eko_params* p, double* ql, double* qr, double* flux)
{
  if (p->riemann == EKO_RIEMANN_HLLD)
    mhd_hlld(p, ql, qr, flux);
  else if (p->riemann == EKO_RIEMANN_HLL)
    mhd_hll(p, ql, qr, flux);
  else if (p->riemann == EKO_RIEMANN_LLF)
    mhd_llf(p, ql, qr, flux);
}

/* mag_riemann2d_hlld (RiemannSolvers_MHD.h:395-627) */
static double mag_riemann2d_hlld(const eko_params* p, double qLLRR[4][NV], const double eLLRR[4])
{
  const double* qLL = qLLRR[ILL];
  const double* qRL = qLLRR[IRL];
  const double* qLR = qLLRR[ILR];
  const double* qRR = qLLRR[IRR];
  const double ELL = eLLRR[ILL], ERL = eLLRR[IRL], ELR = eLLRR[ILR], ERR = eLLRR[IRR];
  const double rLL = qLL[ID], pLL = qLL[IP], uLL = qLL[IU], vLL = qLL[IV], aLL = qLL[IA], bLL = qLL[IB], cLL = qLL[IC];
  const double rLR = qLR[ID], pLR = qLR[IP], uLR = qLR[IU], vLR = qLR[IV], aLR = qLR[IA], bLR = qLR[IB], cLR = qLR[IC];
  const double rRL = qRL[ID], pRL = qRL[IP], uRL = qRL[IU], vRL = qRL[IV], aRL = qRL[IA], bRL = qRL[IB], cRL = qRL[IC];
  const double rRR = qRR[ID], pRR = qRR[IP], uRR = qRR[IU], vRR = qRR[IV], aRR = qRR[IA], bRR = qRR[IB], cRR = qRR[IC];
  double cFastLLx = speed_fast(p, qLL, 0), cFastLRx = speed_fast(p, qLR, 0);
  double cFastRLx = speed_fast(p, qRL, 0), cFastRRx = speed_fast(p, qRR, 0);
  double cFastLLy = speed_fast(p, qLL, 1), cFastLRy = speed_fast(p, qLR, 1);
  double cFastRLy = speed_fast(p, qRL, 1), cFastRRy = speed_fast(p, qRR, 1);
  double SL = fmin4(uLL, uLR, uRL, uRR) - fmax4(cFastLLx, cFastLRx, cFastRLx, cFastRRx);
  double SR = fmax4(uLL, uLR, uRL, uRR) + fmax4(cFastLLx, cFastLRx, cFastRLx, cFastRRx);
  double SB = fmin4(vLL, vLR, vRL, vRR) - fmax4(cFastLLy, cFastLRy, cFastRLy, cFastRRy);
  double ST = fmax4(vLL, vLR, vRL, vRR) + fmax4(cFastLLy, cFastLRy, cFastRLy, cFastRRy);
  double PtotLL = pLL + HALF * (aLL * aLL + bLL * bLL + cLL * cLL);
  double PtotLR = pLR + HALF * (aLR * aLR + bLR * bLR + cLR * cLR);
  double PtotRL = pRL + HALF * (aRL * aRL + bRL * bRL + cRL * cRL);
  double PtotRR = pRR + HALF * (aRR * aRR + bRR * bRR + cRR * cRR);
  double rcLLx = rLL * (uLL - SL), rcRLx = rRL * (SR - uRL), rcLRx = rLR * (uLR - SL), rcRRx = rRR * (SR - uRR);
  double rcLLy = rLL * (vLL - SB), rcLRy = rLR * (ST - vLR), rcRLy = rRL * (vRL - SB), rcRRy = rRR * (ST - vRR);
  double ustar = (rcLLx * uLL + rcLRx * uLR + rcRLx * uRL + rcRRx * uRR + (PtotLL - PtotRL + PtotLR - PtotRR)) /
                 (rcLLx + rcLRx + rcRLx + rcRRx);
  double vstar = (rcLLy * vLL + rcLRy * vLR + rcRLy * vRL + rcRRy * vRR + (PtotLL - PtotLR + PtotRL - PtotRR)) /
                 (rcLLy + rcLRy + rcRLy + rcRRy);
  double rstarLLx = rLL * (SL - uLL) / (SL - ustar);
  double BstarLL = bLL * (SL - uLL) / (SL - ustar);
  double rstarLLy = rLL * (SB - vLL) / (SB - vstar);
  double AstarLL = aLL * (SB - vLL) / (SB - vstar);
  double rstarLL = rLL * (SL - uLL) / (SL - ustar) * (SB - vLL) / (SB - vstar);
  double EstarLLx = ustar * BstarLL - vLL * aLL;
  double EstarLLy = uLL * bLL - vstar * AstarLL;
  double EstarLL = ustar * BstarLL - vstar * AstarLL;
  double rstarLRx = rLR * (SL - uLR) / (SL - ustar);
  double BstarLR = bLR * (SL - uLR) / (SL - ustar);
  double rstarLRy = rLR * (ST - vLR) / (ST - vstar);
  double AstarLR = aLR * (ST - vLR) / (ST - vstar);
  double rstarLR = rLR * (SL - uLR) / (SL - ustar) * (ST - vLR) / (ST - vstar);
  double EstarLRx = ustar * BstarLR - vLR * aLR;
  double EstarLRy = uLR * bLR - vstar * AstarLR;
  double EstarLR = ustar * BstarLR - vstar * AstarLR;
  double rstarRLx = rRL * (SR - uRL) / (SR - ustar);
  double BstarRL = bRL * (SR - uRL) / (SR - ustar);
  double rstarRLy = rRL * (SB - vRL) / (SB - vstar);
  double AstarRL = aRL * (SB - vRL) / (SB - vstar);
  double rstarRL = rRL * (SR - uRL) / (SR - ustar) * (SB - vRL) / (SB - vstar);
  double EstarRLx = ustar * BstarRL - vRL * aRL;
  double EstarRLy = uRL * bRL - vstar * AstarRL;
  double EstarRL = ustar * BstarRL - vstar * AstarRL;
  double rstarRRx = rRR * (SR - uRR) / (SR - ustar);
  double BstarRR = bRR * (SR - uRR) / (SR - ustar);
  double rstarRRy = rRR * (ST - vRR) / (ST - vstar);
  double AstarRR = aRR * (ST - vRR) / (ST - vstar);
  double rstarRR = rRR * (SR - uRR) / (SR - ustar) * (ST - vRR) / (ST - vstar);
  double EstarRRx = ustar * BstarRR - vRR * aRR;
  double EstarRRy = uRR * bRR - vstar * AstarRR;
  double EstarRR = ustar * BstarRR - vstar * AstarRR;
  double calfvenL = fmax5(fabs(aLR) / sqrt(rstarLRx), fabs(AstarLR) / sqrt(rstarLR), fabs(aLL) / sqrt(rstarLLx),
                          fabs(AstarLL) / sqrt(rstarLL), p->smallc);
  double calfvenR = fmax5(fabs(aRR) / sqrt(rstarRRx), fabs(AstarRR) / sqrt(rstarRR), fabs(aRL) / sqrt(rstarRLx),
                          fabs(AstarRL) / sqrt(rstarRL), p->smallc);
  double calfvenB = fmax5(fabs(bLL) / sqrt(rstarLLy), fabs(BstarLL) / sqrt(rstarLL), fabs(bRL) / sqrt(rstarRLy),
                          fabs(BstarRL) / sqrt(rstarRL), p->smallc);
  double calfvenT = fmax5(fabs(bLR) / sqrt(rstarLRy), fabs(BstarLR) / sqrt(rstarLR), fabs(bRR) / sqrt(rstarRRy),
                          fabs(BstarRR) / sqrt(rstarRR), p->smallc);
  double SAL = fmin(ustar - calfvenL, (double)ZERO);
  double SAR = fmax(ustar + calfvenR, (double)ZERO);
  double SAB = fmin(vstar - calfvenB, (double)ZERO);
  double SAT = fmax(vstar + calfvenT, (double)ZERO);
  double AstarT = (SAR * AstarRR - SAL * AstarLR) / (SAR - SAL);
  double AstarB = (SAR * AstarRL - SAL * AstarLL) / (SAR - SAL);
  double BstarR = (SAT * BstarRR - SAB * BstarRL) / (SAT - SAB);
  double BstarL = (SAT * BstarLR - SAB * BstarLL) / (SAT - SAB);
  double E = 0, tmpE = 0;
  int SB_pos = (int)(1 + copysign(ONE, SB)) / 2, SB_neg = 1 - SB_pos;
  int ST_pos = (int)(1 + copysign(ONE, ST)) / 2, ST_neg = 1 - ST_pos;
  int SL_pos = (int)(1 + copysign(ONE, SL)) / 2, SL_neg = 1 - SL_pos;
  int SR_pos = (int)(1 + copysign(ONE, SR)) / 2, SR_neg = 1 - SR_pos;
  tmpE = (SAL * SAB * EstarRR - SAL * SAT * EstarRL - SAR * SAB * EstarLR + SAR * SAT * EstarLL) / (SAR - SAL) /
           (SAT - SAB) -
         SAT * SAB / (SAT - SAB) * (AstarT - AstarB) + SAR * SAL / (SAR - SAL) * (BstarR - BstarL);
  E += (SB_neg * ST_pos * SL_neg * SR_pos) * tmpE;
  tmpE = (SAR * EstarLLx - SAL * EstarRLx + SAR * SAL * (bRL - bLL)) / (SAR - SAL);
  tmpE = SL_pos * ELL + SL_neg * SR_neg * ERL + SL_neg * SR_pos * tmpE;
  E += SB_pos * tmpE;
  tmpE = (SAR * EstarLRx - SAL * EstarRRx + SAR * SAL * (bRR - bLR)) / (SAR - SAL);
  tmpE = SL_pos * ELR + SL_neg * SR_neg * ERR + SL_neg * SR_pos * tmpE;
  E += (SB_neg * ST_neg) * tmpE;
  tmpE = (SAT * EstarLLy - SAB * EstarLRy - SAT * SAB * (aLR - aLL)) / (SAT - SAB);
  E += (SB_neg * ST_pos * SL_pos) * tmpE;
  tmpE = (SAT * EstarRLy - SAB * EstarRRy - SAT * SAB * (aRR - aRL)) / (SAT - SAB);
  E += (SB_neg * ST_pos * SL_neg * SR_neg) * tmpE;
  return E;
}

/* compute_emf<EMFZ> (RiemannSolvers_MHD.h:648-871) */
double eko_compute_emfz(const eko_params* p, double qEdge[4][NV])
{
  const double* qRT = qEdge[IRT];
  const double* qLT = qEdge[ILT];
  const double* qRB = qEdge[IRB];
  const double* qLB = qEdge[ILB];
  double q4[4][NV];
  double* qLL = q4[ILL];
  double* qRL = q4[IRL];
  double* qLR = q4[ILR];
  double* qRR = q4[IRR];
  qLL[ID] = qRT[ID]; qRL[ID] = qLT[ID]; qLR[ID] = qRB[ID]; qRR[ID] = qLB[ID];
  qLL[IP] = qRT[IP]; qRL[IP] = qLT[IP]; qLR[IP] = qRB[IP]; qRR[IP] = qLB[IP];
  qLL[IU] = qRT[IU]; qRL[IU] = qLT[IU]; qLR[IU] = qRB[IU]; qRR[IU] = qLB[IU];
  qLL[IV] = qRT[IV]; qRL[IV] = qLT[IV]; qLR[IV] = qRB[IV]; qRR[IV] = qLB[IV];
  qLL[IA] = HALF * (qRT[IA] + qLT[IA]);
  qRL[IA] = HALF * (qRT[IA] + qLT[IA]);
  qLR[IA] = HALF * (qRB[IA] + qLB[IA]);
  qRR[IA] = HALF * (qRB[IA] + qLB[IA]);
  qLL[IB] = HALF * (qRT[IB] + qRB[IB]);
  qRL[IB] = HALF * (qLT[IB] + qLB[IB]);
  qLR[IB] = HALF * (qRT[IB] + qRB[IB]);
  qRR[IB] = HALF * (qLT[IB] + qLB[IB]);
  qLL[IW] = qRT[IW]; qRL[IW] = qLT[IW]; qLR[IW] = qRB[IW]; qRR[IW] = qLB[IW];
  qLL[IC] = qRT[IC]; qRL[IC] = qLT[IC]; qLR[IC] = qRB[IC]; qRR[IC] = qLB[IC];
  double e[4];
  e[ILL] = qLL[IU] * qLL[IB] - qLL[IV] * qLL[IA];
  e[IRL] = qRL[IU] * qRL[IB] - qRL[IV] * qRL[IA];
  e[ILR] = qLR[IU] * qLR[IB] - qLR[IV] * qLR[IA];
  e[IRR] = qRR[IU] * qRR[IB] - qRR[IV] * qRR[IA];
  return mag_riemann2d_hlld(p, q4, e);
}

/* ---------------------------------------------------------------- array kernels (2D) */

/* ConvertToPrimitivesFunctor2D_MHD (src/muscl/MHDRunFunctors2D.h:85-159): i < isize-1, j < jsize-1 */
void eko_mhd_convert_to_primitives(const eko_params* p, const double* U, double* Q)
{
#pragma omp parallel for collapse(2)
  for (int i = 0; i < p->isize - 1; i++)
    for (int j = 0; j < p->jsize - 1; j++) {
      double u[NV], q[NV], mag[3], c;
      for (int v = 0; v < NV; v++)
        u[v] = U[IX2(p, i, j, v)];
      mag[0] = U[IX2(p, i + 1, j, IA)];
      mag[1] = U[IX2(p, i, j + 1, IB)];
      mag[2] = 0.0;
      eko_mhd_constoprim(p, u, mag, &c, q);
      for (int v = 0; v < NV; v++)
        Q[IX2(p, i, j, v)] = q[v];
    }
}

/* ComputeDtFunctor2D_MHD (MHDRunFunctors2D.h:16-80), from Q */
double eko_mhd_inv_dt(const eko_params* p, const double* Q)
{
  const int gw = p->gw;
  double invDt = 0.0;
#pragma omp parallel for collapse(2) reduction(max : invDt)
  for (int i = gw; i < p->isize - gw; i++)
    for (int j = gw; j < p->jsize - gw; j++) {
      double q[NV], s[3];
      for (int v = 0; v < NV; v++)
        q[v] = Q[IX2(p, i, j, v)];
      eko_params p1 = *p;
      p1.dim = 2;
      eko_mhd_speed_info(&p1, q, s);
      invDt = fmax(invDt, s[0] / p->dx + s[1] / p->dy);
    }
  return invDt;
}

/* ComputeTraceFunctor2D_MHD (MHDRunFunctors2D.h:782-911): i,j in [gw-2, size-gw+1).
 * Out arrays: Qm_x, Qm_y, Qp_x, Qp_y, RT, RB, LT, LB (each isize*jsize*8). */
void eko_mhd_compute_trace(const eko_params* p, const double* U, const double* Q, double* const out[8],
                           double dtdx, double dtdy)
{
  const int gw = p->gw;
#pragma omp parallel for collapse(2)
  for (int i = gw - 2; i < p->isize - gw + 1; i++)
    for (int j = gw - 2; j < p->jsize - gw + 1; j++) {
      double qNb[3][3][NV], bfNb[4][4][3], qm[2][NV], qp[2][NV], qE[4][NV];
      for (int di = 0; di < 3; di++)
        for (int dj = 0; dj < 3; dj++)
          for (int v = 0; v < NV; v++)
            qNb[di][dj][v] = Q[IX2(p, i + di - 1, j + dj - 1, v)];
      for (int di = 0; di < 4; di++)
        for (int dj = 0; dj < 4; dj++) {
          bfNb[di][dj][0] = U[IX2(p, i + di - 1, j + dj - 1, IA)];
          bfNb[di][dj][1] = U[IX2(p, i + di - 1, j + dj - 1, IB)];
          bfNb[di][dj][2] = U[IX2(p, i + di - 1, j + dj - 1, IC)];
        }
      eko_mhd_trace_2d(p, (const double(*)[3][NV])qNb, (const double(*)[4][3])bfNb, dtdx, dtdy, qm, qp, qE);
      const double* src[8] = { qm[0], qm[1], qp[0], qp[1], qE[IRT], qE[IRB], qE[ILT], qE[ILB] };
      for (int a = 0; a < 8; a++)
        for (int v = 0; v < NV; v++)
          out[a][IX2(p, i, j, v)] = src[a][v];
    }
}

static void swapd(double* a, double* b)
{
  double t = *a;
  *a = *b;
  *b = t;
}

/* ComputeFluxesAndStoreFunctor2D_MHD (MHDRunFunctors2D.h:375-470): raw fluxes, y swapped */
void eko_mhd_compute_fluxes(const eko_params* p, double* const tr[8], double* Fx, double* Fy)
{
  const int gw = p->gw;
  const double* Qm_x = tr[0];
  const double* Qm_y = tr[1];
  const double* Qp_x = tr[2];
  const double* Qp_y = tr[3];
#pragma omp parallel for collapse(2)
  for (int i = gw; i < p->isize - gw + 1; i++)
    for (int j = gw; j < p->jsize - gw + 1; j++) {
      double ql[NV], qr[NV], f[NV];
      for (int v = 0; v < NV; v++) {
        ql[v] = Qm_x[IX2(p, i - 1, j, v)];
        qr[v] = Qp_x[IX2(p, i, j, v)];
      }
      eko_riemann_mhd(p, ql, qr, f);
      for (int v = 0; v < NV; v++)
        Fx[IX2(p, i, j, v)] = f[v];
      for (int v = 0; v < NV; v++) {
        ql[v] = Qm_y[IX2(p, i, j - 1, v)];
        qr[v] = Qp_y[IX2(p, i, j, v)];
      }
      swapd(&ql[IU], &ql[IV]);
      swapd(&ql[IA], &ql[IB]);
      swapd(&qr[IU], &qr[IV]);
      swapd(&qr[IA], &qr[IB]);
      eko_riemann_mhd(p, ql, qr, f);
      for (int v = 0; v < NV; v++)
        Fy[IX2(p, i, j, v)] = f[v];
    }
}

/* ComputeEmfAndStoreFunctor2D (MHDRunFunctors2D.h:612-686) */
void eko_mhd_compute_emf(const eko_params* p, double* const tr[8], double* Emf)
{
  const int gw = p->gw;
#pragma omp parallel for collapse(2)
  for (int i = gw; i < p->isize - gw + 1; i++)
    for (int j = gw; j < p->jsize - gw + 1; j++) {
      double qE[4][NV];
      for (int v = 0; v < NV; v++) {
        qE[IRT][v] = tr[4][IX2(p, i - 1, j - 1, v)];
        qE[IRB][v] = tr[5][IX2(p, i - 1, j, v)];
        qE[ILT][v] = tr[6][IX2(p, i, j - 1, v)];
        qE[ILB][v] = tr[7][IX2(p, i, j, v)];
      }
      Emf[(long)i * p->jsize + j] = eko_compute_emfz(p, qE);
    }
}

/* UpdateFunctor2D_MHD + UpdateEmfFunctor2D (MHDRunFunctors2D.h:1539-1699) */
void eko_mhd_update(const eko_params* p, double* U, const double* Fx, const double* Fy, const double* Emf,
                    double dtdx, double dtdy)
{
  const int gw = p->gw;
  static const int vx[6] = { ID, IP, IU, IV, IW, IC };
#pragma omp parallel for collapse(2)
  for (int i = gw; i < p->isize - gw; i++)
    for (int j = gw; j < p->jsize - gw; j++) {
      double ud[NV];
      for (int v = 0; v < NV; v++)
        ud[v] = U[IX2(p, i, j, v)];
      for (int n = 0; n < 6; n++)
        ud[vx[n]] += Fx[IX2(p, i, j, vx[n])] * dtdx;
      for (int n = 0; n < 6; n++)
        ud[vx[n]] -= Fx[IX2(p, i + 1, j, vx[n])] * dtdx;
      /* y fluxes are stored in the rotated frame: IU<->IV (IA/IB not used here) */
      ud[ID] += Fy[IX2(p, i, j, ID)] * dtdy;
      ud[IP] += Fy[IX2(p, i, j, IP)] * dtdy;
      ud[IU] += Fy[IX2(p, i, j, IV)] * dtdy;
      ud[IV] += Fy[IX2(p, i, j, IU)] * dtdy;
      ud[IW] += Fy[IX2(p, i, j, IW)] * dtdy;
      ud[IC] += Fy[IX2(p, i, j, IC)] * dtdy;
      ud[ID] -= Fy[IX2(p, i, j + 1, ID)] * dtdy;
      ud[IP] -= Fy[IX2(p, i, j + 1, IP)] * dtdy;
      ud[IU] -= Fy[IX2(p, i, j + 1, IV)] * dtdy;
      ud[IV] -= Fy[IX2(p, i, j + 1, IU)] * dtdy;
      ud[IW] -= Fy[IX2(p, i, j + 1, IW)] * dtdy;
      ud[IC] -= Fy[IX2(p, i, j + 1, IC)] * dtdy;
      for (int v = 0; v < NV; v++)
        U[IX2(p, i, j, v)] = ud[v];
    }
#pragma omp parallel for collapse(2)
  for (int i = gw; i < p->isize - gw; i++)
    for (int j = gw; j < p->jsize - gw; j++) {
      const long e = (long)i * p->jsize + j;
      U[IX2(p, i, j, IA)] += (Emf[e + 1] - Emf[e]) * dtdy;
      U[IX2(p, i, j, IB)] -= (Emf[e + p->jsize] - Emf[e]) * dtdx;
    }
}

/* InitOrszagTangFunctor2D (src/muscl/MHDInitFunctors2D.h:345-479) */
void eko_mhd_init_orszag_tang(const eko_params* p, double* U)
{
  const int gw = p->gw;
  const double TWOPI = 2 * 3.141592653589793238462643383279502884;
  const double gamma0 = p->gamma0;
  const double B0 = 1.0 / sqrt(2 * TWOPI);
  const double p0 = gamma0 / (2 * TWOPI);
  const double d0 = gamma0 * p0;
  const double v0 = 1.0;
  for (int i = 0; i < p->isize; i++)
    for (int j = 0; j < p->jsize; j++) {
      double xPos = p->xmin + p->dx / 2 + (i - gw) * p->dx;
      double yPos = p->ymin + p->dy / 2 + (j - gw) * p->dy;
      U[IX2(p, i, j, ID)] = d0;
      U[IX2(p, i, j, IU)] = -d0 * v0 * sin(yPos * TWOPI);
      U[IX2(p, i, j, IV)] = d0 * v0 * sin(xPos * TWOPI);
      U[IX2(p, i, j, IW)] = ZERO;
      U[IX2(p, i, j, IA)] = -B0 * sin(yPos * TWOPI);
      U[IX2(p, i, j, IB)] = B0 * sin(2.0 * xPos * TWOPI);
      U[IX2(p, i, j, IC)] = 0.0;
    }
  const double TwoPi = 4.0 * asin(1.0);
  const double p0e = gamma0 / (2.0 * TwoPi);
  for (int i = 0; i < p->isize - 1; i++)
    for (int j = 0; j < p->jsize - 1; j++) {
#define SQR(x) ((x) * (x))
      U[IX2(p, i, j, IP)] = p0e / (gamma0 - 1.0) +
                            0.5 * (SQR(U[IX2(p, i, j, IU)]) / U[IX2(p, i, j, ID)] +
                                   SQR(U[IX2(p, i, j, IV)]) / U[IX2(p, i, j, ID)]) +
                            0.5 * (SQR((U[IX2(p, i, j, IA)] + U[IX2(p, i + 1, j, IA)]) / 2) +
                                   SQR((U[IX2(p, i, j, IB)] + U[IX2(p, i, j + 1, IB)]) / 2));
#undef SQR
    }
}

/* SolverMHDMuscl<2> time loop: next_iteration_impl + godunov_unsplit_impl (impl 0)
 * (src/muscl/SolverMHDMuscl.h:741-801, SolverMHDMuscl.cpp:369-454). dt is computed inside
 * the step from the primitives of the current state. */
int eko_mhd_run(const eko_params* p, double* U, int nStepmax, double tEnd, double* t_io, int* it_io,
                double* dt_last)
{
  const long n = (long)p->isize * p->jsize * NV;
  const long ns = (long)p->isize * p->jsize;
  double* U2 = (double*)malloc(sizeof(double) * n);
  double* Q = (double*)calloc(n, sizeof(double));
  double* Fx = (double*)calloc(n, sizeof(double));
  double* Fy = (double*)calloc(n, sizeof(double));
  double* Emf = (double*)calloc(ns, sizeof(double));
  double* tr[8];
  int ok = U2 && Q && Fx && Fy && Emf;
  for (int a = 0; a < 8; a++) {
    tr[a] = (double*)calloc(n, sizeof(double));
    ok = ok && tr[a];
  }
  if (!ok)
    return 1;
  double t = *t_io, dt = 0.0;
  int it = *it_io;
  eko_make_boundaries(p, U);
  memcpy(U2, U, sizeof(double) * n);
  const int it0 = it;
  while (!(t >= (tEnd - 1e-14) || it >= nStepmax)) {
    double* in = ((it - it0) % 2 == 0) ? U : U2;
    double* out = ((it - it0) % 2 == 0) ? U2 : U;
    eko_make_boundaries(p, in);
    memcpy(out, in, sizeof(double) * n);
    eko_mhd_convert_to_primitives(p, in, Q);
    dt = p->cfl / eko_mhd_inv_dt(p, Q);
    if (t + dt > tEnd)
      dt = tEnd - t;
    const double dtdx = dt / p->dx, dtdy = dt / p->dy;
    eko_mhd_compute_trace(p, in, Q, tr, dtdx, dtdy);
    eko_mhd_compute_fluxes(p, tr, Fx, Fy);
    eko_mhd_compute_emf(p, tr, Emf);
    eko_mhd_update(p, out, Fx, Fy, Emf, dtdx, dtdy);
    ++it;
    t += dt;
  }
  if ((it - it0) % 2 == 1)
    memcpy(U, U2, sizeof(double) * n);
  *t_io = t;
  *it_io = it;
  if (dt_last)
    *dt_last = dt;
  free(U2);
  free(Q);
  free(Fx);
  free(Fy);
  free(Emf);
  for (int a = 0; a < 8; a++)
    free(tr[a]);
  return 0;
}
