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

/* nsteps 2D MHD steps with a fixed dt (the step of godunov_unsplit_impl after compute_dt) */
int eko_mhd_run_dt(const eko_params* p, double* U, int nsteps, double dt)
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
  for (int s = 0; s < nsteps; s++) {
    double* in = (s % 2 == 0) ? U : U2;
    double* out = (s % 2 == 0) ? U2 : U;
    eko_make_boundaries(p, in);
    memcpy(out, in, sizeof(double) * n);
    eko_mhd_convert_to_primitives(p, in, Q);
    const double dtdx = dt / p->dx, dtdy = dt / p->dy;
    eko_mhd_compute_trace(p, in, Q, tr, dtdx, dtdy);
    eko_mhd_compute_fluxes(p, tr, Fx, Fy);
    eko_mhd_compute_emf(p, tr, Emf);
    eko_mhd_update(p, out, Fx, Fy, Emf, dtdx, dtdy);
  }
  if (nsteps % 2 == 1)
    memcpy(U, U2, sizeof(double) * n);
  free(U2);
  free(Q);
  free(Fx);
  free(Fy);
  free(Emf);
  for (int a = 0; a < 8; a++)
    free(tr[a]);
  return 0;
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

/* ================================================================ 3D MHD (impl 0) */
/* SolverMHDMuscl<3>::godunov_unsplit_impl (src/muscl/SolverMHDMuscl.cpp:462-512). Arrays
 * on the reference layout U(i,j,k,var), ghosts included, gw = 3. */

#define IX3(p, i, j, k, v) (((((long)(i)) * (p)->jsize + (j)) * (p)->ksize + (k)) * NV + (v))
#define IV3(p, i, j, k, c) (((((long)(i)) * (p)->jsize + (j)) * (p)->ksize + (k)) * 3 + (c))

/* ConvertToPrimitivesFunctor3D_MHD (MHDRunFunctors3D.h:88-163) */
void eko_mhd3d_convert_to_primitives(const eko_params* p, const double* U, double* Q)
{
#pragma omp parallel for collapse(2)
  for (int i = 0; i < p->isize - 1; i++)
    for (int j = 0; j < p->jsize - 1; j++)
      for (int k = 0; k < p->ksize - 1; k++) {
        double u[NV], q[NV], mag[3], c;
        for (int v = 0; v < NV; v++)
          u[v] = U[IX3(p, i, j, k, v)];
        mag[0] = U[IX3(p, i + 1, j, k, IA)];
        mag[1] = U[IX3(p, i, j + 1, k, IB)];
        mag[2] = U[IX3(p, i, j, k + 1, IC)];
        eko_mhd_constoprim(p, u, mag, &c, q);
        for (int v = 0; v < NV; v++)
          Q[IX3(p, i, j, k, v)] = q[v];
      }
}

/* ComputeDtFunctor3D_MHD (MHDRunFunctors3D.h:16-87) */
double eko_mhd3d_inv_dt(const eko_params* p, const double* Q)
{
  const int gw = p->gw;
  double invDt = 0.0;
#pragma omp parallel for collapse(2) reduction(max : invDt)
  for (int i = gw; i < p->isize - gw; i++)
    for (int j = gw; j < p->jsize - gw; j++)
      for (int k = gw; k < p->ksize - gw; k++) {
        double q[NV], s[3];
        for (int v = 0; v < NV; v++)
          q[v] = Q[IX3(p, i, j, k, v)];
        eko_params p1 = *p;
        p1.dim = 3;
        eko_mhd_speed_info(&p1, q, s);
        invDt = fmax(invDt, s[0] / p->dx + s[1] / p->dy + s[2] / p->dz);
      }
  return invDt;
}

/* ComputeElecFieldFunctor3D (MHDRunFunctors3D.h:278-362): E(i,j,k,{IX,IY,IZ}) */
void eko_mhd3d_elec_field(const eko_params* p, const double* U, const double* Q, double* E)
{
#pragma omp parallel for collapse(2)
  for (int i = 1; i < p->isize - 1; i++)
    for (int j = 1; j < p->jsize - 1; j++)
      for (int k = 1; k < p->ksize - 1; k++) {
        double u, v, w, A, B, C;
        v = 0.25 * (Q[IX3(p, i, j - 1, k - 1, IV)] + Q[IX3(p, i, j - 1, k, IV)] + Q[IX3(p, i, j, k - 1, IV)] +
                    Q[IX3(p, i, j, k, IV)]);
        w = 0.25 * (Q[IX3(p, i, j - 1, k - 1, IW)] + Q[IX3(p, i, j - 1, k, IW)] + Q[IX3(p, i, j, k - 1, IW)] +
                    Q[IX3(p, i, j, k, IW)]);
        B = HALF * (U[IX3(p, i, j, k - 1, IB)] + U[IX3(p, i, j, k, IB)]);
        C = HALF * (U[IX3(p, i, j - 1, k, IC)] + U[IX3(p, i, j, k, IC)]);
        E[IV3(p, i, j, k, 0)] = v * C - w * B;
        u = 0.25 * (Q[IX3(p, i - 1, j, k - 1, IU)] + Q[IX3(p, i - 1, j, k, IU)] + Q[IX3(p, i, j, k - 1, IU)] +
                    Q[IX3(p, i, j, k, IU)]);
        w = 0.25 * (Q[IX3(p, i - 1, j, k - 1, IW)] + Q[IX3(p, i - 1, j, k, IW)] + Q[IX3(p, i, j, k - 1, IW)] +
                    Q[IX3(p, i, j, k, IW)]);
        A = HALF * (U[IX3(p, i, j, k - 1, IA)] + U[IX3(p, i, j, k, IA)]);
        C = HALF * (U[IX3(p, i - 1, j, k, IC)] + U[IX3(p, i, j, k, IC)]);
        E[IV3(p, i, j, k, 1)] = w * A - u * C;
        u = 0.25 * (Q[IX3(p, i - 1, j - 1, k, IU)] + Q[IX3(p, i - 1, j, k, IU)] + Q[IX3(p, i, j - 1, k, IU)] +
                    Q[IX3(p, i, j, k, IU)]);
        v = 0.25 * (Q[IX3(p, i - 1, j - 1, k, IV)] + Q[IX3(p, i - 1, j, k, IV)] + Q[IX3(p, i, j - 1, k, IV)] +
                    Q[IX3(p, i, j, k, IV)]);
        A = HALF * (U[IX3(p, i, j - 1, k, IA)] + U[IX3(p, i, j, k, IA)]);
        B = HALF * (U[IX3(p, i - 1, j, k, IB)] + U[IX3(p, i, j, k, IB)]);
        E[IV3(p, i, j, k, 2)] = u * B - v * A;
      }
}

/* one limited transverse slope of slope_unsplit_mhd_3d (MHDBaseFunctor3D.h:563-668) */
static double slope_bf3(double xst, double b, double bplus, double bminus)
{
  double dlft, drgt, dcen, dsgn, slop, dlim;
  dlft = xst * (b - bminus);
  drgt = xst * (bplus - b);
  dcen = HALF * (bplus - bminus);
  dsgn = (dcen >= ZERO) ? ONE : -ONE;
  slop = fmin(fabs(dlft), fabs(drgt));
  dlim = slop;
  if ((dlft * drgt) <= ZERO)
    dlim = ZERO;
  return dsgn * fmin(dlim, fabs(dcen));
}

/* ComputeMagSlopesFunctor3D (MHDRunFunctors3D.h:441-538): dA[3], dB[3], dC[3] per cell
 * (x, y, z slopes; the normal one is zero) */
void eko_mhd3d_mag_slopes(const eko_params* p, const double* U, double* dA, double* dB, double* dC)
{
  const double xst = fmin(p->slope_type, 2.0);
#pragma omp parallel for collapse(2)
  for (int i = 1; i < p->isize - 1; i++)
    for (int j = 1; j < p->jsize - 1; j++)
      for (int k = 1; k < p->ksize - 1; k++) {
        const double a = U[IX3(p, i, j, k, IA)], b = U[IX3(p, i, j, k, IB)], c = U[IX3(p, i, j, k, IC)];
        dA[IV3(p, i, j, k, 0)] = 0.0;
        dA[IV3(p, i, j, k, 1)] = slope_bf3(xst, a, U[IX3(p, i, j + 1, k, IA)], U[IX3(p, i, j - 1, k, IA)]);
        dA[IV3(p, i, j, k, 2)] = slope_bf3(xst, a, U[IX3(p, i, j, k + 1, IA)], U[IX3(p, i, j, k - 1, IA)]);
        dB[IV3(p, i, j, k, 0)] = slope_bf3(xst, b, U[IX3(p, i + 1, j, k, IB)], U[IX3(p, i - 1, j, k, IB)]);
        dB[IV3(p, i, j, k, 1)] = 0.0;
        dB[IV3(p, i, j, k, 2)] = slope_bf3(xst, b, U[IX3(p, i, j, k + 1, IB)], U[IX3(p, i, j, k - 1, IB)]);
        dC[IV3(p, i, j, k, 0)] = slope_bf3(xst, c, U[IX3(p, i + 1, j, k, IC)], U[IX3(p, i - 1, j, k, IC)]);
        dC[IV3(p, i, j, k, 1)] = slope_bf3(xst, c, U[IX3(p, i, j + 1, k, IC)], U[IX3(p, i, j - 1, k, IC)]);
        dC[IV3(p, i, j, k, 2)] = 0.0;
      }
}

/* trace_unsplit_mhd_3d_simpler (MHDBaseFunctor3D.h:690-1110). dq: half-step slopes are
 * formed here from the full ones; out: qm[3], qp[3], edges [4][3] = (RT,RB,LT,LB) x (X,Y,Z). */
void eko_mhd3d_trace(const eko_params* p, const double* q, double dq[3][NV], const double bfNb[6],
                     const double dbf[12], const double E[3][2][2], double dtdx, double dtdy, double dtdz,
                     double qm[3][NV], double qp[3][NV], double qE[4][3][NV])
{
  const double gamma = p->gamma0, smallR = p->smallr, smallp = p->smallp;
  const double ELL = E[0][0][0], ELR = E[0][0][1], ERL = E[0][1][0], ERR = E[0][1][1];
  const double FLL = E[1][0][0], FLR = E[1][0][1], FRL = E[1][1][0], FRR = E[1][1][1];
  const double GLL = E[2][0][0], GLR = E[2][0][1], GRL = E[2][1][0], GRR = E[2][1][1];
  double r = q[ID], pr = q[IP], u = q[IU], v = q[IV], w = q[IW], A = q[IA], B = q[IB], C = q[IC];
  double AL = bfNb[0], AR = bfNb[1], BL = bfNb[2], BR = bfNb[3], CL = bfNb[4], CR = bfNb[5];
  double drx = dq[0][ID] * HALF, dpx = dq[0][IP] * HALF, dux = dq[0][IU] * HALF, dvx = dq[0][IV] * HALF;
  double dwx = dq[0][IW] * HALF, dCx = dq[0][IC] * HALF, dBx = dq[0][IB] * HALF;
  double dry = dq[1][ID] * HALF, dpy = dq[1][IP] * HALF, duy = dq[1][IU] * HALF, dvy = dq[1][IV] * HALF;
  double dwy = dq[1][IW] * HALF, dCy = dq[1][IC] * HALF, dAy = dq[1][IA] * HALF;
  double drz = dq[2][ID] * HALF, dpz = dq[2][IP] * HALF, duz = dq[2][IU] * HALF, dvz = dq[2][IV] * HALF;
  double dwz = dq[2][IW] * HALF, dAz = dq[2][IA] * HALF, dBz = dq[2][IB] * HALF;
  double dALy = HALF * dbf[0], dALz = HALF * dbf[1], dBLx = HALF * dbf[2], dBLz = HALF * dbf[3];
  double dCLx = HALF * dbf[4], dCLy = HALF * dbf[5];
  double dARy = HALF * dbf[6], dARz = HALF * dbf[7], dBRx = HALF * dbf[8], dBRz = HALF * dbf[9];
  double dCRx = HALF * dbf[10], dCRy = HALF * dbf[11];
  double dAx = HALF * (AR - AL), dBy = HALF * (BR - BL), dCz = HALF * (CR - CL);
  double sr0, su0, sv0, sw0, sp0, sA0, sB0, sC0, sAL0, sAR0, sBL0, sBR0, sCL0, sCR0;
  sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy + (-w * drz - dwz * r) * dtdz;
  su0 = (-u * dux - (dpx + B * dBx + C * dCx) / r) * dtdx + (-v * duy + B * dAy / r) * dtdy +
        (-w * duz + C * dAz / r) * dtdz;
  sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - (dpy + A * dAy + C * dCy) / r) * dtdy +
        (-w * dvz + C * dBz / r) * dtdz;
  sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy +
        (-w * dwz - (dpz + A * dAz + B * dBz) / r) * dtdz;
  sp0 = (-u * dpx - dux * gamma * pr) * dtdx + (-v * dpy - dvy * gamma * pr) * dtdy +
        (-w * dpz - dwz * gamma * pr) * dtdz;
  sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy + (u * dCz + C * duz - w * dAz - A * dwz) * dtdz;
  sB0 = (v * dAx + A * dvx - u * dBx - B * dux) * dtdx + (v * dCz + C * dvz - w * dBz - B * dwz) * dtdz;
  sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (w * dBy + B * dwy - v * dCy - C * dvy) * dtdy;
  sAL0 = +(GLR - GLL) * dtdy * HALF - (FLR - FLL) * dtdz * HALF;
  sAR0 = +(GRR - GRL) * dtdy * HALF - (FRR - FRL) * dtdz * HALF;
  sBL0 = -(GRL - GLL) * dtdx * HALF + (ELR - ELL) * dtdz * HALF;
  sBR0 = -(GRR - GLR) * dtdx * HALF + (ERR - ERL) * dtdz * HALF;
  sCL0 = +(FRL - FLL) * dtdx * HALF - (ERL - ELL) * dtdy * HALF;
  sCR0 = +(FRR - FLR) * dtdx * HALF - (ERR - ELR) * dtdy * HALF;
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
  CL = CL + sCL0;
  CR = CR + sCR0;
#define FLOOR3(s)                    \
  do {                               \
    (s)[ID] = fmax(smallR, (s)[ID]); \
    (s)[IP] = fmax(smallp, (s)[IP]); \
  } while (0)
#define SET8(s, a0, a1, a2, a3, a4, a5, a6, a7) \
  do {                                          \
    (s)[ID] = (a0);                             \
    (s)[IU] = (a1);                             \
    (s)[IV] = (a2);                             \
    (s)[IW] = (a3);                             \
    (s)[IP] = (a4);                             \
    (s)[IA] = (a5);                             \
    (s)[IB] = (a6);                             \
    (s)[IC] = (a7);                             \
    FLOOR3(s);                                  \
  } while (0)
  SET8(qp[0], r - drx, u - dux, v - dvx, w - dwx, pr - dpx, AL, B - dBx, C - dCx);
  SET8(qm[0], r + drx, u + dux, v + dvx, w + dwx, pr + dpx, AR, B + dBx, C + dCx);
  SET8(qp[1], r - dry, u - duy, v - dvy, w - dwy, pr - dpy, A - dAy, BL, C - dCy);
  SET8(qm[1], r + dry, u + duy, v + dvy, w + dwy, pr + dpy, A + dAy, BR, C + dCy);
  SET8(qp[2], r - drz, u - duz, v - dvz, w - dwz, pr - dpz, A - dAz, B - dBz, CL);
  SET8(qm[2], r + drz, u + duz, v + dvz, w + dwz, pr + dpz, A + dAz, B + dBz, CR);
  /* X edges */
  SET8(qE[0][0], r + (+dry + drz), u + (+duy + duz), v + (+dvy + dvz), w + (+dwy + dwz), pr + (+dpy + dpz),
       A + (+dAy + dAz), BR + (+dBRz), CR + (+dCRy));
  SET8(qE[1][0], r + (+dry - drz), u + (+duy - duz), v + (+dvy - dvz), w + (+dwy - dwz), pr + (+dpy - dpz),
       A + (+dAy - dAz), BR + (-dBRz), CL + (+dCLy));
  SET8(qE[2][0], r + (-dry + drz), u + (-duy + duz), v + (-dvy + dvz), w + (-dwy + dwz), pr + (-dpy + dpz),
       A + (-dAy + dAz), BL + (+dBLz), CR + (-dCRy));
  SET8(qE[3][0], r + (-dry - drz), u + (-duy - duz), v + (-dvy - dvz), w + (-dwy - dwz), pr + (-dpy - dpz),
       A + (-dAy - dAz), BL + (-dBLz), CL + (-dCLy));
  /* Y edges */
  SET8(qE[0][1], r + (+drx + drz), u + (+dux + duz), v + (+dvx + dvz), w + (+dwx + dwz), pr + (+dpx + dpz),
       AR + (+dARz), B + (+dBx + dBz), CR + (+dCRx));
  SET8(qE[1][1], r + (+drx - drz), u + (+dux - duz), v + (+dvx - dvz), w + (+dwx - dwz), pr + (+dpx - dpz),
       AR + (-dARz), B + (+dBx - dBz), CL + (+dCLx));
  SET8(qE[2][1], r + (-drx + drz), u + (-dux + duz), v + (-dvx + dvz), w + (-dwx + dwz), pr + (-dpx + dpz),
       AL + (+dALz), B + (-dBx + dBz), CR + (-dCRx));
  SET8(qE[3][1], r + (-drx - drz), u + (-dux - duz), v + (-dvx - dvz), w + (-dwx - dwz), pr + (-dpx - dpz),
       AL + (-dALz), B + (-dBx - dBz), CL + (-dCLx));
  /* Z edges */
  SET8(qE[0][2], r + (+drx + dry), u + (+dux + duy), v + (+dvx + dvy), w + (+dwx + dwy), pr + (+dpx + dpy),
       AR + (+dARy), BR + (+dBRx), C + (+dCx + dCy));
  SET8(qE[1][2], r + (+drx - dry), u + (+dux - duy), v + (+dvx - dvy), w + (+dwx - dwy), pr + (+dpx - dpy),
       AR + (-dARy), BL + (+dBLx), C + (+dCx - dCy));
  SET8(qE[2][2], r + (-drx + dry), u + (-dux + duy), v + (-dvx + dvy), w + (-dwx + dwy), pr + (-dpx + dpy),
       AL + (+dALy), BR + (-dBRx), C + (-dCx + dCy));
  SET8(qE[3][2], r + (-drx - dry), u + (-dux - duy), v + (-dvx - dvy), w + (-dwx - dwy), pr + (-dpx - dpy),
       AL + (-dALy), BL + (-dBLx), C + (-dCx - dCy));
#undef SET8
#undef FLOOR3
}

/* compute_emf<EMFY/EMFX> state mapping (RiemannSolvers_MHD.h:736-818): dir 0 = X, 1 = Y, 2 = Z */
double eko_compute_emf_dir(const eko_params* p, double qEdge[4][NV], int dir)
{
  if (dir == 2)
    return eko_compute_emfz(p, qEdge);
  /* (u, v, A, B, w, c) components per direction */
  const int cu = dir == 1 ? IW : IV, cv = dir == 1 ? IU : IW, ca = dir == 1 ? IC : IB, cb = dir == 1 ? IA : IC;
  const int cw = dir == 1 ? IV : IU, cc = dir == 1 ? IB : IA;
  double t[4][NV];
  for (int e = 0; e < 4; e++) {
    t[e][ID] = qEdge[e][ID];
    t[e][IP] = qEdge[e][IP];
    t[e][IU] = qEdge[e][cu];
    t[e][IV] = qEdge[e][cv];
    t[e][IA] = qEdge[e][ca];
    t[e][IB] = qEdge[e][cb];
    t[e][IW] = qEdge[e][cw];
    t[e][IC] = qEdge[e][cc];
  }
  return eko_compute_emfz(p, t);
}

static void swap2(double* s, int a, int b)
{
  double x = s[a];
  s[a] = s[b];
  s[b] = x;
}

/* One 3D MHD step (godunov_unsplit_impl, impl 0) from `in` (ghosts already filled) into
 * `out` (a copy of `in`), with the given dt. */
static int mhd3d_step(const eko_params* p, const double* in, double* out, double dt)
{
  const long ncell = (long)p->isize * p->jsize * p->ksize;
  const int gw = p->gw;
  double* Q = (double*)calloc(ncell * NV, sizeof(double));
  double* E = (double*)calloc(ncell * 3, sizeof(double));
  double* dA = (double*)calloc(ncell * 3, sizeof(double));
  double* dB = (double*)calloc(ncell * 3, sizeof(double));
  double* dC = (double*)calloc(ncell * 3, sizeof(double));
  double* F[3];
  double* T[18]; /* qm x,y,z; qp x,y,z; edges (RT,RB,LT,LB) x (X,Y,Z) */
  double* M = (double*)calloc(ncell * 3, sizeof(double)); /* EMF x, y, z */
  int ok = Q && E && dA && dB && dC && M;
  for (int a = 0; a < 3; a++) {
    F[a] = (double*)calloc(ncell * NV, sizeof(double));
    ok = ok && F[a];
  }
  for (int a = 0; a < 18; a++) {
    T[a] = (double*)calloc(ncell * NV, sizeof(double));
    ok = ok && T[a];
  }
  if (!ok)
    return 1;
  const double dtdx = dt / p->dx, dtdy = dt / p->dy, dtdz = dt / p->dz;
  eko_mhd3d_convert_to_primitives(p, in, Q);
  eko_mhd3d_elec_field(p, in, Q, E);
  eko_mhd3d_mag_slopes(p, in, dA, dB, dC);
  /* ComputeTraceFunctor3D_MHD (MHDRunFunctors3D.h:543-856) */
#pragma omp parallel for collapse(2)
  for (int i = gw - 2; i < p->isize - gw + 1; i++)
    for (int j = gw - 2; j < p->jsize - gw + 1; j++)
      for (int k = gw - 2; k < p->ksize - gw + 1; k++) {
        double q[NV], dq[3][NV], bf[6], dbf[12], El[3][2][2], qm[3][NV], qp[3][NV], qE[4][3][NV];
        for (int v = 0; v < NV; v++) {
          q[v] = Q[IX3(p, i, j, k, v)];
          dq[0][v] = slope_mhd(p, q[v], Q[IX3(p, i + 1, j, k, v)], Q[IX3(p, i - 1, j, k, v)]);
          dq[1][v] = slope_mhd(p, q[v], Q[IX3(p, i, j + 1, k, v)], Q[IX3(p, i, j - 1, k, v)]);
          dq[2][v] = slope_mhd(p, q[v], Q[IX3(p, i, j, k + 1, v)], Q[IX3(p, i, j, k - 1, v)]);
        }
        bf[0] = in[IX3(p, i, j, k, IA)];
        bf[1] = in[IX3(p, i + 1, j, k, IA)];
        bf[2] = in[IX3(p, i, j, k, IB)];
        bf[3] = in[IX3(p, i, j + 1, k, IB)];
        bf[4] = in[IX3(p, i, j, k, IC)];
        bf[5] = in[IX3(p, i, j, k + 1, IC)];
        dbf[0] = dA[IV3(p, i, j, k, 1)];
        dbf[1] = dA[IV3(p, i, j, k, 2)];
        dbf[2] = dB[IV3(p, i, j, k, 0)];
        dbf[3] = dB[IV3(p, i, j, k, 2)];
        dbf[4] = dC[IV3(p, i, j, k, 0)];
        dbf[5] = dC[IV3(p, i, j, k, 1)];
        dbf[6] = dA[IV3(p, i + 1, j, k, 1)];
        dbf[7] = dA[IV3(p, i + 1, j, k, 2)];
        dbf[8] = dB[IV3(p, i, j + 1, k, 0)];
        dbf[9] = dB[IV3(p, i, j + 1, k, 2)];
        dbf[10] = dC[IV3(p, i, j, k + 1, 0)];
        dbf[11] = dC[IV3(p, i, j, k + 1, 1)];
        El[0][0][0] = E[IV3(p, i, j, k, 0)];
        El[0][0][1] = E[IV3(p, i, j, k + 1, 0)];
        El[0][1][0] = E[IV3(p, i, j + 1, k, 0)];
        El[0][1][1] = E[IV3(p, i, j + 1, k + 1, 0)];
        El[1][0][0] = E[IV3(p, i, j, k, 1)];
        El[1][0][1] = E[IV3(p, i, j, k + 1, 1)];
        El[1][1][0] = E[IV3(p, i + 1, j, k, 1)];
        El[1][1][1] = E[IV3(p, i + 1, j, k + 1, 1)];
        El[2][0][0] = E[IV3(p, i, j, k, 2)];
        El[2][0][1] = E[IV3(p, i, j + 1, k, 2)];
        El[2][1][0] = E[IV3(p, i + 1, j, k, 2)];
        El[2][1][1] = E[IV3(p, i + 1, j + 1, k, 2)];
        eko_mhd3d_trace(p, q, dq, bf, dbf, (const double(*)[2][2])El, dtdx, dtdy, dtdz, qm, qp, qE);
        for (int v = 0; v < NV; v++) {
          const long o = IX3(p, i, j, k, v);
          T[0][o] = qm[0][v];
          T[1][o] = qm[1][v];
          T[2][o] = qm[2][v];
          T[3][o] = qp[0][v];
          T[4][o] = qp[1][v];
          T[5][o] = qp[2][v];
          for (int e = 0; e < 4; e++)
            for (int d = 0; d < 3; d++)
              T[6 + 3 * e + d][o] = qE[e][d][v];
        }
      }
  /* ComputeFluxesAndStoreFunctor3D_MHD (MHDRunFunctors3D.h:1783-1909) and
   * ComputeEmfAndStoreFunctor3D (2100-2238) */
#pragma omp parallel for collapse(2)
  for (int i = gw; i < p->isize - gw + 1; i++)
    for (int j = gw; j < p->jsize - gw + 1; j++)
      for (int k = gw; k < p->ksize - gw + 1; k++) {
        double ql[NV], qr[NV], f[NV], qe[4][NV];
        for (int v = 0; v < NV; v++) {
          ql[v] = T[0][IX3(p, i - 1, j, k, v)];
          qr[v] = T[3][IX3(p, i, j, k, v)];
        }
        eko_riemann_mhd(p, ql, qr, f);
        for (int v = 0; v < NV; v++)
          F[0][IX3(p, i, j, k, v)] = f[v];
        for (int v = 0; v < NV; v++) {
          ql[v] = T[1][IX3(p, i, j - 1, k, v)];
          qr[v] = T[4][IX3(p, i, j, k, v)];
        }
        swap2(ql, IU, IV);
        swap2(ql, IA, IB);
        swap2(qr, IU, IV);
        swap2(qr, IA, IB);
        eko_riemann_mhd(p, ql, qr, f);
        for (int v = 0; v < NV; v++)
          F[1][IX3(p, i, j, k, v)] = f[v];
        for (int v = 0; v < NV; v++) {
          ql[v] = T[2][IX3(p, i, j, k - 1, v)];
          qr[v] = T[5][IX3(p, i, j, k, v)];
        }
        swap2(ql, IU, IW);
        swap2(ql, IA, IC);
        swap2(qr, IU, IW);
        swap2(qr, IA, IC);
        eko_riemann_mhd(p, ql, qr, f);
        for (int v = 0; v < NV; v++)
          F[2][IX3(p, i, j, k, v)] = f[v];
        /* EMF z: RT3(i-1,j-1,k), RB3(i-1,j,k), LT3(i,j-1,k), LB3(i,j,k) */
        for (int v = 0; v < NV; v++) {
          qe[IRT][v] = T[6 + 3 * 0 + 2][IX3(p, i - 1, j - 1, k, v)];
          qe[IRB][v] = T[6 + 3 * 1 + 2][IX3(p, i - 1, j, k, v)];
          qe[ILT][v] = T[6 + 3 * 2 + 2][IX3(p, i, j - 1, k, v)];
          qe[ILB][v] = T[6 + 3 * 3 + 2][IX3(p, i, j, k, v)];
        }
        M[IV3(p, i, j, k, 2)] = eko_compute_emf_dir(p, qe, 2);
        /* EMF y: RT2(i-1,j,k-1), LT2(i,j,k-1) -> IRB, RB2(i-1,j,k) -> ILT, LB2(i,j,k) */
        for (int v = 0; v < NV; v++) {
          qe[IRT][v] = T[6 + 3 * 0 + 1][IX3(p, i - 1, j, k - 1, v)];
          qe[IRB][v] = T[6 + 3 * 2 + 1][IX3(p, i, j, k - 1, v)];
          qe[ILT][v] = T[6 + 3 * 1 + 1][IX3(p, i - 1, j, k, v)];
          qe[ILB][v] = T[6 + 3 * 3 + 1][IX3(p, i, j, k, v)];
        }
        M[IV3(p, i, j, k, 1)] = eko_compute_emf_dir(p, qe, 1);
        /* EMF x: RT(i,j-1,k-1), RB(i,j-1,k), LT(i,j,k-1), LB(i,j,k) */
        for (int v = 0; v < NV; v++) {
          qe[IRT][v] = T[6 + 3 * 0 + 0][IX3(p, i, j - 1, k - 1, v)];
          qe[IRB][v] = T[6 + 3 * 1 + 0][IX3(p, i, j - 1, k, v)];
          qe[ILT][v] = T[6 + 3 * 2 + 0][IX3(p, i, j, k - 1, v)];
          qe[ILB][v] = T[6 + 3 * 3 + 0][IX3(p, i, j, k, v)];
        }
        M[IV3(p, i, j, k, 0)] = eko_compute_emf_dir(p, qe, 0);
      }
  /* UpdateFunctor3D_MHD + UpdateEmfFunctor3D (MHDRunFunctors3D.h:2430-2628) */
#pragma omp parallel for collapse(2)
  for (int i = gw; i < p->isize - gw; i++)
    for (int j = gw; j < p->jsize - gw; j++)
      for (int k = gw; k < p->ksize - gw; k++) {
        double ud[NV];
        for (int v = 0; v < NV; v++)
          ud[v] = out[IX3(p, i, j, k, v)];
        static const int hv[5] = { ID, IP, IU, IV, IW };
        static const int my[5] = { ID, IP, IV, IU, IW };
        static const int mz[5] = { ID, IP, IW, IV, IU };
        for (int n = 0; n < 5; n++)
          ud[hv[n]] += F[0][IX3(p, i, j, k, hv[n])] * dtdx;
        for (int n = 0; n < 5; n++)
          ud[hv[n]] -= F[0][IX3(p, i + 1, j, k, hv[n])] * dtdx;
        for (int n = 0; n < 5; n++)
          ud[hv[n]] += F[1][IX3(p, i, j, k, my[n])] * dtdy;
        for (int n = 0; n < 5; n++)
          ud[hv[n]] -= F[1][IX3(p, i, j + 1, k, my[n])] * dtdy;
        for (int n = 0; n < 5; n++)
          ud[hv[n]] += F[2][IX3(p, i, j, k, mz[n])] * dtdz;
        for (int n = 0; n < 5; n++)
          ud[hv[n]] -= F[2][IX3(p, i, j, k + 1, mz[n])] * dtdz;
        ud[IA] += (M[IV3(p, i, j + 1, k, 2)] - M[IV3(p, i, j, k, 2)]) * dtdy;
        ud[IB] -= (M[IV3(p, i + 1, j, k, 2)] - M[IV3(p, i, j, k, 2)]) * dtdx;
        ud[IA] -= (M[IV3(p, i, j, k + 1, 1)] - M[IV3(p, i, j, k, 1)]) * dtdz;
        ud[IB] += (M[IV3(p, i, j, k + 1, 0)] - M[IV3(p, i, j, k, 0)]) * dtdz;
        ud[IC] += (M[IV3(p, i + 1, j, k, 1)] - M[IV3(p, i, j, k, 1)]) * dtdx;
        ud[IC] -= (M[IV3(p, i, j + 1, k, 0)] - M[IV3(p, i, j, k, 0)]) * dtdy;
        for (int v = 0; v < NV; v++)
          out[IX3(p, i, j, k, v)] = ud[v];
      }
  free(Q);
  free(E);
  free(dA);
  free(dB);
  free(dC);
  free(M);
  for (int a = 0; a < 3; a++)
    free(F[a]);
  for (int a = 0; a < 18; a++)
    free(T[a]);
  return 0;
}

/* nsteps 3D MHD steps with a fixed dt */
int eko_mhd3d_run_dt(const eko_params* p, double* U, int nsteps, double dt)
{
  const long n = (long)p->isize * p->jsize * p->ksize * NV;
  double* U2 = (double*)malloc(sizeof(double) * n);
  if (!U2)
    return 1;
  for (int s = 0; s < nsteps; s++) {
    double* in = (s % 2 == 0) ? U : U2;
    double* out = (s % 2 == 0) ? U2 : U;
    eko_make_boundaries(p, in);
    memcpy(out, in, sizeof(double) * n);
    if (mhd3d_step(p, in, out, dt))
      return 1;
  }
  if (nsteps % 2 == 1)
    memcpy(U, U2, sizeof(double) * n);
  free(U2);
  return 0;
}

/* SolverMHDMuscl<3> time loop (dt from the primitives inside each step). */
int eko_mhd3d_run(const eko_params* p, double* U, int nStepmax, double tEnd, double* t_io, int* it_io,
                  double* dt_last)
{
  const long n = (long)p->isize * p->jsize * p->ksize * NV;
  double* U2 = (double*)malloc(sizeof(double) * n);
  double* Q = (double*)calloc(n, sizeof(double));
  if (!U2 || !Q)
    return 1;
  double t = *t_io, dt = 0.0;
  int it = *it_io;
  const int it0 = it;
  while (!(t >= (tEnd - 1e-14) || it >= nStepmax)) {
    double* in = ((it - it0) % 2 == 0) ? U : U2;
    double* out = ((it - it0) % 2 == 0) ? U2 : U;
    eko_make_boundaries(p, in);
    memcpy(out, in, sizeof(double) * n);
    eko_mhd3d_convert_to_primitives(p, in, Q);
    dt = p->cfl / eko_mhd3d_inv_dt(p, Q);
    if (t + dt > tEnd)
      dt = tEnd - t;
    if (mhd3d_step(p, in, out, dt))
      return 1;
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
  return 0;
}
