/*
 * ek_oracle.c -- CPU restatement of euler_kokkos' hydro MUSCL-Hancock step.
 *
 * TEST INFRASTRUCTURE: this is the oracle the HIP product path is checked against,
 * never the product. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * It follows the reference's impl-0 algorithm (src/muscl/SolverHydroMuscl.cpp:214-390)
 * function by function, on the reference's layout (Kokkos OpenMP LayoutRight with
 * ghosts), in the reference's operation order. Build with -ffp-contract=off.
 * Paths cited below are relative to the euler_kokkos source tree.
 */
#include "ek_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HALF 0.5
#define ONE 1.0
#define ZERO 0.0

int eko_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* HydroParams::init (src/shared/HydroParams.cpp:326-359) */
void eko_params_init(eko_params* p)
{
  p->imin = 0;
  p->jmin = 0;
  p->kmin = 0;
  p->imax = p->nx - 1 + 2 * p->gw;
  p->jmax = p->ny - 1 + 2 * p->gw;
  p->kmax = p->nz - 1 + 2 * p->gw;
  p->isize = p->imax - p->imin + 1;
  p->jsize = p->jmax - p->jmin + 1;
  p->ksize = p->kmax - p->kmin + 1;
  p->dx = (p->xmax - p->xmin) / p->nx;
  p->dy = (p->ymax - p->ymin) / p->ny;
  p->dz = (p->zmax - p->zmin) / p->nz;
  p->smallp = p->smallc * p->smallc / p->gamma0;
  p->smallpp = p->smallr * p->smallp;
  p->gamma6 = (p->gamma0 + ONE) / (2.0 * p->gamma0);
}

/* Kokkos LayoutRight offset of View(i,j[,k],var) (src/shared/kokkos_shared.h:13-28) */
long eko_index(const eko_params* p, int i, int j, int k, int var)
{
  if (p->dim == 2)
    return ((long)i * p->jsize + j) * p->nbvar + var;
  return (((long)i * p->jsize + j) * p->ksize + k) * p->nbvar + var;
}

#define IDX2(p, i, j, v) ((((long)(i)) * (p)->jsize + (j)) * (p)->nbvar + (v))
#define IDX3(p, i, j, k, v) (((((long)(i)) * (p)->jsize + (j)) * (p)->ksize + (k)) * (p)->nbvar + (v))

/* computePrimitives: src/muscl/HydroBaseFunctor2D.h:79-105, HydroBaseFunctor3D.h:79-107 */
void eko_compute_primitives(const eko_params* p, const double* u, double* c, double* q)
{
  const double gamma0 = p->gamma0, smallr = p->smallr, smallp = p->smallp;
  double d, pr, ux, uy, uz = 0.0, eken, e;
  d = fmax(u[EKO_ID], p->smallr);
  (void)smallr;
  ux = u[EKO_IU] / d;
  uy = u[EKO_IV] / d;
  if (p->dim == 3) {
    uz = u[EKO_IW] / d;
    eken = HALF * (ux * ux + uy * uy + uz * uz);
  } else {
    eken = HALF * (ux * ux + uy * uy);
  }
  e = u[EKO_IP] / d - eken;
  pr = fmax((gamma0 - 1.0) * d * e, d * smallp);
  *c = sqrt(gamma0 * (pr) / d);
  q[EKO_ID] = d;
  q[EKO_IP] = pr;
  q[EKO_IU] = ux;
  q[EKO_IV] = uy;
  if (p->dim == 3)
    q[EKO_IW] = uz;
}

/* one direction of slope_unsplit_hydro_{2d,3d}_scalar
 * (src/muscl/HydroBaseFunctor2D.h:423-457, HydroBaseFunctor3D.h:413-461) */
void eko_slope_scalar(const eko_params* p, double q, double qPlus, double qMinus, double* dq)
{
  const double slope_type = p->slope_type;
  double dlft, drgt, dcen, dsgn, slop, dlim;
  dlft = slope_type * (q - qMinus);
  drgt = slope_type * (qPlus - q);
  dcen = HALF * (qPlus - qMinus);
  dsgn = (dcen >= ZERO) ? ONE : -ONE;
  slop = fmin(fabs(dlft), fabs(drgt));
  dlim = slop;
  if ((dlft * drgt) <= ZERO)
    dlim = ZERO;
  *dq = dsgn * fmin(dlim, fabs(dcen));
}

/* slope_unsplit_hydro_{2d,3d}: src/muscl/HydroBaseFunctor2D.h:477-518, 3D.h:484-578.
 * nb[0..5] = x+, x-, y+, y-, z+, z- neighbour states. */
static void slopes_cell(const eko_params* p, const double* q, const double* const nb[6], double* dqX,
                        double* dqY, double* dqZ)
{
  const int nv = p->dim == 3 ? 5 : 4;
  if (p->slope_type == 0) {
    for (int v = 0; v < nv; v++) {
      dqX[v] = ZERO;
      dqY[v] = ZERO;
      if (p->dim == 3)
        dqZ[v] = ZERO;
    }
    return;
  }
  if (p->slope_type == 1 || p->slope_type == 2) {
    for (int v = 0; v < nv; v++) {
      eko_slope_scalar(p, q[v], nb[0][v], nb[1][v], &dqX[v]);
      eko_slope_scalar(p, q[v], nb[2][v], nb[3][v], &dqY[v]);
      if (p->dim == 3)
        eko_slope_scalar(p, q[v], nb[4][v], nb[5][v], &dqZ[v]);
    }
  }
}

/* trace_unsplit_{2d,3d}_along_dir: src/muscl/HydroBaseFunctor2D.h:218-293, 3D.h:276-389.
 * faceId: 0 XMIN, 1 XMAX, 2 YMIN, 3 YMAX, 4 ZMIN, 5 ZMAX (src/shared/enums.h:129-139) */
void eko_trace_along_dir(const eko_params* p, const double* q, const double* dqX, const double* dqY,
                         const double* dqZ, double dtdx, double dtdy, double dtdz, int faceId,
                         double* qface)
{
  const double gamma0 = p->gamma0, smallr = p->smallr;
  const double r = q[EKO_ID], pr = q[EKO_IP], u = q[EKO_IU], v = q[EKO_IV];
  const double drx = dqX[EKO_ID], dpx = dqX[EKO_IP], dux = dqX[EKO_IU], dvx = dqX[EKO_IV];
  const double dry = dqY[EKO_ID], dpy = dqY[EKO_IP], duy = dqY[EKO_IU], dvy = dqY[EKO_IV];
  if (p->dim == 2) {
    double sr0 = -u * drx - v * dry - (dux + dvy) * r;
    double sp0 = -u * dpx - v * dpy - (dux + dvy) * gamma0 * pr;
    double su0 = -u * dux - v * duy - (dpx) / r;
    double sv0 = -u * dvx - v * dvy - (dpy) / r;
    if (faceId == 0) {
      qface[EKO_ID] = r - HALF * drx + sr0 * dtdx * HALF;
      qface[EKO_IP] = pr - HALF * dpx + sp0 * dtdx * HALF;
      qface[EKO_IU] = u - HALF * dux + su0 * dtdx * HALF;
      qface[EKO_IV] = v - HALF * dvx + sv0 * dtdx * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    if (faceId == 1) {
      qface[EKO_ID] = r + HALF * drx + sr0 * dtdx * HALF;
      qface[EKO_IP] = pr + HALF * dpx + sp0 * dtdx * HALF;
      qface[EKO_IU] = u + HALF * dux + su0 * dtdx * HALF;
      qface[EKO_IV] = v + HALF * dvx + sv0 * dtdx * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    if (faceId == 2) {
      qface[EKO_ID] = r - HALF * dry + sr0 * dtdy * HALF;
      qface[EKO_IP] = pr - HALF * dpy + sp0 * dtdy * HALF;
      qface[EKO_IU] = u - HALF * duy + su0 * dtdy * HALF;
      qface[EKO_IV] = v - HALF * dvy + sv0 * dtdy * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    if (faceId == 3) {
      qface[EKO_ID] = r + HALF * dry + sr0 * dtdy * HALF;
      qface[EKO_IP] = pr + HALF * dpy + sp0 * dtdy * HALF;
      qface[EKO_IU] = u + HALF * duy + su0 * dtdy * HALF;
      qface[EKO_IV] = v + HALF * dvy + sv0 * dtdy * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    return;
  }
  {
    const double w = q[EKO_IW];
    const double dwx = dqX[EKO_IW], dwy = dqY[EKO_IW];
    const double drz = dqZ[EKO_ID], dpz = dqZ[EKO_IP], duz = dqZ[EKO_IU], dvz = dqZ[EKO_IV],
                 dwz = dqZ[EKO_IW];
    double sr0 = -u * drx - v * dry - w * drz - (dux + dvy + dwz) * r;
    double sp0 = -u * dpx - v * dpy - w * dpz - (dux + dvy + dwz) * gamma0 * pr;
    double su0 = -u * dux - v * duy - w * duz - (dpx) / r;
    double sv0 = -u * dvx - v * dvy - w * dvz - (dpy) / r;
    double sw0 = -u * dwx - v * dwy - w * dwz - (dpz) / r;
    if (faceId == 0) {
      qface[EKO_ID] = r - HALF * drx + sr0 * dtdx * HALF;
      qface[EKO_IP] = pr - HALF * dpx + sp0 * dtdx * HALF;
      qface[EKO_IU] = u - HALF * dux + su0 * dtdx * HALF;
      qface[EKO_IV] = v - HALF * dvx + sv0 * dtdx * HALF;
      qface[EKO_IW] = w - HALF * dwx + sw0 * dtdx * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    if (faceId == 1) {
      qface[EKO_ID] = r + HALF * drx + sr0 * dtdx * HALF;
      qface[EKO_IP] = pr + HALF * dpx + sp0 * dtdx * HALF;
      qface[EKO_IU] = u + HALF * dux + su0 * dtdx * HALF;
      qface[EKO_IV] = v + HALF * dvx + sv0 * dtdx * HALF;
      qface[EKO_IW] = w + HALF * dwx + sw0 * dtdx * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    if (faceId == 2) {
      qface[EKO_ID] = r - HALF * dry + sr0 * dtdy * HALF;
      qface[EKO_IP] = pr - HALF * dpy + sp0 * dtdy * HALF;
      qface[EKO_IU] = u - HALF * duy + su0 * dtdy * HALF;
      qface[EKO_IV] = v - HALF * dvy + sv0 * dtdy * HALF;
      qface[EKO_IW] = w - HALF * dwy + sw0 * dtdy * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    if (faceId == 3) {
      qface[EKO_ID] = r + HALF * dry + sr0 * dtdy * HALF;
      qface[EKO_IP] = pr + HALF * dpy + sp0 * dtdy * HALF;
      qface[EKO_IU] = u + HALF * duy + su0 * dtdy * HALF;
      qface[EKO_IV] = v + HALF * dvy + sv0 * dtdy * HALF;
      qface[EKO_IW] = w + HALF * dwy + sw0 * dtdy * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    if (faceId == 4) {
      qface[EKO_ID] = r - HALF * drz + sr0 * dtdz * HALF;
      qface[EKO_IP] = pr - HALF * dpz + sp0 * dtdz * HALF;
      qface[EKO_IU] = u - HALF * duz + su0 * dtdz * HALF;
      qface[EKO_IV] = v - HALF * dvz + sv0 * dtdz * HALF;
      qface[EKO_IW] = w - HALF * dwz + sw0 * dtdz * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
    if (faceId == 5) {
      qface[EKO_ID] = r + HALF * drz + sr0 * dtdz * HALF;
      qface[EKO_IP] = pr + HALF * dpz + sp0 * dtdz * HALF;
      qface[EKO_IU] = u + HALF * duz + su0 * dtdz * HALF;
      qface[EKO_IV] = v + HALF * dvz + sv0 * dtdz * HALF;
      qface[EKO_IW] = w + HALF * dwz + sw0 * dtdz * HALF;
      qface[EKO_ID] = fmax(smallr, qface[EKO_ID]);
    }
  }
}

/* cmpflx: src/shared/RiemannSolvers.h:19-48 */
static void cmpflx(const eko_params* p, const double* qgdnv, double* flux)
{
  const int is3d = p->dim == 3;
  flux[EKO_ID] = qgdnv[EKO_ID] * qgdnv[EKO_IU];
  flux[EKO_IU] = flux[EKO_ID] * qgdnv[EKO_IU] + qgdnv[EKO_IP];
  flux[EKO_IV] = flux[EKO_ID] * qgdnv[EKO_IV];
  if (is3d)
    flux[EKO_IW] = flux[EKO_ID] * qgdnv[EKO_IW];
  double entho = ONE / (p->gamma0 - ONE);
  double ekin = HALF * qgdnv[EKO_ID] * (qgdnv[EKO_IU] * qgdnv[EKO_IU] + qgdnv[EKO_IV] * qgdnv[EKO_IV]);
  if (is3d)
    ekin += HALF * qgdnv[EKO_ID] * (qgdnv[EKO_IW] * qgdnv[EKO_IW]);
  double etot = qgdnv[EKO_IP] * entho + ekin;
  flux[EKO_IP] = qgdnv[EKO_IU] * (etot + qgdnv[EKO_IP]);
}

/* riemann_approx: src/shared/RiemannSolvers.h:59-199 */
static void riemann_approx(const eko_params* p, const double* qleft, const double* qright,
                           double* flux)
{
  const double gamma0 = p->gamma0, gamma6 = p->gamma6, smallr = p->smallr, smallc = p->smallc,
               smallp = p->smallp, smallpp = p->smallpp;
  const int is3d = p->dim == 3;
  double qgdnv[5];
  double rl = fmax(qleft[EKO_ID], smallr);
  double ul = qleft[EKO_IU];
  double pl = fmax(qleft[EKO_IP], rl * smallp);
  double rr = fmax(qright[EKO_ID], smallr);
  double ur = qright[EKO_IU];
  double pr = fmax(qright[EKO_IP], rr * smallp);
  double cl = gamma0 * pl * rl;
  double cr = gamma0 * pr * rr;
  double wl = sqrt(cl);
  double wr = sqrt(cr);
  double pstar = fmax(((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr), (double)ZERO);
  double pold = pstar;
  double conv = ONE;
  for (int iter = 0; (iter < 10) && (conv > 1e-6); ++iter) {
    double wwl = sqrt(cl * (ONE + gamma6 * (pold - pl) / pl));
    double wwr = sqrt(cr * (ONE + gamma6 * (pold - pr) / pr));
    double ql = 2.0f * wwl * wwl * wwl / (wwl * wwl + cl);
    double qr = 2.0f * wwr * wwr * wwr / (wwr * wwr + cr);
    double usl = ul - (pold - pl) / wwl;
    double usr = ur + (pold - pr) / wwr;
    double delp = fmax(qr * ql / (qr + ql) * (usl - usr), -pold);
    pold = pold + delp;
    conv = fabs(delp / (pold + smallpp));
  }
  pstar = pold;
  wl = sqrt(cl * (ONE + gamma6 * (pstar - pl) / pl));
  wr = sqrt(cr * (ONE + gamma6 * (pstar - pr) / pr));
  double ustar = HALF * (ul + (pl - pstar) / wl + ur - (pr - pstar) / wr);
  double sgnm = copysign(ONE, ustar);
  double ro, uo, po, wo;
  if (sgnm > ZERO) {
    ro = rl;
    uo = ul;
    po = pl;
    wo = wl;
  } else {
    ro = rr;
    uo = ur;
    po = pr;
    wo = wr;
  }
  double co = fmax(smallc, sqrt(fabs(gamma0 * po / ro)));
  double rstar = fmax((double)(ro / (ONE + ro * (po - pstar) / (wo * wo))), (double)(smallr));
  double cstar = fmax(smallc, sqrt(fabs(gamma0 * pstar / rstar)));
  double spout = co - sgnm * uo;
  double spin = cstar - sgnm * ustar;
  double ushock = wo / ro - sgnm * uo;
  if (pstar >= po) {
    spin = ushock;
    spout = ushock;
  }
  double scr = fmax(spout - spin, smallc + fabs(spout + spin));
  double frac = HALF * (ONE + (spout + spin) / scr);
  if (frac != frac)
    frac = 0.0;
  else
    frac = frac >= 1.0 ? 1.0 : frac <= 0.0 ? 0.0 : frac;
  qgdnv[EKO_ID] = frac * rstar + (ONE - frac) * ro;
  qgdnv[EKO_IU] = frac * ustar + (ONE - frac) * uo;
  qgdnv[EKO_IP] = frac * pstar + (ONE - frac) * po;
  if (spout < ZERO) {
    qgdnv[EKO_ID] = ro;
    qgdnv[EKO_IU] = uo;
    qgdnv[EKO_IP] = po;
  }
  if (spin > ZERO) {
    qgdnv[EKO_ID] = rstar;
    qgdnv[EKO_IU] = ustar;
    qgdnv[EKO_IP] = pstar;
  }
  if (sgnm > ZERO) {
    qgdnv[EKO_IV] = qleft[EKO_IV];
    if (is3d)
      qgdnv[EKO_IW] = qleft[EKO_IW];
  } else {
    qgdnv[EKO_IV] = qright[EKO_IV];
    if (is3d)
      qgdnv[EKO_IW] = qright[EKO_IW];
  }
  cmpflx(p, qgdnv, flux);
}

/* riemann_llf: src/shared/RiemannSolvers.h:217-327 */
static void riemann_llf(const eko_params* p, const double* qleft, const double* qright, double* flux)
{
  const double gamma0 = p->gamma0, smallr = p->smallr, smallp = p->smallp;
  const int is3d = p->dim == 3;
  const double entho = ONE / (gamma0 - ONE);
  double rl = fmax(qleft[EKO_ID], smallr);
  double ul = qleft[EKO_IU];
  double pl = fmax(qleft[EKO_IP], rl * smallp);
  double rr = fmax(qright[EKO_ID], smallr);
  double ur = qright[EKO_IU];
  double pr = fmax(qright[EKO_IP], rr * smallp);
  double cl = sqrt(gamma0 * pl / rl);
  double cr = sqrt(gamma0 * pr / rr);
  double cmax = fmax(fabs(ul) + cl, fabs(ur) + cr);
  double uleft[5], uright[5], fleft[5], fright[5];
  uleft[EKO_ID] = qleft[EKO_ID];
  uright[EKO_ID] = qright[EKO_ID];
  uleft[EKO_IP] = qleft[EKO_IP] * entho + HALF * qleft[EKO_ID] * qleft[EKO_IU] * qleft[EKO_IU];
  uright[EKO_IP] = qright[EKO_IP] * entho + HALF * qright[EKO_ID] * qright[EKO_IU] * qright[EKO_IU];
  uleft[EKO_IP] += HALF * qleft[EKO_ID] * qleft[EKO_IV] * qleft[EKO_IV];
  uright[EKO_IP] += HALF * qright[EKO_ID] * qright[EKO_IV] * qright[EKO_IV];
  if (is3d) {
    uleft[EKO_IP] += HALF * qleft[EKO_ID] * qleft[EKO_IW] * qleft[EKO_IW];
    uright[EKO_IP] += HALF * qright[EKO_ID] * qright[EKO_IW] * qright[EKO_IW];
  }
  uleft[EKO_IU] = qleft[EKO_ID] * qleft[EKO_IU];
  uright[EKO_IU] = qright[EKO_ID] * qright[EKO_IU];
  uleft[EKO_IV] = qleft[EKO_ID] * qleft[EKO_IV];
  uright[EKO_IV] = qright[EKO_ID] * qright[EKO_IV];
  if (is3d) {
    uleft[EKO_IW] = qleft[EKO_ID] * qleft[EKO_IW];
    uright[EKO_IW] = qright[EKO_ID] * qright[EKO_IW];
  }
  fleft[EKO_ID] = uleft[EKO_ID] * qleft[EKO_IU];
  fright[EKO_ID] = uright[EKO_ID] * qright[EKO_IU];
  fleft[EKO_IP] = qleft[EKO_IU] * (uleft[EKO_IP] + qleft[EKO_IP]);
  fright[EKO_IP] = qright[EKO_IU] * (uright[EKO_IP] + qright[EKO_IP]);
  fleft[EKO_IU] = qleft[EKO_IP] + uleft[EKO_IU] * qleft[EKO_IU];
  fright[EKO_IU] = qright[EKO_IP] + uright[EKO_IU] * qright[EKO_IU];
  fleft[EKO_IV] = fleft[EKO_ID] * qleft[EKO_IV];
  fright[EKO_IV] = fright[EKO_ID] * qright[EKO_IV];
  if (is3d) {
    fleft[EKO_IW] = fleft[EKO_ID] * qleft[EKO_IW];
    fright[EKO_IW] = fright[EKO_ID] * qright[EKO_IW];
  }
  for (int nVar = 0; nVar < 4; nVar++)
    flux[nVar] = HALF * (fleft[nVar] + fright[nVar] - cmax * (uright[nVar] - uleft[nVar]));
  if (is3d)
    flux[EKO_IW] = HALF * (fleft[EKO_IW] + fright[EKO_IW] - cmax * (uright[EKO_IW] - uleft[EKO_IW]));
}

/* riemann_hll (hydro, 2D only through riemann_hydro): src/shared/RiemannSolvers.h:343-435 */
static void riemann_hll(const eko_params* p, const double* qleft, const double* qright, double* flux)
{
  const double gamma0 = p->gamma0, smallr = p->smallr, smallp = p->smallp;
  const double entho = ONE / (gamma0 - ONE);
  double rl = fmax(qleft[EKO_ID], smallr);
  double ul = qleft[EKO_IU];
  double pl = fmax(qleft[EKO_IP], rl * smallp);
  double rr = fmax(qright[EKO_ID], smallr);
  double ur = qright[EKO_IU];
  double pr = fmax(qright[EKO_IP], rr * smallp);
  double cl = sqrt(gamma0 * pl / rl);
  double cr = sqrt(gamma0 * pr / rr);
  double SL = fmin(fmin(ul, ur) - fmax(cl, cr), (double)ZERO);
  double SR = fmax(fmax(ul, ur) + fmax(cl, cr), (double)ZERO);
  double uleft[4], uright[4], fleft[4], fright[4];
  uleft[EKO_ID] = qleft[EKO_ID];
  uright[EKO_ID] = qright[EKO_ID];
  uleft[EKO_IP] = qleft[EKO_IP] * entho + HALF * qleft[EKO_ID] * qleft[EKO_IU] * qleft[EKO_IU];
  uright[EKO_IP] = qright[EKO_IP] * entho + HALF * qright[EKO_ID] * qright[EKO_IU] * qright[EKO_IU];
  uleft[EKO_IP] += HALF * qleft[EKO_ID] * qleft[EKO_IV] * qleft[EKO_IV];
  uright[EKO_IP] += HALF * qright[EKO_ID] * qright[EKO_IV] * qright[EKO_IV];
  uleft[EKO_IU] = qleft[EKO_ID] * qleft[EKO_IU];
  uright[EKO_IU] = qright[EKO_ID] * qright[EKO_IU];
  uleft[EKO_IV] = qleft[EKO_ID] * qleft[EKO_IV];
  uright[EKO_IV] = qright[EKO_ID] * qright[EKO_IV];
  fleft[EKO_ID] = uleft[EKO_IU];
  fright[EKO_ID] = uright[EKO_IU];
  fleft[EKO_IP] = qleft[EKO_IU] * (uleft[EKO_IP] + qleft[EKO_IP]);
  fright[EKO_IP] = qright[EKO_IU] * (uright[EKO_IP] + qright[EKO_IP]);
  fleft[EKO_IU] = qleft[EKO_IP] + uleft[EKO_IU] * qleft[EKO_IU];
  fright[EKO_IU] = qright[EKO_IP] + uright[EKO_IU] * qright[EKO_IU];
  fleft[EKO_IV] = fleft[EKO_ID] * qleft[EKO_IV];
  fright[EKO_IV] = fright[EKO_ID] * qright[EKO_IV];
  for (int nVar = 0; nVar < 4; nVar++)
    flux[nVar] =
      (SR * fleft[nVar] - SL * fright[nVar] + SR * SL * (uright[nVar] - uleft[nVar])) / (SR - SL);
}

/* riemann_hllc: src/shared/RiemannSolvers.h:445-568 */
static void riemann_hllc(const eko_params* p, const double* qleft, const double* qright,
                         double* flux)
{
  const double gamma0 = p->gamma0, smallr = p->smallr, smallp = p->smallp, smallc = p->smallc;
  const int is3d = p->dim == 3;
  const double entho = ONE / (gamma0 - ONE);
  double rl = fmax(qleft[EKO_ID], smallr);
  double pl = fmax(qleft[EKO_IP], rl * smallp);
  double ul = qleft[EKO_IU];
  double ecinl = HALF * rl * ul * ul;
  ecinl += HALF * rl * qleft[EKO_IV] * qleft[EKO_IV];
  if (is3d)
    ecinl += HALF * rl * qleft[EKO_IW] * qleft[EKO_IW];
  double etotl = pl * entho + ecinl;
  double ptotl = pl;
  double rr = fmax(qright[EKO_ID], smallr);
  double pr = fmax(qright[EKO_IP], rr * smallp);
  double ur = qright[EKO_IU];
  double ecinr = HALF * rr * ur * ur;
  ecinr += HALF * rr * qright[EKO_IV] * qright[EKO_IV];
  if (is3d)
    ecinr += HALF * rr * qright[EKO_IW] * qright[EKO_IW];
  double etotr = pr * entho + ecinr;
  double ptotr = pr;
  double cfastl = sqrt(fmax(gamma0 * pl / rl, smallc * smallc));
  double cfastr = sqrt(fmax(gamma0 * pr / rr, smallc * smallc));
  double SL = fmin(ul, ur) - fmax(cfastl, cfastr);
  double SR = fmax(ul, ur) + fmax(cfastl, cfastr);
  double rcl = rl * (ul - SL);
  double rcr = rr * (SR - ur);
  double ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  double ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);
  double rstarl = rl * (SL - ul) / (SL - ustar);
  double etotstarl = ((SL - ul) * etotl - ptotl * ul + ptotstar * ustar) / (SL - ustar);
  double rstarr = rr * (SR - ur) / (SR - ustar);
  double etotstarr = ((SR - ur) * etotr - ptotr * ur + ptotstar * ustar) / (SR - ustar);
  double ro, uo, ptoto, etoto;
  if (SL > ZERO) {
    ro = rl;
    uo = ul;
    ptoto = ptotl;
    etoto = etotl;
  } else if (ustar > ZERO) {
    ro = rstarl;
    uo = ustar;
    ptoto = ptotstar;
    etoto = etotstarl;
  } else if (SR > ZERO) {
    ro = rstarr;
    uo = ustar;
    ptoto = ptotstar;
    etoto = etotstarr;
  } else {
    ro = rr;
    uo = ur;
    ptoto = ptotr;
    etoto = etotr;
  }
  flux[EKO_ID] = ro * uo;
  flux[EKO_IU] = ro * uo * uo + ptoto;
  flux[EKO_IP] = (etoto + ptoto) * uo;
  if (flux[EKO_ID] > ZERO)
    flux[EKO_IV] = flux[EKO_ID] * qleft[EKO_IV];
  else
    flux[EKO_IV] = flux[EKO_ID] * qright[EKO_IV];
  if (is3d) {
    if (flux[EKO_ID] > ZERO)
      flux[EKO_IW] = flux[EKO_ID] * qleft[EKO_IW];
    else
      flux[EKO_IW] = flux[EKO_ID] * qright[EKO_IW];
  }
}

/* riemann_hydro dispatch: src/shared/RiemannSolvers.h:573-603 (2D), 608-633 (3D; no HLL) */
void eko_riemann_hydro(const eko_params* p, const double* qleft, const double* qright, double* flux)
{
  switch (p->riemann) {
    case EKO_RIEMANN_APPROX: riemann_approx(p, qleft, qright, flux); break;
    case EKO_RIEMANN_HLL:
      if (p->dim == 2)
        riemann_hll(p, qleft, qright, flux);
      break;
    case EKO_RIEMANN_HLLC: riemann_hllc(p, qleft, qright, flux); break;
    case EKO_RIEMANN_LLF: riemann_llf(p, qleft, qright, flux); break;
    default: break;
  }
}

/* ---------------------------------------------------------------- initial conditions */

/* InitImplodeFunctor2D/3D: src/muscl/HydroInitFunctors2D.h:33-122, HydroInitFunctors3D.h:68-168 */
void eko_init_implode(const eko_params* p, const eko_implode* ip, double* U)
{
  const int gw = p->gw;
  const double xmin = p->xmin, xmax = p->xmax, ymin = p->ymin, zmin = p->zmin;
  const double dx = p->dx, dy = p->dy, dz = p->dz, gamma0 = p->gamma0;
  if (p->dim == 2) {
#pragma omp parallel for collapse(2)
    for (int i = 0; i < p->isize; i++)
      for (int j = 0; j < p->jsize; j++) {
        double x = xmin + dx / 2 + (i - gw) * dx;
        double y = ymin + dy / 2 + (j - gw) * dy;
        int tmp;
        if (ip->shape == 1)
          tmp = x + y * y > 0.5 && x + y * y < 1.5;
        else
          tmp = x + y > (xmin + xmax) / 2. + ymin;
        if (tmp) {
          U[IDX2(p, i, j, EKO_ID)] = ip->rho_out;
          U[IDX2(p, i, j, EKO_IP)] = ip->p_out / (gamma0 - 1.0) +
                                     0.5 * ip->rho_out * (ip->u_out * ip->u_out + ip->v_out * ip->v_out);
          U[IDX2(p, i, j, EKO_IU)] = ip->u_out;
          U[IDX2(p, i, j, EKO_IV)] = ip->v_out;
        } else {
          U[IDX2(p, i, j, EKO_ID)] = ip->rho_in;
          U[IDX2(p, i, j, EKO_IP)] =
            ip->p_in / (gamma0 - 1.0) + 0.5 * ip->rho_in * (ip->u_in * ip->u_in + ip->v_in * ip->v_in);
          U[IDX2(p, i, j, EKO_IU)] = ip->u_in;
          U[IDX2(p, i, j, EKO_IV)] = ip->v_in;
        }
      }
    return;
  }
#pragma omp parallel for collapse(2)
  for (int i = 0; i < p->isize; i++)
    for (int j = 0; j < p->jsize; j++)
      for (int k = 0; k < p->ksize; k++) {
        double x = xmin + dx / 2 + (i - gw) * dx;
        double y = ymin + dy / 2 + (j - gw) * dy;
        double z = zmin + dz / 2 + (k - gw) * dz;
        int tmp;
        if (ip->shape == 1)
          tmp = x + y + z > 0.5 && x + y + z < 2.5;
        else
          tmp = x + y + z > (xmin + xmax) / 2. + ymin + zmin;
        if (tmp) {
          U[IDX3(p, i, j, k, EKO_ID)] = ip->rho_out;
          U[IDX3(p, i, j, k, EKO_IP)] =
            ip->p_out / (gamma0 - 1.0) +
            0.5 * ip->rho_out * (ip->u_out * ip->u_out + ip->v_out * ip->v_out + ip->w_out * ip->w_out);
          U[IDX3(p, i, j, k, EKO_IU)] = ip->u_out;
          U[IDX3(p, i, j, k, EKO_IV)] = ip->v_out;
          U[IDX3(p, i, j, k, EKO_IW)] = ip->w_out;
        } else {
          U[IDX3(p, i, j, k, EKO_ID)] = ip->rho_in;
          U[IDX3(p, i, j, k, EKO_IP)] =
            ip->p_in / (gamma0 - 1.0) +
            0.5 * ip->rho_in * (ip->u_in * ip->u_in + ip->v_in * ip->v_in + ip->w_in * ip->w_in);
          U[IDX3(p, i, j, k, EKO_IU)] = ip->u_in;
          U[IDX3(p, i, j, k, EKO_IV)] = ip->v_in;
          U[IDX3(p, i, j, k, EKO_IW)] = ip->w_in;
        }
      }
}

/* InitBlastFunctor2D/3D: src/muscl/HydroInitFunctors2D.h:127-209, HydroInitFunctors3D.h:173-265 */
void eko_init_blast(const eko_params* p, const eko_blast* bp, double* U)
{
  const int gw = p->gw;
  const double xmin = p->xmin, ymin = p->ymin, zmin = p->zmin;
  const double dx = p->dx, dy = p->dy, dz = p->dz, gamma0 = p->gamma0;
  const double radius2 = bp->radius * bp->radius;
  if (p->dim == 2) {
#pragma omp parallel for collapse(2)
    for (int i = 0; i < p->isize; i++)
      for (int j = 0; j < p->jsize; j++) {
        double x = xmin + dx / 2 + (i - gw) * dx;
        double y = ymin + dy / 2 + (j - gw) * dy;
        double d2 = (x - bp->center_x) * (x - bp->center_x) + (y - bp->center_y) * (y - bp->center_y);
        if (d2 < radius2) {
          U[IDX2(p, i, j, EKO_ID)] = bp->density_in;
          U[IDX2(p, i, j, EKO_IP)] = bp->pressure_in / (gamma0 - 1.0);
        } else {
          U[IDX2(p, i, j, EKO_ID)] = bp->density_out;
          U[IDX2(p, i, j, EKO_IP)] = bp->pressure_out / (gamma0 - 1.0);
        }
        U[IDX2(p, i, j, EKO_IU)] = 0.0;
        U[IDX2(p, i, j, EKO_IV)] = 0.0;
      }
    return;
  }
#pragma omp parallel for collapse(2)
  for (int i = 0; i < p->isize; i++)
    for (int j = 0; j < p->jsize; j++)
      for (int k = 0; k < p->ksize; k++) {
        double x = xmin + dx / 2 + (i - gw) * dx;
        double y = ymin + dy / 2 + (j - gw) * dy;
        double z = zmin + dz / 2 + (k - gw) * dz;
        double d2 = (x - bp->center_x) * (x - bp->center_x) + (y - bp->center_y) * (y - bp->center_y) +
                    (z - bp->center_z) * (z - bp->center_z);
        if (d2 < radius2) {
          U[IDX3(p, i, j, k, EKO_ID)] = bp->density_in;
          U[IDX3(p, i, j, k, EKO_IP)] = bp->pressure_in / (gamma0 - 1.0);
        } else {
          U[IDX3(p, i, j, k, EKO_ID)] = bp->density_out;
          U[IDX3(p, i, j, k, EKO_IP)] = bp->pressure_out / (gamma0 - 1.0);
        }
        U[IDX3(p, i, j, k, EKO_IU)] = 0.0;
        U[IDX3(p, i, j, k, EKO_IV)] = 0.0;
        U[IDX3(p, i, j, k, EKO_IW)] = 0.0;
      }
}

/* ---------------------------------------------------------------- boundaries */

/* MakeBoundariesFunctor2D/3D<face> (src/shared/BoundariesFunctors.h:17-523) applied in
 * SolverBase::make_boundaries_serial order XMIN, XMAX, YMIN, YMAX[, ZMIN, ZMAX]
 * (src/shared/SolverBase.cpp:489-512). Each face covers the full extent of the other
 * axes (ghosts included), so corners come from the later passes. */
static int bc_src(int bc, int side, int g, int n, int gw, double* sgn_normal)
{
  /* g: ghost index along the axis (0..gw-1 on the min side, n+gw..n+2gw-1 on the max side) */
  *sgn_normal = 1.0;
  if (side == 0) {
    if (bc == EKO_BC_DIRICHLET) {
      *sgn_normal = -1.0;
      return 2 * gw - 1 - g;
    } else if (bc == EKO_BC_NEUMANN)
      return gw;
    return n + g;
  }
  if (bc == EKO_BC_DIRICHLET) {
    *sgn_normal = -1.0;
    return 2 * n + 2 * gw - 1 - g;
  } else if (bc == EKO_BC_NEUMANN)
    return n + gw - 1;
  return g - n;
}

void eko_make_boundaries(const eko_params* p, double* U)
{
  const int gw = p->gw, nv = p->nbvar;
  const int normal_var[3] = { EKO_IU, EKO_IV, EKO_IW };
  /* MHD face-B normal components (src/shared/BoundariesFunctors.h:534-1055) */
  const int normal_bvar[3] = { EKO_IA, EKO_IB, EKO_IC };
  const int ndir = p->dim;
  const int n_axis[3] = { p->nx, p->ny, p->nz };
  const int size_axis[3] = { p->isize, p->jsize, p->ksize };
  for (int dir = 0; dir < ndir; dir++) {
    for (int side = 0; side < 2; side++) {
      const int bc = p->bc[2 * dir + side];
      const int n = n_axis[dir];
      const int o1 = dir == 0 ? 1 : 0;         /* other axes */
      const int o2 = dir == 2 ? 1 : (dir == 0 ? 2 : 2);
      const int s1 = size_axis[o1];
      const int s2 = (ndir == 3) ? size_axis[o2] : 1;
#pragma omp parallel for collapse(2)
      for (int a = 0; a < s1; a++)
        for (int b = 0; b < s2; b++)
          for (int gg = 0; gg < gw; gg++) {
            const int g = side == 0 ? gg : n + gw + gg;
            double sgn;
            const int g0 = bc_src(bc, side, g, n, gw, &sgn);
            int dst[3], src[3];
            dst[dir] = g;
            src[dir] = g0;
            dst[o1] = a;
            src[o1] = a;
            if (ndir == 3) {
              dst[o2] = b;
              src[o2] = b;
            } else {
              dst[2] = 0;
              src[2] = 0;
            }
            for (int iVar = 0; iVar < nv; iVar++) {
              double sign = 1.0;
              if (iVar == normal_var[dir] || (p->mhd && iVar == normal_bvar[dir]))
                sign = sgn;
              U[eko_index(p, dst[0], dst[1], dst[2], iVar)] =
                U[eko_index(p, src[0], src[1], src[2], iVar)] * sign;
            }
          }
    }
  }
}

/* ---------------------------------------------------------------- dt */

/* ComputeDtFunctor2D/3D (src/muscl/HydroRunFunctors2D.h:18-82, HydroRunFunctors3D.h:16-87):
 * max over interior cells of (c+|u|)/dx + (c+|v|)/dy [+ (c+|w|)/dz]. */
double eko_compute_inv_dt(const eko_params* p, const double* U)
{
  const int gw = p->gw;
  double invDt = 0.0;
  if (p->dim == 2) {
#pragma omp parallel for collapse(2) reduction(max : invDt)
    for (int i = gw; i < p->isize - gw; i++)
      for (int j = gw; j < p->jsize - gw; j++) {
        double u[4], q[4], c = 0.0;
        for (int v = 0; v < 4; v++)
          u[v] = U[IDX2(p, i, j, v)];
        eko_compute_primitives(p, u, &c, q);
        double vx = c + fabs(q[EKO_IU]);
        double vy = c + fabs(q[EKO_IV]);
        invDt = fmax(invDt, vx / p->dx + vy / p->dy);
      }
    return invDt;
  }
#pragma omp parallel for collapse(2) reduction(max : invDt)
  for (int i = gw; i < p->isize - gw; i++)
    for (int j = gw; j < p->jsize - gw; j++)
      for (int k = gw; k < p->ksize - gw; k++) {
        double u[5], q[5], c = 0.0;
        for (int v = 0; v < 5; v++)
          u[v] = U[IDX3(p, i, j, k, v)];
        eko_compute_primitives(p, u, &c, q);
        double vx = c + fabs(q[EKO_IU]);
        double vy = c + fabs(q[EKO_IV]);
        double vz = c + fabs(q[EKO_IW]);
        invDt = fmax(invDt, vx / p->dx + vy / p->dy + vz / p->dz);
      }
  return invDt;
}

/* ---------------------------------------------------------------- step kernels */

/* ConvertToPrimitivesFunctor2D/3D: src/muscl/HydroRunFunctors2D.h:184-246, 3D.h:200-265 */
void eko_convert_to_primitives(const eko_params* p, const double* U, double* Q)
{
  const int nv = p->nbvar;
  const long ncell = (long)p->isize * p->jsize * (p->dim == 3 ? p->ksize : 1);
#pragma omp parallel for
  for (long c = 0; c < ncell; c++) {
    double cs;
    eko_compute_primitives(p, U + c * nv, &cs, Q + c * nv);
  }
}

static void load_q(const eko_params* p, const double* Q, int i, int j, int k, double* q)
{
  const int nv = p->dim == 3 ? 5 : 4;
  const long o = eko_index(p, i, j, k, 0);
  for (int v = 0; v < nv; v++)
    q[v] = Q[o + v];
}

/* slopes of cell (i,j,k) from its 2*dim neighbours, as gathered by
 * ComputeAndStoreFluxesFunctor2D/3D (src/muscl/HydroRunFunctors2D.h:582-643,
 * HydroRunFunctors3D.h:375-476) */
static void slopes_at(const eko_params* p, const double* Q, int i, int j, int k, double* q,
                      double* dqX, double* dqY, double* dqZ)
{
  double n0[5], n1[5], n2[5], n3[5], n4[5] = { 0 }, n5[5] = { 0 };
  load_q(p, Q, i, j, k, q);
  load_q(p, Q, i + 1, j, k, n0);
  load_q(p, Q, i - 1, j, k, n1);
  load_q(p, Q, i, j + 1, k, n2);
  load_q(p, Q, i, j - 1, k, n3);
  if (p->dim == 3) {
    load_q(p, Q, i, j, k + 1, n4);
    load_q(p, Q, i, j, k - 1, n5);
  }
  const double* nb[6] = { n0, n1, n2, n3, n4, n5 };
  slopes_cell(p, q, nb, dqX, dqY, dqZ);
}

static void swap_d(double* a, double* b)
{
  double t = *a;
  *a = *b;
  *b = t;
}

/* ComputeAndStoreFluxesFunctor2D/3D (src/muscl/HydroRunFunctors2D.h:489-768,
 * HydroRunFunctors3D.h:271-717). Gravity disabled. */
void eko_compute_and_store_fluxes(const eko_params* p, const double* Q, double* Fx, double* Fy,
                                  double* Fz, double dt)
{
  const int gw = p->gw;
  const int is3d = p->dim == 3;
  const double dtdx = dt / p->dx, dtdy = dt / p->dy, dtdz = dt / p->dz;
  const int kb = is3d ? gw : 0, ke = is3d ? p->ksize - gw : 0;
#pragma omp parallel for collapse(2)
  for (int i = gw; i <= p->isize - gw; i++)
    for (int j = gw; j <= p->jsize - gw; j++)
      for (int k = kb; k <= ke; k++) {
        double qLoc[5], dqX[5], dqY[5], dqZ[5] = { 0 };
        double qN[5], dqXn[5], dqYn[5], dqZn[5] = { 0 };
        double qleft[5], qright[5], flux[5];
        slopes_at(p, Q, i, j, k, qLoc, dqX, dqY, dqZ);
        const long o = eko_index(p, i, j, k, 0);
        /* X */
        slopes_at(p, Q, i - 1, j, k, qN, dqXn, dqYn, dqZn);
        eko_trace_along_dir(p, qLoc, dqX, dqY, dqZ, dtdx, dtdy, dtdz, 0, qright);
        eko_trace_along_dir(p, qN, dqXn, dqYn, dqZn, dtdx, dtdy, dtdz, 1, qleft);
        eko_riemann_hydro(p, qleft, qright, flux);
        Fx[o + EKO_ID] = flux[EKO_ID] * dtdx;
        Fx[o + EKO_IP] = flux[EKO_IP] * dtdx;
        Fx[o + EKO_IU] = flux[EKO_IU] * dtdx;
        Fx[o + EKO_IV] = flux[EKO_IV] * dtdx;
        if (is3d)
          Fx[o + EKO_IW] = flux[EKO_IW] * dtdx;
        /* Y */
        slopes_at(p, Q, i, j - 1, k, qN, dqXn, dqYn, dqZn);
        eko_trace_along_dir(p, qLoc, dqX, dqY, dqZ, dtdx, dtdy, dtdz, 2, qright);
        eko_trace_along_dir(p, qN, dqXn, dqYn, dqZn, dtdx, dtdy, dtdz, 3, qleft);
        swap_d(&qleft[EKO_IU], &qleft[EKO_IV]);
        swap_d(&qright[EKO_IU], &qright[EKO_IV]);
        eko_riemann_hydro(p, qleft, qright, flux);
        Fy[o + EKO_ID] = flux[EKO_ID] * dtdy;
        Fy[o + EKO_IP] = flux[EKO_IP] * dtdy;
        Fy[o + EKO_IU] = flux[EKO_IV] * dtdy;
        Fy[o + EKO_IV] = flux[EKO_IU] * dtdy;
        if (is3d)
          Fy[o + EKO_IW] = flux[EKO_IW] * dtdy;
        if (!is3d)
          continue;
        /* Z */
        slopes_at(p, Q, i, j, k - 1, qN, dqXn, dqYn, dqZn);
        eko_trace_along_dir(p, qLoc, dqX, dqY, dqZ, dtdx, dtdy, dtdz, 4, qright);
        eko_trace_along_dir(p, qN, dqXn, dqYn, dqZn, dtdx, dtdy, dtdz, 5, qleft);
        swap_d(&qleft[EKO_IU], &qleft[EKO_IW]);
        swap_d(&qright[EKO_IU], &qright[EKO_IW]);
        eko_riemann_hydro(p, qleft, qright, flux);
        Fz[o + EKO_ID] = flux[EKO_ID] * dtdz;
        Fz[o + EKO_IP] = flux[EKO_IP] * dtdz;
        Fz[o + EKO_IU] = flux[EKO_IW] * dtdz;
        Fz[o + EKO_IV] = flux[EKO_IV] * dtdz;
        Fz[o + EKO_IW] = flux[EKO_IU] * dtdz;
      }
}

/* UpdateFunctor2D/3D: src/muscl/HydroRunFunctors2D.h:773-845, HydroRunFunctors3D.h:722-820 */
void eko_update(const eko_params* p, double* U, const double* Fx, const double* Fy, const double* Fz)
{
  const int gw = p->gw;
  const int is3d = p->dim == 3;
  const int nv = is3d ? 5 : 4;
  const int kb = is3d ? gw : 0, ke = is3d ? p->ksize - gw : 1;
#pragma omp parallel for collapse(2)
  for (int i = gw; i < p->isize - gw; i++)
    for (int j = gw; j < p->jsize - gw; j++)
      for (int k = kb; k < ke; k++) {
        const long o = eko_index(p, i, j, k, 0);
        const long ox = eko_index(p, i + 1, j, k, 0);
        const long oy = eko_index(p, i, j + 1, k, 0);
        for (int v = 0; v < nv; v++)
          U[o + v] += Fx[o + v];
        for (int v = 0; v < nv; v++)
          U[o + v] -= Fx[ox + v];
        for (int v = 0; v < nv; v++)
          U[o + v] += Fy[o + v];
        for (int v = 0; v < nv; v++)
          U[o + v] -= Fy[oy + v];
        if (is3d) {
          const long oz = eko_index(p, i, j, k + 1, 0);
          for (int v = 0; v < nv; v++)
            U[o + v] += Fz[o + v];
          for (int v = 0; v < nv; v++)
            U[o + v] -= Fz[oz + v];
        }
      }
}

/* SolverHydroMuscl<dim>::godunov_unsplit_impl, impl 0 (src/muscl/SolverHydroMuscl.cpp:214-253,
 * 304-339): boundaries(in), deep_copy(out,in), prims(in), fluxes, update(out). */
void eko_godunov_unsplit_impl(const eko_params* p, double* Uin, double* Uout, double* Q, double* Fx,
                              double* Fy, double* Fz, double dt)
{
  const long n = (long)p->isize * p->jsize * (p->dim == 3 ? p->ksize : 1) * p->nbvar;
  eko_make_boundaries(p, Uin);
  memcpy(Uout, Uin, sizeof(double) * n);
  eko_convert_to_primitives(p, Uin, Q);
  eko_compute_and_store_fluxes(p, Q, Fx, Fy, Fz, dt);
  eko_update(p, Uout, Fx, Fy, Fz);
}

/* main loop (src/main.cpp:126-131) with SolverBase::finished/next_iteration/compute_dt
 * (src/shared/SolverBase.cpp:124-195) and SolverHydroMuscl::next_iteration_impl /
 * compute_dt_local / godunov_unsplit (src/muscl/SolverHydroMuscl.h:589-700). */
int eko_run(const eko_params* p, double* U, int nStepmax, double tEnd, double* t_io, int* it_io,
            double* dt_last)
{
  const long n = (long)p->isize * p->jsize * (p->dim == 3 ? p->ksize : 1) * p->nbvar;
  double* U2 = (double*)malloc(sizeof(double) * n);
  double* Q = (double*)malloc(sizeof(double) * n);
  double* Fx = (double*)calloc(n, sizeof(double));
  double* Fy = (double*)calloc(n, sizeof(double));
  double* Fz = p->dim == 3 ? (double*)calloc(n, sizeof(double)) : NULL;
  if (!U2 || !Q || !Fx || !Fy || (p->dim == 3 && !Fz))
    return 1;
  double t = *t_io;
  int it = *it_io;
  double dt = 0.0;
  /* constructor: make_boundaries(U); deep_copy(U2,U) (SolverHydroMuscl.h:280-283) */
  eko_make_boundaries(p, U);
  memcpy(U2, U, sizeof(double) * n);
  const int it0 = it;
  while (!(t >= (tEnd - 1e-14) || it >= nStepmax)) {
    double* cur = ((it - it0) % 2 == 0) ? U : U2;
    double* nxt = ((it - it0) % 2 == 0) ? U2 : U;
    dt = p->cfl / eko_compute_inv_dt(p, cur);
    if (t + dt > tEnd)
      dt = tEnd - t;
    eko_godunov_unsplit_impl(p, cur, nxt, Q, Fx, Fy, Fz, dt);
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
  free(Fz);
  return 0;
}
