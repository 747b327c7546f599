/*
 * wxo.c - CPU restatement (plain C, FP64) of Walicxe3D's per-block Godunov
 * update.  TEST INFRASTRUCTURE: the correctness oracle of the GPU library;
 * never linked into the product.  Every routine cites the reference routine
 * it restates; arithmetic follows the reference's evaluation order
 * (Fortran left-to-right, no contraction: build with -ffp-contract=off).
 *
 * Parity status per function (pins are in tests/test_oracle_*.py):
 *   flow2prim, prim2flow, prim2fluxes, limiter, primfhll/hllc/hlle/hlld,
 *   sound/cfast/cfastX, find_coef, calcTemp: pinned by reference golden
 *   vectors (oracle/_ref unit mode) and closed-form checks.
 *   boundary, calcPrims, fluxes+upwind, Powell source, viscosity, timestep,
 *   cooling: pinned by reference trajectories (oracle/_ref run mode) on
 *   uniform, periodic, reflective and AMR meshes.
 *   Two branches the reference leaves undefined are NOT pinned: the HLLD
 *   UR* region with denr == 0 (source/hlld.f90:289-295 assigns vstl.. and
 *   reads the unset vstr..; here vstr.. take the right state) and Riemann
 *   calls that match no branch (NaN input; reference keeps the previous
 *   flux, here the flux is NaN).
 */
#include "wxo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NC 16
#define NB 20                      /* 16 + 2*2 */
#define NCELL (NB * NB * NB)
#define MAXV 8

/* constants.f90 */
enum { LEFT = 1, RIGHT, FRONT, BACK, BOTTOM, TOP };
enum { BC_REFLECTIVE = 1, BC_FREE_FLOW = 2, BC_PERIODIC = 3 };
enum { NEIGH_SAME = 1, NEIGH_COARSER = 2, NEIGH_FINER = 3, NEIGH_BOUNDARY = 4 };
enum { SOLVER_HLL1 = 2, SOLVER_HLL = 3, SOLVER_HLLC = 4, SOLVER_HLLE = 5, SOLVER_HLLD = 6 };
enum { LIM_NOAVG = -1, LIM_NONE, LIM_MINMOD, LIM_VANLEER, LIM_ALBADA, LIM_UMIST, LIM_WOODWARD,
       LIM_SUPERBEE };
static const double AMU = 1.660538782e-24, KB = 1.380650400e-16, PI_REF = 3.14159265358979;

/* index of cell (i,j,k), i,j,k in [-1,18] */
static inline int cid(int i, int j, int k) { return ((k + 1) * NB + (j + 1)) * NB + (i + 1); }

struct wxo_sim {
  wxo_config c;
  double *tab_T, *tab_C, *winds;
  int nroot[3], maxlev, nblk;
  double dx[3][16];
  int *bid, *lev;
  int *sx, *sy, *sz;                 /* siblingCoords (-1 at level 1) */
  int (*ntype)[6];                   /* per face neighbour type */
  int (*nidx)[6][4];                 /* neighbour block indices */
  int (*periodic)[6];                /* periodic partner index, -1 if none */
  double *U, *UP, *PRIM;
  double dt;                         /* dt used by cooling */
};

/* ------------------------------------------------------------------ */
/* hydro_core.f90                                                      */
/* ------------------------------------------------------------------ */

/* flow variables ahead of the passive scalars (neq): 8 with a field, else 5;
   variables nflow..nv-1 are passive scalars (hydro: nv = 5 + npassive) */
static int nflow(const wxo_config *c) { return c->nv == 8 ? 8 : 5; }

/* flow2prim (source/hydro_core.f90:323-378); returns istat */
int wxo_flow2prim(const wxo_config *c, const double *u, double *p) {
  int istat = 0;
  p[0] = u[0];
  p[1] = u[1] / u[0];
  p[2] = u[2] / u[0];
  p[3] = u[3] / u[0];
  p[4] = (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]) / c->cv;
  if (c->nv == 8 && c->mhd)
    p[4] = p[4] - 0.5 * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]) / c->cv;
  if (p[4] < 1.0e-30) { p[4] = 1.0e-30; istat = 1; }
  if (p[0] < 1.0e-40) { p[0] = 1.0e-40; istat = 2; }
  if (c->nv == 8) { p[5] = u[5]; p[6] = u[6]; p[7] = u[7]; }
  for (int q = nflow(c); q < c->nv; q++) p[q] = u[q];   /* passives (:372-374) */
  return istat;
}

/* prim2flow (source/hydro_core.f90:278-313) */
void wxo_prim2flow(const wxo_config *c, const double *p, double *u) {
  u[0] = p[0];
  u[1] = p[0] * p[1];
  u[2] = p[0] * p[2];
  u[3] = p[0] * p[3];
  double v2 = p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
  u[4] = 0.5 * p[0] * v2 + c->cv * p[4];
  if (c->nv == 8) {
    if (c->mhd) u[4] = u[4] + 0.5 * (p[5] * p[5] + p[6] * p[6] + p[7] * p[7]);
    u[5] = p[5]; u[6] = p[6]; u[7] = p[7];
  }
  for (int q = nflow(c); q < c->nv; q++) u[q] = p[q];   /* passives (:307-309) */
}

/* swapxy / swapxz (source/hydro_core.f90:448-496) */
static void swapxy(int nv, double *v) {
  double t = v[1]; v[1] = v[2]; v[2] = t;
  if (nv == 8) { t = v[5]; v[5] = v[6]; v[6] = t; }
}
static void swapxz(int nv, double *v) {
  double t = v[1]; v[1] = v[3]; v[3] = t;
  if (nv == 8) { t = v[5]; v[5] = v[7]; v[7] = t; }
}

/* prim2fluxes (source/hydro_core.f90:387-442); dim 1,2,3 */
void wxo_prim2fluxes(const wxo_config *c, const double *pin, int dim, double *f) {
  double p[MAXV];
  int nv = c->nv;
  memcpy(p, pin, sizeof(double) * nv);
  if (dim == 2) swapxy(nv, p); else if (dim == 3) swapxz(nv, p);
  f[0] = p[0] * p[1];
  f[1] = p[0] * (p[1] * p[1]) + p[4];
  f[2] = p[0] * p[1] * p[2];
  f[3] = p[0] * p[1] * p[3];
  f[4] = p[1] * (0.5 * p[0] * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) + (c->cv + 1.0) * p[4]);
  if (nv == 8) {
    f[5] = 0.0;
    f[6] = p[1] * p[6] - p[2] * p[5];
    f[7] = p[1] * p[7] - p[3] * p[5];
    if (c->mhd) {
      f[1] = f[1] + 0.5 * (p[6] * p[6] + p[7] * p[7] - p[5] * p[5]);
      f[2] = f[2] - p[5] * p[6];
      f[3] = f[3] - p[5] * p[7];
      f[4] = f[4] + p[1] * (p[5] * p[5] + p[6] * p[6] + p[7] * p[7])
             - p[5] * (p[1] * p[5] + p[2] * p[6] + p[3] * p[7]);
    }
  }
  for (int q = nflow(c); q < nv; q++) f[q] = p[1] * p[q];   /* passives (:429-431) */
  if (dim == 2) swapxy(nv, f); else if (dim == 3) swapxz(nv, f);
}

/* sound (source/hydro_core.f90:503-515) */
static double sound(const wxo_config *c, const double *p) { return sqrt(c->gamma * p[4] / p[0]); }

/* cfast (source/hydro_core.f90:530-543) */
static void cfast(const wxo_config *c, double p, double d, double bx, double by, double bz,
                  double *cfx, double *cfy, double *cfz) {
  double g = c->gamma;
  double b2 = bx * bx + by * by + bz * bz;
  double a = g * p + b2;
  *cfx = sqrt(0.5 * (a + sqrt(a * a - 4.0 * g * p * bx * bx)) / d);
  *cfy = sqrt(0.5 * (a + sqrt(a * a - 4.0 * g * p * by * by)) / d);
  *cfz = sqrt(0.5 * (a + sqrt(a * a - 4.0 * g * p * bz * bz)) / d);
}

/* cfastX (source/hydro_core.f90:553-566) */
static double cfastX(const wxo_config *c, const double *p) {
  double b2 = p[5] * p[5] + p[6] * p[6] + p[7] * p[7];
  double cs2va2 = (c->gamma * p[4] + b2) / p[0];
  return sqrt(0.5 * (cs2va2 + sqrt(cs2va2 * cs2va2 - 4.0 * c->gamma * p[4] * (p[5] * p[5]) / p[0] / p[0])));
}

void wxo_speeds(const wxo_config *c, const double *p, double out[5]) {
  out[0] = sound(c, p);
  out[1] = out[2] = out[3] = out[4] = 0.0;
  if (c->nv == 8) {
    cfast(c, p[4], p[0], p[5], p[6], p[7], &out[1], &out[2], &out[3]);
    out[4] = cfastX(c, p);
  }
}

/* limiter / average (source/hydro_core.f90:613-701) */
static double average(double a, double b, int opt) {
  double s, cc, d, r;
  switch (opt) {
    case LIM_NOAVG: return 0.0;
    case LIM_NONE: return 0.5 * (a + b);
    case LIM_VANLEER:
      if (a * b <= 0.0) return 0.0;
      return a * b * (a + b) / (a * a + b * b);
    case LIM_MINMOD:
      s = copysign(1.0, a);
      return s * fmax(0.0, fmin(fabs(a), s * b));
    case LIM_ALBADA: {
      double eps = 1.0e-7;
      return (a * (b * b + eps) + b * (a * a + eps)) / (a * a + b * b * eps);
    }
    case LIM_UMIST:
      s = copysign(1.0, a);
      cc = 0.25 * a + 0.75 * b;
      d = 0.75 * a + 0.25 * b;
      r = fmin(fmin(fmin(2.0 * fabs(a), 2.0 * s * b), s * cc), s * d);
      return s * fmax(0.0, r);
    case LIM_WOODWARD:
      s = copysign(1.0, a);
      cc = 0.5 * (a + b);
      r = fmin(fmin(2.0 * fabs(a), 2.0 * s * b), s * cc);
      return s * fmax(0.0, r);
    case LIM_SUPERBEE:
      s = copysign(1.0, b);
      cc = fmin(2.0 * fabs(b), s * a);
      d = fmin(fabs(b), 2.0 * s * a);
      return s * fmax(fmax(0.0, cc), d);
    default: return 0.0;
  }
}

void wxo_limiter(const double *pll, double *pl, double *pr, const double *prr, int lim, int n) {
  for (int q = 0; q < n; q++) {
    double dl = pl[q] - pll[q], dm = pr[q] - pl[q], dr = prr[q] - pr[q];
    double al = average(dl, dm, lim), ar = average(dm, dr, lim);
    pl[q] = pl[q] + 0.5 * al;
    pr[q] = pr[q] - 0.5 * ar;
  }
}

/* ------------------------------------------------------------------ */
/* Riemann solvers                                                     */
/* ------------------------------------------------------------------ */

/* primfhll + wavespeedHLL (source/hll.f90:175-260) */
static void primfhll(const wxo_config *c, const double *pL, const double *pR, double *ff) {
  int nv = c->nv;
  double cl = sound(c, pL), cr = sound(c, pR);
  double sl = fmin(pL[1] - cl, pR[1] - cr);
  double sr = fmax(pL[1] + cl, pR[1] + cr);
  double fL[MAXV], fR[MAXV], uL[MAXV], uR[MAXV];
  if (sl >= 0) { wxo_prim2fluxes(c, pL, 1, ff); return; }
  if (sr <= 0) { wxo_prim2fluxes(c, pR, 1, ff); return; }
  wxo_prim2fluxes(c, pL, 1, fL);
  wxo_prim2fluxes(c, pR, 1, fR);
  wxo_prim2flow(c, pL, uL);
  wxo_prim2flow(c, pR, uR);
  for (int q = 0; q < nv; q++) ff[q] = (sr * fL[q] - sl * fR[q] + sl * sr * (uR[q] - uL[q])) / (sr - sl);
}

/* primfhlle + wavespeedHLLE (source/hlle.f90:166-241) */
static void primfhlle(const wxo_config *c, const double *pL, const double *pR, double *ff) {
  int nv = c->nv;
  double cfL = cfastX(c, pL), cfR = cfastX(c, pR);
  double sl = fmin(pL[1] - cfL, pR[1] - cfR);
  double sr = fmax(pL[1] + cfL, pR[1] + cfR);
  double fL[MAXV], fR[MAXV], uL[MAXV], uR[MAXV];
  if (sl > 0) { wxo_prim2fluxes(c, pL, 1, ff); return; }
  if (sr < 0) { wxo_prim2fluxes(c, pR, 1, ff); return; }
  wxo_prim2fluxes(c, pL, 1, fL);
  wxo_prim2fluxes(c, pR, 1, fR);
  wxo_prim2flow(c, pL, uL);
  wxo_prim2flow(c, pR, uR);
  for (int q = 0; q < nv; q++) ff[q] = (sr * fL[q] - sl * fR[q] + sl * sr * (uR[q] - uL[q])) / (sr - sl);
}

/* primfhllc + wavespeedHLLC (source/hllc.f90:209-357) */
static void primfhllc(const wxo_config *c, const double *pL, const double *pR, double *ff) {
  int nv = c->nv;
  double dl1 = pL[0], dr1 = pR[0], ul1 = pL[1], ur1 = pR[1], pl1 = pL[4], pr1 = pR[4];
  double cl = sound(c, pL), cr = sound(c, pR);
  double sl = fmin(ul1 - cl, ur1 - cr);
  double sr = fmax(ul1 + cl, ur1 + cr);
  double sst;
  if (c->solver == SOLVER_HLLC)
    sst = (pr1 - pl1 + dl1 * ul1 * (sl - ul1) - dr1 * ur1 * (sr - ur1)) / (dl1 * (sl - ul1) - dr1 * (sr - ur1));
  else
    sst = 0.0;
  double fK[MAXV], uK[MAXV], ust[MAXV];
  if (sl > 0) { wxo_prim2fluxes(c, pL, 1, ff); return; }
  if (sr < 0) { wxo_prim2fluxes(c, pR, 1, ff); return; }
  const double *pK;
  double sK;
  if (sst >= 0) { pK = pL; sK = sl; }
  else if (sst <= 0) { pK = pR; sK = sr; }
  else { for (int q = 0; q < nv; q++) ff[q] = NAN; return; }
  wxo_prim2fluxes(c, pK, 1, fK);
  wxo_prim2flow(c, pK, uK);
  double rhost = pK[0] * (sK - pK[1]) / (sK - sst);
  double EK = 0.5 * pK[0] * (pK[1] * pK[1] + pK[2] * pK[2] + pK[3] * pK[3]) + c->cv * pK[4];
  ust[0] = rhost;
  ust[1] = rhost * sst;
  ust[2] = rhost * pK[2];
  ust[3] = rhost * pK[3];
  ust[4] = rhost * (EK / pK[0] + (sst - pK[1]) * (sst + pK[4] / (pK[0] * (sK - pK[1]))));
  /* passives (hllc.f90:262-264, 283-285); HLLC is hydro only */
  for (int q = 5; q < nv; q++) ust[q] = rhost * pK[q] / pK[0];
  for (int q = 0; q < nv; q++) ff[q] = fK[q] + sK * (ust[q] - uK[q]);
}

/* primfhlld (source/hlld.f90:166-444), Miyoshi & Kusano (2005) */
static void primfhlld(const wxo_config *c, const double *pl, const double *pr, double *ff) {
  double cv = c->cv;
  double csl = cfastX(c, pl), csr = cfastX(c, pr);
  double sr = fmax(pl[1] + csl, pr[1] + csr);
  double sl = fmin(pl[1] - csl, pr[1] - csr);
  if (sl > 0) { wxo_prim2fluxes(c, pl, 1, ff); return; }
  if (sr < 0) { wxo_prim2fluxes(c, pr, 1, ff); return; }

  double bx = 0.5 * (pl[5] + pr[5]);
  double signBx = copysign(1.0, bx);
  double pTL = pl[4] + 0.5 * (bx * bx + pl[6] * pl[6] + pl[7] * pl[7]);
  double pTR = pr[4] + 0.5 * (bx * bx + pr[6] * pr[6] + pr[7] * pr[7]);
  double slmul = sl - pl[1];
  double srmur = sr - pr[1];
  double rholul = pl[0] * pl[1];
  double rhorur = pr[0] * pr[1];
  double sM = (srmur * rhorur - slmul * rholul - pTR + pTL) / (srmur * pr[0] - slmul * pl[0]);
  double srmsM = sr - sM;
  double slmsM = sl - sM;
  double rhostl = pl[0] * slmul / slmsM;
  double rhostr = pr[0] * srmur / srmsM;
  double sstl = sM - fabs(bx) / sqrt(rhostl);
  double sstr = sM + fabs(bx) / sqrt(rhostr);
  double pst = (srmur * pr[0] * pTL - slmul * pl[0] * pTR + pl[0] * pr[0] * srmur * slmul * (pr[1] - pl[1]))
               / (srmur * pr[0] - slmul * pl[0]);
  double vstl, wstl, bystl, bzstl, vstr, wstr, bystr, bzstr, denl, denr, sMmul, sMmur;
  double el, er, vdotbl, vdotbr, vstdotbstl, vstdotbstr, estl, estr;

  if (sstl >= 0) {  /* UL* */
    el = 0.5 * pl[0] * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) + cv * pl[4]
         + 0.5 * (bx * bx + pl[6] * pl[6] + pl[7] * pl[7]);
    sMmul = sM - pl[1];
    denl = pl[0] * slmul * slmsM - bx * bx;
    if (denl == 0) {
      vstl = pl[2]; wstl = pl[3]; bystl = 0.0; bzstl = 0.0;
    } else {
      vstl = pl[2] - bx * pl[6] * sMmul / denl;
      wstl = pl[3] - bx * pl[7] * sMmul / denl;
      bystl = pl[6] * (pl[0] * (slmul * slmul) - bx * bx) / denl;
      bzstl = pl[7] * (pl[0] * (slmul * slmul) - bx * bx) / denl;
    }
    vdotbl = pl[1] * bx + pl[2] * pl[6] + pl[3] * pl[7];
    vstdotbstl = sM * bx + vstl * bystl + wstl * bzstl;
    estl = (slmul * el - pTL * pl[1] + pst * sM + bx * (vdotbl - vstdotbstl)) / slmsM;
    ff[0] = rhostl * sM;
    ff[1] = rhostl * (sM * sM) + pst - bx * bx;
    ff[2] = rhostl * sM * vstl - bx * bystl;
    ff[3] = rhostl * sM * wstl - bx * bzstl;
    ff[4] = sM * (estl + pst) - bx * (vstdotbstl);
    ff[5] = 0.0;
    ff[6] = bystl * sM - bx * vstl;
    ff[7] = bzstl * sM - bx * wstl;
    return;
  }
  if (sstr <= 0) {  /* UR* */
    er = 0.5 * pr[0] * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) + cv * pr[4]
         + 0.5 * (bx * bx + pr[6] * pr[6] + pr[7] * pr[7]);
    sMmur = sM - pr[1];
    denr = pr[0] * srmur * srmsM - bx * bx;
    if (denr == 0) {
      /* unpinned: the reference leaves vstr.. unset here */
      vstr = pr[2]; wstr = pr[3]; bystr = 0.0; bzstr = 0.0;
    } else {
      vstr = pr[2] - bx * pr[6] * sMmur / denr;
      wstr = pr[3] - bx * pr[7] * sMmur / denr;
      bystr = pr[6] * (pr[0] * (srmur * srmur) - bx * bx) / denr;
      bzstr = pr[7] * (pr[0] * (srmur * srmur) - bx * bx) / denr;
    }
    vdotbr = pr[1] * bx + pr[2] * pr[6] + pr[3] * pr[7];
    vstdotbstr = sM * bx + vstr * bystr + wstr * bzstr;
    estr = (srmur * er - pTR * pr[1] + pst * sM + bx * (vdotbr - vstdotbstr)) / srmsM;
    ff[0] = rhostr * sM;
    ff[1] = rhostr * (sM * sM) + pst - bx * bx;
    ff[2] = rhostr * sM * vstr - bx * bystr;
    ff[3] = rhostr * sM * wstr - bx * bzstr;
    ff[4] = sM * (estr + pst) - bx * (vstdotbstr);
    ff[5] = 0.0;
    ff[6] = bystr * sM - bx * vstr;
    ff[7] = bzstr * sM - bx * wstr;
    return;
  }
  sMmul = sM - pl[1];
  sMmur = sM - pr[1];
  denl = pl[0] * slmul * slmsM - bx * bx;
  denr = pr[0] * srmur * srmsM - bx * bx;
  if (denl == 0) {
    vstl = pl[2]; wstl = pl[3]; bystl = 0.0; bzstl = 0.0;
  } else {
    vstl = pl[2] - bx * pl[6] * sMmul / denl;
    wstl = pl[3] - bx * pl[7] * sMmul / denl;
    bystl = pl[6] * (pl[0] * (slmul * slmul) - bx * bx) / denl;
    bzstl = pl[7] * (pl[0] * (slmul * slmul) - bx * bx) / denl;
  }
  if (denr == 0) {
    vstr = pr[2]; wstr = pr[3]; bystr = 0.0; bzstr = 0.0;
  } else {
    vstr = pr[2] - bx * pr[6] * sMmur / denr;
    wstr = pr[3] - bx * pr[7] * sMmur / denr;
    bystr = pr[6] * (pr[0] * (srmur * srmur) - bx * bx) / denr;
    bzstr = pr[7] * (pr[0] * (srmur * srmur) - bx * bx) / denr;
  }
  double dd = sqrt(rhostl) + sqrt(rhostr);
  double vstst = (sqrt(rhostl) * vstl + sqrt(rhostr) * vstr + (bystr - bystl) * signBx) / dd;
  double wstst = (sqrt(rhostl) * wstl + sqrt(rhostr) * wstr + (bzstr - bzstl) * signBx) / dd;
  double bystst = (sqrt(rhostl) * bystr + sqrt(rhostr) * bystl + sqrt(rhostl * rhostr) * (vstr - vstl) * signBx) / dd;
  double bzstst = (sqrt(rhostl) * bzstr + sqrt(rhostr) * bzstl + sqrt(rhostl * rhostr) * (wstr - wstl) * signBx) / dd;
  double vststdotbstst = sM * bx + vstst * bystst + wstst * bzstst;
  if (sM >= 0) {  /* UL** */
    el = 0.5 * pl[0] * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) + cv * pl[4]
         + 0.5 * (bx * bx + pl[6] * pl[6] + pl[7] * pl[7]);
    vdotbl = pl[1] * bx + pl[2] * pl[6] + pl[3] * pl[7];
    vstdotbstl = sM * bx + vstl * bystl + wstl * bzstl;
    estl = (slmul * el - pTL * pl[1] + pst * sM + bx * (vdotbl - vstdotbstl)) / slmsM;
    double eststl = estl - sqrt(rhostl) * (vstdotbstl - vststdotbstst) * signBx;
    ff[0] = rhostl * sM;
    ff[1] = rhostl * (sM * sM) + pst - bx * bx;
    ff[2] = rhostl * sM * vstst - bx * bystst;
    ff[3] = rhostl * sM * wstst - bx * bzstst;
    ff[4] = sM * (eststl + pst) - bx * (vststdotbstst);
    ff[5] = 0.0;
    ff[6] = bystst * sM - bx * vstst;
    ff[7] = bzstst * sM - bx * wstst;
    return;
  }
  if (sM <= 0) {  /* UR** */
    er = 0.5 * pr[0] * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) + cv * pr[4]
         + 0.5 * (bx * bx + pr[6] * pr[6] + pr[7] * pr[7]);
    vdotbr = pr[1] * bx + pr[2] * pr[6] + pr[3] * pr[7];
    vstdotbstr = sM * bx + vstr * bystr + wstr * bzstr;
    estr = (srmur * er - pTR * pr[1] + pst * sM + bx * (vdotbr - vstdotbstr)) / srmsM;
    double eststr = estr + sqrt(rhostr) * (vstdotbstr - vststdotbstst) * signBx;
    ff[0] = rhostr * sM;
    ff[1] = rhostr * (sM * sM) + pst - bx * bx;
    ff[2] = rhostr * sM * vstst - bx * bystst;
    ff[3] = rhostr * sM * wstst - bx * bzstst;
    ff[4] = sM * (eststr + pst) - bx * (vststdotbstst);
    ff[5] = 0.0;
    ff[6] = bystst * sM - bx * vstst;
    ff[7] = bzstst * sM - bx * wstst;
    return;
  }
  for (int q = 0; q < 8; q++) ff[q] = NAN;
}

void wxo_riemann(const wxo_config *c, int solver, const double *pl, const double *pr, double *f) {
  switch (solver) {
    case SOLVER_HLL1: case SOLVER_HLL: primfhll(c, pl, pr, f); break;
    case SOLVER_HLLC: primfhllc(c, pl, pr, f); break;
    case SOLVER_HLLE: primfhlle(c, pl, pr, f); break;
    case SOLVER_HLLD: primfhlld(c, pl, pr, f); break;
    default: for (int q = 0; q < c->nv; q++) f[q] = NAN;
  }
}

/* ------------------------------------------------------------------ */
/* cooling.f90 (COOL_TABLE) and calcTemp                               */
/* ------------------------------------------------------------------ */

/* calcTemp (source/hydro_core.f90:231-270), EOS_ADIABATIC/SINGLE/TWOTEMP,
   and EOS_H_RATE (:258-264: the neutral hydrogen density in the first
   passive scalar; hydro with passives) */
double wxo_calc_temp(const wxo_config *c, const double *p) {
  if (c->eos == 3) return p[4] / (2.0 * p[0] / c->mu0 - p[5]) * c->p_sc / KB;
  double t = p[4] / p[0] * (c->mu0 * AMU * c->p_sc / c->d_sc / KB);
  if (c->eos == 2 && t > c->ion_thres) t = p[4] / p[0] * (c->mui * AMU * c->p_sc / c->d_sc / KB);
  return t;
}

/* find_coef (source/cooling.f90:378-419) */
double wxo_find_coef(const wxo_config *c, double logT) {
  int n = c->ntab;
  const double *T = c->tab_logT, *C = c->tab_loglam;
  if (logT < T[0]) return 0.0;
  if (logT >= T[n - 1]) return pow(10.0, C[n - 1] + 0.5 * (logT - T[n - 1]));
  for (int i = 1; i < n; i++) {
    if (T[i] > logT) {
      double T0 = T[i - 1], C0 = C[i - 1], T1 = T[i], C1 = C[i];
      return pow(10.0, (C1 - C0) / (T1 - T0) * (logT - T0) + C0);
    }
  }
  return 0.0;
}

/* apply_cooling (source/cooling.f90:285-371) over one block */
static void apply_cooling(wxo_sim *s, int b, double dt) {
  const wxo_config *c = &s->c;
  int nv = c->nv;
  double *P = s->PRIM + (size_t)b * nv * NCELL, *U = s->U + (size_t)b * nv * NCELL;
  for (int k = 1; k <= NC; k++)
    for (int j = 1; j <= NC; j++)
      for (int i = 1; i <= NC; i++) {
        int q = cid(i, j, k);
        double pv[MAXV];
        for (int v = 0; v < nv; v++) pv[v] = P[v * NCELL + q];
        double temp = wxo_calc_temp(c, pv);
        double logT = log10(temp);
        if (logT >= c->tab_logT[0]) {
          double ETH = c->cv * pv[4];
          double vel2 = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
          double EK = 0.5 * pv[0] * vel2;
          double aloss = wxo_find_coef(c, logT);
          double numdens = pv[0] * c->d_sc / (c->mui * AMU);
          double ce = aloss * (numdens * numdens) / (ETH * c->e_sc);
          double cool_factor = exp(-ce * (dt * c->t_sc));
          if (cool_factor < 1.0 - c->cooling_limit) cool_factor = 1.0 - c->cooling_limit;
          double new_temp = temp * cool_factor;
          if (new_temp < 10.0) cool_factor = 10.0 / temp;
          P[4 * NCELL + q] = P[4 * NCELL + q] * cool_factor;
          ETH = c->cv * P[4 * NCELL + q];
          U[4 * NCELL + q] = EK + ETH;
        }
      }
}

/* ------------------------------------------------------------------ */
/* Mesh topology (source/amr.f90:439-1098)                             */
/* ------------------------------------------------------------------ */

typedef struct { int *key, *val; int cap; } imap;

static void imap_init(imap *m, int n) {
  m->cap = 1;
  while (m->cap < 2 * n + 2) m->cap <<= 1;
  m->key = (int *)malloc(sizeof(int) * m->cap);
  m->val = (int *)malloc(sizeof(int) * m->cap);
  for (int i = 0; i < m->cap; i++) m->key[i] = -1;
}
static void imap_put(imap *m, int k, int v) {
  unsigned h = ((unsigned)k * 2654435761u) & (m->cap - 1);
  while (m->key[h] != -1 && m->key[h] != k) h = (h + 1) & (m->cap - 1);
  m->key[h] = k; m->val[h] = v;
}
static int imap_get(const imap *m, int k) {
  if (k < 0) return -1;
  unsigned h = ((unsigned)k * 2654435761u) & (m->cap - 1);
  while (m->key[h] != -1) { if (m->key[h] == k) return m->val[h]; h = (h + 1) & (m->cap - 1); }
  return -1;
}

static long offsets(const wxo_sim *s, int level) {  /* amr.f90:663-677 */
  long o = 0, nr = (long)s->nroot[0] * s->nroot[1] * s->nroot[2];
  for (int l = 2; l <= level; l++) o += nr * (1L << (3 * (l - 2)));
  return o;
}
static int meshlevel(const wxo_sim *s, int bid) {  /* amr.f90:637-658 */
  for (int l = 1; l <= s->maxlev; l++) {
    long nr = (long)s->nroot[0] * s->nroot[1] * s->nroot[2] * (1L << (3 * (l - 1)));
    long lo = offsets(s, l) + 1, hi = offsets(s, l) + nr;
    if (bid >= lo && bid <= hi) return l;
  }
  return -1;
}
static void nblev(const wxo_sim *s, int l, int *nx, int *ny, int *nz) {
  *nx = s->nroot[0] << (l - 1); *ny = s->nroot[1] << (l - 1); *nz = s->nroot[2] << (l - 1);
}
static void bcoords(const wxo_sim *s, int bid, int *x, int *y, int *z) {  /* amr.f90:972-997 */
  int l = meshlevel(s, bid), nx, ny, nz;
  nblev(s, l, &nx, &ny, &nz);
  long loc = bid - offsets(s, l) - 1;
  *x = (int)(loc % nx) + 1;
  *y = (int)((loc / nx) % ny) + 1;
  *z = (int)(loc / ((long)nx * ny)) + 1;
}
static int bid_of(const wxo_sim *s, int l, int x, int y, int z) {
  int nx, ny, nz;
  nblev(s, l, &nx, &ny, &nz);
  return (int)(1 + (x - 1) + (long)(y - 1) * nx + (long)(z - 1) * nx * ny + offsets(s, l));
}

/* Returns the number of periodic domain faces whose same-level partner is
   not active: the periodic fill (userBoundary) copies the block across the
   face at the same level, and the reference's own build of it stops there. */
static int build_topology(wxo_sim *s) {
  int unpaired = 0;
  imap m;
  imap_init(&m, s->nblk);
  for (int b = 0; b < s->nblk; b++) imap_put(&m, s->bid[b], b);
  /* children adjacent to a face of the same-level neighbour (amr.f90:598-628);
     child c = 1 + sx + 2 sy + 4 sz */
  static const int fc[7][4] = {{0}, {2, 4, 6, 8}, {1, 3, 5, 7}, {3, 4, 7, 8},
                               {1, 2, 5, 6}, {5, 6, 7, 8}, {1, 2, 3, 4}};
  for (int b = 0; b < s->nblk; b++) {
    int l = s->lev[b], x, y, z, nx, ny, nz;
    bcoords(s, s->bid[b], &x, &y, &z);
    nblev(s, l, &nx, &ny, &nz);
    if (l > 1) {  /* siblingCoords (amr.f90:831-873) */
      s->sx[b] = (x - 1) % 2; s->sy[b] = (y - 1) % 2; s->sz[b] = (z - 1) % 2;
    } else {
      s->sx[b] = s->sy[b] = s->sz[b] = -1;
    }
    for (int f = 1; f <= 6; f++) {
      int xp = x, yp = y, zp = z, edge = 0;
      switch (f) {  /* neighborLevel (amr.f90:439-523) */
        case LEFT: if (x == 1) edge = 1; else xp = x - 1; break;
        case RIGHT: if (x == nx) edge = 1; else xp = x + 1; break;
        case FRONT: if (y == 1) edge = 1; else yp = y - 1; break;
        case BACK: if (y == ny) edge = 1; else yp = y + 1; break;
        case BOTTOM: if (z == 1) edge = 1; else zp = z - 1; break;
        case TOP: if (z == nz) edge = 1; else zp = z + 1; break;
      }
      s->periodic[b][f - 1] = -1;
      for (int q = 0; q < 4; q++) s->nidx[b][f - 1][q] = -1;
      if (edge) {
        s->ntype[b][f - 1] = NEIGH_BOUNDARY;
        if (s->c.bc[f - 1] == BC_PERIODIC) {  /* userBoundary periodic fill */
          int xs = x, ys = y, zs = z;
          switch (f) {
            case LEFT: xs = nx; break; case RIGHT: xs = 1; break;
            case FRONT: ys = ny; break; case BACK: ys = 1; break;
            case BOTTOM: zs = nz; break; case TOP: zs = 1; break;
          }
          s->periodic[b][f - 1] = imap_get(&m, bid_of(s, l, xs, ys, zs));
          if (s->periodic[b][f - 1] < 0) unpaired++;
        }
        continue;
      }
      int nid = bid_of(s, l, xp, yp, zp);
      int hit = imap_get(&m, nid);
      if (hit >= 0) { s->ntype[b][f - 1] = NEIGH_SAME; s->nidx[b][f - 1][0] = hit; continue; }
      if (l > 1) {  /* father (amr.f90:916-961) */
        int fid = bid_of(s, l - 1, (xp + 1) / 2, (yp + 1) / 2, (zp + 1) / 2);
        hit = imap_get(&m, fid);
        if (hit >= 0) { s->ntype[b][f - 1] = NEIGH_COARSER; s->nidx[b][f - 1][0] = hit; continue; }
      }
      s->ntype[b][f - 1] = NEIGH_FINER;  /* children (amr.f90:882-909) */
      for (int q = 0; q < 4; q++) {
        int cnum = fc[f][q] - 1;
        int cx = 2 * xp - 1 + (cnum & 1), cy = 2 * yp - 1 + ((cnum >> 1) & 1), cz = 2 * zp - 1 + (cnum >> 2);
        s->nidx[b][f - 1][q] = imap_get(&m, bid_of(s, l + 1, cx, cy, cz));
      }
    }
  }
  free(m.key); free(m.val);
  return unpaired;
}

/* ------------------------------------------------------------------ */
/* boundary.f90                                                        */
/* ------------------------------------------------------------------ */

/* layerLimits (source/boundary.f90:700-764) */
static void layer_limits(int face, int depth, int ghost, int *i1, int *i2, int *j1, int *j2, int *k1, int *k2) {
  *i1 = 1; *i2 = NC; *j1 = 1; *j2 = NC; *k1 = 1; *k2 = NC;
  switch (face) {
    case LEFT: *i2 = depth; break;
    case RIGHT: *i1 = NC - depth + 1; break;
    case FRONT: *j2 = depth; break;
    case BACK: *j1 = NC - depth + 1; break;
    case BOTTOM: *k2 = depth; break;
    case TOP: *k1 = NC - depth + 1; break;
  }
  if (ghost) {
    switch (face) {
      case LEFT: *i1 -= depth; *i2 -= depth; break;
      case RIGHT: *i1 += depth; *i2 += depth; break;
      case FRONT: *j1 -= depth; *j2 -= depth; break;
      case BACK: *j1 += depth; *j2 += depth; break;
      case BOTTOM: *k1 -= depth; *k2 -= depth; break;
      case TOP: *k1 += depth; *k2 += depth; break;
    }
  }
}
static int opposite(int f) { return (f % 2 == 1) ? f + 1 : f - 1; }

/* cellPos (source/amr.f90:1037-1065) */
static void cell_pos(const wxo_sim *s, int b, int i, int j, int k, double *x, double *y, double *z) {
  int xb, yb, zb, l = s->lev[b];
  bcoords(s, s->bid[b], &xb, &yb, &zb);
  *x = ((xb - 1) * NC + (i - 1) + 0.5) * s->dx[0][l - 1];
  *y = ((yb - 1) * NC + (j - 1) + 0.5) * s->dx[1][l - 1];
  *z = ((zb - 1) * NC + (k - 1) + 0.5) * s->dx[2][l - 1];
}

/* imposePlaneWind (source/winds.f90:250-340): cells k = ncells_z..nzmax
   (16..18), every i and j, of each block without a TOP neighbour
   (neighborLevel gives none at the domain edge, amr.f90:439-500), whatever
   the plane; the plane only picks the velocity's axis and sign
   (winds.f90:300-313).  Hydro only: the reference leaves the field of the
   state unset. */
static void impose_plane_wind(wxo_sim *s, double *A) {
  const wxo_config *c = &s->c;
  int nv = c->nv;
  const double dens = c->pw[0], vel = c->pw[1], temp = c->pw[2], mu = c->pw[3];
  double vx = 0.0, vy = 0.0, vz = 0.0;
  switch (c->pw_plane) {
    case 1: vx = vel; break;
    case 2: vx = -vel; break;
    case 3: vy = vel; break;
    case 4: vy = -vel; break;
    case 5: vz = vel; break;
    default: vz = -vel; break;
  }
  const double pres = dens / (mu * AMU) * KB * temp;
  double pr[MAXV] = {0}, u[MAXV];
  pr[0] = dens / c->d_sc;
  pr[1] = vx / c->v_sc;
  pr[2] = vy / c->v_sc;
  pr[3] = vz / c->v_sc;
  pr[4] = pres / c->p_sc;
  wxo_prim2flow(c, pr, u);
  for (int b = 0; b < s->nblk; b++) {
    int xb, yb, zb, nx, ny, nz;
    bcoords(s, s->bid[b], &xb, &yb, &zb);
    nblev(s, s->lev[b], &nx, &ny, &nz);
    if (zb != nz) continue;
    double *Ab = A + (size_t)b * nv * NCELL;
    for (int i = -1; i <= 18; i++)
      for (int j = -1; j <= 18; j++)
        for (int k = NC; k <= 18; k++)
          for (int v = 0; v < nv; v++) Ab[v * NCELL + cid(i, j, k)] = u[v];
  }
}

/* imposeSphericalWind (source/winds.f90:123-240), all cells of all blocks */
static void impose_winds(wxo_sim *s, double *A) {
  const wxo_config *c = &s->c;
  int nv = c->nv;
  for (int w = 0; w < s->c.nwind; w++) {
    const double *wp = s->winds + 8 * w;
    double xc = wp[0], yc = wp[1], zc = wp[2], radius = wp[3], mdot = wp[4], vinf = wp[5],
           temp = wp[6], mu = wp[7];
    for (int b = 0; b < s->nblk; b++) {
      double *Ab = A + (size_t)b * nv * NCELL;
      for (int i = -1; i <= 18; i++)
        for (int j = -1; j <= 18; j++)
          for (int k = -1; k <= 18; k++) {
            double x, y, z;
            cell_pos(s, b, i, j, k, &x, &y, &z);
            x = x * c->l_sc; y = y * c->l_sc; z = z * c->l_sc;
            double dist = sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc) + (z - zc) * (z - zc));
            if (dist < radius) {
              double dens = mdot / vinf / dist / dist / (4.0 * PI_REF);
              double vx = vinf * (x - xc) / dist + 0.0;
              double vy = vinf * (y - yc) / dist + 0.0;
              double vz = vinf * (z - zc) / dist + 0.0;
              double pres = dens / (mu * AMU) * KB * temp;
              double pr[MAXV] = {0}, u[MAXV];
              pr[0] = dens / c->d_sc;
              pr[1] = vx / c->v_sc;
              pr[2] = vy / c->v_sc;
              pr[3] = vz / c->v_sc;
              pr[4] = pres / c->p_sc;
              wxo_prim2flow(c, pr, u);
              int q = cid(i, j, k);
              for (int v = 0; v < nv; v++) Ab[v * NCELL + q] = u[v];
            }
          }
    }
  }
}

/* boundary (source/boundary.f90:46-64) -> normalBoundary (:74-686), then
   userBoundary: periodic fill (2 layers) and wind sources */
void wxo_boundary(wxo_sim *s, int depth, int which) {
  int nv = s->c.nv;
  double *A = which == 0 ? s->U : s->UP;
  for (int dir = 1; dir <= 6; dir++) {
    for (int b = 0; b < s->nblk; b++) {
      double *D = A + (size_t)b * nv * NCELL;
      int nt = s->ntype[b][dir - 1];
      int i1, i2, j1, j2, k1, k2, i3, i4, j3, j4, k3, k4;
      if (nt == NEIGH_SAME) {  /* :328-368 */
        const double *S = A + (size_t)s->nidx[b][dir - 1][0] * nv * NCELL;
        layer_limits(dir, depth, 1, &i1, &i2, &j1, &j2, &k1, &k2);
        layer_limits(opposite(dir), depth, 0, &i3, &i4, &j3, &j4, &k3, &k4);
        for (int v = 0; v < nv; v++)
          for (int k = k1; k <= k2; k++)
            for (int j = j1; j <= j2; j++)
              for (int i = i1; i <= i2; i++)
                D[v * NCELL + cid(i, j, k)] = S[v * NCELL + cid(i3 + i - i1, j3 + j - j1, k3 + k - k1)];
      } else if (nt == NEIGH_COARSER) {  /* :372-412 */
        const double *S = A + (size_t)s->nidx[b][dir - 1][0] * nv * NCELL;
        layer_limits(dir, depth, 1, &i1, &i2, &j1, &j2, &k1, &k2);
        int sx = s->sx[b], sy = s->sy[b], sz = s->sz[b];
        for (int v = 0; v < nv; v++)
          for (int i = i1; i <= i2; i++)
            for (int j = j1; j <= j2; j++)
              for (int k = k1; k <= k2; k++) {
                int ip = (i + 1) / 2 + sx * NC / 2, jp = (j + 1) / 2 + sy * NC / 2, kp = (k + 1) / 2 + sz * NC / 2;
                switch (dir) {
                  case LEFT: ip = NC + (i + 1) / 2; break;
                  case RIGHT: ip = (i - NC + 1) / 2; break;
                  case FRONT: jp = NC + (j + 1) / 2; break;
                  case BACK: jp = (j - NC + 1) / 2; break;
                  case BOTTOM: kp = NC + (k + 1) / 2; break;
                  case TOP: kp = (k - NC + 1) / 2; break;
                }
                D[v * NCELL + cid(i, j, k)] = S[v * NCELL + cid(ip, jp, kp)];
              }
      } else if (nt == NEIGH_FINER) {  /* :486-540 */
        layer_limits(dir, depth, 1, &i1, &i2, &j1, &j2, &k1, &k2);
        for (int v = 0; v < nv; v++)
          for (int k = k1; k <= k2; k++)
            for (int j = j1; j <= j2; j++)
              for (int i = i1; i <= i2; i++) D[v * NCELL + cid(i, j, k)] = 0.0;
        for (int q = 0; q < 4; q++) {
          int src = s->nidx[b][dir - 1][q];
          const double *S = A + (size_t)src * nv * NCELL;
          int sx = s->sx[src], sy = s->sy[src], sz = s->sz[src];
          layer_limits(opposite(dir), 2 * depth, 0, &i3, &i4, &j3, &j4, &k3, &k4);
          for (int v = 0; v < nv; v++)
            for (int i = i3; i <= i4; i++)
              for (int j = j3; j <= j4; j++)
                for (int k = k3; k <= k4; k++) {
                  int ip = (i + 1) / 2 + sx * NC / 2, jp = (j + 1) / 2 + sy * NC / 2, kp = (k + 1) / 2 + sz * NC / 2;
                  switch (dir) {
                    case LEFT: ip = (i - NC) / 2; break;
                    case RIGHT: ip = (i + 1) / 2 + NC; break;
                    case FRONT: jp = (j - NC) / 2; break;
                    case BACK: jp = (j + 1) / 2 + NC; break;
                    case BOTTOM: kp = (k - NC) / 2; break;
                    case TOP: kp = (k + 1) / 2 + NC; break;
                  }
                  D[v * NCELL + cid(ip, jp, kp)] = D[v * NCELL + cid(ip, jp, kp)] + S[v * NCELL + cid(i, j, k)] / 8.0;
                }
        }
      }
      int bc = s->c.bc[dir - 1];
      if (nt == NEIGH_BOUNDARY && bc != BC_PERIODIC) {  /* :576-668 */
        layer_limits(dir, depth, 1, &i1, &i2, &j1, &j2, &k1, &k2);
        for (int i = i1; i <= i2; i++)
          for (int j = j1; j <= j2; j++)
            for (int k = k1; k <= k2; k++) {
              int ip = i, jp = j, kp = k;
              if (bc == BC_REFLECTIVE) {
                switch (dir) {
                  case LEFT: ip = 1 - i; break;
                  case RIGHT: ip = 2 * NC - i + 1; break;
                  case FRONT: jp = 1 - j; break;
                  case BACK: jp = 2 * NC - j + 1; break;
                  case BOTTOM: kp = 1 - k; break;
                  case TOP: kp = 2 * NC - k + 1; break;
                }
                int flip = (dir <= 2) ? 1 : (dir <= 4) ? 2 : 3;
                for (int v = 0; v < nv; v++) {
                  double val = D[v * NCELL + cid(ip, jp, kp)];
                  D[v * NCELL + cid(i, j, k)] = (v == flip) ? -1.0 * val : val;
                }
              } else if (bc == BC_FREE_FLOW) {
                switch (dir) {
                  case LEFT: ip = 1; break;
                  case RIGHT: ip = 2 * NC - i + 1; break;
                  case FRONT: jp = 1; break;
                  case BACK: jp = 2 * NC - j + 1; break;
                  case BOTTOM: kp = 1; break;
                  case TOP: kp = 2 * NC - k + 1; break;
                }
                for (int v = 0; v < nv; v++) D[v * NCELL + cid(i, j, k)] = D[v * NCELL + cid(ip, jp, kp)];
              }
            }
      }
    }
  }
  /* userBoundary: periodic wrap at the same level, two layers */
  for (int b = 0; b < s->nblk; b++) {
    double *D = A + (size_t)b * nv * NCELL;
    for (int f = 1; f <= 6; f++) {
      int p = s->periodic[b][f - 1];
      if (p < 0) continue;
      const double *S = A + (size_t)p * nv * NCELL;
      for (int g = 1; g <= 2; g++)
        for (int v = 0; v < nv; v++)
          for (int a1 = 1; a1 <= NC; a1++)
            for (int a2 = 1; a2 <= NC; a2++) {
              int dq, sq;
              switch (f) {
                case LEFT: dq = cid(1 - g, a1, a2); sq = cid(NC + 1 - g, a1, a2); break;
                case RIGHT: dq = cid(NC + g, a1, a2); sq = cid(g, a1, a2); break;
                case FRONT: dq = cid(a1, 1 - g, a2); sq = cid(a1, NC + 1 - g, a2); break;
                case BACK: dq = cid(a1, NC + g, a2); sq = cid(a1, g, a2); break;
                case BOTTOM: dq = cid(a1, a2, 1 - g); sq = cid(a1, a2, NC + 1 - g); break;
                default: dq = cid(a1, a2, NC + g); sq = cid(a1, a2, g); break;
              }
              D[v * NCELL + dq] = S[v * NCELL + sq];
            }
    }
  }
  if (s->c.nwind > 0) impose_winds(s, A);
  if (s->c.pw_plane > 0) impose_plane_wind(s, A);
}

/* ------------------------------------------------------------------ */
/* calcPrimsAll / calcPrimsBlock (source/hydro_core.f90:67-223)        */
/* ------------------------------------------------------------------ */
static int prim_cell(wxo_sim *s, const double *Ub, double *Pb, int q) {
  int nv = s->c.nv;
  double u[MAXV], p[MAXV];
  for (int v = 0; v < nv; v++) u[v] = Ub[v * NCELL + q];
  int st = wxo_flow2prim(&s->c, u, p);
  for (int v = 0; v < nv; v++) Pb[v * NCELL + q] = p[v];
  return st != 0;
}

long wxo_calc_prims(wxo_sim *s, int which, int cells) {
  int nv = s->c.nv;
  long bad = 0;
  const double *A = which == 0 ? s->U : s->UP;
  for (int b = 0; b < s->nblk; b++) {
    const double *Ub = A + (size_t)b * nv * NCELL;
    double *Pb = s->PRIM + (size_t)b * nv * NCELL;
    if (cells == 0 || cells == 1)
      for (int k = 1; k <= NC; k++)
        for (int j = 1; j <= NC; j++)
          for (int i = 1; i <= NC; i++) bad += prim_cell(s, Ub, Pb, cid(i, j, k));
    if (cells == 0 || cells == 2) {
      /* every cell outside the interior whose density is non-zero */
      for (int k = -1; k <= 18; k++)
        for (int j = -1; j <= 18; j++)
          for (int i = -1; i <= 18; i++) {
            if (i >= 1 && i <= NC && j >= 1 && j <= NC && k >= 1 && k <= NC) continue;
            int q = cid(i, j, k);
            if (Ub[q] != 0.0) bad += prim_cell(s, Ub, Pb, q);
          }
    }
  }
  return bad;
}

/* ------------------------------------------------------------------ */
/* HLL*fluxes + upwindStep (+ sources) (source/godunov.f90:199-277,    */
/* source/sources.f90:38-133)                                           */
/* ------------------------------------------------------------------ */
static void gather(const double *Pb, int nv, int q, double *out) {
  for (int v = 0; v < nv; v++) out[v] = Pb[v * NCELL + q];
}

/* flux through the interface on the high side of cell (i,j,k) along dim */
static void interface_flux(wxo_sim *s, const double *Pb, int order, int dim, int i, int j, int k, double *ff) {
  const wxo_config *c = &s->c;
  int nv = c->nv;
  int di = dim == 1, dj = dim == 2, dk = dim == 3;
  double pll[MAXV], pl[MAXV], pr[MAXV], prr[MAXV];
  gather(Pb, nv, cid(i, j, k), pl);
  gather(Pb, nv, cid(i + di, j + dj, k + dk), pr);
  if (order == 2) {
    gather(Pb, nv, cid(i - di, j - dj, k - dk), pll);
    gather(Pb, nv, cid(i + 2 * di, j + 2 * dj, k + 2 * dk), prr);
  }
  if (dim == 2) { swapxy(nv, pl); swapxy(nv, pr); if (order == 2) { swapxy(nv, pll); swapxy(nv, prr); } }
  if (dim == 3) { swapxz(nv, pl); swapxz(nv, pr); if (order == 2) { swapxz(nv, pll); swapxz(nv, prr); } }
  if (order == 2) wxo_limiter(pll, pl, pr, prr, c->limiter, nv);
  wxo_riemann(c, c->solver, pl, pr, ff);
  if (dim == 2) swapxy(nv, ff);
  if (dim == 3) swapxz(nv, ff);
}

static void block_update(wxo_sim *s, int b, int order, double dtp) {
  const wxo_config *c = &s->c;
  int nv = c->nv, l = s->lev[b];
  const double *Pb = s->PRIM + (size_t)b * nv * NCELL;
  const double *Ub = s->U + (size_t)b * nv * NCELL;
  double *UPb = s->UP + (size_t)b * nv * NCELL;
  /* FC/GC/HC over the interfaces upwindStep reads */
  static __thread double *F = NULL;
  if (!F) F = (double *)malloc(sizeof(double) * 3 * MAXV * 17 * 17 * 17);
#define FX(d, v, i, j, k) F[((((d) * MAXV + (v)) * 17 + (k)) * 17 + (j)) * 17 + (i)]
  double ff[MAXV];
  for (int k = 0; k <= NC; k++)
    for (int j = 0; j <= NC; j++)
      for (int i = 0; i <= NC; i++) {
        if (j >= 1 && k >= 1) { interface_flux(s, Pb, order, 1, i, j, k, ff); for (int v = 0; v < nv; v++) FX(0, v, i, j, k) = ff[v]; }
        if (i >= 1 && k >= 1) { interface_flux(s, Pb, order, 2, i, j, k, ff); for (int v = 0; v < nv; v++) FX(1, v, i, j, k) = ff[v]; }
        if (i >= 1 && j >= 1) { interface_flux(s, Pb, order, 3, i, j, k, ff); for (int v = 0; v < nv; v++) FX(2, v, i, j, k) = ff[v]; }
      }
  double dtdx = dtp / s->dx[0][l - 1], dtdy = dtp / s->dx[1][l - 1], dtdz = dtp / s->dx[2][l - 1];
  for (int k = 1; k <= NC; k++)
    for (int j = 1; j <= NC; j++)
      for (int i = 1; i <= NC; i++) {
        int q = cid(i, j, k);
        for (int v = 0; v < nv; v++)
          UPb[v * NCELL + q] = Ub[v * NCELL + q] + dtdx * (FX(0, v, i - 1, j, k) - FX(0, v, i, j, k))
                               + dtdy * (FX(1, v, i, j - 1, k) - FX(1, v, i, j, k))
                               + dtdz * (FX(2, v, i, j, k - 1) - FX(2, v, i, j, k));
        if ((c->eight_wave && nv == 8) || c->user_src) {
          /* source_function (sources.f90:100-133): s = 0, the user terms
             (get_user_source_terms, :124), then divbcorr_8w_source (:62-95);
             upwindStep adds dtp*s when eight_wave or user_source_terms
             (godunov.f90:266-271) */
          double sv[MAXV] = {0};
          double pp[MAXV];
          gather(Pb, nv, q, pp);
          if (c->user_src) {
            /* a uniform acceleration: rho*g on the momenta, rho*(v.g) on E */
            sv[1] = sv[1] + pp[0] * c->grav[0];
            sv[2] = sv[2] + pp[0] * c->grav[1];
            sv[3] = sv[3] + pp[0] * c->grav[2];
            sv[4] = sv[4] + pp[0] * (pp[1] * c->grav[0] + pp[2] * c->grav[1] + pp[3] * c->grav[2]);
          }
          if (c->eight_wave && nv == 8) {
            double divB = (Pb[5 * NCELL + cid(i + 1, j, k)] - Pb[5 * NCELL + cid(i - 1, j, k)]) / (2.0 * s->dx[0][l - 1])
                          + (Pb[6 * NCELL + cid(i, j + 1, k)] - Pb[6 * NCELL + cid(i, j - 1, k)]) / (2.0 * s->dx[1][l - 1])
                          + (Pb[7 * NCELL + cid(i, j, k + 1)] - Pb[7 * NCELL + cid(i, j, k - 1)]) / (2.0 * s->dx[2][l - 1]);
            sv[1] = sv[1] - divB * pp[5];
            sv[2] = sv[2] - divB * pp[6];
            sv[3] = sv[3] - divB * pp[7];
            sv[4] = sv[4] - divB * (pp[1] * pp[5] + pp[2] * pp[6] + pp[3] * pp[7]);
            sv[5] = sv[5] - divB * pp[1];
            sv[6] = sv[6] - divB * pp[2];
            sv[7] = sv[7] - divB * pp[3];
          }
          for (int v = 0; v < nv; v++) UPb[v * NCELL + q] = UPb[v * NCELL + q] + dtp * sv[v];
        }
      }
#undef FX
}

/* Godunov (source/godunov.f90:43-190) */
void wxo_godunov(wxo_sim *s, int order, double dt) {
  double dtp = order == 2 ? dt / 2.0 : dt;
  wxo_boundary(s, 1, 0);
  wxo_calc_prims(s, 0, 2);
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < s->nblk; b++) block_update(s, b, 1, dtp);
  if (order == 2) {
    wxo_boundary(s, 2, 1);
    wxo_calc_prims(s, 1, 0);
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < s->nblk; b++) block_update(s, b, 2, dt);
  }
}

/* doViscousStep (source/hydro_solver.f90:217-250), without the time update */
void wxo_viscous(wxo_sim *s) {
  int nv = s->c.nv;
  double eta = s->c.visc_eta;
#pragma omp parallel for
  for (int b = 0; b < s->nblk; b++) {
    const double *P = s->UP + (size_t)b * nv * NCELL;
    double *U = s->U + (size_t)b * nv * NCELL;
    for (int k = 1; k <= NC; k++)
      for (int j = 1; j <= NC; j++)
        for (int i = 1; i <= NC; i++)
          for (int v = 0; v < nv; v++) {
            const double *A = P + v * NCELL;
            U[v * NCELL + cid(i, j, k)] =
                A[cid(i, j, k)] + eta * (+A[cid(i + 1, j, k)] + A[cid(i - 1, j, k)] + A[cid(i, j + 1, k)] +
                                         A[cid(i, j - 1, k)] + A[cid(i, j, k + 1)] + A[cid(i, j, k - 1)] -
                                         6.0 * A[cid(i, j, k)]);
          }
  }
}

/* getTimestep (source/hydro_solver.f90:129-211) before ramp and clamp */
double wxo_timestep_raw(const wxo_sim *s) {
  const wxo_config *c = &s->c;
  int nv = c->nv;
  double dt_hydro = 1.0e30;
  for (int b = 0; b < s->nblk; b++) {
    const double *P = s->PRIM + (size_t)b * nv * NCELL;
    int l = s->lev[b];
    double dx = s->dx[0][l - 1], dy = s->dx[1][l - 1], dz = s->dx[2][l - 1];
    for (int i = 1; i <= NC; i++)
      for (int j = 1; j <= NC; j++)
        for (int k = 1; k <= NC; k++) {
          int q = cid(i, j, k);
          double p[MAXV];
          gather(P, nv, q, p);
          if (c->mhd && nv == 8) {
            double cfx, cfy, cfz;
            cfast(c, p[4], p[0], p[5], p[6], p[7], &cfx, &cfy, &cfz);
            dt_hydro = fmin(dt_hydro, dx / (fabs(p[1]) + cfx));
            dt_hydro = fmin(dt_hydro, dy / (fabs(p[2]) + cfy));
            dt_hydro = fmin(dt_hydro, dz / (fabs(p[3]) + cfz));
          } else {
            double cs = sound(c, p);
            dt_hydro = fmin(dt_hydro, dx / (fabs(p[1]) + cs));
            dt_hydro = fmin(dt_hydro, dy / (fabs(p[2]) + cs));
            dt_hydro = fmin(dt_hydro, dz / (fabs(p[3]) + cs));
          }
        }
  }
  if (!c->cooling) return dt_hydro * c->cfl;
  return fmin(dt_hydro, 1.0e30) * c->cfl;
}

void wxo_cooling(wxo_sim *s, double dt) {
  if (!s->c.cooling) return;
#pragma omp parallel for
  for (int b = 0; b < s->nblk; b++) apply_cooling(s, b, dt);
}

/* one main-loop iteration (source/main.f90:81-100 with hydroSolver,
   source/hydro_solver.f90:39-121) */
long wxo_step(wxo_sim *s, int *it, double *dt, double *time, int *nextout, double dtout, double t_sc) {
  *it += 1;
  double dtl = wxo_timestep_raw(s);
  if (*it <= 10) dtl = pow(2.0, -(double)(11 - *it)) * dtl;
  double tout = (double)(*nextout) * dtout / t_sc;
  int dump = 0;
  if (*time + dtl >= tout) { dtl = tout - *time; dump = 1; }
  *dt = dtl;
  wxo_godunov(s, s->c.solver == SOLVER_HLL1 ? 1 : 2, dtl);
  *time = *time + dtl;
  wxo_viscous(s);
  long bad = wxo_calc_prims(s, 0, 0);
  wxo_cooling(s, dtl);
  if (dump) *nextout += 1;
  return bad;
}

/* ------------------------------------------------------------------ */
wxo_sim *wxo_create(const wxo_config *cfg, const int nroot[3], int maxlev, const double *dxyz, int nblk,
                    const int *bids) {
  wxo_sim *s = (wxo_sim *)calloc(1, sizeof(wxo_sim));
  s->c = *cfg;
  if (cfg->ntab > 0) {
    s->tab_T = (double *)malloc(sizeof(double) * cfg->ntab);
    s->tab_C = (double *)malloc(sizeof(double) * cfg->ntab);
    memcpy(s->tab_T, cfg->tab_logT, sizeof(double) * cfg->ntab);
    memcpy(s->tab_C, cfg->tab_loglam, sizeof(double) * cfg->ntab);
    s->c.tab_logT = s->tab_T;
    s->c.tab_loglam = s->tab_C;
  }
  if (cfg->nwind > 0) {
    s->winds = (double *)malloc(sizeof(double) * 8 * cfg->nwind);
    memcpy(s->winds, cfg->winds, sizeof(double) * 8 * cfg->nwind);
    s->c.winds = s->winds;
  }
  memcpy(s->nroot, nroot, sizeof(int) * 3);
  s->maxlev = maxlev;
  for (int d = 0; d < 3; d++)
    for (int l = 0; l < maxlev; l++) s->dx[d][l] = dxyz[d * maxlev + l];
  s->nblk = nblk;
  s->bid = (int *)malloc(sizeof(int) * nblk);
  s->lev = (int *)malloc(sizeof(int) * nblk);
  s->sx = (int *)malloc(sizeof(int) * nblk);
  s->sy = (int *)malloc(sizeof(int) * nblk);
  s->sz = (int *)malloc(sizeof(int) * nblk);
  s->ntype = malloc(sizeof(*s->ntype) * nblk);
  s->nidx = malloc(sizeof(*s->nidx) * nblk);
  s->periodic = malloc(sizeof(*s->periodic) * nblk);
  for (int b = 0; b < nblk; b++) { s->bid[b] = bids[b]; s->lev[b] = meshlevel(s, bids[b]); }
  if (build_topology(s) != 0) {  /* a periodic face without a same-level partner */
    wxo_destroy(s);
    return NULL;
  }
  size_t n = (size_t)nblk * cfg->nv * NCELL;
  s->U = (double *)calloc(n, sizeof(double));
  s->UP = (double *)calloc(n, sizeof(double));
  s->PRIM = (double *)calloc(n, sizeof(double));
  return s;
}

void wxo_destroy(wxo_sim *s) {
  if (!s) return;
  free(s->tab_T); free(s->tab_C); free(s->winds);
  free(s->bid); free(s->lev); free(s->sx); free(s->sy); free(s->sz);
  free(s->ntype); free(s->nidx); free(s->periodic);
  free(s->U); free(s->UP); free(s->PRIM);
  free(s);
}

void wxo_set_state(wxo_sim *s, const double *u) {
  int nv = s->c.nv;
  size_t n = (size_t)s->nblk * nv * NCELL;
  memset(s->U, 0, n * sizeof(double));
  for (int b = 0; b < s->nblk; b++)
    for (int v = 0; v < nv; v++)
      for (int k = 1; k <= NC; k++)
        for (int j = 1; j <= NC; j++)
          for (int i = 1; i <= NC; i++)
            s->U[((size_t)b * nv + v) * NCELL + cid(i, j, k)] =
                u[((((size_t)b * nv + v) * NC + (k - 1)) * NC + (j - 1)) * NC + (i - 1)];
  wxo_calc_prims(s, 0, 0);
}

void wxo_get_state(const wxo_sim *s, double *u, double *p) {
  int nv = s->c.nv;
  for (int b = 0; b < s->nblk; b++)
    for (int v = 0; v < nv; v++)
      for (int k = 1; k <= NC; k++)
        for (int j = 1; j <= NC; j++)
          for (int i = 1; i <= NC; i++) {
            size_t o = ((((size_t)b * nv + v) * NC + (k - 1)) * NC + (j - 1)) * NC + (i - 1);
            size_t q = ((size_t)b * nv + v) * NCELL + cid(i, j, k);
            if (u) u[o] = s->U[q];
            if (p) p[o] = s->PRIM[q];
          }
}

double *wxo_array(wxo_sim *s, int which) { return which == 0 ? s->U : which == 1 ? s->UP : s->PRIM; }

void wxo_set_state_carry(wxo_sim *s, const double *u, const double *prim, const int *recompute) {
  int nv = s->c.nv;
  size_t n = (size_t)s->nblk * nv * NCELL;
  memset(s->U, 0, n * sizeof(double));
  memset(s->PRIM, 0, n * sizeof(double));
  for (int b = 0; b < s->nblk; b++) {
    for (int v = 0; v < nv; v++)
      for (int k = 1; k <= NC; k++)
        for (int j = 1; j <= NC; j++)
          for (int i = 1; i <= NC; i++) {
            size_t o = ((((size_t)b * nv + v) * NC + (k - 1)) * NC + (j - 1)) * NC + (i - 1);
            size_t q = ((size_t)b * nv + v) * NCELL + cid(i, j, k);
            s->U[q] = u[o];
            if (!recompute[b]) s->PRIM[q] = prim[o];
          }
    if (recompute[b]) {
      /* calcPrimsBlock(..., CELLS_PHYS) on the new block */
      const double *Ub = s->U + (size_t)b * nv * NCELL;
      double *Pb = s->PRIM + (size_t)b * nv * NCELL;
      for (int k = 1; k <= NC; k++)
        for (int j = 1; j <= NC; j++)
          for (int i = 1; i <= NC; i++) prim_cell(s, Ub, Pb, cid(i, j, k));
    }
  }
}

/* ------------------------------------------------------------------ */
/* markByPhysical (source/amr.f90:142-230), literally: cells in the    */
/* reference's loop order (i outer, k inner), the pressure test before */
/* the density test of each cell, early return on the first cell at or */
/* above the refinement threshold.  maxgrad is the running maximum at  */
/* the point of return (all cells when none triggers).                 */
/* ------------------------------------------------------------------ */
static double rel_grad(const double *A, int i, int j, int k, double two_dx) {
  int q = cid(i, j, k);
  double gradx = fabs(A[cid(i + 1, j, k)] - A[cid(i - 1, j, k)]) / A[q] / two_dx;
  double grady = fabs(A[cid(i, j + 1, k)] - A[cid(i, j - 1, k)]) / A[q] / two_dx;
  double gradz = fabs(A[cid(i, j, k + 1)] - A[cid(i, j, k - 1)]) / A[q] / two_dx;
  double g = gradx > grady ? gradx : grady;
  return g > gradz ? g : gradz;
}

static int mark_block(const wxo_sim *s, int b, double refine_thres, double coarse_thres, double *maxgrad_out) {
  int nv = s->c.nv, ilev = s->lev[b];
  const double *P5 = s->PRIM + ((size_t)b * nv + 4) * NCELL;
  const double *P1 = s->PRIM + ((size_t)b * nv + 0) * NCELL;
  double two_dx = 2 * s->dx[0][ilev - 1];
  double maxgrad = 0.0;
  for (int i = 1; i <= NC; i++)
    for (int j = 1; j <= NC; j++)
      for (int k = 1; k <= NC; k++) {
        for (int pass = 0; pass < 2; pass++) {
          double grad = rel_grad(pass == 0 ? P5 : P1, i, j, k, two_dx);
          if (grad > maxgrad) maxgrad = grad;
          if (grad >= refine_thres) {
            *maxgrad_out = maxgrad;
            return ilev == s->maxlev ? 0 /* FLAG_NONE */ : 1 /* FLAG_REFINE */;
          }
        }
      }
  *maxgrad_out = maxgrad;
  return (maxgrad <= coarse_thres && ilev > 1) ? 2 /* FLAG_COARSE */ : 0;
}

void wxo_mark(const wxo_sim *s, double refine_thres, double coarse_thres, int *flags, double *maxgrad) {
  for (int b = 0; b < s->nblk; b++) {
    double m;
    flags[b] = mark_block(s, b, refine_thres, coarse_thres, &m);
    if (maxgrad) maxgrad[b] = m;
  }
}
