// wx_physics.h - cell and interface physics of Walicxe3D's Godunov update,
// written once as templates over the arithmetic type R:
//   R = double : the plain form; every expression keeps the reference's
//                evaluation order (Fortran left to right) and the library is
//                built with -ffp-contract=off, so results match the FP64
//                reference bit for bit (IEEE division and square root).
//   R = float  : the bulk arithmetic of the fused form.
// Reference routines are cited per function.
#pragma once

#include <hip/hip_runtime.h>

#include "wx_common.h"

namespace wx {

template <typename R> struct Phys {
  int nv;        // 5 or 8
  int mhd;
  int solver;
  int limiter;
  R gamma, cv;
};

template <typename R> __host__ __device__ inline R rabs(R a) { return a < R(0) ? -a : a; }
__host__ __device__ inline double rsqrt_(double a) { return sqrt(a); }
__host__ __device__ inline float rsqrt_(float a) { return sqrtf(a); }
__host__ __device__ inline double rfmin(double a, double b) { return fmin(a, b); }
__host__ __device__ inline float rfmin(float a, float b) { return fminf(a, b); }
__host__ __device__ inline double rfmax(double a, double b) { return fmax(a, b); }
__host__ __device__ inline float rfmax(float a, float b) { return fmaxf(a, b); }
__host__ __device__ inline double rcopysign(double a, double b) { return copysign(a, b); }
__host__ __device__ inline float rcopysign(float a, float b) { return copysignf(a, b); }

// Division in the MHD wave-speed and HLLD code.  FP64: IEEE division (the
// plain form's bit-exact arithmetic).  FP32 on the device: the hardware
// reciprocal (about 1 ulp) times the numerator, two instructions instead of
// the range-scaled sequence; the FP32 form's fluxes carry FP32 rounding
// anyway.  On the host (tests) FP32 division stays IEEE.
__host__ __device__ inline float rcp32(float b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(b);
#else
  return 1.0f / b;
#endif
}
// Square root in the MHD wave-speed and HLLD code.  FP64: IEEE square
// root.  FP32 on the device: the hardware square root (about 1 ulp), without
// the range scaling that keeps the library's sqrtf exact for arguments
// below 2^-96 (wave speeds and densities of a flow are far above that).
__host__ __device__ inline double msqrt(double a) { return sqrt(a); }
__host__ __device__ inline float msqrt(float a) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_sqrtf(a);
#else
  return sqrtf(a);
#endif
}
__host__ __device__ inline double rdiv(double a, double b) { return a / b; }
__host__ __device__ inline float rdiv(float a, float b) {
#ifdef __HIP_DEVICE_COMPILE__
  return a * rcp32(b);
#else
  return a / b;
#endif
}

// flow2prim (source/hydro_core.f90:323-378); returns istat (0, 1 p floor, 2 rho floor)
template <typename R>
__host__ __device__ inline int flow2prim(const Phys<R> &c, const R *u, R *p) {
  int istat = 0;
  p[0] = u[0];
  p[1] = u[1] / u[0];
  p[2] = u[2] / u[0];
  p[3] = u[3] / u[0];
  p[4] = (u[4] - R(0.5) * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]) / c.cv;
  if (c.nv == 8 && c.mhd) p[4] = p[4] - R(0.5) * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]) / c.cv;
  if (p[4] < R(1.0e-30)) { p[4] = R(1.0e-30); istat = 1; }
  if (p[0] < R(1.0e-40)) { p[0] = R(1.0e-40); istat = 2; }
  if (c.nv == 8) { p[5] = u[5]; p[6] = u[6]; p[7] = u[7]; }
  return istat;
}

// prim2flow (source/hydro_core.f90:278-313)
template <typename R>
__host__ __device__ inline void prim2flow(const Phys<R> &c, const R *p, R *u) {
  u[0] = p[0];
  u[1] = p[0] * p[1];
  u[2] = p[0] * p[2];
  u[3] = p[0] * p[3];
  R v2 = p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
  u[4] = R(0.5) * p[0] * v2 + c.cv * p[4];
  if (c.nv == 8) {
    if (c.mhd) u[4] = u[4] + R(0.5) * (p[5] * p[5] + p[6] * p[6] + p[7] * p[7]);
    u[5] = p[5];
    u[6] = p[6];
    u[7] = p[7];
  }
}

// x-direction physical flux of a primitive state (prim2fluxes with
// dimens = DIM_X, source/hydro_core.f90:387-442).  Callers rotate states
// and fluxes for y and z (swapxy/swapxz, :448-496).
template <typename R>
__host__ __device__ inline void xflux(const Phys<R> &c, const R *p, R *f) {
  f[0] = p[0] * p[1];
  f[1] = p[0] * (p[1] * p[1]) + p[4];
  f[2] = p[0] * p[1] * p[2];
  f[3] = p[0] * p[1] * p[3];
  f[4] = p[1] * (R(0.5) * p[0] * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) + (c.cv + R(1.0)) * p[4]);
  if (c.nv == 8) {
    f[5] = R(0);
    f[6] = p[1] * p[6] - p[2] * p[5];
    f[7] = p[1] * p[7] - p[3] * p[5];
    if (c.mhd) {
      f[1] = f[1] + R(0.5) * (p[6] * p[6] + p[7] * p[7] - p[5] * p[5]);
      f[2] = f[2] - p[5] * p[6];
      f[3] = f[3] - p[5] * p[7];
      f[4] = f[4] + p[1] * (p[5] * p[5] + p[6] * p[6] + p[7] * p[7]) -
             p[5] * (p[1] * p[5] + p[2] * p[6] + p[3] * p[7]);
    }
  }
}

// swap of the x and d components of velocity and field (d = 1 -> y, 2 -> z)
template <typename R> __host__ __device__ inline void swapdim(int nv, int d, R *v) {
  if (d == 0) return;
  R t = v[1]; v[1] = v[1 + d]; v[1 + d] = t;
  if (nv == 8) { t = v[5]; v[5] = v[5 + d]; v[5 + d] = t; }
}

// sound (source/hydro_core.f90:503-515)
template <typename R> __host__ __device__ inline R sound(const Phys<R> &c, const R *p) {
  return rsqrt_(c.gamma * p[4] / p[0]);
}

// cfast (source/hydro_core.f90:530-543): fast speed along one axis with
// field component bd
template <typename R>
__host__ __device__ inline R cfast1(const Phys<R> &c, R p, R d, R bx, R by, R bz, R bd) {
  R b2 = bx * bx + by * by + bz * bz;
  R a = c.gamma * p + b2;
  return msqrt(rdiv(R(0.5) * (a + msqrt(a * a - R(4.0) * c.gamma * p * bd * bd)), d));
}

// cfastX (source/hydro_core.f90:553-566)
template <typename R> __host__ __device__ inline R cfastX(const Phys<R> &c, const R *p) {
  R b2 = p[5] * p[5] + p[6] * p[6] + p[7] * p[7];
  R cs2va2 = rdiv(c.gamma * p[4] + b2, p[0]);
  return msqrt(R(0.5) *
               (cs2va2 + msqrt(cs2va2 * cs2va2 - rdiv(rdiv(R(4.0) * c.gamma * p[4] * (p[5] * p[5]), p[0]), p[0]))));
}

// average (source/hydro_core.f90:640-697)
template <typename R> __host__ __device__ inline R average(R a, R b, int opt) {
  R s, cc, d, r;
  switch (opt) {
    case -1: return R(0);
    case 0: return R(0.5) * (a + b);
    case 2:
      if (a * b <= R(0)) return R(0);
      return a * b * (a + b) / (a * a + b * b);
    case 1:
      s = rcopysign(R(1), a);
      return s * rfmax(R(0), rfmin(rabs(a), s * b));
    case 3: {
      R eps = R(1.0e-7);
      return (a * (b * b + eps) + b * (a * a + eps)) / (a * a + b * b * eps);
    }
    case 4:
      s = rcopysign(R(1), a);
      cc = R(0.25) * a + R(0.75) * b;
      d = R(0.75) * a + R(0.25) * b;
      r = rfmin(rfmin(rfmin(R(2.0) * rabs(a), R(2.0) * s * b), s * cc), s * d);
      return s * rfmax(R(0), r);
    case 5:
      s = rcopysign(R(1), a);
      cc = R(0.5) * (a + b);
      r = rfmin(rfmin(R(2.0) * rabs(a), R(2.0) * s * b), s * cc);
      return s * rfmax(R(0), r);
    case 6:
      s = rcopysign(R(1), b);
      cc = rfmin(R(2.0) * rabs(b), s * a);
      d = rfmin(rabs(b), R(2.0) * s * a);
      return s * rfmax(rfmax(R(0), cc), d);
    default: return R(0);
  }
}

// limiter (source/hydro_core.f90:613-638) for one variable
template <typename R>
__host__ __device__ inline void limit1(R pll, R &pl, R &pr, R prr, int lim) {
  R dl = pl - pll, dm = pr - pl, dr = prr - pr;
  R al = average(dl, dm, lim), ar = average(dm, dr, lim);
  pl = pl + R(0.5) * al;
  pr = pr - R(0.5) * ar;
}

// limit1 with the minmod limiter in FP32: minmod(a, b) is the median of a, b
// and 0 (the smaller for a, b > 0, the larger for a, b < 0, zero when the
// signs differ), one v_med3_f32 on the device.  The value equals the
// reference's s*max(0, min(|a|, s*b)) (source/hydro_core.f90:640-697,
// case 1) except for the sign of a zero slope.
__host__ __device__ inline float minmod32(float a, float b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_fmed3f(a, b, 0.0f);
#else
  return fmaxf(fminf(a, b), fminf(fmaxf(a, b), 0.0f));
#endif
}
__host__ __device__ inline void limit1_minmod32(float pll, float &pl, float &pr, float prr) {
  float dl = pl - pll, dm = pr - pl, dr = prr - pr;
  float al = minmod32(dl, dm), ar = minmod32(dm, dr);
  pl = pl + 0.5f * al;
  pr = pr - 0.5f * ar;
}

// HLL / HLLE flux (source/hll.f90:175-215, source/hlle.f90:166-206);
// strict selects the HLLE comparisons (sl > 0, sr < 0)
template <typename R>
__host__ __device__ inline void hll_flux(const Phys<R> &c, const R *pL, const R *pR, R sl, R sr, bool strict,
                                         R *ff) {
  R fL[MAXV], fR[MAXV], uL[MAXV], uR[MAXV];
  bool left = strict ? (sl > R(0)) : (sl >= R(0));
  bool right = strict ? (sr < R(0)) : (sr <= R(0));
  if (left) { xflux(c, pL, ff); return; }
  if (right) { xflux(c, pR, ff); return; }
  xflux(c, pL, fL);
  xflux(c, pR, fR);
  prim2flow(c, pL, uL);
  prim2flow(c, pR, uR);
  for (int q = 0; q < c.nv; q++) ff[q] = (sr * fL[q] - sl * fR[q] + sl * sr * (uR[q] - uL[q])) / (sr - sl);
}

// primfhll + wavespeedHLL (source/hll.f90:175-260)
template <typename R>
__host__ __device__ inline void primfhll(const Phys<R> &c, const R *pL, const R *pR, R *ff) {
  R cl = sound(c, pL), cr = sound(c, pR);
  R sl = rfmin(pL[1] - cl, pR[1] - cr);
  R sr = rfmax(pL[1] + cl, pR[1] + cr);
  hll_flux(c, pL, pR, sl, sr, false, ff);
}

// primfhlle + wavespeedHLLE (source/hlle.f90:166-241)
template <typename R>
__host__ __device__ inline void primfhlle(const Phys<R> &c, const R *pL, const R *pR, R *ff) {
  R cfL = cfastX(c, pL), cfR = cfastX(c, pR);
  R sl = rfmin(pL[1] - cfL, pR[1] - cfR);
  R sr = rfmax(pL[1] + cfL, pR[1] + cfR);
  hll_flux(c, pL, pR, sl, sr, true, ff);
}

// primfhllc + wavespeedHLLC (source/hllc.f90:209-357), Toro 10.33-10.38
template <typename R>
__host__ __device__ inline void primfhllc(const Phys<R> &c, const R *pL, const R *pR, R *ff) {
  R dl1 = pL[0], dr1 = pR[0], ul1 = pL[1], ur1 = pR[1], pl1 = pL[4], pr1 = pR[4];
  R cl = sound(c, pL), cr = sound(c, pR);
  R sl = rfmin(ul1 - cl, ur1 - cr);
  R sr = rfmax(ul1 + cl, ur1 + cr);
  R sst;
  if (c.solver == SOLVER_HLLC)
    sst = (pr1 - pl1 + dl1 * ul1 * (sl - ul1) - dr1 * ur1 * (sr - ur1)) / (dl1 * (sl - ul1) - dr1 * (sr - ur1));
  else
    sst = R(0);
  if (sl > R(0)) { xflux(c, pL, ff); return; }
  if (sr < R(0)) { xflux(c, pR, ff); return; }
  const R *pK;
  R sK;
  if (sst >= R(0)) { pK = pL; sK = sl; }
  else if (sst <= R(0)) { pK = pR; sK = sr; }
  else { for (int q = 0; q < c.nv; q++) ff[q] = R(NAN); return; }
  R fK[MAXV], uK[MAXV], ust[MAXV];
  xflux(c, pK, fK);
  prim2flow(c, pK, uK);
  R rhost = pK[0] * (sK - pK[1]) / (sK - sst);
  R EK = R(0.5) * pK[0] * (pK[1] * pK[1] + pK[2] * pK[2] + pK[3] * pK[3]) + c.cv * pK[4];
  ust[0] = rhost;
  ust[1] = rhost * sst;
  ust[2] = rhost * pK[2];
  ust[3] = rhost * pK[3];
  ust[4] = rhost * (EK / pK[0] + (sst - pK[1]) * (sst + pK[4] / (pK[0] * (sK - pK[1]))));
  for (int q = 5; q < c.nv; q++) ust[q] = R(0);
  for (int q = 0; q < c.nv; q++) ff[q] = fK[q] + sK * (ust[q] - uK[q]);
}

// star state of one side for HLLD (source/hlld.f90:241-257, 289-305, 337-365)
template <typename R>
__host__ __device__ inline void hlld_star(const R *p, R bx, R smuK, R smsM, R sMmuK, R &vst, R &wst, R &byst,
                                          R &bzst) {
  R den = p[0] * smuK * smsM - bx * bx;
  if (den == R(0)) {
    vst = p[2]; wst = p[3]; byst = R(0); bzst = R(0);
  } else {
    vst = p[2] - rdiv(bx * p[6] * sMmuK, den);
    wst = p[3] - rdiv(bx * p[7] * sMmuK, den);
    byst = rdiv(p[6] * (p[0] * (smuK * smuK) - bx * bx), den);
    bzst = rdiv(p[7] * (p[0] * (smuK * smuK) - bx * bx), den);
  }
}

// primfhlld (source/hlld.f90:166-444), Miyoshi & Kusano (2005)
template <typename R>
__host__ __device__ inline void primfhlld(const Phys<R> &c, const R *pl, const R *pr, R *ff) {
  const R cv = c.cv;
  R csl = cfastX(c, pl), csr = cfastX(c, pr);
  R sr = rfmax(pl[1] + csl, pr[1] + csr);
  R sl = rfmin(pl[1] - csl, pr[1] - csr);
  if (sl > R(0)) { xflux(c, pl, ff); return; }
  if (sr < R(0)) { xflux(c, pr, ff); return; }
  R bx = R(0.5) * (pl[5] + pr[5]);
  R signBx = rcopysign(R(1), bx);
  R pTL = pl[4] + R(0.5) * (bx * bx + pl[6] * pl[6] + pl[7] * pl[7]);
  R pTR = pr[4] + R(0.5) * (bx * bx + pr[6] * pr[6] + pr[7] * pr[7]);
  R slmul = sl - pl[1];
  R srmur = sr - pr[1];
  R rholul = pl[0] * pl[1];
  R rhorur = pr[0] * pr[1];
  R sM = rdiv(srmur * rhorur - slmul * rholul - pTR + pTL, srmur * pr[0] - slmul * pl[0]);
  R srmsM = sr - sM;
  R slmsM = sl - sM;
  R rhostl = rdiv(pl[0] * slmul, slmsM);
  R rhostr = rdiv(pr[0] * srmur, srmsM);
  R sstl = sM - rdiv(rabs(bx), msqrt(rhostl));
  R sstr = sM + rdiv(rabs(bx), msqrt(rhostr));
  R pst = rdiv(srmur * pr[0] * pTL - slmul * pl[0] * pTR + pl[0] * pr[0] * srmur * slmul * (pr[1] - pl[1]),
               srmur * pr[0] - slmul * pl[0]);
  R vstl, wstl, bystl, bzstl, vstr, wstr, bystr, bzstr;
  if (sstl >= R(0)) {  // UL*
    R el = R(0.5) * pl[0] * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) + cv * pl[4] +
           R(0.5) * (bx * bx + pl[6] * pl[6] + pl[7] * pl[7]);
    hlld_star(pl, bx, slmul, slmsM, sM - pl[1], vstl, wstl, bystl, bzstl);
    R vdotbl = pl[1] * bx + pl[2] * pl[6] + pl[3] * pl[7];
    R vstdotbstl = sM * bx + vstl * bystl + wstl * bzstl;
    R estl = rdiv(slmul * el - pTL * pl[1] + pst * sM + bx * (vdotbl - vstdotbstl), slmsM);
    ff[0] = rhostl * sM;
    ff[1] = rhostl * (sM * sM) + pst - bx * bx;
    ff[2] = rhostl * sM * vstl - bx * bystl;
    ff[3] = rhostl * sM * wstl - bx * bzstl;
    ff[4] = sM * (estl + pst) - bx * (vstdotbstl);
    ff[5] = R(0);
    ff[6] = bystl * sM - bx * vstl;
    ff[7] = bzstl * sM - bx * wstl;
    return;
  }
  if (sstr <= R(0)) {  // UR*
    R er = R(0.5) * pr[0] * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) + cv * pr[4] +
           R(0.5) * (bx * bx + pr[6] * pr[6] + pr[7] * pr[7]);
    hlld_star(pr, bx, srmur, srmsM, sM - pr[1], vstr, wstr, bystr, bzstr);
    R vdotbr = pr[1] * bx + pr[2] * pr[6] + pr[3] * pr[7];
    R vstdotbstr = sM * bx + vstr * bystr + wstr * bzstr;
    R estr = rdiv(srmur * er - pTR * pr[1] + pst * sM + bx * (vdotbr - vstdotbstr), srmsM);
    ff[0] = rhostr * sM;
    ff[1] = rhostr * (sM * sM) + pst - bx * bx;
    ff[2] = rhostr * sM * vstr - bx * bystr;
    ff[3] = rhostr * sM * wstr - bx * bzstr;
    ff[4] = sM * (estr + pst) - bx * (vstdotbstr);
    ff[5] = R(0);
    ff[6] = bystr * sM - bx * vstr;
    ff[7] = bzstr * sM - bx * wstr;
    return;
  }
  hlld_star(pl, bx, slmul, slmsM, sM - pl[1], vstl, wstl, bystl, bzstl);
  hlld_star(pr, bx, srmur, srmsM, sM - pr[1], vstr, wstr, bystr, bzstr);
  R sql = msqrt(rhostl), sqr = msqrt(rhostr);
  R dd = sql + sqr;
  R vstst = rdiv(sql * vstl + sqr * vstr + (bystr - bystl) * signBx, dd);
  R wstst = rdiv(sql * wstl + sqr * wstr + (bzstr - bzstl) * signBx, dd);
  R sqlr = msqrt(rhostl * rhostr);
  R bystst = rdiv(sql * bystr + sqr * bystl + sqlr * (vstr - vstl) * signBx, dd);
  R bzstst = rdiv(sql * bzstr + sqr * bzstl + sqlr * (wstr - wstl) * signBx, dd);
  R vststdotbstst = sM * bx + vstst * bystst + wstst * bzstst;
  if (sM >= R(0)) {  // UL**
    R el = R(0.5) * pl[0] * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) + cv * pl[4] +
           R(0.5) * (bx * bx + pl[6] * pl[6] + pl[7] * pl[7]);
    R vdotbl = pl[1] * bx + pl[2] * pl[6] + pl[3] * pl[7];
    R vstdotbstl = sM * bx + vstl * bystl + wstl * bzstl;
    R estl = rdiv(slmul * el - pTL * pl[1] + pst * sM + bx * (vdotbl - vstdotbstl), slmsM);
    R eststl = estl - sql * (vstdotbstl - vststdotbstst) * signBx;
    ff[0] = rhostl * sM;
    ff[1] = rhostl * (sM * sM) + pst - bx * bx;
    ff[2] = rhostl * sM * vstst - bx * bystst;
    ff[3] = rhostl * sM * wstst - bx * bzstst;
    ff[4] = sM * (eststl + pst) - bx * (vststdotbstst);
    ff[5] = R(0);
    ff[6] = bystst * sM - bx * vstst;
    ff[7] = bzstst * sM - bx * wstst;
    return;
  }
  if (sM <= R(0)) {  // UR**
    R er = R(0.5) * pr[0] * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) + cv * pr[4] +
           R(0.5) * (bx * bx + pr[6] * pr[6] + pr[7] * pr[7]);
    R vdotbr = pr[1] * bx + pr[2] * pr[6] + pr[3] * pr[7];
    R vstdotbstr = sM * bx + vstr * bystr + wstr * bzstr;
    R estr = rdiv(srmur * er - pTR * pr[1] + pst * sM + bx * (vdotbr - vstdotbstr), srmsM);
    R eststr = estr + sqr * (vstdotbstr - vststdotbstst) * signBx;
    ff[0] = rhostr * sM;
    ff[1] = rhostr * (sM * sM) + pst - bx * bx;
    ff[2] = rhostr * sM * vstst - bx * bystst;
    ff[3] = rhostr * sM * wstst - bx * bzstst;
    ff[4] = sM * (eststr + pst) - bx * (vststdotbstst);
    ff[5] = R(0);
    ff[6] = bystst * sM - bx * vstst;
    ff[7] = bzstst * sM - bx * wstst;
    return;
  }
  for (int q = 0; q < 8; q++) ff[q] = R(NAN);
}

// primfhlld evaluated with one code path for the four inner regions (UL*,
// UR*, UL**, UR**): both star states are formed once, the region's values
// are picked with selects, and only the double-star states sit behind a
// branch.  Every value is the expression primfhlld uses for its region, so
// the two functions agree bit for bit; a wavefront whose lanes fall into
// different regions no longer runs each region's code in turn.
// MERGED: one code path for the two supersonic cases (the flux of the
// upwind state); the separate paths spill less in the second-stage kernel.
template <typename R, bool MERGED = false>
__host__ __device__ inline void primfhlld_sel(const Phys<R> &c, const R *pl, const R *pr, R *ff) {
  const R cv = c.cv;
  R csl = cfastX(c, pl), csr = cfastX(c, pr);
  R sr = rfmax(pl[1] + csl, pr[1] + csr);
  R sl = rfmin(pl[1] - csl, pr[1] - csr);
  if constexpr (MERGED) {
    if (sl > R(0) || sr < R(0)) {
      R pu[8];
      for (int q = 0; q < 8; q++) pu[q] = sl > R(0) ? pl[q] : pr[q];
      xflux(c, pu, ff);
      return;
    }
  } else {
    if (sl > R(0)) { xflux(c, pl, ff); return; }
    if (sr < R(0)) { xflux(c, pr, ff); return; }
  }
  R bx = R(0.5) * (pl[5] + pr[5]);
  R signBx = rcopysign(R(1), bx);
  R pTL = pl[4] + R(0.5) * (bx * bx + pl[6] * pl[6] + pl[7] * pl[7]);
  R pTR = pr[4] + R(0.5) * (bx * bx + pr[6] * pr[6] + pr[7] * pr[7]);
  R slmul = sl - pl[1];
  R srmur = sr - pr[1];
  R rholul = pl[0] * pl[1];
  R rhorur = pr[0] * pr[1];
  R sM = rdiv(srmur * rhorur - slmul * rholul - pTR + pTL, srmur * pr[0] - slmul * pl[0]);
  R srmsM = sr - sM;
  R slmsM = sl - sM;
  R rhostl = rdiv(pl[0] * slmul, slmsM);
  R rhostr = rdiv(pr[0] * srmur, srmsM);
  R sstl = sM - rdiv(rabs(bx), msqrt(rhostl));
  R sstr = sM + rdiv(rabs(bx), msqrt(rhostr));
  R pst = rdiv(srmur * pr[0] * pTL - slmul * pl[0] * pTR + pl[0] * pr[0] * srmur * slmul * (pr[1] - pl[1]),
               srmur * pr[0] - slmul * pl[0]);
  // region, in the order the reference tests it
  const bool s1L = sstl >= R(0);
  const bool s1R = !s1L && sstr <= R(0);
  const bool d2L = !s1L && !s1R && sM >= R(0);
  const bool d2R = !s1L && !s1R && !d2L && sM <= R(0);
  if (!(s1L || s1R || d2L || d2R)) {
    for (int q = 0; q < 8; q++) ff[q] = R(NAN);
    return;
  }
  const bool isL = s1L || d2L;
  const bool single = s1L || s1R;
  R vstl, wstl, bystl, bzstl, vstr, wstr, bystr, bzstr;
  hlld_star(pl, bx, slmul, slmsM, sM - pl[1], vstl, wstl, bystl, bzstl);
  hlld_star(pr, bx, srmur, srmsM, sM - pr[1], vstr, wstr, bystr, bzstr);
  // the side K whose star state the flux is built on
  R pK[8];
  for (int q = 0; q < 8; q++) pK[q] = isL ? pl[q] : pr[q];
  const R sKmuK = isL ? slmul : srmur;
  const R sKmsM = isL ? slmsM : srmsM;
  const R rhostK = isL ? rhostl : rhostr;
  const R pTK = isL ? pTL : pTR;
  const R vstK = isL ? vstl : vstr, wstK = isL ? wstl : wstr;
  const R bystK = isL ? bystl : bystr, bzstK = isL ? bzstl : bzstr;
  R eK = R(0.5) * pK[0] * (pK[1] * pK[1] + pK[2] * pK[2] + pK[3] * pK[3]) + cv * pK[4] +
         R(0.5) * (bx * bx + pK[6] * pK[6] + pK[7] * pK[7]);
  R vdotbK = pK[1] * bx + pK[2] * pK[6] + pK[3] * pK[7];
  R vstdotbstK = sM * bx + vstK * bystK + wstK * bzstK;
  R estK = rdiv(sKmuK * eK - pTK * pK[1] + pst * sM + bx * (vdotbK - vstdotbstK), sKmsM);
  R vq = vstK, wq = wstK, byq = bystK, bzq = bzstK, eq = estK, vdbq = vstdotbstK;
  if (!single) {
    R sql = msqrt(rhostl), sqr = msqrt(rhostr);
    R dd = sql + sqr;
    R vstst = rdiv(sql * vstl + sqr * vstr + (bystr - bystl) * signBx, dd);
    R wstst = rdiv(sql * wstl + sqr * wstr + (bzstr - bzstl) * signBx, dd);
    R sqlr = msqrt(rhostl * rhostr);
    R bystst = rdiv(sql * bystr + sqr * bystl + sqlr * (vstr - vstl) * signBx, dd);
    R bzstst = rdiv(sql * bzstr + sqr * bzstl + sqlr * (wstr - wstl) * signBx, dd);
    R vststdotbstst = sM * bx + vstst * bystst + wstst * bzstst;
    R eststl = estK - sql * (vstdotbstK - vststdotbstst) * signBx;
    R eststr = estK + sqr * (vstdotbstK - vststdotbstst) * signBx;
    vq = vstst;
    wq = wstst;
    byq = bystst;
    bzq = bzstst;
    eq = isL ? eststl : eststr;
    vdbq = vststdotbstst;
  }
  ff[0] = rhostK * sM;
  ff[1] = rhostK * (sM * sM) + pst - bx * bx;
  ff[2] = rhostK * sM * vq - bx * byq;
  ff[3] = rhostK * sM * wq - bx * bzq;
  ff[4] = sM * (eq + pst) - bx * (vdbq);
  ff[5] = R(0);
  ff[6] = byq * sM - bx * vq;
  ff[7] = bzq * sM - bx * wq;
}

// solver dispatch on x-aligned states
template <typename R>
__host__ __device__ inline void riemann(const Phys<R> &c, const R *pl, const R *pr, R *ff) {
  switch (c.solver) {
    case SOLVER_HLLC: primfhllc(c, pl, pr, ff); break;
    case SOLVER_HLLD: primfhlld(c, pl, pr, ff); break;
    case SOLVER_HLLE: primfhlle(c, pl, pr, ff); break;
    default: primfhll(c, pl, pr, ff); break;
  }
}

}  // namespace wx
