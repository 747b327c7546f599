// wx_delta.h - the HLLD and HLLE fluxes of the FP32 MHD form in delta form.
//
// A face flux is written as F = F_K + D: F_K is the physical flux
// (prim2fluxes) of the reconstructed state K on the side of the contact
// that the solution lies on, evaluated in FP64 from FP64 face states; D is
// the Riemann solver's correction, evaluated in the narrow type R (FP32 in
// the kernels) from quantities of the size of the jumps across the fan:
// wave-speed offsets s - u, the contact offset sM - u, and the increments
// U*_K - U_K and U**_K - U*_K of the intermediate states.  Each increment is
// derived algebraically from those offsets (the Rankine-Hugoniot form
// F*_K = F_K + s_K (U*_K - U_K), F**_K = F*_K + s*_K (U**_K - U*_K) of
// Miyoshi & Kusano 2005), so no O(1) quantity is rounded to R on the way:
// the flux difference of a cell keeps the pressure and energy gradients to
// FP64 precision plus R's relative precision on jump-sized terms.  In exact
// arithmetic F equals the reference's flux in every region
// (source/hlld.f90:166-444, source/hlle.f90:166-206).
#pragma once

#include "wx_physics.h"

namespace wx {

// x-flux change when the normal field b of a state (rho,u,v,w,p,b,By,Bz) is
// replaced by b + db (the HLLD fan carries the mean normal field): only the
// terms of prim2fluxes that hold b
template <typename R>
__host__ __device__ inline void bx_shift(const R *p, R b, R db, R *g) {
  g[0] = R(0);
  g[1] = -db * (b + R(0.5) * db);
  g[2] = -db * p[6];
  g[3] = -db * p[7];
  g[4] = -db * (p[2] * p[6] + p[3] * p[7]);
  g[5] = R(0);
  g[6] = -p[2] * db;
  g[7] = -p[3] * db;
}

// Star-state increments of one side K (source/hlld.f90:241-257, 289-305):
// with f = (sM - u_K) / den, den = rho (s - u)(s - sM) - bx^2,
//   v* - v = -bx By f,  w* - w = -bx Bz f,
//   By* - By = By rho (s - u) f,  Bz* - Bz = Bz rho (s - u) f;
// den = 0 takes the reference's degenerate values (v* = v, B* = 0).
// Returns true when den is small against rho (s - u)(s - sM) (a factor
// tau): there the star states grow like 1/den, and a narrow type loses the
// increments to cancellation; the caller evaluates such faces otherwise
// (hlld_flux_ill).
template <typename R>
__host__ __device__ inline bool star_inc(const R *p, R bx, R smu, R smsM, R dsM, R &dv, R &dw, R &dby, R &dbz,
                                         R tau) {
  const R a = p[0] * smu * smsM;
  const R den = a - bx * bx;
  const bool ill = rabs(den) < tau * rabs(a);
  if (den == R(0)) {
    dv = R(0);
    dw = R(0);
    dby = -p[6];
    dbz = -p[7];
  } else {
    const R f = rdiv(dsM, den);
    dv = -bx * p[6] * f;
    dw = -bx * p[7] * f;
    const R g = p[0] * smu * f;
    dby = p[6] * g;
    dbz = p[7] * g;
  }
  return ill;
}

// hlld_delta's outcome: the base side, or a near-degenerate face
enum { DELTA_LEFT = 0, DELTA_RIGHT = 1, DELTA_ILL = 2 };

// HLLD correction D (F = F_K + D) from the face states pl, pr in R and
// their jumps dW = pr - pl (formed where the states are exact, then rounded
// to R).  Returns DELTA_LEFT or DELTA_RIGHT (the base side K), or DELTA_ILL
// when a star state is close to the degenerate den = 0 (tau, see
// star_inc; 0 never flags).
template <typename R>
__host__ __device__ inline int hlld_delta(const Phys<R> &c, const R *pl, const R *pr, const R *dW, R *D,
                                          R tau = R(0)) {
  const R cv = c.cv;
  const R cfl = cfastX(c, pl), cfr = cfastX(c, pr);
  // wave speeds (wavespeedHLLD: Davis estimates with cfastX) as offsets from
  // the side velocities: sl - uL = min(-cfL, (uR - uL) - cfR), and so on
  const R slmul = rfmin(-cfl, dW[1] - cfr);
  const R srmur = rfmax(cfl - dW[1], cfr);
  const R sl = pl[1] + slmul, sr = pr[1] + srmur;
  for (int q = 0; q < 8; q++) D[q] = R(0);
  if (sl > R(0)) return DELTA_LEFT;
  if (sr < R(0)) return DELTA_RIGHT;
  // mean normal field and the total-pressure jump (bx^2 cancels in it)
  const R dbx = R(0.5) * dW[5];
  const R bx = pl[5] + dbx;
  const R sgn = rcopysign(R(1), bx);
  const R dpT = dW[4] + R(0.5) * (dW[6] * (pl[6] + pr[6]) + dW[7] * (pl[7] + pr[7]));
  // contact speed as offsets sM - uL, sM - uR
  const R Dn = srmur * pr[0] - slmul * pl[0];
  const R duL = rdiv(srmur * pr[0] * dW[1] - dpT, Dn);
  const R duR = duL - dW[1];
  const R sM = pl[1] + duL;
  const R slmsM = slmul - duL, srmsM = srmur - duR;
  // rho* - rho and rho*
  const R drL = rdiv(pl[0] * duL, slmsM), drR = rdiv(pr[0] * duR, srmsM);
  const R rsL = pl[0] + drL, rsR = pr[0] + drR;
  const R sqL = msqrt(rsL), sqR = msqrt(rsR);
  const R sstl = sM - rdiv(rabs(bx), sqL);
  const R sstr = sM + rdiv(rabs(bx), sqR);
  // region, in the order the reference tests it
  const bool s1L = sstl >= R(0);
  const bool s1R = !s1L && sstr <= R(0);
  const bool d2L = !s1L && !s1R && sM >= R(0);
  const bool d2R = !s1L && !s1R && !d2L && sM <= R(0);
  if (!(s1L || s1R || d2L || d2R)) {
    for (int q = 0; q < 8; q++) D[q] = R(NAN);
    return DELTA_LEFT;
  }
  const bool isL = s1L || d2L;
  const bool single = s1L || s1R;
  R dvL, dwL, dbyL, dbzL, dvR, dwR, dbyR, dbzR;
  const bool illL = star_inc(pl, bx, slmul, slmsM, duL, dvL, dwL, dbyL, dbzL, tau);
  const bool illR = star_inc(pr, bx, srmur, srmsM, duR, dvR, dwR, dbyR, dbzR, tau);
  // the side K alone in a single-star region, both across the contact
  if (single ? (isL ? illL : illR) : (illL || illR)) return DELTA_ILL;
  // the side K of the solution
  R pK[8];
  for (int q = 0; q < 8; q++) pK[q] = isL ? pl[q] : pr[q];
  const R sK = isL ? sl : sr, sKmuK = isL ? slmul : srmur, sKmsM = isL ? slmsM : srmsM;
  const R duK = isL ? duL : duR, drK = isL ? drL : drR, rsK = isL ? rsL : rsR;
  const R dvK = isL ? dvL : dvR, dwK = isL ? dwL : dwR, dbyK = isL ? dbyL : dbyR, dbzK = isL ? dbzL : dbzR;
  const R vsK = pK[2] + dvK, wsK = pK[3] + dwK, bysK = pK[6] + dbyK, bzsK = pK[7] + dbzK;
  // F_K with the mean normal field, relative to F_K with its own
  bx_shift(pK, pK[5], isL ? dbx : -dbx, D);
  // energy increment E* - E (both with the mean normal field):
  //   [(sM - u)(E + pT + rho (s - u) sM) + bx (v.B - v*.B*)] / (s - sM)
  const R bt2 = bx * bx + pK[6] * pK[6] + pK[7] * pK[7];
  const R eK = R(0.5) * pK[0] * (pK[1] * pK[1] + pK[2] * pK[2] + pK[3] * pK[3]) + cv * pK[4] + R(0.5) * bt2;
  const R pTK = pK[4] + R(0.5) * bt2;
  const R X = -duK * bx - (dvK * bysK + pK[2] * dbyK) - (dwK * bzsK + pK[3] * dbzK);
  const R dE = rdiv(duK * (eK + pTK + pK[0] * sKmuK * sM) + bx * X, sKmsM);
  // + s_K (U*_K - U_K)
  D[0] += sK * drK;
  D[1] += sK * (sK * drK);
  D[2] += sK * (drK * vsK + pK[0] * dvK);
  D[3] += sK * (drK * wsK + pK[0] * dwK);
  D[4] += sK * dE;
  D[6] += sK * dbyK;
  D[7] += sK * dbzK;
  if (!single) {
    // jumps of the star states across the contact
    const R jv = dW[2] + (dvR - dvL), jw = dW[3] + (dwR - dwL);
    const R jby = dW[6] + (dbyR - dbyL), jbz = dW[7] + (dbzR - dbzL);
    const R dd = sqL + sqR;
    // v** - v*L, By** - By*L (source/hlld.f90:320-335)
    const R dvsL = rdiv(sqR * jv + jby * sgn, dd), dwsL = rdiv(sqR * jw + jbz * sgn, dd);
    const R dbysL = rdiv(sqL * jby + sqL * sqR * jv * sgn, dd), dbzsL = rdiv(sqL * jbz + sqL * sqR * jw * sgn, dd);
    const R dvs = isL ? dvsL : dvsL - jv, dws = isL ? dwsL : dwsL - jw;
    const R dbys = isL ? dbysL : dbysL - jby, dbzs = isL ? dbzsL : dbzsL - jbz;
    // v*_K.B*_K - v**.B** (the sM bx terms cancel)
    const R vss = vsK + dvs, wss = wsK + dws;
    const R Y = -(dvs * bysK + vss * dbys) - (dws * bzsK + wss * dbzs);
    const R dEs = isL ? -sqL * Y * sgn : sqR * Y * sgn;
    const R ssK = isL ? sstl : sstr;
    // + s*_K (U**_K - U*_K)
    D[2] += ssK * (rsK * dvs);
    D[3] += ssK * (rsK * dws);
    D[4] += ssK * dEs;
    D[6] += ssK * dbys;
    D[7] += ssK * dbzs;
  }
  return isL ? DELTA_LEFT : DELTA_RIGHT;
}

// fraction of rho (s - u)(s - sM) below which |den| sends a face to FP64
constexpr float DELTA_TAU = 0.05f;

// The flux of a face near the degenerate star state (hlld_delta returned
// DELTA_ILL): the one-code-path HLLD in R on the face states rounded to R.
// There the star states grow like 1/den and the reference's own flux is
// ill-conditioned: the increments of the delta form lose their precision to
// cancellation, while this form keeps the plain solver's accuracy (an FP64
// evaluation would be exact, but its code, even unexecuted, slows the
// kernels by a quarter).
template <typename R>
__host__ __device__ inline void hlld_flux_ill(const Phys<R> &c, const double *wl, const double *wr, double *F) {
  R pl[8], pr[8], f[8];
  for (int q = 0; q < 8; q++) {
    pl[q] = (R)wl[q];
    pr[q] = (R)wr[q];
  }
  primfhlld_sel<R>(c, pl, pr, f);
  for (int q = 0; q < 8; q++) F[q] = (double)f[q];
}

// The HLLD flux from FP64 face states (x-aligned order) in delta form: F_K
// in FP64, D in R.  Returns false and leaves F untouched for a face near the
// degenerate star state (R narrower than FP64 only, rare): the caller
// evaluates that face with hlld_flux_ill.
template <typename R>
__host__ __device__ inline bool hlld_flux_delta(const Phys<double> &c64, const Phys<R> &c, const double *wl,
                                                const double *wr, double *F) {
  R pl[8], pr[8], dW[8], D[8];
  for (int q = 0; q < 8; q++) {
    pl[q] = (R)wl[q];
    pr[q] = (R)wr[q];
    dW[q] = (R)(wr[q] - wl[q]);
  }
  const int side = hlld_delta(c, pl, pr, dW, D, sizeof(R) < 8 ? R(DELTA_TAU) : R(0));
  if (side == DELTA_ILL) return false;
  double wk[8];
  for (int q = 0; q < 8; q++) wk[q] = side == DELTA_RIGHT ? wr[q] : wl[q];
  xflux(c64, wk, F);
  for (int q = 0; q < 8; q++) F[q] += (double)D[q];
  return true;
}

// The HLLE flux (source/hlle.f90:166-206) in delta form: subsonic
// F = F_L + sl (sr (U_R - U_L) - (F_R - F_L)) / (sr - sl), with both
// physical fluxes and both conserved states in FP64 and their differences
// rounded to R.
template <typename R>
__host__ __device__ inline void hlle_flux_delta(const Phys<double> &c64, const Phys<R> &c, const double *wl,
                                                const double *wr, double *F) {
  R pl[8], pr[8];
  for (int q = 0; q < 8; q++) {
    pl[q] = (R)wl[q];
    pr[q] = (R)wr[q];
  }
  const R cfl = cfastX(c, pl), cfr = cfastX(c, pr);
  const R dU1 = (R)(wr[1] - wl[1]);
  const R sl = pl[1] + rfmin(-cfl, dU1 - cfr);
  const R sr = pr[1] + rfmax(cfl - dU1, cfr);
  if (sl > R(0)) { xflux(c64, wl, F); return; }
  if (sr < R(0)) { xflux(c64, wr, F); return; }
  double fr[8], ul[8], ur[8];
  xflux(c64, wl, F);
  xflux(c64, wr, fr);
  prim2flow(c64, wl, ul);
  prim2flow(c64, wr, ur);
  const R w = rdiv(sl, sr - sl);
  for (int q = 0; q < 8; q++) F[q] += (double)(w * (sr * (R)(ur[q] - ul[q]) - (R)(fr[q] - F[q])));
}

}  // namespace wx
