// The delta-form HLLD and HLLE fluxes (wx_delta.h) against the reference's
// form (primfhlld / primfhlle of wx_physics.h, which the unit vectors pin to
// source/hlld.f90 and source/hlle.f90):
//  - with FP64 corrections the two agree to FP64 rounding in every region
//    (the algebra of the increments is exact);
//  - with FP32 corrections on states that differ by small jumps, the error
//    is FP32 rounding of the jump-sized correction, far below FP32
//    rounding of the flux itself (the point of the form).
// Faces near the degenerate HLLD star state (den -> 0) take the plain
// one-code-path HLLD in FP32, as in the kernels: there the error must be the
// plain solver's (it is the same arithmetic); they are counted apart.
// Prints the worst errors.
#include <cmath>
#include <cstdio>
#include <random>

#include "wx_delta.h"

using namespace wx;

int main() {
  Phys<double> c64{8, 1, SOLVER_HLLD, 1, 5.0 / 3.0, 1.5};
  Phys<float> c32{8, 1, SOLVER_HLLD, 1, 5.0f / 3.0f, 1.5f};
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> U(-1, 1);
  double worst64 = 0, worst32_small = 0, worst32_ratio = 0, plain32_small = 0, plain32_large = 0;
  long nan_mismatch = 0, ill_small = 0, ill_large = 0, ill_worse = 0;
  int used[2] = {0, 0};
  for (int solver = 5; solver <= 6; solver++) {
    for (int t = 0; t < 400000; t++) {
      double pl[8], pr[8];
      const double sc = (t % 3 == 0) ? 0.01 : 1.0;
      const double vs = (t % 5 == 0) ? 5.0 : 1.0;
      const bool small = t % 2 == 0;          // nearby states: the jumps of a smooth flow
      pl[0] = 0.1 + std::abs(U(g));
      for (int q = 1; q < 4; q++) pl[q] = U(g) * vs;
      pl[4] = 0.01 + std::abs(U(g));
      for (int q = 5; q < 8; q++) pl[q] = U(g) * sc;
      if (t % 7 == 0) pl[5] = 0.0;
      for (int q = 0; q < 8; q++) {
        if (small) pr[q] = pl[q] * (1.0 + 1e-3 * U(g)) + 1e-4 * U(g) * (q >= 1 && q <= 3 ? 1.0 : 0.0);
        else pr[q] = (q == 0 || q == 4) ? 0.01 + std::abs(U(g)) : U(g) * (q >= 5 ? sc : 1.0);
      }
      if (t % 7 == 0) pr[5] = pl[5];
      double fref[8], f64[8], f32[8];
      bool ill = false;
      float ql[8], qr[8], fp[8];
      for (int q = 0; q < 8; q++) { ql[q] = (float)pl[q]; qr[q] = (float)pr[q]; }
      if (solver == 6) primfhlld(c32, ql, qr, fp);
      else primfhlle(c32, ql, qr, fp);
      if (solver == 6) {
        primfhlld(c64, pl, pr, fref);
        if (!hlld_flux_delta(c64, c64, pl, pr, f64)) return 1;   // FP64 never defers
        if (!hlld_flux_delta(c64, c32, pl, pr, f32)) {
          hlld_flux_ill(c32, pl, pr, f32);
          ill = true;
        }
      } else {
        primfhlle(c64, pl, pr, fref);
        hlle_flux_delta(c64, c64, pl, pr, f64);
        hlle_flux_delta(c64, c32, pl, pr, f32);
      }
      double scale = 0;
      bool anynan = false;
      for (int q = 0; q < 8; q++) {
        if (std::isnan(fref[q])) anynan = true;
        else scale = std::fmax(scale, std::abs(fref[q]));
      }
      if (anynan) {
        for (int q = 0; q < 8; q++)
          if (!std::isnan(f64[q])) nan_mismatch++;
        continue;
      }
      used[solver - 5]++;
      double e64 = 0, e32 = 0, ep = 0;
      for (int q = 0; q < 8; q++) {
        e64 = std::fmax(e64, std::abs(f64[q] - fref[q]));
        e32 = std::fmax(e32, std::abs(f32[q] - fref[q]));
        ep = std::fmax(ep, std::abs((double)fp[q] - fref[q]));
      }
      if (small) plain32_small = std::fmax(plain32_small, ep / scale);
      else plain32_large = std::fmax(plain32_large, ep / scale);
      worst64 = std::fmax(worst64, e64 / scale);
      if (ill) {
        (small ? ill_small : ill_large)++;
        if (e32 > ep) ill_worse++;
        continue;
      }
      if (small) {
        worst32_small = std::fmax(worst32_small, e32 / scale);
      } else {
        worst32_ratio = std::fmax(worst32_ratio, e32 / scale);
      }
    }
  }
  printf("records hlle %d hlld %d nan-mismatch %ld\n", used[0], used[1], nan_mismatch);
  printf("near-degenerate faces (plain FP32 HLLD): small jumps %ld, large jumps %ld, less accurate than the "
         "plain solver %ld\n", ill_small, ill_large, ill_worse);
  printf("fp64 corrections: worst relative %.3e\n", worst64);
  printf("fp32 corrections: small jumps worst relative %.3e, large jumps %.3e\n", worst32_small, worst32_ratio);
  printf("plain fp32 solver: small jumps worst relative %.3e, large jumps %.3e\n", plain32_small, plain32_large);
  // FP64: rounding, amplified by the ill-conditioned star states of strong
  // random jumps; FP32 on small jumps: two orders below the plain FP32
  // solver; on strong random jumps (shocks): the plain solver's magnitude
  int ok = worst64 < 1e-10 && worst32_small < 5e-9 && worst32_small * 100 < plain32_small &&
           worst32_ratio < 1e-3 && nan_mismatch == 0 && ill_worse == 0 && ill_small < used[1] / 100;
  printf(ok ? "delta form ok\n" : "delta form FAILED\n");
  return ok ? 0 : 1;
}
