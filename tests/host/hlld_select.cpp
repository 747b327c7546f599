// Bit-for-bit agreement of primfhlld_sel (one code path for the inner
// regions) with primfhlld (branch per region) on random x-aligned states,
// in double and float; prints how often each region was reached.
#include <cstdio>
#include <cstring>
#include <random>

#include "wx_physics.h"

using namespace wx;

// region of the HLLD fan that primfhlld picks (0 left, 1 right, 2 UL*,
// 3 UR*, 4 UL**, 5 UR**, 6 none)
template <typename R> int region(const Phys<R> &c, const R *pl, const R *pr) {
  R csl = cfastX(c, pl), csr = cfastX(c, pr);
  R sr = rfmax(pl[1] + csl, pr[1] + csr);
  R sl = rfmin(pl[1] - csl, pr[1] - csr);
  if (sl > R(0)) return 0;
  if (sr < R(0)) return 1;
  R bx = R(0.5) * (pl[5] + pr[5]);
  R pTL = pl[4] + R(0.5) * (bx * bx + pl[6] * pl[6] + pl[7] * pl[7]);
  R pTR = pr[4] + R(0.5) * (bx * bx + pr[6] * pr[6] + pr[7] * pr[7]);
  R slmul = sl - pl[1], srmur = sr - pr[1];
  R sM = (srmur * (pr[0] * pr[1]) - slmul * (pl[0] * pl[1]) - pTR + pTL) / (srmur * pr[0] - slmul * pl[0]);
  R rhostl = pl[0] * slmul / (sl - sM), rhostr = pr[0] * srmur / (sr - sM);
  R sstl = sM - rabs(bx) / rsqrt_(rhostl), sstr = sM + rabs(bx) / rsqrt_(rhostr);
  if (sstl >= R(0)) return 2;
  if (sstr <= R(0)) return 3;
  if (sM >= R(0)) return 4;
  if (sM <= R(0)) return 5;
  return 6;
}

template <typename R> int run(const char *tag, int n, unsigned seed) {
  Phys<R> c{8, 1, SOLVER_HLLD, 1, R(5.0 / 3.0), R(1.5)};
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> U(-1, 1);
  int bad = 0, reg[7] = {0};
  for (int t = 0; t < n; t++) {
    R pl[8], pr[8], f1[8], f2[8];
    double sc = (t % 3 == 0) ? 0.01 : 1.0;        // weak fields: star regions
    double vs = (t % 5 == 0) ? 5.0 : 1.0;         // fast flows: outer regions
    pl[0] = R(0.1 + std::abs(U(g)));
    pr[0] = R(0.1 + std::abs(U(g)));
    for (int q = 1; q < 4; q++) {
      pl[q] = R(U(g) * vs);
      pr[q] = R(U(g));
    }
    pl[4] = R(0.01 + std::abs(U(g)));
    pr[4] = R(0.01 + std::abs(U(g)));
    for (int q = 5; q < 8; q++) {
      pl[q] = R(U(g) * sc);
      pr[q] = R(U(g) * sc);
    }
    if (t % 7 == 0) pl[5] = pr[5] = R(0);         // zero normal field
    reg[region(c, pl, pr)]++;
    primfhlld(c, pl, pr, f1);
    primfhlld_sel(c, pl, pr, f2);
    if (memcmp(f1, f2, sizeof f1) != 0) {
      bad++;
      if (bad < 5) printf("%s mismatch at %d\n", tag, t);
    }
  }
  printf("%s regions: left %d right %d UL* %d UR* %d UL** %d UR** %d none %d\n", tag, reg[0], reg[1], reg[2],
         reg[3], reg[4], reg[5], reg[6]);
  return bad;
}

int main() {
  int b1 = run<double>("double", 1000000, 1);
  int b2 = run<float>("float", 1000000, 2);
  printf("double mismatches %d, float mismatches %d\n", b1, b2);
  return (b1 + b2) != 0;
}
