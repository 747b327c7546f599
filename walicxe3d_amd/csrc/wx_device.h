// wx_device.h - device-side configuration and the routines shared by the
// plain and fused forms: ghost-cell source selection (AMR-aware face
// exchange), wind sources, cooling.
#pragma once

#include <hip/hip_runtime.h>

#include "wx_common.h"
#include "wx_physics.h"

namespace wx {

constexpr double AMU = 1.660538782e-24;   // source/constants.f90:36
constexpr double KB = 1.380650400e-16;    // source/constants.f90:37
constexpr double PI_REF = 3.14159265358979;  // source/constants.f90:46
constexpr int MAXLEV = 16;
constexpr int MAXTAB = 512;
constexpr int MAXWIND = 8;

struct DevCfg {
  Phys<double> pd;
  Phys<float> pf;
  int nv, eight_wave, cooling, eos, nwind, ntab, maxlev, nblk;
  int bc[6];
  double visc_eta, cfl, cooling_limit, ion_thres, d_sc, e_sc, t_sc, p_sc, l_sc, v_sc;
  double temp_mu0;   // mu0*AMU*p_sc/d_sc/KB (calcTemp, hydro_core.f90:243)
  double temp_mui;   // mui*AMU*p_sc/d_sc/KB
  double mui_amu;    // mui*AMU (cooling.f90:340)
  double dx[3][MAXLEV];
  const BlockInfo *info;
  const double *tabT;   // device, ntab
  const double *tabC;   // device, ntab
  double winds[MAXWIND][8];
};

// ---------------------------------------------------------------------------
// Face exchange: where the value of a face ghost cell comes from.
// Ghost layer g (1 adjacent, 2 outer) of face f of block b at transverse
// position (a1, a2) in [1,16]^2; axes: n = f/2, (t1, t2) the two others in
// increasing order.  Follows normalBoundary (source/boundary.f90:328-668):
// same-level copy, coarse->fine injection, fine->coarse 2x2x2 average in the
// reference's summation order, free-flow (copy at low faces, mirror at high
// faces) and reflective (mirror, normal momentum negated) domain faces.
// ---------------------------------------------------------------------------
struct GhostSrc {
  int kind;        // 0 copy, 1 average of 8, 2 reflective copy
  int blk;         // source block
  int c[3];        // source cell (copy) or lowest corner of the 2x2x2 cell group
  int flip;        // variable negated (reflective), else -1
};

__host__ __device__ inline void face_axes(int f, int &n, int &t1, int &t2) {
  n = f >> 1;
  t1 = (n == 0) ? 1 : 0;
  t2 = (n == 2) ? 1 : 2;
}

// Returns false when the ghost is left untouched (periodic face without
// partner never happens: the host rejects such meshes).
__host__ __device__ inline bool ghost_source(const BlockInfo *info, const int *bc, int b, int f, int g, int a1,
                                             int a2, GhostSrc &s) {
  const BlockInfo &B = info[b];
  const Face &F = B.face[f];
  int n, t1, t2;
  face_axes(f, n, t1, t2);
  bool low = (f & 1) == 0;
  s.flip = -1;
  // source coordinates along the normal (cn) and the two transverse axes
  int cn, c1, c2;
  bool ok = true;
  switch (F.type) {
    case NEIGH_SAME:
      s.kind = 0;
      s.blk = F.src[0];
      cn = low ? (NC + 1 - g) : g;
      c1 = a1;
      c2 = a2;
      break;
    case NEIGH_COARSER:
      s.kind = 0;
      s.blk = F.src[0];
      cn = low ? NC : 1;
      c1 = (a1 + 1) / 2 + B.sib[t1] * (NC / 2);
      c2 = (a2 + 1) / 2 + B.sib[t2] * (NC / 2);
      break;
    case NEIGH_FINER: {
      int q1 = a1 > NC / 2, q2 = a2 > NC / 2;
      s.kind = 1;
      s.blk = F.src[q1 + 2 * q2];
      cn = low ? (NC + 1 - 2 * g) : (2 * g - 1);
      c1 = 2 * (a1 - q1 * (NC / 2)) - 1;
      c2 = 2 * (a2 - q2 * (NC / 2)) - 1;
      break;
    }
    default: {  // domain boundary
      int t = bc[f];
      s.blk = b;
      c1 = a1;
      c2 = a2;
      s.kind = 0;
      if (t == BC_REFLECTIVE) {
        s.kind = 2;
        cn = low ? g : (NC + 1 - g);
        s.flip = 1 + n;
      } else if (t == BC_FREE_FLOW) {
        cn = low ? 1 : (NC + 1 - g);
      } else {
        cn = 1;
        ok = false;
      }
      break;
    }
  }
  // scatter (cn, c1, c2) to x, y, z without run-time array indexing
  s.c[0] = n == 0 ? cn : c1;
  s.c[1] = n == 0 ? c1 : (n == 1 ? cn : c2);
  s.c[2] = n == 2 ? cn : c2;
  return ok;
}

// Value of variable v of a face ghost from its source, reading an array
// with ghosts ([blk][nv][20^3]) of element type T, computing in R.
template <typename R, typename T>
__device__ inline R ghost_value(const T *A, int nv, const GhostSrc &s, int v) {
  const T *S = A + ((size_t)s.blk * nv + v) * NCELL;
  if (s.kind == 1) {
    R acc = R(0);
    for (int dx = 0; dx < 2; dx++)
      for (int dy = 0; dy < 2; dy++)
        for (int dz = 0; dz < 2; dz++)
          acc = acc + R(S[cid(s.c[0] + dx, s.c[1] + dy, s.c[2] + dz)]) / R(8.0);
    return acc;
  }
  R val = R(S[cid(s.c[0], s.c[1], s.c[2])]);
  if (s.kind == 2 && v == s.flip) val = R(-1.0) * val;
  return val;
}

// ---------------------------------------------------------------------------
// Spherical wind sources (imposeSphericalWind, source/winds.f90:123-240):
// cells whose centre lies within the source radius get the wind state.
// cellPos (source/amr.f90:1037-1065) gives the centre.
// ---------------------------------------------------------------------------
__device__ inline bool wind_state(const DevCfg &c, int b, int i, int j, int k, double *u) {
  const BlockInfo &B = c.info[b];
  int l = B.level - 1;
  double x = ((B.coords[0] - 1) * NC + (i - 1) + 0.5) * c.dx[0][l];
  double y = ((B.coords[1] - 1) * NC + (j - 1) + 0.5) * c.dx[1][l];
  double z = ((B.coords[2] - 1) * NC + (k - 1) + 0.5) * c.dx[2][l];
  x = x * c.l_sc;
  y = y * c.l_sc;
  z = z * c.l_sc;
  bool hit = false;
  for (int w = 0; w < c.nwind; w++) {
    const double *wp = c.winds[w];
    double dist = sqrt((x - wp[0]) * (x - wp[0]) + (y - wp[1]) * (y - wp[1]) + (z - wp[2]) * (z - wp[2]));
    if (dist < wp[3]) {
      double dens = wp[4] / wp[5] / dist / dist / (4.0 * PI_REF);
      double vx = wp[5] * (x - wp[0]) / dist + 0.0;
      double vy = wp[5] * (y - wp[1]) / dist + 0.0;
      double vz = wp[5] * (z - wp[2]) / dist + 0.0;
      double pres = dens / (wp[7] * AMU) * KB * wp[6];
      double p[MAXV] = {0, 0, 0, 0, 0, 0, 0, 0};
      p[0] = dens / c.d_sc;
      p[1] = vx / c.v_sc;
      p[2] = vy / c.v_sc;
      p[3] = vz / c.v_sc;
      p[4] = pres / c.p_sc;
      prim2flow(c.pd, p, u);
      hit = true;
    }
  }
  return hit;
}

// A plane wind (imposePlaneWind, source/winds.f90:250-340), imposed after the
// spherical sources as a userBoundary calls them: cells k = 16..18 (every i,
// j) of each block without a TOP neighbour, which neighborLevel reports at
// the domain's top edge whatever the boundary type (source/amr.f90:439-523);
// the plane only picks the velocity's axis and sign (winds.f90:300-313).
// plane = 0: none.  Hydro only (the reference leaves the state's field
// unset).  Plain FP64 form.
struct PlaneWind {
  int plane;        // PLANE_LEFT=1 .. PLANE_TOP=6 (source/winds.f90:93-98)
  int nroot_z;      // root blocks along z
  double rho, vel, temp, mu;   // g cm^-3, cm s^-1, K, mean mass per particle in AMU
};

__device__ inline bool plane_wind_state(const DevCfg &c, const PlaneWind &pw, int b, int k, double *u) {
  if (pw.plane == 0 || k < NC) return false;
  const BlockInfo &B = c.info[b];
  if (B.coords[2] != (pw.nroot_z << (B.level - 1))) return false;
  double vx = 0.0, vy = 0.0, vz = 0.0;
  switch (pw.plane) {
    case 1: vx = pw.vel; break;
    case 2: vx = -pw.vel; break;
    case 3: vy = pw.vel; break;
    case 4: vy = -pw.vel; break;
    case 5: vz = pw.vel; break;
    default: vz = -pw.vel; break;
  }
  const double pres = pw.rho / (pw.mu * AMU) * KB * pw.temp;
  double p[MAXV] = {0, 0, 0, 0, 0, 0, 0, 0};
  p[0] = pw.rho / c.d_sc;
  p[1] = vx / c.v_sc;
  p[2] = vy / c.v_sc;
  p[3] = vz / c.v_sc;
  p[4] = pres / c.p_sc;
  prim2flow(c.pd, p, u);
  return true;
}

// A uniform acceleration g (code units) as a problem's get_user_source_terms
// (called per cell by source_function, source/sources.f90:100-133, in
// upwindStep, source/godunov.f90:266-271) would add it: s(2:4) += rho*g,
// s(5) += rho*(v.g), before the Powell terms.  The hook receives a cell's
// primitives and in-block indices only, so a source of that interface cannot
// depend on the block's position.  on = 0: none.  Plain FP64 form.
struct UserSource {
  int on;
  double g[3];
};

__device__ inline void user_source(const UserSource &us, const double *pp, double *s) {
  s[1] = s[1] + pp[0] * us.g[0];
  s[2] = s[2] + pp[0] * us.g[1];
  s[3] = s[3] + pp[0] * us.g[2];
  s[4] = s[4] + pp[0] * (pp[1] * us.g[0] + pp[2] * us.g[1] + pp[3] * us.g[2]);
}

// ---------------------------------------------------------------------------
// Cooling (apply_cooling, source/cooling.f90:285-371; find_coef :378-419;
// calcTemp, source/hydro_core.f90:231-270).  Updates p and returns E.
// ---------------------------------------------------------------------------
__device__ inline double find_coef(const DevCfg &c, double logT) {
  const double *T = c.tabT, *C = c.tabC;
  int n = c.ntab;
  if (logT < T[0]) return 0.0;
  if (logT >= T[n - 1]) return pow(10.0, C[n - 1] + 0.5 * (logT - T[n - 1]));
  // first node above logT: binary search gives the same node as the
  // reference's linear scan of a strictly increasing table
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (T[mid] > logT) hi = mid; else lo = mid;
  }
  double T0 = T[hi - 1], C0 = C[hi - 1], T1 = T[hi], C1 = C[hi];
  return pow(10.0, (C1 - C0) / (T1 - T0) * (logT - T0) + C0);
}

__device__ inline double calc_temp(const DevCfg &c, double rho, double p) {
  double t = p / rho * c.temp_mu0;
  if (c.eos == 2 && t > c.ion_thres) t = p / rho * c.temp_mui;
  return t;
}

// One cell; prim holds (rho, u, v, w, p).  Returns true when the cell
// changed, with new p in prim[4] and new total energy in *e.  have_t: the
// caller's temperature t_in replaces calcTemp's (EOS_H_RATE, which needs a
// passive scalar; plain form).
__device__ inline bool cool_cell(const DevCfg &c, double dt, double *prim, double *e, bool have_t = false,
                                 double t_in = 0.0) {
  double temp = have_t ? t_in : calc_temp(c, prim[0], prim[4]);
  double logT = log10(temp);
  if (!(logT >= c.tabT[0])) return false;
  double ETH = c.pd.cv * prim[4];
  double vel2 = prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3];
  double EK = 0.5 * prim[0] * vel2;
  double aloss = find_coef(c, logT);
  double numdens = prim[0] * c.d_sc / c.mui_amu;
  double ce = aloss * (numdens * numdens) / (ETH * c.e_sc);
  double cool_factor = exp(-ce * (dt * c.t_sc));
  if (cool_factor < 1.0 - c.cooling_limit) cool_factor = 1.0 - c.cooling_limit;
  double new_temp = temp * cool_factor;
  if (new_temp < 10.0) cool_factor = 10.0 / temp;
  prim[4] = prim[4] * cool_factor;
  ETH = c.pd.cv * prim[4];
  *e = EK + ETH;
  return true;
}

}  // namespace wx
