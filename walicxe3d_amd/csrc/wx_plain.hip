// wx_plain.hip - the plain FP64 form: one kernel per reference routine,
// arrays with two ghost layers per block ([blk][nv][20^3], i fastest),
// arithmetic in the reference's order (built with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include "wx_device.h"
#include "wx_kernels.h"

namespace wx {

constexpr int NIF = 17 * NC * NC;   // interfaces per direction per block used by upwindStep

// ---------------------------------------------------------------------------
// The device physics routines applied to arrays of records (one thread per
// record), for parity tests of single routines.  op (EVAL_*):
//   flow2prim   in nv -> out nv + 1 (primitives, istat)
//   prim2flow   in nv -> out nv
//   prim2fluxes in nv -> out nv, arg = direction 1..3
//   limiter     in 4 nv (pll, pl, pr, prr) -> out 2 nv (pl, pr), arg = limiter
//   riemann     in 2 nv (pl, pr) -> out nv, arg = solver (HLLD in the
//               reference's branch form)
//   riemann_sel in 2 nv -> out nv: HLLD in its select form (fused FP64 form)
//   speeds      in nv -> out 5 (sound, cfast x/y/z, cfastX)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_eval(Phys<double> P, int op, int arg, int n, const double *in,
                                              double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int nv = P.nv;
  double a[4 * MAXV], f[2 * MAXV];
  const int nin = op == EVAL_LIMITER ? 4 * nv : (op == EVAL_RIEMANN || op == EVAL_RIEMANN_SEL ? 2 * nv : nv);
  for (int q = 0; q < nin; q++) a[q] = in[(size_t)t * nin + q];
  int nout = nv;
  switch (op) {
    case EVAL_FLOW2PRIM:
      f[nv] = (double)flow2prim(P, a, f);
      nout = nv + 1;
      break;
    case EVAL_PRIM2FLOW:
      prim2flow(P, a, f);
      break;
    case EVAL_PRIM2FLUXES:
      swapdim(nv, arg - 1, a);
      xflux(P, a, f);
      swapdim(nv, arg - 1, f);
      break;
    case EVAL_LIMITER:
      for (int q = 0; q < nv; q++) {
        double pl = a[nv + q], pr = a[2 * nv + q];
        limit1(a[q], pl, pr, a[3 * nv + q], arg);
        f[q] = pl;
        f[nv + q] = pr;
      }
      nout = 2 * nv;
      break;
    case EVAL_RIEMANN:
      // P keeps the build's solver: primfhllc forms its contact speed only in
      // HLLC builds (source/hllc.f90:348-353)
      if (arg == SOLVER_HLLC) primfhllc(P, a, a + nv, f);
      else if (arg == SOLVER_HLLE) primfhlle(P, a, a + nv, f);
      else if (arg == SOLVER_HLLD) primfhlld(P, a, a + nv, f);
      else primfhll(P, a, a + nv, f);
      break;
    case EVAL_RIEMANN_SEL:
      primfhlld_sel<double>(P, a, a + nv, f);
      break;
    default:   // EVAL_SPEEDS
      f[0] = sound(P, a);
      f[1] = f[2] = f[3] = f[4] = 0.0;
      if (nv == 8) {
        f[1] = cfast1(P, a[4], a[0], a[5], a[6], a[7], a[5]);
        f[2] = cfast1(P, a[4], a[0], a[5], a[6], a[7], a[6]);
        f[3] = cfast1(P, a[4], a[0], a[5], a[6], a[7], a[7]);
        f[4] = cfastX(P, a);
      }
      nout = 5;
      break;
  }
  for (int q = 0; q < nout; q++) out[(size_t)t * nout + q] = f[q];
}

void plain_eval(const Phys<double> &P, int op, int arg, int n, const double *in, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P, op, arg, n, in, out);
}

// ---------------------------------------------------------------------------
// boundary (source/boundary.f90:46-64): face ghosts of depth `depth`.
// grid (6*depth, nblk), block 256 = transverse 16x16.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_boundary(DevCfg c, double *A, int depth) {
  int b = blockIdx.y;
  int f = blockIdx.x / depth;
  int g = blockIdx.x % depth + 1;
  int a1 = threadIdx.x % NC + 1, a2 = threadIdx.x / NC + 1;
  GhostSrc s;
  if (!ghost_source(c.info, c.bc, b, f, g, a1, a2, s)) return;
  int n, t1, t2;
  face_axes(f, n, t1, t2);
  int d[3];
  d[n] = (f & 1) ? NC + g : 1 - g;
  d[t1] = a1;
  d[t2] = a2;
  int q = cid(d[0], d[1], d[2]);
  int nv = c.nv;
  double val[MAXV];
  for (int v = 0; v < nv; v++) val[v] = ghost_value<double>(A, nv, s, v);
  for (int v = 0; v < nv; v++) A[((size_t)b * nv + v) * NCELL + q] = val[v];
}

// Wind sources over interior and face-ghost cells (userBoundary).
// grid (28, nblk): 16 blocks of 256 interior cells, 12 of face ghosts.
__global__ void __launch_bounds__(256) k_winds(DevCfg c, double *A, PlaneWind pw) {
  int b = blockIdx.y;
  int i, j, k;
  if (blockIdx.x < 16) {
    int t = blockIdx.x * 256 + threadIdx.x;
    i = t % NC + 1; j = (t / NC) % NC + 1; k = t / (NC * NC) + 1;
  } else {
    int f = (blockIdx.x - 16) / 2, g = (blockIdx.x - 16) % 2 + 1;
    int n, t1, t2, d[3];
    face_axes(f, n, t1, t2);
    d[n] = (f & 1) ? NC + g : 1 - g;
    d[t1] = threadIdx.x % NC + 1;
    d[t2] = threadIdx.x / NC + 1;
    i = d[0]; j = d[1]; k = d[2];
  }
  double u[MAXV];
  bool hit = wind_state(c, b, i, j, k, u);
  if (plane_wind_state(c, pw, b, k, u)) hit = true;   // after the spherical sources
  if (hit) {
    int q = cid(i, j, k);
    for (int v = 0; v < c.nv; v++) A[((size_t)b * c.nv + v) * NCELL + q] = u[v];
  }
}

// ---------------------------------------------------------------------------
// calcPrimsAll (source/hydro_core.f90:67-223): interior (cells ALL/PHYS)
// and face ghosts with non-zero density (ALL/GHOST).  Edge and corner
// ghosts are never read by the update and are skipped.
// grid (28, nblk)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_prims(DevCfg c, const double *A, double *P, int cells,
                                               unsigned long long *bad) {
  int b = blockIdx.y;
  int i, j, k;
  bool ghost = blockIdx.x >= 16;
  if (!ghost) {
    if (cells == 2) return;
    int t = blockIdx.x * 256 + threadIdx.x;
    i = t % NC + 1; j = (t / NC) % NC + 1; k = t / (NC * NC) + 1;
  } else {
    if (cells == 1) return;
    int f = (blockIdx.x - 16) / 2, g = (blockIdx.x - 16) % 2 + 1;
    int n, t1, t2, d[3];
    face_axes(f, n, t1, t2);
    d[n] = (f & 1) ? NC + g : 1 - g;
    d[t1] = threadIdx.x % NC + 1;
    d[t2] = threadIdx.x / NC + 1;
    i = d[0]; j = d[1]; k = d[2];
  }
  int nv = c.nv, q = cid(i, j, k);
  const double *Ab = A + (size_t)b * nv * NCELL;
  double u[MAXV], p[MAXV];
  if (ghost && Ab[q] == 0.0) return;
  for (int v = 0; v < nv; v++) u[v] = Ab[v * NCELL + q];
  int st = flow2prim(c.pd, u, p);
  for (int v = c.pd.nv; v < nv; v++) p[v] = u[v];   // passive scalars (hydro_core.f90:372-374)
  double *Pb = P + (size_t)b * nv * NCELL;
  for (int v = 0; v < nv; v++) Pb[v * NCELL + q] = p[v];
  if (st && bad && !ghost) atomicAdd(bad, 1ull);
}

// ---------------------------------------------------------------------------
// Passive scalars of the hydro solvers, variables pd.nv..nv-1, on states
// already turned to the x direction: a passive's physical flux is u*s
// (prim2fluxes, hydro_core.f90:429-431) and its conserved value s
// (prim2flow :307-309).  HLL and HLL1 take the HLL average
// (hll.f90:186-212); HLLC the star state rho* s_K / rho_K of the side the
// contact leaves (hllc.f90:232-296: 262-264, 283-285).  The wave speeds
// repeat those of the solver, with the same arithmetic.
// ---------------------------------------------------------------------------
__device__ inline void passive_fluxes(const Phys<double> &c, int nv, const double *pL, const double *pR,
                                      double *ff) {
  const double cl = sound(c, pL), cr = sound(c, pR);
  if (c.solver == SOLVER_HLLC) {
    const double dl1 = pL[0], dr1 = pR[0], ul1 = pL[1], ur1 = pR[1], pl1 = pL[4], pr1 = pR[4];
    const double sl = rfmin(ul1 - cl, ur1 - cr);
    const double sr = rfmax(ul1 + cl, ur1 + cr);
    const double sst =
        (pr1 - pl1 + dl1 * ul1 * (sl - ul1) - dr1 * ur1 * (sr - ur1)) / (dl1 * (sl - ul1) - dr1 * (sr - ur1));
    if (sl > 0.0) {
      for (int q = c.nv; q < nv; q++) ff[q] = pL[1] * pL[q];
      return;
    }
    if (sr < 0.0) {
      for (int q = c.nv; q < nv; q++) ff[q] = pR[1] * pR[q];
      return;
    }
    const double *pK;
    double sK;
    if (sst >= 0.0) { pK = pL; sK = sl; }
    else if (sst <= 0.0) { pK = pR; sK = sr; }
    else {
      for (int q = c.nv; q < nv; q++) ff[q] = NAN;
      return;
    }
    const double rhost = pK[0] * (sK - pK[1]) / (sK - sst);
    for (int q = c.nv; q < nv; q++) ff[q] = pK[1] * pK[q] + sK * (rhost * pK[q] / pK[0] - pK[q]);
    return;
  }
  const double sl = rfmin(pL[1] - cl, pR[1] - cr);
  const double sr = rfmax(pL[1] + cl, pR[1] + cr);
  if (sl >= 0.0) {
    for (int q = c.nv; q < nv; q++) ff[q] = pL[1] * pL[q];
    return;
  }
  if (sr <= 0.0) {
    for (int q = c.nv; q < nv; q++) ff[q] = pR[1] * pR[q];
    return;
  }
  for (int q = c.nv; q < nv; q++)
    ff[q] = (sr * (pL[1] * pL[q]) - sl * (pR[1] * pR[q]) + sl * sr * (pR[q] - pL[q])) / (sr - sl);
}

// ---------------------------------------------------------------------------
// HLL*fluxes (hll.f90:52-167, hllc.f90:55-201, hlle.f90:54-158,
// hlld.f90:54-158): the 3 x 17 x 16 x 16 interfaces upwindStep reads.
// grid (3*17, nblk), block 256; F layout [blk][dim][nv][17*16*16].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fluxes(DevCfg c, const double *P, double *F, int order) {
  int b = blockIdx.y;
  int dim = blockIdx.x / 17;
  int s = blockIdx.x % 17;          // position along dim: interface (s | s+1)
  int a1 = threadIdx.x % NC + 1, a2 = threadIdx.x / NC + 1;
  int n, t1, t2;
  face_axes(2 * dim, n, t1, t2);
  int base[3];
  base[n] = s;
  base[t1] = a1;
  base[t2] = a2;
  int nv = c.nv;
  const double *Pb = P + (size_t)b * nv * NCELL;
  int st = (dim == 0) ? 1 : (dim == 1) ? NB : NB * NB;
  int q0 = cid(base[0], base[1], base[2]);
  double pl[MAXV], pr[MAXV], ff[MAXV];
  for (int v = 0; v < nv; v++) {
    pl[v] = Pb[v * NCELL + q0];
    pr[v] = Pb[v * NCELL + q0 + st];
  }
  swapdim(nv, dim, pl);
  swapdim(nv, dim, pr);
  if (order == 2) {
    double pll[MAXV], prr[MAXV];
    for (int v = 0; v < nv; v++) {
      pll[v] = Pb[v * NCELL + q0 - st];
      prr[v] = Pb[v * NCELL + q0 + 2 * st];
    }
    swapdim(nv, dim, pll);
    swapdim(nv, dim, prr);
    for (int v = 0; v < nv; v++) limit1(pll[v], pl[v], pr[v], prr[v], c.pd.limiter);
  }
  riemann(c.pd, pl, pr, ff);
  if (nv > c.pd.nv) passive_fluxes(c.pd, nv, pl, pr, ff);
  swapdim(nv, dim, ff);
  // interface index: position s fastest along dim, then t1, then t2
  int idx = ((a2 - 1) * NC + (a1 - 1)) * 17 + s;
  double *Fb = F + ((size_t)b * 3 + dim) * nv * NIF;
  for (int v = 0; v < nv; v++) Fb[(size_t)v * NIF + idx] = ff[v];
}

__device__ inline int fidx(int s, int a1, int a2) { return ((a2 - 1) * NC + (a1 - 1)) * 17 + s; }

// ---------------------------------------------------------------------------
// upwindStep (source/godunov.f90:199-277) with source_function
// (source/sources.f90:100-133: the user terms, then Powell's :38-95).
// grid (16, nblk).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_upwind(DevCfg c, const double *U, const double *P, const double *F,
                                                double *UP, double dtp, UserSource us) {
  int b = blockIdx.y;
  int t = blockIdx.x * 256 + threadIdx.x;
  int i = t % NC + 1, j = (t / NC) % NC + 1, k = t / (NC * NC) + 1;
  int nv = c.nv, l = c.info[b].level - 1;
  double dx = c.dx[0][l], dy = c.dx[1][l], dz = c.dx[2][l];
  double dtdx = dtp / dx, dtdy = dtp / dy, dtdz = dtp / dz;
  const double *Fx = F + ((size_t)b * 3 + 0) * nv * NIF;
  const double *Fy = F + ((size_t)b * 3 + 1) * nv * NIF;
  const double *Fz = F + ((size_t)b * 3 + 2) * nv * NIF;
  int q = cid(i, j, k);
  const double *Ub = U + (size_t)b * nv * NCELL;
  double *UPb = UP + (size_t)b * nv * NCELL;
  // x: (s=i, t1=j, t2=k); y: (s=j, t1=i, t2=k); z: (s=k, t1=i, t2=j)
  int ix0 = fidx(i - 1, j, k), ix1 = fidx(i, j, k);
  int iy0 = fidx(j - 1, i, k), iy1 = fidx(j, i, k);
  int iz0 = fidx(k - 1, i, j), iz1 = fidx(k, i, j);
  double upv[MAXV];
  for (int v = 0; v < nv; v++) {
    size_t o = (size_t)v * NIF;
    upv[v] = Ub[v * NCELL + q] + dtdx * (Fx[o + ix0] - Fx[o + ix1]) + dtdy * (Fy[o + iy0] - Fy[o + iy1]) +
             dtdz * (Fz[o + iz0] - Fz[o + iz1]);
  }
  const bool powell = c.eight_wave && nv == 8;
  if (powell || us.on) {
    const double *Pb = P + (size_t)b * nv * NCELL;
    double pp[MAXV];
    for (int v = 0; v < nv; v++) pp[v] = Pb[v * NCELL + q];
    double s[MAXV] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (us.on) user_source(us, pp, s);
    if (powell) {
      double divB = (Pb[5 * NCELL + q + 1] - Pb[5 * NCELL + q - 1]) / (2.0 * dx) +
                    (Pb[6 * NCELL + q + NB] - Pb[6 * NCELL + q - NB]) / (2.0 * dy) +
                    (Pb[7 * NCELL + q + NB * NB] - Pb[7 * NCELL + q - NB * NB]) / (2.0 * dz);
      s[1] = s[1] - divB * pp[5];
      s[2] = s[2] - divB * pp[6];
      s[3] = s[3] - divB * pp[7];
      s[4] = s[4] - divB * (pp[1] * pp[5] + pp[2] * pp[6] + pp[3] * pp[7]);
      s[5] = s[5] - divB * pp[1];
      s[6] = s[6] - divB * pp[2];
      s[7] = s[7] - divB * pp[3];
    }
    for (int v = 0; v < nv; v++) upv[v] = upv[v] + dtp * s[v];
  }
  for (int v = 0; v < nv; v++) UPb[v * NCELL + q] = upv[v];
}

// ---------------------------------------------------------------------------
// doViscousStep (source/hydro_solver.f90:217-250).  grid (16, nblk).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_viscous(DevCfg c, const double *UP, double *U) {
  int b = blockIdx.y;
  int t = blockIdx.x * 256 + threadIdx.x;
  int i = t % NC + 1, j = (t / NC) % NC + 1, k = t / (NC * NC) + 1;
  int q = cid(i, j, k);
  double eta = c.visc_eta;
  for (int v = 0; v < c.nv; v++) {
    const double *A = UP + ((size_t)b * c.nv + v) * NCELL;
    U[((size_t)b * c.nv + v) * NCELL + q] =
        A[q] + eta * (+A[q + 1] + A[q - 1] + A[q + NB] + A[q - NB] + A[q + NB * NB] + A[q - NB * NB] - 6.0 * A[q]);
  }
}

// ---------------------------------------------------------------------------
// getTimestep (source/hydro_solver.f90:129-211): min over interior cells.
// Positive doubles order like their bit patterns, so the global minimum is
// an unsigned 64-bit atomicMin on the bits.  grid (16, nblk).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_timestep(DevCfg c, const double *P, unsigned long long *dtbits) {
  int b = blockIdx.y;
  int t = blockIdx.x * 256 + threadIdx.x;
  int i = t % NC + 1, j = (t / NC) % NC + 1, k = t / (NC * NC) + 1;
  int nv = c.nv, l = c.info[b].level - 1;
  const double *Pb = P + (size_t)b * nv * NCELL;
  int q = cid(i, j, k);
  double p[MAXV];
  for (int v = 0; v < nv; v++) p[v] = Pb[v * NCELL + q];
  double dt = 1.0e30;
  if (c.pd.mhd && nv == 8) {
    double cfx = cfast1(c.pd, p[4], p[0], p[5], p[6], p[7], p[5]);
    double cfy = cfast1(c.pd, p[4], p[0], p[5], p[6], p[7], p[6]);
    double cfz = cfast1(c.pd, p[4], p[0], p[5], p[6], p[7], p[7]);
    dt = fmin(dt, c.dx[0][l] / (fabs(p[1]) + cfx));
    dt = fmin(dt, c.dx[1][l] / (fabs(p[2]) + cfy));
    dt = fmin(dt, c.dx[2][l] / (fabs(p[3]) + cfz));
  } else {
    double cs = sound(c.pd, p);
    dt = fmin(dt, c.dx[0][l] / (fabs(p[1]) + cs));
    dt = fmin(dt, c.dx[1][l] / (fabs(p[2]) + cs));
    dt = fmin(dt, c.dx[2][l] / (fabs(p[3]) + cs));
  }
  // wavefront min, then one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
  if ((threadIdx.x & 63) == 0) atomicMin(dtbits, (unsigned long long)__double_as_longlong(dt));
}

// ---------------------------------------------------------------------------
// cooling (source/cooling.f90:37-118, apply_cooling :285-371).  grid (16, nblk)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cooling(DevCfg c, double *U, double *P, double dt, double mu0) {
  int b = blockIdx.y;
  int t = blockIdx.x * 256 + threadIdx.x;
  int i = t % NC + 1, j = (t / NC) % NC + 1, k = t / (NC * NC) + 1;
  int nv = c.nv, q = cid(i, j, k);
  double *Pb = P + (size_t)b * nv * NCELL;
  double pr[5];
  for (int v = 0; v < 5; v++) pr[v] = Pb[v * NCELL + q];
  // EOS_H_RATE (calcTemp, hydro_core.f90:258-264): P = (n_HI + n_HII) k T
  // with the neutral hydrogen density in the first passive scalar
  const bool hrate = c.eos == 3;
  double th = 0.0;
  if (hrate) th = pr[4] / (2.0 * pr[0] / mu0 - Pb[5 * NCELL + q]) * c.p_sc / KB;
  double e;
  if (cool_cell(c, dt, pr, &e, hrate, th)) {
    Pb[4 * NCELL + q] = pr[4];
    U[((size_t)b * nv + 4) * NCELL + q] = e;
  }
}

// ---------------------------------------------------------------------------
// markByPhysical (source/amr.f90:142-230), the per-block part: the largest
// |q(i+1) - q(i-1)| / q(i) / (2 dx) along x, y and z over the interior
// cells, for q = pressure and q = density, read from PRIM with its depth-1
// face ghosts.  dx is the level's x spacing in all three directions, as in
// the reference.  The reference returns at the first cell whose gradient
// reaches the refinement threshold; which flag it returns then depends only
// on the level, so the flag is a function of this maximum (set on the host).
// grid nblk, block 256: lane (i, j) walks k.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mark(DevCfg c, const double *P, double *maxgrad) {
  __shared__ double part[4];
  const int b = blockIdx.x;
  const int i = (threadIdx.x & (NC - 1)) + 1, j = threadIdx.x / NC + 1;
  const double two_dx = 2.0 * c.dx[0][c.info[b].level - 1];
  double m = 0.0;
  for (int s = 0; s < 2; s++) {
    const double *A = P + ((size_t)b * c.nv + (s == 0 ? 4 : 0)) * NCELL;
    for (int k = 1; k <= NC; k++) {
      const int q = cid(i, j, k);
      const double a = A[q];
      const double gx = fabs(A[q + 1] - A[q - 1]) / a / two_dx;
      const double gy = fabs(A[q + NB] - A[q - NB]) / a / two_dx;
      const double gz = fabs(A[q + NB * NB] - A[q - NB * NB]) / a / two_dx;
      m = fmax(m, fmax(gx, fmax(gy, gz)));
    }
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) maxgrad[b] = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
}

// ---------------------------------------------------------------------------
// The same for the fused form, whose state has no ghost layers
// ([blk][nv][16^3] FP64): density and pressure of the depth-1 face ghosts
// are formed from the sources the exchange reads (same-level copy,
// coarse->fine injection, fine->coarse 8-cell average, domain faces) and
// the wind sources, with the plain form's arithmetic, then flow2prim.
// ---------------------------------------------------------------------------
__device__ inline double ghost_value_int(const double *U, int nv, const GhostSrc &s, int v) {
  const double *S = U + ((size_t)s.blk * nv + v) * NINT;
  if (s.kind == 1) {
    double acc = 0.0;
    for (int dx = 0; dx < 2; dx++)
      for (int dy = 0; dy < 2; dy++)
        for (int dz = 0; dz < 2; dz++)
          acc = acc + S[((s.c[2] + dz - 1) * NC + (s.c[1] + dy - 1)) * NC + (s.c[0] + dx - 1)] / 8.0;
    return acc;
  }
  double val = S[((s.c[2] - 1) * NC + (s.c[1] - 1)) * NC + (s.c[0] - 1)];
  if (s.kind == 2 && v == s.flip) val = -1.0 * val;
  return val;
}

// density and pressure of ghost (a1, a2) of face f, layer 1, of block b
__device__ inline void ghost_rho_p(const DevCfg &c, const double *U, int b, int f, int a1, int a2, double &rho,
                                   double &pres) {
  GhostSrc s;
  double u[MAXV], p[MAXV];
  if (!ghost_source(c.info, c.bc, b, f, 1, a1, a2, s)) {
    rho = pres = 0.0;
    return;
  }
  for (int v = 0; v < c.nv; v++) u[v] = ghost_value_int(U, c.nv, s, v);
  int n, t1, t2, d[3];
  face_axes(f, n, t1, t2);
  d[n] = (f & 1) ? NC + 1 : 0;
  d[t1] = a1;
  d[t2] = a2;
  wind_state(c, b, d[0], d[1], d[2], u);
  flow2prim(c.pd, u, p);
  rho = p[0];
  pres = p[4];
}

// Recorded ghost layer: G[blk][2 (rho, p)][6 faces][256].  grid (6, nblk).
__global__ void __launch_bounds__(256) k_ghost_prims_fused(DevCfg c, const double *U, double *G) {
  const int b = blockIdx.y, f = blockIdx.x;
  double rho, pres;
  ghost_rho_p(c, U, b, f, threadIdx.x % NC + 1, threadIdx.x / NC + 1, rho, pres);
  G[(((size_t)b * 2 + 0) * 6 + f) * 256 + threadIdx.x] = rho;
  G[(((size_t)b * 2 + 1) * 6 + f) * 256 + threadIdx.x] = pres;
}

// markByPhysical on the fused state: density and pressure of the block and
// its depth-1 face ghosts (from G, or formed from U when G is null) staged
// in LDS as an 18^3 cube, then the maximum of k_mark.  grid nblk, 1024
// threads (93 KB of LDS: one workgroup of 16 waves per CU).
__global__ void __launch_bounds__(1024) k_mark_fused(DevCfg c, const double *U, const double *G, double *maxgrad) {
  constexpr int E = NC + 2;
  __shared__ double sr[E * E * E], sp[E * E * E];
  __shared__ double part[16];
  const int b = blockIdx.x, t = threadIdx.x, nv = c.nv;
  for (int q = t; q < NINT; q += 1024) {
    double u[MAXV], p[MAXV];
    for (int v = 0; v < nv; v++) u[v] = U[((size_t)b * nv + v) * NINT + q];
    flow2prim(c.pd, u, p);
    const int i = q % NC + 1, j = (q / NC) % NC + 1, k = q / (NC * NC) + 1;
    sr[(k * E + j) * E + i] = p[0];
    sp[(k * E + j) * E + i] = p[4];
  }
  for (int r = t; r < 6 * 256; r += 1024) {
    const int f = r >> 8, a1 = (r & 255) % NC + 1, a2 = (r & 255) / NC + 1;
    double rho, pres;
    if (G) {
      rho = G[(((size_t)b * 2 + 0) * 6 + f) * 256 + (r & 255)];
      pres = G[(((size_t)b * 2 + 1) * 6 + f) * 256 + (r & 255)];
    } else {
      ghost_rho_p(c, U, b, f, a1, a2, rho, pres);
    }
    int n, t1, t2, d[3];
    face_axes(f, n, t1, t2);
    d[n] = (f & 1) ? NC + 1 : 0;
    d[t1] = a1;
    d[t2] = a2;
    sr[(d[2] * E + d[1]) * E + d[0]] = rho;
    sp[(d[2] * E + d[1]) * E + d[0]] = pres;
  }
  __syncthreads();
  const double two_dx = 2.0 * c.dx[0][c.info[b].level - 1];
  double m = 0.0;
  for (int s2 = 0; s2 < 2; s2++) {
    const double *A = s2 == 0 ? sp : sr;
    for (int q = t; q < NINT; q += 1024) {
      const int o = ((q / (NC * NC) + 1) * E + (q / NC) % NC + 1) * E + q % NC + 1;
      const double a = A[o];
      const double gx = fabs(A[o + 1] - A[o - 1]) / a / two_dx;
      const double gy = fabs(A[o + E] - A[o - E]) / a / two_dx;
      const double gz = fabs(A[o + E * E] - A[o - E * E]) / a / two_dx;
      m = fmax(m, fmax(gx, fmax(gy, gz)));
    }
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((t & 63) == 0) part[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    double r = part[0];
    for (int w = 1; w < 16; w++) r = fmax(r, part[w]);
    maxgrad[b] = r;
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
void plain_boundary(const DevCfg &c, const PlaneWind &pw, double *A, int depth, hipStream_t st) {
  hipLaunchKernelGGL(k_boundary, dim3(6 * depth, c.nblk), dim3(256), 0, st, c, A, depth);
  if (c.nwind > 0 || pw.plane != 0) hipLaunchKernelGGL(k_winds, dim3(28, c.nblk), dim3(256), 0, st, c, A, pw);
}
void plain_prims(const DevCfg &c, const double *A, double *P, int cells, unsigned long long *bad,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_prims, dim3(28, c.nblk), dim3(256), 0, st, c, A, P, cells, bad);
}
void plain_fluxes(const DevCfg &c, const double *P, double *F, int order, hipStream_t st) {
  hipLaunchKernelGGL(k_fluxes, dim3(3 * 17, c.nblk), dim3(256), 0, st, c, P, F, order);
}
void plain_upwind(const DevCfg &c, const UserSource &us, const double *U, const double *P, const double *F,
                  double *UP, double dtp, hipStream_t st) {
  hipLaunchKernelGGL(k_upwind, dim3(16, c.nblk), dim3(256), 0, st, c, U, P, F, UP, dtp, us);
}
void plain_viscous(const DevCfg &c, const double *UP, double *U, hipStream_t st) {
  hipLaunchKernelGGL(k_viscous, dim3(16, c.nblk), dim3(256), 0, st, c, UP, U);
}
void plain_timestep(const DevCfg &c, const double *P, unsigned long long *dtbits, hipStream_t st) {
  hipLaunchKernelGGL(k_timestep, dim3(16, c.nblk), dim3(256), 0, st, c, P, dtbits);
}
void plain_cooling(const DevCfg &c, double *U, double *P, double dt, double mu0, hipStream_t st) {
  hipLaunchKernelGGL(k_cooling, dim3(16, c.nblk), dim3(256), 0, st, c, U, P, dt, mu0);
}
void plain_mark(const DevCfg &c, const double *P, double *maxgrad, hipStream_t st) {
  hipLaunchKernelGGL(k_mark, dim3(c.nblk), dim3(256), 0, st, c, P, maxgrad);
}
void fused_ghost_prims(const DevCfg &c, const double *U, double *G, hipStream_t st) {
  hipLaunchKernelGGL(k_ghost_prims_fused, dim3(6, c.nblk), dim3(256), 0, st, c, U, G);
}
void fused_mark(const DevCfg &c, const double *U, const double *G, double *maxgrad, hipStream_t st) {
  hipLaunchKernelGGL(k_mark_fused, dim3(c.nblk), dim3(1024), 0, st, c, U, G, maxgrad);
}
size_t plain_flux_elems(int nblk, int nv) { return (size_t)nblk * 3 * nv * NIF; }

}  // namespace wx
