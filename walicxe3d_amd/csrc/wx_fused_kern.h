// wx_fused_kern.h - the fused stage kernel and its device helpers, shared by
// the two translation units that instantiate it: wx_fused.hip (hydro kernels
// and the FP64 variant, built without FMA contraction so that the FP64
// variant keeps the reference's rounding) and wx_fused_mhd.hip (the FP32 MHD
// kernels, built with contraction).  Each unit links its own device code, so
// the two builds of these inline helpers never mix.
#pragma once

#include <hip/hip_runtime.h>

#include "wx_fused.h"

namespace wx {

__device__ __forceinline__ int lin(int i, int j, int k) {
  WX_EMU_CHECK(i >= 1 && i <= NC && j >= 1 && j <= NC && k >= 1 && k <= NC);
  return ((k - 1) * NC + (j - 1)) * NC + (i - 1);
}

struct FusedArgs {
  DevCfg c;
  const double *U;   // stage 1 input (FP64) / stage 2 base
  double *Uout;      // stage 2 output (in place)
  const void *UPin;  // stage 2 input (storage type ST)
  void *UPout;       // stage 1 output
  const int *order;
  StepState *step;
  int nblk;
  int nslab;         // z slabs per block, one workgroup each (FP32 MHD kernels without viscosity; else 1)
};

template <typename RF> __device__ __forceinline__ const Phys<RF> &phys(const DevCfg &c);
template <> __device__ __forceinline__ const Phys<float> &phys<float>(const DevCfg &c) { return c.pf; }
template <> __device__ __forceinline__ const Phys<double> &phys<double>(const DevCfg &c) { return c.pd; }

// Riemann solver chosen at compile time (one solver's code per kernel)
#ifdef WX_NOINLINE_RIEMANN
#define WX_RIEMANN_INLINE __noinline__
#else
#define WX_RIEMANN_INLINE __forceinline__
#endif
// (S1 = first stage: the FP32 HLLD there takes its supersonic cases on
// one code path)
template <int SOLVER, typename RF, bool S1 = false>
__device__ WX_RIEMANN_INLINE void riemann_t(const Phys<RF> &P, const RF *pl, const RF *pr, RF *ff) {
  if constexpr (SOLVER == SOLVER_HLLC) primfhllc(P, pl, pr, ff);
  else if constexpr (SOLVER == SOLVER_HLLD) {
#ifdef WX_HLLD_BRANCHES
    primfhlld(P, pl, pr, ff);
#else
    primfhlld_sel<RF, S1 && sizeof(RF) == 4>(P, pl, pr, ff);
#endif
  }
  else if constexpr (SOLVER == SOLVER_HLLE) primfhlle(P, pl, pr, ff);
  else primfhll(P, pl, pr, ff);
}

// a value that is the same in every lane, read from the first active lane
// into scalar registers (the value itself off the device)
__device__ __forceinline__ int uni(int x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_readfirstlane(x);
#else
  return x;
#endif
}
__device__ __forceinline__ float uni(float x) { return __int_as_float(uni(__float_as_int(x))); }
__device__ __forceinline__ double uni(double x) {
  const long long v = __double_as_longlong(x);
  const unsigned lo = (unsigned)uni((int)(unsigned)v);
  const unsigned hi = (unsigned)uni((int)(unsigned)(v >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// limiter of one variable: fixed at compile time (minmod in FP32 as a
// median) or read at run time
constexpr int LIM_RUNTIME = -100;
template <int LIM, typename RF>
__device__ __forceinline__ void limit_t(const Phys<RF> &P, RF pll, RF &pl, RF &pr, RF prr) {
  if constexpr (LIM == 1 && sizeof(RF) == 4) limit1_minmod32(pll, pl, pr, prr);
  else limit1(pll, pl, pr, prr, LIM == LIM_RUNTIME ? P.limiter : LIM);
}

// exchange of the x components with those along d (0 x, 1 y, 2 z) of
// velocity and field, written with selects so that the arrays stay in
// registers when d is not a compile-time constant (swapxy/swapxz,
// source/hydro_core.f90:448-496)
template <int NV, typename RF> __device__ __forceinline__ void rot(int d, RF *v) {
  RF a = v[1], b = v[2], c = v[3];
  v[1] = d == 0 ? a : (d == 1 ? b : c);
  v[2] = d == 1 ? a : b;
  v[3] = d == 2 ? a : c;
  if (NV == 8) {
    RF e = v[5], f = v[6], g = v[7];
    v[5] = d == 0 ? e : (d == 1 ? f : g);
    v[6] = d == 1 ? e : f;
    v[7] = d == 2 ? e : g;
  }
}

// conserved state of cell (i,j,k) of block b: interior or face ghost,
// then the wind sources; T is the storage type of the input array.
// WIN = 0: interior cells keep the state before the wind is re-imposed (the
// reference's stage 1 reuses the interior primitives of updatePrims, computed
// before boundary() re-imposes the sources; source/godunov.f90:77-83).
// DEL: W holds the half-step state as FP32 increments over U^n (a.U), and
// the value of an element is U^n + increment, formed in FP64.
// WINDS = false leaves the wind code out (kernels for meshes without sources).
template <typename T, int NV, bool WIN, bool DEL = false, bool WINDS = true>
__device__ __forceinline__ void load_cell(const FusedArgs &a, const T *W, int b, int i, int j, int k, double *u) {
  bool in = (unsigned)(i - 1) < NC && (unsigned)(j - 1) < NC && (unsigned)(k - 1) < NC;
  if (in) {
    const T *Wb = W + (size_t)b * NV * NINT + lin(i, j, k);
    if constexpr (DEL) {
      const double *Ub = a.U + (size_t)b * NV * NINT + lin(i, j, k);
#pragma unroll
      for (int v = 0; v < NV; v++) u[v] = Ub[v * NINT] + (double)Wb[v * NINT];
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++) u[v] = (double)Wb[v * NINT];
    }
  } else {
    int f, g, a1, a2;
    if (i < 1) { f = LEFT; g = 1 - i; a1 = j; a2 = k; }
    else if (i > NC) { f = RIGHT; g = i - NC; a1 = j; a2 = k; }
    else if (j < 1) { f = FRONT; g = 1 - j; a1 = i; a2 = k; }
    else if (j > NC) { f = BACK; g = j - NC; a1 = i; a2 = k; }
    else if (k < 1) { f = BOTTOM; g = 1 - k; a1 = i; a2 = j; }
    else { f = TOP; g = k - NC; a1 = i; a2 = j; }
    GhostSrc s;
    ghost_source(a.c.info, a.c.bc, b, f, g, a1, a2, s);
    const int ci = s.c[0], cj = s.c[1], ck = s.c[2];
    WX_EMU_CHECK(s.blk >= 0 && s.blk < a.nblk);
    WX_EMU_CHECK(s.kind != 1 || (ci < NC && cj < NC && ck < NC));   // the 2x2x2 group inside the block
    const T *S = W + (size_t)s.blk * NV * NINT;
    if constexpr (DEL) {
      const double *SU = a.U + (size_t)s.blk * NV * NINT;
      const int q0 = lin(ci, cj, ck);
      // value of element q of the source block
      auto el = [&](int q) -> double { return SU[q] + (double)S[q]; };
      if (s.kind == 1) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const int q = v * NINT + q0;
          // 2x2x2 average in the reference's summation order (x, then y, then z fastest)
          double acc = 0.0;
          acc = acc + el(q) / 8.0;
          acc = acc + el(q + NC * NC) / 8.0;
          acc = acc + el(q + NC) / 8.0;
          acc = acc + el(q + NC + NC * NC) / 8.0;
          acc = acc + el(q + 1) / 8.0;
          acc = acc + el(q + 1 + NC * NC) / 8.0;
          acc = acc + el(q + 1 + NC) / 8.0;
          acc = acc + el(q + 1 + NC + NC * NC) / 8.0;
          u[v] = acc;
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          double val = el(v * NINT + q0);
          u[v] = (s.kind == 2 && v == s.flip) ? -val : val;
        }
      }
    } else if (s.kind == 1) {
      const int q0 = lin(ci, cj, ck);
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const T *Sv = S + v * NINT + q0;
        // 2x2x2 average in the reference's summation order (x, then y, then z fastest)
        double acc = 0.0;
        acc = acc + (double)Sv[0] / 8.0;
        acc = acc + (double)Sv[NC * NC] / 8.0;
        acc = acc + (double)Sv[NC] / 8.0;
        acc = acc + (double)Sv[NC + NC * NC] / 8.0;
        acc = acc + (double)Sv[1] / 8.0;
        acc = acc + (double)Sv[1 + NC * NC] / 8.0;
        acc = acc + (double)Sv[1 + NC] / 8.0;
        acc = acc + (double)Sv[1 + NC + NC * NC] / 8.0;
        u[v] = acc;
      }
    } else {
      const int q0 = lin(ci, cj, ck);
#pragma unroll
      for (int v = 0; v < NV; v++) {
        double val = (double)S[v * NINT + q0];
        u[v] = (s.kind == 2 && v == s.flip) ? -val : val;
      }
    }
  }
  if (WINDS && a.c.nwind > 0 && (WIN || !in)) {
    double w[MAXV];
    if (wind_state(a.c, b, i, j, k, w)) {
#pragma unroll
      for (int v = 0; v < NV; v++) u[v] = w[v];
    }
  }
}

// MHD primitives for the FP32 kernels (flow2prim, source/hydro_core.f90:
// 323-378, with its floors): one reciprocal of the density for the
// velocities; the pressure from the FP64 state in FP64 when fp64 (the
// kinetic and magnetic energies cancel against E there), else in FP32.
__device__ __forceinline__ void prims_mhd32(const DevCfg &c, const double *u, float *p, bool fp64) {
  const float rho = (float)u[0];
  const float ir = rcp32(rho);
  float pr;
  if (fp64) {
    double ke = 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0];
    double me = 0.5 * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]);
    pr = (float)((u[4] - ke - me) / c.pd.cv);
  } else {
    const float m1 = (float)u[1], m2 = (float)u[2], m3 = (float)u[3];
    const float b1 = (float)u[5], b2 = (float)u[6], b3 = (float)u[7];
    pr = ((float)u[4] - 0.5f * (m1 * m1 + m2 * m2 + m3 * m3) * ir - 0.5f * (b1 * b1 + b2 * b2 + b3 * b3)) *
         rcp32(c.pf.cv);
  }
  p[0] = rho < 1.0e-40f ? 1.0e-40f : rho;
  p[1] = (float)u[1] * ir;
  p[2] = (float)u[2] * ir;
  p[3] = (float)u[3] * ir;
  p[4] = pr < 1.0e-30f ? 1.0e-30f : pr;
  p[5] = (float)u[5];
  p[6] = (float)u[6];
  p[7] = (float)u[7];
}

// primitives in the flux precision RF; fp64 = recover pressure in FP64
template <int NV, typename RF>
__device__ __forceinline__ void prims_r(const DevCfg &c, const double *u, RF *p, bool fp64) {
  if constexpr (NV == 8 && sizeof(RF) == 4) {
    if (c.pf.mhd) {
      prims_mhd32(c, u, p, fp64);
      return;
    }
  }
  if (fp64) {
    double pd[MAXV];
    flow2prim(c.pd, u, pd);
#pragma unroll
    for (int v = 0; v < NV; v++) p[v] = (RF)pd[v];
  } else {
    RF ur[MAXV];
#pragma unroll
    for (int v = 0; v < NV; v++) ur[v] = (RF)u[v];
    flow2prim(phys<RF>(c), ur, p);
  }
}

// CFL candidate of one cell in FP64 (getTimestep, source/hydro_solver.f90:150-182)
template <int NV>
__device__ __forceinline__ double dt_cell(const DevCfg &c, const double *p, double dx, double dy, double dz) {
  double r;
  if (NV == 8 && c.pd.mhd) {
    double cfx = cfast1(c.pd, p[4], p[0], p[5], p[6], p[7], p[5]);
    double cfy = cfast1(c.pd, p[4], p[0], p[5], p[6], p[7], p[6]);
    double cfz = cfast1(c.pd, p[4], p[0], p[5], p[6], p[7], p[7]);
    r = fmin(fmin(dx / (fabs(p[1]) + cfx), dy / (fabs(p[2]) + cfy)), dz / (fabs(p[3]) + cfz));
  } else {
    double cs = sound(c.pd, p);
    r = fmin(fmin(dx / (fabs(p[1]) + cs), dy / (fabs(p[2]) + cs)), dz / (fabs(p[3]) + cs));
  }
  return r;
}

// CFL candidate of one cell from FP32 primitives
template <int NV>
__device__ __forceinline__ float dt_cell_f(const DevCfg &c, const float *p, float dx, float dy, float dz) {
  float r;
  if (NV == 8 && c.pf.mhd) {
    float cfx = cfast1(c.pf, p[4], p[0], p[5], p[6], p[7], p[5]);
    float cfy = cfast1(c.pf, p[4], p[0], p[5], p[6], p[7], p[6]);
    float cfz = cfast1(c.pf, p[4], p[0], p[5], p[6], p[7], p[7]);
    r = fminf(fminf(rdiv(dx, fabsf(p[1]) + cfx), rdiv(dy, fabsf(p[2]) + cfy)), rdiv(dz, fabsf(p[3]) + cfz));
  } else {
    float cs = sound(c.pf, p);
    r = fminf(fminf(dx / (fabsf(p[1]) + cs), dy / (fabsf(p[2]) + cs)), dz / (fabsf(p[3]) + cs));
  }
  return r;
}

// stage-2 epilogue of one cell: write U^{n+1}, primitives, cooling, dt
// candidate.  With cooling (or in the FP64 variant) the primitives and the
// CFL candidate are FP64; otherwise the candidate comes from FP32
// primitives (the min itself stays exact).
template <int NV, typename RF>
__device__ __forceinline__ void finish_cell(const FusedArgs &a, int b, int i, int j, int k, double dt, double *u,
                                            double dx, double dy, double dz, double &dtmin) {
  const DevCfg &c = a.c;
  double *Uo = a.Uout + (size_t)b * NV * NINT + lin(i, j, k);
  if (sizeof(RF) == 4 && !c.cooling) {
#pragma unroll
    for (int v = 0; v < NV; v++) Uo[v * NINT] = u[v];
    float uf[MAXV], pf[MAXV];
    if constexpr (NV == 8) {
      prims_r<NV, float>(c, u, pf, false);
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++) uf[v] = (float)u[v];
      flow2prim(c.pf, uf, pf);
    }
    dtmin = fmin(dtmin, (double)dt_cell_f<NV>(c, pf, (float)dx, (float)dy, (float)dz));
    return;
  }
  double pd[MAXV];
  flow2prim(c.pd, u, pd);
  if (c.cooling) {
    double e;
    if (cool_cell(c, dt, pd, &e)) u[4] = e;
  }
#pragma unroll
  for (int v = 0; v < NV; v++) Uo[v * NINT] = u[v];
  dtmin = fmin(dtmin, dt_cell<NV>(c, pd, dx, dy, dz));
}

// ---------------------------------------------------------------------------
// Stage kernel.  S2 = second stage; VISC = artificial viscosity active;
// RF = flux/reconstruction arithmetic; ST = storage type of UP; SOLVER =
// Riemann solver; LIM = limiter fixed at compile time, or LIM_RUNTIME.
// Launch bound: 256 threads, at least three waves per SIMD (four for the MHD
// first stage).
// ---------------------------------------------------------------------------
#ifndef WX_STAGE_MIN_WAVES
#define WX_STAGE_MIN_WAVES 3
#endif
// the MHD first stage (no limiter, fewer live values) runs best at four
#ifndef WX_MHD_STAGE1_MIN_WAVES
#define WX_MHD_STAGE1_MIN_WAVES 4
#endif
template <int NV, bool S2, bool VISC, typename RF, typename ST, int SOLVER, int LIM>
__global__ void __launch_bounds__(256, (NV == 8 && !S2) ? WX_MHD_STAGE1_MIN_WAVES : WX_STAGE_MIN_WAVES)
    k_stage(FusedArgs a) {
  constexpr int ORD = S2 ? 2 : 1;
  constexpr int H = ORD;
  constexpr int W = NC + 2 * H;
  // LDS row pitch: odd (bank spread of y-direction reads); in the FP32 MHD
  // second stage without viscosity 49, which is 17 mod 32, so a half-wave's
  // x-face reads (rows of 17 interfaces) fall in distinct banks
  constexpr int WP = (NV == 8 && S2 && !VISC && sizeof(RF) == 4) ? 49 : (W | 1);
  constexpr int NW = S2 ? 4 : 3;               // planes in the column window
  constexpr bool SV = S2 && VISC;
  constexpr int NXF = (NC + 1) * NC;           // x (or y) interfaces per plane
  __shared__ RF sp[NV][W * WP];                 // primitives of the plane (cross region)
  __shared__ RF sfx[NV][NC * (NC + 1)];        // x fluxes [v][(j-1)*17 + s]
  __shared__ RF sfy[NV][(NC + 1) * NC];        // y fluxes, see YT
  // y-flux layout: YT (MHD kernels) stores [v][(i-1)*17 + s] like sfx, so the
  // stores of consecutive lanes (consecutive s) and the update's reads (lanes
  // along i, row pitch 17) hit distinct LDS banks; otherwise [v][s*16 + (i-1)]
  constexpr bool YT = NV == 8;
  __shared__ RF sun[SV ? 2 : 1][SV ? NV : 1][NC + 2][NC + 2];  // UP_new planes + half-step ghost ring
  __shared__ double sred[4];

  const DevCfg &c = a.c;
  const Phys<RF> &P = phys<RF>(c);
  const ST *UPin = static_cast<const ST *>(a.UPin);
  ST *UPout = static_cast<ST *>(a.UPout);
  const int tid = threadIdx.x;
  const int ti = tid & 15, tj = tid >> 4;
  const int i = ti + 1, j = tj + 1;
  const double dt = NV == 8 ? uni(a.step->dt) : a.step->dt;
  const double dtp = S2 ? dt : dt / 2.0;
  double dtmin = 1.0e30;

  for (int bi = blockIdx.x; bi < a.nblk; bi += gridDim.x) {
    // the block and its level are the same in every lane (and every lane is
    // active here, at the top of the block loop): in the MHD kernels they are
    // read into scalar registers, so that block addresses and per-level
    // values stay scalar
    const int b = NV == 8 ? uni(a.order[bi]) : a.order[bi];
    WX_EMU_CHECK(b >= 0 && b < a.nblk);
    const int l = NV == 8 ? uni(c.info[b].level - 1) : c.info[b].level - 1;
    const double dxl = c.dx[0][l], dyl = c.dx[1][l], dzl = c.dx[2][l];
    double dtdx = dtp / dxl, dtdy = dtp / dyl, dtdz = dtp / dzl;
    RF idx2 = (RF)(1.0 / (2.0 * dxl)), idy2 = (RF)(1.0 / (2.0 * dyl)), idz2 = (RF)(1.0 / (2.0 * dzl));
    if constexpr (NV == 8) {
      dtdx = uni(dtdx); dtdy = uni(dtdy); dtdz = uni(dtdz);
      idx2 = uni(idx2); idy2 = uni(idy2); idz2 = uni(idz2);
    }

    // column window of primitives; at iteration k (after the shift)
    // S2: win = planes {k-1, k, k+1, k+2};  S1: win = {k-1, k, k+1}
    RF win[NW][NV];
    double ucol[MAXV];
    double ughost_lo[MAXV], ughost_hi[MAXV];   // conserved half-step ghosts at planes 0 and 17
    RF fzp[NV];                                // z flux (k-1 | k)
    double unew_prev[MAXV], unew_prev2[MAXV];  // UP_new at planes k-1, k-2 (viscosity)
    // preload so that the shift at k = 0 produces the window of plane 0
#pragma unroll
    for (int w = 1; w < NW; w++) {
      const int kk = w - 2;                    // S2: -1, 0, 1 ; S1: -1, 0
      if (!S2 && kk < 0) continue;
      if constexpr (S2) load_cell<ST, NV, true>(a, UPin, b, i, j, kk, ucol);
      else load_cell<double, NV, false>(a, a.U, b, i, j, kk, ucol);
      if (SV && kk == 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) ughost_lo[v] = ucol[v];
      }
      prims_r<NV, RF>(c, ucol, win[w], !S2);
    }

    for (int k = 0; k <= NC; k++) {
      // shift the window and load the new top plane of the column
#pragma unroll
      for (int w = 0; w < NW - 1; w++)
#pragma unroll
        for (int v = 0; v < NV; v++) win[w][v] = win[w + 1][v];
      {
        const int kn = S2 ? k + 2 : k + 1;
        if constexpr (S2) load_cell<ST, NV, true>(a, UPin, b, i, j, kn, ucol);
        else load_cell<double, NV, false>(a, a.U, b, i, j, kn, ucol);
        if (SV && kn == NC + 1) {
#pragma unroll
          for (int v = 0; v < NV; v++) ughost_hi[v] = ucol[v];
        }
        prims_r<NV, RF>(c, ucol, win[NW - 1], !S2);
      }
      // z flux (k | k+1): one code path for every plane, the bottom face included
      RF fzc[NV];
      {
        RF pll[NV], pl[NV], pr[NV], prr[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          pl[v] = win[1][v];
          pr[v] = win[2][v];
          if constexpr (S2) { pll[v] = win[0][v]; prr[v] = win[3][v]; }
        }
        rot<NV>(2, pl);
        rot<NV>(2, pr);
        if constexpr (S2) {
          rot<NV>(2, pll);
          rot<NV>(2, prr);
#pragma unroll
          for (int v = 0; v < NV; v++) limit_t<LIM>(P, pll[v], pl[v], pr[v], prr[v]);
        }
        riemann_t<SOLVER, RF, !S2>(P, pl, pr, fzc);
        rot<NV>(2, fzc);
      }
      if (k == 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) fzp[v] = fzc[v];
        continue;
      }
      __syncthreads();   // previous plane's LDS readers are done
      // own primitives of plane k into LDS; halo threads add the face halo
#pragma unroll
      for (int v = 0; v < NV; v++) sp[v][(j - 1 + H) * WP + (i - 1 + H)] = win[1][v];
      if (tid < 4 * H * NC) {
        const int part = tid / (2 * H * NC);
        const int r = tid % (2 * H * NC);
        const int side = r / (H * NC);
        const int g = (r % (H * NC)) / NC + 1;
        const int aa = r % NC + 1;
        int hi, hj;
        if (part == 0) { hi = side ? NC + g : 1 - g; hj = aa; }
        else { hj = side ? NC + g : 1 - g; hi = aa; }
        double uh[MAXV];
        RF ph[MAXV];
        if constexpr (S2) load_cell<ST, NV, true>(a, UPin, b, hi, hj, k, uh);
        else load_cell<double, NV, false>(a, a.U, b, hi, hj, k, uh);
        prims_r<NV, RF>(c, uh, ph, !S2);
#pragma unroll
        for (int v = 0; v < NV; v++) sp[v][(hj - 1 + H) * WP + (hi - 1 + H)] = ph[v];
        if (SV && g == 1) {
#pragma unroll
          for (int v = 0; v < NV; v++) sun[k & 1][SV ? v : 0][hj][hi] = (RF)uh[v];
        }
      }
      __syncthreads();
      // x and y fluxes of plane k: 272 + 272 interfaces over 256 threads
      for (int t = tid; t < 2 * NXF; t += 256) {
        const int d = t >= NXF;               // 0: x, 1: y
        const int r = d ? t - NXF : t;
        const int s = r % (NC + 1);           // interface (s | s+1) along d
        const int aa = r / (NC + 1);          // row (0-based) across d
        // LDS position of cell s-1 along d, and the stride along d
        const int base = d ? (s - 2 + H) * WP + (aa + H) : (aa + H) * WP + (s - 2 + H);
        const int st = d ? WP : 1;
        RF pll[NV], pl[NV], pr[NV], prr[NV], ff[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          pl[v] = sp[v][base + st];
          pr[v] = sp[v][base + 2 * st];
          if constexpr (S2) { pll[v] = sp[v][base]; prr[v] = sp[v][base + 3 * st]; }
        }
        rot<NV>(d, pl);
        rot<NV>(d, pr);
        if constexpr (S2) {
          rot<NV>(d, pll);
          rot<NV>(d, prr);
#pragma unroll
          for (int v = 0; v < NV; v++) limit_t<LIM>(P, pll[v], pl[v], pr[v], prr[v]);
        }
        riemann_t<SOLVER, RF, !S2>(P, pl, pr, ff);
        rot<NV>(d, ff);
        if (d == 0) {
#pragma unroll
          for (int v = 0; v < NV; v++) sfx[v][aa * (NC + 1) + s] = ff[v];
        } else {
#pragma unroll
          for (int v = 0; v < NV; v++) sfy[v][YT ? aa * (NC + 1) + s : s * NC + aa] = ff[v];
        }
      }
      __syncthreads();
      // conservative update of cell (i, j, k), accumulated in FP64
      double unew[MAXV];
      {
        const double *Ub = a.U + (size_t)b * NV * NINT + lin(i, j, k);
        double ub[MAXV];
#pragma unroll
        for (int v = 0; v < NV; v++) ub[v] = Ub[v * NINT];
        if (c.nwind > 0) {
          double w[MAXV];
          if (wind_state(c, b, i, j, k, w))
#pragma unroll
            for (int v = 0; v < NV; v++) ub[v] = w[v];
        }
        // flux differences formed in FP64 (exact for FP32 operands), so the
        // sum over cells telescopes
#pragma unroll
        for (int v = 0; v < NV; v++)
          unew[v] = ub[v] + dtdx * ((double)sfx[v][tj * (NC + 1) + ti] - (double)sfx[v][tj * (NC + 1) + ti + 1]) +
                    dtdy * ((double)sfy[v][YT ? ti * (NC + 1) + tj : tj * NC + ti] -
                            (double)sfy[v][YT ? ti * (NC + 1) + tj + 1 : (tj + 1) * NC + ti]) +
                    dtdz * ((double)fzp[v] - (double)fzc[v]);
        if (NV == 8 && c.eight_wave) {
          // Powell source (source/sources.f90:38-95) from the stage primitives
          const int q = (j - 1 + H) * WP + (i - 1 + H);
          RF divB = (sp[5][q + 1] - sp[5][q - 1]) * idx2 + (sp[6][q + WP] - sp[6][q - WP]) * idy2 +
                    (win[2][7] - win[0][7]) * idz2;
          const RF *pp = win[1];
          unew[1] += dtp * (double)(-divB * pp[5]);
          unew[2] += dtp * (double)(-divB * pp[6]);
          unew[3] += dtp * (double)(-divB * pp[7]);
          unew[4] += dtp * (double)(-divB * (pp[1] * pp[5] + pp[2] * pp[6] + pp[3] * pp[7]));
          unew[5] += dtp * (double)(-divB * pp[1]);
          unew[6] += dtp * (double)(-divB * pp[2]);
          unew[7] += dtp * (double)(-divB * pp[3]);
        }
      }
#pragma unroll
      for (int v = 0; v < NV; v++) fzp[v] = fzc[v];

      if constexpr (!S2) {
        ST *UPo = UPout + (size_t)b * NV * NINT + lin(i, j, k);
#pragma unroll
        for (int v = 0; v < NV; v++) UPo[v * NINT] = (ST)unew[v];
        continue;
      }

      // ---- stage 2 epilogue: viscosity (one plane behind), prims, cooling, dt
      if constexpr (SV) {
#pragma unroll
        for (int v = 0; v < NV; v++) sun[k & 1][v][j][i] = (RF)unew[v];
        __syncthreads();
        if (k > 1) {
          const int pb = (k - 1) & 1;
          double ufin[MAXV];
#pragma unroll
          for (int v = 0; v < NV; v++) {
            double below = (k - 2 == 0) ? ughost_lo[v] : unew_prev2[v];
            double lap = (double)sun[pb][v][j][i + 1] + (double)sun[pb][v][j][i - 1] +
                         (double)sun[pb][v][j + 1][i] + (double)sun[pb][v][j - 1][i] + below + unew[v] -
                         6.0 * unew_prev[v];
            ufin[v] = unew_prev[v] + c.visc_eta * lap;
          }
          finish_cell<NV, RF>(a, b, i, j, k - 1, dt, ufin, dxl, dyl, dzl, dtmin);
        }
#pragma unroll
        for (int v = 0; v < NV; v++) { unew_prev2[v] = unew_prev[v]; unew_prev[v] = unew[v]; }
      } else {
        finish_cell<NV, RF>(a, b, i, j, k, dt, unew, dxl, dyl, dzl, dtmin);
      }
    }
    if constexpr (SV) {
      // plane 16: its upper neighbour is the half-step ghost at plane 17
      __syncthreads();
      const int pb = NC & 1;
      double ufin[MAXV];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        double lap = (double)sun[pb][v][j][i + 1] + (double)sun[pb][v][j][i - 1] + (double)sun[pb][v][j + 1][i] +
                     (double)sun[pb][v][j - 1][i] + unew_prev2[v] + ughost_hi[v] - 6.0 * unew_prev[v];
        ufin[v] = unew_prev[v] + c.visc_eta * lap;
      }
      finish_cell<NV, RF>(a, b, i, j, NC, dt, ufin, dxl, dyl, dzl, dtmin);
    }
    __syncthreads();
  }
  if constexpr (S2) {
    // exact min: wavefront shuffles, then one 64-bit atomic per workgroup on
    // the bit pattern (positive doubles order like their bits)
    for (int off = 32; off > 0; off >>= 1) dtmin = fmin(dtmin, __shfl_xor(dtmin, off));
    if ((tid & 63) == 0) sred[tid >> 6] = dtmin;
    __syncthreads();
    if (tid == 0) {
      double m = fmin(fmin(sred[0], sred[1]), fmin(sred[2], sred[3]));
      atomicMin(&a.step->dtmin_bits, (unsigned long long)__double_as_longlong(m));
    }
  }
}

// second-stage kernels with the limiter fixed at compile time (minmod, in
// the FP32 MHD form) or read at run time
template <typename RF, typename ST, int NV, int SOLVER, bool VISC>
inline void launch_stage2(const FusedArgs &a, int grid, hipStream_t st) {
  constexpr int MM = 1;  // LIMITER_MINMOD (source/constants.f90)
  if constexpr (sizeof(RF) == 4 && NV == 8) {
    if (a.c.pf.limiter == MM) {
      hipLaunchKernelGGL((k_stage<NV, true, VISC, RF, ST, SOLVER, MM>), dim3(grid), dim3(256), 0, st, a);
      return;
    }
  }
  hipLaunchKernelGGL((k_stage<NV, true, VISC, RF, ST, SOLVER, LIM_RUNTIME>), dim3(grid), dim3(256), 0, st, a);
}

template <typename RF, typename ST, int NV, int SOLVER>
inline void launch_stage_s(const FusedArgs &a, int grid, bool visc, int stage, hipStream_t st) {
  if (stage == 1)
    hipLaunchKernelGGL((k_stage<NV, false, false, RF, ST, SOLVER, LIM_RUNTIME>), dim3(grid), dim3(256), 0, st, a);
  else if (visc)
    launch_stage2<RF, ST, NV, SOLVER, true>(a, grid, st);
  else
    launch_stage2<RF, ST, NV, SOLVER, false>(a, grid, st);
}


// FP32 MHD stage kernels (HLLD, HLLE), defined in wx_fused_mhd.hip: the
// delta-form kernels, or with plain32 the plain FP32 flux kernels above
void launch_stage_mhd32(const FusedArgs &a, int grid, int solver, bool visc, int stage, int plain32, hipStream_t st);

}  // namespace wx
