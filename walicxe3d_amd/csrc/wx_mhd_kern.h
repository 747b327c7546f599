// wx_mhd_kern.h - the stage kernel of the FP32 MHD form (HLLD, HLLE).
//
// Same block-per-workgroup structure as k_stage (wx_fused_kern.h): thread
// (i, j) owns a z-column and marches k; the column's cells sit in a
// register window, the plane k and its x/y face halo in LDS (the second
// stage, short of registers, keeps plane k only there).  The
// arithmetic is split so that FP32 rounding never touches an O(1) part of
// a face flux:
//   - primitives of every cell in FP64 (from the FP64 state, or in the
//     second stage from U^n plus the FP32 half-step increment);
//   - MUSCL slopes in FP32 from FP64 differences, face states in FP64;
//   - face flux = FP64 physical flux of the upwind state + the Riemann
//     correction in FP32 from jump-sized quantities (wx_delta.h), stored to
//     LDS in FP64;
//   - U + dt * (flux differences + Powell source) accumulated in FP64.
// Faces near the degenerate HLLD star state take the plain FP32 HLLD
// (wx_delta.h).
// Stage 1 writes the half-step state as FP32 increments over U^n; stage 2
// reads U^n of neighbour blocks for its halo and so writes U^{n+1} into the
// second state buffer (the two swap roles every step).
#pragma once

#include "wx_delta.h"
#include "wx_fused_kern.h"

namespace wx {

// FP64 primitives of one cell (flow2prim, source/hydro_core.f90:323-378,
// with its floors).  Every operation is written out (no contraction), so
// the two blocks that share a face derive the same values from a cell.
__device__ __forceinline__ void prims64(double icv, const double *u, double *w) {
#pragma clang fp contract(off)
  const double rho = u[0];
  // density reciprocal: hardware FP32 reciprocal and one Newton step
  double r = (double)rcp32((float)rho);
  r = __builtin_fma(r, __builtin_fma(-rho, r, 1.0), r);
  const double v1 = u[1] * r, v2 = u[2] * r, v3 = u[3] * r;
  const double ke = 0.5 * __builtin_fma(u[3], v3, __builtin_fma(u[2], v2, u[1] * v1));
  const double me = 0.5 * __builtin_fma(u[7], u[7], __builtin_fma(u[6], u[6], u[5] * u[5]));
  const double p = (u[4] - ke - me) * icv;
  w[0] = rho < 1.0e-40 ? 1.0e-40 : rho;
  w[1] = v1;
  w[2] = v2;
  w[3] = v3;
  w[4] = p < 1.0e-30 ? 1.0e-30 : p;
  w[5] = u[5];
  w[6] = u[6];
  w[7] = u[7];
}

// MUSCL face values of one variable (limiter, source/hydro_core.f90:613-697)
// from FP64 cell values: the slopes in FP32 (FP64 differences rounded), the
// face values cell +- slope / 2 in FP64
template <int LIM>
__device__ __forceinline__ void rec64(const Phys<float> &P, double xll, double xl, double xr, double xrr, double &l,
                                      double &r) {
  const float dl = (float)(xl - xll), dm = (float)(xr - xl), dr = (float)(xrr - xr);
  float al, ar;
  if constexpr (LIM == 1) {
    al = minmod32(dl, dm);
    ar = minmod32(dm, dr);
  } else {
    const int lim = LIM == LIM_RUNTIME ? P.limiter : LIM;
    al = average(dl, dm, lim);
    ar = average(dm, dr, lim);
  }
  l = xl + 0.5 * (double)al;
  r = xr - 0.5 * (double)ar;
}

// Exchange of the x components of velocity and field with those along D
// (swapxy / swapxz, source/hydro_core.f90:448-496)
template <int D> __device__ __forceinline__ void rot8(double *v) {
  if constexpr (D != 0) {
    double t = v[1];
    v[1] = v[1 + D];
    v[1 + D] = t;
    t = v[5];
    v[5] = v[5 + D];
    v[5 + D] = t;
  }
}

// rot8<1> where d holds (the y direction), per lane: selects, no branch
__device__ __forceinline__ void rot8_if(bool d, double *v) {
  const double v1 = v[1], v2 = v[2], v5 = v[5], v6 = v[6];
  v[1] = d ? v2 : v1;
  v[2] = d ? v1 : v2;
  v[5] = d ? v6 : v5;
  v[6] = d ? v5 : v6;
}

// Flux across one face along D from its two face states (unrotated) in the
// delta form of wx_delta.h; a face near the degenerate HLLD star state
// takes the one-code-path HLLD in FP32 (hlld_flux_ill)
template <int D, int SOLVER>
__device__ __forceinline__ void face_flux(const DevCfg &c, double *wl, double *wr, double *F) {
  rot8<D>(wl);
  rot8<D>(wr);
  if constexpr (SOLVER == SOLVER_HLLD) {
    if (!hlld_flux_delta(c.pd, c.pf, wl, wr, F)) hlld_flux_ill(c.pf, wl, wr, F);
  } else {
    hlle_flux_delta(c.pd, c.pf, wl, wr, F);
  }
  rot8<D>(F);
}

#ifndef WX_MHD_MIN_WAVES
#define WX_MHD_MIN_WAVES 2
#endif

template <bool S2, bool VISC, int SOLVER, int LIM, bool WIND>
__global__ void __launch_bounds__(256, WX_MHD_MIN_WAVES) k_mhd(FusedArgs a) {
  constexpr int NV = 8;
  constexpr int H = S2 ? 2 : 1;            // halo depth of the stencil
  constexpr int W = NC + 2 * H;
  constexpr int WPD = W | 1;               // LDS row pitch (FP64 elements)
  constexpr int NW = S2 ? 4 : 3;           // planes in the column window
  constexpr bool SV = S2 && VISC;
  constexpr int NXF = (NC + 1) * NC;       // x (or y) faces of a plane
  __shared__ double sw[NV][W * WPD];       // FP64 primitives of plane k (cross region)
  __shared__ double sfx[NV][NXF];          // x fluxes [v][row * 17 + s]
  __shared__ double sfy[NV][NXF];          // y fluxes [v][col * 17 + s]
  __shared__ float sun[SV ? 2 : 1][SV ? NV : 1][NC + 2][NC + 2];  // UP_new planes + half-step ghost ring
  __shared__ double sred[4];

  const DevCfg &c = a.c;
  const Phys<float> &P = c.pf;
  const float *UPin = static_cast<const float *>(a.UPin);
  float *UPout = static_cast<float *>(a.UPout);
  const int tid = threadIdx.x;
  const int ti = tid & 15, tj = tid >> 4;
  const int i = ti + 1, j = tj + 1;
  const double dt = uni(a.step->dt);
  const double dtp = S2 ? dt : dt / 2.0;
  const double icv = uni(1.0 / c.pd.cv);
  double dtmin = 1.0e30;

  // x/y face halo of a plane, H layers on each of the 4 sides: lanes
  // tid < NH load one cell each.  The first stage issues these loads at the
  // top of the plane (their latency overlaps the z face); the second, whose
  // registers are full there, after the first barrier.
  constexpr int NH = 4 * H * NC;
  constexpr bool EARLY_HALO = !S2;
  // second stage: plane k of the column window lives in this lane's slot of
  // sw (where the x/y faces read it) rather than in registers, from the
  // point where it is stored until it is read back as plane k-1
  constexpr bool LW = S2;
  const int own = (j - 1 + H) * WPD + (i - 1 + H);
  const bool hth = tid < NH;
  // halo cell (hi, hj) of lane tid, layer g
  auto halo_cell = [&](int &hi, int &hj, int &g) {
    const int part = tid / (2 * H * NC);
    const int r = tid % (2 * H * NC);
    const int side = r / (H * NC);
    g = (r % (H * NC)) / NC + 1;
    const int aa = r % NC + 1;
    if (part == 0) { hi = side ? NC + g : 1 - g; hj = aa; }
    else { hj = side ? NC + g : 1 - g; hi = aa; }
  };

  // Work items: block bi of the order list, z slab s of it (planes
  // s L + 1 .. (s + 1) L, L = NC / nslab; the viscous kernels run whole
  // blocks, nslab = 1).  A slab starts one plane early, at its bottom z face,
  // which the slab below computes too, with the same code and inputs.
  const int nslab = uni(a.nslab);
  WX_EMU_CHECK(nslab >= 1 && NC % nslab == 0 && (!SV || nslab == 1));
  const int L = NC / nslab;
  for (int item = blockIdx.x; item < a.nblk * nslab; item += gridDim.x) {
    const int bi = item / nslab;
    const int k0 = (item - bi * nslab) * L + 1, k1 = k0 + L - 1;
    // block, level and per-block constants are the same in every lane:
    // scalar registers
    const int b = uni(a.order[bi]);
    WX_EMU_CHECK(b >= 0 && b < a.nblk);
    const int l = uni(c.info[b].level - 1);
    const double dxl = c.dx[0][l], dyl = c.dx[1][l], dzl = c.dx[2][l];
    const double dtdx = uni(dtp / dxl), dtdy = uni(dtp / dyl), dtdz = uni(dtp / dzl);
    const double idx2 = uni(1.0 / (2.0 * dxl)), idy2 = uni(1.0 / (2.0 * dyl)), idz2 = uni(1.0 / (2.0 * dzl));

    // column window of FP64 primitives: at iteration k (after the shift)
    // S2: planes {k-1, k, k+1, k+2};  S1: planes {k-1, k, k+1}
    double wq[NW][NV];
    double ucol[NV];
    double ughost_lo[SV ? NV : 1], ughost_hi[SV ? NV : 1];   // half-step ghosts at planes 0 and 17
    double fzp[NV];                                          // z flux (k-1 | k)
    double unew_prev[SV ? NV : 1], unew_prev2[SV ? NV : 1];  // UP_new at planes k-1, k-2 (viscosity)
#pragma unroll
    for (int w = 0; w < NW; w++)
#pragma unroll
      for (int v = 0; v < NV; v++) wq[w][v] = 0.0;

    // Iteration k loads plane k + NW - 2 into the window.  The first
    // iterations only fill it (S2: planes k0-2 .. k0, S1: plane k0-1), so
    // that one call site of load_cell, with its ghost paths, serves every
    // plane: the kernel's code stays small.
    for (int k = k0 - (S2 ? 4 : 2); k <= k1; k++) {
      if constexpr (LW) {
        // plane k-1 from this lane's slot of the plane store, where the
        // previous iteration left it
#pragma unroll
        for (int v = 0; v < NV; v++) wq[0][v] = sw[v][own];
#pragma unroll
        for (int w = 1; w < NW - 1; w++)
#pragma unroll
          for (int v = 0; v < NV; v++) wq[w][v] = wq[w + 1][v];
      } else {
#pragma unroll
        for (int w = 0; w < NW - 1; w++)
#pragma unroll
          for (int v = 0; v < NV; v++) wq[w][v] = wq[w + 1][v];
      }
      {
        const int kn = S2 ? k + 2 : k + 1;
        if constexpr (S2) load_cell<float, NV, true, true, WIND>(a, UPin, b, i, j, kn, ucol);
        else load_cell<double, NV, false, false, WIND>(a, a.U, b, i, j, kn, ucol);
        if constexpr (SV) {
          if (kn == 0)
#pragma unroll
            for (int v = 0; v < NV; v++) ughost_lo[v] = ucol[v];
          if (kn == NC + 1)
#pragma unroll
            for (int v = 0; v < NV; v++) ughost_hi[v] = ucol[v];
        }
        prims64(icv, ucol, wq[NW - 1]);
      }
      if constexpr (LW) {
        // while the window fills, the slot takes plane k here (later after
        // the first barrier, with the rest of the plane)
        if (k < k0)
#pragma unroll
          for (int v = 0; v < NV; v++) sw[v][own] = wq[1][v];
      }
      double wh[NV];                       // halo cell primitives
      float uhf[SV ? NV : 1];              // halo cell state (viscosity ring)
      auto halo_load = [&]() {
        int hi, hj, hg;
        halo_cell(hi, hj, hg);
        double uh[NV];
        if constexpr (S2) load_cell<float, NV, true, true, WIND>(a, UPin, b, hi, hj, k, uh);
        else load_cell<double, NV, false, false, WIND>(a, a.U, b, hi, hj, k, uh);
        prims64(icv, uh, wh);
        if constexpr (SV) {
#pragma unroll
          for (int v = 0; v < NV; v++) uhf[v] = (float)uh[v];
        }
      };
      if (EARLY_HALO && hth && k >= k0) halo_load();
      if (k < k0 - 1) continue;
      // z flux (k | k+1): one code path for every plane, the bottom face included
      double fzc[NV];
      {
        double wl[NV], wr[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          if constexpr (S2) rec64<LIM>(P, wq[0][v], wq[1][v], wq[2][v], wq[3][v], wl[v], wr[v]);
          else { wl[v] = wq[1][v]; wr[v] = wq[2][v]; }
        }
        face_flux<2, SOLVER>(c, wl, wr, fzc);
      }
      if (k == k0 - 1) {
#pragma unroll
        for (int v = 0; v < NV; v++) fzp[v] = fzc[v];
        continue;
      }
      __syncthreads();   // the previous plane's LDS readers are done
      // own primitives of plane k into LDS; halo threads add the face halo
#pragma unroll
      for (int v = 0; v < NV; v++) sw[v][own] = wq[1][v];
      if (hth) {
        if (!EARLY_HALO) halo_load();
        int hi, hj, hg;
        halo_cell(hi, hj, hg);
#pragma unroll
        for (int v = 0; v < NV; v++) sw[v][(hj - 1 + H) * WPD + (hi - 1 + H)] = wh[v];
        if constexpr (SV) {
          if (hg == 1)
#pragma unroll
            for (int v = 0; v < NV; v++) sun[k & 1][v][hj][hi] = uhf[v];
        }
      }
      __syncthreads();
      // x faces s = 0..15 of plane k, one per lane (row tj, face ti)
      {
        const int base = (tj + H) * WPD + (ti - 1 + H);       // cell s = ti
        double wl[NV], wr[NV], ff[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double *x = &sw[v][base];
          if constexpr (S2) rec64<LIM>(P, x[-1], x[0], x[1], x[2], wl[v], wr[v]);
          else { wl[v] = x[0]; wr[v] = x[1]; }
        }
        face_flux<0, SOLVER>(c, wl, wr, ff);
#pragma unroll
        for (int v = 0; v < NV; v++) sfx[v][tj * (NC + 1) + ti] = ff[v];
      }
      // y faces s = 0..15, one per lane
      {
        const int t = (tid + 192) & 255;
        const int sy = t & 15, col = t >> 4;
        const int base = (sy - 1 + H) * WPD + (col + H);      // cell s
        double wl[NV], wr[NV], ff[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double *x = &sw[v][base];
          if constexpr (S2) rec64<LIM>(P, x[-WPD], x[0], x[WPD], x[2 * WPD], wl[v], wr[v]);
          else { wl[v] = x[0]; wr[v] = x[WPD]; }
        }
        face_flux<1, SOLVER>(c, wl, wr, ff);
#pragma unroll
        for (int v = 0; v < NV; v++) sfy[v][col * (NC + 1) + sy] = ff[v];
      }
      // the 16 x and the 16 y faces s = 16 together on the last 32 lanes (x
      // on the first 16, y on the others): one solve per lane, the states
      // rotated by lane (selects) into the x-aligned order of face_flux<0>
      if (tid >= 256 - 2 * NC) {
        const int e = tid - (256 - 2 * NC);
        const bool d = e >= NC;                                // y face
        const int aa = d ? e - NC : e;                         // row or column
        const int st = d ? WPD : 1;
        const int base = d ? (NC - 1 + H) * WPD + (aa + H) : (aa + H) * WPD + (NC - 1 + H);
        double wl[NV], wr[NV], ff[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double *x = &sw[v][base];
          if constexpr (S2) rec64<LIM>(P, x[-st], x[0], x[st], x[2 * st], wl[v], wr[v]);
          else { wl[v] = x[0]; wr[v] = x[st]; }
        }
        rot8_if(d, wl);
        rot8_if(d, wr);
        face_flux<0, SOLVER>(c, wl, wr, ff);
        rot8_if(d, ff);
#pragma unroll
        for (int v = 0; v < NV; v++) (d ? sfy[v] : sfx[v])[aa * (NC + 1) + NC] = ff[v];
      }
      __syncthreads();
      // conservative update of cell (i, j, k) in FP64
      double unew[NV];
      {
        const double *Ub = a.U + (size_t)b * NV * NINT + lin(i, j, k);
        double ub[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) ub[v] = Ub[v * NINT];
        // base of the update: U^n, or the wind state in a cell that a source
        // covers (each stage's boundary() imposes the sources through
        // userBoundary: source/godunov.f90:79,141, source/boundary.f90:61)
        double u0[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) u0[v] = ub[v];
        if constexpr (WIND) {
          double w[MAXV];
          if (c.nwind > 0 && wind_state(c, b, i, j, k, w))
#pragma unroll
            for (int v = 0; v < NV; v++) u0[v] = w[v];
        }
#pragma unroll
        for (int v = 0; v < NV; v++)
          unew[v] = u0[v] + dtdx * (sfx[v][tj * (NC + 1) + ti] - sfx[v][tj * (NC + 1) + ti + 1]) +
                    dtdy * (sfy[v][ti * (NC + 1) + tj] - sfy[v][ti * (NC + 1) + tj + 1]) +
                    dtdz * (fzp[v] - fzc[v]);
        if (c.eight_wave) {
          // Powell source (source/sources.f90:38-95) from the stage primitives
          const int q = own;
          const double divB = (sw[5][q + 1] - sw[5][q - 1]) * idx2 + (sw[6][q + WPD] - sw[6][q - WPD]) * idy2 +
                              (wq[2][7] - wq[0][7]) * idz2;
          double ppl[NV];
          const double *pp = wq[1];
          if constexpr (LW) {
#pragma unroll
            for (int v = 0; v < NV; v++) ppl[v] = sw[v][q];
            pp = ppl;
          }
          unew[1] += dtp * (-divB * pp[5]);
          unew[2] += dtp * (-divB * pp[6]);
          unew[3] += dtp * (-divB * pp[7]);
          unew[4] += dtp * (-divB * (pp[1] * pp[5] + pp[2] * pp[6] + pp[3] * pp[7]));
          unew[5] += dtp * (-divB * pp[1]);
          unew[6] += dtp * (-divB * pp[2]);
          unew[7] += dtp * (-divB * pp[3]);
        }
        if constexpr (!S2) {
          // the half-step state as FP32 increments over U^n
          float *UPo = UPout + (size_t)b * NV * NINT + lin(i, j, k);
#pragma unroll
          for (int v = 0; v < NV; v++) UPo[v * NINT] = (float)(unew[v] - ub[v]);   // over the stored U^n
        }
      }
#pragma unroll
      for (int v = 0; v < NV; v++) fzp[v] = fzc[v];
      if constexpr (!S2) continue;

      // ---- stage 2 epilogue: viscosity (one plane behind), prims, dt
      if constexpr (SV) {
#pragma unroll
        for (int v = 0; v < NV; v++) sun[k & 1][v][j][i] = (float)unew[v];
        __syncthreads();
        if (k > 1) {
          const int pb = (k - 1) & 1;
          double ufin[NV];
#pragma unroll
          for (int v = 0; v < NV; v++) {
            double below = (k - 2 == 0) ? ughost_lo[v] : unew_prev2[v];
            double lap = (double)sun[pb][v][j][i + 1] + (double)sun[pb][v][j][i - 1] +
                         (double)sun[pb][v][j + 1][i] + (double)sun[pb][v][j - 1][i] + below + unew[v] -
                         6.0 * unew_prev[v];
            ufin[v] = unew_prev[v] + c.visc_eta * lap;
          }
          finish_cell<NV, float>(a, b, i, j, k - 1, dt, ufin, dxl, dyl, dzl, dtmin);
        }
#pragma unroll
        for (int v = 0; v < NV; v++) { unew_prev2[v] = unew_prev[v]; unew_prev[v] = unew[v]; }
      } else {
        finish_cell<NV, float>(a, b, i, j, k, dt, unew, dxl, dyl, dzl, dtmin);
      }
    }
    if constexpr (SV) {
      // plane 16: its upper neighbour is the half-step ghost at plane 17
      __syncthreads();
      const int pb = NC & 1;
      double ufin[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        double lap = (double)sun[pb][v][j][i + 1] + (double)sun[pb][v][j][i - 1] + (double)sun[pb][v][j + 1][i] +
                     (double)sun[pb][v][j - 1][i] + unew_prev2[v] + ughost_hi[v] - 6.0 * unew_prev[v];
        ufin[v] = unew_prev[v] + c.visc_eta * lap;
      }
      finish_cell<NV, float>(a, b, i, j, NC, dt, ufin, dxl, dyl, dzl, dtmin);
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

template <int SOLVER, bool WIND>
inline void launch_mhd_w(const FusedArgs &a, int grid, bool visc, int stage, hipStream_t st) {
  constexpr int MM = 1;  // LIMITER_MINMOD (source/constants.f90)
  const bool mm = a.c.pf.limiter == MM && !WIND;
  if (stage == 1)
    hipLaunchKernelGGL((k_mhd<false, false, SOLVER, LIM_RUNTIME, WIND>), dim3(grid), dim3(256), 0, st, a);
  else if (visc)
    hipLaunchKernelGGL((k_mhd<true, true, SOLVER, LIM_RUNTIME, WIND>), dim3(grid), dim3(256), 0, st, a);
  else if (mm)
    hipLaunchKernelGGL((k_mhd<true, false, SOLVER, MM, false>), dim3(grid), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((k_mhd<true, false, SOLVER, LIM_RUNTIME, WIND>), dim3(grid), dim3(256), 0, st, a);
}

// meshes with wind sources take the kernels that carry the wind code (every
// limiter at run time); the others leave it out
template <int SOLVER>
inline void launch_mhd(const FusedArgs &a, int grid, bool visc, int stage, hipStream_t st) {
  if (a.c.nwind > 0) launch_mhd_w<SOLVER, true>(a, grid, visc, stage, st);
  else launch_mhd_w<SOLVER, false>(a, grid, visc, stage, st);
}

}  // namespace wx
