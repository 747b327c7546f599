// wx_kernels.h - host-side launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "wx_device.h"

namespace wx {

// plain FP64 form (wx_plain.hip)
void plain_boundary(const DevCfg &c, const PlaneWind &pw, double *A, int depth, hipStream_t st);
void plain_prims(const DevCfg &c, const double *A, double *P, int cells, unsigned long long *bad, hipStream_t st);
void plain_fluxes(const DevCfg &c, const double *P, double *F, int order, hipStream_t st);
void plain_upwind(const DevCfg &c, const UserSource &us, const double *U, const double *P, const double *F,
                  double *UP, double dtp, hipStream_t st);
void plain_viscous(const DevCfg &c, const double *UP, double *U, hipStream_t st);
void plain_timestep(const DevCfg &c, const double *P, unsigned long long *dtbits, hipStream_t st);
void plain_cooling(const DevCfg &c, double *U, double *P, double dt, double mu0, hipStream_t st);
// per-block largest relative pressure / density gradient (markByPhysical)
void plain_mark(const DevCfg &c, const double *P, double *maxgrad, hipStream_t st);
// the same on the fused state ([blk][nv][16^3]): recorded ghost layer
// G[blk][rho, p][6][256], and the per-block maximum (G may be null)
void fused_ghost_prims(const DevCfg &c, const double *U, double *G, hipStream_t st);
void fused_mark(const DevCfg &c, const double *U, const double *G, double *maxgrad, hipStream_t st);
size_t plain_flux_elems(int nblk, int nv);
// device physics routines over arrays of records (wx_eval_physics)
enum { EVAL_FLOW2PRIM = 0, EVAL_PRIM2FLOW, EVAL_PRIM2FLUXES, EVAL_LIMITER, EVAL_RIEMANN, EVAL_RIEMANN_SEL,
       EVAL_SPEEDS };
void plain_eval(const Phys<double> &P, int op, int arg, int n, const double *in, double *out, hipStream_t st);

}  // namespace wx
