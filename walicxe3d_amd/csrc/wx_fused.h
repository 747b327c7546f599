// wx_fused.h - the fused FP32 form: one kernel per stage over every block of
// every level, halos gathered from neighbour interiors on the fly, FP64
// conserved accumulation, device-side dt, HIP graph per mesh.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "wx_device.h"

namespace wx {

// Device-resident step scalars (the host never reads them between steps).
struct StepState {
  double dt;          // dt of the current step
  double time;        // simulation time at the start of the step
  double dtout;       // output interval (parameters dtout)
  double t_sc;        // time unit (parameters t_sc)
  unsigned long long dtmin_bits;  // running min of dx/(|v|+c) over cells
  int it;             // iteration counter
  int nextout;        // next output number
  int dump;           // the clamp fired this step
  int pad;
};

struct FusedState {
  double *U = nullptr;       // [blk][nv][4096] FP64 conserved state (the current one)
  // FP32 MHD form: UP holds FP32 increments over U^n, so the second stage
  // needs U^n of neighbour blocks after their own update: it writes U^{n+1}
  // into this second buffer, and the two swap roles every step
  double *U2 = nullptr;
  int two_u = 0;
  void *UP = nullptr;        // [blk][nv][4096] half-step state (FP32, or FP64 in the fp64 variant)
  int fp64 = 0;              // flux arithmetic and UP storage in FP64
  int *order = nullptr;      // block processing order (device)
  StepState *step = nullptr; // device
  int nblk = 0, nv = 0;
  int grid = 0;              // stage-kernel grid size (0 = one workgroup per block)
  int xcd_map = 1;           // XCD-aware block order
  int curve = 0;             // block order along the Morton (0) or Hilbert (1) curve
  int mhd_plain32 = 0;       // FP32 MHD: plain FP32 fluxes (1) instead of the delta form (0)
  int nslab = 1;             // FP32 MHD without viscosity: z slabs per block, one workgroup each
  int nroot[3] = {1, 1, 1};  // root blocks per axis (Hilbert curve order)
  hipGraph_t graph = nullptr;      // one step from the current U
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph2 = nullptr;     // with two U buffers: the step from the other one
  hipGraphExec_t exec2 = nullptr;
  const double *graph_u = nullptr, *graph2_u = nullptr;   // the U each graph reads
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double t_s1 = 0, t_s2 = 0, t_tot = 0;
  int timing_runs = 0;
  std::vector<int> host_order;
  std::vector<BlockInfo> host_info;
  int maxlev = 0;
};

int fused_alloc(FusedState &f, const DevCfg &c, const std::vector<BlockInfo> &info, hipStream_t st);
void fused_free(FusedState &f);
// State in from U ([blk][nv][16^3] FP64, device): copied, or with adopt the
// buffer itself becomes the state (the library frees it); dt re-initialised.
int fused_upload(FusedState &f, const DevCfg &c, double *U, bool adopt, hipStream_t st);
// PRIM of the state into P ([blk][nv][16^3] FP64, device)
int fused_prims(FusedState &f, const DevCfg &c, double *P, hipStream_t st);
int fused_timestep(FusedState &f, const DevCfg &c, hipStream_t st, double *dt_raw);
int fused_run(FusedState &f, const DevCfg &c, hipStream_t st, int nsteps, int *it, double *dt, double *time,
              int *nextout, double dtout, double t_sc);
int fused_time_stages(FusedState &f, const DevCfg &c, hipStream_t st, int reps);
int fused_set_tuning(FusedState &f, int kind, int value);
void fused_last_timing(const FusedState &f, double *s1, double *s2, double *tot);
const char *fused_error();

}  // namespace wx
