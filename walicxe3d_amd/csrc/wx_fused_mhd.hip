// wx_fused_mhd.hip - instantiation of the FP32 MHD stage kernels (HLLD and
// HLLE, 8 variables): the delta-form kernels of wx_mhd_kern.h (default) and
// the plain FP32 kernels of wx_fused_kern.h (tuning option).  Built with FMA
// contraction where a pragma does not turn it off: a face flux stays
// identical for the two blocks that share it because both compute it at the
// same code site (one per direction), and the per-cell FP64 primitives,
// which two sites compute for the same cell, are written without
// contraction.
#include <hip/hip_runtime.h>

#include "wx_mhd_kern.h"

namespace wx {

void launch_stage_mhd32(const FusedArgs &a, int grid, int solver, bool visc, int stage, int plain32, hipStream_t st) {
  if (plain32) {
    if (solver == SOLVER_HLLD) launch_stage_s<float, float, 8, SOLVER_HLLD>(a, grid, visc, stage, st);
    else launch_stage_s<float, float, 8, SOLVER_HLLE>(a, grid, visc, stage, st);
    return;
  }
  if (solver == SOLVER_HLLD) launch_mhd<SOLVER_HLLD>(a, grid, visc, stage, st);
  else launch_mhd<SOLVER_HLLE>(a, grid, visc, stage, st);
}

}  // namespace wx
