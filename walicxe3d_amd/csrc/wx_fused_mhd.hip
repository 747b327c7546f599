// wx_fused_mhd.hip - instantiation of the FP32 MHD stage kernels (HLLD and
// HLLE, 8 variables).  Built with FMA contraction: the fluxes of these
// kernels are FP32 arithmetic whose rounding the form does not pin, and a
// face flux stays identical for the two blocks that share it because both
// compute it with the same instructions (one code path per direction).
#include <hip/hip_runtime.h>

#include "wx_fused_kern.h"

namespace wx {

void launch_stage_mhd32(const FusedArgs &a, int grid, int solver, bool visc, int stage, hipStream_t st) {
  if (solver == SOLVER_HLLD) launch_stage_s<float, float, 8, SOLVER_HLLD>(a, grid, visc, stage, st);
  else launch_stage_s<float, float, 8, SOLVER_HLLE>(a, grid, visc, stage, st);
}

}  // namespace wx
