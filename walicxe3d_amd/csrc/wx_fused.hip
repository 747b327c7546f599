// wx_fused.hip - the fused FP32 form of the Godunov update.
//
// One stage = one kernel launch over all blocks of all levels.  A workgroup
// of 256 threads owns a block; thread (ti, tj) owns the column (i, j) and
// marches in k.  Per k-plane the primitives of the plane and its x/y face
// halo sit in LDS (x and y fluxes), the column's primitives of the planes
// around k sit in registers (z fluxes, Powell term).  Face halos are read
// straight from the neighbour blocks' interiors with the AMR exchange rules
// of wx_device.h - no ghost cells are stored in HBM.
//
// Precision: bulk flux/reconstruction arithmetic in FP32; conserved state U
// and its accumulation U + dt*dF in FP64; stage-1 pressure recovery from the
// FP64 state in FP64; the stage-2 epilogue (viscosity, primitives, cooling,
// dt candidate) in FP64 on U^{n+1}; dt reduced as an exact min.
//
// Stage 1 (first order, dt/2): U^n (FP64) -> UP (FP32).
// Stage 2 (second order, dt): UP, U^n -> U^{n+1} (in place), with the
// artificial viscosity (its face neighbours are the half-step ghosts, as in
// the reference), updatePrims, cooling and the next dt fused in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "wx_fused.h"
#include "wx_hilbert.h"
#include "wx_fused_kern.h"

namespace wx {

static thread_local std::string f_err;
const char *fused_error() { return f_err.c_str(); }

// initial dt candidate from U (after an upload)
template <int NV>
__global__ void __launch_bounds__(256) k_dt_init(FusedArgs a) {
  const DevCfg &c = a.c;
  int b = blockIdx.x;
  int l = c.info[b].level - 1;
  double m = 1.0e30;
  for (int q = threadIdx.x; q < NINT; q += 256) {
    double u[MAXV], pd[MAXV];
    for (int v = 0; v < NV; v++) u[v] = a.U[((size_t)b * NV + v) * NINT + q];
    flow2prim(c.pd, u, pd);
    m = fmin(m, dt_cell<NV>(c, pd, c.dx[0][l], c.dx[1][l], c.dx[2][l]));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMin(&a.step->dtmin_bits, (unsigned long long)__double_as_longlong(m));
}

// dt of the step: CFL, start-up ramp, output clamp
// (source/hydro_solver.f90:183-207); resets the running minimum
__global__ void k_dt_finalize(StepState *s, double cfl, int cooling) {
  double m = __longlong_as_double((long long)s->dtmin_bits);
  double dtl = cooling ? fmin(m, 1.0e30) * cfl : m * cfl;
  int it = s->it + 1;
  s->it = it;
  if (it <= 10) dtl = ldexp(1.0, -(11 - it)) * dtl;
  double tout = (double)s->nextout * s->dtout / s->t_sc;
  s->dump = 0;
  if (s->time + dtl >= tout) {
    dtl = tout - s->time;
    s->dump = 1;
  }
  s->dt = dtl;
  s->dtmin_bits = ~0ull;
}

__global__ void k_step_post(StepState *s) {
  s->time = s->time + s->dt;
  if (s->dump) s->nextout += 1;
}

// PRIM of the state in the same layout ([blk][nv][16^3]), for downloads
template <int NV>
__global__ void k_prims_int(const DevCfg c, int nblk, const double *src, double *dstP) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nblk * NINT) return;
  int q = (int)(t % NINT);
  size_t b = t / NINT;
  double u[MAXV], p[MAXV];
  for (int v = 0; v < NV; v++) u[v] = src[(b * NV + v) * NINT + q];
  flow2prim(c.pd, u, p);
  for (int v = 0; v < NV; v++) dstP[(b * NV + v) * NINT + q] = p[v];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int hip_fail(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  f_err = std::string(what) + ": " + hipGetErrorString(e);
  return 1;
}
#define FCHK(x) do { if (hip_fail((x), #x)) return 1; } while (0)

// Morton key of a block's position at the finest level: neighbours in space
// end up close in the processing order (shared L2 lines for halo reads)
static unsigned long long morton(const BlockInfo &B, int maxlev) {
  unsigned long long key = 0;
  int shift = maxlev - B.level;
  unsigned x = (unsigned)(B.coords[0] - 1) << shift, y = (unsigned)(B.coords[1] - 1) << shift,
           z = (unsigned)(B.coords[2] - 1) << shift;
  for (int bit = 0; bit < 20; bit++) {
    key |= (unsigned long long)((x >> bit) & 1) << (3 * bit);
    key |= (unsigned long long)((y >> bit) & 1) << (3 * bit + 1);
    key |= (unsigned long long)((z >> bit) & 1) << (3 * bit + 2);
  }
  return key;
}

// Block processing order: Morton (or Hilbert) order of the blocks'
// positions; with the XCD map, workgroup w runs on XCD w % 8, so each XCD
// gets a contiguous range of the curve (neighbour halos then hit that XCD's
// L2).
static int build_order(FusedState &f) {
  const int nb = f.nblk;
  std::vector<std::pair<unsigned long long, int>> keys(nb);
  if (f.curve == 1) {
    const int order = hilbert_order(f.nroot, f.maxlev);
    for (int b = 0; b < nb; b++)
      keys[b] = {block_hilbert_key(f.host_info[b].coords, f.host_info[b].level, f.maxlev, order), b};
  } else {
    for (int b = 0; b < nb; b++) keys[b] = {morton(f.host_info[b], f.maxlev), b};
  }
  std::sort(keys.begin(), keys.end());
  const int nx = 8;
  std::vector<int> ord(nb, -1);
  if (f.xcd_map && nb >= 64) {
    int full = (nb / nx) * nx;
    int perx = full / nx;
    for (int r = 0; r < full; r++) {
      int xcd = r / perx, within = r % perx;
      ord[within * nx + xcd] = keys[r].second;
    }
    for (int r = full; r < nb; r++) ord[r] = keys[r].second;
  } else {
    for (int r = 0; r < nb; r++) ord[r] = keys[r].second;
  }
  f.host_order = ord;
  FCHK(hipMemcpy(f.order, ord.data(), nb * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

int fused_alloc(FusedState &f, const DevCfg &c, const std::vector<BlockInfo> &info, hipStream_t st) {
  fused_free(f);
  f.nblk = c.nblk;
  f.nv = c.nv;
  f.two_u = !f.fp64 && c.nv == 8;
  size_t n = (size_t)c.nblk * c.nv * NINT;
  FCHK(hipMalloc(&f.U, n * sizeof(double)));
  if (f.two_u) FCHK(hipMalloc(&f.U2, n * sizeof(double)));
  FCHK(hipMalloc(&f.UP, n * (f.fp64 ? sizeof(double) : sizeof(float))));
  FCHK(hipMalloc(&f.order, c.nblk * sizeof(int)));
  FCHK(hipMalloc(&f.step, sizeof(StepState)));
  f.host_info = info;
  f.maxlev = c.maxlev;
  if (build_order(f)) return 1;
  for (int e = 0; e < 4; e++) FCHK(hipEventCreate(&f.ev[e]));
  (void)st;
  return 0;
}

static void drop_graphs(FusedState &f) {
  if (f.exec) hipGraphExecDestroy(f.exec);
  if (f.graph) hipGraphDestroy(f.graph);
  if (f.exec2) hipGraphExecDestroy(f.exec2);
  if (f.graph2) hipGraphDestroy(f.graph2);
  f.exec = f.exec2 = nullptr;
  f.graph = f.graph2 = nullptr;
  f.graph_u = f.graph2_u = nullptr;
}

void fused_free(FusedState &f) {
  drop_graphs(f);
  hipFree(f.U);
  hipFree(f.U2);
  hipFree(f.UP);
  hipFree(f.order);
  hipFree(f.step);
  for (int e = 0; e < 4; e++)
    if (f.ev[e]) hipEventDestroy(f.ev[e]);
  for (int e = 0; e < 4; e++) f.ev[e] = nullptr;
  f.U = nullptr;
  f.U2 = nullptr;
  f.UP = nullptr;
  f.order = nullptr;
  f.step = nullptr;
}

static FusedArgs make_args(FusedState &f, const DevCfg &c) {
  FusedArgs a;
  a.c = c;
  a.U = f.U;
  a.Uout = f.U2 ? f.U2 : f.U;
  a.UPin = f.UP;
  a.UPout = f.UP;
  a.order = f.order;
  a.step = f.step;
  a.nblk = f.nblk;
  a.nslab = 1;
  return a;
}

static void launch_dt_init(FusedState &f, const DevCfg &c, hipStream_t st) {
  FusedArgs a = make_args(f, c);
  if (c.nv == 8) hipLaunchKernelGGL(k_dt_init<8>, dim3(f.nblk), dim3(256), 0, st, a);
  else hipLaunchKernelGGL(k_dt_init<5>, dim3(f.nblk), dim3(256), 0, st, a);
}

int fused_upload(FusedState &f, const DevCfg &c, double *U, bool adopt, hipStream_t st) {
  if (adopt) {   // ownership first, so that no return path leaks U
    hipFree(f.U);
    f.U = U;
  } else {
    FCHK(hipMemcpyAsync(f.U, U, (size_t)f.nblk * f.nv * NINT * sizeof(double), hipMemcpyDeviceToDevice, st));
  }
  StepState s{};
  s.dtmin_bits = ~0ull;
  FCHK(hipMemcpyAsync(f.step, &s, sizeof(s), hipMemcpyHostToDevice, st));
  launch_dt_init(f, c, st);
  FCHK(hipGetLastError());
  FCHK(hipStreamSynchronize(st));
  return 0;
}

int fused_prims(FusedState &f, const DevCfg &c, double *P, hipStream_t st) {
  size_t n = (size_t)f.nblk * NINT;
  if (f.nv == 8)
    hipLaunchKernelGGL(k_prims_int<8>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c, f.nblk, f.U, P);
  else
    hipLaunchKernelGGL(k_prims_int<5>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c, f.nblk, f.U, P);
  FCHK(hipGetLastError());
  FCHK(hipStreamSynchronize(st));
  return 0;
}

int fused_timestep(FusedState &f, const DevCfg &c, hipStream_t st, double *dt_raw) {
  StepState s;
  FCHK(hipMemcpyAsync(&s, f.step, sizeof(s), hipMemcpyDeviceToHost, st));
  FCHK(hipStreamSynchronize(st));
  double m;
  std::memcpy(&m, &s.dtmin_bits, sizeof(m));
  *dt_raw = c.cooling ? std::fmin(m, 1.0e30) * c.cfl : m * c.cfl;
  return 0;
}

template <typename RF, typename ST>
static void launch_stage_t(const FusedArgs &a, int grid, int nv, int solver, bool visc, int stage, int plain32,
                           hipStream_t st) {
  if (nv == 8) {
    if constexpr (sizeof(RF) == 4) {
      launch_stage_mhd32(a, grid, solver, visc, stage, plain32, st);
    } else {
      if (solver == SOLVER_HLLD) launch_stage_s<RF, ST, 8, SOLVER_HLLD>(a, grid, visc, stage, st);
      else launch_stage_s<RF, ST, 8, SOLVER_HLLE>(a, grid, visc, stage, st);
    }
  } else {
    if (solver == SOLVER_HLLC) launch_stage_s<RF, ST, 5, SOLVER_HLLC>(a, grid, visc, stage, st);
    else launch_stage_s<RF, ST, 5, SOLVER_HLL>(a, grid, visc, stage, st);
  }
}

static void launch_stage(FusedState &f, const DevCfg &c, hipStream_t st, int stage) {
  FusedArgs a = make_args(f, c);
  bool visc = c.visc_eta != 0.0;
  // z slabs: the delta-form MHD kernels without viscosity (the viscous
  // epilogue works one plane behind and needs whole columns)
  if (!f.fp64 && c.nv == 8 && !f.mhd_plain32 && !visc) a.nslab = f.nslab;
  const int items = f.nblk * a.nslab;
  int grid = f.grid > 0 ? std::min(f.grid, items) : items;
  if (f.fp64) launch_stage_t<double, double>(a, grid, c.nv, c.pd.solver, visc, stage, 0, st);
  else launch_stage_t<float, float>(a, grid, c.nv, c.pd.solver, visc, stage, f.mhd_plain32, st);
}

static void enqueue_step(FusedState &f, const DevCfg &c, hipStream_t st) {
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(1), 0, st, f.step, c.cfl, c.cooling);
  launch_stage(f, c, st, 1);
  launch_stage(f, c, st, 2);
  hipLaunchKernelGGL(k_step_post, dim3(1), dim3(1), 0, st, f.step);
}

// after a step of the two-buffer form the new state is in the other buffer
static void swap_u(FusedState &f) {
  if (f.U2) std::swap(f.U, f.U2);
}

static int capture_step(FusedState &f, const DevCfg &c, hipStream_t st, hipGraph_t *g, hipGraphExec_t *e) {
  FCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  enqueue_step(f, c, st);
  hipError_t err = hipStreamEndCapture(st, g);
  FCHK(err);
  FCHK(hipGraphInstantiate(e, *g, nullptr, nullptr, 0));
  return 0;
}

// The step graphs: ex[0] steps from the current U, ex[1] from the other
// buffer of the two-buffer form (the same graph otherwise).  Captured again
// whenever a buffer moved since the capture.
static int step_graphs(FusedState &f, const DevCfg &c, hipStream_t st, hipGraphExec_t ex[2]) {
  const bool ok = f.exec && f.graph_u == f.U && (!f.U2 || (f.exec2 && f.graph2_u == f.U2));
  if (!ok) {
    drop_graphs(f);
    if (capture_step(f, c, st, &f.graph, &f.exec)) return 1;
    f.graph_u = f.U;
    if (f.U2) {
      swap_u(f);
      const int rc = capture_step(f, c, st, &f.graph2, &f.exec2);
      swap_u(f);
      if (rc) return 1;
      f.graph2_u = f.U2;
    }
  }
  ex[0] = f.exec;
  ex[1] = f.U2 ? f.exec2 : f.exec;
  return 0;
}

int fused_run(FusedState &f, const DevCfg &c, hipStream_t st, int nsteps, int *it, double *dt, double *time,
              int *nextout, double dtout, double t_sc) {
  if (nsteps == 0) return 0;
  // host scalars -> device (keeps the running dt minimum)
  StepState s;
  FCHK(hipMemcpyAsync(&s, f.step, sizeof(s), hipMemcpyDeviceToHost, st));
  FCHK(hipStreamSynchronize(st));
  s.it = *it;
  s.time = *time;
  s.nextout = *nextout;
  s.dtout = dtout;
  s.t_sc = t_sc;
  FCHK(hipMemcpyAsync(f.step, &s, sizeof(s), hipMemcpyHostToDevice, st));
  hipGraphExec_t ex[2];
  if (step_graphs(f, c, st, ex)) return 1;
  FCHK(hipEventRecord(f.ev[0], st));
  for (int n = 0; n < nsteps; n++) {
    FCHK(hipGraphLaunch(ex[n & 1], st));
    swap_u(f);
  }
  FCHK(hipEventRecord(f.ev[1], st));
  FCHK(hipMemcpyAsync(&s, f.step, sizeof(s), hipMemcpyDeviceToHost, st));
  FCHK(hipStreamSynchronize(st));
  float ms = 0;
  hipEventElapsedTime(&ms, f.ev[0], f.ev[1]);
  f.t_tot = nsteps > 0 ? ms / nsteps : 0.0;
  *it = s.it;
  *dt = s.dt;
  *time = s.time;
  *nextout = s.nextout;
  return 0;
}

// Times the two stage kernels separately (HIP events on the library stream)
// by replaying one step outside the graph; used by the benchmark.
int fused_time_stages(FusedState &f, const DevCfg &c, hipStream_t st, int reps) {
  float a1 = 0, a2 = 0;
  for (int r = 0; r < reps; r++) {
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(1), 0, st, f.step, c.cfl, c.cooling);
    FCHK(hipEventRecord(f.ev[0], st));
    launch_stage(f, c, st, 1);
    FCHK(hipEventRecord(f.ev[1], st));
    launch_stage(f, c, st, 2);
    FCHK(hipEventRecord(f.ev[2], st));
    hipLaunchKernelGGL(k_step_post, dim3(1), dim3(1), 0, st, f.step);
    swap_u(f);
    FCHK(hipEventSynchronize(f.ev[2]));
    float m1, m2;
    hipEventElapsedTime(&m1, f.ev[0], f.ev[1]);
    hipEventElapsedTime(&m2, f.ev[1], f.ev[2]);
    a1 += m1;
    a2 += m2;
  }
  f.t_s1 = a1 / reps;
  f.t_s2 = a2 / reps;
  f.timing_runs = reps;
  return 0;
}

// kind 0: stage-kernel grid (0 = one workgroup per block, else a persistent
//         grid of that many workgroups striding over the block list)
// kind 1: XCD-aware block order (1) or one range for all XCDs (0)
// kind 2: block order along the Morton (0) or the Hilbert (1) curve
// kind 3: FP32 MHD fluxes in delta form (0) or plain FP32 (1)
// kind 4: z slabs per block (1, 2, 4, 8), one workgroup each, in the
//         delta-form MHD kernels without viscosity (no effect elsewhere)
int fused_set_tuning(FusedState &f, int kind, int value) {
  if (kind < 0 || kind > 4) return 1;
  if ((kind == 2 || kind == 3) && (value < 0 || value > 1)) return 1;
  if (kind == 4 && value != 1 && value != 2 && value != 4 && value != 8) return 1;
  drop_graphs(f);
  if (kind == 3) {
    f.mhd_plain32 = value;
    return 0;
  }
  if (kind == 4) {
    f.nslab = value;
    return 0;
  }
  if (kind == 0) {
    if (value < 0) return 1;
    f.grid = value;
    return 0;
  }
  if (kind == 1)
    f.xcd_map = value != 0;
  else
    f.curve = value;
  return f.nblk > 0 ? build_order(f) : 0;
}

void fused_last_timing(const FusedState &f, double *s1, double *s2, double *tot) {
  if (s1) *s1 = f.t_s1;
  if (s2) *s2 = f.t_s2;
  if (tot) *tot = f.t_tot;
}

}  // namespace wx
