/*
 * walicxe_hip.h - C ABI of the MI355X (gfx950) implementation of
 * Walicxe3D's per-block Godunov update.
 *
 * Plain C types only; every call returns an int status:
 *    0            success
 *   >0            success, with that many cells floored (pressure/density)
 *   WX_E* (<0)    failure (see wx_last_error for text)
 * The device context is opaque.  One context drives one GPU (the current HIP
 * device at wx_init).  Calls on one context must not overlap.
 *
 * Block state is exchanged either in the library layout (WX_LAYOUT_BLOCKS:
 * [nblk][nv][16][16][16], i fastest, blocks in the order given to
 * wx_set_mesh) or in the reference's Fortran layout (WX_LAYOUT_FORTRAN:
 * U(nbMaxProc, nv, -1:18, -1:18, -1:18), block slot fastest, with the
 * slot->bID map `localBlocks`).
 *
 * Each entry point names the reference routine whose interface it replaces
 * (paths relative to the Walicxe3D source tree).
 */
#ifndef WALICXE_HIP_H
#define WALICXE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define WX_OK 0
#define WX_EINVAL (-1)
#define WX_EHIP (-2)
#define WX_ENOMEM (-3)
#define WX_ESTATE (-4)
#define WX_ETOPO (-5)
#define WX_EIO (-6)

/* numerical forms */
#define WX_FORM_PLAIN_FP64 0   /* one kernel per reference routine, FP64, reference order */
#define WX_FORM_FUSED_FP32 1   /* one fused kernel per stage, FP32 bulk, FP64 accumulation */
#define WX_FORM_FUSED_FP64 2   /* the fused kernels with FP64 bulk arithmetic (precision reference) */

#define WX_LAYOUT_BLOCKS 0
#define WX_LAYOUT_FORTRAN 1

/* which array */
#define WX_U 0
#define WX_UP 1
/* cell ranges (source/constants.f90:119-121) */
#define WX_CELLS_ALL 0
#define WX_CELLS_PHYS 1
#define WX_CELLS_GHOST 2

/* Compile-time parameters of the reference (source/parameters.f90) as a
   run-time struct.  Enumerations use the reference's values
   (source/constants.f90:52-104). */
typedef struct wx_config {
  int nv;            /* neqtot: 5 hydro, 8 MHD; 6 or 7 hydro with one or two
                        passive scalars (plain FP64 form, no wind sources) */
  int mhd;           /* parameters `mhd` */
  int solver;        /* SOLVER_HLL1=2, HLL=3, HLLC=4, HLLE=5, HLLD=6 */
  int limiter;       /* LIMITER_NO_AVERAGE=-1 ... SUPERBEE=6 */
  int eight_wave;    /* Powell divergence source (sources.f90) */
  int cooling;       /* COOL_NONE=0, COOL_TABLE=1 */
  int eos;           /* EOS_ADIABATIC=0, SINGLE_SPECIE=1, TWOTEMP=2 */
  int bc[6];         /* LEFT..TOP: BC_REFLECTIVE=1, FREE_FLOW=2, PERIODIC=3 */
  double gamma, cv, cfl, visc_eta;
  double cooling_limit, mu0, mui, ion_thres;
  double d_sc, e_sc, t_sc, p_sc, l_sc, v_sc;
  int ntab;                  /* cooling table (loadcooldata, cooling.f90:135) */
  const double *tab_logT;
  const double *tab_loglam;  /* log10(Lambda) */
  int nwind;                 /* spherical wind sources re-imposed at every */
  const double *winds;       /* boundary call (winds.f90:123): 8 per source:
                                xc,yc,zc,radius,mdot,vinf,temp,mu (cgs)   */
  int form;                  /* WX_FORM_* */
} wx_config;

/* Set-up.  Replaces the allocation part of initmain (source/init.f90:41-480). */
int wx_init(const wx_config *cfg, void **ctx);
int wx_finalize(void *ctx);
const char *wx_last_error(void);
/* Library version string. */
const char *wx_version(void);

/* Mesh: root blocks per dimension, number of levels, grid spacings
   dxyz[3*maxlev] (dx(1..maxlev), dy(..), dz(..): basegrid,
   source/init.f90:726-730) and the active block IDs.  Builds the face
   neighbour table that source/amr.f90:539-631 `neighbors` computes per
   query.  Must be called again whenever the block list changes.  Returns
   WX_ETOPO when a periodic domain face wraps onto a block that is not
   active at the same level (the periodic fill copies same-level blocks). */
int wx_set_mesh(void *ctx, const int nroot[3], int maxlev, const double *dxyz, int nblk,
                const int *bids);

/* A plane wind (imposePlaneWind, source/winds.f90:250-340, as a problem's
   userBoundary, source/boundary.f90:61, would call it after its spherical
   sources): at every boundary call, cells k = 16..18 (the top interior plane
   and the two ghost layers, every i and j) of each block at the domain's top
   edge take the state rho [g cm^-3], temp [K], mu, velocity vel [cm s^-1]
   along +x (plane PLANE_LEFT=1), -x (PLANE_RIGHT=2), +y (PLANE_FRONT=3),
   -y (PLANE_BACK=4), +z (PLANE_BOTTOM=5) or -z (PLANE_TOP=6).  As in the
   reference, the cells are those of the top edge whatever the plane.
   plane = 0 removes it.  Plain FP64 form and hydro only (the reference leaves
   the field of the state unset): other contexts get WX_EINVAL.  No
   counterpart call in the reference: it replaces the user module's call. */
int wx_set_plane_wind(void *ctx, int plane, double rho, double vel, double temp, double mu);

/* A uniform acceleration (gx, gy, gz) in code units, the source term a
   problem's get_user_source_terms (source/sources.f90:124, called per cell
   from upwindStep, source/godunov.f90:266-271, with user_source_terms on)
   adds from the cell's primitives: s(2:4) += rho*g, s(5) += rho*(v.g), ahead
   of the Powell terms, so that each stage adds dtp*s.  The hook's arguments
   (the cell's primitives and its in-block indices) carry no block position,
   and this is the one such source the library runs.  on = 0 removes it.
   Plain FP64 form (hydro and MHD): fused contexts get WX_EINVAL.  No
   counterpart call in the reference: it replaces the user module's code. */
int wx_set_uniform_gravity(void *ctx, int on, double gx, double gy, double gz);

/* State transfer.  Upload sets U's interior, zeroes its ghosts and computes
   PRIM (updatePrims, source/hydro_core.f90:43-59).  For WX_LAYOUT_FORTRAN
   pass nbmax = nbMaxProc and local_blocks = localBlocks; slots whose bID is
   not in the mesh are ignored. */
int wx_upload_state(void *ctx, const double *u, int layout, int nbmax, const int *local_blocks);
/* Overwrite the interior primitives (plain form).  The reference's stage 1
   reads the interior PRIM left by updatePrims and cooling, whose pressure
   update (source/cooling.f90:361-363) is not flow2prim of the cooled U; a
   host that runs cooling itself passes its PRIM here after the upload. */
int wx_upload_prims(void *ctx, const double *prim, int layout, int nbmax, const int *local_blocks);
/* Download interior U and PRIM (either may be NULL). */
int wx_download_state(void *ctx, double *u, double *prim, int layout, int nbmax,
                      const int *local_blocks);

/* getTimestep (source/hydro_solver.f90:129-211) without the start-up ramp
   and output clamp, which need `it`, `time` and `nextout` (host scalars):
   dt_raw = CFL * min over interior cells of dx/(|v|+c). */
int wx_get_timestep(void *ctx, double *dt_raw);
/* Godunov(order) (source/godunov.f90:43-190): boundary, primitives, fluxes
   and conservative update of every block, one or two stages. */
int wx_godunov(void *ctx, int order, double dt);
/* doViscousStep (source/hydro_solver.f90:217-250) without `time += dt`. */
int wx_viscous_step(void *ctx);
/* updatePrims (source/hydro_core.f90:43-59). */
int wx_update_prims(void *ctx);
/* cooling (source/cooling.f90:37-118), COOL_TABLE. */
int wx_cooling(void *ctx, double dt);
/* One iteration of the main loop (source/main.f90:81-100): it += 1, dt from
   getTimestep with ramp and clamp (tout = nextout*dtout/t_sc), hydroSolver,
   updatePrims, cooling; advances *time and, when the clamp fired,
   *nextout.  dtout and t_sc are the reference's parameters. */
int wx_step(void *ctx, int *it, double *dt, double *time, int *nextout, double dtout, double t_sc);
/* Advance nsteps iterations with no host synchronisation between steps (the
   ramp and clamp run on the device); fused form only, captured in a HIP
   graph per mesh. */
int wx_run(void *ctx, int nsteps, int *it, double *dt, double *time, int *nextout, double dtout,
           double t_sc);

/* Fine-grained entry points for parity tests of single routines. */
/* boundary(depth, uvars) (source/boundary.f90:46-64), which = WX_U|WX_UP */
int wx_boundary(void *ctx, int depth, int which);
/* calcPrimsAll(uvars, PRIM, cells) (source/hydro_core.f90:67-101) */
int wx_calc_prims(void *ctx, int which, int cells);
/* HLL*fluxes(b, order) + upwindStep(b, dtp) for every block
   (source/godunov.f90:95-125,199-277) */
int wx_fluxes_update(void *ctx, int order, double dtp);
/* Raw device arrays with ghosts ([nblk][nv][20][20][20]) copied to the host;
   which = 0 U, 1 UP, 2 PRIM.  Plain form only. */
int wx_copy_raw(void *ctx, int which, double *out);

/* Change of the active block list with the state kept on the device
   (admesh's topology change, source/amr.f90:44-133, with the data motion of
   refineBlock, :284-384, and coarsenBlock, :1430-1650): a block present in
   both lists is copied, a new block under an old ancestor takes the
   ancestor's cells by injection, a new block whose 8 children were active
   takes their 2x2x2 means (z outer, x inner, then /8).  Then as wx_set_mesh
   followed by an upload of that state (face table rebuilt, the step graph
   recaptured on the next wx_run).  The plain form carries PRIM as the
   reference does: kept by copied blocks, injected into refined children,
   recomputed from U in new coarse blocks; the fused form recomputes its
   state from U.  Returns WX_ETOPO, with the mesh and state untouched, when a
   new block would need a restriction over more than one level or the new
   list is not a mesh wx_set_mesh accepts (a jump of two levels across a
   face, a periodic face without a partner). */
int wx_remesh(void *ctx, int nblk, const int *bids);

/* Refinement flags (WX_FLAG_*) of admesh */
#define WX_FLAG_NONE 0
#define WX_FLAG_REFINE 1
#define WX_FLAG_COARSE 2

/* markByPhysical (source/amr.f90:142-230) of every block, on the device:
   the largest relative gradient |q(i+1)-q(i-1)|/q(i)/(2 dx) of pressure and
   density over the interior (maxgrad, may be NULL) and the flag it gives:
   REFINE at or above refine_thres (NONE at the maximum level), else COARSE
   at or below coarse_thres above level 1, else NONE.  The plain form reads
   PRIM with the ghost layer its last step left (the exchange of the state the
   step started from), as the reference's admesh does after a step.  The
   fused form keeps no ghost layers: it reads the layer wx_record_ghosts
   recorded, or else exchanges the current state. */
int wx_mark_blocks(void *ctx, double refine_thres, double coarse_thres, int *flags, double *maxgrad);

/* Fused form: records the depth-1 PRIM ghost layer of the current state (the
   boundary + calcPrimsAll(CELLS_GHOST) of the reference's first stage,
   source/godunov.f90:78-83), which wx_mark_blocks / wx_admesh then read, so
   that called right before the step that admesh follows the flags are the
   reference's (its markByPhysical reads that layer).  Kept until the next
   upload or mesh change.  No-op in the plain form. */
int wx_record_ghosts(void *ctx);

/* The mesh change of admesh (source/amr.f90:44-133) for given flags, host
   only: checkProximity (:1109-1254), then every block flagged REFINE
   replaced by its 8 children (refineBlock, :284-384) and every family of 8
   siblings all flagged COARSE by their father (coarseBlocks, :1430-1650).
   The new list keeps the old order (children in place of their father, x
   fastest; a father in place of the first of its children).  flags_out (may
   be NULL) gets the flags after the proximity rule.  *nblk_new is set even
   when the list does not fit in cap (WX_ENOMEM). */
int wx_admesh_plan(const int nroot[3], int maxlev, int nblk, const int *bids, const int *flags,
                   int *flags_out, int cap, int *nblk_new, int *bids_new);

/* admesh (source/amr.f90:44-133) on the device state: wx_mark_blocks,
   wx_admesh_plan, then wx_remesh to the new list (skipped when the list is
   unchanged).  The new list is returned in bids_new (cap entries).  Returns
   WX_ENOMEM or WX_ETOPO (a list beyond cap, or a mesh the face table cannot
   hold, such as a jump of two levels) with the mesh and state untouched. */
int wx_admesh(void *ctx, double refine_thres, double coarse_thres, int cap, int *nblk_new, int *bids_new);

/* Hilbert key of every block (getHilbertKey, source/amr.f90:2031-2061, with
   HKey3, source/hilbert.f90:38-127): the key of the block's lowest
   finest-level cell on the reference's 3-D Hilbert curve; sorting by key
   gives the reference's HilbertOrder (amr.f90:1930-1996), the order its
   load balance deals blocks to ranks in.  Host only. */
int wx_hilbert_keys(const int nroot[3], int maxlev, int nblk, const int *bids, int *keys);

/* The device implementations of single physics routines applied to n
   records of host arrays (FP64, the plain form's arithmetic), for parity
   tests against the reference's own per-routine outputs.  op:
     WX_EVAL_FLOW2PRIM   flow2prim (source/hydro_core.f90:323-378):
                         in nv -> out nv + 1 (primitives, then istat)
     WX_EVAL_PRIM2FLOW   prim2flow (:278-313): in nv -> out nv
     WX_EVAL_PRIM2FLUXES prim2fluxes (:387-442) along arg = 1, 2, 3 (x, y, z):
                         in nv -> out nv
     WX_EVAL_LIMITER     limiter (:613-701) with option arg: in 4 nv (pll, pl,
                         pr, prr) -> out 2 nv (pl, pr)
     WX_EVAL_RIEMANN     primfhll / primfhllc / primfhlle / primfhlld
                         (source/hll.f90:175, hllc.f90:209, hlle.f90:166,
                         hlld.f90:166) for solver arg: in 2 nv (pl, pr) ->
                         out nv
     WX_EVAL_RIEMANN_SEL primfhlld in the one-code-path form of the fused
                         FP64 kernels: in 2 nv -> out nv
     WX_EVAL_SPEEDS      sound, cfast (x, y, z), cfastX (:503-566): in nv ->
                         out 5
   cfg supplies nv, mhd, solver (primfhllc forms its contact speed only in
   HLLC builds, source/hllc.f90:348-353), limiter, gamma and cv.  A
   diagnostic entry point with no counterpart routine of its own in the
   reference. */
#define WX_EVAL_FLOW2PRIM 0
#define WX_EVAL_PRIM2FLOW 1
#define WX_EVAL_PRIM2FLUXES 2
#define WX_EVAL_LIMITER 3
#define WX_EVAL_RIEMANN 4
#define WX_EVAL_RIEMANN_SEL 5
#define WX_EVAL_SPEEDS 6
int wx_eval_physics(const wx_config *cfg, int op, int arg, int n, const double *in, double *out);

/* wx_eval_physics without any device allocation: the same records, checks
   and results (bit for bit), staged through a fixed array of
   WX_EVAL_NOALLOC_DOUBLES doubles (4 MiB) that each GPU the library runs on
   holds until the process ends.  n * (record in + record out doubles) must
   not exceed it (WX_EINVAL otherwise: at most 10922 limiter records of 8
   variables).  Every call on one GPU shares that GPU's array: calls that
   overlap on one GPU (from several host threads) race.  No counterpart in
   the reference; for callers that cannot allocate device memory (harnesses
   that forbid hipMalloc, tight evaluation loops). */
#define WX_EVAL_NOALLOC_DOUBLES (1 << 19)
int wx_eval_physics_noalloc(const wx_config *cfg, int op, int arg, int n, const double *in, double *out);

/* Checkpoint files of the reference, byte for byte.  writeBin
   (source/output.f90:156-214): the device state to a stream file (int32
   block count; per block int32 bID and U(1:neqtot,1:16,1:16,1:16), variable
   fastest), neqtot = flow variables + npassive: the passive scalars the
   context carries (nv = 6, 7), then zero columns for the rest; npassive
   below the carried count is refused. */
int wx_write_bin(void *ctx, const char *path, int npassive);
/* warmstart's read of that file (source/init.f90:582-610) into host arrays:
   bids[cap] and u[cap][nv][16][16][16] (WX_LAYOUT_BLOCKS), passive columns
   skipped; then wx_set_mesh + wx_upload_state.  *nblk is set even when the
   file holds more than cap blocks (WX_ENOMEM).  Host only. */
int wx_read_bin(const char *path, int nv, int npassive, int cap, int *nblk, int *bids, double *u);
/* writeState (source/output.f90:219-268): time*t_sc, it, noutput as
   '(es22.15,i8,i5)', then the data directory.  Host only. */
int wx_write_state(const char *path, double time, double t_sc, int it, int noutput, const char *datadir);
/* warmstart's read of it (source/init.f90:545-560): time = t/t_sc,
   nextout = noutput + 1, and the data directory (len bytes). */
int wx_read_state(const char *path, double t_sc, double *time, int *it, int *nextout, char *datadir, int len);

/* The VTK part of writeOutput (source/output.f90:38-151) on one rank:
   updatePrims, then write3DVTKBlocks (:273-750: datadir/Blocks000.NNNN.vtk,
   every cell a hexahedron, rho, vel, P, B, T and npassive passive scalars
   (those the context carries, zero for the rest; npassive below the
   carried count is refused), FP32 big-endian; master.visit replaced at output 0
   and appended to after) and write3DVTKGrid (:755-895: datadir/Grid.NNNN.vtk,
   one hexahedron per block with rank, bID and Hilbert key).  units is the
   reference's units_type: 1 PHYS_UNITS (scaled by l_sc, d_sc, v_sc, p_sc,
   b_sc), 2 CODE_UNITS. */
int wx_write_vtk(void *ctx, const char *datadir, int noutput, int npassive, int units);

/* Launch and order choices of the fused form (autotuner): kind 0 = stage
   kernel grid (0 one workgroup per block, else a persistent grid of that
   many workgroups), kind 1 = block order split into one curve range per XCD
   (1) or not (0), kind 2 = block order along the Morton (0) or Hilbert (1)
   curve, kind 3 = the FP32 MHD kernels' fluxes in delta form (0: FP64
   physical flux plus FP32 Riemann correction, the default) or entirely in
   FP32 (1: faster, with FP32 rounding of every flux term), kind 4 = z slabs
   per block (1, 2, 4 or 8), one workgroup each, in the delta-form MHD
   kernels without viscosity (more workgroups for meshes of few blocks;
   results unchanged; no effect elsewhere). */
int wx_set_tuning(void *ctx, int kind, int value);
/* Advances `reps` more iterations launching each stage kernel on its own
   between HIP events on the library's stream (fused form; no reference
   counterpart: measurement only).  The step scalars advance on the device
   exactly as in wx_run. */
int wx_time_stages(void *ctx, int reps);
/* Average stage-kernel durations (ms) of the last wx_time_stages and the
   average time per step of the last wx_run; zero if none. */
int wx_last_timing(void *ctx, double *stage1_ms, double *stage2_ms, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
