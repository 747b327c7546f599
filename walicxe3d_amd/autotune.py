"""Launch-shape autotuner for the fused form.

Tunes, per workload (mesh size, variable count, solver), the choices the
fused stage kernels leave open:
  - grid: one workgroup per block, or a persistent grid of k x 256
    workgroups (256 CUs) striding over the block list;
  - block order: along the Morton or the Hilbert curve, as one range or as
    8 ranges mapped to XCDs (workgroup w runs on XCD w % 8, so each XCD's L2
    serves its own halo reads);
  - sub-block tiling: each 16^3 block split into 1, 2, 4 or 8 z slabs, one
    workgroup each (FP32 MHD kernels without viscosity), so that a mesh of
    few blocks still fills the 256 CUs.
Each candidate runs a few timed steps (HIP events inside wx_run) on the live
state; the state advances, so tune on a scratch context.
"""
from __future__ import annotations

import itertools
import time

NUM_CUS = 256


def slab_choices(solver):
    """z slabs per block the solver's kernels can use: the FP32 MHD form
    without viscosity splits blocks; every other kernel runs whole blocks."""
    from walicxe3d_amd import hip
    w = solver.w
    if solver.form != hip.FUSED_FP32 or w.nv != 8 or w.visc_eta != 0.0:
        return (1,)
    return (1, 2, 4, 8)


def candidates(nblk, slabs=(1,)):
    out = []
    for nslab in slabs:
        items = nblk * nslab
        grids = [0] + [k * NUM_CUS for k in (2, 4, 8) if k * NUM_CUS < items]
        out += [(g, x, c, nslab) for g, x, c in itertools.product(grids, (1, 0), (0, 1))]
    return out


CONFIRM = 4   # candidates measured a second time, with 4x the steps


def tune(solver, steps=3, verbose=False):
    """Try every candidate on `solver` (a hip.Walicxe3D in a fused form),
    then the CONFIRM fastest again with 4x the steps; returns (best, table)
    with best = {"grid": g, "xcd_map": x, "curve": c (0 Morton, 1 Hilbert),
    "nslab": s, "ms": t, "ms_confirm": t2}, chosen by ms_confirm."""
    table = []
    for grid, xcd, curve, nslab in candidates(solver.nblk, slab_choices(solver)):
        solver.set_tuning(0, grid)
        solver.set_tuning(1, xcd)
        solver.set_tuning(2, curve)
        solver.set_tuning(4, nslab)
        solver.run(1)                      # graph capture, warm caches
        t0 = time.perf_counter()
        solver.run(steps)
        ms = 1e3 * (time.perf_counter() - t0) / steps
        table.append({"grid": grid, "xcd_map": xcd, "curve": curve, "nslab": nslab, "ms": ms})
        if verbose:
            print("  grid %5d xcd %d curve %d slabs %d: %.3f ms/step" % (grid, xcd, curve, nslab, ms), flush=True)
    # the fastest few again, with more steps each: one short measurement
    # does not separate choices that differ by less than its own noise
    top = sorted(table, key=lambda r: r["ms"])[:CONFIRM]
    for r in top:
        solver.set_tuning(0, r["grid"])
        solver.set_tuning(1, r["xcd_map"])
        solver.set_tuning(2, r["curve"])
        solver.set_tuning(4, r["nslab"])
        solver.run(1)
        t0 = time.perf_counter()
        solver.run(4 * steps)
        r["ms_confirm"] = 1e3 * (time.perf_counter() - t0) / (4 * steps)
    best = min(top, key=lambda r: r["ms_confirm"])
    solver.set_tuning(0, best["grid"])
    solver.set_tuning(1, best["xcd_map"])
    solver.set_tuning(2, best["curve"])
    solver.set_tuning(4, best["nslab"])
    return best, table
