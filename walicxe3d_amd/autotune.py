"""Launch-shape autotuner for the fused form.

Tunes, per workload (mesh size, variable count, solver), the choices the
fused stage kernels leave open:
  - grid: one workgroup per block, or a persistent grid of k x 256
    workgroups (256 CUs) striding over the block list;
  - block order: along the Morton or the Hilbert curve, as one range or as
    8 ranges mapped to XCDs (workgroup w runs on XCD w % 8, so each XCD's L2
    serves its own halo reads).
Each candidate runs a few timed steps (HIP events inside wx_run) on the live
state; the state advances, so tune on a scratch context.
"""
from __future__ import annotations

import itertools
import time

NUM_CUS = 256


def candidates(nblk):
    grids = [0] + [k * NUM_CUS for k in (2, 4, 8) if k * NUM_CUS < nblk]
    return list(itertools.product(grids, (1, 0), (0, 1)))


def tune(solver, steps=3, verbose=False):
    """Try every candidate on `solver` (a hip.Walicxe3D in a fused form);
    returns (best, table) with best = {"grid": g, "xcd_map": x,
    "curve": c (0 Morton, 1 Hilbert), "ms": t}."""
    table = []
    for grid, xcd, curve in candidates(solver.nblk):
        solver.set_tuning(0, grid)
        solver.set_tuning(1, xcd)
        solver.set_tuning(2, curve)
        solver.run(1)                      # graph capture, warm caches
        t0 = time.perf_counter()
        solver.run(steps)
        ms = 1e3 * (time.perf_counter() - t0) / steps
        table.append({"grid": grid, "xcd_map": xcd, "curve": curve, "ms": ms})
        if verbose:
            print("  grid %5d xcd %d curve %d: %.3f ms/step" % (grid, xcd, curve, ms), flush=True)
    best = min(table, key=lambda r: r["ms"])
    solver.set_tuning(0, best["grid"])
    solver.set_tuning(1, best["xcd_map"])
    solver.set_tuning(2, best["curve"])
    return best, table
