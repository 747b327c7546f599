"""The reference's main program (source/main.f90:32-134) on the device.

    python -m walicxe3d_amd.run --workload amr_adapt --form fused --datadir DATA \\
        --bin --vtk [--steps N] [--no-admesh] [--warm DATA/State.0003.dat]

Follows the reference's order: initial condition (or warm start from a
State file and its blocks file: warmstart, source/init.f90:527-623), output 0,
then while time <= tfin: the step (getTimestep with ramp and output clamp,
hydroSolver, updatePrims, cooling), admesh (source/amr.f90:44-133), and
writeOutput (source/output.f90:38-151: blocks file, VTK files, State file)
whenever the output clamp fired.  Load balance has nothing to do on one
rank.  In the fused form the ghost layer admesh reads is recorded before each
step, so its decisions are the reference's; in the plain form the VTK output
is followed by updatePrims, as writeOutput does, so later steps are those of a
reference run with VTK output on.
"""
from __future__ import annotations

import argparse
import os
import sys
import time as _time

import numpy as np

FORMS = {"plain": 0, "fused": 1, "fused64": 2}


def write_output(s, w, datadir, noutput, binout, vtkout, npassive=0):
    """writeOutput(noutput) of the current state."""
    from walicxe3d_amd import hip
    if binout:
        s.writeBin(os.path.join(datadir, "Blocks000.%04d.bin" % noutput), npassive)
    if vtkout:
        s.write_vtk(datadir, noutput, npassive)
        if s.form == hip.PLAIN_FP64:
            s.updatePrims()
    if binout or vtkout:
        hip.write_state(os.path.join(datadir, "State.%04d.dat" % noutput), s.time.value, w.t_sc,
                        s.it.value, noutput, datadir if datadir.endswith("/") else datadir + "/")


def simulate(w, form="fused", steps=None, admesh=True, datadir=None, binout=False, vtkout=False, warm=None,
             refine_thres=0.3, coarse_thres=0.03, log=None):
    """Runs the main loop; returns the hip.Walicxe3D context (state on the
    device) after the last step."""
    from walicxe3d_amd import hip
    from walicxe3d_amd import workloads as W
    say = log or (lambda *a: None)
    f = FORMS[form] if isinstance(form, str) else form
    if warm:
        t, it, nextout, olddir = hip.read_state(warm, w.t_sc)
        nout = nextout - 1
        # the blocks file lies in the data folder the State file names
        # (relative names from where the run started), else beside it
        path = os.path.join(olddir, "Blocks000.%04d.bin" % nout)
        if not os.path.exists(path):
            path = os.path.join(os.path.dirname(warm), "Blocks000.%04d.bin" % nout)
        bids, U0 = hip.read_bin(path, w.nv)
    else:
        _, blocks, U0 = W.initial_state(w)
        bids = np.asarray(blocks, dtype=np.int32)
    s = hip.Walicxe3D(w, form=f, bids=bids)
    s.upload(U0)
    if warm:
        s.it.value, s.time.value, s.nextout.value = it, t, nextout
    if datadir:
        os.makedirs(datadir, exist_ok=True)
        if not warm:
            write_output(s, w, datadir, 0, binout, vtkout)
            s.nextout.value = 1
    n = 0
    while s.time.value <= w.tfin / w.t_sc and (steps is None or n < steps):
        if admesh and f != hip.PLAIN_FP64:
            s.record_ghosts()
        before = s.nextout.value
        s.step()
        n += 1
        if admesh:
            s.admesh(refine_thres, coarse_thres)
        if s.nextout.value != before:
            say("it %d  t = %.6e  blocks %d  output %d" % (s.it.value, s.time.value * w.t_sc, s.nblk, before))
            if datadir:
                write_output(s, w, datadir, before, binout, vtkout)
    return s


def main(argv=None):
    from walicxe3d_amd import workloads as W
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="amr_adapt")
    ap.add_argument("--form", choices=list(FORMS), default="fused")
    ap.add_argument("--steps", type=int, default=None, help="stop after this many steps (default: run to tfin)")
    ap.add_argument("--no-admesh", action="store_true", help="keep the initial mesh")
    ap.add_argument("--refine-thres", type=float, default=0.3)
    ap.add_argument("--coarse-thres", type=float, default=0.03)
    ap.add_argument("--datadir", default=None)
    ap.add_argument("--bin", action="store_true", help="write the blocks and State files at each output")
    ap.add_argument("--vtk", action="store_true", help="write the VTK files at each output")
    ap.add_argument("--warm", default=None, help="State.NNNN.dat to start from")
    ap.add_argument("--tfin", type=float, default=None, help="final time (physical units, as the reference's tfin)")
    ap.add_argument("--dtout", type=float, default=None, help="time between outputs (physical units)")
    args = ap.parse_args(argv)
    w = W.get(args.workload)
    if args.tfin is not None or args.dtout is not None:
        import dataclasses
        w = dataclasses.replace(w, tfin=args.tfin if args.tfin is not None else w.tfin,
                                dtout=args.dtout if args.dtout is not None else w.dtout)
    t0 = _time.time()
    s = simulate(w, args.form, args.steps, not args.no_admesh, args.datadir, args.bin, args.vtk, args.warm,
                 args.refine_thres, args.coarse_thres, log=lambda *a: print(*a, flush=True))
    print("done: it %d, t = %.6e, %d blocks, %.1f s" % (s.it.value, s.time.value * w.t_sc, s.nblk,
                                                          _time.time() - t0), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
