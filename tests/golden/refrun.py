"""Run the reference oracle binary (oracle/_ref/<workload>/harness) on a
workload's initial condition and read its dumps back.  Used only by
tests/golden/make_golden.py in the build container, where the reference
sources exist; the GPU tests read the committed fixtures instead."""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np

from walicxe3d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REFDIR = os.path.join(ROOT, "oracle", "_ref")
MPIEXEC = "/opt/conda/bin/mpiexec"


def harness_path(name: str) -> str:
    return os.path.join(REFDIR, name, "harness")


def neq_ref(w):
    """neqtot of the reference builds: the flow variables and one passive
    scalar (a workload's own tracer, or zero everywhere)."""
    assert w.npassive <= 1, "the reference builds carry one passive scalar"
    return (8 if w.mhd else 5) + 1


def write_ic_with_passive(path, w, blocks, U):
    """The oracle builds carry one passive scalar: the workload's tracer, or
    zero everywhere."""
    nb = U.shape[0]
    if w.npassive:
        U1 = U
    else:
        U1 = np.concatenate([U, np.zeros((nb, 1) + U.shape[2:])], axis=1)
    refine = w.refine_order()
    wp = W.wind_params(w) + W.plane_wind_record(w) + W.gravity_record(w)
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", W.IC_MAGIC, neq_ref(w), len(refine), nb, len(wp)))
        if refine:
            f.write(np.asarray(refine, dtype="<i4").tobytes())
        for rec in wp:
            f.write(np.asarray(rec, dtype="<f8").tobytes())
        for n, b in enumerate(blocks):
            f.write(struct.pack("<i", b))
            f.write(np.ascontiguousarray(U1[n], dtype="<f8").tobytes())


def read_topology(path):
    with open(path, "rb") as f:
        nact, maxlev, nbx, nby, nbz = struct.unpack("<5i", f.read(20))
        rec = np.frombuffer(f.read(8 * nact), dtype="<i4").reshape(nact, 2)
        dx = np.frombuffer(f.read(8 * 3 * maxlev), dtype="<f8").reshape(3, maxlev)
        consts = np.frombuffer(f.read(80), dtype="<f8")
    return dict(bids=rec[:, 0].copy(), levels=rec[:, 1].copy(), maxlev=maxlev,
                nroot=(nbx, nby, nbz), dx=dx.copy(),
                consts=dict(zip(["gamma", "cv", "cfl", "visc", "d_sc", "e_sc", "t_sc", "p_sc",
                                 "l_sc", "v_sc"], consts.tolist())))


def read_state(path, nblk, nv_ref):
    a = np.fromfile(path, dtype="<f8")
    return a.reshape(nblk, nv_ref, 16, 16, 16)


def read_ints(path):
    a = np.fromfile(path, dtype="<i4")
    return a[1:1 + int(a[0])].copy()


def run(w, nsteps, dumps, workdir=None, unit_op=None, unit_in=None, timeout=3600, tstop=-1.0, variant="",
        nprocs=1, amr=False, bindump=False, vtkdump=False, env=None):
    """Run the harness; returns dict with topo, steps (it, dt, time, tout),
    and states {step: (U, PRIM)} with the passive scalar stripped.  With
    nprocs > 1 the multi-rank build runs under mpiexec and only the step log
    comes back (no topology, no states).  With amr the adaptive-mesh build
    runs admesh after every step; then also flags {step: (bids, flags)}
    (markByPhysical before admesh) and blocks {step: bids after admesh, in
    the order of that step's state dump}.  With bindump the reference's
    own checkpoint files of each dumped step come back as
    files {step: (Blocks000.NNNN.bin bytes, State.NNNN.dat text)}; with
    vtkdump vtk {step: (Blocks000.NNNN.vtk bytes, Grid.NNNN.vtk bytes)} and
    visit (master.visit text at the end).  env: extra environment variables
    of the harness process (e.g. WALICXE_GPU_FORM for the drop-in builds)."""
    if nprocs > 1:
        variant = variant + "_np%d" % nprocs
    if amr:
        variant = variant + "_amr"
    exe = harness_path(w.name + variant)
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    own = workdir is None
    wd = workdir or tempfile.mkdtemp(prefix="wxref_")
    if w.cooling:
        lt, lam = W.cooling_table()
        W.write_cooling_table(os.path.join(wd, "cooltable.dat"), lt, lam)
    if unit_op is not None:
        if unit_op == 12:   # the mesh routines need the grid set up from ic.bin
            slots, blocks, U = W.initial_state(w)
            write_ic_with_passive(os.path.join(wd, "ic.bin"), w, blocks, U)
        with open(os.path.join(wd, "harness.nml"), "w") as f:
            f.write("&ctl\n mode='unit'\n op=%d\n/\n" % unit_op)
        with open(os.path.join(wd, "unit_in.bin"), "wb") as f:
            f.write(struct.pack("<i", unit_in.shape[0]))
            f.write(np.ascontiguousarray(unit_in, dtype="<f8").tobytes())
        subprocess.run([exe], cwd=wd, check=True, timeout=timeout, capture_output=True)
        return np.fromfile(os.path.join(wd, "unit_out.bin"), dtype="<f8")
    slots, blocks, U = W.initial_state(w)
    write_ic_with_passive(os.path.join(wd, "ic.bin"), w, blocks, U)
    dl = ",".join(str(d) for d in dumps)
    with open(os.path.join(wd, "harness.nml"), "w") as f:
        f.write("&ctl\n mode='run'\n nsteps=%d\n ndump=%d\n dumps=%s\n tstop=%r\n%s/\n"
                % (nsteps, len(dumps), dl or "-1", float(tstop),
                   (" amr=.true.\n" if amr else "") + (" bindump=.true.\n" if bindump else "")
                   + (" vtkdump=.true.\n" if vtkdump else "")))
    cmd = [exe] if nprocs == 1 else [MPIEXEC, "-n", str(nprocs), exe]
    r = subprocess.run(cmd, cwd=wd, timeout=timeout, capture_output=True, text=True,
                       env=dict(os.environ, **env) if env else None)
    if r.returncode != 0:
        raise RuntimeError("reference harness failed (%d): %s %s" % (r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    if nprocs > 1:
        steps = np.loadtxt(os.path.join(wd, "steps.txt"), ndmin=2)
        if own:
            import shutil
            shutil.rmtree(wd, ignore_errors=True)
        return dict(steps=steps)
    topo = read_topology(os.path.join(wd, "topo.bin"))
    steps = np.loadtxt(os.path.join(wd, "steps.txt"), ndmin=2)
    nblk = len(topo["bids"])
    states = {}
    if tstop > 0 and len(steps) and int(steps[-1, 0]) not in dumps:
        dumps = list(dumps) + [int(steps[-1, 0])]
    flags, blocks_after = {}, {}
    if amr:
        for s in range(1, len(steps) + 1):
            fl = np.fromfile(os.path.join(wd, "flags_%05d.bin" % s), dtype="<i4")[1:].reshape(-1, 2)
            flags[s] = (fl[:, 0].copy(), fl[:, 1].copy())
            blocks_after[s] = read_ints(os.path.join(wd, "blocks_%05d.bin" % s))
    for d in dumps:
        nb = len(blocks_after[d]) if (amr and d > 0) else nblk
        u = read_state(os.path.join(wd, "u_%05d.bin" % d), nb, neq_ref(w))[:, : w.nv]
        p = read_state(os.path.join(wd, "prim_%05d.bin" % d), nb, neq_ref(w))[:, : w.nv]
        states[d] = (u.copy(), p.copy())
    out = dict(topo=topo, steps=steps, states=states, ic_blocks=blocks, ic_U=U, dumps=list(dumps))
    if bindump:
        files = {}
        for d in dumps:
            with open(os.path.join(wd, "Blocks000.%04d.bin" % d), "rb") as f:
                b = f.read()
            with open(os.path.join(wd, "State.%04d.dat" % d)) as f:
                files[d] = (b, f.read())
        out["files"] = files
    if vtkdump:
        vtk = {}
        for d in dumps:
            with open(os.path.join(wd, "Blocks000.%04d.vtk" % d), "rb") as f:
                b = f.read()
            with open(os.path.join(wd, "Grid.%04d.vtk" % d), "rb") as f:
                vtk[d] = (b, f.read())
        out["vtk"] = vtk
        with open(os.path.join(wd, "master.visit")) as f:
            out["visit"] = f.read()
    if amr:
        out.update(flags=flags, blocks=blocks_after)
    if own:
        import shutil
        shutil.rmtree(wd, ignore_errors=True)
    return out
