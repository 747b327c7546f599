#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the fused FP32 form against the plain FP64
form on the same GPU, with the L1 error of every variable against FP64, the
HBM roofline fraction of the dominant kernel and a CPU baseline.

A step is one full timestep (two stages, viscosity, primitives, cooling, dt)
of the whole block list; a zone-update is one interior cell advanced one
step.  Default workload: Orszag-Tang vortex 256^3 (4096 blocks of 16^3,
HLLD, periodic).  Inputs are resident in HBM when the timed region starts.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload ot256]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s/GPU (batched FP32 vs plain FP64, same GPU) at stated L1 err; % HBM BW"
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(workload):
    """HBM bytes per stage-2 launch measured with rocprofv3 PMC counters
    (summary committed under profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "stage2_%s_traffic.txt" % workload)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except ValueError:
                pass
    return vals.get("hbm_bytes_per_launch")


def cpu_baseline(budget_s=20.0):
    """The reference itself (oracle/_ref: the Walicxe3D sources compiled with
    amdflang -O2 and MPICH) on the 64^3 Orszag-Tang workload, run with
    mpiexec over 16 host cores (the host share of one GPU) when that build is
    present, else on one rank; start-up subtracted.  Falls back to timing the
    C restatement when no reference build exists."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from walicxe3d_amd import workloads as W
    w = W.get("ot64")
    zones_per_step = len(w.blocks()) * 4096
    try:
        import refrun
        for nprocs, nsteps in ((16, 60), (1, 6)):
            exe = refrun.harness_path(w.name + ("_np%d" % nprocs if nprocs > 1 else ""))
            if os.path.exists(exe) and (nprocs == 1 or os.path.exists(refrun.MPIEXEC)):
                break
        else:
            raise FileNotFoundError(refrun.harness_path(w.name))
        t0 = time.time()
        refrun.run(w, 1, [], nprocs=nprocs)          # start-up cost (MPI, init, IC read)
        t_init = time.time() - t0
        t0 = time.time()
        refrun.run(w, nsteps + 1, [], nprocs=nprocs)
        t = time.time() - t0 - t_init
        return {"value": zones_per_step * nsteps / t, "unit": "zone-updates/s", "cores": nprocs,
                "kind": "reference",
                "sample": "Walicxe3D reference (oracle/_ref, amdflang -O2, MPICH, %d MPI rank%s) on Orszag-Tang "
                          "64^3 (64 blocks, HLLD, periodic), %d steps, start-up subtracted"
                          % (nprocs, "s" if nprocs > 1 else "", nsteps)}
    except Exception as e:  # reference binary not available: time the C restatement
        log("reference binary unavailable (%s); timing the C restatement" % e)
        from oracle import wxo
        slots, blocks, U0 = W.initial_state(w)
        s = wxo.Sim(w, blocks)
        s.set_state(U0)
        t0 = time.time()
        n = 0
        while time.time() - t0 < budget_s and n < 40:
            s.step()
            n += 1
        t = time.time() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        return {"value": len(blocks) * 4096 * n / t, "unit": "zone-updates/s", "cores": threads, "kind": "port",
                "sample": "C restatement (OpenMP) on Orszag-Tang 64^3, %d steps" % n}


def dropin_rate(name="ot64", nsteps=20):
    """The reference's own driver and modules with the Fortran drop-in of
    module hydro_solver in its fused FP32 form (oracle/_ref/<name>_gpu, the
    harness of the drop-in GPU test; WALICXE_GPU_FORM=fused): zone-updates/s
    of whole main-loop iterations (hydroSolver on the device, U over PCIe each
    step, updatePrims and cooling on the host), start-up subtracted.  None
    when that build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    try:
        import refrun
        from walicxe3d_amd import workloads as W
        w = W.get(name)
        if not os.path.exists(refrun.harness_path(name + "_gpu")):
            return None
        env = {"WALICXE_GPU_FORM": "fused"}
        t0 = time.time()
        refrun.run(w, 1, [], variant="_gpu", env=env, timeout=300)
        t_one = time.time() - t0
        t0 = time.time()
        refrun.run(w, nsteps + 1, [], variant="_gpu", env=env, timeout=300)
        t = time.time() - t0 - t_one
        zones = len(w.blocks()) * 4096
        return {"value": zones * nsteps / t, "unit": "zone-updates/s",
                "sample": "Walicxe3D driver + Fortran drop-in (fused FP32 form) on %s, %d steps, U over PCIe "
                          "every step, updatePrims/cooling on the host, start-up subtracted" % (name, nsteps)}
    except Exception as e:  # the drop-in is reported, never required
        log("drop-in timing skipped: %s" % e)
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ot256")
    ap.add_argument("--err-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch autotuner (profiling runs)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the comparison forms (fused FP64, plain-FP32 MHD fluxes, Fortran drop-in)")
    ap.add_argument("--remesh-every", type=int, default=0,
                    help="AMR workloads: change the block list every N steps (device remesh, graph recapture), "
                         "alternating with a 6%% wider refinement")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the block list of this path lives on a single device "
              "(multi-GPU sharding is not built); run with --gpus 1", file=sys.stderr)
        return 2

    from walicxe3d_amd import hip
    from walicxe3d_amd import workloads as W

    w = W.get(args.workload)
    t0 = time.time()
    slots, blocks, U0 = W.initial_state(w)
    bids = np.asarray(blocks, dtype=np.int32)
    nblk = len(bids)
    zones = nblk * 4096
    log("workload %s: %d blocks, %d zones, IC %.1fs" % (w.name, nblk, zones, time.time() - t0))

    # ---- fused FP32 form ----------------------------------------------------
    f = hip.Walicxe3D(w, form=hip.FUSED_FP32, bids=bids)
    f.upload(U0)
    best = {"grid": 0, "xcd_map": 0, "curve": 0, "nslab": 1}
    if args.no_tune:
        f.set_tuning(0, best["grid"])
        f.set_tuning(1, best["xcd_map"])
        f.set_tuning(2, best["curve"])
        f.set_tuning(4, best["nslab"])
    else:
        from walicxe3d_amd import autotune
        best, table = autotune.tune(f, steps=3)
        log("autotune: %s" % ", ".join("grid %d xcd %d curve %d slabs %d %.3f ms" % (
            r["grid"], r["xcd_map"], r["curve"], r["nslab"], r["ms"]) for r in table))
    f.run(args.warmup)
    remesh = None
    if args.remesh_every > 0:
        import dataclasses
        wB = dataclasses.replace(w, refine_radii=tuple(r * 1.06 for r in w.refine_radii))
        meshes = [bids, np.asarray(wB.blocks(), dtype=np.int32)]
        # one untimed round trip so both meshes' allocations and graphs have been built once
        f.remesh(meshes[1])
        f.run(1)
        f.remesh(meshes[0])
        f.run(1)
        cur, done, zsum, t_rm, n_rm = 0, 0, 0, 0.0, 0
        t0 = time.perf_counter()
        while done < args.steps:
            n = min(args.remesh_every, args.steps - done)
            f.run(n)
            zsum += len(meshes[cur]) * 4096 * n
            done += n
            if done < args.steps:
                tr = time.perf_counter()
                cur = 1 - cur
                f.remesh(meshes[cur])
                t_rm += time.perf_counter() - tr
                n_rm += 1
        t_fused = time.perf_counter() - t0
        fused_zps = zsum / t_fused
        remesh = {"every": args.remesh_every, "count": n_rm, "ms_per_remesh": 1e3 * t_rm / max(n_rm, 1),
                  "blocks": [len(m_) for m_ in meshes],
                  "zone_updates_per_s_without_remesh": zsum / max(t_fused - t_rm, 1e-12)}
        log("remesh every %d steps: %d remeshes, %.1f ms each" % (args.remesh_every, n_rm, remesh["ms_per_remesh"]))
        if cur != 0:
            f.remesh(meshes[0])
    else:
        t0 = time.perf_counter()
        f.run(args.steps)                      # returns after a stream synchronize
        t_fused = time.perf_counter() - t0
        fused_zps = zones * args.steps / t_fused
    s1, s2 = f.time_stages(3)
    log("fused: %.3f ms/step, %.3e zone-updates/s; stage1 %.3f ms, stage2 %.3f ms"
        % (1e3 * t_fused / args.steps, fused_zps, s1, s2))
    f.close()

    # ---- plain FP64 form ----------------------------------------------------
    plain = None
    if not args.no_plain:
        p = hip.Walicxe3D(w, form=hip.PLAIN_FP64, bids=bids)
        p.upload(U0)
        for _ in range(args.warmup):
            p.step()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            p.step()
        t_plain = time.perf_counter() - t0
        plain_zps = zones * args.steps / t_plain
        log("plain: %.3f ms/step, %.3e zone-updates/s" % (1e3 * t_plain / args.steps, plain_zps))
        p.close()
        plain = {"value": plain_zps, "ms_per_step": 1e3 * t_plain / args.steps, "dtype": "f64",
                 "speedup_fused_over_plain": fused_zps / plain_zps}

    # ---- comparison forms on the same workload ------------------------------
    def timed(form, tuning=()):
        g = hip.Walicxe3D(w, form=form, bids=bids)
        for kind, value in tuning:
            g.set_tuning(kind, value)
        g.upload(U0)
        g.run(args.warmup)
        t0 = time.perf_counter()
        g.run(args.steps)
        t = time.perf_counter() - t0
        g.close()
        return zones * args.steps / t, 1e3 * t / args.steps

    fused64 = plain32 = None
    if not args.no_extra:
        v, ms = timed(hip.FUSED_FP64, [(0, best["grid"]), (1, best["xcd_map"]), (2, best["curve"])])
        fused64 = {"value": v, "ms_per_step": ms, "dtype": "f64",
                   "speedup_fused_fp32_over_fused_fp64": fused_zps / v}
        log("fused FP64: %.3f ms/step, %.3e zone-updates/s" % (ms, v))
        if w.nv == 8:
            v, ms = timed(hip.FUSED_FP32, [(0, best["grid"]), (1, best["xcd_map"]), (2, best["curve"]), (3, 1),
                                           (4, best["nslab"])])
            plain32 = {"value": v, "ms_per_step": ms,
                       "form": "fused FP32 with every flux term in FP32 (tuning kind 3 = 1)"}
            log("fused FP32, plain FP32 fluxes: %.3f ms/step, %.3e zone-updates/s" % (ms, v))

    # ---- L1 error of the fused form against FP64 after err-steps steps -----
    err = None
    if not args.no_plain:
        a = hip.Walicxe3D(w, form=hip.FUSED_FP32, bids=bids)
        a.upload(U0)
        a.run(args.err_steps)
        ua = a.download(prim=False)
        a.close()
        b = hip.Walicxe3D(w, form=hip.PLAIN_FP64, bids=bids)
        b.upload(U0)
        for _ in range(args.err_steps):
            b.step()
        ub = b.download(prim=False)
        b.close()
        d = np.abs(ua - ub).sum(axis=(0, 2, 3, 4))
        s = np.abs(ub).sum(axis=(0, 2, 3, 4))
        rel = np.where(s > 0, d / np.maximum(s, 1e-300), d)
        names = ["rho", "mx", "my", "mz", "E", "bx", "by", "bz"][: w.nv]
        err = {"steps": args.err_steps, "reference": "plain FP64 GPU form (bit-exact with the FP64 oracle)",
               "rel_l1_all_conserved": float("%.3e" % (d.sum() / s.sum())),
               "rel_l1": {n: float("%.3e" % v) for n, v in zip(names, rel)}, "max": float("%.3e" % rel.max())}
        log("fused rel L1 vs FP64 after %d steps: %s" % (args.err_steps, err["rel_l1"]))
        if plain32 is not None:
            c = hip.Walicxe3D(w, form=hip.FUSED_FP32, bids=bids)
            c.set_tuning(3, 1)
            c.upload(U0)
            c.run(args.err_steps)
            uc = c.download(prim=False)
            c.close()
            dc = np.abs(uc - ub).sum(axis=(0, 2, 3, 4))
            relc = np.where(s > 0, dc / np.maximum(s, 1e-300), dc)
            plain32["rel_l1"] = {n: float("%.3e" % v) for n, v in zip(names, relc)}
            plain32["rel_l1_max"] = float("%.3e" % relc.max())
            log("plain-FP32-flux rel L1 after %d steps: %s" % (args.err_steps, plain32["rel_l1"]))

    # ---- roofline of the dominant kernel (stage 2) ---------------------------
    nv = w.nv
    b_stage2 = zones * nv * (4 + 8 + 8)      # read UP (f32) + U (f64), write U (f64)
    b_stage1 = zones * nv * (8 + 4)          # read U (f64), write UP (f32)
    dom_ms, dom_bytes, dom = (s2, b_stage2, "stage2") if s2 >= s1 else (s1, b_stage1, "stage1")
    achieved = dom_bytes / (dom_ms * 1e-3)
    kname = ("k_mhd<%s>" % ("S2" if dom == "stage2" else "S1")) if nv == 8 else \
        ("k_stage<%d,%s>" % (nv, "S2" if dom == "stage2" else "S1"))
    roof = {"bound": "hbm", "kernel": kname,
            "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": traffic_from_profiles(w.name), "algorithmic_bytes_per_launch": dom_bytes,
            "avg_launch_ms": dom_ms}
    step_bytes = zones * nv * 32
    cpu = None if args.no_cpu else cpu_baseline()
    dropin = None if (args.no_extra or args.no_cpu) else dropin_rate()
    out = {
        "metric": METRIC,
        "value": fused_zps,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_fused / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s: %s" % (w.name, w.description), "blocks": nblk, "zones": zones,
                   "form": ("fused FP32: FP64 primitives, face states and physical fluxes, FP32 Riemann correction "
                            "(delta form), FP64 accumulation" if nv == 8 else
                            "fused FP32 bulk, FP64 conserved accumulation"), "solver": w.solver,
                   "limiter": w.limiter, "step_hbm_roofline_frac": (fused_zps * step_bytes / zones) / HBM_PEAK,
                   "tuning": {"grid": best["grid"], "xcd_map": best["xcd_map"],
                              "curve": "hilbert" if best["curve"] else "morton",
                              "z_slabs_per_block": best["nslab"]}, "remesh": remesh},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_fp64": plain,
        "fused_fp64": fused64,
        "fused_fp32_plain_flux": plain32,
        "dropin_fortran": dropin,
        "l1_err_vs_fp64": err,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
