"""Profiling driver for rocprofv3 counter runs of the fused stage kernels.

Run mode (the program to put after `rocprofv3 ... --`):
    python3 -m walicxe3d_amd.profile_run run --workload ot256 --reps 2
uploads the workload, advances one step through the step graph, then
launches each stage kernel on its own `reps` times (so counters are
collected per dispatch) and finally copies a 1 GiB device buffer (a plain
copy kernel of known size, used to scale the memory counters).

Mesh-adaptation mode:
    python3 -m walicxe3d_amd.profile_run admesh --workload winds32k --reps 3
times (wall clock around each synchronous call) wx_mark_blocks with and
without a recorded ghost layer, wx_admesh_plan, and wx_admesh with a
coarsening threshold that merges part of the mesh, in the fused form, and
wx_mark_blocks in the plain form; under `rocprofv3 --kernel-trace --stats`
the per-kernel times of the same calls.

Sensitivity mode:
    python3 -m walicxe3d_amd.profile_run sensitivity --workload ot256 --steps 1000 --rel 1e-15
runs the plain FP64 form twice from the workload's initial state, the second
time with every conserved value scaled by (1 + rel), and prints the relative
L1 difference of the two states per variable after `steps` steps (the
measure bench.py applies to the fused FP32 form): how much the flow itself
amplifies a rounding-size difference.

Summary mode:
    python3 -m walicxe3d_amd.profile_run summarize --fetch DIR --write DIR \
        --workload ot256 --out profiles/stage2_ot256_traffic.txt [--cmd "..."]
reads the two counter_collection.csv files (FETCH_SIZE and WRITE_SIZE, KiB
per dispatch), takes the copy kernel's readings as the scale (bytes moved /
counter), and writes per-launch HBM traffic of the stage kernels next to
their algorithmic bytes.

Counter table mode:
    python3 -m walicxe3d_amd.profile_run counters --dirs DIR [DIR ...] \
        --workload ot256 --out profiles/valu_mix_ot256.txt [--cmd "..."] [--rc "..."]
the mean of every counter per dispatch of each stage kernel, over the
counter_collection.csv files under the given directories, and the same per
zone (x 64 lanes / zones: the instructions each lane issues per cell).

Kernel-trace mode:
    python3 -m walicxe3d_amd.profile_run stats --csv FILE --out profiles/... [--cmd "..."] [--rc "..."]
formats rocprofv3's --stats kernel summary (durations in milliseconds).
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import os
import sys

COPY_BYTES = 1 << 30

if __package__ in (None, ""):   # run as a script: make the package importable
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def _run(args):
    import numpy as np
    import torch

    from walicxe3d_amd import hip
    from walicxe3d_amd import workloads as W

    w = W.get(args.workload)
    slots, blocks, U0 = W.initial_state(w)
    g = hip.Walicxe3D(w, form=hip.FUSED_FP32, bids=np.asarray(blocks, dtype=np.int32))
    g.upload(U0)
    g.run(1)
    g.time_stages(args.reps)
    g.close()
    x = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print("profile_run done: %s, %d stage launches each, %d-byte copy" % (w.name, args.reps, COPY_BYTES),
          flush=True)
    return 0


def _admesh(args):
    import json
    import time

    import numpy as np

    from walicxe3d_amd import hip
    from walicxe3d_amd import workloads as W

    w = W.get(args.workload)
    m = w.mesh()
    slots, blocks, U0 = W.initial_state(w)
    b0 = np.asarray(blocks, dtype=np.int32)
    out = dict(workload=w.name, blocks=len(b0), reps=args.reps)

    def timed(fn):
        best = None
        for _ in range(args.reps):
            t = time.perf_counter()
            r = fn()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        return r, 1e3 * best

    g = hip.Walicxe3D(w, form=hip.FUSED_FP32, bids=b0)
    g.upload(U0)
    g.run(1)
    (f, grad), out["fused_mark_exchange_ms"] = timed(lambda: g.markByPhysical())
    _, out["fused_record_ghosts_ms"] = timed(g.record_ghosts)
    g.run(1)
    (f, grad), out["fused_mark_recorded_ms"] = timed(lambda: g.markByPhysical())
    (plan, _), out["plan_ms"] = timed(lambda: hip.admesh_plan(m, g.bids, f))
    out["plan_blocks_default_thresholds"] = int(len(plan))
    # a coarsening threshold at the 40th percentile of the block gradients,
    # no refinement: part of the families merge
    ct = float(np.percentile(grad, 40))
    t = time.perf_counter()
    new = g.admesh(1e300, ct)
    out["fused_admesh_ms"] = 1e3 * (time.perf_counter() - t)
    out["admesh_coarse_thres"] = ct
    out["admesh_blocks_after"] = int(len(new))
    t = time.perf_counter()
    g.run(1)
    out["fused_first_step_after_ms"] = 1e3 * (time.perf_counter() - t)
    g.close()
    p = hip.Walicxe3D(w, form=hip.PLAIN_FP64, bids=b0)
    p.upload(U0)
    p.step()
    _, out["plain_mark_ms"] = timed(lambda: p.markByPhysical())
    p.close()
    print("profile_run admesh " + json.dumps(out), flush=True)
    return 0


def _read(d):
    paths = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not paths:
        raise FileNotFoundError("no counter_collection.csv under %s" % d)
    rows = []
    for p in paths:
        rows += list(csv.DictReader(open(p)))
    return rows


def _per_kernel(rows, counter):
    acc = collections.defaultdict(list)
    for r in rows:
        if r["Counter_Name"] == counter:
            acc[r["Kernel_Name"]].append((int(r["Grid_Size"]), int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    return acc


def _copy_reading(acc):
    """Counter value of the 1 GiB device copy that ends the run: the last of
    the copy dispatches with the largest grid (the state upload is a copy of
    the same grid, of another size)."""
    best = None
    for k, v in acc.items():
        if "copyBuffer" in k:
            for grid, disp, val in v:
                if best is None or (grid, disp) > best[:2]:
                    best = (grid, disp, val)
    if best is None:
        raise RuntimeError("copy kernel not found in the counter data")
    return best[2]


def _summarize(args):
    from walicxe3d_amd import workloads as W

    w = W.get(args.workload)
    zones = len(w.blocks()) * 4096
    fetch = _per_kernel(_read(args.fetch), "FETCH_SIZE")
    write = _per_kernel(_read(args.write), "WRITE_SIZE")
    fscale = COPY_BYTES / (1024.0 * _copy_reading(fetch))
    wscale = COPY_BYTES / (1024.0 * _copy_reading(write))
    alg = {"stage1": zones * w.nv * (8 + 4), "stage2": zones * w.nv * (4 + 8 + 8)}
    lines = ["# HBM traffic of the fused stage kernels, %s (%d blocks, %d zones, nv=%d)" %
             (w.name, len(w.blocks()), zones, w.nv),
             "# command: %s" % (args.cmd or "(see walicxe3d_amd/profile_run.py)"),
             "# rocprofv3 exit status: %s" % (args.rc or "not recorded"),
             "# counters: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), one rocprofv3 --pmc run each",
             "# scale: the same counters on a %d-byte device copy (copyBuffer): bytes per counter byte" % COPY_BYTES,
             "fetch_scale = %.4f" % fscale,
             "write_scale = %.4f" % wscale]
    out = {}
    for k in fetch:
        # the second-stage flag: first template argument of k_mhd, second of k_stage
        st = _stage_of(k)
        if st is None:
            continue
        f = sum(v for _, _, v in fetch[k]) / len(fetch[k])
        wr = sum(v for _, _, v in write.get(k, [(0, 0, 0.0)])) / max(1, len(write.get(k, [])))
        hbm = 1024.0 * (f * fscale + wr * wscale)
        out[st] = (k, f, wr, hbm)
    for st in ("stage1", "stage2"):
        if st not in out:
            continue
        k, f, wr, hbm = out[st]
        lines += ["", "[%s] %s" % (st, k),
                  "%s_fetch_kib = %.1f" % (st, f),
                  "%s_write_kib = %.1f" % (st, wr),
                  "%s_hbm_bytes_per_launch = %.4e" % (st, hbm),
                  "%s_algorithmic_bytes_per_launch = %.4e" % (st, alg[st]),
                  "%s_traffic_over_algorithmic = %.3f" % (st, hbm / alg[st])]
    if "stage2" in out:
        lines += ["", "hbm_bytes_per_launch = %.4e   (stage 2, the dominant kernel)" % out["stage2"][3]]
    text = "\n".join(lines) + "\n"
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    sys.stdout.write(text)
    return 0


def _header(args, title):
    lines = ["# " + title]
    if args.cmd:
        lines.append("# command: " + args.cmd)
    if args.rc:
        lines.append("# rocprofv3 exit status: " + args.rc)
    return lines


def _stage_of(k):
    """'stage1' / 'stage2' for a fused stage kernel name, else None."""
    if "k_stage" not in k and "k_mhd" not in k:
        return None
    targs = k.split("<", 1)[1].split(",")
    return "stage2" if "true" in targs[0 if "k_mhd" in k else 1] else "stage1"


def _counters(args):
    from walicxe3d_amd import workloads as W

    w = W.get(args.workload)
    zones = len(w.blocks()) * 4096
    rows = []
    for d in args.dirs:
        rows += _read(d)
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in rows:
        st = _stage_of(r["Kernel_Name"])
        if st:
            per[(st, r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    lines = _header(args, "Counters of the fused stage kernels, %s (%d blocks, %d zones, nv=%d), mean per dispatch"
                    % (w.name, len(w.blocks()), zones, w.nv))
    lines.append("# per zone = counter x 64 lanes / zones (instructions each lane of a wave issues per cell)")
    for (st, k) in sorted(per):
        lines += ["", "[%s] %s" % (st, k)]
        for cn in sorted(per[(st, k)]):
            v = per[(st, k)][cn]
            m = sum(v) / len(v)
            lines.append("  %-28s %12.4e   per zone %9.1f" % (cn, m, m * 64.0 / zones))
    text = "\n".join(lines) + "\n"
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    sys.stdout.write(text)
    return 0


def _stats(args):
    rows = list(csv.DictReader(open(args.csv)))
    lines = _header(args, "rocprofv3 kernel summary (%s); durations in milliseconds" % os.path.basename(args.csv))
    lines.append("%-80s %6s %10s %9s %9s %9s %6s" % ("kernel", "calls", "total_ms", "avg_ms", "min_ms", "max_ms",
                                                   "pct"))
    for r in rows:
        name = r["Name"][:80]
        lines.append("%-80s %6s %10.2f %9.3f %9.3f %9.3f %6.2f" % (
            name, r["Calls"], float(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e6,
            float(r["MinNs"]) / 1e6, float(r["MaxNs"]) / 1e6, float(r["Percentage"])))
    text = "\n".join(lines) + "\n"
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    sys.stdout.write(text)
    return 0


def _sensitivity(args):
    import json

    import numpy as np

    from walicxe3d_amd import hip
    from walicxe3d_amd import workloads as W

    w = W.get(args.workload)
    slots, blocks, U0 = W.initial_state(w)
    bids = np.asarray(blocks, dtype=np.int32)
    out = []
    for scale in (1.0, 1.0 + args.rel):
        g = hip.Walicxe3D(w, form=hip.PLAIN_FP64, bids=bids)
        g.upload(U0 * scale)
        for n in range(args.steps):
            g.step()
            if (n + 1) % 100 == 0:
                print("  scale %.17g: step %d" % (scale, n + 1), flush=True)
        out.append(g.download(prim=False))
        g.close()
    d = np.abs(out[1] - out[0]).sum(axis=(0, 2, 3, 4))
    s = np.abs(out[0]).sum(axis=(0, 2, 3, 4))
    rel = np.where(s > 0, d / np.maximum(s, 1e-300), d)
    names = ["rho", "mx", "my", "mz", "E", "bx", "by", "bz"][: w.nv]
    print("profile_run sensitivity " + json.dumps({
        "workload": w.name, "steps": args.steps, "initial_rel_perturbation": args.rel,
        "rel_l1_all_conserved": float("%.3e" % (d.sum() / s.sum())),
        "rel_l1": {n: float("%.3e" % v) for n, v in zip(names, rel)}}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    r = sub.add_parser("run")
    r.add_argument("--workload", default="ot256")
    r.add_argument("--reps", type=int, default=2)
    a = sub.add_parser("admesh")
    a.add_argument("--workload", default="winds32k")
    a.add_argument("--reps", type=int, default=3)
    t = sub.add_parser("sensitivity")
    t.add_argument("--workload", default="ot256")
    t.add_argument("--steps", type=int, default=1000)
    t.add_argument("--rel", type=float, default=1e-15)
    s = sub.add_parser("summarize")
    s.add_argument("--fetch", required=True)
    s.add_argument("--write", required=True)
    s.add_argument("--workload", default="ot256")
    s.add_argument("--out")
    s.add_argument("--cmd")
    s.add_argument("--rc")
    c = sub.add_parser("counters")
    c.add_argument("--dirs", nargs="+", required=True)
    c.add_argument("--workload", default="ot256")
    c.add_argument("--out")
    c.add_argument("--cmd")
    c.add_argument("--rc")
    k = sub.add_parser("stats")
    k.add_argument("--csv", required=True)
    k.add_argument("--out")
    k.add_argument("--cmd")
    k.add_argument("--rc")
    args = ap.parse_args(argv)
    if args.mode == "counters":
        return _counters(args)
    if args.mode == "stats":
        return _stats(args)
    if args.mode == "admesh":
        return _admesh(args)
    if args.mode == "sensitivity":
        return _sensitivity(args)
    return _run(args) if args.mode == "run" else _summarize(args)


if __name__ == "__main__":
    sys.exit(main())
