#!/usr/bin/env python3
"""Build recipe for the reference oracle binaries (test infrastructure only).

Compiles the unmodified Walicxe3D Fortran sources where they lie under
/root/reference/source, together with this repository's per-problem
`parameters` module (oracle/ref/parameters.f90.in, filled per workload),
`userconds` module (oracle/ref/user.f90) and driver (oracle/ref/harness.f90),
into oracle/_ref/<workload>/harness.  Nothing is copied out of the reference
tree; objects, module files and binaries are written only under oracle/_ref/.

Toolchain: amdflang (ROCm) with the image's MPICH (mpif.h, libmpi,
libmpifort).  The reference's Makefile flags for a double-precision MPI build
(Makefile:101-128: -cpp -DMPIP -DDOUBLEP -fdefault-real-8 [-DBFIELD]) are used.

Usage:  python oracle/build_ref.py [workload ...]     (default: all)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = "/root/reference"
OUT = os.path.join(HERE, "_ref")
sys.path.insert(0, ROOT)

from walicxe3d_amd import workloads  # noqa: E402

FC = os.environ.get("WX_FC", "amdflang")
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib"

# Object order of the reference's Makefile (Makefile:49-84), with the
# problem-supplied modules (parameters, user) taken from this repository.
SOURCES = [
    ("ref", "constants.f90"), ("own", "parameters.f90"), ("ref", "globals.f90"),
    ("ref", "tictoc.f90"), ("ref", "clean_quit.f90"), ("ref", "hydro_core.f90"),
    ("ref", "utils.f90"), ("ref", "hilbert.f90"), ("ref", "amr.f90"),
    ("ref", "loadbalance.f90"), ("ref", "uniformISM.f90"), ("ref", "snr.f90"),
    ("ref", "winds.f90"), ("ref", "orbits.f90"), ("own", "user.f90"),
    ("ref", "hrate.f90"), ("ref", "cooling_h.f90"), ("ref", "cooling_schure.f90"),
    ("ref", "cooling.f90"), ("ref", "init.f90"), ("ref", "output.f90"),
    ("ref", "boundary.f90"), ("ref", "lax.f90"), ("ref", "hll.f90"),
    ("ref", "hllc.f90"), ("ref", "hlle.f90"), ("ref", "hlld.f90"),
    ("ref", "sources.f90"), ("ref", "godunov.f90"), ("ref", "hydro_solver.f90"),
    ("ref", "report.f90"), ("own", "harness.f90"),
]

BC = {"reflective": "BC_REFLECTIVE", "free": "BC_FREE_FLOW", "periodic": "BC_PERIODIC"}
SOLVER = {"hll1": "SOLVER_HLL1", "hll": "SOLVER_HLL", "hllc": "SOLVER_HLLC",
          "hlle": "SOLVER_HLLE", "hlld": "SOLVER_HLLD"}
LIMITER = {"noavg": "LIMITER_NO_AVERAGE", "none": "LIMITER_NONE", "minmod": "LIMITER_MINMOD",
           "vanleer": "LIMITER_VANLEER", "albada": "LIMITER_ALBADA", "umist": "LIMITER_UMIST",
           "woodward": "LIMITER_WOODWARD", "superbee": "LIMITER_SUPERBEE"}
EOS = {"adiabatic": "EOS_ADIABATIC", "single": "EOS_SINGLE_SPECIE", "twotemp": "EOS_TWOTEMP",
       "hrate": "EOS_H_RATE"}


def flit(x):
    """Fortran literal for a float (exact round trip through repr)."""
    if isinstance(x, str):
        return x
    s = repr(float(x))
    if "e" not in s and "." not in s:
        s += ".0"
    return s


def amr_blocks(w):
    """Block slots of a build whose mesh adapts every step: room for every
    block of the finest level (capped), as refineBlock aborts when the slots
    run out (source/amr.f90:318-324)."""
    finest = w.nroot[0] * w.nroot[1] * w.nroot[2] * 8 ** (w.maxlev - 1)
    return min(finest + 64, 6000)


def parameters_text(w, nprocs=1, amr=False):
    nblk_max = amr_blocks(w) if amr else w.max_blocks_ref()
    nv = (8 if w.mhd else 5) + 1
    block_ram = 20 ** 3 * nv * 8
    # RAM_per_proc such that nbMaxProc (source/parameters.f90:342-343) >= needed
    ram_mb = int(((nblk_max + 2) * block_ram * 3 / 0.95 + 6 * block_ram) / (1024 * 1024)) + 2
    bcs = w.bc
    sub = {
        "NPROCS": str(nprocs), "TFIN": flit(w.tfin), "DTOUT": flit(w.dtout), "RAM_MB": flit(float(ram_mb)),
        "XPHYS": flit(w.phys[0]), "YPHYS": flit(w.phys[1]), "ZPHYS": flit(w.phys[2]),
        "NBX": str(w.nroot[0]), "NBY": str(w.nroot[1]), "NBZ": str(w.nroot[2]),
        "MAXLEV": str(w.maxlev),
        "MAXCX": str(w.nroot[0] * 16 * 2 ** (w.maxlev - 1)),
        "MAXCY": str(w.nroot[1] * 16 * 2 ** (w.maxlev - 1)),
        "MAXCZ": str(w.nroot[2] * 16 * 2 ** (w.maxlev - 1)),
        "BC_LEFT": BC[bcs[0]], "BC_RIGHT": BC[bcs[1]], "BC_FRONT": BC[bcs[2]],
        "BC_BACK": BC[bcs[3]], "BC_BOTTOM": BC[bcs[4]], "BC_TOP": BC[bcs[5]],
        "MHD": ".true." if w.mhd else ".false.",
        "SOLVER": SOLVER[w.solver], "LIMITER": LIMITER[w.limiter],
        "CFL": flit(w.cfl), "VISC": flit(w.visc_eta),
        "EIGHTWAVE": ".true." if w.eight_wave else ".false.",
        "USERSRC": ".true." if w.gravity else ".false.",
        "COOLING": "COOL_TABLE" if w.cooling else "COOL_NONE",
        "COOLFILE": "cooltable.dat", "COOL_LIMIT": flit(w.cooling_limit),
        "GAMMA": w.gamma_expr, "MU0": flit(w.mu0), "MUI": flit(w.mui),
        "ION_THRES": flit(w.ion_thres), "EOS": EOS[w.eos],
        "L_SC": w.l_sc_expr, "D_SC": w.d_sc_expr, "V_SC": w.v_sc_expr,
        "UNITS": "PHYS_UNITS" if w.units == "phys" else "CODE_UNITS",
    }
    with open(os.path.join(HERE, "ref", "parameters.f90.in")) as f:
        text = f.read()
    for k, v in sub.items():
        text = text.replace("@%s@" % k, v)
    import re
    left = re.findall(r"@[A-Z_0-9]+@", text)
    assert not left, "unfilled tokens %s" % left
    return text


FORTRAN_LAYER = os.path.join(ROOT, "walicxe3d_amd", "fortran")
GPU_LIBDIR = os.path.join(ROOT, "walicxe3d_amd")

# GPU drop-in variant: the same build with the reference's module
# `hydro_solver` replaced by this repository's GPU implementation
# (walicxe3d_amd/fortran/hydro_solver_gpu.f90) linked to libwalicxe_hip.so.
def gpu_sources():
    out = []
    for where, fn in SOURCES:
        if fn == "hydro_solver.f90":
            out.append(("gpu", "walicxe_hip.f90"))
            out.append(("gpu", "hydro_solver_gpu.f90"))
        else:
            out.append((where, fn))
    return out


def build(name, gpu=False, nprocs=1, amr=False):
    w = workloads.get(name)
    bdir = os.path.join(OUT, name + ("_gpu" if gpu else "") + ("_np%d" % nprocs if nprocs > 1 else "")
                        + ("_amr" if amr else ""))
    os.makedirs(bdir, exist_ok=True)
    with open(os.path.join(bdir, "parameters.f90"), "w") as f:
        f.write(parameters_text(w, nprocs, amr))
    for fn in ("user.f90", "harness.f90"):
        shutil.copyfile(os.path.join(HERE, "ref", fn), os.path.join(bdir, fn))
    flags = ["-cpp", "-DMPIP", "-DDOUBLEP", "-fdefault-real-8", os.environ.get("WX_FOPT", "-O2"),
             "-ffp-contract=off", "-I" + MPI_INC, "-J", bdir, "-w"]
    if w.mhd:
        flags.append("-DBFIELD")
    if gpu:
        flags.append("-DWX_DROPIN")
    objs = []
    for where, fn in (gpu_sources() if gpu else SOURCES):
        if where == "ref":
            src = os.path.join(REF, "source", fn)
        elif where == "gpu":
            src = os.path.join(FORTRAN_LAYER, fn)
        else:
            src = os.path.join(bdir, fn)
        obj = os.path.join(bdir, fn.replace(".f90", ".o"))
        cmd = [FC] + flags + ["-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
            raise SystemExit("compile failed: %s" % fn)
        objs.append(obj)
    exe = os.path.join(bdir, "harness")
    cmd = [FC] + objs + ["-o", exe, "-L" + MPI_LIB, "-lmpifort", "-lmpi",
                         "-Wl,-rpath," + MPI_LIB]
    if gpu:
        cmd += ["-L" + GPU_LIBDIR, "-lwalicxe_hip", "-Wl,-rpath,$ORIGIN/../../../walicxe3d_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit("link failed")
    return exe


def main(argv):
    if not os.path.isdir(os.path.join(REF, "source")):
        print("reference sources not present; nothing to build")
        return 0
    gpu = "--gpu" in argv
    argv = [a for a in argv if not a.startswith("--")]
    names = argv or workloads.oracle_builds()
    jobs = [(n, False, 1, False) for n in names]
    if gpu or not argv:
        # drop-in builds used by tests/test_gpu_fortran.py
        jobs += [(n, True, 1, False) for n in (argv or DROPIN_WORKLOADS)]
    if not argv or "--baseline" in sys.argv:
        # multi-rank builds for the CPU baseline of bench.py (mpiexec -n N)
        jobs += [(n, False, np_, False) for n, np_ in BASELINE_BUILDS]
    if not argv or "--amr" in sys.argv:
        # adaptive-mesh builds (admesh every step) for the AMR fixtures
        jobs += [(n, False, 1, True) for n in (argv or AMR_WORKLOADS)]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=4) as ex:
        for (n, g, np_, amr), exe in zip(jobs, ex.map(lambda j: build(*j), jobs)):
            print("building", n, "(gpu drop-in)" if g else "", "(%d ranks)" % np_ if np_ > 1 else "",
                  "(adaptive mesh)" if amr else "", "->", exe)
    return 0


DROPIN_WORKLOADS = ["sedov", "ot64", "amr_small", "blast_reflect", "plane_wind", "sedov_gravity", "sedov_passive",
                    "amr_small_passive"]
# (workload, MPI ranks): 16 ranks = the host cores one GPU of the box comes with
BASELINE_BUILDS = [("ot64", 16)]
# workloads whose fixtures include the reference's mesh adaptation
AMR_WORKLOADS = ["amr_adapt"]


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
