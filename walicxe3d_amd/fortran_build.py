"""Build of the Fortran ISO_C_BINDING layer (walicxe3d_amd/fortran) with
amdflang: the binding module walicxe_hip and a layout probe.  The drop-in
module hydro_solver_gpu.f90 depends on the host Walicxe3D build's own
`parameters`/`globals` modules and is compiled inside that build, in place
of source/hydro_solver.f90."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fortran")
OUT = os.path.join(HERE, "_build", "fortran")
FC = os.environ.get("WX_FC", "amdflang")


def build(verbose=False):
    if shutil.which(FC) is None:
        raise FileNotFoundError("Fortran compiler %s not found" % FC)
    os.makedirs(OUT, exist_ok=True)
    obj = os.path.join(OUT, "walicxe_hip.o")
    cmds = [[FC, "-c", os.path.join(SRC, "walicxe_hip.f90"), "-J", OUT, "-o", obj],
            [FC, os.path.join(SRC, "wx_layout_probe.f90"), obj, "-I", OUT, "-J", OUT,
             "-o", os.path.join(OUT, "wx_layout_probe"), "-L" + HERE, "-lwalicxe_hip",
             "-Wl,-rpath,$ORIGIN/../.."]]
    for c in cmds:
        if verbose:
            print(" ".join(c))
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("Fortran build failed: %s\n%s%s" % (" ".join(c), r.stdout, r.stderr))
    return OUT
