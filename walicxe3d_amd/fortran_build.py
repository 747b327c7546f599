"""Build of the Fortran ISO_C_BINDING layer (walicxe3d_amd/fortran)."""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose=False):
    src = os.path.join(HERE, "fortran")
    if not any(f.endswith(".f90") for f in os.listdir(src)):
        raise FileNotFoundError("no Fortran sources in %s" % src)
    raise FileNotFoundError("Fortran layer build not available")
