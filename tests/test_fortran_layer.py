"""The Fortran ISO_C_BINDING layer: the bind(C) configuration type has the C
struct's layout (checked against the C compiler and the ctypes mirror)."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from conftest import ROOT
from walicxe3d_amd.hip import WxConfig

PROBE = os.path.join(ROOT, "walicxe3d_amd", "_build", "fortran", "wx_layout_probe")

C_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "walicxe_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(wx_config), offsetof(wx_config, gamma),
         offsetof(wx_config, ntab), offsetof(wx_config, tab_logT), offsetof(wx_config, winds),
         offsetof(wx_config, form));
  return 0;
}
"""


def c_layout(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    src = tmp_path / "probe.c"
    src.write_text(C_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]


def test_ctypes_mirror_matches_c(tmp_path):
    want = c_layout(tmp_path)
    got = [C.sizeof(WxConfig), WxConfig.gamma.offset, WxConfig.ntab.offset, WxConfig.tab_logT.offset,
           WxConfig.winds.offset, WxConfig.form.offset]
    assert got == want


def test_fortran_bind_c_type_matches_c(tmp_path):
    if not os.path.exists(PROBE):
        pytest.skip("Fortran layer not built (python -m walicxe3d_amd.build)")
    want = c_layout(tmp_path)
    out = subprocess.run([PROBE], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == want


def test_dropin_module_keeps_reference_interface():
    """The drop-in module exports the reference's hydro_solver interface."""
    src = open(os.path.join(ROOT, "walicxe3d_amd", "fortran", "hydro_solver_gpu.f90")).read().lower()
    assert "module hydro_solver" in src
    for name in ("subroutine hydrosolver ()", "subroutine gettimestep (dt_glob)", "subroutine doviscousstep ()"):
        assert name in src, name
