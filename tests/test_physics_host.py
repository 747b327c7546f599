"""Host-side checks of the device physics header (walicxe3d_amd/csrc/
wx_physics.h), compiled for the CPU with hipcc: the select-based HLLD used by
the fused kernels agrees bit for bit with primfhlld (the reference's branch
structure, source/hlld.f90:166-444) in FP64 and FP32 over random states that
reach every region."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_hlld_select_matches_branches(tmp_path):
    exe = str(tmp_path / "hlld_select")
    cmd = [HIPCC, "-x", "hip", "-O2", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I" + os.path.join(ROOT, "walicxe3d_amd", "csrc"), "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "host", "hlld_select.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "double mismatches 0, float mismatches 0" in r.stdout
    # every region of the solver was reached
    for reg in ("left", "right", "UL*", "UR*", "UL**", "UR**"):
        assert ("%s 0 " % reg) not in r.stdout, r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_delta_form_solvers(tmp_path):
    """The delta-form HLLD and HLLE of the FP32 MHD kernels (wx_delta.h):
    exact algebra against the reference form in FP64, and FP32 error on
    small jumps two orders below the plain FP32 solver's."""
    exe = str(tmp_path / "hlld_delta")
    cmd = [HIPCC, "-x", "hip", "-O2", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I" + os.path.join(ROOT, "walicxe3d_amd", "csrc"), "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "host", "hlld_delta.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "delta form ok" in r.stdout, r.stdout + r.stderr
