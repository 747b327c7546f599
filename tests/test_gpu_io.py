"""The reference's output files from the device state, byte for byte
against the files the reference itself wrote (tests/golden/io.npz: SHA-256
and size): wx_write_bin (writeBin, source/output.f90:156-214), read back by
wx_read_bin (warmstart, source/init.f90:582-610), and wx_write_vtk
(write3DVTKBlocks, write3DVTKGrid and master.visit, output.f90:273-895)."""
import hashlib
import sys

import numpy as np
import pytest

from conftest import ic_in_order, load_golden
from walicxe3d_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _digest(path):
    data = open(path, "rb").read()
    return len(data), hashlib.sha256(data).hexdigest()


# (workload, form, steps before writing): the written state must be the
# reference's bit for bit, so the fused FP32 form writes its initial state
CASES = [("ot64", "plain", 0), ("ot64", "plain", 3), ("ot64", "fused_fp32", 0), ("amr_small", "plain", 0),
         ("amr_small", "fused_fp32", 0)]


@pytest.mark.parametrize("name,form,steps", CASES)
def test_blocks_file_is_the_reference_file(tmp_path, name, form, steps):
    from walicxe3d_amd import hip
    fx = load_golden("io.npz")
    w = W.get(name)
    bids = np.asarray(w.blocks(), dtype=np.int32)
    say("case write_bin", name, form, steps)
    g = hip.Walicxe3D(w, form=hip.PLAIN_FP64 if form == "plain" else hip.FUSED_FP32, bids=bids)
    g.upload(ic_in_order(w, bids))
    for _ in range(steps):
        g.step()
    p = str(tmp_path / "Blocks000.bin")
    g.writeBin(p, npassive=1)          # the oracle builds carry one passive scalar
    size, sha = _digest(p)
    assert size == int(fx["%s_d%d_size" % (name, steps)])
    assert sha == str(fx["%s_d%d_sha256" % (name, steps)])
    rb, ru = hip.read_bin(p, w.nv, npassive=1)
    assert np.array_equal(rb, bids)
    assert np.array_equal(ru, g.download(prim=False))
    g.close()


# writeOutput's VTK files (wx_write_vtk): the reference's bytes.  The AMR
# case is in physical units (units_type PHYS_UNITS), Orszag-Tang in code units
# with the magnetic field; output numbers are the step counts.
VTK_CASES = [("ot64", "plain", [0, 3]), ("amr_small", "plain", [0]), ("amr_small", "fused_fp32", [0])]


@pytest.mark.parametrize("name,form,outputs", VTK_CASES)
def test_vtk_files_are_the_reference_files(tmp_path, name, form, outputs):
    from walicxe3d_amd import hip
    fx = load_golden("io.npz")
    w = W.get(name)
    bids = np.asarray(w.blocks(), dtype=np.int32)
    say("case write_vtk", name, form, outputs)
    g = hip.Walicxe3D(w, form=hip.PLAIN_FP64 if form == "plain" else hip.FUSED_FP32, bids=bids)
    g.upload(ic_in_order(w, bids))
    done = 0
    for d in outputs:
        for _ in range(d - done):
            g.step()
        done = d
        g.write_vtk(str(tmp_path), d, npassive=1)
        for kind, fn in (("vtk", "Blocks000.%04d.vtk" % d), ("grid", "Grid.%04d.vtk" % d)):
            size, sha = _digest(str(tmp_path / fn))
            assert size == int(fx["%s_d%d_%s_size" % (name, d, kind)]), fn
            assert sha == str(fx["%s_d%d_%s_sha256" % (name, d, kind)]), fn
    if len(outputs) == len(fx["%s_dumps" % name]):
        assert (tmp_path / "master.visit").read_text() == str(fx["%s_visit" % name])
    g.close()
