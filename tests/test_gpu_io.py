"""The reference's checkpoint files from the device state (wx_write_bin,
writeBin of source/output.f90:156-214), byte for byte against the files the
reference itself wrote (tests/golden/io.npz: SHA-256 and size), and read
back (wx_read_bin, warmstart of source/init.f90:582-610)."""
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
