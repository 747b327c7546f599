"""The reference's checkpoint files on the CPU: writeState's text
(wx_write_state) against the files the reference wrote
(tests/golden/io.npz), and warmstart's read of them (wx_read_state).  The
blocks files (wx_write_bin / wx_read_bin) are checked byte for byte against
the reference's in tests/test_emulated_kernels.py (CPU build of the library)
and tests/test_gpu_io.py."""
import pytest

from conftest import load_golden
from walicxe3d_amd import hip
from walicxe3d_amd import workloads as W

RUNS = [("ot64", 0), ("ot64", 3), ("amr_small", 0), ("amr_small", 2)]


@pytest.mark.parametrize("name,d", RUNS)
def test_state_file_is_the_reference_text(tmp_path, name, d):
    fx = load_golden("io.npz")
    w = W.get(name)
    time = 0.0 if d == 0 else float(fx["%s_steps" % name][d - 1, 2])
    p = str(tmp_path / ("State.%04d.dat" % d))
    hip.write_state(p, time, w.t_sc, d, d, "./")
    assert open(p).read() == str(fx["%s_d%d_state" % (name, d)])


@pytest.mark.parametrize("name,d", RUNS)
def test_state_file_read_as_warmstart_does(tmp_path, name, d):
    fx = load_golden("io.npz")
    w = W.get(name)
    text = str(fx["%s_d%d_state" % (name, d)])
    p = str(tmp_path / "State.dat")
    with open(p, "w") as f:
        f.write(text)
    time, it, nextout, datadir = hip.read_state(p, w.t_sc)
    assert it == d and nextout == d + 1 and datadir == "./"
    assert time == float(text[:22]) / w.t_sc


def test_blocks_file_errors(tmp_path):
    p = str(tmp_path / "bad.bin")
    with open(p, "wb") as f:
        f.write(b"\x02\x00\x00\x00" + b"\x00" * 100)
    with pytest.raises(hip.WxError):
        hip.read_bin(p, 5)
