"""The reference's main program on the device (walicxe3d_amd/run.py:
source/main.f90:32-134 with admesh and writeOutput), against the reference's
own adaptive run, and its warm start (source/init.f90:527-623) from the
files it wrote."""
import dataclasses
import os
import sys

import numpy as np
import pytest

from conftest import compare_dump, load_golden
from walicxe3d_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_main_loop_with_admesh_matches_the_reference():
    from walicxe3d_amd import run
    fx = load_golden("amr_amr_adapt.npz")
    w = W.get("amr_adapt")
    n = len(fx["steps"])
    say("case main loop amr_adapt plain", n)
    s = run.simulate(w, "plain", steps=n)
    assert s.it.value == n
    assert s.time.value == pytest.approx(fx["steps"][-1, 2], rel=1e-13)
    order = fx["b%d" % n]
    assert sorted(s.bids.tolist()) == sorted(order.tolist())
    u, p = s.download()
    pos = {int(b): k for k, b in enumerate(s.bids)}
    sel = [pos[int(b)] for b in order]
    dev_sum, dev_s = compare_dump(w, fx, n, u[sel], p[sel], order)
    assert dev_sum <= 1e-12 and dev_s <= 1e-12, (dev_sum, dev_s)
    s.close()


def test_warm_start_continues_the_run(tmp_path):
    """Orszag-Tang (no cooling: PRIM is flow2prim(U), which the warm start
    recomputes): 3 steps, an output, 3 more steps, against a warm start from
    that output's files; U bit-identical, time as the State file rounds it."""
    from walicxe3d_amd import hip, run
    w = W.get("ot64")
    say("case warm start ot64")
    a = run.simulate(w, "plain", steps=3, admesh=False)
    run.write_output(a, w, str(tmp_path), 7, True, False)
    for _ in range(3):
        a.step()
    b = run.simulate(w, "plain", steps=3, admesh=False, warm=str(tmp_path / "State.0007.dat"))
    assert b.it.value == a.it.value and b.nextout.value == 8
    assert b.time.value == pytest.approx(a.time.value, rel=1e-14)
    assert np.array_equal(a.download(prim=False), b.download(prim=False))
    a.close()
    b.close()


def test_outputs_when_the_clamp_fires(tmp_path):
    """amr_adapt in the fused form with admesh, dtout set inside the first
    steps: output 0 and the clamped outputs are written (blocks, VTK, State)
    and their State files carry the step and time of the output."""
    from walicxe3d_amd import hip, run
    fx = load_golden("amr_amr_adapt.npz")
    w0 = W.get("amr_adapt")
    dtout = 0.7 * fx["steps"][3, 2] * w0.t_sc      # falls within the first few steps
    w = dataclasses.replace(w0, dtout=dtout, tfin=2.5 * dtout)
    say("case outputs amr_adapt fused dtout", dtout)
    s = run.simulate(w, "fused", datadir=str(tmp_path), binout=True, vtkout=True)
    for k in (0, 1, 2):
        for name in ("Blocks000.%04d.bin", "Blocks000.%04d.vtk", "Grid.%04d.vtk", "State.%04d.dat"):
            assert os.path.exists(str(tmp_path / (name % k))), name % k
        t, it, nextout, d = hip.read_state(str(tmp_path / ("State.%04d.dat" % k)), w.t_sc)
        assert nextout == k + 1
        if k:
            assert t == pytest.approx(k * dtout / w.t_sc, rel=1e-12)   # the clamp lands on the output time
    assert s.time.value > 2.5 * dtout / w.t_sc
    lines = (tmp_path / "master.visit").read_text().splitlines()
    assert lines[0].startswith("!NBLOCKS") and lines[1:4] == ["Blocks000.%04d.vtk" % k for k in (0, 1, 2)]
    s.close()
