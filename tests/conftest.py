import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer-running check")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "libwxo.so")
    src = os.path.join(ROOT, "oracle", "wxo.c")
    if not os.path.exists(lib) or os.path.getmtime(lib) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return lib


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from oracle import wxo
    return wxo


def load_golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated" % name)
    return dict(np.load(path, allow_pickle=False))


def ic_in_order(w, bids):
    """The workload's initial condition ordered like `bids`."""
    from walicxe3d_amd import workloads as W
    slots, blocks, U = W.initial_state(w)
    idx = {b: i for i, b in enumerate(blocks)}
    return U[[idx[int(b)] for b in bids]]


def compare_dump(w, fx, d, u, p, bids, rtol=0.0):
    """Compare a state against the dumped summary `d` of a trajectory
    fixture; returns (max relative deviation of block sums, of samples)."""
    sums = u.reshape(len(bids), w.nv, -1).sum(axis=2)
    ref = fx["d%d_sums" % d]
    scale = np.maximum(fx["d%d_abss" % d], 1e-300)
    dev_sum = float(np.max(np.abs(sums - ref) / scale))
    pos = {int(b): i for i, b in enumerate(bids)}
    bi = np.array([pos[int(b)] for b in fx["d%d_sample_bid" % d]])
    ci = fx["d%d_sample_cell" % d]
    su = u.reshape(len(bids), w.nv, -1)[bi, :, ci]
    sp = p.reshape(len(bids), w.nv, -1)[bi, :, ci]
    ru, rp = fx["d%d_sample_u" % d], fx["d%d_sample_p" % d]
    # per variable: largest deviation relative to the variable's mean magnitude
    den_u = np.maximum(np.abs(ru).mean(axis=0), 1e-300)
    den_p = np.maximum(np.abs(rp).mean(axis=0), 1e-300)
    dev_s = float(max(np.max(np.abs(su - ru).max(axis=0) / den_u), np.max(np.abs(sp - rp).max(axis=0) / den_p)))
    return dev_sum, dev_s


def rel_l1(a, b):
    """Per-variable relative L1 error sum|a-b| / sum|b| (max over variables);
    a, b shaped [nblk][nv][...]."""
    nv = a.shape[1]
    d = np.abs(a - b).reshape(a.shape[0], nv, -1).sum(axis=(0, 2))
    s = np.abs(b).reshape(a.shape[0], nv, -1).sum(axis=(0, 2))
    return np.where(s > 0, d / np.maximum(s, 1e-300), d)


def mixed_level_ot(bc=None):
    """Orszag-Tang 64^3 (64 blocks at level 3 of 3) with two families merged
    into their level-2 fathers, whose state is the 2x2x2 restriction of the
    initial condition: (workload, block IDs, state).  Every coarse block has
    finer neighbours across its faces.  `bc` replaces the workload's periodic
    boundaries (on a periodic domain some of those faces wrap around)."""
    import dataclasses
    import numpy as np
    from walicxe3d_amd import regrid
    from walicxe3d_amd import workloads as W
    w = W.get("ot64")
    if bc is not None:
        w = dataclasses.replace(w, bc=(bc,) * 6)
    m = w.mesh()
    b1 = np.asarray(w.blocks(), dtype=np.int32)
    fathers = [m.bid(2, 1, 1, 1), m.bid(2, 2, 2, 1)]
    gone = {c for f in fathers for c in m.children(f)}
    b2 = np.asarray([b for b in b1.tolist() if b not in gone] + fathers, dtype=np.int32)
    return w, b2, regrid.remap(m, b1, ic_in_order(w, b1), b2)
