"""wx_eval_physics_noalloc (no device allocation: staging in a fixed device
array) against wx_eval_physics on the same records, bit for bit, on calls
in a row of decreasing size (the array keeps the previous call's contents)."""
import numpy as np
import pytest

from walicxe3d_amd import workloads as W

pytestmark = pytest.mark.gpu


def same(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(np.where(nan, 0, a), np.where(nan, 0, b))


def states(n, seed):
    g = np.random.default_rng(seed)
    p = np.empty((n, 8))
    p[:, 0] = g.uniform(0.1, 2.0, n)
    p[:, 1:4] = g.uniform(-2, 2, (n, 3))
    p[:, 4] = g.uniform(0.05, 2.0, n)
    p[:, 5:8] = g.uniform(-1.5, 1.5, (n, 3))
    return p


def test_noalloc_eval_matches_the_allocating_entry_point():
    from walicxe3d_amd import hip
    cfg = hip.make_config(W.get("ot64"), hip.PLAIN_FP64)
    print("case noalloc eval: HLLD 20000 records, limiter 777, flow2prim 5, HLLD 3", flush=True)
    calls = [(hip.EVAL_RIEMANN, 6, np.concatenate([states(20000, 1), states(20000, 2)], axis=1), 8),
             (hip.EVAL_LIMITER, 1, np.concatenate([states(777, 3 + q) for q in range(4)], axis=1), 16),
             (hip.EVAL_FLOW2PRIM, 0, states(5, 7), 9),
             (hip.EVAL_RIEMANN, 6, np.concatenate([states(3, 8), states(3, 9)], axis=1), 8)]
    for op, arg, rec, nout in calls:
        ref = hip.eval_physics(cfg, op, arg, rec, nout)
        got = hip.eval_physics(cfg, op, arg, rec, nout, noalloc=True)
        assert same(got, ref), (op, len(rec))
        assert np.isfinite(ref).mean() > 0.5
