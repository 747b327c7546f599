"""The device implementations of single physics routines (wx_eval_physics,
FP64, the plain form's arithmetic) against the reference's own per-routine
outputs (tests/golden/unit_*.npz, written by tests/golden/make_golden.py from
the compiled Walicxe3D sources): flow2prim, prim2flow, prim2fluxes along
each axis, all eight limiter options, the HLL, HLLC (in HLLC and non-HLLC
builds), HLLE and HLLD solvers, the HLLD of the fused FP64 kernels in its
one-code-path form, and the wave speeds.  The bar is bit for bit."""
import numpy as np
import pytest

from conftest import load_golden
from walicxe3d_amd import workloads as W

pytestmark = pytest.mark.gpu

UNIT_BUILDS = ["sedov", "blast_reflect", "briowu"]


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("build", UNIT_BUILDS)
def test_device_routines_match_reference_vectors(build):
    from walicxe3d_amd import hip
    fx = load_golden("unit_%s.npz" % build)
    w = W.get(build)
    cfg = hip.make_config(w, hip.PLAIN_FP64)
    nv = w.nv
    say("case device unit vectors", build)
    checked = []
    if "flow2prim_in" in fx:
        out = hip.eval_physics(cfg, hip.EVAL_FLOW2PRIM, 0, fx["flow2prim_in"], nv + 1)
        assert np.array_equal(out[:, :nv], fx["flow2prim_out"])
        assert np.array_equal(out[:, nv].astype(int), fx["flow2prim_istat"])
        checked.append("flow2prim")
    if "prim2flow_in" in fx:
        out = hip.eval_physics(cfg, hip.EVAL_PRIM2FLOW, 0, fx["prim2flow_in"], nv)
        assert np.array_equal(out, fx["prim2flow_out"])
        checked.append("prim2flow")
    if "prim2fluxes_in" in fx:
        for d in (1, 2, 3):
            sel = fx["prim2fluxes_dim"] == d
            assert sel.any()
            out = hip.eval_physics(cfg, hip.EVAL_PRIM2FLUXES, d, fx["prim2fluxes_in"][sel], nv)
            assert np.array_equal(out, fx["prim2fluxes_out"][sel]), d
        checked.append("prim2fluxes")
    if "limiter_in" in fx:
        lims = sorted(set(int(x) for x in fx["limiter_lim"]))
        assert lims == list(range(-1, 7)), lims          # every option of the reference
        for lim in lims:
            sel = fx["limiter_lim"] == lim
            out = hip.eval_physics(cfg, hip.EVAL_LIMITER, lim, fx["limiter_in"][sel], 2 * nv)
            assert np.array_equal(out.reshape(-1, 2, nv), fx["limiter_out"][sel]), lim
        checked.append("limiter")
    for key, solver in (("hll", 3), ("hllc", 4), ("hlle", 5), ("hlld", 6)):
        if key + "_pl" not in fx:
            continue
        rec = np.concatenate([fx[key + "_pl"], fx[key + "_pr"]], axis=1)
        out = hip.eval_physics(cfg, hip.EVAL_RIEMANN, solver, rec, nv)
        # NaN where the reference's routine matches no branch (its output is
        # undefined there, source/hllc.f90:290-296, hlld.f90:440-444)
        ref = fx[key + "_out"]
        both_nan = np.isnan(out) & np.isnan(ref)
        assert np.array_equal(np.where(both_nan, 0, out), np.where(both_nan, 0, ref)), key
        if key == "hlld":
            sel = hip.eval_physics(cfg, hip.EVAL_RIEMANN_SEL, 0, rec, nv)
            both_nan = np.isnan(sel) & np.isnan(ref)
            assert np.array_equal(np.where(both_nan, 0, sel), np.where(both_nan, 0, ref)), "hlld select form"
        checked.append(key)
    if "speeds_in" in fx:
        out = hip.eval_physics(cfg, hip.EVAL_SPEEDS, 0, fx["speeds_in"], 5)
        assert np.array_equal(out[:, 0], fx["speeds_out"][:, 0])
        if nv == 8:
            assert np.array_equal(out, fx["speeds_out"])
        checked.append("speeds")
    say("  checked", checked)
    assert checked
