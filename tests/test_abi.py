"""The C ABI: the library loads and exports every entry point declared in
include/walicxe_hip.h; argument checks that need no GPU."""
import ctypes as C
import os
import re

import numpy as np

import pytest

from conftest import ROOT
from walicxe3d_amd import hip

HEADER = os.path.join(ROOT, "include", "walicxe_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(wx_\w+)\s*\(", text, flags=re.M)


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 20
    assert set(names) == set(hip.SIGNATURES), set(names) ^ set(hip.SIGNATURES)


def test_library_exports_every_symbol():
    if not os.path.exists(hip.LIBPATH):
        pytest.skip("library not built")
    lib = C.CDLL(hip.LIBPATH)
    for n in declared():
        assert hasattr(lib, n), n


def test_version_and_error_strings():
    if not os.path.exists(hip.LIBPATH):
        pytest.skip("library not built")
    L = hip.load()
    assert b"gfx950" in L.wx_version()
    assert isinstance(L.wx_last_error(), bytes)


def test_rejects_bad_config_before_touching_a_device():
    """Argument validation runs before any device query."""
    if not os.path.exists(hip.LIBPATH):
        pytest.skip("library not built")
    from walicxe3d_amd import workloads as W
    L = hip.load()
    cfg = hip.make_config(W.get("sedov"))
    cfg.nv = 9
    h = C.c_void_p()
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1
    assert b"nv" in L.wx_last_error()
    # passive scalars (nv = 6, 7): plain form, hydro, no wind sources
    cfg = hip.make_config(W.get("sedov_passive"), form=hip.FUSED_FP32)
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1
    assert b"plain FP64 form" in L.wx_last_error()
    cfg = hip.make_config(W.get("sedov_passive"))
    cfg.mhd = 1
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1
    assert b"hydro" in L.wx_last_error()
    cfg = hip.make_config(W.get("sedov"))
    cfg.solver = 6   # HLLD needs MHD
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1
    cfg = hip.make_config(W.get("sedov"), form=hip.FUSED_FP32)
    cfg.solver = 2   # first-order HLL runs only in the plain form
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1
    assert L.wx_set_mesh(None, None, 1, None, 1, None) == -1
    # an unknown EOS; EOS_H_RATE without the passive scalar it reads
    cfg = hip.make_config(W.get("amr_small"))
    cfg.eos = 4
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1
    assert b"eos" in L.wx_last_error()
    cfg.eos = 3
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1
    assert b"passive scalar" in L.wx_last_error()
    # the fused form runs HLLE/HLLD (and MHD) for 8 variables only
    cfg = hip.make_config(W.get("ot64"), form=hip.FUSED_FP32)
    cfg.solver = 4
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1
    cfg = hip.make_config(W.get("ot64"), form=hip.FUSED_FP32)
    cfg.mhd = 0
    assert L.wx_init(C.byref(cfg), C.byref(h)) == -1


def test_download_checks_its_layout_arguments():
    """wx_download_state refuses a Fortran layout without localBlocks or
    nbMaxProc, and an unknown layout, before it touches any state."""
    if not os.path.exists(hip.LIBPATH):
        pytest.skip("library not built")
    L = hip.load()
    buf = (C.c_double * 8)()
    assert L.wx_download_state(None, buf, None, hip.LAYOUT_FORTRAN, 0, None) == -1
    assert b"localBlocks" in L.wx_last_error()
    lb = (C.c_int * 1)(1)
    assert L.wx_download_state(None, buf, None, hip.LAYOUT_FORTRAN, 0, lb) == -1
    assert L.wx_download_state(None, buf, None, 7, 1, lb) == -1
    assert b"layout" in L.wx_last_error()
    assert L.wx_step(None, None, None, None, None, 1.0, 1.0) == -1


def test_plane_wind_refuses_a_missing_context():
    """wx_set_plane_wind checks its context before anything else (the
    form and state checks need one, on a GPU)."""
    if not os.path.exists(hip.LIBPATH):
        pytest.skip("library not built")
    L = hip.load()
    assert L.wx_set_plane_wind(None, 6, 1e-23, 1e7, 1e4, 1.3) == -1
    assert b"null context" in L.wx_last_error()


def test_uniform_gravity_refuses_a_missing_context():
    """wx_set_uniform_gravity checks its context before anything else."""
    if not os.path.exists(hip.LIBPATH):
        pytest.skip("library not built")
    L = hip.load()
    assert L.wx_set_uniform_gravity(None, 1, 0.0, 0.0, -1.0) == -1
    assert b"null context" in L.wx_last_error()


def test_noalloc_eval_refuses_more_than_its_staging():
    """wx_eval_physics_noalloc stages through a fixed array: a call whose
    records exceed it is refused before any device call."""
    if not os.path.exists(hip.LIBPATH):
        pytest.skip("library not built")
    from walicxe3d_amd import workloads as W
    L = hip.load()
    cfg = hip.make_config(W.get("ot64"))
    n = 10923                                  # 48 doubles a limiter record of 8 variables
    a = np.zeros((n, 32))
    out = np.zeros((n, 16))
    dp = C.POINTER(C.c_double)
    rc = L.wx_eval_physics_noalloc(C.byref(cfg), hip.EVAL_LIMITER, 1, n, a.ctypes.data_as(dp), out.ctypes.data_as(dp))
    assert rc == -1
    assert b"staging" in L.wx_last_error()
    assert L.wx_eval_physics_noalloc(C.byref(cfg), 99, 0, 1, a.ctypes.data_as(dp), out.ctypes.data_as(dp)) == -1
