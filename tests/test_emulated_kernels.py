"""The library's HIP sources compiled for the CPU (tests/emu) and run under
AddressSanitizer and UBSan's array-bounds checks: every kernel of the plain
and fused forms executes on real workloads (AMR faces of every kind,
cooling, winds, reflective / free-flow / periodic boundaries, hydro and
MHD), and any out-of-range index aborts the run with its source line - into
a device array (ASan), into a fixed-size array such as an LDS plane or a
lane's register array (bounds), or outside the ranges the kernels state
for cell, block and ghost-source indices (WX_EMU_CHECK).  Lanes waiting at
different barriers abort too, and the fused kernels must give bit-identical
states when the lanes of every barrier interval run in a shuffled order.
The emulation itself is checked against the oracle: the fused FP64 form
must reproduce it bit for bit on Orszag-Tang, as it does on the GPU."""
import fcntl
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from walicxe3d_amd import hip
from walicxe3d_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "emu")
BUILD = os.path.join(EMU, "_build")
CXX = shutil.which("g++")

# (workload, form, steps)
CASES = [("ot64", hip.FUSED_FP32, 2), ("sedov", hip.FUSED_FP32, 2), ("blast_reflect", hip.FUSED_FP32, 2),
         ("amr_small", hip.FUSED_FP32, 2), ("mhd_hlle", hip.FUSED_FP32, 2), ("winds4", hip.FUSED_FP32, 1),
         ("snr_amr4", hip.FUSED_FP32, 1), ("amr_small", hip.PLAIN_FP64, 1), ("mhd_hlle", hip.PLAIN_FP64, 1),
         ("winds4", hip.PLAIN_FP64, 1)]


def _sources():
    import glob
    return (glob.glob(os.path.join(ROOT, "walicxe3d_amd", "csrc", "*")) + glob.glob(os.path.join(EMU, "*.cpp"))
            + glob.glob(os.path.join(EMU, "hip", "*")) + glob.glob(os.path.join(ROOT, "include", "*")))


def _build():
    # Parallel test workers share the build folder: one build at a time, a
    # binary newer than every source is reused, and a new one replaces the
    # old by rename (a worker still running the old keeps its file).
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, "emu")
    with open(os.path.join(BUILD, ".lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        if os.path.exists(exe) and os.path.getmtime(exe) > max(os.path.getmtime(f) for f in _sources()):
            return exe
        tmp = exe + ".%d" % os.getpid()
        _build_into(tmp)
        os.replace(tmp, exe)
        return exe


def _build_into(exe):
    flags = ["-std=c++20", "-O1", "-g", "-fsanitize=address,bounds", "-fno-sanitize-recover=bounds",
             "-fno-omit-frame-pointer", "-ffp-contract=off", "-DWX_EMU_CHECKS",
             "-Wno-unknown-pragmas", "-I" + EMU, "-I" + os.path.join(ROOT, "walicxe3d_amd", "csrc"),
             "-I" + os.path.join(ROOT, "include")]
    objs = []
    for src in ("wx_api", "wx_fused", "wx_fused_mhd", "wx_plain", "wx_io"):
        o = os.path.join(BUILD, src + ".o")
        cmd = [CXX, "-x", "c++"] + flags + ["-c", os.path.join(ROOT, "walicxe3d_amd", "csrc", src + ".hip"),
                                            "-o", o]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        objs.append(o)
    cmd = [CXX] + flags + [os.path.join(EMU, "emu_rt.cpp"), os.path.join(EMU, "emu_main.cpp")] + objs + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _write_input(path, w, form, remesh_to=None, admesh=None, state=None):
    if state is None:
        slots, blocks, U0 = W.initial_state(w)
    else:
        blocks, U0 = state
    cfg = hip.make_config(w, form)
    raw = bytes(cfg)
    lt, lam = W.cooling_table() if w.cooling else (np.zeros(0), np.zeros(0))
    wp = np.asarray(W.wind_params(w), dtype=np.float64).reshape(-1)
    m = w.mesh()
    with open(path, "wb") as f:
        f.write(struct.pack("<8i", w.nv, len(blocks), len(lt), len(wp) // 8, m.maxlev, *m.nroot))
        f.write(struct.pack("<i", len(raw)))
        f.write(raw)
        for a in (lt, lam, wp, np.array(m.dx, dtype=np.float64).reshape(-1)):
            f.write(np.asarray(a, dtype="<f8").tobytes())
        f.write(np.asarray(blocks, dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(U0, dtype="<f8").tobytes())
        f.write(struct.pack("<2d", w.dtout, w.t_sc))
        if remesh_to is not None:
            f.write(struct.pack("<i", len(remesh_to)))
            f.write(np.asarray(remesh_to, dtype="<i4").tobytes())
        elif admesh is not None:
            f.write(struct.pack("<i2d", -1, *admesh))
    return blocks, U0


@pytest.fixture(scope="module")
def emu():
    if CXX is None:
        pytest.skip("g++ not available")
    return _build()


def _run(emu, args, tmp, **env_extra):
    env = dict(os.environ, ASAN_OPTIONS="detect_stack_use_after_return=0:detect_leaks=0", **env_extra)
    r = subprocess.run([emu] + args, capture_output=True, text=True, timeout=900, env=env, cwd=tmp)
    assert r.returncode == 0 and "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, \
        (r.stdout + r.stderr)[-4000:]
    return r.stdout


def test_every_kernel_in_bounds(emu, tmp_path):
    import dataclasses
    args = []
    for n, (name, form, steps) in enumerate(CASES):
        p = str(tmp_path / ("in%d.bin" % n))
        _write_input(p, W.get(name), form)
        args += [p, str(steps)]
    # the device remesh (refine, then coarsen back) in both forms
    w = W.get("winds4")
    finer = [int(b) for b in dataclasses.replace(w, refine_radii=(0.3, 0.15, 0.08, 0.04)).blocks()]
    for n, form in enumerate((hip.FUSED_FP32, hip.PLAIN_FP64)):
        p = str(tmp_path / ("remesh%d.bin" % n))
        _write_input(p, w, form, remesh_to=finer)
        args += [p, "1"]
    out = _run(emu, args, str(tmp_path))
    assert out.count("it ") == len(CASES) + 2, out


def test_fused_kernels_do_not_depend_on_lane_order(emu, tmp_path):
    """The fused stage kernels with the lanes of every barrier interval run
    in two seeded random orders give the state of the in-order run bit for
    bit: no lane reads LDS (or device) data that another lane writes in the
    same interval (a race the GPU would run in some order of its own)."""
    cases = [("ot64", hip.FUSED_FP32), ("amr_small", hip.FUSED_FP32), ("snr_amr4", hip.FUSED_FP32),
             ("ot64_amr", hip.FUSED_FP32), ("blast_reflect", hip.FUSED_FP64)]
    for name, form in cases:
        outs = []
        for seed in (None, "1", "77"):
            p = str(tmp_path / ("%s_%s.bin" % (name, seed)))
            _write_input(p, W.get(name), form)
            env = {"EMU_ORDER_SEED": seed} if seed else {}
            _run(emu, [p, "1"], str(tmp_path), **env)
            outs.append(np.fromfile(p + ".out"))
        assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2]), name


def test_z_slabs_leave_the_state_unchanged(emu, tmp_path):
    """The FP32 MHD kernels with each block split into 2, 4 and 8 z slabs
    (one workgroup each) stay in bounds and give the whole-block state bit
    for bit (a periodic and a two-level mesh, and the kernels that carry the
    wind code on a mesh with wind sources).  Elsewhere the setting has no
    effect (viscous MHD, hydro)."""
    import dataclasses
    cases = [(W.get(n), "1 2 4 8", "2") for n in ("ot64", "ot64_amr", "mhd_hlle", "sedov")]
    # the kernels that carry the wind code (inviscid: in slabs), one step
    cases.append((dataclasses.replace(W.get("mhd_winds4"), visc_eta=0.0), "1 8", "1"))
    for w, slabs, steps in cases:
        outs = []
        for nslab in slabs.split():
            p = str(tmp_path / ("%s_%s.bin" % (w.name, nslab)))
            _write_input(p, w, hip.FUSED_FP32)
            _run(emu, [p, steps], str(tmp_path), EMU_NSLAB=nslab)
            outs.append(np.fromfile(p + ".out"))
        for o in outs[1:]:
            assert np.array_equal(o, outs[0]), w.name


def test_emulation_reproduces_fused_fp64(emu, tmp_path, oracle):
    w = W.get("ot64")
    p = str(tmp_path / "ot64.bin")
    blocks, U0 = _write_input(p, w, hip.FUSED_FP64)
    vtkdir = tmp_path / "vtk"
    vtkdir.mkdir()
    _run(emu, [p, "3"], str(tmp_path), EMU_VTK_DIR=str(vtkdir))
    o = oracle.Sim(w, blocks)
    o.set_state(U0)
    for _ in range(3):
        o.step()
    uo, _ = o.get_state()
    ue = np.fromfile(p + ".out").reshape(uo.shape)
    assert np.array_equal(ue, uo)
    # wx_write_bin of that state is the reference's own writeBin file of step 3
    # byte for byte, and wx_read_bin reads it back
    import hashlib
    from conftest import load_golden
    fx = load_golden("io.npz")
    data = open(p + ".wxbin", "rb").read()
    assert len(data) == int(fx["ot64_d3_size"])
    assert hashlib.sha256(data).hexdigest() == str(fx["ot64_d3_sha256"])
    bids, ur = hip.read_bin(p + ".wxbin", 8, npassive=1)
    assert np.array_equal(bids, np.asarray(blocks)) and np.array_equal(ur, uo)
    # wx_write_vtk: the reference's VTK files of outputs 0 and 3 and its
    # master.visit, byte for byte
    for d in (0, 3):
        for kind, name in (("vtk", "Blocks000.%04d.vtk" % d), ("grid", "Grid.%04d.vtk" % d)):
            data = (vtkdir / name).read_bytes()
            assert len(data) == int(fx["ot64_d%d_%s_size" % (d, kind)]), name
            assert hashlib.sha256(data).hexdigest() == str(fx["ot64_d%d_%s_sha256" % (d, kind)]), name
    assert (vtkdir / "master.visit").read_text() == str(fx["ot64_visit"])


def test_emulated_mhd_with_wind_sources_match_the_oracle(emu, tmp_path, oracle):
    """The FP32 MHD stage kernels that carry the wind code: in a cell a
    source covers, each stage's update starts from the wind state (the
    reference's boundary() imposes it before each stage), viscous and
    inviscid; within 1e-5 per variable of the oracle after one step (without
    the wind state the momenta are off by O(1) at once)."""
    import dataclasses
    from conftest import rel_l1
    for visc in (5.0e-3, 0.0):
        w = dataclasses.replace(W.get("mhd_winds4"), visc_eta=visc)
        p = str(tmp_path / ("mhd_winds_%g.bin" % visc))
        blocks, U0 = _write_input(p, w, hip.FUSED_FP32)
        _run(emu, [p, "1"], str(tmp_path))
        o = oracle.Sim(w, blocks)
        o.set_state(U0)
        o.step()
        uo, _ = o.get_state()
        e = rel_l1(np.fromfile(p + ".out").reshape(uo.shape), uo)
        assert np.all(e <= 1e-5), (visc, e)


def test_emulated_uniform_gravity_reproduces_the_oracle(emu, tmp_path, oracle):
    """The user source term (a uniform acceleration) in the plain FP64 form,
    hydro and MHD (ahead of the Powell terms): bit-exact to the oracle after
    two steps."""
    for name in ("sedov_gravity", "ot64_gravity"):
        w = W.get(name)
        p = str(tmp_path / (name + ".bin"))
        blocks, U0 = _write_input(p, w, hip.PLAIN_FP64)
        _run(emu, [p, "2"], str(tmp_path), EMU_GRAVITY="%r %r %r" % tuple(w.gravity))
        o = oracle.Sim(w, blocks)
        o.set_state(U0)
        o.step()
        o.step()
        uo, _ = o.get_state()
        assert np.array_equal(np.fromfile(p + ".out").reshape(uo.shape), uo), name


def test_emulated_passive_scalars_reproduce_the_oracle(emu, tmp_path, oracle):
    """Passive scalars in the plain FP64 form (HLLC, HLL with reflecting
    walls, HLLC on AMR levels with cooling; one tracer, and two on the
    blast): bit-exact to the oracle, tracers included."""
    import dataclasses
    cases = [(W.get("sedov_passive"), 2), (dataclasses.replace(W.get("sedov_passive"), npassive=2), 2),
             (W.get("blast_reflect_passive"), 2), (W.get("amr_small_passive"), 1), (W.get("amr_small_hrate"), 2)]
    for n, (w, steps) in enumerate(cases):
        p = str(tmp_path / ("passive%d.bin" % n))
        blocks, U0 = _write_input(p, w, hip.PLAIN_FP64)
        vtkdir = tmp_path / ("vtk%d" % n)
        vtkdir.mkdir()
        _run(emu, [p, str(steps)], str(tmp_path), EMU_VTK_DIR=str(vtkdir) + "/", EMU_VTK_UNITS="2")
        o = oracle.Sim(w, blocks)
        o.set_state(U0)
        for _ in range(steps):
            o.step()
        uo, _ = o.get_state()
        u = np.fromfile(p + ".out").reshape(uo.shape)
        assert np.array_equal(u, uo), (w.name, w.npassive, [float(np.abs(u[:, v] - uo[:, v]).max()) for v in range(w.nv)])
        # the checkpoint file (writeBin layout: per block the bID, then
        # [cell][neqtot], variable fastest) holds the tracers as its passives
        raw = open(p + ".wxbin", "rb").read()
        nb = int(np.frombuffer(raw[:4], dtype="<i4")[0])
        assert nb == u.shape[0]
        rec = 4 + 8 * w.nv * 4096
        for b in (0, nb - 1):
            data = np.frombuffer(raw[4 + b * rec + 4: 4 + (b + 1) * rec], dtype="<f8").reshape(4096, w.nv)
            assert np.array_equal(data.T, u[b].reshape(w.nv, 4096))
        # the VTK blocks file's passive1 field (write3DVTKBlocks,
        # output.f90:711-737) holds the tracer as FP32, big-endian
        if w.npassive == 1:
            vtk = (vtkdir / ("Blocks000.%04d.vtk" % steps)).read_bytes()
            key = b"SCALARS passive1 float 1\nLOOKUP_TABLE default\n"
            at = vtk.index(key) + len(key)
            field = np.frombuffer(vtk[at: at + 4 * u.shape[0] * 4096], dtype=">f4")
            assert np.array_equal(field, u[:, 5].reshape(-1).astype(np.float32))


def test_emulated_admesh_reproduces_the_oracle(emu, tmp_path, oracle):
    """wx_admesh after every step (device flagging, host plan, remesh with
    PRIM carried) in the plain FP64 form, against the oracle's restatement of
    the reference's admesh: same final block set, bit-identical state.  The
    fused form runs the same calls for the bounds check."""
    from oracle import amr
    w = W.get("amr_adapt")
    nsteps = 3
    p = str(tmp_path / "adapt_plain.bin")
    _write_input(p, w, hip.PLAIN_FP64, admesh=(0.3, 0.03))
    q = str(tmp_path / "adapt_fused.bin")
    _write_input(q, w, hip.FUSED_FP32, admesh=(0.3, 0.03))
    out = _run(emu, [p, str(nsteps), q, str(nsteps)], str(tmp_path))
    assert out.count("it ") == 2, out
    o = amr.trajectory(w, nsteps, [nsteps])
    ob = [int(b) for b in o["blocks"][nsteps]]
    eb = [int(b) for b in np.fromfile(p + ".blocks", dtype=np.int32)]
    assert sorted(eb) == sorted(ob)
    uo, _ = o["states"][nsteps]
    ue = np.fromfile(p + ".out").reshape((len(eb),) + uo.shape[1:])
    pos = {b: n for n, b in enumerate(eb)}
    assert np.array_equal(ue[[pos[b] for b in ob]], uo)


def test_mhd_fused_on_mixed_levels(emu, tmp_path, oracle):
    """The FP32 MHD stage kernels across coarse-fine faces (Orszag-Tang on
    two levels, free boundaries: the ot64_amr workload): every access in
    bounds, and the state after 3 steps within the Orszag-Tang tolerances of
    the oracle (1e-5 on the near-zero z-momentum and B_z, 1e-7 on the
    other variables).  On the periodic domain the library refuses such a mesh
    (faces wrap onto blocks of another level)."""
    from conftest import mixed_level_ot, rel_l1
    w, b2, U2 = mixed_level_ot()
    p = str(tmp_path / "periodic.bin")
    _write_input(p, w, hip.FUSED_FP32, state=(b2, U2))
    r = subprocess.run([emu, p, "1"], capture_output=True, text=True, timeout=300, cwd=str(tmp_path),
                       env=dict(os.environ, ASAN_OPTIONS="detect_stack_use_after_return=0:detect_leaks=0"))
    assert r.returncode != 0 and "not active at the same level" in r.stdout, r.stdout[-2000:]
    w = W.get("ot64_amr")                     # the same mesh on free boundaries
    p = str(tmp_path / "mixed.bin")
    b2, U2 = _write_input(p, w, hip.FUSED_FP32)
    _run(emu, [p, "3"], str(tmp_path))
    o = oracle.Sim(w, b2)
    o.set_state(U2)
    for _ in range(3):
        o.step()
    uo, _ = o.get_state()
    ue = np.fromfile(p + ".out").reshape(uo.shape)
    e = rel_l1(ue, uo)
    assert np.all(np.isfinite(ue))
    for v in range(w.nv):
        assert e[v] <= (1e-5 if v in (3, 7) else 1e-7), (v, e)
