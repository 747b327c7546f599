"""ctypes binding of libwalicxe_hip.so (include/walicxe_hip.h) and a host-side
mirror of the reference's integrator interface.

The class `Walicxe3D` exposes the reference's routine names
(source/hydro_solver.f90 `hydroSolver`, `getTimestep`, `doViscousStep`;
source/godunov.f90 `Godunov`; source/boundary.f90 `boundary`;
source/hydro_core.f90 `calcPrimsAll`, `updatePrims`; source/cooling.f90
`cooling`) on device-resident state.  There is no CPU fallback: without the
built library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libwalicxe_hip.so")
if os.environ.get("WX_LIB"):
    # another build of the same C ABI (A/B measurements); said loudly, so that
    # no result is taken for the in-tree library's by mistake
    LIBPATH = os.environ["WX_LIB"]
    import sys as _sys
    _sys.stderr.write("walicxe3d_amd.hip: WX_LIB is set, loading %s instead of the in-tree library\n" % LIBPATH)

PLAIN_FP64 = 0
FUSED_FP32 = 1
FUSED_FP64 = 2
LAYOUT_BLOCKS = 0
LAYOUT_FORTRAN = 1
U, UP = 0, 1
CELLS_ALL, CELLS_PHYS, CELLS_GHOST = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class WxConfig(C.Structure):
    _fields_ = [("nv", C.c_int), ("mhd", C.c_int), ("solver", C.c_int), ("limiter", C.c_int),
                ("eight_wave", C.c_int), ("cooling", C.c_int), ("eos", C.c_int), ("bc", C.c_int * 6),
                ("gamma", C.c_double), ("cv", C.c_double), ("cfl", C.c_double), ("visc_eta", C.c_double),
                ("cooling_limit", C.c_double), ("mu0", C.c_double), ("mui", C.c_double),
                ("ion_thres", C.c_double), ("d_sc", C.c_double), ("e_sc", C.c_double),
                ("t_sc", C.c_double), ("p_sc", C.c_double), ("l_sc", C.c_double), ("v_sc", C.c_double),
                ("ntab", C.c_int), ("tab_logT", _dp), ("tab_loglam", _dp),
                ("nwind", C.c_int), ("winds", _dp), ("form", C.c_int)]


# every exported symbol of include/walicxe_hip.h with its signature
SIGNATURES = {
    "wx_init": (C.c_int, [C.POINTER(WxConfig), C.POINTER(C.c_void_p)]),
    "wx_finalize": (C.c_int, [C.c_void_p]),
    "wx_last_error": (C.c_char_p, []),
    "wx_version": (C.c_char_p, []),
    "wx_set_mesh": (C.c_int, [C.c_void_p, _ip, C.c_int, _dp, C.c_int, _ip]),
    "wx_upload_state": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, _ip]),
    "wx_upload_prims": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, _ip]),
    "wx_download_state": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int, C.c_int, _ip]),
    "wx_get_timestep": (C.c_int, [C.c_void_p, _dp]),
    "wx_godunov": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "wx_viscous_step": (C.c_int, [C.c_void_p]),
    "wx_update_prims": (C.c_int, [C.c_void_p]),
    "wx_cooling": (C.c_int, [C.c_void_p, C.c_double]),
    "wx_step": (C.c_int, [C.c_void_p, _ip, _dp, _dp, _ip, C.c_double, C.c_double]),
    "wx_run": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp, _dp, _ip, C.c_double, C.c_double]),
    "wx_boundary": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "wx_calc_prims": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "wx_fluxes_update": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "wx_copy_raw": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "wx_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "wx_set_plane_wind": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
    "wx_set_uniform_gravity": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]),
    "wx_last_timing": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "wx_time_stages": (C.c_int, [C.c_void_p, C.c_int]),
    "wx_remesh": (C.c_int, [C.c_void_p, C.c_int, _ip]),
    "wx_mark_blocks": (C.c_int, [C.c_void_p, C.c_double, C.c_double, _ip, _dp]),
    "wx_admesh_plan": (C.c_int, [_ip, C.c_int, C.c_int, _ip, _ip, _ip, C.c_int, _ip, _ip]),
    "wx_admesh": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int, _ip, _ip]),
    "wx_record_ghosts": (C.c_int, [C.c_void_p]),
    "wx_hilbert_keys": (C.c_int, [_ip, C.c_int, C.c_int, _ip, _ip]),
    "wx_write_bin": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "wx_read_bin": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _ip, _ip, _dp]),
    "wx_write_state": (C.c_int, [C.c_char_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_char_p]),
    "wx_read_state": (C.c_int, [C.c_char_p, C.c_double, _dp, _ip, _ip, C.c_char_p, C.c_int]),
    "wx_write_vtk": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int]),
    "wx_eval_physics": (C.c_int, [C.POINTER(WxConfig), C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "wx_eval_physics_noalloc": (C.c_int, [C.POINTER(WxConfig), C.c_int, C.c_int, C.c_int, _dp, _dp]),
}

EVAL_FLOW2PRIM, EVAL_PRIM2FLOW, EVAL_PRIM2FLUXES, EVAL_LIMITER, EVAL_RIEMANN, EVAL_RIEMANN_SEL, EVAL_SPEEDS = range(7)


def eval_physics(cfg, op, arg, records, nout, noalloc=False):
    """The device implementation of one physics routine on every record
    (rows of `records`); returns an array [n][nout] (wx_eval_physics, or
    wx_eval_physics_noalloc with noalloc)."""
    a = np.ascontiguousarray(records, dtype=np.float64).reshape(len(records), -1)
    out = np.empty((len(a), nout))
    name = "wx_eval_physics_noalloc" if noalloc else "wx_eval_physics"
    _check(getattr(load(), name)(C.byref(cfg), op, arg, len(a), a.ctypes.data_as(_dp), out.ctypes.data_as(_dp)),
           name)
    return out

FLAG_NONE, FLAG_REFINE, FLAG_COARSE = 0, 1, 2

# refineThres / coarseThres of the reference's parameters (source/parameters.f90:119-120)
REFINE_THRES = 0.3
COARSE_THRES = 0.03

_lib = None


def load():
    """Load the library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libwalicxe_hip.so is not built: run `python -m walicxe3d_amd.build`")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class WxError(RuntimeError):
    pass


def _check(rc, what):
    if rc < 0:
        raise WxError("%s failed (%d): %s" % (what, rc, load().wx_last_error().decode()))
    return rc


def _p(a):
    return a.ctypes.data_as(_dp)


def admesh_plan(mesh, bids, flags):
    """checkProximity + refine / coarsen list of admesh for given flags
    (wx_admesh_plan): returns (new bids, flags after the proximity rule)."""
    b = np.ascontiguousarray(bids, dtype=np.int32)
    f = np.ascontiguousarray(flags, dtype=np.int32)
    assert b.shape == f.shape and b.ndim == 1
    fo = np.empty_like(f)
    cap = 8 * len(b)
    out = np.empty(cap, dtype=np.int32)
    n = C.c_int(0)
    nroot = (C.c_int * 3)(*mesh.nroot)
    _check(load().wx_admesh_plan(nroot, mesh.maxlev, len(b), b.ctypes.data_as(_ip), f.ctypes.data_as(_ip),
                                 fo.ctypes.data_as(_ip), cap, C.byref(n), out.ctypes.data_as(_ip)),
           "wx_admesh_plan")
    return out[:n.value].copy(), fo


def hilbert_keys(mesh, bids):
    """getHilbertKey of every block (wx_hilbert_keys)."""
    b = np.ascontiguousarray(bids, dtype=np.int32)
    k = np.empty_like(b)
    nroot = (C.c_int * 3)(*mesh.nroot)
    _check(load().wx_hilbert_keys(nroot, mesh.maxlev, len(b), b.ctypes.data_as(_ip), k.ctypes.data_as(_ip)),
           "wx_hilbert_keys")
    return k


def read_bin(path, nv, npassive=0, cap=None):
    """The reference's blocks file (wx_read_bin): (bids, U [nblk][nv][16]^3)."""
    n = C.c_int(0)
    if cap is None:   # the header gives the count
        with open(path, "rb") as f:
            cap = int(np.frombuffer(f.read(4), dtype="<i4")[0])
    bids = np.empty(max(cap, 1), dtype=np.int32)
    u = np.empty((max(cap, 1), nv, 16, 16, 16))
    _check(load().wx_read_bin(os.fsencode(path), nv, npassive, cap, C.byref(n), bids.ctypes.data_as(_ip), _p(u)),
           "wx_read_bin")
    return bids[:n.value].copy(), u[:n.value].copy()


def write_state(path, time, t_sc, it, noutput, datadir="./"):
    """writeState's file (wx_write_state); time in code units."""
    _check(load().wx_write_state(os.fsencode(path), time, t_sc, it, noutput, datadir.encode()), "wx_write_state")


def read_state(path, t_sc):
    """warmstart's read (wx_read_state): (time in code units, it, nextout, datadir)."""
    t, it, nx = C.c_double(0.0), C.c_int(0), C.c_int(0)
    buf = C.create_string_buffer(256)
    _check(load().wx_read_state(os.fsencode(path), t_sc, C.byref(t), C.byref(it), C.byref(nx), buf, 256),
           "wx_read_state")
    return t.value, it.value, nx.value, buf.value.decode()


def make_config(w, form=PLAIN_FP64, table=None, winds=None):
    """wx_config for a workload (walicxe3d_amd.workloads.Workload)."""
    from walicxe3d_amd import workloads as W
    c = WxConfig()
    c.nv = w.nv
    c.mhd = int(w.mhd)
    c.solver = W.SOLVERS[w.solver]
    c.limiter = W.LIMITERS[w.limiter]
    c.eight_wave = int(w.eight_wave)
    c.cooling = int(w.cooling)
    c.eos = W.EOSS[w.eos]
    for f in range(6):
        c.bc[f] = W.BCS[w.bc[f]]
    c.gamma, c.cv, c.cfl, c.visc_eta = w.gamma, w.cv, w.cfl, w.visc_eta
    c.cooling_limit, c.mu0, c.mui, c.ion_thres = w.cooling_limit, w.mu0, w.mui, w.ion_thres
    c.d_sc, c.e_sc, c.t_sc, c.p_sc, c.l_sc, c.v_sc = w.d_sc, w.p_sc, w.t_sc, w.p_sc, w.l_sc, w.v_sc
    keep = []
    if w.cooling:
        lt, lam = table if table is not None else W.cooling_table()
        lt = np.ascontiguousarray(lt, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        keep += [lt, lam]
        c.ntab, c.tab_logT, c.tab_loglam = len(lt), _p(lt), _p(lam)
    wp = winds if winds is not None else W.wind_params(w)
    if wp:
        arr = np.ascontiguousarray(np.asarray(wp, dtype=np.float64).reshape(-1))
        keep.append(arr)
        c.nwind, c.winds = len(wp), _p(arr)
    c.form = form
    c._keep = keep
    return c


class Walicxe3D:
    """Device-resident integrator for one mesh, with the reference's routine
    names.  `it`, `time`, `dt` and `nextout` mirror the reference globals
    (source/globals.f90:44-48)."""

    def __init__(self, w, form=PLAIN_FP64, bids=None):
        L = load()
        self.w = w
        self.form = form
        self.cfg = make_config(w, form)
        h = C.c_void_p()
        _check(L.wx_init(C.byref(self.cfg), C.byref(h)), "wx_init")
        self._h = h
        if w.plane_wind:
            # imposePlaneWind of the problem's userBoundary (plain FP64 form,
            # hydro); on refusal the context is released by close()
            plane, rho, vel, temp, mu = w.plane_wind
            _check(L.wx_set_plane_wind(h, int(plane), rho, vel, temp, mu), "wx_set_plane_wind")
        if w.gravity:
            # the problem's get_user_source_terms (plain FP64 form)
            gx, gy, gz = w.gravity
            _check(L.wx_set_uniform_gravity(h, 1, gx, gy, gz), "wx_set_uniform_gravity")
        self.set_mesh(bids if bids is not None else w.blocks())
        self.it = C.c_int(0)
        self.dt = C.c_double(0.0)
        self.time = C.c_double(0.0)
        self.nextout = C.c_int(1)

    def set_mesh(self, bids):
        """(Re)send the active block list (after a regrid): rebuilds the face
        table and, in the fused form, recaptures the step graph on the next
        run.  The state must be uploaded again afterwards."""
        m = self.w.mesh()
        self.bids = np.ascontiguousarray(bids, dtype=np.int32)
        dx = np.ascontiguousarray(np.array(m.dx, dtype=np.float64).reshape(-1))
        nroot = (C.c_int * 3)(*m.nroot)
        _check(load().wx_set_mesh(self._h, nroot, m.maxlev, _p(dx), len(self.bids),
                                  self.bids.ctypes.data_as(_ip)), "wx_set_mesh")
        self.nblk = len(self.bids)

    def remesh(self, bids):
        """New active block list with the state carried over on the device
        (copy / octant injection / 2x2x2 restriction)."""
        b = np.ascontiguousarray(bids, dtype=np.int32)
        rc = _check(load().wx_remesh(self._h, len(b), b.ctypes.data_as(_ip)), "wx_remesh")
        self.bids = b
        self.nblk = len(b)
        return rc

    def writeBin(self, path, npassive=0):
        """The reference's blocks file of the current state (wx_write_bin)."""
        return _check(load().wx_write_bin(self._h, os.fsencode(path), npassive), "wx_write_bin")

    def write_vtk(self, datadir, noutput, npassive=0, units=None):
        """writeOutput's VTK files of the current state (wx_write_vtk);
        units "phys" or "code" (default: the workload's)."""
        units = units or self.w.units
        return _check(load().wx_write_vtk(self._h, os.fsencode(datadir), noutput, npassive,
                                          1 if units == "phys" else 2), "wx_write_vtk")

    def record_ghosts(self):
        """Fused form: keep the ghost layer of the current state for the next
        markByPhysical / admesh (call before the step they follow)."""
        return _check(load().wx_record_ghosts(self._h), "wx_record_ghosts")

    def markByPhysical(self, refine_thres=REFINE_THRES, coarse_thres=COARSE_THRES):
        """Flags and largest relative gradient of every block (wx_mark_blocks)."""
        f = np.empty(self.nblk, dtype=np.int32)
        g = np.empty(self.nblk)
        _check(load().wx_mark_blocks(self._h, refine_thres, coarse_thres, f.ctypes.data_as(_ip), _p(g)),
               "wx_mark_blocks")
        return f, g

    def admesh(self, refine_thres=REFINE_THRES, coarse_thres=COARSE_THRES, cap=None):
        """The reference's admesh on the device state (wx_admesh); returns
        the new block list (also kept in self.bids)."""
        cap = cap if cap is not None else 8 * self.nblk
        out = np.empty(cap, dtype=np.int32)
        n = C.c_int(0)
        rc = _check(load().wx_admesh(self._h, refine_thres, coarse_thres, cap, C.byref(n),
                                     out.ctypes.data_as(_ip)), "wx_admesh")
        self.bids = out[:n.value].copy()
        self.nblk = n.value
        return self.bids

    def close(self):
        if getattr(self, "_h", None):
            load().wx_finalize(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # ---- state -------------------------------------------------------------
    def upload(self, Ub):
        Ub = np.ascontiguousarray(Ub, dtype=np.float64)
        assert Ub.shape == (self.nblk, self.w.nv, 16, 16, 16), Ub.shape
        return _check(load().wx_upload_state(self._h, _p(Ub), LAYOUT_BLOCKS, 0, None), "wx_upload_state")

    def upload_prims(self, Pb):
        """Interior PRIM [nblk][nv][16][16][16] (plain form), e.g. the
        pressure cooling left, which U does not give back exactly."""
        Pb = np.ascontiguousarray(Pb, dtype=np.float64)
        assert Pb.shape == (self.nblk, self.w.nv, 16, 16, 16)
        return _check(load().wx_upload_prims(self._h, _p(Pb), LAYOUT_BLOCKS, 0, None), "wx_upload_prims")

    def download(self, prim=True):
        u = np.empty((self.nblk, self.w.nv, 16, 16, 16))
        p = np.empty_like(u) if prim else None
        _check(load().wx_download_state(self._h, _p(u), _p(p) if prim else None, LAYOUT_BLOCKS, 0, None),
               "wx_download_state")
        return (u, p) if prim else u

    def raw(self, which):
        n = self.nblk * self.w.nv * 20 ** 3
        out = np.empty(n)
        _check(load().wx_copy_raw(self._h, which, _p(out)), "wx_copy_raw")
        return out.reshape(self.nblk, self.w.nv, 20, 20, 20)

    # ---- the reference's routines --------------------------------------------
    def getTimestep(self):
        """Raw CFL timestep (before the start-up ramp and output clamp)."""
        d = C.c_double()
        _check(load().wx_get_timestep(self._h, C.byref(d)), "wx_get_timestep")
        return d.value

    def Godunov(self, order, dt):
        return _check(load().wx_godunov(self._h, order, dt), "wx_godunov")

    def doViscousStep(self):
        return _check(load().wx_viscous_step(self._h), "wx_viscous_step")

    def updatePrims(self):
        return _check(load().wx_update_prims(self._h), "wx_update_prims")

    def cooling(self, dt):
        return _check(load().wx_cooling(self._h, dt), "wx_cooling")

    def boundary(self, depth, which=U):
        return _check(load().wx_boundary(self._h, depth, which), "wx_boundary")

    def calcPrimsAll(self, which, cells):
        return _check(load().wx_calc_prims(self._h, which, cells), "wx_calc_prims")

    def fluxes_update(self, order, dtp):
        return _check(load().wx_fluxes_update(self._h, order, dtp), "wx_fluxes_update")

    def step(self):
        """One main-loop iteration: hydroSolver, updatePrims, cooling."""
        return _check(load().wx_step(self._h, C.byref(self.it), C.byref(self.dt), C.byref(self.time),
                                     C.byref(self.nextout), self.w.dtout, self.w.t_sc), "wx_step")

    def run(self, nsteps):
        return _check(load().wx_run(self._h, nsteps, C.byref(self.it), C.byref(self.dt), C.byref(self.time),
                                    C.byref(self.nextout), self.w.dtout, self.w.t_sc), "wx_run")

    def set_tuning(self, kind, value):
        return _check(load().wx_set_tuning(self._h, kind, value), "wx_set_tuning")

    def time_stages(self, reps):
        """Average stage-kernel durations (ms) from HIP events on the
        library's stream, fused form only."""
        _check(load().wx_time_stages(self._h, reps), "wx_time_stages")
        a, b, t = C.c_double(), C.c_double(), C.c_double()
        load().wx_last_timing(self._h, C.byref(a), C.byref(b), C.byref(t))
        return a.value, b.value

    def last_run_ms_per_step(self):
        a, b, t = C.c_double(), C.c_double(), C.c_double()
        load().wx_last_timing(self._h, C.byref(a), C.byref(b), C.byref(t))
        return t.value
