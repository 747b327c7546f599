"""Mesh adaptation (admesh, source/amr.f90:44-133) on the CPU.

* The oracle's restatement (oracle/amr.py + wxo_mark) against the reference's
  own run with admesh after every step (tests/golden/amr_amr_adapt.npz):
  markByPhysical flags, the block list in slot order, dt/time and the state,
  all bit-exact.
* The library's host plan (wx_admesh_plan: checkProximity, refine list,
  coarsened families) against the same fixture and the oracle.
* Hand-built meshes whose outcome follows from the reference's rules by
  inspection, one per rule, along each axis, through both the oracle and the
  library.
* Hilbert keys (wx_hilbert_keys) against getHilbertKey of the reference for
  every block ID of two meshes (tests/golden/hkeys.npz), and the curve's own
  property: consecutive keys are face neighbours.
"""
import numpy as np
import pytest

from conftest import compare_dump, ic_in_order, load_golden
from walicxe3d_amd import hip
from walicxe3d_amd import workloads as W

NONE, REFINE, COARSE = 0, 1, 2


def _fixture():
    return load_golden("amr_amr_adapt.npz")


def test_oracle_reproduces_the_reference_admesh(oracle):
    from oracle import amr
    fx = _fixture()
    w = W.get("amr_adapt")
    U0 = ic_in_order(w, fx["ic_bids"])
    assert np.sum(U0) == pytest.approx(fx["ic_sum"][0], rel=1e-15, abs=0)
    n = len(fx["steps"])
    dumps = [int(d) for d in fx["dumps"]]
    o = amr.trajectory(w, n, dumps)
    for s in range(1, n + 1):
        it, dt, time = o["steps"][s - 1]
        assert dt == fx["steps"][s - 1, 1] and time == fx["steps"][s - 1, 2], s
        b, f = o["flags"][s]
        assert np.array_equal(b, fx["f%d_bids" % s]), s
        assert np.array_equal(f, fx["f%d_flags" % s]), s
        # slot order included: the put/pop bookkeeping is the reference's
        assert np.array_equal(o["blocks"][s], fx["b%d" % s]), s
    # the trajectory exercises every outcome of markByPhysical and a change
    # of the block count in both directions
    allf = np.concatenate([fx["f%d_flags" % s] for s in range(1, n + 1)])
    assert {NONE, REFINE, COARSE} <= set(allf.tolist())
    counts = [len(fx["f1_bids"])] + [len(fx["b%d" % s]) for s in range(1, n + 1)]
    assert max(np.diff(counts)) > 0 and min(np.diff(counts)) < 0, counts
    for d in dumps:
        u, p = o["states"][d]
        dev_sum, dev_s = compare_dump(w, fx, d, u, p, o["blocks"][d])
        assert dev_sum == 0.0 and dev_s == 0.0, (d, dev_sum, dev_s)


def test_library_plan_matches_the_reference(oracle):
    from oracle import amr
    fx = _fixture()
    w = W.get("amr_adapt")
    m = w.mesh()
    for s in range(1, len(fx["steps"]) + 1):
        b, f = fx["f%d_bids" % s], fx["f%d_flags" % s]
        new, fo = hip.admesh_plan(m, b, f)
        assert sorted(new.tolist()) == sorted(fx["b%d" % s].tolist()), s
        fr = [int(x) for x in f]
        amr.check_proximity(m, [int(x) for x in b], fr, m.maxlev - 1)
        assert fo.tolist() == fr, s


# --------------------------------------------------------------------------
# hand-built cases: two root blocks side by side along `axis`, maxlev 3
# --------------------------------------------------------------------------
def _mesh(axis):
    nroot = [1, 1, 1]
    nroot[axis] = 2
    return W.Mesh(nroot, 3, [float(n) for n in nroot], 1.0)


def _at(m, level, axis, a, t1=1, t2=1):
    """Block at `level` with coordinate a along the axis, t1/t2 across."""
    c = [t1, t2]
    c.insert(axis, a)
    return m.bid(level, *c)


def _plan_both(m, bids, flags):
    """(new list, flags after proximity) from the library and the oracle,
    checked to agree."""
    from oracle import amr
    new, fo = hip.admesh_plan(m, bids, flags)
    nv = 5
    U = np.zeros((len(bids), nv, 16, 16, 16))
    slots, fl, _, _, _ = amr.admesh(m, list(bids), list(flags), U, U)
    assert sorted(b for b in slots if b != -1) == sorted(new.tolist())
    assert fl == fo.tolist()
    return set(new.tolist()), dict(zip(bids, fo.tolist()))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_refining_block_forces_its_coarser_neighbour(oracle, axis):
    m = _mesh(axis)
    r1, r2 = 1, 2
    fam = m.children(r1)
    x = _at(m, 2, axis, 2)                       # child of root 1 touching root 2
    bids = fam + [r2]
    flags = [REFINE if b == x else NONE for b in bids]
    new, fo = _plan_both(m, bids, flags)
    assert fo[r2] == REFINE
    assert new == (set(fam) - {x}) | set(m.children(x)) | set(m.children(r2))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_refining_block_stops_a_same_level_neighbour_coarsening(oracle, axis):
    m = _mesh(axis)
    fam1, fam2 = m.children(1), m.children(2)
    y = _at(m, 2, axis, 3)                       # root 2's child touching root 1's family
    nbr = _at(m, 2, axis, 2)                     # its same-level neighbour in root 1's family
    bids = fam1 + fam2
    flags = [COARSE if b in fam1 else (REFINE if b == y else NONE) for b in bids]
    new, fo = _plan_both(m, bids, flags)
    assert fo[nbr] == NONE and sum(fo[b] == COARSE for b in fam1) == 7
    # a family with 7 members flagged stays; y is refined
    assert new == set(fam1) | (set(fam2) - {y}) | set(m.children(y))


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("finer_all_coarse", [False, True])
def test_finer_family_decides_a_coarsening_neighbour(oracle, axis, finer_all_coarse):
    m = _mesh(axis)
    fam1, fam2 = m.children(1), m.children(2)
    z = _at(m, 2, axis, 3)                       # refined child of root 2, touching root 1
    zkids = m.children(z)
    wblk = _at(m, 2, axis, 2)                    # root 1's child facing z's children
    bids = fam1 + [b for b in fam2 if b != z] + zkids
    odd = zkids[5]                               # one finer block that does not coarsen
    flags = [NONE if (b == odd and not finer_all_coarse) else COARSE for b in bids]
    new, fo = _plan_both(m, bids, flags)
    if finer_all_coarse:
        # z's children merge back into z; root 1's family merges into root 1;
        # root 2's family misses z (not active before the change)
        assert fo[wblk] == COARSE
        assert new == {1} | (set(fam2) - {z}) | {z}
    else:
        assert fo[wblk] == NONE
        assert new == set(bids)


def test_root_blocks_never_coarsen_and_maximum_level_cannot_refine(oracle):
    m = _mesh(0)
    new, fo = _plan_both(m, [1, 2], [COARSE, COARSE])
    assert new == {1, 2}
    fam = m.children(1)
    deep = m.children(fam[0])                    # level 3 = maximum level
    bids = deep + fam[1:] + [2]
    flags = [REFINE if b == deep[0] else NONE for b in bids]
    with pytest.raises(hip.WxError):
        hip.admesh_plan(m, bids, flags)


def test_plan_reports_the_size_when_the_capacity_is_short():
    import ctypes as C
    m = _mesh(0)
    bids = np.array([1, 2], dtype=np.int32)
    flags = np.array([REFINE, NONE], dtype=np.int32)
    n = C.c_int(0)
    out = np.zeros(4, dtype=np.int32)
    nroot = (C.c_int * 3)(*m.nroot)
    rc = hip.load().wx_admesh_plan(nroot, m.maxlev, 2, bids.ctypes.data_as(hip._ip),
                                   flags.ctypes.data_as(hip._ip), None, 4, C.byref(n),
                                   out.ctypes.data_as(hip._ip))
    assert rc < 0 and n.value == 9


# --------------------------------------------------------------------------
# Hilbert keys (getHilbertKey / HKey3): the reference's keys, and the curve
# property that fixes any table entry: consecutive keys are face neighbours
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["amr_adapt", "briowu"])
def test_hilbert_keys_match_the_reference(name):
    fx = load_golden("hkeys.npz")
    m = W.get(name).mesh()
    keys = hip.hilbert_keys(m, fx["%s_bids" % name])
    assert np.array_equal(keys, fx["%s_keys" % name])


@pytest.mark.parametrize("maxlev", [2, 3, 4, 5])
def test_hilbert_curve_steps_between_face_neighbours(maxlev):
    m = W.Mesh((1, 1, 1), maxlev, (1.0, 1.0, 1.0), 1.0)
    finest = list(range(m.minid[-1], m.maxid[-1] + 1))
    keys = hip.hilbert_keys(m, finest)
    n = len(finest)
    assert sorted(keys.tolist()) == list(range(n))
    pos = np.array([m.coords(b) for b in finest])[np.argsort(keys)]
    steps = np.abs(np.diff(pos, axis=0)).sum(axis=1)
    assert np.all(steps == 1)
    # a coarser block takes the key of its lowest finest-level cell
    for b in range(m.minid[-2], m.maxid[-2] + 1):
        x, y, z = m.coords(b)
        low = m.bid(maxlev, 2 * x - 1, 2 * y - 1, 2 * z - 1)
        assert hip.hilbert_keys(m, [b])[0] == hip.hilbert_keys(m, [low])[0]


def test_oracle_needs_same_level_periodic_partners(oracle):
    """The periodic fill copies the block across a periodic face at the same
    level (the reference has no periodic fill of its own,
    source/boundary.f90:594); a mesh whose periodic faces wrap onto a block
    of another level is refused, the same mesh with free boundaries is not."""
    from conftest import mixed_level_ot
    w, bids, _ = mixed_level_ot()
    with pytest.raises(ValueError):
        oracle.Sim(w, bids)
    w, bids, _ = mixed_level_ot("free")
    oracle.Sim(w, bids)
