"""Host regrid (walicxe3d_amd/regrid.py): copy, octant injection and 2x2x2
restriction between two block lists of the same octree."""
import dataclasses

import numpy as np

from walicxe3d_amd import regrid
from walicxe3d_amd import workloads as W


def _integral(mesh, bids, U):
    vol = np.array([mesh.dx[0][mesh.level(int(b)) - 1] * mesh.dx[1][mesh.level(int(b)) - 1] *
                    mesh.dx[2][mesh.level(int(b)) - 1] for b in bids])
    return np.einsum("b,bv->v", vol, U.sum(axis=(2, 3, 4)))


def _meshes():
    wa = W.get("winds4")
    wb = dataclasses.replace(wa, refine_radii=(0.3, 0.15, 0.08, 0.04))
    a = [int(b) for b in wa.blocks()]
    b = [int(x) for x in wb.blocks()]
    return wa, a, b


def test_identity():
    w, a, _ = _meshes()
    U = np.random.default_rng(1).random((len(a), 5, 16, 16, 16))
    assert np.array_equal(regrid.remap(w.mesh(), a, U, a), U)


def test_refine_then_coarsen_returns_the_state_and_conserves():
    w, a, b = _meshes()
    m = w.mesh()
    assert set(a) != set(b)
    _, _, Ua = W.initial_state(w)
    Ub = regrid.remap(m, a, Ua, b)
    Ua2 = regrid.remap(m, b, Ub, a)
    assert np.allclose(Ua2, Ua, rtol=1e-15, atol=0.0)
    ia, ib = _integral(m, a, Ua), _integral(m, b, Ub)
    assert np.allclose(ia, ib, rtol=1e-13, atol=0.0)


def test_injection_and_restriction_by_hand():
    w = W.get("sedov")                  # one root, 3 levels
    m = w.mesh()
    parent = m.bid(2, 1, 1, 1)          # a level-2 block
    kids = m.children(parent)
    U = np.arange(5 * 16 ** 3, dtype=np.float64).reshape(1, 5, 16, 16, 16)
    Uk = regrid.remap(m, [parent], U, kids)
    # child with octant (1, 0, 1): x high half, y low, z high
    c = kids[1 + 4]
    cell = Uk[kids.index(c)][:, 3, 5, 7]          # k=3, j=5, i=7 of the child
    assert np.array_equal(cell, U[0][:, 8 + 3 // 2, 0 + 5 // 2, 8 + 7 // 2])
    back = regrid.remap(m, kids, Uk, [parent])
    assert np.array_equal(back, U)
