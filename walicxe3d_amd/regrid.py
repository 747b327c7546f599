"""Host-side regrid: carry a block state from one active block list to
another on the same octree (what the reference's refineBlock / coarsenBlock,
source/amr.f90:284-384 and :1430-1650, do to the data when admesh changes the
topology).

A block present in both lists is copied.  A new block whose ancestor was
active takes the ancestor's cells by injection (each fine cell gets the value
of the coarse cell that contains it, the reference's octant copy).  A new
block whose descendants were active gets the 2x2x2 average of each cell
group, recursively.  Both directions conserve the volume integral of every
conserved variable exactly up to the rounding of the averages.

The GPU library takes the new list with `wx_set_mesh` (which rebuilds the
face table and recaptures the step graph) and the new state with
`wx_upload_state`.
"""
from __future__ import annotations

from typing import Dict, Sequence

import numpy as np

NC = 16


def _inject(parent: np.ndarray, octant_path, ) -> np.ndarray:
    """Cells of a descendant of `parent` ([nv][k][j][i]) reached by the list
    of octants (sx, sy, sz), by piecewise-constant injection."""
    out = parent
    for sx, sy, sz in octant_path:
        h = NC // 2
        sub = out[:, sz * h:(sz + 1) * h, sy * h:(sy + 1) * h, sx * h:(sx + 1) * h]
        out = np.repeat(np.repeat(np.repeat(sub, 2, axis=1), 2, axis=2), 2, axis=3)
    return out


def _restrict(children: Sequence[np.ndarray]) -> np.ndarray:
    """Parent block from its 8 children (x fastest, then y, then z), each cell
    the mean of the 2x2x2 fine cells it covers."""
    nv = children[0].shape[0]
    out = np.empty((nv, NC, NC, NC))
    h = NC // 2
    for c, u in enumerate(children):
        sx, sy, sz = c & 1, (c >> 1) & 1, (c >> 2) & 1
        g = u.reshape(nv, h, 2, h, 2, h, 2)
        acc = np.zeros((nv, h, h, h))
        for dk in (0, 1):              # z outer, x inner: the device's order
            for dj in (0, 1):
                for di in (0, 1):
                    acc = acc + g[:, :, dk, :, dj, :, di]
        avg = acc / 8.0
        out[:, sz * h:(sz + 1) * h, sy * h:(sy + 1) * h, sx * h:(sx + 1) * h] = avg
    return out


def remap(mesh, old_bids: Sequence[int], U_old: np.ndarray, new_bids: Sequence[int]) -> np.ndarray:
    """State on `new_bids` ([nblk_new][nv][16][16][16]) from `U_old` on
    `old_bids`.  Raises if a new block is covered neither by an old block, an
    old ancestor nor old descendants."""
    old: Dict[int, int] = {int(b): n for n, b in enumerate(old_bids)}
    new_bids = [int(b) for b in new_bids]
    nv = U_old.shape[1]
    U_new = np.empty((len(new_bids), nv, NC, NC, NC))
    same_new = [n for n, b in enumerate(new_bids) if b in old]
    if same_new:
        U_new[same_new] = U_old[[old[new_bids[n]] for n in same_new]]
    cache: Dict[int, np.ndarray] = {}

    def from_descendants(b: int) -> np.ndarray:
        if b in old:
            return U_old[old[b]]
        if b in cache:
            return cache[b]
        if mesh.level(b) >= mesh.maxlev:
            raise ValueError("block %d: no active ancestor or descendants" % b)
        u = _restrict([from_descendants(c) for c in mesh.children(b)])
        cache[b] = u
        return u

    for n, b in enumerate(new_bids):
        if b in old:
            continue
        # nearest active ancestor
        path = []
        a = b
        while a != -1 and a not in old:
            f = mesh.father(a)
            if f == -1:
                a = -1
                break
            x, y, z = mesh.coords(a)
            path.append(((x - 1) & 1, (y - 1) & 1, (z - 1) & 1))
            a = f
        if a != -1:
            U_new[n] = _inject(U_old[old[a]], reversed(path))
        else:
            U_new[n] = from_descendants(b)
    return U_new
