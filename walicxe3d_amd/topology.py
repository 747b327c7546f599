"""Face-neighbour classification of an active block list (host side).

Same rules as the reference's `neighbors` (source/amr.f90:539-631): across a
face a block sees a same-level block, its coarser cousin (the father of the
same-level position), four finer blocks (the children of the same-level
position that touch the face, in the order of source/amr.f90:598-628), or the
domain boundary.  Only used for workload validation in Python; the GPU library
builds its own table in C++ (walicxe3d_amd/csrc/topology.cpp).
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

SAME, COARSER, FINER, BOUNDARY = 1, 2, 3, 4
LEFT, RIGHT, FRONT, BACK, BOTTOM, TOP = 1, 2, 3, 4, 5, 6

# children (1-based child numbers) adjacent to the face of the *neighbour*
_FINER_CHILDREN = {LEFT: (2, 4, 6, 8), RIGHT: (1, 3, 5, 7), FRONT: (3, 4, 7, 8),
                   BACK: (1, 2, 5, 6), BOTTOM: (5, 6, 7, 8), TOP: (1, 2, 3, 4)}


def same_level_neighbour(mesh, bid: int, face: int, periodic: Sequence[bool] = (False,) * 6) -> int:
    lev = mesh.level(bid)
    x, y, z = mesh.coords(bid)
    nx, ny, nz = mesh.nb(lev)
    c = [x, y, z]
    n = [nx, ny, nz]
    d = (face - 1) // 2
    step = -1 if face % 2 == 1 else 1
    c[d] += step
    if c[d] < 1 or c[d] > n[d]:
        if not periodic[face - 1]:
            return -1
        c[d] = n[d] if c[d] < 1 else 1
    return mesh.bid(lev, c[0], c[1], c[2])


def face_table(mesh, active: Sequence[int], periodic: Sequence[bool] = (False,) * 6
               ) -> Dict[Tuple[int, int], Tuple[int, List[int]]]:
    """{(bid, face): (type, [neighbour bids])}.  Raises if the 2:1 rule is
    broken (a face neighbour two levels apart)."""
    aset = set(active)
    out = {}
    for b in active:
        for f in range(1, 7):
            nid = same_level_neighbour(mesh, b, f, periodic)
            if nid == -1:
                out[(b, f)] = (BOUNDARY, [])
                continue
            if nid in aset:
                out[(b, f)] = (SAME, [nid])
                continue
            fid = mesh.father(nid) if mesh.level(nid) > 1 else -1
            if fid != -1 and fid in aset:
                out[(b, f)] = (COARSER, [fid])
                continue
            if mesh.level(nid) >= mesh.maxlev:
                raise ValueError("no neighbour for block %d face %d" % (b, f))
            ch = mesh.children(nid)
            sel = [ch[i - 1] for i in _FINER_CHILDREN[f]]
            for s in sel:
                if s not in aset:
                    raise ValueError("2:1 rule broken at block %d face %d" % (b, f))
            out[(b, f)] = (FINER, sel)
    return out
