"""The autotuner's search space (no GPU): launch grid, XCD ranges, curve and
z slabs per block, with slabs offered only where the kernels use them (the
FP32 MHD form without viscosity)."""
import types

from walicxe3d_amd import autotune, hip
from walicxe3d_amd import workloads as W


def fake(name, form):
    return types.SimpleNamespace(w=W.get(name), form=form)


def test_slab_choices_follow_the_kernels():
    assert autotune.slab_choices(fake("ot64", hip.FUSED_FP32)) == (1, 2, 4, 8)
    assert autotune.slab_choices(fake("briowu", hip.FUSED_FP32)) == (1, 2, 4, 8)
    assert autotune.slab_choices(fake("mhd_hlle", hip.FUSED_FP32)) == (1,)     # viscosity: whole columns
    assert autotune.slab_choices(fake("sedov", hip.FUSED_FP32)) == (1,)        # hydro kernels
    assert autotune.slab_choices(fake("ot64", hip.FUSED_FP64)) == (1,)
    assert autotune.slab_choices(fake("ot64", hip.PLAIN_FP64)) == (1,)


def test_candidates_cover_every_dimension():
    c = autotune.candidates(64, (1, 2, 4, 8))
    assert {n for *_, n in c} == {1, 2, 4, 8}
    # persistent grids only below the work-item count: 64 blocks x 8 slabs = 512 items
    assert {g for g, _, _, n in c if n == 8} == {0}
    assert {g for g, _, _, n in c if n == 1} == {0}
    c = autotune.candidates(4096, (1,))
    assert {g for g, *_ in c} == {0, 512, 1024, 2048}
    assert len(c) == 4 * 2 * 2
