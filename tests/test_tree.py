"""The broadphase tree and continuous collision against tree shapes.

The device builds no tree; the host keeps the reference's tree
(src/dynamic_tree.c) and replays each step's leaf moves into it
(src/broad_phase.c:505-539), and the continuous pass meets its tree
candidates in that tree's depth-first order (src/solver.c:3081-3278,
3349-3430).  Checked on the CPU twin of the step (world.c, m3_tree.c and the
serial step, the functions the device kernels share) against the live
reference every step; the GPU test repeats it on the HIP path.

The tree itself is pinned by the reference's tree height
(m3World_GetCounters) and by structural invariants: every internal box is the
union of its children's, heights are one more than the taller child, and the
free chain plus the reachable nodes cover the pool exactly once.
"""

from __future__ import annotations

import ctypes as C
import os

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, ref_api

from maul3d_amd import api, scenes

TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")


def _twin():
    if not os.path.exists(TWIN32):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    return api.Api(C.CDLL(TWIN32))


def _moved(m3, a, b) -> int:
    """Bodies whose position differs between two worlds of the same scene."""
    n = m3.m3World_GetCounters(a).bodyCount
    out = 0
    for s in range(n):
        p = m3.m3Body_GetPosition(api.BodyId(s + 1, a.index1 - 1, 0))
        q = m3.m3Body_GetPosition(api.BodyId(s + 1, b.index1 - 1, 0))
        out += (p.x, p.y, p.z) != (q.x, q.y, q.z)
    return out


@pytest.mark.parametrize("bullets", [0, 3])
def test_cpu_twin_continuous_against_tree_shapes_matches_reference_every_step(bullets):
    ref = ref_api()
    tw = _twin()
    a = scenes.ccd_arena(ref, bullets=bullets)
    b = scenes.ccd_arena(tw, bullets=bullets)
    assert hexh(tw.m3World_Hash(b)) == hexh(ref.m3World_Hash(a))
    for step in range(1, 241):
        ref.m3World_Step(a, DT, SUBSTEPS)
        tw.m3World_Step(b, DT, SUBSTEPS)
        assert hexh(tw.m3World_Hash(b)) == hexh(ref.m3World_Hash(a)), f"step {step}"
        if step % 40 == 0:
            # The tree the host replays has the reference's height.
            assert tw.m3World_GetCounters(b).treeHeight == ref.m3World_GetCounters(a).treeHeight
    ref.m3DestroyWorld(a)
    tw.m3DestroyWorld(b)


def test_continuous_pass_acts_on_the_arena():
    """The control: with continuous collision off the reference's bodies
    end up elsewhere (the scene has no plane, so every pull-back is against
    a tree shape)."""
    ref = ref_api()
    a = scenes.ccd_arena(ref)
    b = scenes.ccd_arena(ref)
    ref.m3World_EnableContinuous(b, False)
    for _ in range(30):
        ref.m3World_Step(a, DT, SUBSTEPS)
        ref.m3World_Step(b, DT, SUBSTEPS)
    assert _moved(ref, a, b) > 10
    ref.m3DestroyWorld(a)
    ref.m3DestroyWorld(b)
