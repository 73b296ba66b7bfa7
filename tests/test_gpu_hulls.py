"""Point-cloud hulls on the device, against the reference.

The hull-rain-with-chains scene (maul3d_amd.scenes.hull_chains: 4000
QuickHull hulls of 8-64 ellipsoid points plus 500 chains of 20 capsule
links on spherical joints with 45-degree cones; 14,501 bodies) is checked
every 10th step for 600 steps against tests/golden/hashes.json.  A dense
200-hull column (hull-on-hull manifolds of up to 124-face hulls) is checked
every step against the live reference build when it is present, and
through a snapshot rollback.
"""

from __future__ import annotations

import pytest

from tests.helpers import DT, SUBSTEPS, golden, hexh, product_api, ref_api

from maul3d_amd import scenes

pytestmark = pytest.mark.gpu


def test_hull_chains_fixture_every_10_steps():
    fx = golden()["scenes"]["hull_chains"]
    gpu = product_api()
    w = scenes.hull_chains(gpu)
    assert hexh(gpu.m3World_Hash(w)) == fx["init"]
    for i in range(1, fx["steps"] + 1):
        gpu.m3World_Step(w, DT, SUBSTEPS)
        want = fx["hashes"].get(str(i))
        if want is not None:
            assert hexh(gpu.m3World_Hash(w)) == want, f"hull_chains diverged at step {i}"
    gpu.m3DestroyWorld(w)


def test_hull_pile_live_every_step():
    ref = ref_api()
    gpu = product_api()
    wr = scenes.hull_pile(ref)
    wg = scenes.hull_pile(gpu)
    assert ref.m3World_Hash(wr) == gpu.m3World_Hash(wg)
    for i in range(300):
        ref.m3World_Step(wr, DT, SUBSTEPS)
        gpu.m3World_Step(wg, DT, SUBSTEPS)
        assert ref.m3World_Hash(wr) == gpu.m3World_Hash(wg), f"step {i + 1}"
    ref.m3DestroyWorld(wr)
    gpu.m3DestroyWorld(wg)


def test_hull_pile_rollback_matches_linear_run():
    gpu = product_api()
    linear = scenes.hull_pile(gpu, hulls=120)
    rolled = scenes.hull_pile(gpu, hulls=120)
    for _ in range(150):
        gpu.m3World_Step(linear, DT, SUBSTEPS)
        gpu.m3World_Step(rolled, DT, SUBSTEPS)
    snap = gpu.snapshot(rolled)
    for _ in range(9):
        gpu.m3World_Step(rolled, DT, SUBSTEPS)
    assert gpu.restore(rolled, snap)
    for i in range(40):
        gpu.m3World_Step(linear, DT, SUBSTEPS)
        gpu.m3World_Step(rolled, DT, SUBSTEPS)
        assert gpu.m3World_Hash(linear) == gpu.m3World_Hash(rolled), f"step {i + 1} after restore"
    gpu.m3DestroyWorld(linear)
    gpu.m3DestroyWorld(rolled)
