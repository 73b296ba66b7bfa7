"""Spherical joints with cone limits on the device, against the reference.

The chain scene (maul3d_amd.scenes.chains: 500 chains of 20 capsule links on
spherical joints with 45-degree cones, anchored to static bodies, plus 4000
falling boxes; 14,501 bodies, 10,000 joints) is checked every 10th step for
600 steps against tests/golden/hashes.json; smaller chain scenes are checked
every step against the live reference build when it is present, through a
snapshot rollback and through batched stepping.
"""

from __future__ import annotations

import ctypes as C

import pytest

from tests.helpers import DT, SUBSTEPS, golden, hexh, product_api, ref_api

from maul3d_amd import api, scenes

pytestmark = pytest.mark.gpu


def test_chains_fixture_every_10_steps():
    fx = golden()["scenes"]["chains"]
    gpu = product_api()
    w = scenes.chains(gpu)
    assert hexh(gpu.m3World_Hash(w)) == fx["init"]
    for i in range(1, fx["steps"] + 1):
        gpu.m3World_Step(w, DT, SUBSTEPS)
        want = fx["hashes"].get(str(i))
        if want is not None:
            assert hexh(gpu.m3World_Hash(w)) == want, f"chains diverged at step {i}"
    gpu.m3DestroyWorld(w)


def test_small_chains_live_every_step():
    ref = ref_api()
    gpu = product_api()
    kw = dict(count=12, links=15, boxes=40, seed=0xBEEF)
    wr = scenes.chains(ref, **kw)
    wg = scenes.chains(gpu, **kw)
    for i in range(150):
        ref.m3World_Step(wr, DT, SUBSTEPS)
        gpu.m3World_Step(wg, DT, SUBSTEPS)
        assert ref.m3World_Hash(wr) == gpu.m3World_Hash(wg), f"step {i + 1}"
    ref.m3DestroyWorld(wr)
    gpu.m3DestroyWorld(wg)


def test_chains_rollback_matches_linear_run():
    gpu = product_api()
    kw = dict(count=20, links=20, boxes=100)
    linear = scenes.chains(gpu, **kw)
    rolled = scenes.chains(gpu, **kw)
    for _ in range(60):
        gpu.m3World_Step(linear, DT, SUBSTEPS)
        gpu.m3World_Step(rolled, DT, SUBSTEPS)
    snap = gpu.snapshot(rolled)
    for _ in range(7):
        gpu.m3World_Step(rolled, DT, SUBSTEPS)
    assert gpu.restore(rolled, snap)
    for i in range(30):
        gpu.m3World_Step(linear, DT, SUBSTEPS)
        gpu.m3World_Step(rolled, DT, SUBSTEPS)
        assert gpu.m3World_Hash(linear) == gpu.m3World_Hash(rolled), f"step {i + 1} after restore"
    gpu.m3DestroyWorld(linear)
    gpu.m3DestroyWorld(rolled)


def test_batched_chain_worlds_equal_single_world_path():
    import maul3d_amd

    gpu = product_api()
    lib = maul3d_amd.library()
    kws = [dict(count=4 + k, links=8 + k, boxes=10 * k, seed=0x100 + k) for k in range(6)]
    seq = [scenes.chains(gpu, **kw) for kw in kws]
    bat = [scenes.chains(gpu, **kw) for kw in kws]
    ids = (api.WorldId * len(bat))(*bat)
    for step in range(60):
        for w in seq:
            gpu.m3World_Step(w, DT, SUBSTEPS)
        assert lib.m3World_StepMany(ids, len(bat), C.c_float(DT), SUBSTEPS) == len(bat)
        for a, b in zip(seq, bat):
            assert gpu.m3World_Hash(a) == gpu.m3World_Hash(b), f"step {step + 1}"
    for w in seq + bat:
        gpu.m3DestroyWorld(w)
