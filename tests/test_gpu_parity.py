"""Parity of the MI355X step with the reference engine.

The bar is bit-exact: the curated world hash (body state, shapes, every
pair key and every 188-byte manifold) after each checked step must equal
the reference's.  Hashes come from the committed fixtures
(tests/golden/hashes.json, written by the reference build) and, where the
reference build is present, from live reference runs on fresh seeds.
"""

from __future__ import annotations

import ctypes as C

import pytest

from tests.helpers import DT, SUBSTEPS, golden, hexh, product_api, ref_api

from maul3d_amd import scenes

pytestmark = pytest.mark.gpu


def _run_against_fixture(name: str):
    fx = golden()["scenes"][name]
    gpu = product_api()
    w = scenes.SCENES[name](gpu)
    assert hexh(gpu.m3World_Hash(w)) == fx["init"]
    checked = 0
    for i in range(1, fx["steps"] + 1):
        gpu.m3World_Step(w, DT, SUBSTEPS)
        want = fx["hashes"].get(str(i))
        if want is not None:
            got = hexh(gpu.m3World_Hash(w))
            assert got == want, f"{name}: world hash diverged at step {i}"
            checked += 1
    gpu.m3DestroyWorld(w)
    return checked


def test_sphere_pyramid_pin_every_step():
    """bench/pins.txt:1 scene; every one of 300 steps checked."""
    assert _run_against_fixture("bench_pyramid") == 300
    assert golden()["scenes"]["bench_pyramid"]["hashes"]["300"] == "975f319f2f44cc5d"


def test_box_pyramid_every_step():
    """Classic 20-row, 210-box pyramid: every step of 300."""
    assert _run_against_fixture("box_pyramid") == 300


def test_smoke10k_pin():
    """bench/pins.txt:7 scene (spheres, boxes, capsules, rolling resistance)."""
    assert _run_against_fixture("smoke10k") == 15
    assert golden()["scenes"]["smoke10k"]["hashes"]["150"] == "bc84e0074315a0b3"


def test_rain10k_600_steps():
    """The 10k-body mixed rain: every 10th step of 600."""
    assert _run_against_fixture("rain") == 60


def test_fresh_seed_live_against_reference():
    """A seed no fixture covers, checked live against the reference build."""
    ref = ref_api()
    gpu = product_api()
    wr = scenes.rain(ref, count=1200, seed=0xBADC0DE, extent=12.0)
    wg = scenes.rain(gpu, count=1200, seed=0xBADC0DE, extent=12.0)
    for i in range(1, 241):
        ref.m3World_Step(wr, DT, SUBSTEPS)
        gpu.m3World_Step(wg, DT, SUBSTEPS)
        if i % 20 == 0:
            assert hexh(gpu.m3World_Hash(wg)) == hexh(ref.m3World_Hash(wr)), f"step {i}"
    ref.m3DestroyWorld(wr)
    gpu.m3DestroyWorld(wg)


def test_twin_worlds_are_identical():
    gpu = product_api()
    a = scenes.rain(gpu, count=800, seed=77, extent=8.0)
    b = scenes.rain(gpu, count=800, seed=77, extent=8.0)
    for _ in range(90):
        gpu.m3World_Step(a, DT, SUBSTEPS)
        gpu.m3World_Step(b, DT, SUBSTEPS)
    assert gpu.m3World_Hash(a) == gpu.m3World_Hash(b)
    assert gpu.snapshot(a) == gpu.snapshot(b)


def test_rollback_restore_resimulate_matches_linear_run():
    """Snapshot at step 100, restore into another world, resimulate 8 steps:
    the hash equals the linear run at step 108 and the reference's."""
    ref = ref_api()
    gpu = product_api()
    wl = scenes.rain(gpu, count=1500, seed=4242, extent=10.0)
    wr = scenes.rain(ref, count=1500, seed=4242, extent=10.0)
    for _ in range(100):
        gpu.m3World_Step(wl, DT, SUBSTEPS)
        ref.m3World_Step(wr, DT, SUBSTEPS)
    snap = gpu.snapshot(wl)
    for _ in range(8):
        gpu.m3World_Step(wl, DT, SUBSTEPS)
        ref.m3World_Step(wr, DT, SUBSTEPS)
    linear = gpu.m3World_Hash(wl)
    assert hexh(linear) == hexh(ref.m3World_Hash(wr))
    wb = scenes.rain(gpu, count=1500, seed=1, extent=10.0)  # same capacities
    assert gpu.restore(wb, snap)
    for _ in range(8):
        gpu.m3World_Step(wb, DT, SUBSTEPS)
    assert gpu.m3World_Hash(wb) == linear


def test_toss_worlds_match_fixture():
    """64 differentiable-toss worlds (32 boxes each, 120 steps)."""
    fx = golden()["toss_120"]
    gpu = product_api()
    worlds = [scenes.toss(gpu, seed=0x7055 + k) for k in range(64)]
    for _ in range(120):
        for w in worlds:
            gpu.m3World_Step(w, DT, SUBSTEPS)
    for k, w in enumerate(worlds):
        assert hexh(gpu.m3World_Hash(w)) == fx[str(k)], f"toss world {k}"
        gpu.m3DestroyWorld(w)


def test_step_many_equals_sequential():
    import maul3d_amd
    from maul3d_amd import api

    gpu = product_api()
    lib = maul3d_amd.library()
    seq = [scenes.toss(gpu, seed=900 + k) for k in range(6)]
    bat = [scenes.toss(gpu, seed=900 + k) for k in range(6)]
    ids = (api.WorldId * 6)(*bat)
    for _ in range(30):
        for w in seq:
            gpu.m3World_Step(w, DT, SUBSTEPS)
        assert lib.m3World_StepMany(ids, 6, C.c_float(DT), SUBSTEPS) == 6
    for a, b in zip(seq, bat):
        assert gpu.m3World_Hash(a) == gpu.m3World_Hash(b)
