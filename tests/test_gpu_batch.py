"""Batched stepping: many independent worlds share every launch of the step.

The batched-rollback set (maul3d_amd.scenes.batch_world, 256 worlds of 50 to
2000 bodies: pyramids, box stacks and sphere/box rain) is warmed for 100
steps with m3World_StepMany, snapshotted, restored and resimulated for 8
more; every world's hash after the warm-up and after each resimulated step
must equal the reference engine's (tests/golden/batch_rollback.json, written
by the reference build world by world).
"""

from __future__ import annotations

import ctypes as C
import json
import os

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, product_api

from maul3d_amd import api, scenes

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(ROOT, "tests", "golden", "batch_rollback.json")


def _fixture() -> dict:
    with open(FIXTURE) as f:
        return json.load(f)


def _step_many(lib, worlds) -> None:
    ids = (api.WorldId * len(worlds))(*worlds)
    assert lib.m3World_StepMany(ids, len(worlds), C.c_float(DT), SUBSTEPS) == len(worlds)


def test_batch_rollback_all_worlds():
    import maul3d_amd

    fx = _fixture()
    gpu = product_api()
    lib = maul3d_amd.library()
    n = len(fx["worlds"])
    worlds = [scenes.batch_world(gpu, k) for k in range(n)]
    for _ in range(fx["warm"]):
        _step_many(lib, worlds)
    snaps = []
    for k, w in enumerate(worlds):
        assert hexh(gpu.m3World_Hash(w)) == fx["worlds"][str(k)]["warm"], f"world {k} after warm-up"
        snaps.append(gpu.snapshot(w))
    # Disturb every world, then roll back.
    for _ in range(3):
        _step_many(lib, worlds)
    for w, snap in zip(worlds, snaps):
        assert gpu.restore(w, snap)
    for i in range(fx["resim"]):
        _step_many(lib, worlds)
        for k, w in enumerate(worlds):
            want = fx["worlds"][str(k)]["resim"][i]
            assert hexh(gpu.m3World_Hash(w)) == want, f"world {k} resim step {i + 1}"
    for w in worlds:
        gpu.m3DestroyWorld(w)


def test_batch_equals_single_world_path():
    """Uneven worlds, batched and stepped one by one, hash alike every step."""
    import maul3d_amd

    gpu = product_api()
    lib = maul3d_amd.library()
    picks = [3, 17, 40, 41, 99]
    seq = [scenes.batch_world(gpu, k) for k in picks]
    bat = [scenes.batch_world(gpu, k) for k in picks]
    for step in range(40):
        for w in seq:
            gpu.m3World_Step(w, DT, SUBSTEPS)
        _step_many(lib, bat)
        for a, b in zip(seq, bat):
            assert gpu.m3World_Hash(a) == gpu.m3World_Hash(b), f"step {step + 1}"
    for w in seq + bat:
        gpu.m3DestroyWorld(w)


def test_batch_interleaved_with_single_steps():
    """A world stepped alone between batched calls stays in order."""
    import maul3d_amd

    gpu = product_api()
    lib = maul3d_amd.library()
    ref_worlds = [scenes.toss(gpu, seed=0x7055 + k) for k in range(4)]
    mixed = [scenes.toss(gpu, seed=0x7055 + k) for k in range(4)]
    for step in range(24):
        for w in ref_worlds:
            gpu.m3World_Step(w, DT, SUBSTEPS)
        if step % 5 == 2:
            for w in mixed:
                gpu.m3World_Step(w, DT, SUBSTEPS)
        else:
            _step_many(lib, mixed)
        if step % 7 == 0:
            gpu.m3World_Hash(mixed[1])  # a read between batched steps
    for a, b in zip(ref_worlds, mixed):
        assert gpu.m3World_Hash(a) == gpu.m3World_Hash(b)
    for w in ref_worlds + mixed:
        gpu.m3DestroyWorld(w)
