"""m3World_Autotune on the MI355X: it times every candidate launch shape,
leaves the worlds as they were (snapshot, restore), and the steps that
follow with the chosen shape stay bit-exact with the reference build."""

from __future__ import annotations

import ctypes as C

import pytest

from tests.helpers import DT, SUBSTEPS, hexh, product_api, ref_api

from maul3d_amd import api, scenes

pytestmark = pytest.mark.gpu


def test_autotune_batch_keeps_worlds_and_parity():
    ref = ref_api()
    gpu = product_api()
    lib = gpu.lib
    n = 6
    wr = [scenes.batch_world(ref, k) for k in range(n)]
    wg = [scenes.batch_world(gpu, k) for k in range(n)]
    ids = (api.WorldId * n)(*wg)
    try:
        for _ in range(20):
            assert lib.m3World_StepMany(ids, n, C.c_float(DT), SUBSTEPS) == n
        before = [gpu.m3World_Hash(w) for w in wg]
        res = api.AutotuneResult()
        got = gpu.m3World_Autotune(ids, n, DT, SUBSTEPS, 3, C.byref(res))
        assert got == 4 and res.candidates == 4
        assert sorted(res.candidateA[:4]) == [64, 128, 256, 512]
        assert res.batchThreads in (64, 128, 256, 512)
        assert all(ms > 0.0 for ms in res.candidateMs[:4])
        assert lib.m3gpu_get_batch_threads() == res.batchThreads
        assert [gpu.m3World_Hash(w) for w in wg] == before, "autotune must leave the worlds as they were"
        for _ in range(25):
            for w in wr:
                ref.m3World_Step(w, DT, SUBSTEPS)
        for i in range(5):
            assert lib.m3World_StepMany(ids, n, C.c_float(DT), SUBSTEPS) == n
        for k in range(n):
            assert hexh(gpu.m3World_Hash(wg[k])) == hexh(ref.m3World_Hash(wr[k])), f"world {k}"
    finally:
        lib.m3gpu_set_batch_threads(0)
        for w in wr:
            ref.m3DestroyWorld(w)
        for w in wg:
            gpu.m3DestroyWorld(w)


def test_autotune_single_world_keeps_world_and_parity():
    ref = ref_api()
    gpu = product_api()
    wr = scenes.rain(ref, count=1200, seed=0xA070, extent=12.0)
    wg = scenes.rain(gpu, count=1200, seed=0xA070, extent=12.0)
    try:
        for _ in range(60):
            ref.m3World_Step(wr, DT, SUBSTEPS)
            gpu.m3World_Step(wg, DT, SUBSTEPS)
        before = gpu.m3World_Hash(wg)
        assert hexh(before) == hexh(ref.m3World_Hash(wr))
        res = api.AutotuneResult()
        got = gpu.m3World_Autotune(C.byref(wg), 1, DT, SUBSTEPS, 2, C.byref(res))
        assert 1 <= got <= 12 and res.candidates == got
        assert res.substepBlocks >= 1 and res.tailLimit in (512, 1024, 2048)
        assert gpu.m3World_Hash(wg) == before, "autotune must leave the world as it was"
        for i in range(1, 31):
            ref.m3World_Step(wr, DT, SUBSTEPS)
            gpu.m3World_Step(wg, DT, SUBSTEPS)
            assert hexh(gpu.m3World_Hash(wg)) == hexh(ref.m3World_Hash(wr)), f"step {i} after tuning"
    finally:
        ref.m3DestroyWorld(wr)
        gpu.m3DestroyWorld(wg)
