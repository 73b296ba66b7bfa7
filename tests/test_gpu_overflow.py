"""The coloring's overflow color on the MI355X path, live against the
reference build: a plank carrying more constraints than there are graph
colors (maul3d_amd.scenes.crowded_plank; the reference reports 17 colors,
the 16 graph colors plus overflow, from about step 30).  Covers the
overflow branches of the coloring, the per-body warm start, the substep
tail (barrier-free small colors, then the serial overflow color) and the
batched block kernel.
"""

from __future__ import annotations

import ctypes as C

import pytest

from tests.helpers import DT, SUBSTEPS, hexh, product_api, ref_api

from maul3d_amd import api, scenes

pytestmark = pytest.mark.gpu

STEPS = 150


def test_overflow_color_single_world_every_step():
    ref = ref_api()
    gpu = product_api()
    wr = scenes.crowded_plank(ref)
    wg = scenes.crowded_plank(gpu)
    overflow_steps = 0
    for i in range(1, STEPS + 1):
        ref.m3World_Step(wr, DT, SUBSTEPS)
        gpu.m3World_Step(wg, DT, SUBSTEPS)
        assert hexh(gpu.m3World_Hash(wg)) == hexh(ref.m3World_Hash(wr)), f"step {i}"
        if ref.m3World_GetCounters(wr).colorCount > 16:
            overflow_steps += 1
    ref.m3DestroyWorld(wr)
    gpu.m3DestroyWorld(wg)
    assert overflow_steps >= 100


def test_overflow_color_batched_every_step():
    ref = ref_api()
    gpu = product_api()
    lib = gpu.lib
    seeds = (0xC0FFEE, 0xFACADE)
    wr = [scenes.crowded_plank(ref, seed=s) for s in seeds]
    wg = [scenes.crowded_plank(gpu, seed=s) for s in seeds]
    ids = (api.WorldId * len(wg))(*wg)
    for i in range(1, STEPS + 1):
        for w in wr:
            ref.m3World_Step(w, DT, SUBSTEPS)
        assert lib.m3World_StepMany(ids, len(wg), C.c_float(DT), SUBSTEPS) == len(wg)
        for k in range(len(wg)):
            assert hexh(gpu.m3World_Hash(wg[k])) == hexh(ref.m3World_Hash(wr[k])), f"world {k} step {i}"
    for w in wr:
        ref.m3DestroyWorld(w)
    for w in wg:
        gpu.m3DestroyWorld(w)
