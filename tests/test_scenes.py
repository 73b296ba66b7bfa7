"""Scene-level parity on the CPU: the reference's own bench pins
(bench/pins.txt, committed as tests/golden/pins.json) reproduced by this
repo's scene builders on the reference build and on the CPU twin of this
build's step, and the restitution scene stepped against the live reference.

The twin is world.c (the file the product links) over the serial step of
maul3d_amd/csrc/m3_serial.h, which calls the same stage functions as the
device kernels; the GPU tests repeat these checks on the HIP path."""

from __future__ import annotations

import ctypes as C
import json
import os
import re
import subprocess

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, ref_api

from maul3d_amd import api, scenes

TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")
BENCH = os.path.join(ROOT, "oracle", "_ref", "maul3d_bench")


def pins() -> dict:
    with open(os.path.join(ROOT, "tests", "golden", "pins.json")) as f:
        return json.load(f)["pins"]


def _twin():
    if not os.path.exists(TWIN32):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    return api.Api(C.CDLL(TWIN32))


def _run(m3, world, steps: int) -> str:
    for _ in range(steps):
        m3.m3World_Step(world, DT, SUBSTEPS)
    h = hexh(m3.m3World_Hash(world))
    m3.m3DestroyWorld(world)
    return h


def test_reference_bench_binary_reproduces_every_pin():
    """The reference's bench program, built from its sources by
    oracle/Makefile, prints the seven pinned hashes: the build of the checker
    is the reference's own (bench/bench_main.c:693-711, 675-690)."""
    if not os.path.exists(BENCH):
        pytest.skip("oracle/_ref/maul3d_bench not built")
    out = subprocess.run([BENCH], capture_output=True, text=True, timeout=1800, cwd="/tmp").stdout
    got = {m.group(1): m.group(2) for m in re.finditer(r"M3_BENCH (\w+)\s+steps=\d+ .*?hash=([0-9a-f]{16})", out)}
    assert got == pins()


def test_hulljam_scene_builder_reproduces_pin_on_reference():
    assert _run(ref_api(), scenes.hulljam(ref_api()), 300) == pins()["hulljam"]


def test_cpu_twin_hulljam_reproduces_pin():
    tw = _twin()
    assert _run(tw, scenes.hulljam(tw), 300) == pins()["hulljam"]


def test_cpu_twin_restitution_scene_matches_reference_every_step():
    ref = ref_api()
    tw = _twin()
    a = scenes.bounce(ref)
    b = scenes.bounce(tw)
    for step in range(1, 241):
        ref.m3World_Step(a, DT, SUBSTEPS)
        tw.m3World_Step(b, DT, SUBSTEPS)
        assert hexh(tw.m3World_Hash(b)) == hexh(ref.m3World_Hash(a)), f"step {step}"
    ref.m3DestroyWorld(a)
    tw.m3DestroyWorld(b)


def test_restitution_changes_the_bounce_trajectory():
    """The control: the same draws with every coefficient 0 diverge from the
    restitution scene on the reference, so the pass is exercised."""
    ref = ref_api()
    a = scenes.bounce(ref)
    z = scenes.bounce(ref, restitution=False)
    assert ref.m3World_Hash(a) != ref.m3World_Hash(z)  # shape records differ
    ups = 0
    for _ in range(90):
        ref.m3World_Step(a, DT, SUBSTEPS)
        ref.m3World_Step(z, DT, SUBSTEPS)
    c = ref.m3World_GetCounters(a)
    for b in range(1, c.bodyCount):
        bid = api.BodyId(b + 1, a.index1 - 1, 0)
        v = ref.m3Body_GetLinearVelocity(bid)
        ups += v.y > 1.0
    assert ups > 0
    ref.m3DestroyWorld(a)
    ref.m3DestroyWorld(z)
