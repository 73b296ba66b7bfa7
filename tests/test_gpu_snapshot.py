"""Snapshots in the reference's layout on the MI355X path.

With the tree mirror on from creation, the device world's snapshot bytes
equal the reference's: on the 10k rain at steps 1, 100 and 600 and on the
hull rain with chains at steps 1 and 60 (size and SHA-256 from the reference
build, tests/golden/snapshots.json), and on small scenes against the live
reference.  Snapshots restore across the two libraries in both directions and
both continue to the same hashes.
"""

from __future__ import annotations

import hashlib
import json
import os

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, product_api, ref_api

from maul3d_amd import scenes

pytestmark = pytest.mark.gpu


def _fixture() -> dict:
    with open(os.path.join(ROOT, "tests", "golden", "snapshots.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["rain", "hull_chains"])
def test_snapshot_bytes_equal_reference_fixture(name):
    fx = _fixture()[name]
    gpu = product_api()
    w = scenes.SCENES[name](gpu)
    assert gpu.m3World_EnableTreeMirror(w, True)
    last = max(int(k) for k in fx)
    for step in range(1, last + 1):
        gpu.m3World_Step(w, DT, SUBSTEPS)
        want = fx.get(str(step))
        if want is not None:
            snap = gpu.snapshot(w)
            assert len(snap) == want["bytes"], f"{name} step {step}"
            assert hashlib.sha256(snap).hexdigest() == want["sha256"], f"{name} step {step}"
    gpu.m3DestroyWorld(w)


@pytest.mark.parametrize("name", ["bench_pyramid", "ccd_arena"])
def test_snapshot_bytes_equal_live_reference(name):
    ref = ref_api()
    gpu = product_api()
    a = scenes.SCENES[name](ref)
    b = scenes.SCENES[name](gpu)
    assert gpu.m3World_EnableTreeMirror(b, True)
    assert gpu.snapshot(b) == ref.snapshot(a)
    for step in range(1, 121):
        ref.m3World_Step(a, DT, SUBSTEPS)
        gpu.m3World_Step(b, DT, SUBSTEPS)
        if step in (1, 30, 120):
            assert gpu.snapshot(b) == ref.snapshot(a), f"step {step}"
    ref.m3DestroyWorld(a)
    gpu.m3DestroyWorld(b)


@pytest.mark.parametrize("direction", ["reference_to_device", "device_to_reference"])
def test_snapshot_restores_across_libraries(direction):
    ref = ref_api()
    gpu = product_api()
    src, dst = (ref, gpu) if direction == "reference_to_device" else (gpu, ref)
    a = scenes.rain(src, count=3000)
    for _ in range(80):
        src.m3World_Step(a, DT, SUBSTEPS)
    snap = src.snapshot(a)
    b = scenes.rain(dst, count=3000)
    assert dst.restore(b, snap)
    assert dst.m3World_Hash(b) == src.m3World_Hash(a)
    for step in range(1, 81):
        src.m3World_Step(a, DT, SUBSTEPS)
        dst.m3World_Step(b, DT, SUBSTEPS)
        assert hexh(dst.m3World_Hash(b)) == hexh(src.m3World_Hash(a)), f"step {step}"
    src.m3DestroyWorld(a)
    dst.m3DestroyWorld(b)
