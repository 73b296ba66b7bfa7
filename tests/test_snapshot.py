"""Snapshots in the reference's layout (src/snapshot.c, format 54).

With the tree mirror on from creation, this build's snapshot bytes equal the
reference's after the same steps: header, every body, shape, joint and pair
block, the broadphase tree's node array and the proxy ids, and the pair table
past the live count.  A snapshot from either library restores in the other
and both continue to the same hashes.  Restore is atomic: a snapshot that
is truncated, carries a malformed tree, or holds state of a subsystem this
build does not simulate is refused and the world is left as it was.

Checked on the CPU twin (world.c and m3_tree.c as the product links them)
against the live reference; tests/test_gpu_snapshot.py repeats it on the
device.
"""

from __future__ import annotations

import ctypes as C
import os
import struct

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, ref_api

from maul3d_amd import api, scenes

TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")
HEADER = 192


def _twin():
    if not os.path.exists(TWIN32):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    return api.Api(C.CDLL(TWIN32))


def _mirror(m3, world) -> None:
    assert m3.m3World_EnableTreeMirror(world, True)


SCENES = {
    "bench_pyramid": scenes.bench_pyramid,
    "rain1500": lambda m3: scenes.rain(m3, count=1500),
    "ccd_arena": scenes.ccd_arena,
    "hulljam": scenes.hulljam,
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_snapshot_bytes_equal_reference(name):
    ref = ref_api()
    tw = _twin()
    a = SCENES[name](ref)
    b = SCENES[name](tw)
    _mirror(tw, b)
    assert tw.snapshot(b) == ref.snapshot(a)
    for step in range(1, 121):
        ref.m3World_Step(a, DT, SUBSTEPS)
        tw.m3World_Step(b, DT, SUBSTEPS)
        if step in (1, 30, 120):
            assert tw.snapshot(b) == ref.snapshot(a), f"step {step}"
    ref.m3DestroyWorld(a)
    tw.m3DestroyWorld(b)


@pytest.mark.parametrize("direction", ["reference_to_build", "build_to_reference"])
def test_snapshot_restores_across_libraries(direction):
    """Snapshot at step 60 on one library, restore into a fresh world of the
    other, step both 60 more: equal hashes every step."""
    ref = ref_api()
    tw = _twin()
    src, dst = (ref, tw) if direction == "reference_to_build" else (tw, ref)
    a = scenes.rain(src, count=1500)
    if src is tw:
        _mirror(tw, a)
    for _ in range(60):
        src.m3World_Step(a, DT, SUBSTEPS)
    snap = src.snapshot(a)
    b = scenes.rain(dst, count=1500)
    assert dst.restore(b, snap)
    assert dst.m3World_Hash(b) == src.m3World_Hash(a)
    for step in range(1, 61):
        src.m3World_Step(a, DT, SUBSTEPS)
        dst.m3World_Step(b, DT, SUBSTEPS)
        assert hexh(dst.m3World_Hash(b)) == hexh(src.m3World_Hash(a)), f"step {step}"
    src.m3DestroyWorld(a)
    dst.m3DestroyWorld(b)


def test_snapshot_without_mirror_restores_and_continues():
    """Without the mirror the tree block is the canonical rebuild of the
    current leaves and the table tail is zero: a different byte string that
    still restores in the reference and continues on its trajectory (the
    pair set does not depend on the tree's shape)."""
    ref = ref_api()
    tw = _twin()
    a = scenes.rain(ref, count=1500)
    b = scenes.rain(tw, count=1500)
    for _ in range(40):
        ref.m3World_Step(a, DT, SUBSTEPS)
        tw.m3World_Step(b, DT, SUBSTEPS)
    assert not tw.m3World_IsTreeMirrorCurrent(b)
    snap = tw.snapshot(b)
    assert tw.m3World_IsTreeMirrorCurrent(b)  # rebuilt for the snapshot
    c = scenes.rain(ref, count=1500)
    assert ref.restore(c, snap)
    for step in range(1, 41):
        ref.m3World_Step(a, DT, SUBSTEPS)
        ref.m3World_Step(c, DT, SUBSTEPS)
        assert ref.m3World_Hash(c) == ref.m3World_Hash(a), f"step {step}"
    for w in (a, c):
        ref.m3DestroyWorld(w)
    tw.m3DestroyWorld(b)


def _corrupt(snap: bytes, offset: int, data: bytes) -> bytes:
    return snap[:offset] + data + snap[offset + len(data) :]


def test_restore_refuses_hostile_snapshots_and_leaves_world():
    ref = ref_api()
    tw = _twin()
    a = scenes.bench_pyramid(ref)
    for _ in range(20):
        ref.m3World_Step(a, DT, SUBSTEPS)
    good = ref.snapshot(a)
    b = scenes.bench_pyramid(tw)
    tw.m3World_Step(b, DT, SUBSTEPS)
    before = tw.m3World_Hash(b)
    hdr = struct.unpack_from("<IIQQ", good, 0)
    tree_root_at = 4 + 4 + 8 + 8 + 4 * 11  # header field treeRoot (src/snapshot.c:84-100)
    cases = {
        "truncated": good[:-1],
        "longer": good + b"\0",
        "magic": _corrupt(good, 0, b"XXXX"),
        "tree root out of range": _corrupt(good, tree_root_at, struct.pack("<i", 1 << 20)),
        # the world scalars follow 291 bytes per body slot; wind speed is a
        # scalar this build does not simulate
        "wind": _corrupt(good, HEADER + 291 * 256 + 42, struct.pack("<f", 3.0)),
    }
    assert hdr[0] == 0x4D33534E
    for label, bad in cases.items():
        assert not tw.restore(b, bad), label
        assert tw.m3World_Hash(b) == before, label
    # A tree whose root points at a free node.
    root = struct.unpack_from("<i", good, tree_root_at)[0]
    free_list = struct.unpack_from("<i", good, tree_root_at + 4)[0]
    assert root >= 0 and free_list >= 0
    bad = _corrupt(good, tree_root_at, struct.pack("<i", free_list))
    assert not tw.restore(b, bad)
    assert tw.m3World_Hash(b) == before
    assert tw.restore(b, good)
    assert tw.m3World_Hash(b) == ref.m3World_Hash(a)
    ref.m3DestroyWorld(a)
    tw.m3DestroyWorld(b)
