"""CPU-side tests of the product library: exports, host-parity creation,
hashing, snapshot validation and the API's refusal paths.  Nothing here
steps a world, so no GPU is touched."""

from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from tests.helpers import ROOT, hexh, product_api, ref_api

from maul3d_amd import api, scenes


def _declared(header: str, macro: str) -> list[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    names = re.findall(macro + r"[^;]*?\b(m3\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    import maul3d_amd

    lib = maul3d_amd.library()
    declared = _declared("maul3d_amd.h", r"M3_API") + _declared("m3gpu.h", r"M3GPU_API")
    assert len(declared) > 30
    missing = [n for n in declared if not hasattr(lib, n)]
    assert missing == []


@pytest.mark.parametrize("scene", ["bench_pyramid", "box_pyramid", "rain", "smoke10k", "sweepers"])
def test_initial_world_hash_matches_reference(scene):
    """Creation-time parity: ids, mass properties, box hulls, shape records
    and the curated hash all agree with the reference before any step."""
    ref = ref_api()
    gpu = product_api()
    kw = {"count": 1500} if scene == "rain" else {}
    wr = scenes.SCENES[scene](ref, **kw)
    wg = scenes.SCENES[scene](gpu, **kw)
    assert hexh(gpu.m3World_Hash(wg)) == hexh(ref.m3World_Hash(wr))
    cr = ref.m3World_GetCounters(wr)
    cg = gpu.m3World_GetCounters(wg)
    assert (cg.bodyCount, cg.shapeCount, cg.hullCount) == (cr.bodyCount, cr.shapeCount, cr.hullCount)
    ref.m3DestroyWorld(wr)
    gpu.m3DestroyWorld(wg)


def test_body_readback_matches_reference():
    ref = ref_api()
    gpu = product_api()
    wr = scenes.rain(ref, count=200)
    wg = scenes.rain(gpu, count=200)
    for k in range(1, 201, 37):
        br = api.BodyId(k + 1, wr.index1 - 1, 0)
        bg = api.BodyId(k + 1, wg.index1 - 1, 0)
        pr, pg = ref.m3Body_GetPosition(br), gpu.m3Body_GetPosition(bg)
        assert (pr.x, pr.y, pr.z) == (pg.x, pg.y, pg.z)
        qr, qg = ref.m3Body_GetRotation(br), gpu.m3Body_GetRotation(bg)
        assert (qr.x, qr.y, qr.z, qr.w) == (qg.x, qg.y, qg.z, qg.w)
    ref.m3DestroyWorld(wr)
    gpu.m3DestroyWorld(wg)


def test_snapshot_roundtrip_before_stepping():
    gpu = product_api()
    w1 = scenes.box_pyramid(gpu, rows=6)
    snap = gpu.snapshot(w1)
    w2 = scenes.box_pyramid(gpu, rows=3)  # same capacities, different content
    assert gpu.m3World_Hash(w2) != gpu.m3World_Hash(w1)
    assert gpu.restore(w2, snap)
    assert gpu.m3World_Hash(w2) == gpu.m3World_Hash(w1)
    assert gpu.snapshot(w2) == snap
    gpu.m3DestroyWorld(w1)
    gpu.m3DestroyWorld(w2)


def test_restore_is_atomic_on_bad_bytes():
    gpu = product_api()
    w1 = scenes.box_pyramid(gpu, rows=5)
    w2 = scenes.box_pyramid(gpu, rows=4)
    before = gpu.m3World_Hash(w2)
    snap = bytearray(gpu.snapshot(w1))
    for pos in (0, 9, len(snap) // 2, len(snap) - 1):
        bad = bytearray(snap)
        bad[pos] ^= 0x5A
        assert not gpu.restore(w2, bytes(bad))
        assert gpu.m3World_Hash(w2) == before
    assert not gpu.restore(w2, bytes(snap[:-1]))
    assert gpu.m3World_Hash(w2) == before
    # Capacity mismatch is refused too.
    w3 = scenes.rain(gpu, count=10)
    assert not gpu.restore(w3, bytes(snap))
    for w in (w1, w2, w3):
        gpu.m3DestroyWorld(w)


def test_refusals_match_reference_contract():
    gpu = product_api()
    bad = api.WorldId(0, 0)
    assert gpu.m3World_Hash(bad) == 0
    assert gpu.m3World_SnapshotSize(bad) == -1
    d = gpu.m3DefaultWorldDef()
    d.internalValue ^= 1
    assert gpu.m3CreateWorld(C.byref(d)).index1 == 0
    w = scenes.box_pyramid(gpu, rows=2)
    h0 = gpu.m3World_Hash(w)
    # dt <= 0 and substeps outside [1, 256] are silent no-ops (never reach the device).
    gpu.m3World_Step(w, 0.0, 4)
    gpu.m3World_Step(w, -1.0, 4)
    gpu.m3World_Step(w, 1.0 / 60.0, 0)
    gpu.m3World_Step(w, 1.0 / 60.0, 257)
    assert gpu.m3World_Hash(w) == h0
    bd = gpu.m3DefaultBodyDef()
    bd.rotation = api.Quat(0.0, 0.0, 0.0, 2.0)
    assert gpu.m3CreateBody(w, C.byref(bd)).index1 == 0
    sd = gpu.m3DefaultShapeDef()
    body = gpu.m3CreateBody(w, C.byref(gpu.m3DefaultBodyDef()))
    assert gpu.m3CreateBoxShape(body, C.byref(sd), api.Vec3(0.0, 1.0, 1.0)).index1 == 0
    gpu.m3DestroyWorld(w)
    assert not gpu.m3World_IsValid(w)


def test_hash64_is_fnv1a():
    """FNV-1a 64 known answers (the textbook vectors)."""
    gpu = product_api()
    init = 0xCBF29CE484222325
    assert gpu.m3Hash64(init, b"", 0) == 0xCBF29CE484222325
    assert gpu.m3Hash64(init, b"a", 1) == 0xAF63DC4C8601EC8C
    assert gpu.m3Hash64(init, b"foobar", 6) == 0x85944171F73967E8


def test_joint_creation_matches_reference():
    """Joint frames (quaternion from the axis basis), the joint lists, the
    wake-on-create rule and the hash's joint block agree before stepping;
    the default def is byte-identical, cookie included."""
    ref = ref_api()
    gpu = product_api()
    assert bytes(ref.m3DefaultJointDef()) == bytes(gpu.m3DefaultJointDef())
    for kw in (dict(count=4, links=6, boxes=0), dict(count=30, links=20, boxes=200)):
        wr = scenes.chains(ref, **kw)
        wg = scenes.chains(gpu, **kw)
        assert hexh(gpu.m3World_Hash(wg)) == hexh(ref.m3World_Hash(wr))
        ref.m3DestroyWorld(wr)
        gpu.m3DestroyWorld(wg)


def test_joint_refusals_and_destroy():
    ref = ref_api()
    gpu = product_api()
    for lib in (ref, gpu):
        w = scenes.box_pyramid(lib, rows=3)
        sbd = lib.m3DefaultBodyDef()
        static = lib.m3CreateBody(w, C.byref(sbd))
        static2 = lib.m3CreateBody(w, C.byref(sbd))
        dbd = lib.m3DefaultBodyDef()
        dbd.type = api.DYNAMIC
        dbd.position = api.Pos3(0.0, 5.0, 0.0)
        dyn = lib.m3CreateBody(w, C.byref(dbd))
        jd = lib.m3DefaultJointDef()
        jd.bodyA, jd.bodyB = static, static2
        assert lib.m3CreateJoint(C.byref(jd)).index1 == 0  # two immovables
        jd.bodyA, jd.bodyB = dyn, dyn
        assert lib.m3CreateJoint(C.byref(jd)).index1 == 0  # same body
        jd.bodyA, jd.bodyB = static, dyn
        bad = lib.m3DefaultJointDef()
        bad.bodyA, bad.bodyB = static, dyn
        bad.internalValue ^= 4
        assert lib.m3CreateJoint(C.byref(bad)).index1 == 0  # cookie
        bad = lib.m3DefaultJointDef()
        bad.bodyA, bad.bodyB = static, dyn
        bad.enableCone = True
        bad.coneAngle = -0.5
        assert lib.m3CreateJoint(C.byref(bad)).index1 == 0  # a negative cone
        j = lib.m3CreateJoint(C.byref(jd))
        assert j.index1 != 0 and lib.m3Joint_IsValid(j)
        lib.m3DestroyJoint(j)
        assert not lib.m3Joint_IsValid(j)
        j2 = lib.m3CreateJoint(C.byref(jd))
        assert j2.index1 == j.index1 and j2.generation == j.generation + 1
        lib.m3DestroyWorld(w)
    # Kinds other than spherical are refused on this build.
    w = scenes.box_pyramid(gpu, rows=2)
    s = gpu.m3CreateBody(w, C.byref(gpu.m3DefaultBodyDef()))
    dbd = gpu.m3DefaultBodyDef()
    dbd.type = api.DYNAMIC
    d = gpu.m3CreateBody(w, C.byref(dbd))
    jd = gpu.m3DefaultJointDef()
    jd.bodyA, jd.bodyB = s, d
    jd.type = 3  # fixed
    assert gpu.m3CreateJoint(C.byref(jd)).index1 == 0
    gpu.m3DestroyWorld(w)


def test_joint_world_snapshot_roundtrip():
    gpu = product_api()
    w = scenes.chains(gpu, count=10, links=10, boxes=20)
    h0 = gpu.m3World_Hash(w)
    snap = gpu.snapshot(w)
    w2 = scenes.chains(gpu, count=10, links=10, boxes=20)
    jd = gpu.m3DefaultJointDef()
    assert gpu.restore(w2, snap)
    assert gpu.m3World_Hash(w2) == h0
    gpu.m3DestroyWorld(w)
    gpu.m3DestroyWorld(w2)
