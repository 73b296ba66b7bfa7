"""Refusal paths of the host world (maul3d_amd/csrc/world.c, the file the
product links), exercised on the CPU twin of the step (oracle/_ref,
libm3twin32.so: world.c over a CPU implementation of the m3gpu ABI).

The reference's contract (include/maul3d/base.h:56-62, src/solver.c:4322-4330,
src/world.c:1623-1669): creation refuses a def it cannot serve with a null
id; m3World_Step is void and never ends the process; a journal record that
cannot be applied refuses the whole replay, which leaves the world as it was.
Nothing here may abort the host.
"""

from __future__ import annotations

import ctypes as C
import os
import struct

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, ref_api

from maul3d_amd import api, scenes

TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")


def _twin():
    if not os.path.exists(TWIN32):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    return api.Api(C.CDLL(TWIN32))


def _world(m3, continuous: bool = True, bodies: int = 64, shapes: int = 64):
    d = m3.m3DefaultWorldDef()
    d.bodyCapacity = bodies
    d.shapeCapacity = shapes
    d.enableContinuous = continuous
    w = m3.m3CreateWorld(C.byref(d))
    assert w.index1 != 0
    return w


def _ground_and_ball(m3, world, static_box: bool):
    sd = m3.m3DefaultShapeDef()
    bd = m3.m3DefaultBodyDef()
    ground = m3.m3CreateBody(world, C.byref(bd))
    if static_box:
        assert m3.m3CreateBoxShape(ground, C.byref(sd), api.Vec3(20.0, 0.5, 20.0)).index1
    else:
        plane = scenes.Plane(api.Vec3(0.0, 1.0, 0.0), 0.0)
        assert m3.m3CreatePlaneShape(ground, C.byref(sd), C.byref(plane)).index1
    bd.type = scenes.DYNAMIC
    bd.position = api.Pos3(0.0, 3.0, 0.0)
    ball = m3.m3CreateBody(world, C.byref(bd))
    assert m3.m3CreateSphereShape(ball, C.byref(sd), C.byref(scenes.Sphere(api.Vec3(0, 0, 0), 0.5))).index1
    return ball


@pytest.mark.parametrize("lib", ["product", "twin"])
def test_oversized_shape_capacity_refused_at_creation(lib):
    """The device packs two shape slots into a 32-bit key: a def above 65535
    shapes gets the null id at m3CreateWorld (no device is touched)."""
    if lib == "product":
        import maul3d_amd

        m3 = maul3d_amd.get_api()
    else:
        m3 = _twin()
    d = m3.m3DefaultWorldDef()
    d.shapeCapacity = 70000
    assert m3.m3CreateWorld(C.byref(d)).index1 == 0
    d.shapeCapacity = 65535
    w = m3.m3CreateWorld(C.byref(d))
    assert w.index1 != 0
    m3.m3DestroyWorld(w)


def _journal(m3, world, fn) -> bytes:
    buf = C.create_string_buffer(1 << 16)
    assert m3.m3World_JournalBegin(world, buf, len(buf))
    fn()
    n = m3.m3World_JournalEnd(world)
    assert n >= 0
    return buf.raw[:n]


def _replay(m3, world, data: bytes) -> bool:
    b = C.create_string_buffer(data, max(len(data), 1))
    return bool(m3.m3World_JournalReplay(world, b, len(data)))


def test_hostile_step_records_refuse_and_leave_world():
    """A step record with dt <= 0 or substeps outside [1, 256] refuses the
    replay (src/world.c:1635-1642) instead of being skipped."""
    m3 = _twin()
    world = _world(m3)
    _ground_and_ball(m3, world, static_box=False)
    m3.m3World_Step(world, DT, SUBSTEPS)
    before = m3.m3World_Hash(world)
    for dt, sub in ((0.0, 4), (-DT, 4), (float("nan"), 4), (DT, 0), (DT, 1 << 30)):
        rec = struct.pack("<ii", 1, 8) + struct.pack("<fi", dt, sub)
        assert not _replay(m3, world, rec)
        assert m3.m3World_Hash(world) == before
    ok = struct.pack("<ii", 1, 8) + struct.pack("<fi", DT, SUBSTEPS)
    assert _replay(m3, world, ok)
    assert m3.m3World_Hash(world) != before
    m3.m3DestroyWorld(world)


def test_hostile_shape_records_refuse_and_leave_world():
    """A create-shape record from the reference, with its geometry patched to
    a NaN, zero or negative radius, refuses the replay (the per-type wall of
    src/shape.c:433-470), and the unpatched record replays."""
    ref = ref_api()
    m3 = _twin()
    rw = _world(ref)
    sd = ref.m3DefaultShapeDef()
    bd = ref.m3DefaultBodyDef()
    bd.type = scenes.DYNAMIC
    body = ref.m3CreateBody(rw, C.byref(bd))
    sphere = scenes.Sphere(api.Vec3(0.0, 0.0, 0.0), 0.75)
    journal = _journal(ref, rw, lambda: ref.m3CreateSphereShape(body, C.byref(sd), C.byref(sphere)))
    ref.m3DestroyWorld(rw)
    op, size = struct.unpack_from("<ii", journal, 0)
    assert op == 6 and 8 + size == len(journal)
    radius = struct.pack("<f", 0.75)
    at = journal.index(radius)  # geom.s of the record's m3ShapeGeom
    tw = _world(m3)
    b2 = m3.m3CreateBody(tw, C.byref(bd))
    assert b2.index1 == body.index1
    before = m3.m3World_Hash(tw)
    for bad in (float("nan"), 0.0, -1.0, float("inf")):
        patched = journal[:at] + struct.pack("<f", bad) + journal[at + 4 :]
        assert not _replay(m3, tw, patched)
        assert m3.m3World_Hash(tw) == before
    assert _replay(m3, tw, journal)
    assert m3.m3World_Hash(tw) != before
    m3.m3DestroyWorld(tw)


def test_device_fault_stops_the_world_without_ending_the_process():
    """A world whose device step fails (here: the twin refuses joints, which
    it does not simulate, with a failure code) is faulted: the step is a
    no-op, the hash reads 0, the snapshot refuses, and the error text is
    reported; other worlds keep stepping."""
    m3 = _twin()
    d = m3.m3DefaultWorldDef()
    d.bodyCapacity = 8
    d.shapeCapacity = 8
    d.jointCapacity = 4
    w = m3.m3CreateWorld(C.byref(d))
    healthy = _world(m3, continuous=False)
    _ground_and_ball(m3, healthy, static_box=False)
    sd = m3.m3DefaultShapeDef()
    bd = m3.m3DefaultBodyDef()
    bd.type = scenes.DYNAMIC
    a = m3.m3CreateBody(w, C.byref(bd))
    bd.position = api.Pos3(1.0, 0.0, 0.0)
    b = m3.m3CreateBody(w, C.byref(bd))
    for body in (a, b):
        m3.m3CreateSphereShape(body, C.byref(sd), C.byref(scenes.Sphere(api.Vec3(0, 0, 0), 0.25)))
    jd = m3.m3DefaultJointDef()
    jd.bodyA = a
    jd.bodyB = b
    assert m3.m3CreateJoint(C.byref(jd)).index1
    assert m3.m3World_GetDeviceError(w) is None
    m3.m3World_Step(w, DT, SUBSTEPS)
    err = m3.m3World_GetDeviceError(w)
    assert err is not None and b"m3gpu_step" in err
    assert m3.m3World_Hash(w) == 0
    assert m3.m3World_SnapshotSize(w) == -1
    m3.m3World_Step(w, DT, SUBSTEPS)  # a no-op, not a crash
    h0 = m3.m3World_Hash(healthy)
    m3.m3World_Step(healthy, DT, SUBSTEPS)
    assert m3.m3World_Hash(healthy) != h0
    assert m3.m3World_GetDeviceError(healthy) is None
    m3.m3DestroyWorld(w)
    m3.m3DestroyWorld(healthy)


def test_autotune_leaves_no_journal_records():
    """The autotuner's trial steps are restored away, so they must not reach
    an active journal."""
    m3 = _twin()
    world = _world(m3, continuous=False)
    _ground_and_ball(m3, world, static_box=False)
    before = m3.m3World_Hash(world)
    res = api.AutotuneResult()
    ids = (api.WorldId * 1)(world)
    journal = _journal(
        m3, world, lambda: m3.m3World_Autotune(ids, 1, DT, SUBSTEPS, 2, C.byref(res))
    )
    assert journal == b""
    assert m3.m3World_Hash(world) == before
    m3.m3DestroyWorld(world)


def test_step_many_journals_each_stepped_world():
    """m3World_StepMany steps every valid world and records a step in each
    stepped world's journal, once per listing; a world whose continuous pass
    meets tree shapes (a static box here) steps on its own path and records
    like the others."""
    m3 = _twin()
    good = _world(m3, continuous=True)
    _ground_and_ball(m3, good, static_box=False)
    boxed = _world(m3, continuous=True)
    _ground_and_ball(m3, boxed, static_box=True)
    h_boxed = m3.m3World_Hash(boxed)
    bufs = []
    for w in (good, boxed):
        buf = C.create_string_buffer(4096)
        assert m3.m3World_JournalBegin(w, buf, len(buf))
        bufs.append(buf)
    ids = (api.WorldId * 3)(good, boxed, good)
    assert m3.m3World_StepMany(ids, 3, DT, SUBSTEPS) == 3
    assert m3.m3World_JournalEnd(good) == 2 * 16  # two step records: op, size, dt, substeps
    assert m3.m3World_JournalEnd(boxed) == 16
    assert m3.m3World_Hash(boxed) != h_boxed
    m3.m3DestroyWorld(good)
    m3.m3DestroyWorld(boxed)


def test_reference_journal_with_static_box_replays_exactly():
    """A reference journal of a world with a static box floor and continuous
    collision on replays on this build's host world and lands on the
    reference's hash (the continuous pass sweeps against the tree shape)."""
    ref = ref_api()
    m3 = _twin()
    rw = _world(ref)
    _ground_and_ball(ref, rw, static_box=True)
    journal = _journal(ref, rw, lambda: [ref.m3World_Step(rw, DT, SUBSTEPS) for _ in range(30)])
    want = ref.m3World_Hash(rw)
    ref.m3DestroyWorld(rw)
    tw = _world(m3)
    _ground_and_ball(m3, tw, static_box=True)
    assert _replay(m3, tw, journal)
    assert m3.m3World_Hash(tw) == want
    m3.m3DestroyWorld(tw)
