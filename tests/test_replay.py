"""The M3J1 replay container and the divergence finder
(include/maul3d/replay.h, src/replay.c, src/diff.c).

The container's framing equals the reference's byte for byte; a session
recorded on this build (snapshot + journal + final hash) verifies on the
reference and the other way round: decode, restore the snapshot into a fresh
world, replay the journal, compare the final hash.  The divergence finder
returns the reference's rows for the same two worlds.  Checked on the CPU
twin; tests/test_gpu_replay.py repeats the session on the device.
"""

from __future__ import annotations

import ctypes as C
import os

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, ref_api

from maul3d_amd import api, scenes

TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")


class JournalInfo(C.Structure):
    _fields_ = [("opCount", C.c_int32), ("stepCount", C.c_int32)]


class ReplayView(C.Structure):
    _fields_ = [
        ("snapshot", C.c_void_p),
        ("snapshotBytes", C.c_int32),
        ("journal", C.c_void_p),
        ("journalBytes", C.c_int32),
        ("finalHash", C.c_uint64),
        ("info", JournalInfo),
    ]


class BodyDiff(C.Structure):
    _fields_ = [
        ("body", api.BodyId),
        ("positionError", C.c_float),
        ("rotationError", C.c_float),
        ("velocityError", C.c_float),
        ("onlyInA", C.c_bool),
        ("onlyInB", C.c_bool),
    ]


def bind(m3):
    lib = m3.lib
    lib.m3JournalDescribe.restype = C.c_bool
    lib.m3JournalDescribe.argtypes = [C.c_void_p, C.c_int32, C.POINTER(JournalInfo)]
    lib.m3ReplayEncodeSize.restype = C.c_int32
    lib.m3ReplayEncodeSize.argtypes = [C.c_int32, C.c_int32]
    lib.m3ReplayEncode.restype = C.c_int32
    lib.m3ReplayEncode.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p, C.c_int32]
    lib.m3ReplayDecode.restype = C.c_bool
    lib.m3ReplayDecode.argtypes = [C.c_void_p, C.c_int32, C.POINTER(ReplayView)]
    lib.m3World_DiffReport.restype = C.c_int32
    lib.m3World_DiffReport.argtypes = [api.WorldId, api.WorldId, C.POINTER(BodyDiff), C.c_int32, C.POINTER(C.c_int32)]
    return lib


def encode(m3, snap: bytes, journal: bytes, final: int) -> bytes:
    lib = bind(m3)
    n = lib.m3ReplayEncodeSize(len(snap), len(journal))
    assert n == 32 + len(snap) + len(journal)
    out = C.create_string_buffer(n)
    assert lib.m3ReplayEncode(snap, len(snap), journal, len(journal), final, out, n) == n
    return out.raw


def decode(m3, data: bytes):
    lib = bind(m3)
    buf = C.create_string_buffer(data, len(data))
    view = ReplayView()
    if not lib.m3ReplayDecode(buf, len(data), C.byref(view)):
        return None
    snap = C.string_at(view.snapshot, view.snapshotBytes)
    journal = C.string_at(view.journal, view.journalBytes) if view.journalBytes else b""
    return snap, journal, view.finalHash, (view.info.opCount, view.info.stepCount)


def _twin():
    if not os.path.exists(TWIN32):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    return api.Api(C.CDLL(TWIN32))


def record_session(m3, mirror: bool):
    """A session on the rain: snapshot at step 10, then a journal over 40
    steps with velocity changes and forces, then the final hash."""
    w = scenes.rain(m3, count=1200)
    if mirror:
        assert m3.m3World_EnableTreeMirror(w, True)
    for _ in range(10):
        m3.m3World_Step(w, DT, SUBSTEPS)
    snap = m3.snapshot(w)
    buf = C.create_string_buffer(1 << 16)
    assert m3.m3World_JournalBegin(w, buf, len(buf))
    for step in range(40):
        if step % 10 == 0:
            bid = api.BodyId(2 + step, w.index1 - 1, 0)
            m3.m3Body_SetLinearVelocity(bid, api.Vec3(1.0, 5.0, -2.0))
            m3.m3Body_ApplyForce(api.BodyId(3 + step, w.index1 - 1, 0), api.Vec3(0.0, 400.0, 0.0))
        m3.m3World_Step(w, DT, SUBSTEPS)
    n = m3.m3World_JournalEnd(w)
    final = m3.m3World_Hash(w)
    m3.m3DestroyWorld(w)
    return snap, buf.raw[:n], final


def verify(m3, container: bytes) -> bool:
    """The reference's m3replay verify: decode, restore, replay, compare."""
    dec = decode(m3, container)
    assert dec is not None
    snap, journal, final, _ = dec
    w = scenes.rain(m3, count=1200)
    assert m3.restore(w, snap)
    jb = C.create_string_buffer(journal, max(len(journal), 1))
    assert m3.m3World_JournalReplay(w, jb, len(journal))
    ok = m3.m3World_Hash(w) == final
    m3.m3DestroyWorld(w)
    return ok


def test_container_framing_equals_reference():
    ref = ref_api()
    tw = _twin()
    snap, journal, final = record_session(ref, mirror=False)
    a = encode(ref, snap, journal, final)
    b = encode(tw, snap, journal, final)
    assert a == b
    for m3 in (ref, tw):
        dec = decode(m3, a)
        assert dec is not None and dec[0] == snap and dec[1] == journal and dec[2] == final
        assert dec[3] == (len([1 for _ in range(40)]) + 8, 40)  # 40 steps, 4 velocity + 4 force records
    bad = {
        "truncated": a[:-1],
        "magic": b"X" + a[1:],
        "step count lies": a[:16] + (41).to_bytes(4, "little") + a[20:],
        "journal framing": a[: 32 + len(snap)] + (0).to_bytes(4, "little") + a[32 + len(snap) + 4 :],
    }
    for label, data in bad.items():
        assert decode(ref, data) is None, label
        assert decode(tw, data) is None, label


@pytest.mark.parametrize("recorder", ["reference", "build"])
def test_session_container_verifies_on_the_other_library(recorder):
    ref = ref_api()
    tw = _twin()
    rec, other = (ref, tw) if recorder == "reference" else (tw, ref)
    snap, journal, final = record_session(rec, mirror=rec is tw)
    container = encode(rec, snap, journal, final)
    assert verify(other, container)
    assert verify(rec, container)


def _diff_rows(m3, a, b):
    lib = bind(m3)
    out = (BodyDiff * 64)()
    n = C.c_int32(0)
    total = lib.m3World_DiffReport(a, b, out, 64, C.byref(n))
    rows = [
        (r.body.index1, r.body.generation, r.positionError, r.rotationError, r.velocityError, r.onlyInA, r.onlyInB)
        for r in out[: n.value]
    ]
    return total, rows


def test_diff_report_equals_reference():
    ref = ref_api()
    tw = _twin()
    results = []
    for m3 in (ref, tw):
        a = scenes.rain(m3, count=1200)
        b = scenes.rain(m3, count=1200)
        for step in range(40):
            if step == 20:
                m3.m3Body_SetLinearVelocity(api.BodyId(500, b.index1 - 1, 0), api.Vec3(0.0, 3.0, 0.0))
            m3.m3World_Step(a, DT, SUBSTEPS)
            m3.m3World_Step(b, DT, SUBSTEPS)
        results.append(_diff_rows(m3, a, b))
        m3.m3DestroyWorld(a)
        m3.m3DestroyWorld(b)
    assert results[0][0] > 0
    assert results[0] == results[1]
