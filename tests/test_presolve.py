"""The pre-solve veto (src/solver.c:184-248, include/maul3d/world.h:315-340).

One-way platforms: bodies launched up through static box platforms whose
contacts from below are vetoed by a pure callback.  Checked on the CPU twin
(world.c's pre-solve pass over the serial step's candidates) against the live
reference every step, with the same callback registered on both; the journal
records each step's vetoes before the step, and a bare replay (no callback)
of either library's journal lands on the recorded hash.  The GPU test repeats
it on the device.
"""

from __future__ import annotations

import ctypes as C
import os

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, ref_api

from maul3d_amd import api, scenes

TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")

PreSolveFn = C.CFUNCTYPE(C.c_bool, api.ShapeId, api.ShapeId, api.Pos3, api.Vec3, C.c_void_p)


def install(m3, world, slots, calls=None):
    """Registers the scene's one-way rule; returns the ctypes callback (the
    caller keeps it alive)."""
    rule = scenes.one_way_rule(slots)

    def cb(a, b, point, normal, ctx):
        keep = rule(a.index1 - 1, b.index1 - 1, normal.y)
        if calls is not None:
            calls.append((a.index1, b.index1, point.x, point.y, point.z, normal.x, normal.y, normal.z, keep))
        return keep

    fn = PreSolveFn(cb)
    m3.lib.m3World_SetPreSolveCallback.argtypes = [api.WorldId, PreSolveFn, C.c_void_p]
    m3.lib.m3World_SetPreSolveCallback.restype = None
    m3.lib.m3World_SetPreSolveCallback(world, fn, None)
    return fn


def _twin():
    if not os.path.exists(TWIN32):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    return api.Api(C.CDLL(TWIN32))


def run_against_reference(other, steps=180):
    ref = ref_api()
    a, sa = scenes.platforms(ref)
    b, sb = scenes.platforms(other)
    assert sa == sb
    ca, cb = [], []
    fa = install(ref, a, sa, ca)
    fb = install(other, b, sb, cb)
    vetoes = 0
    for step in range(1, steps + 1):
        na, nb = len(ca), len(cb)
        ref.m3World_Step(a, DT, SUBSTEPS)
        other.m3World_Step(b, DT, SUBSTEPS)
        # The callback saw the same candidates with the same arguments.
        assert cb[nb:] == ca[na:], f"step {step}"
        assert hexh(other.m3World_Hash(b)) == hexh(ref.m3World_Hash(a)), f"step {step}"
        vetoes += sum(1 for c in ca[na:] if not c[-1])
    assert vetoes > 20 and len(ca) > vetoes  # contacts were vetoed and kept
    del fa, fb
    ref.m3DestroyWorld(a)
    other.m3DestroyWorld(b)


def test_cpu_twin_presolve_matches_reference_every_step():
    run_against_reference(_twin())


def record_with_vetoes(m3, steps=120):
    w, slots = scenes.platforms(m3)
    fn = install(m3, w, slots)
    buf = C.create_string_buffer(1 << 20)
    assert m3.m3World_JournalBegin(w, buf, len(buf))
    for _ in range(steps):
        m3.m3World_Step(w, DT, SUBSTEPS)
    n = m3.m3World_JournalEnd(w)
    final = m3.m3World_Hash(w)
    del fn
    m3.m3DestroyWorld(w)
    return buf.raw[:n], final


def bare_replay(m3, journal: bytes) -> int:
    w, _ = scenes.platforms(m3)
    data = C.create_string_buffer(journal, len(journal))
    assert m3.m3World_JournalReplay(w, data, len(journal))
    h = m3.m3World_Hash(w)
    m3.m3DestroyWorld(w)
    return h


def _ops(journal: bytes) -> list[int]:
    out = []
    i = 0
    while i < len(journal):
        out.append(int.from_bytes(journal[i : i + 4], "little"))
        i += 8 + int.from_bytes(journal[i + 4 : i + 8], "little")
    return out


def test_vetoes_are_journaled_and_replay_bare():
    ref = ref_api()
    tw = _twin()
    jr, hr = record_with_vetoes(ref)
    jt, ht = record_with_vetoes(tw)
    assert jt == jr  # the same records, annexes included, byte for byte
    assert 79 in _ops(jr)
    assert ht == hr
    assert bare_replay(tw, jr) == hr
    assert bare_replay(ref, jt) == hr
