"""Shape material setters (include/maul3d/shape.h:286-296, src/shape.c:
1388-1541): friction, restitution and rolling resistance bind at the next
step's prepare; a density change rebuilds the body's mass properties when
asked (src/shape.c:182-300).  Changed mid-run on a mixed rain and checked
against the live reference every step, on the CPU twin and on the MI355X.
Journal records of these ops are covered by the cross replays of
tests/test_journal.py."""

from __future__ import annotations

import ctypes as C
import os

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, product_api, ref_api

from maul3d_amd import api, scenes

TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")


def _twin():
    if not os.path.exists(TWIN32):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    return api.Api(C.CDLL(TWIN32))


def _shape(w, slot: int) -> api.ShapeId:
    return api.ShapeId(slot + 1, w.index1 - 1, 0)


def _edit(m3, w, step: int) -> None:
    """Every 15 steps a different group of shapes gets new materials."""
    n = m3.m3World_GetCounters(w).shapeCount
    k = step // 15
    for s in range(1 + k % 5, n, 5):
        sid = _shape(w, s)
        if s % 4 == 0:
            m3.m3Shape_SetFriction(sid, 0.05 + 0.1 * (k % 7))
        elif s % 4 == 1:
            m3.m3Shape_SetRestitution(sid, 0.2 * (k % 5))
        elif s % 4 == 2:
            m3.m3Shape_SetRollingResistance(sid, 0.05 * (k % 3))
        else:
            m3.m3Shape_SetDensity(sid, 300.0 + 400.0 * (k % 4), k % 2 == 0)


def _run_against_reference(other, steps: int = 240) -> str:
    ref = ref_api()
    a = scenes.rain(ref, count=300, extent=6.0)
    b = scenes.rain(other, count=300, extent=6.0)
    assert hexh(other.m3World_Hash(b)) == hexh(ref.m3World_Hash(a))
    for step in range(1, steps + 1):
        if step % 15 == 0:
            _edit(ref, a, step)
            _edit(other, b, step)
        ref.m3World_Step(a, DT, SUBSTEPS)
        other.m3World_Step(b, DT, SUBSTEPS)
        assert hexh(other.m3World_Hash(b)) == hexh(ref.m3World_Hash(a)), f"step {step}"
    final = hexh(ref.m3World_Hash(a))
    ref.m3DestroyWorld(a)
    other.m3DestroyWorld(b)
    return final


def test_cpu_twin_material_edits_match_reference_every_step():
    edited = _run_against_reference(_twin())
    ref = ref_api()
    w = scenes.rain(ref, count=300, extent=6.0)
    for _ in range(240):
        ref.m3World_Step(w, DT, SUBSTEPS)
    assert hexh(ref.m3World_Hash(w)) != edited  # the edits changed the trajectory
    ref.m3DestroyWorld(w)


def test_setters_validate_and_getters_read_back():
    T = _twin()
    w = scenes.bench_pyramid(T)
    sid = _shape(w, 3)
    body = T.m3Shape_GetBody(sid)
    assert T.m3Shape_IsValid(sid) and T.m3Body_IsValid(body)
    m0 = T.m3Body_GetMass(body)
    d0 = T.m3Shape_GetDensity(sid)
    T.m3Shape_SetDensity(sid, 2.0 * d0, False)
    assert T.m3Body_GetMass(body) == m0  # books not rebuilt
    T.m3Shape_SetDensity(sid, 2.0 * d0, True)
    assert T.m3Body_GetMass(body) == pytest.approx(2.0 * m0, rel=1e-6)
    for bad in (-1.0, float("nan"), float("inf")):
        T.m3Shape_SetFriction(sid, bad)
        T.m3Shape_SetDensity(sid, bad, True)
    T.m3Shape_SetDensity(sid, 0.0, True)
    assert T.m3Shape_GetDensity(sid) == C.c_float(2.0 * d0).value
    T.m3Shape_SetFriction(sid, 0.25)
    T.m3Shape_SetRestitution(sid, 0.75)
    T.m3Shape_SetRollingResistance(sid, 0.125)
    assert (T.m3Shape_GetFriction(sid), T.m3Shape_GetRestitution(sid), T.m3Shape_GetRollingResistance(sid)) == (
        0.25, 0.75, 0.125)
    bad = api.ShapeId(9999, w.index1 - 1, 0)
    assert not T.m3Shape_IsValid(bad) and T.m3Shape_GetFriction(bad) == 0.0
    assert T.m3Shape_GetBody(bad).index1 == 0
    T.m3DestroyWorld(w)


@pytest.mark.gpu
def test_material_edits_match_live_reference_every_step():
    _run_against_reference(product_api())
