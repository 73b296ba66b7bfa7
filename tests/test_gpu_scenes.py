"""More scene-level parity of the MI355X step: the reference's hull-jam pin
(bench/pins.txt:2), restitution against the live reference every step, and
the capacity stall (src/solver.c:3726-3730: a full pair table stalls the
step, which does not advance)."""

from __future__ import annotations

import ctypes as C
import json
import os

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, hexh, product_api, ref_api

from maul3d_amd import api, scenes

pytestmark = pytest.mark.gpu


def _pins() -> dict:
    with open(os.path.join(ROOT, "tests", "golden", "pins.json")) as f:
        return json.load(f)["pins"]


def test_hulljam_pin():
    """bench/pins.txt:2: 80 boxes and 12-point QuickHull rocks in a column,
    300 steps."""
    gpu = product_api()
    w = scenes.hulljam(gpu)
    for _ in range(300):
        gpu.m3World_Step(w, DT, SUBSTEPS)
    assert hexh(gpu.m3World_Hash(w)) == _pins()["hulljam"]
    gpu.m3DestroyWorld(w)


def test_restitution_matches_live_reference_every_step():
    """Spheres, boxes and capsules with restitution up to 0.95 bouncing on a
    plane and each other: the serial restitution pass (src/solver.c:2879-2907)
    runs on the device every step; every step's hash equals the reference's."""
    ref = ref_api()
    gpu = product_api()
    a = scenes.bounce(ref)
    b = scenes.bounce(gpu)
    for step in range(1, 241):
        ref.m3World_Step(a, DT, SUBSTEPS)
        gpu.m3World_Step(b, DT, SUBSTEPS)
        assert hexh(gpu.m3World_Hash(b)) == hexh(ref.m3World_Hash(a)), f"step {step}"
    ref.m3DestroyWorld(a)
    gpu.m3DestroyWorld(b)


def _crowd(m3, spread: bool):
    """41 shape slots (pair table 8 x 41 = 328 keys); 40 unit spheres, all at
    one point (780 overlapping pairs + 40 plane pairs: the table overflows)
    or spread 2 m apart (no overflow)."""
    d = m3.m3DefaultWorldDef()
    d.bodyCapacity = 41
    d.shapeCapacity = 41
    w = m3.m3CreateWorld(C.byref(d))
    gd = m3.m3DefaultBodyDef()
    ground = m3.m3CreateBody(w, C.byref(gd))
    sd = m3.m3DefaultShapeDef()
    assert m3.m3CreatePlaneShape(ground, C.byref(sd), C.byref(scenes.Plane(api.Vec3(0, 1, 0), 0.0))).index1
    bd = m3.m3DefaultBodyDef()
    bd.type = scenes.DYNAMIC
    bodies = []
    for k in range(40):
        bd.position = api.Pos3(2.0 * (k % 8) if spread else 0.0, 3.0 + (2.0 * (k // 8) if spread else 0.0), 0.0)
        b = m3.m3CreateBody(w, C.byref(bd))
        assert m3.m3CreateSphereShape(b, C.byref(sd), C.byref(scenes.Sphere(api.Vec3(0, 0, 0), 0.5))).index1
        bodies.append(b)
    return w, bodies


def _state(m3, bodies):
    out = []
    for b in bodies:
        p = m3.m3Body_GetPosition(b)
        v = m3.m3Body_GetLinearVelocity(b)
        out.append((p.x, p.y, p.z, v.x, v.y, v.z))
    return out


def test_pair_capacity_stall_does_not_advance_and_resumes():
    """A full pair table stalls the step: the world does not advance (the
    step count and every body's pose and velocity stay as they were, as on
    the reference).  A restored snapshot of a world that fits steps on, equal
    to the reference.  This build keeps the previous pair list on a stall;
    the reference keeps the partial list it had emitted (its pair keys and
    count, which the hash covers, differ after a stall)."""
    ref = ref_api()
    gpu = product_api()
    rw, rbodies = _crowd(ref, spread=False)
    gw, gbodies = _crowd(gpu, spread=False)
    assert hexh(gpu.m3World_Hash(gw)) == hexh(ref.m3World_Hash(rw))
    h0 = gpu.m3World_Hash(gw)
    s0 = _state(gpu, gbodies)
    assert s0 == _state(ref, rbodies)
    for _ in range(3):
        ref.m3World_Step(rw, DT, SUBSTEPS)
        gpu.m3World_Step(gw, DT, SUBSTEPS)
    assert gpu.m3World_Hash(gw) == h0  # nothing advanced (step count, bodies, pairs)
    assert _state(gpu, gbodies) == s0
    assert _state(ref, rbodies) == s0  # the reference did not advance either
    assert gpu.m3World_GetDeviceError(gw) is None
    # Resume: the snapshot of a same-capacity world that fits.
    fit, _ = _crowd(gpu, spread=True)
    snap = gpu.snapshot(fit)
    assert gpu.restore(gw, snap)
    rfit, _ = _crowd(ref, spread=True)
    for step in range(1, 31):
        gpu.m3World_Step(gw, DT, SUBSTEPS)
        ref.m3World_Step(rfit, DT, SUBSTEPS)
        assert hexh(gpu.m3World_Hash(gw)) == hexh(ref.m3World_Hash(rfit)), f"step {step}"
    for w in (gw, fit):
        gpu.m3DestroyWorld(w)
    for w in (rw, rfit):
        ref.m3DestroyWorld(w)


@pytest.mark.parametrize("bullets", [0, 3])
def test_continuous_against_tree_shapes_matches_live_reference_every_step(bullets):
    """A static box floor and walls, static hull rocks and capsules, fast
    bodies and (second case) bullets: the device's continuous pass sweeps
    against the tree shapes in the host tree's depth-first order, with the
    fraction threaded through the candidates (src/solver.c:3081-3430)."""
    ref = ref_api()
    gpu = product_api()
    a = scenes.ccd_arena(ref, bullets=bullets)
    b = scenes.ccd_arena(gpu, bullets=bullets)
    for step in range(1, 241):
        ref.m3World_Step(a, DT, SUBSTEPS)
        gpu.m3World_Step(b, DT, SUBSTEPS)
        assert hexh(gpu.m3World_Hash(b)) == hexh(ref.m3World_Hash(a)), f"step {step}"
    assert gpu.m3World_GetCounters(b).treeHeight == ref.m3World_GetCounters(a).treeHeight
    assert gpu.m3World_GetDeviceError(b) is None
    ref.m3DestroyWorld(a)
    gpu.m3DestroyWorld(b)


def test_kinematic_sweepers_match_live_reference():
    """Kinematic paddles sliding and spinning through a 1600-body pile that
    spans several solver cells: a contact's kinematic side moves its position
    deltas every substep, so those contacts must see them in step with the
    reference.  Hash checked every 10 steps of 240."""
    ref = ref_api()
    gpu = product_api()
    wr = scenes.sweepers(ref)
    wg = scenes.sweepers(gpu)
    for i in range(1, 241):
        ref.m3World_Step(wr, DT, SUBSTEPS)
        gpu.m3World_Step(wg, DT, SUBSTEPS)
        if i % 10 == 0:
            assert hexh(gpu.m3World_Hash(wg)) == hexh(ref.m3World_Hash(wr)), f"step {i}"
    ref.m3DestroyWorld(wr)
    gpu.m3DestroyWorld(wg)
