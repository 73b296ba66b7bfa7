"""Event streams of the last step: contact and sensor begin/end, hits, body
moves and joint breaks (include/maul3d/world.h:206-313; emitted by
src/solver.c:3732-3815, 2991-3026 and 4295-4306).

The device builds the streams when the host reads them; each stream must
equal the reference's record for record, in the same order, with the same
bits in every floating-point field.  Fixture: tests/golden/events.json (per
step of the event rain: world hash and a count + SHA-256 prefix of each
stream, from the reference build, with the scene's mid-run changes of the
hit threshold and of shapes' hit opt-in).
"""

from __future__ import annotations

import ctypes as C
import json
import os

import pytest

from tests.helpers import (
    DT,
    EVENT_STEPS,
    ROOT,
    SUBSTEPS,
    EventActions,
    event_digest,
    hexh,
    product_api,
    ref_api,
)

from maul3d_amd import scenes


def _fixture() -> dict:
    with open(os.path.join(ROOT, "tests", "golden", "events.json")) as f:
        return json.load(f)


def test_fixture_covers_every_stream():
    """The scene exercises every stream the device builds (the joint break
    stream stays empty: joints never break in this build)."""
    fx = _fixture()
    totals = {}
    for rec in fx["per_step"].values():
        for name, val in rec["events"].items():
            if isinstance(val, list):
                totals[name] = totals.get(name, 0) + val[0]
    for name in (
        "ContactBeginEvents",
        "ContactEndEvents",
        "SensorBeginEvents",
        "SensorEndEvents",
        "HitEvents",
        "BodyMoveEvents",
    ):
        assert totals[name] > 0, name
    assert totals["JointBreakEvents"] == 0


def test_reference_reproduces_event_fixture_prefix():
    """The digest pipeline and the scene are stable: the reference build
    reproduces the first 130 steps of the fixture (the hit threshold change
    at step 120 included)."""
    ref = ref_api()
    fx = _fixture()
    w = scenes.event_rain(ref)
    acts = EventActions(ref)
    assert hexh(ref.m3World_Hash(w)) == fx["init"]
    for step in range(1, 131):
        ref.m3World_Step(w, DT, SUBSTEPS)
        want = fx["per_step"][str(step)]
        assert hexh(ref.m3World_Hash(w)) == want["hash"], step
        assert event_digest(ref.events(w)) == want["events"], step
        acts.after_step(w, step)
    ref.m3DestroyWorld(w)


def test_cpu_twin_events_match_fixture_every_step():
    """The float CPU twin (oracle/twin: the same store and hit-record code and
    the device's event rule on the serial step) reproduces every step's
    streams of the fixture, bit for bit, before any GPU run."""
    path = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    from maul3d_amd import api

    twin = api.Api(C.CDLL(path))
    fx = _fixture()
    w = scenes.event_rain(twin)
    acts = EventActions(twin)
    for step in range(1, EVENT_STEPS + 1):
        twin.m3World_Step(w, DT, SUBSTEPS)
        want = fx["per_step"][str(step)]
        got = event_digest(twin.events(w))
        assert got == want["events"], (step, {k: (got[k], want["events"][k]) for k in got if got[k] != want["events"][k]})
        assert hexh(twin.m3World_Hash(w)) == want["hash"], step
        acts.after_step(w, step)
    twin.m3DestroyWorld(w)


def test_cpu_twin_restore_clears_events():
    path = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref/libm3twin32.so not built")
    from maul3d_amd import api

    twin = api.Api(C.CDLL(path))
    fx = _fixture()
    w = scenes.event_rain(twin)
    for _ in range(50):
        twin.m3World_Step(w, DT, SUBSTEPS)
    snap = twin.snapshot(w)
    twin.m3World_Step(w, DT, SUBSTEPS)
    assert twin.restore(w, snap)
    assert all(len(v) == 0 for v in twin.events(w).values() if isinstance(v, list))
    twin.m3World_Step(w, DT, SUBSTEPS)
    assert event_digest(twin.events(w)) == fx["per_step"]["51"]["events"]
    twin.m3DestroyWorld(w)


def test_product_exports_event_api_with_reference_layouts():
    from maul3d_amd import api

    assert C.sizeof(api.ContactEvent) == 16
    assert C.sizeof(api.HitEvent) == 56
    assert C.sizeof(api.BodyMoveEvent) == 56
    gpu = product_api()
    for name in api.EVENT_STREAMS:
        assert hasattr(gpu, name), name


@pytest.mark.gpu
def test_event_streams_match_fixture_every_step():
    gpu = product_api()
    fx = _fixture()
    w = scenes.event_rain(gpu)
    acts = EventActions(gpu)
    assert hexh(gpu.m3World_Hash(w)) == fx["init"]
    fell = 0
    for step in range(1, EVENT_STEPS + 1):
        gpu.m3World_Step(w, DT, SUBSTEPS)
        ev = gpu.events(w)
        fell += sum(1 for m in ev["BodyMoveEvents"] if m[3])
        want = fx["per_step"][str(step)]
        got = event_digest(ev)
        assert got == want["events"], (step, {k: (got[k], want["events"][k]) for k in got if got[k] != want["events"][k]})
        assert hexh(gpu.m3World_Hash(w)) == want["hash"], step
        acts.after_step(w, step)
    assert fell > 0  # some island fell asleep inside the window
    gpu.m3DestroyWorld(w)


@pytest.mark.gpu
def test_event_streams_equal_live_reference_every_step():
    """Record for record and bit for bit against the reference build."""
    ref = ref_api()
    gpu = product_api()
    wr = scenes.event_rain(ref, count=250, seed=0xE7E48)
    wg = scenes.event_rain(gpu, count=250, seed=0xE7E48)
    ar = EventActions(ref)
    ag = EventActions(gpu)
    for step in range(1, 241):
        ref.m3World_Step(wr, DT, SUBSTEPS)
        gpu.m3World_Step(wg, DT, SUBSTEPS)
        er = ref.events(wr)
        eg = gpu.events(wg)
        for name in er:
            assert eg[name] == er[name], (step, name)
        assert gpu.m3World_Hash(wg) == ref.m3World_Hash(wr), step
        ar.after_step(wr, step)
        ag.after_step(wg, step)
    ref.m3DestroyWorld(wr)
    gpu.m3DestroyWorld(wg)


@pytest.mark.gpu
def test_restore_clears_events_and_resimulation_repeats_them():
    gpu = product_api()
    fx = _fixture()
    w = scenes.event_rain(gpu)
    for _ in range(50):
        gpu.m3World_Step(w, DT, SUBSTEPS)
    snap = gpu.snapshot(w)
    first = []
    for _ in range(10):
        gpu.m3World_Step(w, DT, SUBSTEPS)
        first.append(event_digest(gpu.events(w)))
    assert gpu.restore(w, snap)
    cleared = gpu.events(w)
    assert all(len(v) == 0 for k, v in cleared.items() if isinstance(v, list))
    for k in range(10):
        gpu.m3World_Step(w, DT, SUBSTEPS)
        assert event_digest(gpu.events(w)) == first[k], k
        assert event_digest(gpu.events(w)) == fx["per_step"][str(51 + k)]["events"], k
    gpu.m3DestroyWorld(w)


@pytest.mark.gpu
def test_batched_event_worlds_equal_single_world_path():
    gpu = product_api()
    seeds = (0xE7E47, 0xE7E49, 0xE7E4A)
    batch = [scenes.event_rain(gpu, count=300, seed=s) for s in seeds]
    single = [scenes.event_rain(gpu, count=300, seed=s) for s in seeds]
    ids = (type(batch[0]) * len(batch))(*batch)
    lib = gpu.lib
    for step in range(1, 91):
        assert lib.m3World_StepMany(ids, len(batch), C.c_float(DT), SUBSTEPS) == len(batch)
        for k in range(len(batch)):
            gpu.m3World_Step(single[k], DT, SUBSTEPS)
        for k in range(len(batch)):
            assert event_digest(gpu.events(batch[k])) == event_digest(gpu.events(single[k])), (step, k)
    for w in batch + single:
        gpu.m3DestroyWorld(w)
