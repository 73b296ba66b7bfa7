"""Writes the golden world-hash fixtures from the reference engine.

Run in the build container after ``make -C oracle ref`` (which compiles the
reference's own C sources into oracle/_ref/libmaul3d_ref.so).  For each
scene of maul3d_amd.scenes it records the world hash before stepping and
after selected steps of m3World_Step(1/60, 4), plus the counters the tests
use to size their checks.  The output, hashes.json, is data: inputs are the
seeded scene definitions, outputs the reference's hashes.  ``events``
writes events.json: per step of the event rain, the hash and a digest of
each event stream.  ``snapshots`` writes snapshots.json: the size and
SHA-256 of the reference's snapshot bytes at selected steps of two scenes.
"""

from __future__ import annotations

import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from maul3d_amd import api, scenes  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "libmaul3d_ref.so")

# scene -> (steps, record every k steps)
PLAN = {
    "bench_pyramid": (300, 1),
    "box_pyramid": (300, 1),
    "smoke10k": (150, 10),
    "rain": (600, 10),
    "chains": (600, 10),
    "hull_chains": (600, 10),
}


def run_scene(ref: api.Api, name: str, steps: int, every: int) -> dict:
    world = scenes.SCENES[name](ref)
    out = {"steps": steps, "every": every, "init": "%016x" % ref.m3World_Hash(world), "hashes": {}}
    t0 = time.time()
    for i in range(steps):
        ref.m3World_Step(world, 1.0 / 60.0, 4)
        if (i + 1) % every == 0 or i + 1 == steps:
            out["hashes"][str(i + 1)] = "%016x" % ref.m3World_Hash(world)
    c = ref.m3World_GetCounters(world)
    out["final"] = {
        "bodies": c.bodyCount,
        "contacts": c.contactCount,
        "awake": c.awakeCount,
        "islands": c.islandCount,
        "colors": c.colorCount,
        "cpu_seconds": round(time.time() - t0, 2),
    }
    ref.m3DestroyWorld(world)
    return out


def toss_fixture(ref: api.Api, worlds: int, steps: int) -> dict:
    out = {}
    for k in range(worlds):
        w = scenes.toss(ref, seed=0x7055 + k)
        for _ in range(steps):
            ref.m3World_Step(w, 1.0 / 60.0, 4)
        out[str(k)] = "%016x" % ref.m3World_Hash(w)
        ref.m3DestroyWorld(w)
    return out


def batch_rollback_fixture(ref: api.Api, worlds: int, warm: int, resim: int) -> dict:
    """Per world of the batched-rollback set: the hash after ``warm`` steps,
    then snapshot, restore and ``resim`` more steps (hash after each)."""
    out = {"warm": warm, "resim": resim, "worlds": {}}
    for k in range(worlds):
        w = scenes.batch_world(ref, k)
        for _ in range(warm):
            ref.m3World_Step(w, 1.0 / 60.0, 4)
        snap = ref.snapshot(w)
        rec = {"warm": "%016x" % ref.m3World_Hash(w), "bodies": ref.m3World_GetCounters(w).bodyCount}
        if not ref.restore(w, snap):
            raise RuntimeError(f"reference restore refused world {k}")
        hashes = []
        for _ in range(resim):
            ref.m3World_Step(w, 1.0 / 60.0, 4)
            hashes.append("%016x" % ref.m3World_Hash(w))
        rec["resim"] = hashes
        out["worlds"][str(k)] = rec
        ref.m3DestroyWorld(w)
        if k % 32 == 31:
            print("batch world", k + 1, flush=True)
    return out


def events_fixture(ref: api.Api) -> dict:
    """Per step of the event rain: the world hash and the digest (count and
    SHA-256 prefix) of every event stream, with the scene's mid-run changes
    (tests/helpers.py EventActions)."""
    from tests.helpers import EVENT_STEPS, EventActions, event_digest

    world = scenes.event_rain(ref)
    acts = EventActions(ref)
    out = {"steps": EVENT_STEPS, "init": "%016x" % ref.m3World_Hash(world), "per_step": {}}
    for i in range(1, EVENT_STEPS + 1):
        ref.m3World_Step(world, 1.0 / 60.0, 4)
        out["per_step"][str(i)] = {
            "hash": "%016x" % ref.m3World_Hash(world),
            "events": event_digest(ref.events(world)),
        }
        acts.after_step(world, i)
    ref.m3DestroyWorld(world)
    return out


# scene -> steps after which the snapshot's digest is recorded
SNAPSHOT_PLAN = {"rain": (1, 100, 600), "hull_chains": (1, 60)}


def snapshots_fixture(ref: api.Api) -> dict:
    """The reference's snapshot bytes (size and SHA-256) of the 10k rain and
    the hull-rain-with-chains scene at selected steps."""
    import hashlib

    out = {}
    for name, marks in SNAPSHOT_PLAN.items():
        world = scenes.SCENES[name](ref)
        rec = {}
        for i in range(1, max(marks) + 1):
            ref.m3World_Step(world, 1.0 / 60.0, 4)
            if i in marks:
                snap = ref.snapshot(world)
                rec[str(i)] = {"bytes": len(snap), "sha256": hashlib.sha256(snap).hexdigest()}
        out[name] = rec
        ref.m3DestroyWorld(world)
        print("snapshots", name, flush=True)
    return out


def main() -> None:
    ref = api.Api(C.CDLL(REF))
    if len(sys.argv) > 1 and sys.argv[1] == "snapshots":
        data = snapshots_fixture(ref)
        with open(os.path.join(HERE, "snapshots.json"), "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        print("wrote snapshots.json")
        return
    if len(sys.argv) > 2 and sys.argv[1] == "scene":
        # Refresh one scene's entry in hashes.json.
        path = os.path.join(HERE, "hashes.json")
        with open(path) as f:
            data = json.load(f)
        name = sys.argv[2]
        steps, every = PLAN[name]
        data["scenes"][name] = run_scene(ref, name, steps, every)
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        print("updated", name)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "events":
        data = events_fixture(ref)
        with open(os.path.join(HERE, "events.json"), "w") as f:
            json.dump(data, f, indent=0, sort_keys=True)
        print("wrote events.json")
        return
    if len(sys.argv) > 1 and sys.argv[1] == "batch":
        data = batch_rollback_fixture(ref, 256, 100, 8)
        with open(os.path.join(HERE, "batch_rollback.json"), "w") as f:
            json.dump(data, f, indent=0, sort_keys=True)
        print("wrote batch_rollback.json")
        return
    data = {"dt": "1/60", "substeps": 4, "scenes": {}}
    for name, (steps, every) in PLAN.items():
        print("scene", name, flush=True)
        data["scenes"][name] = run_scene(ref, name, steps, every)
    data["toss_120"] = toss_fixture(ref, 64, 120)
    with open(os.path.join(HERE, "hashes.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("wrote hashes.json")


if __name__ == "__main__":
    main()
