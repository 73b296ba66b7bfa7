"""Benchmark: body-substeps/s of the soft-step contact solver on one MI355X.

Workload: the 10k-body mixed rain, one world (10,000 dynamic bodies: 40%
boxes, 40% spheres, 20% capsules, seeded), m3World_Step(1/60, 4).  The
headline ``value`` always times the scene's full 600-step trajectory (fall,
pile-up, settle) from step 0, with the world resident in HBM, and the
step-600 world hash must equal the committed reference fixture (a missing
fixture or a mismatch exits non-zero).  ``steps`` and ``ms_per_step`` describe
that timed trajectory.

``--warmup W --steps K`` do not shorten the headline: they set the untimed
warm-up (W steps on a separate world of the same scene, discarded) and the
``window`` sub-record (steps W+1..W+K of the trajectory timed on that
warm-up world; its hash is checked at the next step that has a fixture).

The same line carries more measurements: ``batched`` (256 independent
worlds restored from a snapshot and resimulated 8 steps through
m3World_StepMany, against the same worlds stepped one by one),
``hull_chains`` (4000 QuickHull hulls plus 500 chains of 20 capsule links on
spherical joints, 600 steps), ``toss_adjoint`` (64 differentiable toss
worlds: forward time, adjoint time and their ratio) and ``events_read``.

Prints one JSON line.  ``roofline`` reports the dominant kernel's
algorithmic HBM bytes over its event-timed average duration, measured over
the same 600-step trajectory; ``cpu_baseline`` times the reference engine
(built from its C sources into oracle/_ref) on the host cores over the same
600 steps, serial and through its task hooks on a host thread pool, and
reports the faster.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DT = 1.0 / 60.0
SUBSTEPS = 4
BODIES = 10000
TRAJECTORY = 600  # the 10k rain's full trajectory (steps)
HBM_PEAK_GBS = 8000.0


def parse() -> argparse.Namespace:
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20, help="steps of the timed window sub-record")
    p.add_argument("--warmup", type=int, default=5, help="untimed warm-up steps (separate world)")
    p.add_argument(
        "--cpu-steps",
        type=int,
        default=TRAJECTORY,
        help="reference steps timed on the host (steps 1..N of the same trajectory)",
    )
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-roofline", action="store_true")
    p.add_argument("--no-batch", action="store_true", help="skip the batched-rollback measurement")
    p.add_argument("--batch-worlds", type=int, default=256)
    p.add_argument("--batch-only", action="store_true", help="run only the batched-rollback measurement")
    p.add_argument("--no-hull-chains", action="store_true", help="skip the hull-rain-with-chains measurement")
    p.add_argument("--no-adjoint", action="store_true", help="skip the toss adjoint measurement")
    p.add_argument("--no-events", action="store_true", help="skip the event-stream read measurement")
    p.add_argument("--events-only", action="store_true", help="run only the event-stream read measurement")
    p.add_argument("--adjoint-only", action="store_true", help="run only the toss adjoint measurement")
    p.add_argument(
        "--hull-chains-only", action="store_true", help="run only the hull-rain-with-chains measurement"
    )
    p.add_argument(
        "--copy-calibration",
        action="store_true",
        help="run only a plain 256 MiB device copy (calibrates HBM byte counters)",
    )
    return p.parse_args()


def golden_hashes() -> dict:
    """Reference world hashes of the 10k rain (every 10th step)."""
    with open(os.path.join(ROOT, "tests", "golden", "hashes.json")) as f:
        return json.load(f)["scenes"]["rain"]["hashes"]


def run_trajectory() -> dict:
    """The headline: a fresh 10k rain stepped through its whole 600-step
    trajectory, timed from the first step to the last (the scene is uploaded
    and resident before the clock starts); the final hash is checked."""
    import maul3d_amd
    from maul3d_amd import scenes

    gpu = maul3d_amd.get_api()
    lib = maul3d_amd.library()
    world = scenes.rain(gpu, count=BODIES)
    lib.m3World_Sync(world)  # upload: the world is resident before timing
    t0 = time.perf_counter()
    for _ in range(TRAJECTORY):
        gpu.m3World_Step(world, DT, SUBSTEPS)
    lib.m3World_Sync(world)
    wall = time.perf_counter() - t0
    h = "%016x" % gpu.m3World_Hash(world)
    gpu.m3DestroyWorld(world)
    return {"wall": wall, "hash": h}


def run_window(warmup: int, steps: int, fixtures: dict) -> dict:
    """The requested window: ``warmup`` untimed steps, then ``steps`` timed
    steps of the same trajectory on its own world; afterwards the world is
    stepped (untimed) to the next step with a reference fixture and its hash
    compared there."""
    import maul3d_amd
    from maul3d_amd import scenes

    gpu = maul3d_amd.get_api()
    lib = maul3d_amd.library()
    world = scenes.rain(gpu, count=BODIES)
    lib.m3World_Sync(world)
    for _ in range(warmup):
        gpu.m3World_Step(world, DT, SUBSTEPS)
    lib.m3World_Sync(world)
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu.m3World_Step(world, DT, SUBSTEPS)
    lib.m3World_Sync(world)
    wall = time.perf_counter() - t0
    at = warmup + steps
    check = next((k for k in range(at, TRAJECTORY + 1) if str(k) in fixtures), None)
    exact = None
    if check is not None:
        for _ in range(check - at):
            gpu.m3World_Step(world, DT, SUBSTEPS)
        exact = "%016x" % gpu.m3World_Hash(world) == fixtures[str(check)]
    gpu.m3DestroyWorld(world)
    return {
        "steps": f"{warmup + 1}-{at}" if steps > 0 else "none",
        "warmup": warmup,
        "timed_steps": steps,
        "ms_per_step": 1000.0 * wall / steps if steps > 0 else None,
        "value": BODIES * SUBSTEPS * steps / wall if steps > 0 and wall > 0 else None,
        "hash_checked_at_step": check,
        "hash_bit_exact": exact,
    }


def run_roofline() -> dict | None:
    """Event-timed per-kernel averages over a second, profiled run of the
    same 600-step trajectory (every step waits for its stage events)."""
    import maul3d_amd
    from maul3d_amd import profile, scenes

    gpu = maul3d_amd.get_api()
    lib = maul3d_amd.library()
    world = scenes.rain(gpu, count=BODIES)
    lib.m3World_Sync(world)
    stats = profile.profile_steps(world, TRAJECTORY)
    gpu.m3DestroyWorld(world)
    return stats


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo) and the logical CPUs this
    process may run on."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return f"{name} ({n} logical CPUs available to this process)"


def host_threads() -> int:
    """Threads for the threaded reference: the host cores this process may
    use, at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _time_reference(ref, steps: int) -> tuple[float, str]:
    from maul3d_amd import scenes

    world = scenes.rain(ref, count=BODIES)
    t0 = time.perf_counter()
    for _ in range(steps):
        ref.m3World_Step(world, DT, SUBSTEPS)
    wall = time.perf_counter() - t0
    h = "%016x" % ref.m3World_Hash(world)
    ref.m3DestroyWorld(world)
    return wall, h


def run_cpu(steps: int) -> dict | None:
    """The reference engine on the host, serial (its default, no task hooks)
    and threaded (its task hooks served by oracle/task_pool.c: parallel
    narrowphase and colored solve, include/maul3d/world.h:14-26), on steps
    1..steps of the same rain.  ``value`` is the faster of the two."""
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libmaul3d_ref.so")
    if not os.path.exists(ref_path):
        return None
    from maul3d_amd import api

    ref = api.Api(C.CDLL(ref_path))
    serial_wall, serial_hash = _time_reference(ref, steps)
    rate = lambda wall: BODIES * SUBSTEPS * steps / wall  # noqa: E731
    out = {
        "value": rate(serial_wall),
        "unit": "body-substeps/s",
        "cores": 1,
        "kind": "reference",
        "cpu": cpu_model(),
        "steps": steps,
        "serial_value": rate(serial_wall),
        "serial_ms_per_step": 1000.0 * serial_wall / steps,
        "serial_hash": serial_hash,
        "sample": f"reference engine (oracle/_ref) on steps 1-{steps} of the same 10k rain; "
        f"serial (no task hooks) {serial_wall:.1f} s",
    }
    pool_path = os.path.join(ROOT, "oracle", "libm3pool.so")
    if not os.path.exists(pool_path):
        return out
    threads = host_threads()
    pool_lib = C.CDLL(pool_path)
    for name in ("m3pool_create", "m3pool_enqueue_fn", "m3pool_finish_fn"):
        getattr(pool_lib, name).restype = C.c_void_p
    pool_lib.m3pool_create.argtypes = [C.c_int32]
    pool_lib.m3pool_destroy.argtypes = [C.c_void_p]
    pool = pool_lib.m3pool_create(threads)
    if not pool:
        return out
    ref.task_hooks = (pool_lib.m3pool_enqueue_fn(), pool_lib.m3pool_finish_fn(), pool)
    try:
        thr_wall, thr_hash = _time_reference(ref, steps)
    finally:
        ref.task_hooks = None
        pool_lib.m3pool_destroy(pool)
    out["threaded_value"] = rate(thr_wall)
    out["threaded_ms_per_step"] = 1000.0 * thr_wall / steps
    out["threaded_cores"] = threads
    out["threaded_hash_equals_serial"] = thr_hash == serial_hash
    out["sample"] += f", threaded on {threads} host threads {thr_wall:.1f} s"
    if thr_wall < serial_wall:
        out["value"] = rate(thr_wall)
        out["cores"] = threads
    return out


def run_events_read(steps: int = 20, warm: int = 200) -> dict:
    """Cost of reading the event streams on the 10k rain: after ``warm``
    steps (bodies landing), each of ``steps`` steps is synced, then every
    stream is read (the device builds them at the first read).  Reports the
    read time per step and the stream sizes of the last step."""
    import maul3d_amd
    from maul3d_amd import scenes

    gpu = maul3d_amd.get_api()
    lib = maul3d_amd.library()
    world = scenes.rain(gpu, count=BODIES)
    for _ in range(warm):
        gpu.m3World_Step(world, DT, SUBSTEPS)
    total = 0.0
    sizes = {}
    for _ in range(steps):
        gpu.m3World_Step(world, DT, SUBSTEPS)
        lib.m3World_Sync(world)
        t0 = time.perf_counter()
        for name in ("ContactBeginEvents", "ContactEndEvents", "SensorBeginEvents", "SensorEndEvents",
                     "HitEvents", "BodyMoveEvents"):
            n = C.c_int32(0)
            getattr(gpu, "m3World_" + name)(world, C.byref(n))
            sizes[name] = n.value
        total += time.perf_counter() - t0
    gpu.m3DestroyWorld(world)
    return {
        "workload": f"10k rain, steps {warm + 1}-{warm + steps}: every event stream read after each step",
        "ms_per_read": 1000.0 * total / steps,
        "last_step_sizes": sizes,
    }


def batch_fixture() -> dict | None:
    try:
        with open(os.path.join(ROOT, "tests", "golden", "batch_rollback.json")) as f:
            return json.load(f)
    except OSError:
        return None


def run_batched(worlds: int) -> dict:
    """Batched rollback: ``worlds`` independent worlds of 50-2000 bodies
    (pyramids, box stacks, sphere/box rain) warmed 100 steps, snapshotted,
    restored and resimulated 8 steps.  Timed: the 8 resimulated steps, once
    through m3World_StepMany (every launch shared by all worlds) and once
    world by world through m3World_Step, waiting for each world before the
    next (the sequential single-world path).  Restores happen before each
    timed region; hashes are checked against the reference fixture.  Before
    timing, m3World_Autotune picks the batched substep kernel's workgroup
    size on the restored state (8-step trials per candidate)."""
    import maul3d_amd
    from maul3d_amd import api, scenes

    gpu = maul3d_amd.get_api()
    lib = maul3d_amd.library()
    fx = batch_fixture()
    ws = [scenes.batch_world(gpu, k) for k in range(worlds)]
    dyn = sum(gpu.m3World_GetCounters(w).bodyCount - 1 for w in ws)  # minus each ground
    ids = (api.WorldId * worlds)(*ws)
    warm = 100
    resim = 8
    for _ in range(warm):
        lib.m3World_StepMany(ids, worlds, C.c_float(DT), SUBSTEPS)
    snaps = [gpu.snapshot(w) for w in ws]

    def check() -> bool | None:
        if fx is None or worlds > len(fx["worlds"]):
            return None
        return all(
            "%016x" % gpu.m3World_Hash(w) == fx["worlds"][str(k)]["resim"][resim - 1]
            for k, w in enumerate(ws)
        )

    def restore_all() -> None:
        for w, snap in zip(ws, snaps):
            if not gpu.restore(w, snap):
                raise RuntimeError("restore refused")
            lib.m3World_Sync(w)

    # Launch autotuning on the restored state (every candidate starts from
    # the snapshot; the worlds come back restored), outside the timed region.
    tune = api.AutotuneResult()
    t_tune = time.perf_counter()
    ntune = gpu.m3World_Autotune(ids, worlds, DT, SUBSTEPS, resim, C.byref(tune))
    t_tune = time.perf_counter() - t_tune
    autotune = None
    if ntune > 0:
        autotune = {
            "threads_per_world_workgroup": tune.batchThreads,
            "candidates_ms_per_step": {
                str(tune.candidateA[k]): round(tune.candidateMs[k], 3) for k in range(ntune)
            },
            "trial_steps": resim,
            "seconds": round(t_tune, 2),
        }

    restore_all()
    t0 = time.perf_counter()
    for _ in range(resim):
        lib.m3World_StepMany(ids, worlds, C.c_float(DT), SUBSTEPS)
    for w in ws:
        lib.m3World_Sync(w)
    t_batch = time.perf_counter() - t0
    ok_batch = check()

    restore_all()
    t0 = time.perf_counter()
    for w in ws:
        for _ in range(resim):
            gpu.m3World_Step(w, DT, SUBSTEPS)
        lib.m3World_Sync(w)
    t_seq = time.perf_counter() - t0
    ok_seq = check()
    for w in ws:
        gpu.m3DestroyWorld(w)
    work = dyn * SUBSTEPS * resim
    return {
        "workload": f"{worlds} worlds, {dyn} dynamic bodies; restore at step {warm}, "
        f"{resim} steps resimulated",
        "value": work / t_batch,
        "unit": "body-substeps/s",
        "ms_per_step": 1000.0 * t_batch / resim,
        "sequential_value": work / t_seq,
        "sequential_ms_per_step": 1000.0 * t_seq / resim,
        "speedup_vs_sequential": t_seq / t_batch,
        "hash_bit_exact": (ok_batch and ok_seq) if ok_batch is not None else None,
        "autotune": autotune,
    }


def run_hull_chains(steps: int) -> dict:
    """Hull rain plus joint chains: 4000 QuickHull hulls and 500 chains of 20
    capsule links on spherical joints with 45-degree cones (14,000 dynamic
    bodies, 10,000 joints), one world, ``steps`` steps timed with the world
    resident; the final hash is checked against the reference fixture."""
    import maul3d_amd
    from maul3d_amd import scenes

    gpu = maul3d_amd.get_api()
    lib = maul3d_amd.library()
    world = scenes.hull_chains(gpu)
    dyn = gpu.m3World_GetCounters(world).bodyCount - 501  # ground and chain anchors are static
    lib.m3World_Sync(world)
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu.m3World_Step(world, DT, SUBSTEPS)
    lib.m3World_Sync(world)
    wall = time.perf_counter() - t0
    h = "%016x" % gpu.m3World_Hash(world)
    gpu.m3DestroyWorld(world)
    try:
        with open(os.path.join(ROOT, "tests", "golden", "hashes.json")) as f:
            want = json.load(f)["scenes"]["hull_chains"]["hashes"].get(str(steps))
    except OSError:
        want = None
    return {
        "workload": f"hull rain + chains: 4000 hulls, 500 chains x 20 links, {dyn} dynamic bodies, "
        f"10,000 spherical joints; {steps} steps",
        "value": dyn * SUBSTEPS * steps / wall,
        "unit": "body-substeps/s",
        "ms_per_step": 1000.0 * wall / steps,
        "world_hash": h,
        "hash_bit_exact": (h == want) if want else None,
    }


def run_toss_adjoint(worlds: int = 64, steps: int = 120) -> dict:
    """Differentiable toss: ``worlds`` worlds of 32 boxes thrown onto a plane,
    ``steps`` steps.  Forward: the worlds stepped together through
    m3World_StepMany (hashes checked against the reference fixture).
    Adjoint: m3World_Adjoint over the same worlds from the same initial
    state, loss |COM_T - target|^2 with gradients for every body's initial
    velocities, mass, friction and applied force, checkpoints every 10 steps
    with every earlier segment recomputed from its checkpoint (and, as a
    second figure, with all segment tapes kept resident).  Both timed from
    the same resident initial state, after one untimed warm-up call each."""
    import maul3d_amd
    from maul3d_amd import api, scenes

    gpu = maul3d_amd.get_api()
    lib = maul3d_amd.library()
    ws = [scenes.toss(gpu, seed=0x7055 + k) for k in range(worlds)]
    ids = (api.WorldId * worlds)(*ws)
    snaps = [gpu.snapshot(w) for w in ws]
    nb = 40

    def restore() -> None:
        for w, snap in zip(ws, snaps):
            if not gpu.restore(w, snap):
                raise RuntimeError("restore refused")
            lib.m3World_Sync(w)

    def forward() -> float:
        restore()
        t0 = time.perf_counter()
        for _ in range(steps):
            lib.m3World_StepMany(ids, worlds, C.c_float(DT), SUBSTEPS)
        for w in ws:
            lib.m3World_Sync(w)
        return time.perf_counter() - t0

    forward()
    t_fwd = forward()
    try:
        with open(os.path.join(ROOT, "tests", "golden", "hashes.json")) as f:
            fx = json.load(f)["toss_120"]
        exact = all("%016x" % gpu.m3World_Hash(w) == fx[str(k)] for k, w in enumerate(ws)) if steps == 120 else None
    except OSError:
        exact = None
    restore()
    defs = (api.AdjointDef * worlds)()
    res = (api.AdjointResult * worlds)()
    keep = []
    for k in range(worlds):
        d = gpu.m3DefaultAdjointDef()
        d.steps = steps
        d.checkpointEvery = 10
        d.target = api.Vec3(0.0, 0.5, 0.0)
        defs[k] = d
        arrs = [(api.Vec3 * nb)(), (api.Vec3 * nb)(), (api.Vec3 * nb)(), (C.c_float * nb)(), (C.c_float * nb)()]
        keep.append(arrs)
        res[k].gradLinearVelocity = C.cast(arrs[0], C.POINTER(api.Vec3))
        res[k].gradAngularVelocity = C.cast(arrs[1], C.POINTER(api.Vec3))
        res[k].gradForce = C.cast(arrs[2], C.POINTER(api.Vec3))
        res[k].gradMass = C.cast(arrs[3], C.POINTER(C.c_float))
        res[k].gradFriction = C.cast(arrs[4], C.POINTER(C.c_float))

    def adjoint(resident: str) -> float:
        os.environ["MAUL3D_AMD_ADJOINT_RESIDENT"] = resident
        t0 = time.perf_counter()
        if gpu.m3World_Adjoint(ids, worlds, defs, res) != worlds:
            raise RuntimeError("m3World_Adjoint refused the worlds")
        return time.perf_counter() - t0

    adjoint("1")
    t_adj = adjoint("1")
    adjoint("1000")
    t_res = adjoint("1000")
    os.environ.pop("MAUL3D_AMD_ADJOINT_RESIDENT", None)
    unsupported = sum(res[k].unsupported for k in range(worlds))
    for w in ws:
        gpu.m3DestroyWorld(w)
    return {
        "workload": f"{worlds} toss worlds x 32 boxes, {steps} steps, loss |COM_T - target|^2, gradients "
        "for initial linear/angular velocity, mass, friction, applied force",
        "forward_ms": 1000.0 * t_fwd,
        "adjoint_ms": 1000.0 * t_adj,
        "adjoint_forward_ratio": t_adj / t_fwd,
        "checkpoint_every": 10,
        "adjoint_resident_tapes_ms": 1000.0 * t_res,
        "adjoint_forward_ratio_resident_tapes": t_res / t_fwd,
        "forward_hash_bit_exact": exact,
        "unsupported": unsupported,
    }


def pmc_traffic() -> dict:
    """HBM bytes per substep launch from the committed counter runs
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each scaled by the same
    counter's reading for a copy of known size), with the window they cover
    and the modelled algorithmic bytes for that window."""
    path = os.path.join(ROOT, "profiles", "k_substeps_rain10k_pmc.md")
    fields = {}
    try:
        with open(path) as f:
            for ln in f:
                key, sep, val = ln.partition(":")
                if sep and key.strip() in ("traffic_bytes_per_launch", "algorithmic_bytes_per_launch", "window"):
                    fields[key.strip()] = val.strip()
    except OSError:
        return {"traffic": None}
    if "traffic_bytes_per_launch" not in fields:
        return {"traffic": None}
    return {
        "traffic": int(fields["traffic_bytes_per_launch"]),
        "traffic_window": fields.get("window"),
        "traffic_algorithmic_bytes": int(fields.get("algorithmic_bytes_per_launch", "0")) or None,
        "traffic_source": "profiles/k_substeps_rain10k_pmc.md",
    }


def copy_calibration() -> None:
    """A plain elementwise copy of known size, the scale for FETCH_SIZE and
    WRITE_SIZE counter readings of the step kernels."""
    import torch

    n = 64 << 20
    src = torch.ones(n, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}))


def main() -> None:
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(
            "bench.py: this build runs on exactly one GPU (worlds shard across GPUs "
            "with no exchange; no multi-GPU run is part of this build)",
            file=sys.stderr,
        )
        sys.exit(2)
    if args.batch_only:
        print(json.dumps({"batched": run_batched(args.batch_worlds)}))
        return
    if args.hull_chains_only:
        print(json.dumps({"hull_chains": run_hull_chains(args.steps)}))
        return
    if args.events_only:
        print(json.dumps({"events_read": run_events_read()}))
        return
    if args.adjoint_only:
        print(json.dumps({"toss_adjoint": run_toss_adjoint()}))
        return
    fixtures = golden_hashes()
    want = fixtures.get(str(TRAJECTORY))
    if want is None:
        print(f"bench.py: no reference fixture for step {TRAJECTORY} of the 10k rain", file=sys.stderr)
        sys.exit(1)
    # Untimed warm-up and the requested window, on a world of their own.
    window = run_window(max(args.warmup, 0), max(args.steps, 0), fixtures)
    res = run_trajectory()
    value = BODIES * SUBSTEPS * TRAJECTORY / res["wall"]
    line = {
        "metric": "body-substeps/sec (10k rain, single world)",
        "value": value,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": TRAJECTORY,
        "warmup": max(args.warmup, 0),
        "ms_per_step": 1000.0 * res["wall"] / TRAJECTORY,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded 10k rain scene)",
        "config": {
            "workload": "10k-body rain, single world: 10,000 dynamic bodies (40% boxes, "
            "40% spheres, 20% capsules), dt=1/60, 4 substeps, the full 600-step trajectory",
            "bodies": BODIES,
            "substeps": SUBSTEPS,
            "trajectory_steps": TRAJECTORY,
        },
        "world_hash": res["hash"],
        "hash_expected": want,
        "hash_bit_exact": res["hash"] == want,
        "window": window,
    }
    if not args.no_roofline:
        try:
            prof = run_roofline()
        except Exception as exc:  # reported, never silently dropped
            prof = None
            line["roofline_error"] = repr(exc)
        if prof is not None:
            line["roofline"] = prof["roofline"]
            line["roofline"]["window"] = f"steps 1-{TRAJECTORY} of the 10k rain"
            line["roofline"]["counts_per_step"] = prof["counts_per_step"]
            line["roofline"].update(pmc_traffic())
            line["kernels"] = prof["kernels"]
    if not args.no_batch:
        line["batched"] = run_batched(args.batch_worlds)
    if not args.no_hull_chains:
        line["hull_chains"] = run_hull_chains(600)
    if not args.no_adjoint:
        line["toss_adjoint"] = run_toss_adjoint()
    if not args.no_events:
        line["events_read"] = run_events_read()
    if not args.no_cpu:
        line["cpu_baseline"] = run_cpu(args.cpu_steps)
    print(json.dumps(line), flush=True)
    if not line["hash_bit_exact"]:
        print("bench.py: the step-600 world hash differs from the reference fixture", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
