"""Stage profile of the device step and the roofline of its dominant kernel.

Stage times are HIP events recorded on the world's own stream around every
launch (m3World_ProfileBegin/End).  Algorithmic bytes are the minimum HBM
traffic each stage must move for the work counts the device reports,
per unit:

* solve / relax, per constraint and substep: 370 B + 48 B per contact point
  (constraint record read, accumulators written, both bodies' velocities
  read and written, both bodies' substep position/rotation deltas read)
* warm start, per constraint and substep: 270 B + 28 B per point
* integrate velocities, per moving body and substep: 142 B
* integrate positions, per moving body and substep: 173 B
* narrowphase, per pair: 572 B (key, two transforms, centers of mass and
  shape records read, the previous manifold read, the 188 B manifold written)
* prepare, per constraint: 454 B read + the constraint record written
* substeps (the fused substep loop): the sum of the five substep stages
"""

from __future__ import annotations

import ctypes as C

from . import api

CONSTRAINT_BYTES = 460  # sizeof(m3Constraint) on the device


class StageTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("ms", C.c_double)]


class GpuProfile(C.Structure):
    _fields_ = [
        ("stageCount", C.c_int32),
        ("stage", StageTime * 64),
        ("steps", C.c_int64),
        ("sumPairs", C.c_int64),
        ("sumOldPairs", C.c_int64),
        ("sumConstraints", C.c_int64),
        ("sumPoints", C.c_int64),
        ("sumMovers", C.c_int64),
        ("sumSubsteps", C.c_int64),
        ("sumShapes", C.c_int64),
        ("sumBodies", C.c_int64),
        ("sumOverflow", C.c_int64),
    ]


def stage_bytes(name: str, p: GpuProfile) -> float | None:
    """Algorithmic bytes of one stage summed over the profiled steps."""
    sub = p.sumSubsteps / max(p.steps, 1)
    c, q, m = p.sumConstraints, p.sumPoints, p.sumMovers
    if name in ("solve", "relax"):
        return sub * (370.0 * c + 48.0 * q)
    if name == "warm_start":
        return sub * (270.0 * c + 28.0 * q)
    if name == "integrate_velocities":
        return sub * 142.0 * m
    if name == "integrate_positions":
        return sub * 173.0 * m
    if name == "substeps":
        return sub * ((370.0 * 2 + 270.0) * c + (48.0 * 2 + 28.0) * q + (142.0 + 173.0) * m)
    if name == "narrowphase":
        return 572.0 * p.sumPairs
    if name == "prepare":
        return (454.0 + CONSTRAINT_BYTES) * c
    return None


def _lib():
    import maul3d_amd

    lib = maul3d_amd.library()
    lib.m3World_ProfileBegin.restype = C.c_int32
    lib.m3World_ProfileBegin.argtypes = [api.WorldId]
    lib.m3World_ProfileEnd.restype = C.c_int32
    lib.m3World_ProfileEnd.argtypes = [api.WorldId, C.POINTER(GpuProfile)]
    return lib


def profile_steps(world: api.WorldId, steps: int, dt: float = 1.0 / 60.0, substeps: int = 4,
                  hbm_peak_gbs: float = 8000.0) -> dict:
    import maul3d_amd

    gpu = maul3d_amd.get_api()
    lib = _lib()
    if lib.m3World_ProfileBegin(world) != 0:
        raise RuntimeError("m3World_ProfileBegin failed")
    for _ in range(steps):
        gpu.m3World_Step(world, dt, substeps)
    prof = GpuProfile()
    if lib.m3World_ProfileEnd(world, C.byref(prof)) != 0:
        raise RuntimeError("m3World_ProfileEnd failed")
    kernels = {}
    for k in range(prof.stageCount):
        st = prof.stage[k]
        name = st.name.decode()
        kernels[name] = {
            "launches": int(st.launches),
            "ms_total": round(st.ms, 3),
            "us_per_launch": round(1000.0 * st.ms / max(st.launches, 1), 2),
        }
    modelled = {n: v for n, v in kernels.items() if stage_bytes(n, prof) is not None}
    dominant = max(modelled, key=lambda n: modelled[n]["ms_total"])
    dom = kernels[dominant]
    bytes_total = stage_bytes(dominant, prof)
    per_launch = bytes_total / max(dom["launches"], 1)
    avg_s = dom["ms_total"] / 1000.0 / max(dom["launches"], 1)
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    total_ms = sum(v["ms_total"] for v in kernels.values())
    return {
        "kernels": kernels,
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": hbm_peak_gbs,
            "unit": "GB/s",
            "frac": round(achieved / hbm_peak_gbs, 5),
            "traffic": None,
            "bytes_per_launch": round(per_launch, 1),
            "us_per_launch": dom["us_per_launch"],
            "share_of_step_time": round(dom["ms_total"] / total_ms, 3) if total_ms else None,
        },
        "counts_per_step": {
            "pairs": prof.sumPairs / max(prof.steps, 1),
            "constraints": prof.sumConstraints / max(prof.steps, 1),
            "points": prof.sumPoints / max(prof.steps, 1),
            "movers": prof.sumMovers / max(prof.steps, 1),
            "overflow_constraints": prof.sumOverflow / max(prof.steps, 1),
        },
    }
