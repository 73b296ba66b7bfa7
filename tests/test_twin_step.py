"""The float instantiation of this build's step functions (the CPU twin in
oracle/twin: the same m3_step.h stage functions the device kernels call,
driven by the serial step) against the reference build, hash by hash.

The device-only parts (parallel coloring, the substep kernel's per-body
warm start, barriers, gathering) are checked on the GPU by
test_gpu_parity.py; this test pins, on the CPU, the arithmetic those
kernels share with the twin: narrowphase, prepare, the register-resident
contact solve with its branch-free point loop, integration, sleep.
"""

from __future__ import annotations

import ctypes as C
import os

import pytest

from tests.helpers import ROOT

from maul3d_amd import api, scenes

REF = os.path.join(ROOT, "oracle", "_ref", "libmaul3d_ref.so")
TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")


def _lib(path):
    if not os.path.exists(path):
        pytest.skip(f"{os.path.basename(path)} not built (python -c 'import __graft_entry__ as g; g.build()')")
    return api.Api(C.CDLL(path))


CASES = [
    ("toss0", lambda a: scenes.toss(a, 0), 120),
    ("toss3", lambda a: scenes.toss(a, 3), 120),
    ("box_pyramid", scenes.box_pyramid, 150),
    ("batch7", lambda a: scenes.batch_world(a, 7), 200),
    ("hull_pile", lambda a: scenes.hull_pile(a, hulls=60), 200),
    ("crowded_plank", scenes.crowded_plank, 150),
]


@pytest.mark.parametrize("name,make,steps", CASES, ids=[c[0] for c in CASES])
def test_twin_step_matches_reference_every_step(name, make, steps):
    ref = _lib(REF)
    twin = _lib(TWIN32)
    wr = make(ref)
    wt = make(twin)
    try:
        for i in range(steps):
            ref.m3World_Step(wr, 1.0 / 60.0, 4)
            twin.m3World_Step(wt, 1.0 / 60.0, 4)
            assert ref.m3World_Hash(wr) == twin.m3World_Hash(wt), f"{name}: hashes differ at step {i + 1}"
    finally:
        ref.m3DestroyWorld(wr)
        twin.m3DestroyWorld(wt)
