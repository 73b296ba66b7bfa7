"""m3gpu_events_noalloc against m3gpu_events (include/m3gpu.h): the same
event streams read twice in a row on the MI355X, the second time from a
smaller world, through the C ABI.  The allocation-free variant's scratch is
one file-scope device array per GPU that keeps the last call's contents, so
the second, smaller call also shows that nothing of the first leaks in."""

from __future__ import annotations

import ctypes as C

import pytest

from tests.helpers import DT, SUBSTEPS, product_api

from maul3d_amd import scenes

pytestmark = pytest.mark.gpu

PAIRS = 32768
BODIES = 8192


class Hit(C.Structure):
    _fields_ = [("key", C.c_uint64), ("point", C.c_double * 3), ("normal", C.c_float * 3), ("speed", C.c_float)]


class Move(C.Structure):
    _fields_ = [("body", C.c_int32), ("fellAsleep", C.c_uint8), ("pad", C.c_uint8 * 3), ("p", C.c_double * 3), ("q", C.c_float * 4)]


class Events(C.Structure):
    _fields_ = [
        ("beginKeys", C.POINTER(C.c_uint64)),
        ("endKeys", C.POINTER(C.c_uint64)),
        ("sensorBeginKeys", C.POINTER(C.c_uint64)),
        ("sensorEndKeys", C.POINTER(C.c_uint64)),
        ("hits", C.POINTER(Hit)),
        ("moves", C.POINTER(Move)),
        ("counts", C.c_int32 * 7),
    ]


def read(fn, dev, threshold: float):
    keys = [(C.c_uint64 * PAIRS)() for _ in range(4)]
    hits = (Hit * PAIRS)()
    moves = (Move * BODIES)()
    out = Events(*[C.cast(k, C.POINTER(C.c_uint64)) for k in keys], C.cast(hits, C.POINTER(Hit)),
                 C.cast(moves, C.POINTER(Move)))
    assert fn(dev, C.c_float(threshold), C.byref(out)) == 0
    counts = list(out.counts)
    streams = [list(keys[k][: counts[k]]) for k in range(4)]
    streams.append(bytes(C.string_at(hits, counts[4] * C.sizeof(Hit))))
    streams.append(bytes(C.string_at(moves, counts[5] * C.sizeof(Move))))
    return counts, streams


def test_events_noalloc_equal_events_twice_in_a_row():
    gpu = product_api()
    lib = gpu.lib
    for fn in (lib.m3gpu_events, lib.m3gpu_events_noalloc):
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_float, C.POINTER(Events)]
    seen = []
    for count in (400, 150):  # the second input smaller
        w = scenes.event_rain(gpu, count=count)
        for _ in range(70):
            gpu.m3World_Step(w, DT, SUBSTEPS)
        dev = gpu.m3World_DeviceWorld(w)
        assert dev
        want = read(lib.m3gpu_events, dev, 1.0)
        got = read(lib.m3gpu_events_noalloc, dev, 1.0)
        assert got == want
        assert sum(want[0][:6]) > 0  # the step produced events to compare
        seen.append(want[0])
        gpu.m3DestroyWorld(w)
    assert seen[1][5] < seen[0][5]  # fewer moves: the second world is smaller
