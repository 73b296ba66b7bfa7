"""M3J1 sessions on the MI355X path: a session recorded on the device
(snapshot + journal + final hash) verifies on the reference, and a reference
session verifies on the device (include/maul3d/replay.h, src/replay.c); the
divergence finder returns the reference's rows for two device worlds."""

from __future__ import annotations

import pytest

from tests.helpers import product_api, ref_api
from tests.test_replay import _diff_rows, encode, record_session, verify

from maul3d_amd import api, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("recorder", ["reference", "device"])
def test_session_container_verifies_on_the_other_library(recorder):
    ref = ref_api()
    gpu = product_api()
    rec, other = (ref, gpu) if recorder == "reference" else (gpu, ref)
    snap, journal, final = record_session(rec, mirror=rec is gpu)
    container = encode(rec, snap, journal, final)
    assert verify(other, container)
    assert verify(rec, container)


def test_diff_report_equals_reference():
    from tests.helpers import DT, SUBSTEPS

    results = []
    for m3 in (ref_api(), product_api()):
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
