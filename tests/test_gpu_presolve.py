"""The pre-solve veto on the MI355X path: the device pauses after its island
wake pass, the host calls the callback over the device's candidates in pair
order, and the vetoed pairs make no constraint (src/solver.c:184-248).
Checked against the live reference every step, and through the journal."""

from __future__ import annotations

import pytest

from tests.helpers import product_api, ref_api
from tests.test_presolve import bare_replay, record_with_vetoes, run_against_reference

pytestmark = pytest.mark.gpu


def test_presolve_matches_live_reference_every_step():
    run_against_reference(product_api())


def test_device_vetoes_are_journaled_and_replay_bare():
    ref = ref_api()
    gpu = product_api()
    jr, hr = record_with_vetoes(ref)
    jg, hg = record_with_vetoes(gpu)
    assert jg == jr
    assert hg == hr
    assert bare_replay(gpu, jr) == hr
    assert bare_replay(ref, jg) == hr
