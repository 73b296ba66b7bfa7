"""Shared test plumbing: library handles, fixtures and step loops."""

from __future__ import annotations

import ctypes as C
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libmaul3d_ref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "hashes.json")
DT = 1.0 / 60.0
SUBSTEPS = 4


def golden() -> dict:
    with open(GOLDEN) as f:
        return json.load(f)


def ref_api():
    """The reference engine built from its own sources (the checker)."""
    import pytest

    from maul3d_amd import api

    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (make -C oracle ref)")
    return api.Api(C.CDLL(REF_LIB))


def product_api():
    import maul3d_amd

    return maul3d_amd.get_api()


def hexh(h: int) -> str:
    return "%016x" % h
