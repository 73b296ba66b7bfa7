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


POOL_LIB = os.path.join(ROOT, "oracle", "libm3pool.so")


def threaded_ref_api(threads: int):
    """The reference engine with its task hooks (include/maul3d/world.h:14-26)
    served by a pool of ``threads`` host threads; returns (api, destroy)."""
    import pytest

    if not os.path.exists(POOL_LIB):
        pytest.skip("oracle/libm3pool.so not built (make -C oracle restated)")
    ref = ref_api()
    pool_lib = C.CDLL(POOL_LIB)
    for name in ("m3pool_create", "m3pool_enqueue_fn", "m3pool_finish_fn"):
        getattr(pool_lib, name).restype = C.c_void_p
    pool_lib.m3pool_create.argtypes = [C.c_int32]
    pool_lib.m3pool_destroy.argtypes = [C.c_void_p]
    pool = pool_lib.m3pool_create(threads)
    assert pool
    ref.task_hooks = (pool_lib.m3pool_enqueue_fn(), pool_lib.m3pool_finish_fn(), pool)
    return ref, lambda: pool_lib.m3pool_destroy(pool)


def product_api():
    import maul3d_amd

    return maul3d_amd.get_api()


def hexh(h: int) -> str:
    return "%016x" % h
