"""maul3d_amd: Maul3D's soft-step contact solver, MI355X-native.

The product is ``libmaul3d_amd.so`` (built in-tree by :mod:`maul3d_amd.build`):
the Maul3D C API (``m3CreateWorld``, ``m3World_Step``, ``m3World_Snapshot``,
``m3World_Restore``, ``m3World_Hash`` ...) over a C17 host world and
hand-written gfx950 kernels.  This module only locates and loads it; there
is no Python or CPU fallback for the step.
"""

from __future__ import annotations

import ctypes as C
import os

from . import api, scenes  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmaul3d_amd.so")

_lib: C.CDLL | None = None
_api: api.Api | None = None


def library() -> C.CDLL:
    """Load the built library, failing loudly when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m maul3d_amd.build` "
                "(the step has no CPU fallback)"
            )
        _lib = C.CDLL(LIB_PATH)
        _lib.m3World_StepMany.restype = C.c_int32
        _lib.m3World_StepMany.argtypes = [C.POINTER(api.WorldId), C.c_int32, C.c_float, C.c_int32]
        _lib.m3World_Sync.restype = None
        _lib.m3World_Sync.argtypes = [api.WorldId]
        _lib.m3gpu_device_count.restype = C.c_int
        _lib.m3gpu_device_count.argtypes = []
    return _lib


def get_api() -> api.Api:
    global _api
    if _api is None:
        _api = api.Api(library())
    return _api


def device_count() -> int:
    return int(library().m3gpu_device_count())
