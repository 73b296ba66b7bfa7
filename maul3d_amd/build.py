"""Builds libmaul3d_amd.so in-tree: the C17 host world (gcc) plus the gfx950
device layer (hipcc), linked into one shared library beside this file.

Flags: the host world uses the reference FP law (-ffp-contract=off, no fast
math); the HIP code uses the same contraction rule, keeps IEEE division and
square root (hipcc's default correctly-rounded forms) and disables packed
FP32 instruction selection.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmaul3d_amd.so")
ARCH = os.environ.get("MAUL3D_AMD_ARCH", "gfx950")

HOST_FLAGS = [
    "-std=c17",
    "-O2",
    "-fPIC",
    "-fvisibility=hidden",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-unsafe-math-optimizations",
    "-fno-trapping-math",
    "-Wall",
    "-Wextra",
    "-Wno-unused-parameter",
]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-std=c++17",
    "-O3",
    "-fPIC",
    "-fvisibility=hidden",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-parameter",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    world_o = os.path.join(BUILD, "world.o")
    hull_o = os.path.join(BUILD, "hull_build.o")
    tree_o = os.path.join(BUILD, "m3_tree.o")
    replay_o = os.path.join(BUILD, "replay.o")
    dev_o = os.path.join(BUILD, "m3_device.o")
    _run(["gcc", *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, "world.c"), "-o", world_o])
    _run(["gcc", *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, "hull_build.c"), "-o", hull_o])
    _run(["gcc", *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, "m3_tree.c"), "-o", tree_o])
    _run(["gcc", *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, "replay.c"), "-o", replay_o])
    _run([_hipcc(), *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, "m3_device.hip"), "-o", dev_o])
    tmp = LIB + ".tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp,
          world_o, hull_o, tree_o, replay_o, dev_o])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
