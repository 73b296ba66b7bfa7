"""Trajectory gradients on the device against the CPU twin.

The device runs the serial step and the hand-written reverse pass one
thread per world at float precision; the float twin (oracle/twin) runs the
same stage functions on the CPU (and is itself bit-exact with the reference
engine's trajectories), so losses must agree exactly and gradients to float
rounding.  Gradients of the same code at double precision are checked
against FP64 finite differences in tests/test_adjoint.py.
"""

from __future__ import annotations

import ctypes as C
import os

import pytest

from tests.helpers import DT, ROOT, SUBSTEPS, product_api

from maul3d_amd import api, scenes

pytestmark = pytest.mark.gpu

TWIN32 = os.path.join(ROOT, "oracle", "_ref", "libm3twin32.so")
NB = 40
NAMES = ("gradLinearVelocity", "gradAngularVelocity", "gradForce", "gradMass", "gradFriction")


def _twin():
    if not os.path.exists(TWIN32):
        pytest.skip("oracle twin not built (make -C oracle/twin)")
    return api.Api(C.CDLL(TWIN32))


def run(T, worlds, steps=120, forward=False):
    n = len(worlds)
    defs = (api.AdjointDef * n)()
    res = (api.AdjointResult * n)()
    keep = []
    for k in range(n):
        d = T.m3DefaultAdjointDef()
        d.steps = steps
        d.checkpointEvery = 10
        d.target = api.Vec3(0.5 * k, 0.4, -1.0)
        d.forwardOnly = 1 if forward else 0
        defs[k] = d
        if not forward:
            arrs = {}
            for name in NAMES[:3]:
                a = (api.Vec3 * NB)()
                arrs[name] = a
                setattr(res[k], name, C.cast(a, C.POINTER(api.Vec3)))
            for name in NAMES[3:]:
                a = (C.c_float * NB)()
                arrs[name] = a
                setattr(res[k], name, C.cast(a, C.POINTER(C.c_float)))
            keep.append(arrs)
    ids = (api.WorldId * n)(*worlds)
    assert T.m3World_Adjoint(ids, n, defs, res) == n
    return res, keep


def flat(arrs):
    out = []
    for name in NAMES[:3]:
        for v in arrs[name]:
            out += [v.x, v.y, v.z]
    for name in NAMES[3:]:
        out += list(arrs[name])
    return out


def test_device_gradients_equal_float_twin():
    gpu = product_api()
    T = _twin()
    seeds = [0, 1, 2, 3]
    wg = [scenes.toss(gpu, s) for s in seeds]
    wt = [scenes.toss(T, s) for s in seeds]
    rg, ag = run(gpu, wg)
    rt, at = run(T, wt)
    for k in range(len(seeds)):
        assert rg[k].loss == rt[k].loss, k
        assert rg[k].unsupported == rt[k].unsupported == 0
        fg, ft = flat(ag[k]), flat(at[k])
        for a, b in zip(fg, ft):
            assert abs(a - b) <= 1e-5 * max(1.0, abs(b)), (k, a, b)
    for w in wg:
        gpu.m3DestroyWorld(w)
    for w in wt:
        T.m3DestroyWorld(w)


def test_device_adjoint_leaves_worlds_stepping_like_before():
    gpu = product_api()
    a = scenes.toss(gpu, 9)
    b = scenes.toss(gpu, 9)
    for _ in range(20):
        gpu.m3World_Step(a, DT, SUBSTEPS)
        gpu.m3World_Step(b, DT, SUBSTEPS)
    h = gpu.m3World_Hash(a)
    run(gpu, [a], steps=50)
    assert gpu.m3World_Hash(a) == h
    for i in range(30):
        gpu.m3World_Step(a, DT, SUBSTEPS)
        gpu.m3World_Step(b, DT, SUBSTEPS)
        assert gpu.m3World_Hash(a) == gpu.m3World_Hash(b), i
    gpu.m3DestroyWorld(a)
    gpu.m3DestroyWorld(b)


def test_device_losses_of_64_worlds_equal_float_twin():
    gpu = product_api()
    T = _twin()
    wg = [scenes.toss(gpu, s) for s in range(64)]
    wt = [scenes.toss(T, s) for s in range(64)]
    rg, _ = run(gpu, wg, forward=True)
    rt, _ = run(T, wt, forward=True)
    assert [rg[k].loss for k in range(64)] == [rt[k].loss for k in range(64)]
    for w in wg:
        gpu.m3DestroyWorld(w)
    for w in wt:
        T.m3DestroyWorld(w)


def test_trajectory_loss_on_device_equals_float_twin():
    """The torch wrapper on the MI355X path: edited masses and frictions are
    written into the world (read back), and its loss and every gradient
    equal the float twin's for the same inputs."""
    torch = pytest.importorskip("torch")
    from maul3d_amd.autograd import trajectory_loss, world_parameters

    out = []
    for m3 in (product_api(), _twin()):
        w = scenes.toss(m3, 6)
        nb = 33
        mass, fric = world_parameters(m3, w, nb)
        mass[5] *= 1.5
        fric[9] = 0.2
        lin = torch.zeros(nb, 3)
        lin[:, 1] = 2.0
        ang = torch.zeros(nb, 3)
        force = torch.zeros(nb, 3)
        force[3] = torch.tensor([0.0, 5.0, 1.0])
        for t in (lin, ang, force, mass, fric):
            t.requires_grad_(True)
        loss = trajectory_loss(m3, w, (0.5, 0.3, -1.0), lin, ang, force, mass, fric, steps=60)
        loss.backward()
        now_m, now_f = world_parameters(m3, w, nb)
        assert float(now_m[5]) == pytest.approx(float(mass[5]), rel=1e-6)
        assert float(now_f[9]) == pytest.approx(0.2, rel=1e-7)
        out.append((float(loss.detach()), [t.grad.clone() for t in (lin, ang, force, mass, fric)]))
        m3.m3DestroyWorld(w)
    (lg, gg), (lt, gt) = out
    assert lg == lt
    for a, b in zip(gg, gt):
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-5)
