"""The torch.autograd wrapper of the adjoint (maul3d_amd.autograd), on the
FP64 CPU twin: backward hands out exactly what m3World_Adjoint returns, and a
descent step on the initial velocities lowers the loss."""

from __future__ import annotations

import ctypes as C
import os

import pytest

from tests.helpers import ROOT

from maul3d_amd import api, scenes

torch = pytest.importorskip("torch")

TWIN64 = os.path.join(ROOT, "oracle", "_ref", "libm3twin64.so")
NB = 33


def _twin():
    if not os.path.exists(TWIN64):
        pytest.skip("oracle twin not built (make -C oracle/twin)")
    return api.Api(C.CDLL(TWIN64))


def _state(T, w):
    lin, ang = [], []
    for b in range(NB):
        body = api.BodyId(b + 1, w.index1 - 1, 0)
        v = T.m3Body_GetLinearVelocity(body)
        a = T.m3Body_GetAngularVelocity(body)
        lin.append([v.x, v.y, v.z])
        ang.append([a.x, a.y, a.z])
    return torch.tensor(lin), torch.tensor(ang)


def test_backward_returns_adjoint_gradients_and_descends():
    from maul3d_amd.autograd import _adjoint, trajectory_loss

    T = _twin()
    w = scenes.toss(T, 2)
    lin, ang = _state(T, w)
    force = torch.zeros(NB, 3)
    mass = torch.ones(NB)
    fric = torch.ones(NB)
    target = (0.5, 0.3, -1.0)
    steps = 30
    for t in (lin, ang, force, mass, fric):
        t.requires_grad_(True)
    loss = trajectory_loss(T, w, target, lin, ang, force, mass, fric, steps=steps, checkpoint_every=10)
    loss.backward()
    ref = _adjoint(T, w, NB, target, [[0.0, 0.0, 0.0]] * NB, steps, 10, 1.0 / 60.0, 4)
    assert float(loss) == ref[0]
    assert torch.equal(lin.grad, torch.tensor(ref[1]))
    assert torch.equal(ang.grad, torch.tensor(ref[2]))
    assert torch.equal(mass.grad, torch.tensor(ref[4]))
    # one small descent step on the initial linear velocities lowers the loss
    with torch.no_grad():
        lin2 = lin - 0.05 * lin.grad
    loss2 = trajectory_loss(T, w, target, lin2, ang.detach(), force.detach(), mass.detach(),
                            fric.detach(), steps=steps)
    assert float(loss2) < float(loss)
    T.m3DestroyWorld(w)
