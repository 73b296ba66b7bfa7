"""Trajectory gradients on the device against the CPU twin and FP64 differences.

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


def test_device_gradients_track_fp64_differences_on_all_64_worlds():
    """The device's gradients (precision 32, the default) on the full toss
    configuration, all 64 worlds and 120 steps, against FP64 central
    differences of the FP64 twin's loss on two random parameters per world
    (contact-stable samples: steps eps and 4 eps agree to 1e-5).  The device
    differentiates its own float trajectory, which drifts from the double
    one over 120 contact-rich steps, so a sample may miss the FP64 slope by
    more than 1e-3: at least two thirds of the stable samples must match to
    1e-3 and every one to 5e-2 (the float twin, bit-equal to the device,
    gives 7 of 9 within 1e-3 and a worst of 2.7e-2 on these samples)."""
    import random

    from tests.test_adjoint import EPS, KINDS, TWIN64
    from tests.test_adjoint import adjoint as twin_adjoint

    if not os.path.exists(TWIN64):
        pytest.skip("oracle twin not built (make -C oracle/twin)")
    T64 = api.Api(C.CDLL(TWIN64))
    gpu = product_api()
    seeds = list(range(64))
    wg = [scenes.toss(gpu, s) for s in seeds]
    n = len(seeds)
    defs = (api.AdjointDef * n)()
    res = (api.AdjointResult * n)()
    keep = []
    for k in range(n):
        d = gpu.m3DefaultAdjointDef()
        d.steps = 120
        d.checkpointEvery = 10
        d.target = api.Vec3(1.0, 0.5, -2.0)
        defs[k] = d
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
    ids = (api.WorldId * n)(*wg)
    assert gpu.m3World_Adjoint(ids, n, defs, res) == n
    rng = random.Random(0xF32)
    errs = []
    for k, s in enumerate(seeds):
        assert res[k].unsupported == 0
        w64 = scenes.toss(T64, s)
        for _ in range(2):
            kind = rng.choice([1, 2, 3, 4, 5])
            body = rng.randrange(1, 33)
            comp = rng.randrange(3) if kind in (1, 2, 5) else 0
            eps = EPS[kind]
            L = lambda e: twin_adjoint(T64, w64, 120, (kind, body, comp, e), True, every=10)[0].loss  # noqa: E731
            fd = (L(eps) - L(-eps)) / (2 * eps)
            fd4 = (L(4 * eps) - L(-4 * eps)) / (8 * eps)
            scale = max(abs(fd), 1e-6)
            if abs(fd - fd4) / scale > 1e-5:
                continue
            a = keep[k][KINDS[kind]][body]
            dev = (a.x, a.y, a.z)[comp] if kind in (1, 2, 5) else a
            errs.append((abs(dev - fd) / scale, s, KINDS[kind], body, comp, dev, fd))
        T64.m3DestroyWorld(w64)
    for w in wg:
        gpu.m3DestroyWorld(w)
    assert len(errs) >= 6, f"only {len(errs)} contact-stable samples"
    assert max(e[0] for e in errs) < 5e-2, max(errs)
    assert 3 * sum(e[0] < 1e-3 for e in errs) >= 2 * len(errs), sorted(errs, reverse=True)[:4]


@pytest.mark.skip(reason="precision 64 is switched off on the device: its first MI355X run ended in a "
                  "GPU memory fault whose cause is not yet shown (its driver runs clean on the CPU)")
def test_fp64_device_gradients_match_fp64_differences_on_all_64_worlds():
    """Precision 64 (the trajectory and its reverse in double precision on
    the device) on the full toss configuration: all 64 worlds, 120 steps.
    Every gradient of every world equals the FP64 twin's adjoint to float
    output rounding, the losses agree, and on two random parameters per
    world the device gradient matches FP64 central differences (taken on
    the FP64 twin, steps eps and 4 eps agreeing to 1e-5: contact-stable) to
    1e-3 relative."""
    import random

    from tests.test_adjoint import EPS, KINDS, TWIN64
    from tests.test_adjoint import adjoint as twin_adjoint

    if not os.path.exists(TWIN64):
        pytest.skip("oracle twin not built (make -C oracle/twin)")
    T64 = api.Api(C.CDLL(TWIN64))
    gpu = product_api()
    seeds = list(range(64))
    wg = [scenes.toss(gpu, s) for s in seeds]
    n = len(seeds)
    defs = (api.AdjointDef * n)()
    res = (api.AdjointResult * n)()
    keep = []
    for k in range(n):
        d = gpu.m3DefaultAdjointDef()
        d.steps = 120
        d.checkpointEvery = 10
        d.target = api.Vec3(1.0, 0.5, -2.0)
        d.precision = 64
        defs[k] = d
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
    ids = (api.WorldId * n)(*wg)
    assert gpu.m3World_Adjoint(ids, n, defs, res) == n
    rng = random.Random(0xF64)
    stable = 0
    for k, s in enumerate(seeds):
        w64 = scenes.toss(T64, s)
        r64, g64 = twin_adjoint(T64, w64, 120, every=10)
        assert res[k].unsupported == r64.unsupported == 0
        assert res[k].loss == pytest.approx(r64.loss, rel=1e-12, abs=1e-12), s
        for a, b in zip(flat(keep[k]), flat(g64)):
            assert abs(a - b) <= 2e-6 * max(1.0, abs(b)), (s, a, b)
        for _ in range(2):
            kind = rng.choice([1, 2, 3, 4, 5])
            body = rng.randrange(1, 33)
            comp = rng.randrange(3) if kind in (1, 2, 5) else 0
            eps = EPS[kind]
            L = lambda e: twin_adjoint(T64, w64, 120, (kind, body, comp, e), True, every=10)[0].loss  # noqa: E731
            fd = (L(eps) - L(-eps)) / (2 * eps)
            fd4 = (L(4 * eps) - L(-4 * eps)) / (8 * eps)
            scale = max(abs(fd), 1e-6)
            if abs(fd - fd4) / scale > 1e-5:
                continue
            stable += 1
            name = KINDS[kind]
            a = keep[k][name][body]
            dev = (a.x, a.y, a.z)[comp] if kind in (1, 2, 5) else a
            assert abs(dev - fd) / scale < 1e-3, (s, name, body, comp, dev, fd)
        T64.m3DestroyWorld(w64)
    assert stable >= 6, f"only {stable} contact-stable samples"
    for w in wg:
        gpu.m3DestroyWorld(w)


def test_precision_64_is_refused_without_faulting_the_world():
    """While the double-precision device pass is switched off, a request
    for it is refused (-1) and the world keeps stepping."""
    gpu = product_api()
    w = scenes.toss(gpu, 3)
    d = gpu.m3DefaultAdjointDef()
    d.steps = 10
    d.precision = 64
    r = api.AdjointResult()
    assert gpu.m3World_Adjoint((api.WorldId * 1)(w), 1, C.byref(d), C.byref(r)) == -1
    assert gpu.m3World_GetDeviceError(w) is None
    h = gpu.m3World_Hash(w)
    gpu.m3World_Step(w, DT, SUBSTEPS)
    assert gpu.m3World_Hash(w) != h
    gpu.m3DestroyWorld(w)
