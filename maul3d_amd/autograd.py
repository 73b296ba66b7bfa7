"""torch.autograd wrapper of m3World_Adjoint (the gradient harness's entry).

``trajectory_loss(api, world, target, linear_velocity, angular_velocity,
force, mass, friction)`` returns the loss |COM_T - target|^2 of ``steps``
steps from the world's current poses with the given initial velocities,
masses and frictions (written into the world's body slots) and a force
applied to each body every step; its backward hands out the gradients the
adjoint kernels computed in the same call.  Tensors are (bodyCapacity, 3) or
(bodyCapacity,) over body slots, float32, on the CPU.

A dynamic body's mass is written by scaling the densities of all its shapes
(``m3Shape_SetDensity``, the last one rebuilding the body's mass
properties), so its shapes keep their proportions: the change the adjoint
differentiates (inverse mass and inverse inertia scaling as 1/m).  Its
friction is written to every one of its shapes.  Entries of static bodies
and empty slots are ignored.  ``world_parameters`` reads the current values.
The world keeps the written velocities, masses and frictions; its poses are
left as they were.
"""

from __future__ import annotations

import ctypes as C

from .api import AdjointResult, BodyId, ShapeId, Vec3, WorldId


def _adjoint(api, world: WorldId, nb: int, target, force, steps: int, every: int, dt: float, substeps: int):
    d = api.m3DefaultAdjointDef()
    d.timeStep = dt
    d.subStepCount = substeps
    d.steps = steps
    d.checkpointEvery = every
    d.target = Vec3(*[float(x) for x in target])
    farr = (Vec3 * nb)(*[Vec3(*map(float, row)) for row in force])
    d.force = C.cast(farr, C.POINTER(Vec3))
    r = AdjointResult()
    gl = (Vec3 * nb)()
    ga = (Vec3 * nb)()
    gf = (Vec3 * nb)()
    gm = (C.c_float * nb)()
    gu = (C.c_float * nb)()
    r.gradLinearVelocity = C.cast(gl, C.POINTER(Vec3))
    r.gradAngularVelocity = C.cast(ga, C.POINTER(Vec3))
    r.gradForce = C.cast(gf, C.POINTER(Vec3))
    r.gradMass = C.cast(gm, C.POINTER(C.c_float))
    r.gradFriction = C.cast(gu, C.POINTER(C.c_float))
    ids = (WorldId * 1)(world)
    if api.m3World_Adjoint(ids, 1, C.byref(d), C.byref(r)) != 1:
        raise RuntimeError("m3World_Adjoint refused the world")
    vec = lambda a: [[a[b].x, a[b].y, a[b].z] for b in range(nb)]  # noqa: E731
    return r.loss, vec(gl), vec(ga), vec(gf), list(gm), list(gu), r.unsupported, r.ccdPulled


def _body_shapes(api, world: WorldId, nb: int) -> dict[int, list]:
    """Body slot -> its live shapes' ids, in shape-slot order."""
    out: dict[int, list] = {}
    for s in range(api.m3World_GetCounters(world).shapeCount):
        sid = ShapeId(s + 1, world.index1 - 1, 0)
        if not api.m3Shape_IsValid(sid):
            continue
        b = api.m3Shape_GetBody(sid).index1 - 1
        if 0 <= b < nb:
            out.setdefault(b, []).append(sid)
    return out


def world_parameters(api, world: WorldId, nb: int):
    """(mass, friction) tensors of shape (nb,) over body slots: each dynamic
    body's mass and its first shape's friction; 0 elsewhere."""
    import torch

    mass = [0.0] * nb
    fric = [0.0] * nb
    for b, sids in _body_shapes(api, world, nb).items():
        m = api.m3Body_GetMass(BodyId(b + 1, world.index1 - 1, 0))
        if m > 0.0:
            mass[b] = m
            fric[b] = api.m3Shape_GetFriction(sids[0])
    return torch.tensor(mass), torch.tensor(fric)


def _write_parameters(api, world: WorldId, nb: int, mass, friction) -> None:
    for b, sids in _body_shapes(api, world, nb).items():
        body = BodyId(b + 1, world.index1 - 1, 0)
        now = api.m3Body_GetMass(body)
        if not now > 0.0:
            continue
        want = float(mass[b])
        if want > 0.0 and want != now:
            scale = want / now
            for k, sid in enumerate(sids):
                api.m3Shape_SetDensity(sid, api.m3Shape_GetDensity(sid) * scale, k == len(sids) - 1)
        mu = float(friction[b])
        for sid in sids:
            if api.m3Shape_GetFriction(sid) != C.c_float(mu).value:
                api.m3Shape_SetFriction(sid, mu)


def trajectory_loss(api, world: WorldId, target, linear_velocity, angular_velocity, force, mass, friction,
                    steps: int = 120, checkpoint_every: int = 10, dt: float = 1.0 / 60.0, substeps: int = 4):
    import torch

    class _Trajectory(torch.autograd.Function):
        @staticmethod
        def forward(ctx, lin, ang, frc, m, mu):
            nb = lin.shape[0]
            _write_parameters(api, world, nb, m.tolist(), mu.tolist())
            for b in range(nb):
                body = BodyId(b + 1, world.index1 - 1, 0)
                if api.m3Body_IsValid(body):
                    api.m3Body_SetLinearVelocity(body, Vec3(*map(float, lin[b].tolist())))
                    api.m3Body_SetAngularVelocity(body, Vec3(*map(float, ang[b].tolist())))
            loss, gl, ga, gf, gm, gu, unsupported, ccd = _adjoint(
                api, world, nb, target, frc.tolist(), steps, checkpoint_every, dt, substeps
            )
            ctx.unsupported = unsupported
            ctx.save_for_backward(
                torch.tensor(gl), torch.tensor(ga), torch.tensor(gf), torch.tensor(gm), torch.tensor(gu)
            )
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            gl, ga, gf, gm, gu = ctx.saved_tensors
            s = gout.to(torch.float32)
            return s * gl, s * ga, s * gf, s * gm, s * gu

    return _Trajectory.apply(linear_velocity, angular_velocity, force, mass, friction)
