"""Seeded scene builders over the public ``m3*`` API.

Every builder takes an :class:`maul3d_amd.api.Api` and returns the world id,
so one scene definition feeds the MI355X library and the CPU checker with
bit-identical inputs.  Randomness is SplitMix64 with the 53-bit uniform draw
(restated from bench/bench_main.c:43-55); every float handed to the API is the
round-to-nearest cast of that double, as the C harness does.
"""

from __future__ import annotations

import ctypes as C
import math

from .api import (
    DYNAMIC,
    KINEMATIC,
    STATIC,
    Api,
    BodyDef,
    Capsule,
    Plane,
    Pos3,
    Quat,
    ShapeDef,
    Sphere,
    Vec3,
    WorldId,
)

_MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream and its [lo, hi) double draw (bench/bench_main.c:43-55)."""

    def __init__(self, seed: int):
        self.state = seed & _MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * (float(self.next() >> 11) / 9007199254740992.0)

    def unit_quat(self) -> tuple[float, float, float, float]:
        """Uniform rotation by rejection from the 4-cube, normalized in double."""
        while True:
            x = self.range(-1.0, 1.0)
            y = self.range(-1.0, 1.0)
            z = self.range(-1.0, 1.0)
            w = self.range(-1.0, 1.0)
            s = x * x + y * y + z * z + w * w
            if 1.0e-4 < s <= 1.0:
                inv = 1.0 / math.sqrt(s)
                return (x * inv, y * inv, z * inv, w * inv)


def _world(api: Api, bodies: int, shapes: int, joints: int | None = None) -> WorldId:
    d = api.m3DefaultWorldDef()
    d.bodyCapacity = bodies
    d.shapeCapacity = shapes
    if joints is not None:
        d.jointCapacity = joints
    hooks = getattr(api, "task_hooks", None)
    if hooks is not None:
        d.enqueueTask, d.finishTask, d.userTaskContext = hooks
    w = api.m3CreateWorld(C.byref(d))
    if w.index1 == 0:
        raise RuntimeError("m3CreateWorld refused the def")
    return w


def _ground(api: Api, world: WorldId, friction: float = 0.6) -> None:
    gd = api.m3DefaultBodyDef()
    ground = api.m3CreateBody(world, C.byref(gd))
    sd = api.m3DefaultShapeDef()
    sd.friction = friction
    plane = Plane(Vec3(0.0, 1.0, 0.0), 0.0)
    if api.m3CreatePlaneShape(ground, C.byref(sd), C.byref(plane)).index1 == 0:
        raise RuntimeError("plane refused")


def _check(shape_id, what: str) -> None:
    if shape_id.index1 == 0:
        raise RuntimeError(f"{what} refused")


def bench_pyramid(api: Api) -> WorldId:
    """The reference bench's sphere pyramid (bench/bench_main.c:68-95).

    Its hash after 300 steps of (1/60, 4) is pinned at 975f319f2f44cc5d in
    bench/pins.txt:1.
    """
    world = _world(api, 256, 256)
    _ground(api, world)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    sd = api.m3DefaultShapeDef()
    sd.friction = 0.6
    for layer in range(6):
        n = 6 - layer
        for i in range(n):
            for j in range(n):
                bd.position = Pos3(
                    float(i) - 0.5 * float(n) + 0.05 * float(layer),
                    0.5 + 1.02 * float(layer),
                    float(j) - 0.5 * float(n),
                )
                body = api.m3CreateBody(world, C.byref(bd))
                ball = Sphere(Vec3(0.0, 0.0, 0.0), 0.5)
                _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
    return world


def hulljam(api: Api) -> WorldId:
    """The reference bench's hull jam (bench/bench_main.c:97-135): 80 dynamic
    bodies dropped in a column over a plane, every third a 12-point QuickHull
    rock (a 0.25 m box where the hull is refused), the rest 0.3 x 0.25 x 0.35
    boxes, friction 0.5.

    Its hash after 300 steps of (1/60, 4) is pinned at 7bb4425d3adc3a51 in
    bench/pins.txt:2.
    """
    world = _world(api, 256, 256)
    _ground(api, world)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    sd = api.m3DefaultShapeDef()
    sd.friction = 0.5
    rng = SplitMix64(0xBEEF)
    for k in range(80):
        x = rng.range(-2.0, 2.0)
        z = rng.range(-2.0, 2.0)
        bd.position = Pos3(x, 0.6 + 0.5 * float(k), z)
        body = api.m3CreateBody(world, C.byref(bd))
        if k % 3 == 0:
            cloud = (Vec3 * 12)(
                *[Vec3(rng.range(-0.3, 0.3), rng.range(-0.3, 0.3), rng.range(-0.3, 0.3)) for _ in range(12)]
            )
            if api.m3CreateHullShape(body, C.byref(sd), cloud, 12).index1 == 0:
                _check(api.m3CreateBoxShape(body, C.byref(sd), Vec3(0.25, 0.25, 0.25)), "box")
        else:
            _check(api.m3CreateBoxShape(body, C.byref(sd), Vec3(0.3, 0.25, 0.35)), "box")
    return world


def bounce(api: Api, count: int = 160, seed: int = 0xB0A7CE, restitution: bool = True) -> WorldId:
    """Restitution exercise: ``count`` spheres, boxes and capsules with
    restitution U[0, 0.95] (friction U[0.2, 0.8]) dropped from 2-14 m onto a
    plane of restitution 0.5 inside a 6 m square, so they strike the plane
    and each other above the 1 m/s restitution threshold.  Every contact's
    restitution is max(a, b) (src/solver.c:156-404) and the serial
    restitution pass (src/solver.c:2879-2907) runs every step.  With
    ``restitution`` false every coefficient is 0 (same draws), the control
    that shows the pass changes the trajectory."""
    world = _world(api, count + 1, count + 1)
    gd = api.m3DefaultBodyDef()
    ground = api.m3CreateBody(world, C.byref(gd))
    sd = api.m3DefaultShapeDef()
    sd.friction = 0.6
    sd.restitution = 0.5 if restitution else 0.0
    _check(api.m3CreatePlaneShape(ground, C.byref(sd), C.byref(Plane(Vec3(0.0, 1.0, 0.0), 0.0))), "plane")
    rng = SplitMix64(seed)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    for k in range(count):
        bd.position = Pos3(rng.range(-3.0, 3.0), rng.range(2.0, 14.0), rng.range(-3.0, 3.0))
        q = rng.unit_quat()
        bd.rotation = Quat(*q)
        bd.linearVelocity = Vec3(rng.range(-2.0, 2.0), rng.range(-4.0, 0.0), rng.range(-2.0, 2.0))
        body = api.m3CreateBody(world, C.byref(bd))
        e = rng.range(0.0, 0.95)
        sd.restitution = e if restitution else 0.0
        sd.friction = rng.range(0.2, 0.8)
        kind = k % 3
        if kind == 0:
            ball = Sphere(Vec3(0.0, 0.0, 0.0), rng.range(0.2, 0.5))
            _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
        elif kind == 1:
            he = Vec3(rng.range(0.2, 0.5), rng.range(0.2, 0.5), rng.range(0.2, 0.5))
            _check(api.m3CreateBoxShape(body, C.byref(sd), he), "box")
        else:
            r = rng.range(0.15, 0.35)
            hh = rng.range(0.2, 0.5)
            cap = Capsule(Vec3(-hh, 0.0, 0.0), Vec3(hh, 0.0, 0.0), r)
            _check(api.m3CreateCapsuleShape(body, C.byref(sd), C.byref(cap)), "capsule")
    return world


def ccd_arena(api: Api, seed: int = 0xCCD5, movers: int = 60, bullets: int = 3) -> WorldId:
    """Continuous collision against tree shapes (src/solver.c:3081-3430): a
    static box floor and four static box walls (no plane), static 12-point
    hull rocks and static capsules inside, ``movers`` spheres, boxes and
    capsules thrown at 15-45 m/s, and ``bullets`` small bullet spheres
    (include/maul3d/body.h:32-37) at 60-90 m/s, which also sweep against the
    dynamic bodies.  Fast bodies are pulled back to their first time of
    impact with the static geometry each step."""
    world = _world(api, movers + bullets + 32, movers + bullets + 32)
    rng = SplitMix64(seed)
    sd = api.m3DefaultShapeDef()
    sd.friction = 0.6
    gd = api.m3DefaultBodyDef()
    gd.position = Pos3(0.0, -0.5, 0.0)
    floor = api.m3CreateBody(world, C.byref(gd))
    _check(api.m3CreateBoxShape(floor, C.byref(sd), Vec3(16.0, 0.5, 16.0)), "floor")
    for x, z, hx, hz in ((15.5, 0.0, 0.5, 16.0), (-15.5, 0.0, 0.5, 16.0), (0.0, 15.5, 16.0, 0.5), (0.0, -15.5, 16.0, 0.5)):
        gd.position = Pos3(x, 3.0, z)
        wall = api.m3CreateBody(world, C.byref(gd))
        _check(api.m3CreateBoxShape(wall, C.byref(sd), Vec3(hx, 3.0, hz)), "wall")
    for k in range(8):
        gd.position = Pos3(rng.range(-10.0, 10.0), rng.range(0.5, 2.0), rng.range(-10.0, 10.0))
        rock = api.m3CreateBody(world, C.byref(gd))
        if k % 2 == 0:
            cloud = (Vec3 * 12)(
                *[Vec3(rng.range(-1.0, 1.0), rng.range(-1.0, 1.0), rng.range(-1.0, 1.0)) for _ in range(12)]
            )
            if api.m3CreateHullShape(rock, C.byref(sd), cloud, 12).index1 == 0:
                _check(api.m3CreateBoxShape(rock, C.byref(sd), Vec3(0.8, 0.8, 0.8)), "rock")
        else:
            cap = Capsule(Vec3(-1.2, 0.0, 0.0), Vec3(1.2, 0.0, 0.0), rng.range(0.3, 0.6))
            _check(api.m3CreateCapsuleShape(rock, C.byref(sd), C.byref(cap)), "post")
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    for k in range(movers + bullets):
        bullet = k >= movers
        bd.position = Pos3(rng.range(-12.0, 12.0), rng.range(2.0, 10.0), rng.range(-12.0, 12.0))
        bd.rotation = Quat(*rng.unit_quat())
        speed = rng.range(60.0, 90.0) if bullet else rng.range(15.0, 45.0)
        d = (rng.range(-1.0, 1.0), rng.range(-1.0, -0.2), rng.range(-1.0, 1.0))
        n = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        bd.linearVelocity = Vec3(speed * d[0] / n, speed * d[1] / n, speed * d[2] / n)
        bd.angularVelocity = Vec3(rng.range(-5.0, 5.0), rng.range(-5.0, 5.0), rng.range(-5.0, 5.0))
        bd.isBullet = bullet
        body = api.m3CreateBody(world, C.byref(bd))
        kind = 0 if bullet else k % 3
        if kind == 0:
            ball = Sphere(Vec3(0.0, 0.0, 0.0), 0.1 if bullet else rng.range(0.2, 0.4))
            _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
        elif kind == 1:
            he = Vec3(rng.range(0.15, 0.4), rng.range(0.15, 0.4), rng.range(0.15, 0.4))
            _check(api.m3CreateBoxShape(body, C.byref(sd), he), "box")
        else:
            hh = rng.range(0.2, 0.4)
            cap = Capsule(Vec3(-hh, 0.0, 0.0), Vec3(hh, 0.0, 0.0), rng.range(0.12, 0.25))
            _check(api.m3CreateCapsuleShape(body, C.byref(sd), C.byref(cap)), "capsule")
    bd.isBullet = False
    return world


def platforms(api: Api, seed: int = 0x0DE7A, movers: int = 80) -> tuple[WorldId, list[int]]:
    """One-way platforms for the pre-solve veto (src/solver.c:214-248): a
    plane, five static box platforms (opted in with enablePreSolveEvents) at
    2-7 m, and ``movers`` spheres and boxes launched upward from under them at
    8-14 m/s, which pass up through a platform when its contact is vetoed and
    land on it on the way down.  Returns the world and the platforms' shape
    slots (the callback's only state)."""
    world = _world(api, movers + 8, movers + 8)
    _ground(api, world)
    rng = SplitMix64(seed)
    sd = api.m3DefaultShapeDef()
    sd.friction = 0.6
    gd = api.m3DefaultBodyDef()
    slots = []
    for k in range(5):
        gd.position = Pos3(-8.0 + 4.0 * k, 2.0 + 1.2 * k, 0.0)
        plat = api.m3CreateBody(world, C.byref(gd))
        sd.enablePreSolveEvents = True
        sid = api.m3CreateBoxShape(plat, C.byref(sd), Vec3(1.8, 0.15, 3.0))
        _check(sid, "platform")
        slots.append(sid.index1 - 1)
    sd.enablePreSolveEvents = False
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    for k in range(movers):
        lane = k % 5
        bd.position = Pos3(-8.0 + 4.0 * lane + rng.range(-1.5, 1.5), rng.range(0.4, 1.2), rng.range(-2.5, 2.5))
        bd.linearVelocity = Vec3(rng.range(-0.5, 0.5), rng.range(8.0, 14.0), rng.range(-0.5, 0.5))
        body = api.m3CreateBody(world, C.byref(bd))
        if k % 2 == 0:
            _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(Sphere(Vec3(0, 0, 0), rng.range(0.15, 0.3)))), "sphere")
        else:
            he = Vec3(rng.range(0.12, 0.25), rng.range(0.12, 0.25), rng.range(0.12, 0.25))
            _check(api.m3CreateBoxShape(body, C.byref(sd), he), "box")
    bd.linearVelocity = Vec3(0.0, 0.0, 0.0)
    return world, slots


def one_way_rule(platform_slots: list[int]):
    """The pure pre-solve callback of ``platforms``: a contact with a
    platform holds only when the other shape is above it (the normal, from A
    to B, points away from the platform's top)."""
    plats = set(platform_slots)

    def rule(a: int, b: int, normal_y: float) -> bool:
        if a in plats:
            return normal_y > 0.5
        if b in plats:
            return normal_y < -0.5
        return True

    return rule


def smoke10k(api: Api) -> WorldId:
    """The reference bench's 10k mixed rain (bench/bench_main.c:612-650).

    Hash after 150 steps pinned at bc84e0074315a0b3 (bench/pins.txt:7).
    """
    world = _world(api, 16384, 16384)
    _ground(api, world)
    sd = api.m3DefaultShapeDef()
    sd.friction = 0.6
    sd.rollingResistance = 0.05
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    rng = SplitMix64(0x10A11)
    for k in range(10000):
        x = rng.range(-46.0, 46.0)
        y = rng.range(8.0, 60.0)
        z = rng.range(-46.0, 46.0)
        bd.position = Pos3(x, y, z)
        body = api.m3CreateBody(world, C.byref(bd))
        if k % 3 == 0:
            ball = Sphere(Vec3(0.0, 0.0, 0.0), rng.range(0.2, 0.4))
            _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
        elif k % 3 == 1:
            hx = rng.range(0.15, 0.35)
            hy = rng.range(0.15, 0.35)
            hz = rng.range(0.15, 0.35)
            _check(api.m3CreateBoxShape(body, C.byref(sd), Vec3(hx, hy, hz)), "box")
        else:
            cap = Capsule(Vec3(-0.2, 0.0, 0.0), Vec3(0.2, 0.0, 0.0), rng.range(0.12, 0.2))
            _check(api.m3CreateCapsuleShape(body, C.byref(sd), C.byref(cap)), "capsule")
    return world


def box_pyramid(api: Api, rows: int = 20) -> WorldId:
    """Classic pyramid: ``rows`` rows of unit cubes on a static plane.

    Row r holds rows-r boxes of half-extent 0.5 (density 1000, friction 0.6,
    restitution 0) centred at x = i - (rows-1-r)/2, y = 0.5 + r, z = 0.
    20 rows = 210 dynamic boxes.
    """
    n_boxes = rows * (rows + 1) // 2
    cap = max(256, n_boxes + 8)
    world = _world(api, cap, cap)
    _ground(api, world)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    sd.restitution = 0.0
    for r in range(rows):
        count = rows - r
        for i in range(count):
            bd.position = Pos3(float(i) - 0.5 * float(rows - 1 - r), 0.5 + float(r), 0.0)
            body = api.m3CreateBody(world, C.byref(bd))
            _check(api.m3CreateBoxShape(body, C.byref(sd), Vec3(0.5, 0.5, 0.5)), "box")
    return world


def rain(api: Api, count: int = 10000, seed: int = 0x5EED10C, extent: float = 50.0) -> WorldId:
    """Mixed rain: ``count`` dynamic bodies dropped onto a plane.

    Body k is a box when k % 5 < 2 (half-extents U[0.25,1] each), a sphere when
    k % 5 is 2 or 3 (radius U[0.25,0.75]) and a capsule when k % 5 == 4
    (radius U[0.2,0.5], half-height U[0.25,1] along local x).  Positions are
    uniform in [-extent,extent] x [2,52] x [-extent,extent], rotations uniform,
    velocities zero, density 1000, friction 0.6.  Draw order per body:
    position x, y, z, rotation, then the shape's sizes.
    """
    cap = count + 64
    world = _world(api, cap, cap)
    _ground(api, world)
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    rng = SplitMix64(seed)
    for k in range(count):
        x = rng.range(-extent, extent)
        y = rng.range(2.0, 52.0)
        z = rng.range(-extent, extent)
        qx, qy, qz, qw = rng.unit_quat()
        bd.position = Pos3(x, y, z)
        bd.rotation = Quat(qx, qy, qz, qw)
        body = api.m3CreateBody(world, C.byref(bd))
        kind = k % 5
        if kind < 2:
            hx = rng.range(0.25, 1.0)
            hy = rng.range(0.25, 1.0)
            hz = rng.range(0.25, 1.0)
            _check(api.m3CreateBoxShape(body, C.byref(sd), Vec3(hx, hy, hz)), "box")
        elif kind < 4:
            ball = Sphere(Vec3(0.0, 0.0, 0.0), rng.range(0.25, 0.75))
            _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
        else:
            r = rng.range(0.2, 0.5)
            hh = rng.range(0.25, 1.0)
            cap_ = Capsule(Vec3(-hh, 0.0, 0.0), Vec3(hh, 0.0, 0.0), r)
            _check(api.m3CreateCapsuleShape(body, C.byref(sd), C.byref(cap_)), "capsule")
    return world


def toss(api: Api, seed: int, boxes: int = 32) -> WorldId:
    """One differentiable-toss world: ``boxes`` boxes launched onto a plane.

    Half-extents U[0.2,0.6], mass U[0.5,5] kg (density = m / (8 hx hy hz)),
    linear velocity U[-5,5]^3, angular velocity U[-3,3]^3, friction U[0.3,0.9],
    positions on a 4-wide grid 1.5 m apart at heights U[1,3].
    """
    world = _world(api, boxes + 8, boxes + 8)
    _ground(api, world)
    rng = SplitMix64(seed)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    for k in range(boxes):
        hx = rng.range(0.2, 0.6)
        hy = rng.range(0.2, 0.6)
        hz = rng.range(0.2, 0.6)
        mass = rng.range(0.5, 5.0)
        fx = C.c_float(hx).value
        fy = C.c_float(hy).value
        fz = C.c_float(hz).value
        sd = api.m3DefaultShapeDef()
        sd.density = mass / (8.0 * fx * fy * fz)
        sd.friction = rng.range(0.3, 0.9)
        bd.position = Pos3(1.5 * float(k % 4) - 2.25, rng.range(1.0, 3.0), 1.5 * float(k // 4) - 5.25)
        bd.linearVelocity = Vec3(rng.range(-5, 5), rng.range(-5, 5), rng.range(-5, 5))
        bd.angularVelocity = Vec3(rng.range(-3, 3), rng.range(-3, 3), rng.range(-3, 3))
        body = api.m3CreateBody(world, C.byref(bd))
        _check(api.m3CreateBoxShape(body, C.byref(sd), Vec3(hx, hy, hz)), "box")
    return world


def batch_world(api: Api, index: int, seed: int = 0xB47C4) -> WorldId:
    """World ``index`` of the batched-rollback set (uneven sizes and kinds).

    Each world draws from its own stream SplitMix64(seed + 0x9E3779B9 * index):
    a body budget N = 50 + floor(U[0,1951)) and a kind floor(U[0,3)):

    * 0, pyramid: the largest r-row pyramid of unit cubes (half-extent 0.5,
      centres x = i - (r-1-row)/2, y = 0.5 + row, z = 0) with r(r+1)/2 <= N;
    * 1, box stacks: height h = 10 + floor(U[0,31)), N // h columns of unit
      cubes on a square grid 1.5 m apart, y = 0.5 + level;
    * 2, rain: N bodies, even k a sphere (radius U[0.25,0.5]), odd k a box
      (half-extents U[0.25,0.5] each), positions U[-e,e] x U[1,11] x U[-e,e]
      with e = max(3, 0.6 sqrt(N)), rotations uniform.

    All dynamic, density 1000, friction 0.6, on a plane; capacities N + 8.
    """
    rng = SplitMix64(seed + 0x9E3779B9 * index)
    n = 50 + int(rng.range(0.0, 1951.0))
    kind = int(rng.range(0.0, 3.0))
    world = _world(api, n + 8, n + 8)
    _ground(api, world)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    unit = Vec3(0.5, 0.5, 0.5)
    if kind == 0:
        rows = int((math.sqrt(8.0 * n + 1.0) - 1.0) / 2.0)
        while rows * (rows + 1) // 2 > n:
            rows -= 1
        for r in range(rows):
            for i in range(rows - r):
                bd.position = Pos3(float(i) - 0.5 * float(rows - 1 - r), 0.5 + float(r), 0.0)
                body = api.m3CreateBody(world, C.byref(bd))
                _check(api.m3CreateBoxShape(body, C.byref(sd), unit), "box")
    elif kind == 1:
        h = 10 + int(rng.range(0.0, 31.0))
        columns = max(1, n // h)
        side = int(math.ceil(math.sqrt(columns)))
        for c in range(columns):
            x = 1.5 * float(c % side) - 0.75 * float(side - 1)
            z = 1.5 * float(c // side) - 0.75 * float(side - 1)
            for level in range(h):
                bd.position = Pos3(x, 0.5 + float(level), z)
                body = api.m3CreateBody(world, C.byref(bd))
                _check(api.m3CreateBoxShape(body, C.byref(sd), unit), "box")
    else:
        e = max(3.0, 0.6 * math.sqrt(float(n)))
        for k in range(n):
            x = rng.range(-e, e)
            y = rng.range(1.0, 11.0)
            z = rng.range(-e, e)
            qx, qy, qz, qw = rng.unit_quat()
            bd.position = Pos3(x, y, z)
            bd.rotation = Quat(qx, qy, qz, qw)
            body = api.m3CreateBody(world, C.byref(bd))
            if k % 2 == 0:
                ball = Sphere(Vec3(0.0, 0.0, 0.0), rng.range(0.25, 0.5))
                _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
            else:
                half = Vec3(rng.range(0.25, 0.5), rng.range(0.25, 0.5), rng.range(0.25, 0.5))
                _check(api.m3CreateBoxShape(body, C.byref(sd), half), "box")
        bd.rotation = Quat(0.0, 0.0, 0.0, 1.0)
    return world


def _add_chains(api: Api, world: WorldId, count: int, links: int, height: float) -> None:
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    capsule = Capsule(Vec3(-0.25, 0.0, 0.0), Vec3(0.25, 0.0, 0.0), 0.1)
    for c in range(count):
        x0 = 12.0 * float(c % 10) - 54.0
        z0 = 1.5 * float(c // 10) - 37.5
        ad = api.m3DefaultBodyDef()
        ad.type = STATIC
        ad.position = Pos3(x0, height, z0)
        anchor = api.m3CreateBody(world, C.byref(ad))
        bodies = []
        bd = api.m3DefaultBodyDef()
        bd.type = DYNAMIC
        for k in range(links):
            bd.position = Pos3(x0 + 0.5 * float(k) + 0.25, height, z0)
            body = api.m3CreateBody(world, C.byref(bd))
            _check(api.m3CreateCapsuleShape(body, C.byref(sd), C.byref(capsule)), "capsule")
            bodies.append(body)
        jd = api.m3DefaultJointDef()
        jd.enableCone = True
        jd.coneAngle = 0.7853982
        jd.localAxisA = Vec3(1.0, 0.0, 0.0)
        jd.localAxisB = Vec3(1.0, 0.0, 0.0)
        for k in range(links):
            jd.bodyA = anchor if k == 0 else bodies[k - 1]
            jd.bodyB = bodies[k]
            jd.localAnchorA = Vec3(0.0, 0.0, 0.0) if k == 0 else Vec3(0.25, 0.0, 0.0)
            jd.localAnchorB = Vec3(-0.25, 0.0, 0.0)
            if api.m3CreateJoint(C.byref(jd)).index1 == 0:
                raise RuntimeError("joint refused")


def chains(
    api: Api,
    count: int = 500,
    links: int = 20,
    boxes: int = 4000,
    seed: int = 0xC4A1,
    height: float = 14.0,
) -> WorldId:
    """Capsule chains on spherical joints with 45-degree cone limits, plus
    optional falling boxes.

    Chain c hangs from a static anchor body at (12 (c % 10) - 54, height,
    1.5 (c // 10) - 37.5).  Its ``links`` capsules (radius 0.1, segment from
    -0.25 to 0.25 along local x, density 1000, friction 0.6) start
    horizontal along +x, link k centred 0.5 k + 0.25 from the anchor.  Joint
    k pins the previous body (the anchor for k = 0) to link k at their shared
    point; both frame z-axes are local +x (the chain direction), the cone
    half-angle is 0.7853982 rad and jointed bodies do not collide.  Creation
    order: the plane, then per chain the anchor, its links, then its joints.
    ``boxes`` boxes (half-extents U[0.25,0.75]) then fall from
    U[-55,55] x U[20,40] x U[-40,40] with uniform rotations (stream
    SplitMix64(seed)).
    """
    n_bodies = count * (links + 1) + boxes + 1
    world = _world(api, n_bodies + 8, n_bodies + 8, joints=count * links + 8)
    _ground(api, world)
    _add_chains(api, world, count, links, height)
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    rng = SplitMix64(seed)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    for _ in range(boxes):
        x = rng.range(-55.0, 55.0)
        y = rng.range(20.0, 40.0)
        z = rng.range(-40.0, 40.0)
        qx, qy, qz, qw = rng.unit_quat()
        bd.position = Pos3(x, y, z)
        bd.rotation = Quat(qx, qy, qz, qw)
        body = api.m3CreateBody(world, C.byref(bd))
        half = Vec3(rng.range(0.25, 0.75), rng.range(0.25, 0.75), rng.range(0.25, 0.75))
        _check(api.m3CreateBoxShape(body, C.byref(sd), half), "box")
    return world


def ellipsoid_cloud(rng: SplitMix64) -> list[tuple[float, float, float]]:
    """k = 8 + floor(U[0,57)) points on an ellipsoid with semi-axes U[0.3,1.2]:
    directions by rejection from the unit cube (|d|^2 in (1e-4, 1]),
    normalized in double, scaled per axis."""
    k = 8 + int(rng.range(0.0, 57.0))
    a = rng.range(0.3, 1.2)
    b = rng.range(0.3, 1.2)
    c = rng.range(0.3, 1.2)
    pts = []
    for _ in range(k):
        while True:
            x = rng.range(-1.0, 1.0)
            y = rng.range(-1.0, 1.0)
            z = rng.range(-1.0, 1.0)
            r2 = x * x + y * y + z * z
            if 1.0e-4 < r2 <= 1.0:
                break
        inv = 1.0 / math.sqrt(r2)
        pts.append((a * x * inv, b * y * inv, c * z * inv))
    return pts


def hull_chains(
    api: Api,
    hulls: int = 4000,
    count: int = 500,
    links: int = 20,
    seed: int = 0x4C4,
) -> WorldId:
    """Hull rain plus joint chains: ``count`` chains exactly as in
    :func:`chains` (no boxes), then ``hulls`` convex hulls built by
    m3CreateHullShape from :func:`ellipsoid_cloud` point sets (stream
    SplitMix64(seed)), dropped with uniform rotations from
    U[-50,50] x U[16,66] x U[-50,50] (stream SplitMix64(seed ^ 0xD40F)),
    density 1000, friction 0.6.  Every cloud of the default seed is one the
    reference's QuickHull builds inside its fixed work pools.
    """
    n_bodies = count * (links + 1) + hulls + 1
    world = _world(api, n_bodies + 8, n_bodies + 8, joints=count * links + 8)
    _ground(api, world)
    _add_chains(api, world, count, links, 14.0)
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    clouds = SplitMix64(seed)
    place = SplitMix64(seed ^ 0xD40F)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    for _ in range(hulls):
        pts = ellipsoid_cloud(clouds)
        arr = (Vec3 * len(pts))(*[Vec3(*p) for p in pts])
        x = place.range(-50.0, 50.0)
        y = place.range(16.0, 66.0)
        z = place.range(-50.0, 50.0)
        qx, qy, qz, qw = place.unit_quat()
        bd.position = Pos3(x, y, z)
        bd.rotation = Quat(qx, qy, qz, qw)
        body = api.m3CreateBody(world, C.byref(bd))
        _check(api.m3CreateHullShape(body, C.byref(sd), arr, len(pts)), "hull")
    return world


def hull_pile(api: Api, hulls: int = 200, seed: int = 0x4C4) -> WorldId:
    """A dense hull column: the first ``hulls`` clouds of the hull-rain cloud
    stream SplitMix64(seed), hull i dropped at U[-4,4] x (2 + 1.2 i) x
    U[-4,4] with a uniform rotation (stream SplitMix64(seed ^ 0x9111)),
    density 1000, friction 0.6.  Hull-on-hull manifolds dominate."""
    world = _world(api, hulls + 8, hulls + 8)
    _ground(api, world)
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    clouds = SplitMix64(seed)
    place = SplitMix64(seed ^ 0x9111)
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    for i in range(hulls):
        pts = ellipsoid_cloud(clouds)
        arr = (Vec3 * len(pts))(*[Vec3(*p) for p in pts])
        x = place.range(-4.0, 4.0)
        z = place.range(-4.0, 4.0)
        qx, qy, qz, qw = place.unit_quat()
        bd.position = Pos3(x, 2.0 + 1.2 * i, z)
        bd.rotation = Quat(qx, qy, qz, qw)
        body = api.m3CreateBody(world, C.byref(bd))
        _check(api.m3CreateHullShape(body, C.byref(sd), arr, len(pts)), "hull")
    return world


def crowded_plank(api: Api, n: int = 7, seed: int = 0xC0FFEE) -> WorldId:
    """A dynamic 8 x 0.5 x 8 plank on a plane under two layers of n x n unit
    spheres dropped from jittered heights: from about step 30 the plank
    carries more constraints than there are graph colors, so the coloring's
    overflow color (solved serially, after every colored constraint) is in
    use for the rest of the first 150 steps.
    """
    world = _world(api, 256, 256)
    _ground(api, world)
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    bd.position = Pos3(0.0, 0.25, 0.0)
    plank = api.m3CreateBody(world, C.byref(bd))
    _check(api.m3CreateBoxShape(plank, C.byref(sd), Vec3(4.0, 0.25, 4.0)), "plank")
    rng = SplitMix64(seed)
    for i in range(n):
        for j in range(n):
            for layer in range(2):
                x = -3.0 + i * 1.0 + rng.range(-0.2, 0.2)
                y = 1.0 + layer * 1.1 + rng.range(0.0, 2.0)
                z = -3.0 + j * 1.0 + rng.range(-0.2, 0.2)
                bd.position = Pos3(x, y, z)
                body = api.m3CreateBody(world, C.byref(bd))
                ball = Sphere(Vec3(0.0, 0.0, 0.0), 0.5)
                _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
    return world


def event_rain(api: Api, count: int = 400, seed: int = 0xE7E47, extent: float = 8.0) -> WorldId:
    """A rain for the event streams: ``count`` bodies as in ``rain`` (boxes,
    spheres and capsules by index, positions in [-extent,extent] x [2,30] x
    [-extent,extent]) falling through four static sensor boxes (half-extents
    2.5 x 0.75 x 2.5 centred at y = 9) onto a plane.  Every third body's
    shape opts into hit events."""
    cap = count + 64
    world = _world(api, cap, cap)
    _ground(api, world)
    sens = api.m3DefaultShapeDef()
    sens.isSensor = True
    gd = api.m3DefaultBodyDef()
    for cx, cz in ((-4.0, -4.0), (4.0, -4.0), (-4.0, 4.0), (4.0, 4.0)):
        gd.position = Pos3(cx, 9.0, cz)
        holder = api.m3CreateBody(world, C.byref(gd))
        _check(api.m3CreateBoxShape(holder, C.byref(sens), Vec3(2.5, 0.75, 2.5)), "sensor box")
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    rng = SplitMix64(seed)
    for k in range(count):
        x = rng.range(-extent, extent)
        y = rng.range(2.0, 30.0)
        z = rng.range(-extent, extent)
        qx, qy, qz, qw = rng.unit_quat()
        bd.position = Pos3(x, y, z)
        bd.rotation = Quat(qx, qy, qz, qw)
        body = api.m3CreateBody(world, C.byref(bd))
        sd.enableHitEvents = k % 3 == 0
        kind = k % 5
        if kind < 2:
            he = Vec3(rng.range(0.25, 1.0), rng.range(0.25, 1.0), rng.range(0.25, 1.0))
            _check(api.m3CreateBoxShape(body, C.byref(sd), he), "box")
        elif kind < 4:
            ball = Sphere(Vec3(0.0, 0.0, 0.0), rng.range(0.25, 0.75))
            _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
        else:
            r = rng.range(0.2, 0.5)
            hh = rng.range(0.25, 1.0)
            cap_ = Capsule(Vec3(-hh, 0.0, 0.0), Vec3(hh, 0.0, 0.0), r)
            _check(api.m3CreateCapsuleShape(body, C.byref(sd), C.byref(cap_)), "capsule")
    return world


def sweepers(api: Api, count: int = 1600, seed: int = 0x5EE9, extent: float = 12.0) -> WorldId:
    """Kinematic paddles ploughing a pile: ``count`` boxes and spheres
    dropped from 1-9 m onto a plane over [-extent, extent]^2, four
    kinematic slabs sliding along x at 3 m/s while they spin about y, and a
    kinematic post turning in place at the centre.  Every contact with a
    kinematic body reads position deltas that the body's own substeps move
    (src/solver.c:3938-4203 integrates kinematic movers with the rest)."""
    world = _world(api, count + 16, count + 16)
    _ground(api, world)
    sd = api.m3DefaultShapeDef()
    sd.density = 1000.0
    sd.friction = 0.6
    kd = api.m3DefaultBodyDef()
    kd.type = KINEMATIC
    for k in range(4):
        kd.position = Pos3(-extent + 1.0, 0.7, -9.0 + 6.0 * k)
        kd.linearVelocity = Vec3(3.0, 0.0, 0.0)
        kd.angularVelocity = Vec3(0.0, 0.8 if k % 2 == 0 else -0.8, 0.0)
        paddle = api.m3CreateBody(world, C.byref(kd))
        _check(api.m3CreateBoxShape(paddle, C.byref(sd), Vec3(2.5, 0.6, 0.3)), "paddle")
    kd.position = Pos3(0.0, 1.5, 0.0)
    kd.linearVelocity = Vec3(0.0, 0.0, 0.0)
    kd.angularVelocity = Vec3(0.0, 2.0, 0.0)
    post = api.m3CreateBody(world, C.byref(kd))
    _check(api.m3CreateBoxShape(post, C.byref(sd), Vec3(3.0, 1.5, 0.4)), "post")
    bd = api.m3DefaultBodyDef()
    bd.type = DYNAMIC
    rng = SplitMix64(seed)
    for k in range(count):
        x = rng.range(-extent, extent)
        y = rng.range(1.0, 9.0)
        z = rng.range(-extent, extent)
        bd.position = Pos3(x, y, z)
        body = api.m3CreateBody(world, C.byref(bd))
        if k % 2 == 0:
            he = Vec3(rng.range(0.15, 0.4), rng.range(0.15, 0.4), rng.range(0.15, 0.4))
            _check(api.m3CreateBoxShape(body, C.byref(sd), he), "box")
        else:
            ball = Sphere(Vec3(0.0, 0.0, 0.0), rng.range(0.15, 0.4))
            _check(api.m3CreateSphereShape(body, C.byref(sd), C.byref(ball)), "sphere")
    return world


SCENES = {
    "bench_pyramid": bench_pyramid,
    "smoke10k": smoke10k,
    "box_pyramid": box_pyramid,
    "rain": rain,
    "chains": chains,
    "hull_chains": hull_chains,
    "hull_pile": hull_pile,
    "crowded_plank": crowded_plank,
    "event_rain": event_rain,
    "hulljam": hulljam,
    "bounce": bounce,
    "ccd_arena": ccd_arena,
    "sweepers": sweepers,
}
