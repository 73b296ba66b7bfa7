"""ctypes mirror of the Maul3D public C API.

The same bindings drive any shared library that exports the ``m3*`` surface:
this package's ``libmaul3d_amd.so`` (the MI355X build) and, in tests only,
the reference engine compiled as the parity checker.  Struct layouts follow
the reference public headers field for field:

* ``m3Vec3``/``m3Pos3``/``m3Quat``/``m3Transform``  include/maul3d/math.h:27-55
* ``m3WorldId``/``m3BodyId``/``m3ShapeId``          include/maul3d/base.h:69-104
* ``m3WorldDef``                                     include/maul3d/world.h:43-76
* ``m3BodyDef``                                      include/maul3d/body.h:21-39
* ``m3ShapeDef``                                     include/maul3d/shape.h:40-74
* ``m3Counters``/``m3Profile``                       include/maul3d/world.h:531-591
"""

from __future__ import annotations

import ctypes as C

c_f = C.c_float
c_d = C.c_double
c_i32 = C.c_int32
c_u16 = C.c_uint16
c_u64 = C.c_uint64


class Vec3(C.Structure):
    _fields_ = [("x", c_f), ("y", c_f), ("z", c_f)]


class Pos3(C.Structure):
    _fields_ = [("x", c_d), ("y", c_d), ("z", c_d)]


class Quat(C.Structure):
    _fields_ = [("x", c_f), ("y", c_f), ("z", c_f), ("w", c_f)]


class Transform(C.Structure):
    _fields_ = [("p", Pos3), ("q", Quat)]


class WorldId(C.Structure):
    _fields_ = [("index1", c_i32), ("generation", c_u16)]


class BodyId(C.Structure):
    _fields_ = [("index1", c_i32), ("world0", c_u16), ("generation", c_u16)]


class ShapeId(C.Structure):
    _fields_ = [("index1", c_i32), ("world0", c_u16), ("generation", c_u16)]


class JointId(C.Structure):
    _fields_ = [("index1", c_i32), ("world0", c_u16), ("generation", c_u16)]


class WorldDef(C.Structure):
    _fields_ = [
        ("gravity", Vec3),
        ("bodyCapacity", c_i32),
        ("shapeCapacity", c_i32),
        ("meshCapacity", c_i32),
        ("jointCapacity", c_i32),
        ("voxelCapacity", c_i32),
        ("characterCapacity", c_i32),
        ("vehicleCapacity", c_i32),
        ("softBodyCapacity", c_i32),
        ("workerCount", c_i32),
        ("enqueueTask", C.c_void_p),
        ("finishTask", C.c_void_p),
        ("userTaskContext", C.c_void_p),
        ("contactHertz", c_f),
        ("contactDampingRatio", c_f),
        ("contactPushMaxSpeed", c_f),
        ("restitutionThreshold", c_f),
        ("maximumLinearSpeed", c_f),
        ("enableSleeping", c_i32),
        ("enableContinuous", c_i32),
        ("hitEventThreshold", c_f),
        ("internalValue", c_i32),
    ]


class BodyDef(C.Structure):
    _fields_ = [
        ("type", c_i32),
        ("position", Pos3),
        ("rotation", Quat),
        ("linearVelocity", Vec3),
        ("angularVelocity", Vec3),
        ("gravityScale", c_f),
        ("linearDamping", c_f),
        ("angularDamping", c_f),
        ("userData", c_u64),
        ("isBullet", C.c_bool),
        ("internalValue", c_i32),
    ]


class ShapeDef(C.Structure):
    _fields_ = [
        ("density", c_f),
        ("friction", c_f),
        ("restitution", c_f),
        ("categoryBits", c_u64),
        ("maskBits", c_u64),
        ("groupIndex", c_i32),
        ("rollingResistance", c_f),
        ("userData", c_u64),
        ("isSensor", C.c_bool),
        ("enableHitEvents", C.c_bool),
        ("enablePreSolveEvents", C.c_bool),
        ("localPosition", Vec3),
        ("localRotation", Quat),
        ("internalValue", c_i32),
    ]


class JointDef(C.Structure):
    """m3JointDef field for field (include/maul3d/joint.h:80-163), 224 bytes."""

    _fields_ = [
        ("type", c_i32),
        ("bodyA", BodyId),
        ("bodyB", BodyId),
        ("localAnchorA", Vec3),
        ("localAnchorB", Vec3),
        ("localAxisA", Vec3),
        ("localAxisB", Vec3),
        ("enableLimit", C.c_bool),
        ("lowerLimit", c_f),
        ("upperLimit", c_f),
        ("enableMotor", C.c_bool),
        ("motorSpeed", c_f),
        ("maxMotorEffort", c_f),
        ("enableCone", C.c_bool),
        ("coneAngle", c_f),
        ("collideConnected", C.c_bool),
        ("genericLinear", C.c_uint8 * 3),
        ("genericAngular", C.c_uint8 * 3),
        ("genericMotorAxis", C.c_uint8),
        ("genericReserved", C.c_uint8),
        ("genericLinearLower", c_f * 3),
        ("genericLinearUpper", c_f * 3),
        ("genericAngularLower", c_f * 3),
        ("genericAngularUpper", c_f * 3),
        ("ratio", c_f),
        ("groundAnchorA", Pos3),
        ("groundAnchorB", Pos3),
        ("internalValue", c_i32),
    ]


class Sphere(C.Structure):
    _fields_ = [("center", Vec3), ("radius", c_f)]


class Plane(C.Structure):
    _fields_ = [("normal", Vec3), ("offset", c_f)]


class Capsule(C.Structure):
    _fields_ = [("point1", Vec3), ("point2", Vec3), ("radius", c_f)]


class Counters(C.Structure):
    _fields_ = [
        ("bodyCount", c_i32),
        ("shapeCount", c_i32),
        ("jointCount", c_i32),
        ("contactCount", c_i32),
        ("awakeCount", c_i32),
        ("islandCount", c_i32),
        ("colorCount", c_i32),
        ("characterCount", c_i32),
        ("vehicleCount", c_i32),
        ("softBodyCount", c_i32),
        ("voxelChunkCount", c_i32),
        ("hullCount", c_i32),
        ("meshCount", c_i32),
        ("treeHeight", c_i32),
        ("scratchPeak", c_i32),
        ("scratchCapacity", c_i32),
        ("snapshotBytes", c_i32),
        ("allocCalls", C.c_int64),
        ("freeCalls", C.c_int64),
    ]


class Profile(C.Structure):
    _fields_ = [
        (n, c_f)
        for n in (
            "step",
            "vehicles",
            "broadphase",
            "narrowphase",
            "events",
            "prepare",
            "solve",
            "continuous",
            "sleep",
            "characters",
            "softBodies",
        )
    ]


STATIC, KINEMATIC, DYNAMIC = 0, 1, 2

# name -> (restype, argtypes)
_SIGNATURES = {
    "m3DefaultWorldDef": (WorldDef, []),
    "m3CreateWorld": (WorldId, [C.POINTER(WorldDef)]),
    "m3DestroyWorld": (None, [WorldId]),
    "m3World_IsValid": (C.c_bool, [WorldId]),
    "m3World_Step": (None, [WorldId, c_f, c_i32]),
    "m3World_Hash": (c_u64, [WorldId]),
    "m3World_SnapshotSize": (c_i32, [WorldId]),
    "m3World_Snapshot": (c_i32, [WorldId, C.c_void_p, c_i32]),
    "m3World_Restore": (C.c_bool, [WorldId, C.c_void_p, c_i32]),
    "m3World_GetCounters": (Counters, [WorldId]),
    "m3World_GetProfile": (Profile, [WorldId]),
    "m3World_GetGravity": (Vec3, [WorldId]),
    "m3World_SetGravity": (None, [WorldId, Vec3]),
    "m3World_EnableSleeping": (None, [WorldId, C.c_bool]),
    "m3World_EnableContinuous": (None, [WorldId, C.c_bool]),
    "m3DefaultBodyDef": (BodyDef, []),
    "m3CreateBody": (BodyId, [WorldId, C.POINTER(BodyDef)]),
    "m3DestroyBody": (None, [BodyId]),
    "m3Body_IsValid": (C.c_bool, [BodyId]),
    "m3Body_GetPosition": (Pos3, [BodyId]),
    "m3Body_GetRotation": (Quat, [BodyId]),
    "m3Body_GetLinearVelocity": (Vec3, [BodyId]),
    "m3Body_GetAngularVelocity": (Vec3, [BodyId]),
    "m3Body_SetLinearVelocity": (None, [BodyId, Vec3]),
    "m3Body_SetAngularVelocity": (None, [BodyId, Vec3]),
    "m3Body_IsAwake": (C.c_bool, [BodyId]),
    "m3Body_SetAwake": (None, [BodyId, C.c_bool]),
    "m3Body_ApplyForce": (None, [BodyId, Vec3]),
    "m3Body_ApplyTorque": (None, [BodyId, Vec3]),
    "m3DefaultShapeDef": (ShapeDef, []),
    "m3CreateSphereShape": (ShapeId, [BodyId, C.POINTER(ShapeDef), C.POINTER(Sphere)]),
    "m3CreatePlaneShape": (ShapeId, [BodyId, C.POINTER(ShapeDef), C.POINTER(Plane)]),
    "m3CreateBoxShape": (ShapeId, [BodyId, C.POINTER(ShapeDef), Vec3]),
    "m3CreateCapsuleShape": (ShapeId, [BodyId, C.POINTER(ShapeDef), C.POINTER(Capsule)]),
    "m3CreateHullShape": (ShapeId, [BodyId, C.POINTER(ShapeDef), C.POINTER(Vec3), c_i32]),
    "m3Hash64": (c_u64, [c_u64, C.c_void_p, c_i32]),
    "m3DefaultJointDef": (JointDef, []),
    "m3CreateJoint": (JointId, [C.POINTER(JointDef)]),
    "m3DestroyJoint": (None, [JointId]),
    "m3Joint_IsValid": (C.c_bool, [JointId]),
}


class AdjointDef(C.Structure):
    """m3AdjointDef (this build's trajectory-gradient request)."""

    _fields_ = [
        ("timeStep", C.c_float),
        ("subStepCount", c_i32),
        ("steps", c_i32),
        ("checkpointEvery", c_i32),
        ("target", Vec3),
        ("force", C.POINTER(Vec3)),
        ("forwardOnly", c_i32),
        ("offsetKind", c_i32),
        ("offsetBody", c_i32),
        ("offsetComp", c_i32),
        ("offsetValue", C.c_double),
        ("precision", c_i32),
    ]


class AdjointResult(C.Structure):
    """m3AdjointResult: loss, scope counters and caller-owned gradient arrays."""

    _fields_ = [
        ("loss", C.c_double),
        ("unsupported", c_i32),
        ("ccdPulled", c_i32),
        ("gradLinearVelocity", C.POINTER(Vec3)),
        ("gradAngularVelocity", C.POINTER(Vec3)),
        ("gradMass", C.POINTER(C.c_float)),
        ("gradFriction", C.POINTER(C.c_float)),
        ("gradForce", C.POINTER(Vec3)),
    ]


AUTOTUNE_MAX_CANDIDATES = 16


class AutotuneResult(C.Structure):
    """m3AutotuneResult (this build's launch autotuner)."""

    _fields_ = [
        ("batchThreads", c_i32),
        ("substepBlocks", c_i32),
        ("tailLimit", c_i32),
        ("candidates", c_i32),
        ("candidateA", c_i32 * AUTOTUNE_MAX_CANDIDATES),
        ("candidateB", c_i32 * AUTOTUNE_MAX_CANDIDATES),
        ("candidateMs", C.c_float * AUTOTUNE_MAX_CANDIDATES),
    ]


_SIGNATURES["m3World_Autotune"] = (
    c_i32,
    [C.POINTER(WorldId), c_i32, C.c_float, c_i32, c_i32, C.POINTER(AutotuneResult)],
)
_SIGNATURES["m3DefaultAdjointDef"] = (AdjointDef, [])
_SIGNATURES["m3World_Adjoint"] = (
    c_i32,
    [C.POINTER(WorldId), c_i32, C.POINTER(AdjointDef), C.POINTER(AdjointResult)],
)


class ContactEvent(C.Structure):  # include/maul3d/world.h:213-217
    _fields_ = [("shapeA", ShapeId), ("shapeB", ShapeId)]


class HitEvent(C.Structure):  # include/maul3d/world.h:273-280
    _fields_ = [
        ("shapeA", ShapeId),
        ("shapeB", ShapeId),
        ("point", Pos3),
        ("normal", Vec3),
        ("approachSpeed", c_f),
    ]


class BodyMoveEvent(C.Structure):  # include/maul3d/world.h:295-300
    _fields_ = [("body", BodyId), ("transform", Transform), ("fellAsleep", C.c_bool)]


class JointBreakEvent(C.Structure):  # include/maul3d/world.h:308-311
    _fields_ = [("joint", c_i32), ("world0", c_u16), ("generation", c_u16)]


EVENT_STREAMS = {
    "m3World_ContactBeginEvents": ContactEvent,
    "m3World_ContactEndEvents": ContactEvent,
    "m3World_SensorBeginEvents": ContactEvent,
    "m3World_SensorEndEvents": ContactEvent,
    "m3World_HitEvents": HitEvent,
    "m3World_BodyMoveEvents": BodyMoveEvent,
    "m3World_JointBreakEvents": JointBreakEvent,
}
for _name, _rec in EVENT_STREAMS.items():
    _SIGNATURES[_name] = (C.POINTER(_rec), [WorldId, C.POINTER(c_i32)])
_SIGNATURES["m3World_HitEventsDropped"] = (c_i32, [WorldId])
_SIGNATURES["m3World_JournalBegin"] = (C.c_bool, [WorldId, C.c_void_p, c_i32])
_SIGNATURES["m3World_JournalEnd"] = (c_i32, [WorldId])
_SIGNATURES["m3World_JournalReplay"] = (C.c_bool, [WorldId, C.c_void_p, c_i32])
_SIGNATURES["m3World_SetHitEventThreshold"] = (None, [WorldId, c_f])
_SIGNATURES["m3Shape_EnableHitEvents"] = (None, [ShapeId, C.c_bool])
_SIGNATURES["m3Shape_AreHitEventsEnabled"] = (C.c_bool, [ShapeId])
_SIGNATURES["m3World_StepMany"] = (c_i32, [C.POINTER(WorldId), c_i32, c_f, c_i32])
_SIGNATURES["m3World_GetDeviceError"] = (C.c_char_p, [WorldId])
_SIGNATURES["m3World_EnableTreeMirror"] = (C.c_bool, [WorldId, C.c_bool])
_SIGNATURES["m3World_IsTreeMirrorCurrent"] = (C.c_bool, [WorldId])
_SIGNATURES["m3World_RebuildBroadphase"] = (None, [WorldId])
_SIGNATURES["m3World_DeviceWorld"] = (C.c_void_p, [WorldId])
_SIGNATURES["m3Shape_SetFriction"] = (None, [ShapeId, c_f])
_SIGNATURES["m3Shape_GetFriction"] = (c_f, [ShapeId])
_SIGNATURES["m3Shape_SetRestitution"] = (None, [ShapeId, c_f])
_SIGNATURES["m3Shape_GetRestitution"] = (c_f, [ShapeId])
_SIGNATURES["m3Shape_SetRollingResistance"] = (None, [ShapeId, c_f])
_SIGNATURES["m3Shape_GetRollingResistance"] = (c_f, [ShapeId])
_SIGNATURES["m3Shape_SetDensity"] = (None, [ShapeId, c_f, C.c_bool])
_SIGNATURES["m3Shape_GetDensity"] = (c_f, [ShapeId])
_SIGNATURES["m3Shape_IsValid"] = (C.c_bool, [ShapeId])
_SIGNATURES["m3Shape_GetBody"] = (BodyId, [ShapeId])
_SIGNATURES["m3Body_GetMass"] = (c_f, [BodyId])


def _id_tuple(sid) -> tuple[int, int]:
    """(index1, generation): the world slot is left out, so ids compare
    across libraries and processes."""
    return (sid.index1, sid.generation)


class Api:
    """Typed view of one library's ``m3*`` exports."""

    def __init__(
        self,
        lib: C.CDLL,
        optional: tuple[str, ...] = (
            "m3DestroyBody",
            "m3Body_SetAwake",
            "m3CreateHullShape",
            "m3World_GetProfile",
            "m3DefaultAdjointDef",
            "m3World_Adjoint",
            "m3World_Autotune",
            "m3World_StepMany",
            "m3World_GetDeviceError",
            "m3World_EnableTreeMirror",
            "m3World_IsTreeMirrorCurrent",
            "m3World_DeviceWorld",
            "m3Body_GetMass",
        ),
    ):
        self.lib = lib
        # (enqueueTask, finishTask, userTaskContext) pointers written into
        # every m3WorldDef the scene builders make through this Api, or None
        # for serial worlds (the reference default, include/maul3d/world.h:60).
        self.task_hooks: tuple[int, int, int] | None = None
        for name, (res, args) in _SIGNATURES.items():
            try:
                fn = getattr(lib, name)
            except AttributeError:
                if name in optional:
                    continue
                raise
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    # convenience wrappers -------------------------------------------------
    def events(self, world: WorldId) -> dict:
        """The last step's event streams as plain tuples (bit patterns of the
        floating-point fields, so comparisons are exact)."""
        import struct

        def bits_d(x: float) -> int:
            return struct.unpack("<Q", struct.pack("<d", x))[0]

        def bits_f(x: float) -> int:
            return struct.unpack("<I", struct.pack("<f", x))[0]

        out = {}
        for name, rec in EVENT_STREAMS.items():
            n = c_i32(0)
            ptr = getattr(self, name)(world, C.byref(n))
            items = []
            for i in range(n.value):
                e = ptr[i]
                if rec is ContactEvent:
                    items.append((_id_tuple(e.shapeA), _id_tuple(e.shapeB)))
                elif rec is HitEvent:
                    items.append(
                        (
                            _id_tuple(e.shapeA),
                            _id_tuple(e.shapeB),
                            tuple(bits_d(c) for c in (e.point.x, e.point.y, e.point.z)),
                            tuple(bits_f(c) for c in (e.normal.x, e.normal.y, e.normal.z)),
                            bits_f(e.approachSpeed),
                        )
                    )
                elif rec is BodyMoveEvent:
                    t = e.transform
                    items.append(
                        (
                            _id_tuple(e.body),
                            tuple(bits_d(c) for c in (t.p.x, t.p.y, t.p.z)),
                            tuple(bits_f(c) for c in (t.q.x, t.q.y, t.q.z, t.q.w)),
                            bool(e.fellAsleep),
                        )
                    )
                else:
                    items.append((e.joint, e.generation))
            out[name.replace("m3World_", "")] = items
        out["HitEventsDropped"] = self.m3World_HitEventsDropped(world)
        return out

    def snapshot(self, world: WorldId) -> bytes:
        size = self.m3World_SnapshotSize(world)
        if size < 0:
            raise RuntimeError("m3World_SnapshotSize refused the world")
        buf = C.create_string_buffer(size)
        written = self.m3World_Snapshot(world, buf, size)
        if written != size:
            raise RuntimeError(f"m3World_Snapshot wrote {written} of {size} bytes")
        return buf.raw

    def restore(self, world: WorldId, data: bytes) -> bool:
        buf = C.create_string_buffer(data, len(data))
        return bool(self.m3World_Restore(world, buf, len(data)))
