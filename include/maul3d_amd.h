// maul3d_amd: the Maul3D public C surface for the MI355X build.
//
// This header declares the subset of the reference API that the soft-step
// contact-solver path needs, with the same names, layouts and error
// behaviour, so a host that links the reference can link
// libmaul3d_amd.so instead.  Each declaration names the reference
// declaration it replaces.  The step runs on one MI355X through the
// m3gpu_* layer (include/m3gpu.h); everything here is plain C.
#ifndef MAUL3D_AMD_H
#define MAUL3D_AMD_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M3_API __attribute__((visibility("default"))) extern

// ---- math types (include/maul3d/math.h:27-55) -----------------------------
#ifndef M3_CORE_H
typedef float m3real;
typedef struct m3Vec3 { m3real x, y, z; } m3Vec3;
typedef struct m3Pos3 { double x, y, z; } m3Pos3;
typedef struct m3Quat { m3real x, y, z, w; } m3Quat;
typedef struct m3Transform { m3Pos3 p; m3Quat q; } m3Transform;
typedef struct m3Mat3 { m3Vec3 cx, cy, cz; } m3Mat3;
#endif

// ---- ids (include/maul3d/base.h:69-104) ------------------------------------
typedef struct m3WorldId { int32_t index1; uint16_t generation; } m3WorldId;
typedef struct m3BodyId { int32_t index1; uint16_t world0; uint16_t generation; } m3BodyId;
typedef struct m3ShapeId { int32_t index1; uint16_t world0; uint16_t generation; } m3ShapeId;
typedef struct m3JointId { int32_t index1; uint16_t world0; uint16_t generation; } m3JointId;

typedef enum m3Result
{
    m3_success = 0,
    m3_errorInvalid = 1,
    m3_errorCapacity = 2,
    m3_errorConfig = 3,
} m3Result;

typedef enum m3BodyType
{
    m3_staticBody = 0,
    m3_kinematicBody = 1,
    m3_dynamicBody = 2,
} m3BodyType;

typedef enum m3ShapeType
{
    m3_sphereShape = 0,
    m3_planeShape = 1,
    m3_hullShape = 2,
    m3_capsuleShape = 3,
} m3ShapeType;

// ---- defs (include/maul3d/world.h:43-76, body.h:21-39, shape.h:40-74) ----
typedef void m3TaskFn(int32_t startIndex, int32_t endIndex, void* taskContext);
typedef void* m3EnqueueTaskFn(m3TaskFn* task, int32_t itemCount, int32_t minRange,
                              void* taskContext, void* userContext);
typedef void m3FinishTaskFn(void* userTask, void* userContext);

typedef struct m3WorldDef
{
    m3Vec3 gravity;
    int32_t bodyCapacity;
    int32_t shapeCapacity;
    int32_t meshCapacity;
    int32_t jointCapacity;
    int32_t voxelCapacity;
    int32_t characterCapacity;
    int32_t vehicleCapacity;
    int32_t softBodyCapacity;
    int32_t workerCount;
    m3EnqueueTaskFn* enqueueTask; // accepted for source compatibility; the
    m3FinishTaskFn* finishTask;   // device step never calls host task hooks
    void* userTaskContext;
    float contactHertz;
    float contactDampingRatio;
    float contactPushMaxSpeed;
    float restitutionThreshold;
    float maximumLinearSpeed;
    int32_t enableSleeping;
    int32_t enableContinuous;
    float hitEventThreshold;
    int32_t internalValue;
} m3WorldDef;

typedef struct m3BodyDef
{
    int32_t type;
    m3Pos3 position;
    m3Quat rotation;
    m3Vec3 linearVelocity;
    m3Vec3 angularVelocity;
    float gravityScale;
    float linearDamping;
    float angularDamping;
    uint64_t userData;
    bool isBullet;
    int32_t internalValue;
} m3BodyDef;

typedef struct m3ShapeDef
{
    float density;
    float friction;
    float restitution;
    uint64_t categoryBits;
    uint64_t maskBits;
    int32_t groupIndex;
    float rollingResistance;
    uint64_t userData;
    bool isSensor;
    bool enableHitEvents;
    bool enablePreSolveEvents;
    m3Vec3 localPosition;
    m3Quat localRotation;
    int32_t internalValue;
} m3ShapeDef;

// Joint kinds (include/maul3d/joint.h:12-62).  This build steps the
// spherical joint with its cone and twist limits; m3CreateJoint refuses the
// other kinds.
typedef enum m3JointType
{
    m3_sphericalJoint = 0,
    m3_revoluteJoint = 1,
    m3_prismaticJoint = 2,
    m3_fixedJoint = 3,
    m3_distanceJoint = 4,
    m3_genericJoint = 5,
    m3_wheelJoint = 6,
    m3_filterJoint = 7,
    m3_parallelJoint = 8,
    m3_motorJoint = 9,
    m3_gearJoint = 10,
    m3_pulleyJoint = 11,
} m3JointType;

// The joint definition, field for field (include/maul3d/joint.h:80-163);
// build it with m3DefaultJointDef.
typedef struct m3JointDef
{
    int32_t type;
    m3BodyId bodyA;
    m3BodyId bodyB;
    m3Vec3 localAnchorA;
    m3Vec3 localAnchorB;
    m3Vec3 localAxisA;
    m3Vec3 localAxisB;
    bool enableLimit; // spherical: twist range [lowerLimit, upperLimit]
    m3real lowerLimit;
    m3real upperLimit;
    bool enableMotor;
    m3real motorSpeed;
    m3real maxMotorEffort;
    bool enableCone; // spherical: swing within coneAngle of A's frame z-axis
    m3real coneAngle;
    bool collideConnected;
    uint8_t genericLinear[3];
    uint8_t genericAngular[3];
    uint8_t genericMotorAxis;
    uint8_t genericReserved;
    m3real genericLinearLower[3];
    m3real genericLinearUpper[3];
    m3real genericAngularLower[3];
    m3real genericAngularUpper[3];
    m3real ratio;
    m3Pos3 groundAnchorA;
    m3Pos3 groundAnchorB;
    int32_t internalValue;
} m3JointDef;

typedef struct m3Sphere { m3Vec3 center; m3real radius; } m3Sphere;
typedef struct m3Plane { m3Vec3 normal; m3real offset; } m3Plane;
typedef struct m3Capsule { m3Vec3 point1; m3Vec3 point2; m3real radius; } m3Capsule;

// include/maul3d/world.h:531-552
typedef struct m3Counters
{
    int32_t bodyCount;
    int32_t shapeCount;
    int32_t jointCount;
    int32_t contactCount;
    int32_t awakeCount;
    int32_t islandCount;
    int32_t colorCount;
    int32_t characterCount;
    int32_t vehicleCount;
    int32_t softBodyCount;
    int32_t voxelChunkCount;
    int32_t hullCount;
    int32_t meshCount;
    int32_t treeHeight;
    int32_t scratchPeak;
    int32_t scratchCapacity;
    int32_t snapshotBytes;
    int64_t allocCalls;
    int64_t freeCalls;
} m3Counters;

// include/maul3d/world.h:578-591 (device-timed per stage)
typedef struct m3Profile
{
    float step;
    float vehicles;
    float broadphase;
    float narrowphase;
    float events;
    float prepare;
    float solve;
    float continuous;
    float sleep;
    float characters;
    float softBodies;
} m3Profile;

// ---- API -------------------------------------------------------------------
// Citations are include/maul3d/<file>:<line> of the reference declaration.
M3_API int m3GetVersion(void);                                          // base.h:39
M3_API uint64_t m3Hash64(uint64_t h, const void* bytes, int32_t count); // base.h:112

M3_API m3WorldDef m3DefaultWorldDef(void);                         // world.h:82
M3_API m3WorldId m3CreateWorld(const m3WorldDef* def);             // world.h:86
M3_API void m3DestroyWorld(m3WorldId worldId);                     // world.h:87
M3_API bool m3World_IsValid(m3WorldId worldId);                    // world.h:88
M3_API void m3World_SetGravity(m3WorldId worldId, m3Vec3 gravity); // world.h:92
M3_API m3Vec3 m3World_GetGravity(m3WorldId worldId);               // world.h:134
M3_API void m3World_EnableSleeping(m3WorldId worldId, bool flag);  // world.h:157
M3_API void m3World_EnableContinuous(m3WorldId worldId, bool flag); // world.h:163

// The drop-in boundary: world.h:195 (Step), :203-205 (Snapshot/Restore),
// :510 (Hash).  Step is void and silently refuses a bad id, dt <= 0 or
// substeps outside [1, 256]; a capacity failure stalls the world without
// advancing stepCount.  Snapshot returns the size written or -1; Restore is
// atomic and returns false without touching the world on bad bytes.
M3_API void m3World_Step(m3WorldId worldId, float dt, int32_t substeps);
M3_API int32_t m3World_SnapshotSize(m3WorldId worldId);
M3_API int32_t m3World_Snapshot(m3WorldId worldId, void* out, int32_t capacity);
M3_API bool m3World_Restore(m3WorldId worldId, const void* data, int32_t size);
M3_API uint64_t m3World_Hash(m3WorldId worldId);
M3_API m3Counters m3World_GetCounters(m3WorldId worldId);       // world.h:554
M3_API m3Profile m3World_GetProfile(m3WorldId worldId);         // world.h:593

// ---- event streams of the last step (include/maul3d/world.h:206-313) -------
// Same records, order and lifetime as the reference: pointers stay valid until
// the next step, restore or world destruction; restore clears the streams.
// The device builds them when first read after a step (or before a host
// mutation of the world); joints never break in this build, so the joint
// break stream is always empty, and hits never overflow (at most one per
// constraint).
typedef struct m3ContactEvent // world.h:213-217
{
    m3ShapeId shapeA; // the lower shape index of the pair
    m3ShapeId shapeB;
} m3ContactEvent;
typedef struct m3HitEvent // world.h:273-280
{
    m3ShapeId shapeA;
    m3ShapeId shapeB;
    m3Pos3 point;
    m3Vec3 normal;
    m3real approachSpeed;
} m3HitEvent;
typedef struct m3BodyMoveEvent // world.h:295-300
{
    m3BodyId body;
    m3Transform transform;
    bool fellAsleep;
} m3BodyMoveEvent;
typedef struct m3JointBreakEvent // world.h:308-311
{
    m3JointId joint;
} m3JointBreakEvent;
M3_API const m3ContactEvent* m3World_ContactBeginEvents(m3WorldId worldId, int32_t* count); // world.h:223
M3_API const m3ContactEvent* m3World_ContactEndEvents(m3WorldId worldId, int32_t* count);   // world.h:224
M3_API const m3ContactEvent* m3World_SensorBeginEvents(m3WorldId worldId, int32_t* count);  // world.h:265
M3_API const m3ContactEvent* m3World_SensorEndEvents(m3WorldId worldId, int32_t* count);    // world.h:266
M3_API const m3HitEvent* m3World_HitEvents(m3WorldId worldId, int32_t* count);              // world.h:282
M3_API int32_t m3World_HitEventsDropped(m3WorldId worldId);                                 // world.h:285
M3_API void m3World_SetHitEventThreshold(m3WorldId worldId, float value);                   // world.h:288
M3_API const m3BodyMoveEvent* m3World_BodyMoveEvents(m3WorldId worldId, int32_t* count);    // world.h:302
M3_API const m3JointBreakEvent* m3World_JointBreakEvents(m3WorldId worldId, int32_t* count); // world.h:313
M3_API void m3Shape_EnableHitEvents(m3ShapeId shapeId, bool flag);                          // shape.h:300
// ---- journal (include/maul3d/world.h:600-606) ------------------------------
// The reference's record stream (op, byte count, payload; op codes and record
// layouts of src/world_internal.h:108-192, 543-581), for the calls this build
// supports: step, body/shape/hull/joint creation, joint destruction, velocity
// setters, forces and torques, gravity, the sleep and continuous toggles, the
// hit threshold and shape hit opt-in.  Replay is atomic: a journal holding
// any other op, or whose ids do not come out as recorded, is refused and the
// world is left as it was.
M3_API bool m3World_JournalBegin(m3WorldId worldId, void* buffer, int32_t capacity);       // world.h:600
M3_API int32_t m3World_JournalEnd(m3WorldId worldId);                                      // world.h:601
M3_API bool m3World_JournalReplay(m3WorldId worldId, const void* data, int32_t size);      // world.h:606
M3_API bool m3Shape_AreHitEventsEnabled(m3ShapeId shapeId);                                 // shape.h:301
// Shape materials (include/maul3d/shape.h:286-296): non-finite or negative
// values (density: not positive) are ignored; journaled as ops 40-43.
M3_API void m3Shape_SetFriction(m3ShapeId shapeId, float friction);                         // shape.h:286
M3_API float m3Shape_GetFriction(m3ShapeId shapeId);                                        // shape.h:287
M3_API void m3Shape_SetRestitution(m3ShapeId shapeId, float restitution);                   // shape.h:288
M3_API float m3Shape_GetRestitution(m3ShapeId shapeId);                                     // shape.h:289
M3_API void m3Shape_SetRollingResistance(m3ShapeId shapeId, float value);                   // shape.h:290
M3_API float m3Shape_GetRollingResistance(m3ShapeId shapeId);                               // shape.h:291
M3_API void m3Shape_SetDensity(m3ShapeId shapeId, float density, bool updateBodyMass);      // shape.h:295
M3_API float m3Shape_GetDensity(m3ShapeId shapeId);                                         // shape.h:296
M3_API bool m3Shape_IsValid(m3ShapeId shapeId);                                             // shape.h:267
M3_API m3BodyId m3Shape_GetBody(m3ShapeId shapeId);                                         // shape.h:271
// A dynamic body's mass (1 / its inverse mass), 0 for other bodies and bad
// ids.  No counterpart in the reference (which exposes only densities); the
// gradient harness reads it to write a mass through the shapes' densities.
M3_API float m3Body_GetMass(m3BodyId bodyId);

// Extensions of this build (no reference counterpart): step many worlds with
// one call (the batched rollback and toss workloads; returns the number of
// valid worlds stepped or -1 on bad arguments), and wait for a world's
// queued device work after uploading pending host changes.
M3_API int32_t m3World_StepMany(const m3WorldId* worlds, int32_t count, float dt,
                                int32_t substeps);
M3_API void m3World_Sync(m3WorldId worldId);
// Device fault report (no reference counterpart: the reference step cannot
// fail).  A world whose device call failed (a HIP error, a substep barrier
// that timed out) stops advancing: m3World_Step is a no-op, m3World_Hash
// returns 0, m3World_Snapshot returns -1 and m3World_Restore false.  Returns
// the failure text, or null for a healthy world or a bad id.
M3_API const char* m3World_GetDeviceError(m3WorldId worldId);
// The world's device mirror, for callers that drive the m3gpu layer
// (include/m3gpu.h) directly, e.g. m3gpu_events; null before the
// world's first step or after a device fault.  Nothing is allocated.
struct m3gpuWorld;
M3_API struct m3gpuWorld* m3World_DeviceWorld(m3WorldId worldId);
// Pre-solve veto (include/maul3d/world.h:315-340, shape.h:302-303): the
// callback runs on the host in pair order for every touching pair with an
// awake dynamic side where either shape opted in; returning false drops the
// contact for that step.  The device step pauses after its island wake pass
// while the host calls it.  Vetoes are journaled before their step and a
// replayed journal applies them with the callback silent.
typedef bool m3PreSolveFn(m3ShapeId shapeA, m3ShapeId shapeB, m3Pos3 point, m3Vec3 normal,
                          void* context);
M3_API void m3World_SetPreSolveCallback(m3WorldId worldId, m3PreSolveFn* fn, void* context); // world.h:340
M3_API void m3Shape_EnablePreSolve(m3ShapeId shapeId, bool flag);                            // shape.h:302
M3_API bool m3Shape_IsPreSolveEnabled(m3ShapeId shapeId);                                   // shape.h:303
// The broadphase tree (no reference counterpart: the reference always keeps
// its tree; the device step does not need one).  With the mirror on, the
// host replays every step's leaf moves into the reference's tree
// (src/broad_phase.c:505-539), so snapshots carry the reference's exact tree
// bytes and the world continues on the tree the reference would hold; it
// costs a host synchronisation and the tree work per step.  Off (the
// default), a snapshot carries the canonical rebuild of the current leaves
// (what m3World_RebuildBroadphase leaves).  A world whose continuous pass
// meets tree shapes (a static non-plane shape or a bullet) keeps the mirror
// current whatever the setting.  Returns false for a bad or faulted world.
M3_API bool m3World_EnableTreeMirror(m3WorldId worldId, bool flag);
// True when the host tree equals the reference's for this world's history.
M3_API bool m3World_IsTreeMirrorCurrent(m3WorldId worldId);
M3_API void m3World_RebuildBroadphase(m3WorldId worldId); // world.h:100

// The M3J1 replay container and the divergence finder
// (include/maul3d/replay.h:15-87): the same structs, framing and ordering.
typedef struct m3JournalInfo // replay.h:25-29
{
    int32_t opCount;
    int32_t stepCount;
} m3JournalInfo;

typedef struct m3ReplayView // replay.h:35-43
{
    const void* snapshot;
    int32_t snapshotBytes;
    const void* journal;
    int32_t journalBytes;
    uint64_t finalHash;
    m3JournalInfo info;
} m3ReplayView;

typedef struct m3BodyDiff // replay.h:63-71
{
    m3BodyId body;
    float positionError;
    float rotationError;
    float velocityError;
    bool onlyInA;
    bool onlyInB;
} m3BodyDiff;

M3_API bool m3JournalDescribe(const void* journal, int32_t bytes, m3JournalInfo* out); // replay.h:31
M3_API int32_t m3ReplayEncodeSize(int32_t snapshotBytes, int32_t journalBytes);        // replay.h:46
M3_API int32_t m3ReplayEncode(const void* snapshot, int32_t snapshotBytes, const void* journal,
                              int32_t journalBytes, uint64_t finalHash, void* out,
                              int32_t capacity);                                       // replay.h:52
M3_API bool m3ReplayDecode(const void* data, int32_t bytes, m3ReplayView* out);        // replay.h:58
M3_API int32_t m3World_DiffReport(m3WorldId worldA, m3WorldId worldB, m3BodyDiff* out,
                                  int32_t capacity, int32_t* outCount);                // replay.h:80
// Trajectory gradients (no reference counterpart; the reference has no
// differentiable path).  For each world, from its current state: `steps`
// steps of (timeStep, subStepCount), `force[b]` applied to dynamic body slot
// b every step (null: none), loss = |COM_T - target|^2 of the final
// mass-weighted centre of mass, and the loss's gradients with respect to
// each body slot's initial linear and angular velocity, mass, friction (its
// first shape's) and applied force, from hand-written adjoint kernels that
// replay checkpointed segments of `checkpointEvery` steps.  Worlds are left
// in their current state.  Gradients are taken on the piece of the step the
// forward point lies in (clamps, contacts and sleep as the forward took
// them).  Scope: dynamic hulls on static planes, no joints, restitution and
// rolling resistance zero; `unsupported` counts what fell outside it and
// `ccdPulled` the continuous pull-backs passed through as the identity.
// `offset*` shift one parameter before the run (finite-difference checks).
// Returns the number of worlds processed or -1 on bad arguments.
typedef struct m3AdjointDef
{
    float timeStep;
    int32_t subStepCount;
    int32_t steps;
    int32_t checkpointEvery;
    m3Vec3 target;
    const m3Vec3* force; // body slots
    int32_t forwardOnly;
    int32_t offsetKind; // 0 none, 1 linear velocity, 2 angular velocity, 3 mass, 4 friction, 5 force
    int32_t offsetBody; // body slot
    int32_t offsetComp;
    double offsetValue;
    // 32 (the default): the gradients of the float trajectory the step
    // runs, exact to the world's own steps.  64: the trajectory and its
    // reverse recomputed in double precision from the world's state, the
    // gradients FP64 finite differences converge to (one workgroup per
    // world; all worlds of a call share the precision).
    int32_t precision;
} m3AdjointDef;

typedef struct m3AdjointResult
{
    double loss;
    int32_t unsupported;
    int32_t ccdPulled;
    m3Vec3* gradLinearVelocity; // caller arrays of bodyCapacity entries (body slots)
    m3Vec3* gradAngularVelocity;
    float* gradMass;
    float* gradFriction;
    m3Vec3* gradForce;
} m3AdjointResult;

M3_API m3AdjointDef m3DefaultAdjointDef(void);
M3_API int32_t m3World_Adjoint(const m3WorldId* worlds, int32_t count, const m3AdjointDef* defs,
                               m3AdjointResult* results);

// Launch autotuning (no reference counterpart; the reference solver has no
// launch shapes).  Times candidate launch shapes on the worlds' current
// state for `trialSteps` steps each and keeps the fastest: for several
// worlds (a batch stepped with m3World_StepMany) the threads per world
// workgroup of the batched substep kernel (64, 128, 256, 512; process-wide);
// for one world its cooperative substep workgroups (default, 24, 32, 48)
// times the constraint budget of the trailing colors solved in workgroup 0
// (512, 1024, 2048).  Every world is snapshotted first and restored before
// each candidate and at the end, so the call leaves the worlds as they were;
// the shapes change a step's speed, never its result.  Returns the number of
// candidates timed, or -1 on bad arguments or a failed snapshot/restore.
#define M3_AUTOTUNE_MAX_CANDIDATES 16
typedef struct m3AutotuneResult
{
    int32_t batchThreads;  // chosen (batch), else 0
    int32_t substepBlocks; // chosen (one world): workgroups its steps launch, else 0
    int32_t tailLimit;     // chosen (one world), else 0
    int32_t candidates;
    int32_t candidateA[M3_AUTOTUNE_MAX_CANDIDATES]; // threads (batch) or workgroups (0: default)
    int32_t candidateB[M3_AUTOTUNE_MAX_CANDIDATES]; // 0 (batch) or tail budget
    float candidateMs[M3_AUTOTUNE_MAX_CANDIDATES];  // mean ms per step over the trial
} m3AutotuneResult;

M3_API int32_t m3World_Autotune(const m3WorldId* worlds, int32_t count, float timeStep,
                                int32_t subStepCount, int32_t trialSteps,
                                m3AutotuneResult* result);

// Stage profile of the steps between Begin and End (struct in m3gpu.h).
struct m3gpuProfile;
M3_API int32_t m3World_ProfileBegin(m3WorldId worldId);
M3_API int32_t m3World_ProfileEnd(m3WorldId worldId, struct m3gpuProfile* out);

M3_API m3BodyDef m3DefaultBodyDef(void);                               // body.h:43
M3_API m3BodyId m3CreateBody(m3WorldId worldId, const m3BodyDef* def); // body.h:49
M3_API bool m3Body_IsValid(m3BodyId bodyId);                           // body.h:55
M3_API m3Pos3 m3Body_GetPosition(m3BodyId bodyId);                     // body.h:56
M3_API m3Quat m3Body_GetRotation(m3BodyId bodyId);                     // body.h:57
M3_API m3Vec3 m3Body_GetLinearVelocity(m3BodyId bodyId);               // body.h:58
M3_API m3Vec3 m3Body_GetAngularVelocity(m3BodyId bodyId);              // body.h:59
M3_API void m3Body_SetLinearVelocity(m3BodyId bodyId, m3Vec3 velocity); // body.h:122
M3_API void m3Body_SetAngularVelocity(m3BodyId bodyId, m3Vec3 velocity); // body.h:123
M3_API bool m3Body_IsAwake(m3BodyId bodyId);                           // body.h:103
M3_API void m3Body_ApplyForce(m3BodyId bodyId, m3Vec3 force);          // body.h:115
M3_API void m3Body_ApplyTorque(m3BodyId bodyId, m3Vec3 torque);        // body.h:116

M3_API m3ShapeDef m3DefaultShapeDef(void);                      // shape.h:78
M3_API m3ShapeId m3CreateSphereShape(m3BodyId bodyId, const m3ShapeDef* def,
                                     const m3Sphere* sphere);   // shape.h:86
M3_API m3ShapeId m3CreatePlaneShape(m3BodyId bodyId, const m3ShapeDef* def,
                                    const m3Plane* plane);      // shape.h:91
M3_API m3ShapeId m3CreateBoxShape(m3BodyId bodyId, const m3ShapeDef* def,
                                  m3Vec3 halfExtents);          // shape.h:98
M3_API m3ShapeId m3CreateCapsuleShape(m3BodyId bodyId, const m3ShapeDef* def,
                                      const m3Capsule* capsule); // shape.h:112

M3_API m3ShapeId m3CreateHullShape(m3BodyId bodyId, const m3ShapeDef* def, const m3Vec3* points,
                                   int32_t count); // shape.h:148, 4..256 points

M3_API m3JointDef m3DefaultJointDef(void);                 // joint.h:168
M3_API m3JointId m3CreateJoint(const m3JointDef* def);    // joint.h:174 (spherical only)
M3_API void m3DestroyJoint(m3JointId jointId);            // joint.h:175
M3_API bool m3Joint_IsValid(m3JointId jointId);           // joint.h:176

#ifdef __cplusplus
}
#endif

#endif // MAUL3D_AMD_H
