// m3gpu: the thin C ABI between the C17 host world and the gfx950 kernels.
//
// The host world (maul3d_amd/csrc/world.c) owns creation, ids, hashing and
// snapshots; this layer owns the device mirror of the world state and runs
// the whole step on the GPU: broadphase pair list, narrowphase manifolds and
// warm-start carry, island wake, contact prepare, graph coloring, the
// substep loop (integrate / warm start / solve / integrate positions /
// relax), restitution, impulse store, continuous collision against planes
// and the island sleep pass.  It replaces the body of m3StepInternal
// (src/solver.c:3662-4320) behind m3World_Step (src/solver.c:4322).
//
// Every entry point takes plain pointers and sizes.  Return codes: 0 = ok,
// M3GPU_STALL = the step refused for capacity (the world did not advance),
// M3GPU_REFUSED = the world holds something the device step does not
// simulate (nothing was enqueued), negative = a HIP error (the message is
// available from m3gpu_last_error).
#ifndef M3GPU_H
#define M3GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M3GPU_API __attribute__((visibility("default"))) extern

#define M3GPU_STALL 1
#define M3GPU_REFUSED 2

// Largest shapeCapacity of a device world: pair keys are packed as two
// 16-bit shape slots on the device.
#define M3_DEVICE_MAX_SHAPES 65535

// Field groups for upload/download.
#define M3GPU_FIELDS_STATIC  0x1u // mass, shapes, hulls, filters, materials
#define M3GPU_FIELDS_DYNAMIC 0x2u // transforms, velocities, awake, sleep timers, forces
#define M3GPU_FIELDS_PAIRS   0x4u // pair keys, manifolds, frozen keys, leaf boxes
#define M3GPU_FIELDS_ALL     0x7u
// The pair table in the reference's layout (pairCapacity keys and
// manifolds, the tail past the current count holding what earlier steps
// wrote there); only while m3gpu_set_pair_table has it on.
#define M3GPU_FIELDS_PAIR_TABLE 0x8u

typedef struct m3gpuWorld m3gpuWorld;

typedef struct m3gpuCaps
{
    int32_t bodyCapacity;
    int32_t shapeCapacity;
    int32_t pairCapacity;
    int32_t hullCapacity;
    int32_t jointCapacity;
} m3gpuCaps;

// Host arrays of one world.  Element types are the records of m3_core.h
// (sizes noted); counts are the id-pool high-water marks.
typedef struct m3gpuHostState
{
    int32_t maxBody;
    int32_t maxShape;
    int32_t maxHull;
    // bodies (maxBody each)
    void* transforms;        // m3Transform, 40 B
    void* linearVelocities;  // m3Vec3, 12 B
    void* angularVelocities; // m3Vec3
    float* invMass;
    void* invInertiaLocal; // m3Mat3, 36 B
    void* inertiaLocal;    // m3Mat3
    void* localCenters;    // m3Vec3
    float* gravityScales;
    float* linearDamping;
    float* angularDamping;
    uint8_t* bodyAlive;
    uint8_t* types;
    uint8_t* awake;
    float* sleepTimes;
    uint8_t* bulletFlags;
    float* minExtents;
    float* maxExtents;
    void* bodyForce;  // m3Vec3
    void* bodyTorque; // m3Vec3
    uint8_t* bodyEnabled;
    uint8_t* bodyLocks;
    float* bodySleepThreshold;
    uint8_t* bodyCanSleep;
    uint8_t* bodyHasTarget;
    void* bodyTarget; // m3Transform
    int32_t* bodyShapeHead;
    // shapes (maxShape each)
    uint8_t* shapeAlive;
    int32_t* shapeBody;
    uint8_t* shapeType;
    void* shapeGeom; // m3ShapeGeom, 32 B
    float* shapeFriction;
    float* shapeRestitution;
    float* shapeRollingResistance;
    uint8_t* shapeSensor;
    uint8_t* shapeHitEvents;
    uint8_t* shapePreSolve;
    uint8_t* shapeHasOffset;
    void* shapeLocalPos;   // m3Vec3
    void* shapeLocalRot;   // m3Quat
    void* shapeSurfaceVel; // m3Vec3
    uint64_t* shapeCategory;
    uint64_t* shapeMask;
    int32_t* shapeGroup;
    int32_t* shapeNext;
    int32_t* shapeHullIndex;
    double* leafBoxes; // 6 doubles per shape: lo xyz, hi xyz (the proxy's stored box)
    // hulls (maxHull)
    void* hullData; // m3HullData, 5808 B
    // pairs
    int32_t* pairCount;
    uint64_t* pairKeys;  // pairCapacity
    void* manifolds;     // m3Manifold, 188 B, pairCapacity
    int32_t* sleepingPairCount;
    uint64_t* sleepingPairKeys; // pairCapacity
    uint8_t* frozenDirty;
    uint8_t* pairsFullQuery;
    uint64_t* stepCount; // advanced on the device by every step that completes
    // joints (static group unless noted): slot arrays in the reference's
    // layout, the body joint lists, and the joint components the host
    // derives (joints linked through non-static bodies; component k owns
    // jointCompList[jointCompStart[k] .. jointCompStart[k + 1]), ascending)
    int32_t maxJoint;
    int32_t jointCompCount;
    uint8_t* jointAlive;
    uint8_t* jointType;
    int32_t* jointBodyA;
    int32_t* jointBodyB;
    void* jointLocalA; // m3Vec3
    void* jointLocalB;
    uint8_t* jointCollide;
    void* jointFrameQA; // m3Quat
    void* jointFrameQB;
    uint8_t* jointFlags;
    void* jointLimits;         // m3Vec3: twist lower, twist upper, cone
    void* jointImpulse;        // m3Vec3, dynamic group
    void* jointPerpImpulse;    // m3Vec3, dynamic group
    void* jointLimitImpulse;   // m3Vec3, dynamic group
    void* jointAngularImpulse; // m3Vec3, dynamic group
    void* jointSpringImpulse;  // m3Vec3, dynamic group
    int32_t* jointNextA;
    int32_t* jointNextB;
    int32_t* bodyJointHead;
    int32_t* jointCompStart; // jointCompCount + 1
    int32_t* jointCompList;  // maxJoint
} m3gpuHostState;

typedef struct m3gpuStepParams
{
    float dt;
    int32_t substeps;
    float gravity[3];
    float contactHertz;
    float contactDampingRatio;
    float contactPushMaxSpeed;
    float restitutionThreshold;
    float maximumLinearSpeed;
    float maximumAngularSpeed;
    int32_t sleepEnabled;
    int32_t continuousEnabled;
} m3gpuStepParams;

typedef struct m3gpuStepStats
{
    int32_t pairCount;
    int32_t constraintCount;
    int32_t colorCount;
    int32_t islandCount;
    int32_t awakeCount;
    float msBroadphase;
    float msNarrowphase;
    float msPrepare;
    float msSolve;
    float msContinuous;
    float msSleep;
    float msStep;
} m3gpuStepStats;

M3GPU_API int m3gpu_device_count(void);
M3GPU_API const char* m3gpu_last_error(void);

M3GPU_API m3gpuWorld* m3gpu_world_create(const m3gpuCaps* caps);
M3GPU_API void m3gpu_world_destroy(m3gpuWorld* world);
// Releases the process-wide device state (batch launch graphs and views,
// the adjoint arena, their streams); later calls recreate it on demand.
// Every world must be destroyed first.
M3GPU_API void m3gpu_shutdown(void);

// Copies the selected field groups host -> device (upload) or device -> host
// (download).  Both block until the copy is complete.
M3GPU_API int m3gpu_upload(m3gpuWorld* world, const m3gpuHostState* host, uint32_t fields);
M3GPU_API int m3gpu_download(m3gpuWorld* world, m3gpuHostState* host, uint32_t fields);

// One full step on the device.  When stats is non-null the call waits for
// the step and fills it (counts, per-stage device milliseconds); with null
// stats the step is only enqueued and the caller may return at once.
M3GPU_API int m3gpu_step(m3gpuWorld* world, const m3gpuStepParams* params,
                         m3gpuStepStats* stats);

// Keep (or stop keeping) the pair table in the reference's layout: each
// step writes its pair list over the table's first pairCount entries and
// leaves the rest, as the reference's single pair array does.  The table is
// allocated on first use; its content comes from an upload of
// M3GPU_FIELDS_PAIR_TABLE.
M3GPU_API int m3gpu_set_pair_table(m3gpuWorld* world, int32_t on);

// 1 when m3gpu_step would run this world with these params in one call, 0
// when it would return M3GPU_REFUSED.
M3GPU_API int m3gpu_step_supported(const m3gpuWorld* world, const m3gpuStepParams* params);

// 1 when this step's continuous pass meets shapes of the broadphase tree
// (continuous collision on, and a static non-plane shape or a bullet body
// present).  The reference sweeps those in its tree's depth-first order
// (src/solver.c:3402), so the host, which keeps the tree, steps such a world
// in two calls: m3gpu_step_begin enqueues everything before the continuous
// pass, m3gpu_step_leaves waits for the step's leaf boxes (maxShape x 6
// doubles: lo xyz, hi xyz; valid until the next step), the host replays the
// tree's leaf moves from them and passes every shape's depth-first rank
// (-1: not in the tree) to m3gpu_step_end, which enqueues the rest.
M3GPU_API int m3gpu_step_needs_tree(const m3gpuWorld* world, const m3gpuStepParams* params);

// Flags of m3gpu_step_begin.
#define M3GPU_STEP_LEAVES 0x1u   // copy the step's leaf boxes for m3gpu_step_leaves
#define M3GPU_STEP_PRESOLVE 0x2u // stop after the island wake pass for the pre-solve callback

// The pre-solve callback's inputs for one candidate pair
// (src/solver.c:214-236): the pair's key and index in this step's list, the
// deepest contact point and the normal (A to B).
typedef struct m3gpuPreSolve
{
    uint64_t key;
    int32_t pair;
    int32_t pad;
    double point[3];
    float normal[3];
    float pad2;
} m3gpuPreSolve;

// m3gpu_step_begin enqueues the step up to the continuous pass.  With
// M3GPU_STEP_PRESOLVE it stops after the island wake pass instead:
// m3gpu_step_presolve waits and returns the candidates in pair order (host
// memory owned by the device layer, valid until the next step), and
// m3gpu_step_vetoes takes the keys the host vetoed (ascending) and enqueues
// the rest up to the continuous pass.  Without it, `vetoKeys` (ascending;
// null or count 0 for none) are vetoed directly, as a replayed journal's
// recorded vetoes are (src/solver.c:184-212).
M3GPU_API int m3gpu_step_begin(m3gpuWorld* world, const m3gpuStepParams* params, uint32_t flags,
                               const uint64_t* vetoKeys, int32_t vetoCount);
M3GPU_API int m3gpu_step_presolve(m3gpuWorld* world, const m3gpuPreSolve** candidates, int32_t* count);
M3GPU_API int m3gpu_step_vetoes(m3gpuWorld* world, const m3gpuStepParams* params, const uint64_t* vetoKeys,
                                int32_t vetoCount);
M3GPU_API const double* m3gpu_step_leaves(m3gpuWorld* world);
M3GPU_API int m3gpu_step_end(m3gpuWorld* world, const m3gpuStepParams* params, const int32_t* ranks);

// Steps several worlds back to back on the device, each with its own
// params (every world must be supported); returns 0, or -1 on a HIP error.
M3GPU_API int m3gpu_step_many(m3gpuWorld* const* worlds, const m3gpuStepParams* params,
                              int32_t count);

// Launch tuning (no counterpart in the reference, whose solver has no
// launch shapes; used by m3World_Autotune).  None of these settings changes
// what a step computes.
//
// Workgroups of one world's cooperative substep grid (0: one per 256
// bodies, at most 40) and the constraint budget of the trailing colors its
// workgroup 0 solves without grid barriers (0: 1024).  Fails when blocks
// exceeds the co-resident limit of the cooperative launch.
M3GPU_API int m3gpu_set_substep_tuning(m3gpuWorld* world, int32_t blocks, int32_t tailLimit);
// The values the world's next step launches with.
M3GPU_API int m3gpu_get_substep_tuning(const m3gpuWorld* world, int32_t* blocks,
                                       int32_t* tailLimit);
// Threads per world workgroup of the batched substep kernel, process-wide
// (0: MAUL3D_AMD_BATCH_THREADS, else 256; otherwise a multiple of 64 in
// [64, 512]).  Waits for queued batch work and drops the captured batch
// launch graphs, which hold the old size.
M3GPU_API int m3gpu_set_batch_threads(int32_t threads);
M3GPU_API int32_t m3gpu_get_batch_threads(void);

// Event streams of the world's last step (src/solver.c:3732-3815,
// 2991-3026, 4295-4306), computed on the device when called: contact and
// sensor begin/end events as pair keys in ascending key order, hit events
// (hit threshold = the world's at that step) in pair order and body move
// events in ascending body order.  The caller's arrays hold pairCapacity
// keys or hits and bodyCapacity moves.  After a capacity stall every count
// is zero and `stalled` is set.
typedef struct m3gpuHit
{
    uint64_t key; // shape A << 32 | shape B
    double point[3];
    float normal[3];
    float approachSpeed;
} m3gpuHit;
typedef struct m3gpuMove
{
    int32_t body;
    uint8_t fellAsleep;
    uint8_t pad[3];
    double p[3]; // post-step transform (m3Transform: f64 position, f32 quaternion)
    float q[4];
} m3gpuMove;
typedef struct m3gpuEvents
{
    uint64_t* beginKeys;
    uint64_t* endKeys;
    uint64_t* sensorBeginKeys;
    uint64_t* sensorEndKeys;
    m3gpuHit* hits;
    m3gpuMove* moves;
    int32_t counts[7]; // begin, end, sensor begin, sensor end, hits, moves, stalled
} m3gpuEvents;
M3GPU_API int m3gpu_events(m3gpuWorld* world, float hitThreshold, m3gpuEvents* out);
// m3gpu_events without allocating (no counterpart in the reference): for
// callers that cannot allocate (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  m3gpu_events allocates a scratch per
// world on its first call; this variant uses instead one file-scope device
// array per GPU (about 3.1 MB, held until the process ends) and so serves
// worlds of at most 32768 pair slots (4096 shapes) and 8192 bodies; larger
// worlds get -1 with a message.  Calls that overlap on one GPU share the
// array and race.  Same outputs as m3gpu_events.
M3GPU_API int m3gpu_events_noalloc(m3gpuWorld* world, float hitThreshold, m3gpuEvents* out);
// Waits for all queued device work of this world.
M3GPU_API int m3gpu_sync(m3gpuWorld* world);

// Waits for the world's last step and reports its counts and stage times.
M3GPU_API int m3gpu_last_stats(m3gpuWorld* world, m3gpuStepStats* stats);

// Stage profile: between begin and end every step of the world waits for
// its own completion and adds each stage's HIP-event time (events recorded
// on the world's stream around each launch) and its work counts.
#define M3GPU_MAX_STAGES 64
typedef struct m3gpuStageTime
{
    char name[32];
    int64_t launches;
    double ms;
} m3gpuStageTime;

typedef struct m3gpuProfile
{
    int32_t stageCount;
    m3gpuStageTime stage[M3GPU_MAX_STAGES];
    int64_t steps;
    int64_t sumPairs;       // pair list entries, summed over profiled steps
    int64_t sumOldPairs;    // previous-step pairs (the warm-start carry source)
    int64_t sumConstraints; // contact constraints
    int64_t sumPoints;      // contact points over all constraints
    int64_t sumMovers;      // awake dynamic and kinematic bodies
    int64_t sumSubsteps;
    int64_t sumShapes; // shape high-water mark
    int64_t sumBodies; // body high-water mark
    int64_t sumOverflow; // constraints in the serial overflow color
} m3gpuProfile;

M3GPU_API int m3gpu_profile_begin(m3gpuWorld* world);
M3GPU_API int m3gpu_profile_end(m3gpuWorld* world, m3gpuProfile* out);

// Trajectory gradients.  From a world's current state, `steps` steps with
// `params` are simulated (every step applies `force` to each dynamic body),
// checkpointing the state every `checkpointEvery` steps; the loss
// |COM_T - target|^2 of the final mass-weighted centre of mass and its
// gradients with respect to every body's initial linear and angular
// velocity, mass, friction (of its first shape) and applied force are
// returned by the reverse pass, which replays each segment from its
// checkpoint.  The world's state is left as it was.  Scope: dynamic hulls
// on static planes, no joints, restitution and rolling resistance zero.
// `offsetKind` (0 none, 1 linear velocity, 2 angular velocity, 3 mass,
// 4 friction, 5 force) shifts one parameter of body `offsetBody`, component
// `offsetComp`, by `offsetValue` before the run (finite differences).
typedef struct m3gpuAdjointRequest
{
    int32_t steps;
    int32_t checkpointEvery;
    int32_t forwardOnly; // loss only, no reverse pass
    float target[3];
    const float* force; // maxBody * 3, or null for none
    int32_t offsetKind;
    int32_t offsetBody;
    int32_t offsetComp;
    double offsetValue;
    // outputs (maxBody entries; may be null when forwardOnly)
    double* loss;
    float* gradLinearVelocity;  // maxBody * 3
    float* gradAngularVelocity; // maxBody * 3
    float* gradMass;            // maxBody
    float* gradFriction;        // maxBody
    float* gradForce;           // maxBody * 3
    int32_t* unsupported; // touching pairs or constraints outside the scope
    int32_t* ccdPulled;   // continuous pull-backs passed through as identity
    // 32: the float trajectory on the device's step kernels.  64: the step
    // and its reverse in double precision, from `host` (the world's current
    // host arrays, required then); every request of a call shares it.
    int32_t precision;
    const m3gpuHostState* host;
} m3gpuAdjointRequest;

M3GPU_API int m3gpu_adjoint_many(m3gpuWorld* const* worlds, const m3gpuStepParams* params,
                                 const m3gpuAdjointRequest* requests, int32_t count);

#ifdef __cplusplus
}
#endif

#endif // M3GPU_H
