// The host world of the MI355X build: id pools, world registry, creation
// with closed-form mass properties and canonical box hulls, the curated
// world hash, snapshot/restore, and the hand-off of every step to the
// device layer (include/m3gpu.h).  C17, compiled under the reference FP law
// (-ffp-contract=off, no fast-math) so creation-time floats are the
// reference's bits.
//
// Host state is authoritative for everything the step never writes (mass,
// shapes, hulls, filters).  The device is authoritative for what the step
// writes (transforms, velocities, sleep state, pairs, manifolds); the host
// copy is refreshed lazily on the first read after a step.

// clock_gettime (the autotuner's trial timing) under -std=c17.
#define _POSIX_C_SOURCE 199309L

#include "m3_core.h"

#include "maul3d_amd.h"
#include "m3gpu.h"
#include "m3_hull.h"
#include "m3_tree.h"

#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

// Larger than the reference's 64-slot table (src/world_internal.h:37) so one
// process can hold a 256-world rollback batch.
#define M3_MAX_WORLDS 1024

#define M3_COOKIE 0x4D33u
#define M3_WORLD_COOKIE ((int32_t)(M3_COOKIE ^ ((int32_t)sizeof(m3WorldDef) << 8)))
#define M3_BODY_COOKIE ((int32_t)(M3_COOKIE ^ ((int32_t)sizeof(m3BodyDef) << 8) ^ 1))
#define M3_SHAPE_COOKIE ((int32_t)(M3_COOKIE ^ ((int32_t)sizeof(m3ShapeDef) << 8) ^ 2))
#define M3_JOINT_COOKIE ((int32_t)(M3_COOKIE ^ ((int32_t)sizeof(m3JointDef) << 8) ^ 3))
_Static_assert(sizeof(m3JointDef) == 224, "m3JointDef layout");

// Tuning defaults (src/world_internal.h:1024-1046).
#define M3_CONTACT_HERTZ_DEFAULT 30.0f
#define M3_CONTACT_DAMPING_RATIO_DEFAULT 10.0f
#define M3_CONTACT_PUSH_MAX_SPEED_DEFAULT 3.0f
#define M3_RESTITUTION_THRESHOLD_DEFAULT 1.0f
#define M3_MAX_LINEAR_SPEED_DEFAULT 400.0f
#define M3_MAX_ANGULAR_SPEED_DEFAULT 800.0f
#define M3_HIT_EVENT_THRESHOLD_DEFAULT 1.0f
#define M3_SLEEP_VELOCITY_DEFAULT 0.05f

// ---------------------------------------------------------------------------
// FIFO id pools (src/allocator.c:178-218): recycled slots first in free
// order, then fresh slots; slot order defines pair keys.
// ---------------------------------------------------------------------------

typedef struct m3IdPool
{
    int32_t capacity;
    int32_t maxIndex;
    int32_t freeHead;
    int32_t freeCount;
    int32_t retiredCount;
    uint16_t* generations;
    uint8_t* alive;
    int32_t* freeQueue;
} m3IdPool;

static void* Zalloc(size_t n)
{
    void* p = calloc(n > 0 ? n : 1, 1);
    if (p == NULL)
    {
        fprintf(stderr, "maul3d_amd: out of host memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}

static m3IdPool PoolCreate(int32_t capacity)
{
    m3IdPool p;
    memset(&p, 0, sizeof(p));
    p.capacity = capacity;
    p.generations = (uint16_t*)Zalloc((size_t)capacity * sizeof(uint16_t));
    p.alive = (uint8_t*)Zalloc((size_t)capacity);
    p.freeQueue = (int32_t*)Zalloc((size_t)capacity * sizeof(int32_t));
    return p;
}

static void PoolDestroy(m3IdPool* p)
{
    free(p->generations);
    free(p->alive);
    free(p->freeQueue);
    memset(p, 0, sizeof(*p));
}

static int32_t PoolAlloc(m3IdPool* p)
{
    if (p->freeCount > 0)
    {
        int32_t index = p->freeQueue[p->freeHead];
        p->freeHead = (p->freeHead + 1) % p->capacity;
        p->freeCount -= 1;
        p->alive[index] = 1;
        return index;
    }
    if (p->maxIndex < p->capacity)
    {
        int32_t index = p->maxIndex;
        p->maxIndex += 1;
        p->alive[index] = 1;
        return index;
    }
    return -1;
}

static void PoolFree(m3IdPool* p, int32_t index)
{
    if (index < 0 || index >= p->capacity || p->alive[index] == 0)
    {
        return;
    }
    p->alive[index] = 0;
    if (p->generations[index] == 0xFFFF)
    {
        p->retiredCount += 1;
        return;
    }
    p->generations[index] += 1;
    p->freeQueue[(p->freeHead + p->freeCount) % p->capacity] = index;
    p->freeCount += 1;
}

static int PoolValid(const m3IdPool* p, int32_t index, uint16_t generation)
{
    return index >= 0 && index < p->capacity && p->alive[index] != 0 &&
           p->generations[index] == generation;
}

// ---------------------------------------------------------------------------
// World
// ---------------------------------------------------------------------------

typedef struct m3World
{
    m3Vec3 gravity;
    float contactHertz;
    float contactDampingRatio;
    float contactPushMaxSpeed;
    float restitutionThreshold;
    float maximumLinearSpeed;
    float maximumAngularSpeed;
    uint8_t sleepEnabled;
    uint8_t continuousEnabled;
    float hitEventThreshold;
    uint64_t stepCount;
    int32_t bodyCapacity;
    int32_t shapeCapacity;
    int32_t jointCapacity;
    // capacities of the def's subsystems this build does not simulate; they
    // size the empty blocks of the reference's snapshot layout
    int32_t meshCapacity;
    int32_t voxelCapacity;
    int32_t characterCapacity;
    int32_t vehicleCapacity;
    int32_t softBodyCapacity;
    uint16_t generation;
    uint16_t worldIndex0;

    m3IdPool bodyPool;
    m3Transform* transforms;
    m3Vec3* linearVelocities;
    m3Vec3* angularVelocities;
    float* invMass;
    m3Mat3* invInertiaLocal;
    m3Mat3* inertiaLocal;
    m3Vec3* localCenters;
    float* gravityScales;
    float* linearDamping;
    float* angularDamping;
    uint8_t* types;
    uint8_t* awake;
    float* sleepTimes;
    uint8_t* bulletFlags;
    float* minExtents;
    float* maxExtents;
    uint64_t* userData;
    m3Vec3* bodyForce;
    m3Vec3* bodyTorque;
    uint8_t* bodyEnabled;
    uint8_t* bodyLocks;
    float* bodySleepThreshold;
    uint8_t* bodyCanSleep;
    uint8_t* bodyHasTarget;
    m3Transform* bodyTarget;
    int32_t* bodyShapeHead;

    m3IdPool shapePool;
    int32_t* shapeBody;
    uint8_t* shapeType;
    m3ShapeGeom* shapeGeom;
    float* shapeDensity;
    float* shapeFriction;
    float* shapeRestitution;
    uint8_t* shapeHitEvents;
    m3Vec3* shapeLocalPos;
    m3Quat* shapeLocalRot;
    uint8_t* shapeHasOffset;
    m3Vec3* shapeSurfaceVel;
    uint8_t* shapePreSolve;
    float* shapeRollingResistance;
    uint64_t* shapeCategory;
    uint64_t* shapeMask;
    int32_t* shapeGroup;
    uint64_t* shapeUserData;
    int32_t* shapeNext;
    int32_t* shapeHullIndex;
    int32_t* shapeMeshIndex;
    uint8_t* shapeSensor;
    double* leafBoxes; // the stored proxy box of every tree shape

    // The reference's broadphase tree, mirrored on the host (m3_tree.h);
    // proxyIds[s] is shape s's leaf node (-1: planes).  Shape creation
    // inserts; while the mirror is on (m3World_EnableTreeMirror, or a
    // continuous pass that sweeps against tree shapes) every step replays
    // the device's leaf moves in shape order.  A step taken without it
    // leaves the tree stale: the next use rebuilds it canonically from the
    // current leaves (the reference's m3World_RebuildBroadphase).
    m3Tree tree;
    int32_t* proxyIds;
    int32_t* treeRanks;
    uint8_t treeMirror;
    uint8_t treeStale;
    // The pair table in the reference's layout (the whole capacity, the
    // tail past pairCount holding what earlier steps wrote there) is known
    // while every step since creation or the last restore ran with the
    // mirror; the device keeps it current from the first such step on.
    uint8_t pairTableLive;
    uint8_t pairTableOnDevice;

    // Pre-solve veto (src/solver.c:184-248, include/maul3d/world.h:315-340):
    // the host's callback (never state), this step's vetoed keys (the
    // journal annex), and a replayed journal's recorded vetoes for the next
    // step.
    m3PreSolveFn* preSolveFn;
    void* preSolveContext;
    uint64_t* stepVetoKeys;
    int32_t stepVetoCount;
    uint64_t* replayVetoKeys;
    int32_t replayVetoCount;

    m3IdPool hullPool;
    m3HullData* hullData;
    int32_t* hullRefCounts;

    // Joints (src/world_internal.h joint arrays; spherical on this build).
    m3IdPool jointPool;
    uint8_t* jointType;
    int32_t* jointBodyA;
    int32_t* jointBodyB;
    m3Vec3* jointLocalA;
    m3Vec3* jointLocalB;
    uint8_t* jointCollide;
    m3Vec3* jointImpulse;
    m3Vec3* jointPerpImpulse;
    m3Vec3* jointLimitImpulse;
    m3Vec3* jointAngularImpulse;
    m3Quat* jointFrameQA;
    m3Quat* jointFrameQB;
    uint8_t* jointFlags;
    m3Vec3* jointMotor;
    m3Vec3* jointBreak;
    m3Vec3* jointSpring;
    float* jointTargetScalar;
    m3Quat* jointTargetQ;
    m3Vec3* jointSpringImpulse;
    m3Vec3* jointLimits;
    int32_t* jointNextA;
    int32_t* jointNextB;
    int32_t* bodyJointHead;
    // the device's joint schedule: components of joints linked through
    // non-static bodies, each a run of ascending slots
    int32_t jointCompCount;
    int32_t* jointCompStart;
    int32_t* jointCompList;
    int32_t* jointScratch;

    uint64_t* pairKeys;
    m3Manifold* manifolds;
    int32_t pairCount;
    int32_t pairCapacity;
    uint64_t* sleepingPairKeys;
    int32_t sleepingPairCount;
    uint8_t frozenDirty;
    uint8_t pairsFullQuery;

    // device mirror
    m3gpuWorld* dev;
    // A device call failed (HIP error, a substep barrier that timed out):
    // the world keeps its last host state, steps become no-ops and
    // device-backed reads refuse; m3World_GetDeviceError returns the text.
    uint8_t faulted;
    char faultText[256];
    uint32_t devDirty;  // field groups the host changed since the last upload
    uint32_t hostStale; // field groups the device changed since the last download

    // journal (src/world.c:1348-1400): op, byte count and payload records
    // into the caller's buffer
    uint8_t* journalBuffer;
    int32_t journalCapacity;
    int32_t journalCursor;
    uint8_t journalActive;
    uint8_t journalOverflow;

    // event streams of the last step: 0 none (creation, restore), 1 to be
    // built on the device, 2 built into the arrays below
    int32_t eventsState;
    float eventThreshold; // the hit threshold the last step ran with
    m3ContactEvent* beginEvents;
    m3ContactEvent* endEvents;
    m3ContactEvent* sensorBeginEvents;
    m3ContactEvent* sensorEndEvents;
    m3HitEvent* hitEvents;
    m3BodyMoveEvent* moveEvents;
    int32_t eventCounts[6];
    uint64_t* eventKeys; // 4 x pairCapacity
    m3gpuHit* eventHits;
    m3gpuMove* eventMoves;

    m3Profile profile;
    int32_t lastIslandCount;
    int32_t lastColorCount;
    int32_t lastAwakeCount;
    int32_t lastContactCount;
} m3World;

static m3World* s_worlds[M3_MAX_WORLDS];
static uint16_t s_worldGenerations[M3_MAX_WORLDS];

static m3World* WorldFromId(m3WorldId id)
{
    int32_t index = id.index1 - 1;
    if (index < 0 || index >= M3_MAX_WORLDS || s_worlds[index] == NULL ||
        s_worldGenerations[index] != id.generation)
    {
        return NULL;
    }
    return s_worlds[index];
}

static m3World* WorldFromIndex0(uint16_t index0)
{
    return index0 < M3_MAX_WORLDS ? s_worlds[index0] : NULL;
}

static int FiniteF(float f) { return f == f && f - f == 0.0f; }
static int FiniteV3(m3Vec3 v) { return FiniteF(v.x) && FiniteF(v.y) && FiniteF(v.z); }
static int FiniteQ(m3Quat q) { return FiniteF(q.x) && FiniteF(q.y) && FiniteF(q.z) && FiniteF(q.w); }
static int FiniteD(double d) { return d == d && d - d == 0.0; }
static int FiniteP3(m3Pos3 p) { return FiniteD(p.x) && FiniteD(p.y) && FiniteD(p.z); }

static void HostView(m3World* w, m3gpuHostState* h)
{
    memset(h, 0, sizeof(*h));
    h->maxBody = w->bodyPool.maxIndex;
    h->maxShape = w->shapePool.maxIndex;
    h->maxHull = w->hullPool.maxIndex;
    h->transforms = w->transforms;
    h->linearVelocities = w->linearVelocities;
    h->angularVelocities = w->angularVelocities;
    h->invMass = w->invMass;
    h->invInertiaLocal = w->invInertiaLocal;
    h->inertiaLocal = w->inertiaLocal;
    h->localCenters = w->localCenters;
    h->gravityScales = w->gravityScales;
    h->linearDamping = w->linearDamping;
    h->angularDamping = w->angularDamping;
    h->bodyAlive = w->bodyPool.alive;
    h->types = w->types;
    h->awake = w->awake;
    h->sleepTimes = w->sleepTimes;
    h->bulletFlags = w->bulletFlags;
    h->minExtents = w->minExtents;
    h->maxExtents = w->maxExtents;
    h->bodyForce = w->bodyForce;
    h->bodyTorque = w->bodyTorque;
    h->bodyEnabled = w->bodyEnabled;
    h->bodyLocks = w->bodyLocks;
    h->bodySleepThreshold = w->bodySleepThreshold;
    h->bodyCanSleep = w->bodyCanSleep;
    h->bodyHasTarget = w->bodyHasTarget;
    h->bodyTarget = w->bodyTarget;
    h->bodyShapeHead = w->bodyShapeHead;
    h->shapeAlive = w->shapePool.alive;
    h->shapeBody = w->shapeBody;
    h->shapeType = w->shapeType;
    h->shapeGeom = w->shapeGeom;
    h->shapeFriction = w->shapeFriction;
    h->shapeRestitution = w->shapeRestitution;
    h->shapeRollingResistance = w->shapeRollingResistance;
    h->shapeSensor = w->shapeSensor;
    h->shapeHitEvents = w->shapeHitEvents;
    h->shapePreSolve = w->shapePreSolve;
    h->shapeHasOffset = w->shapeHasOffset;
    h->shapeLocalPos = w->shapeLocalPos;
    h->shapeLocalRot = w->shapeLocalRot;
    h->shapeSurfaceVel = w->shapeSurfaceVel;
    h->shapeCategory = w->shapeCategory;
    h->shapeMask = w->shapeMask;
    h->shapeGroup = w->shapeGroup;
    h->shapeNext = w->shapeNext;
    h->shapeHullIndex = w->shapeHullIndex;
    h->leafBoxes = w->leafBoxes;
    h->hullData = w->hullData;
    h->pairCount = &w->pairCount;
    h->pairKeys = w->pairKeys;
    h->manifolds = w->manifolds;
    h->sleepingPairCount = &w->sleepingPairCount;
    h->sleepingPairKeys = w->sleepingPairKeys;
    h->frozenDirty = &w->frozenDirty;
    h->pairsFullQuery = &w->pairsFullQuery;
    h->stepCount = &w->stepCount;
    h->maxJoint = w->jointPool.maxIndex;
    h->jointCompCount = w->jointCompCount;
    h->jointAlive = w->jointPool.alive;
    h->jointType = w->jointType;
    h->jointBodyA = w->jointBodyA;
    h->jointBodyB = w->jointBodyB;
    h->jointLocalA = w->jointLocalA;
    h->jointLocalB = w->jointLocalB;
    h->jointCollide = w->jointCollide;
    h->jointFrameQA = w->jointFrameQA;
    h->jointFrameQB = w->jointFrameQB;
    h->jointFlags = w->jointFlags;
    h->jointLimits = w->jointLimits;
    h->jointImpulse = w->jointImpulse;
    h->jointPerpImpulse = w->jointPerpImpulse;
    h->jointLimitImpulse = w->jointLimitImpulse;
    h->jointAngularImpulse = w->jointAngularImpulse;
    h->jointSpringImpulse = w->jointSpringImpulse;
    h->jointNextA = w->jointNextA;
    h->jointNextB = w->jointNextB;
    h->bodyJointHead = w->bodyJointHead;
    h->jointCompStart = w->jointCompStart;
    h->jointCompList = w->jointCompList;
}

// The joint schedule: joints that share a non-static body (directly or
// through other joints) form a component; the device runs each component
// serially in ascending slot order and the components in parallel, which
// is the reference's serial slot order restricted to joints that can see
// each other's writes (joints never change static or kinematic velocities:
// those bodies have zero inverse mass and inertia in the joint rows).
static int32_t JointFind(int32_t* parent, int32_t x)
{
    while (parent[x] != x)
    {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

static void BuildJointComponents(m3World* w)
{
    int32_t nj = w->jointPool.maxIndex;
    int32_t nb = w->bodyPool.maxIndex;
    int32_t* parent = w->jointScratch;        // bodies
    int32_t* compOf = w->jointScratch + nb;   // per joint: component root body, -1 none
    for (int32_t b = 0; b < nb; ++b)
    {
        parent[b] = b;
    }
    for (int32_t j = 0; j < nj; ++j)
    {
        compOf[j] = -1;
        if (w->jointPool.alive[j] == 0)
        {
            continue;
        }
        int32_t a = w->jointBodyA[j];
        int32_t b = w->jointBodyB[j];
        int movA = w->types[a] != (uint8_t)M3_STATIC;
        int movB = w->types[b] != (uint8_t)M3_STATIC;
        if (movA && movB)
        {
            int32_t ra = JointFind(parent, a);
            int32_t rb = JointFind(parent, b);
            if (ra != rb)
            {
                parent[ra < rb ? rb : ra] = ra < rb ? ra : rb;
            }
        }
    }
    // Components in order of their first joint; joints ascending within.
    int32_t count = 0;
    int32_t* label = (int32_t*)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(int32_t));
    int32_t* sizes = (int32_t*)calloc((size_t)(nj > 0 ? nj : 1) + 1, sizeof(int32_t));
    if (label == NULL || sizes == NULL)
    {
        fprintf(stderr, "maul3d_amd: out of host memory\n");
        abort();
    }
    for (int32_t b = 0; b < nb; ++b)
    {
        label[b] = -1;
    }
    for (int32_t j = 0; j < nj; ++j)
    {
        if (w->jointPool.alive[j] == 0)
        {
            continue;
        }
        int32_t a = w->jointBodyA[j];
        int32_t body = w->types[a] != (uint8_t)M3_STATIC ? a : w->jointBodyB[j];
        int32_t root = JointFind(parent, body);
        if (label[root] < 0)
        {
            label[root] = count;
            count += 1;
        }
        compOf[j] = label[root];
        sizes[label[root] + 1] += 1;
    }
    for (int32_t k = 0; k < count; ++k)
    {
        sizes[k + 1] += sizes[k];
    }
    for (int32_t k = 0; k <= count; ++k)
    {
        w->jointCompStart[k] = sizes[k];
    }
    for (int32_t j = 0; j < nj; ++j)
    {
        if (compOf[j] >= 0)
        {
            w->jointCompList[sizes[compOf[j]]] = j;
            sizes[compOf[j]] += 1;
        }
    }
    w->jointCompCount = count;
    free(label);
    free(sizes);
}

// A failed device call faults the world instead of ending the process: the
// reference's step is void and never fails (src/solver.c:4322-4330), so the
// nearest contract is a world that stops advancing.  The message is kept for
// m3World_GetDeviceError and printed once.
static void DeviceFail(m3World* w, const char* what)
{
    if (w->faulted == 0)
    {
        snprintf(w->faultText, sizeof(w->faultText), "%s failed: %s", what, m3gpu_last_error());
        fprintf(stderr, "maul3d_amd: %s\n", w->faultText);
    }
    w->faulted = 1;
}

// Bring the host copy up to date with the device (after a step).  Returns 0
// when the world is faulted (the host copy is then the last good one).
static int SyncHost(m3World* w)
{
    if (w->faulted != 0)
    {
        return 0;
    }
    if (w->hostStale == 0 || w->dev == NULL)
    {
        return 1;
    }
    m3gpuHostState h;
    HostView(w, &h);
    if (m3gpu_download(w->dev, &h, w->hostStale) != 0)
    {
        DeviceFail(w, "m3gpu_download");
        return 0;
    }
    w->hostStale = 0;
    return 1;
}

// ---------------------------------------------------------------------------
// Journal (src/world.c:1348-1400): every mutating call of the supported
// surface appends (op, byte count, payload) to the caller's buffer, with the
// reference's op codes (src/world_internal.h:108-192) and record layouts, so
// a journal written by either library replays on the other.
// ---------------------------------------------------------------------------

enum
{
    M3_OP_STEP = 1,
    M3_OP_CREATE_BODY = 2,
    M3_OP_SET_LINEAR_VELOCITY = 4,
    M3_OP_SET_ANGULAR_VELOCITY = 5,
    M3_OP_CREATE_SHAPE = 6,
    M3_OP_CREATE_HULL_SHAPE = 7,
    M3_OP_CREATE_JOINT = 9,
    M3_OP_DESTROY_JOINT = 10,
    M3_OP_APPLY_FORCE = 26,
    M3_OP_APPLY_TORQUE = 27,
    M3_OP_SET_GRAVITY = 39,
    M3_OP_SET_SHAPE_FRICTION = 40,
    M3_OP_SET_SHAPE_RESTITUTION = 41,
    M3_OP_SET_SHAPE_ROLLING = 42,
    M3_OP_SET_SHAPE_DENSITY = 43,
    M3_OP_ENABLE_SLEEPING = 47,
    M3_OP_ENABLE_CONTINUOUS = 48,
    M3_OP_SET_HIT_EVENT_THRESHOLD = 49,
    M3_OP_ENABLE_SHAPE_HIT_EVENTS = 50,
    M3_OP_ENABLE_SHAPE_PRE_SOLVE = 51,
    M3_OP_REBUILD_BROADPHASE = 73,
    M3_OP_STEP_VETOES = 79,
};

#define M3_HULL_MAX_INPUT 256

typedef struct m3OpStep { float dt; int32_t substeps; } m3OpStep;
typedef struct m3OpCreateBody { m3BodyDef def; m3BodyId expected; } m3OpCreateBody;
typedef struct m3OpBodyVec { m3BodyId id; m3Vec3 v; } m3OpBodyVec;
typedef struct m3OpCreateShape // src/world_internal.h:543-551
{
    m3ShapeDef def;
    m3ShapeGeom geom;
    m3BodyId body;
    m3ShapeId expected;
    uint8_t type;
    uint8_t pad[7];
} m3OpCreateShape;
typedef struct m3OpCreateHullShape // src/world_internal.h:574-581
{
    m3ShapeDef def;
    m3BodyId body;
    m3ShapeId expected;
    int32_t count;
    m3Vec3 points[M3_HULL_MAX_INPUT];
} m3OpCreateHullShape;
typedef struct m3OpCreateJoint { m3JointDef def; m3JointId expected; } m3OpCreateJoint;
typedef struct m3OpShapeFlag { m3ShapeId id; int32_t on; } m3OpShapeFlag;
typedef struct m3OpShapeScalar { m3ShapeId id; float value; } m3OpShapeScalar; // src/shape.c:1439-1443
typedef struct m3OpShapeDensity { m3ShapeId id; float value; int32_t updateMass; } m3OpShapeDensity;

static void JournalRecord(m3World* w, int32_t op, const void* payload, int32_t bytes)
{
    if (w->journalActive == 0)
    {
        return;
    }
    const int32_t need = 8 + bytes;
    if (w->journalCursor + need > w->journalCapacity)
    {
        // Overflow latches and stops recording; JournalEnd reports -1.
        w->journalOverflow = 1;
        w->journalActive = 0;
        return;
    }
    uint8_t* out = w->journalBuffer + w->journalCursor;
    memcpy(out, &op, 4);
    memcpy(out + 4, &bytes, 4);
    memcpy(out + 8, payload, (size_t)bytes);
    w->journalCursor += need;
}

// Event streams (src/solver.c:3732-3815, 2991-3026, 4295-4306): the device
// builds the last step's streams into host arrays that stay valid until the
// next step or restore.  Contact ids carry the shapes' generations.
static m3ShapeId ShapeIdOf(const m3World* w, int32_t s)
{
    m3ShapeId id = {s + 1, w->worldIndex0, w->shapePool.generations[s]};
    return id;
}

static void BuildEvents(m3World* w)
{
    if (w->eventsState != 1)
    {
        return;
    }
    w->eventsState = 2;
    memset(w->eventCounts, 0, sizeof(w->eventCounts));
    if (w->dev == NULL)
    {
        return;
    }
    const int32_t np = w->pairCapacity;
    const int32_t nb = w->bodyCapacity;
    if (w->eventKeys == NULL)
    {
        w->eventKeys = (uint64_t*)Zalloc(4 * (size_t)np * sizeof(uint64_t));
        w->eventHits = (m3gpuHit*)Zalloc((size_t)np * sizeof(m3gpuHit));
        w->eventMoves = (m3gpuMove*)Zalloc((size_t)nb * sizeof(m3gpuMove));
        w->beginEvents = (m3ContactEvent*)Zalloc((size_t)np * sizeof(m3ContactEvent));
        w->endEvents = (m3ContactEvent*)Zalloc((size_t)np * sizeof(m3ContactEvent));
        w->sensorBeginEvents = (m3ContactEvent*)Zalloc((size_t)np * sizeof(m3ContactEvent));
        w->sensorEndEvents = (m3ContactEvent*)Zalloc((size_t)np * sizeof(m3ContactEvent));
        w->hitEvents = (m3HitEvent*)Zalloc((size_t)np * sizeof(m3HitEvent));
        w->moveEvents = (m3BodyMoveEvent*)Zalloc((size_t)nb * sizeof(m3BodyMoveEvent));
    }
    m3gpuEvents out;
    memset(&out, 0, sizeof(out));
    out.beginKeys = w->eventKeys;
    out.endKeys = w->eventKeys + np;
    out.sensorBeginKeys = w->eventKeys + 2 * (size_t)np;
    out.sensorEndKeys = w->eventKeys + 3 * (size_t)np;
    out.hits = w->eventHits;
    out.moves = w->eventMoves;
    if (w->faulted != 0 || m3gpu_events(w->dev, w->eventThreshold, &out) != 0)
    {
        if (w->faulted == 0)
        {
            DeviceFail(w, "m3gpu_events");
        }
        return;
    }
    const uint64_t* keys[4] = {out.beginKeys, out.endKeys, out.sensorBeginKeys, out.sensorEndKeys};
    m3ContactEvent* lists[4] = {w->beginEvents, w->endEvents, w->sensorBeginEvents,
                                w->sensorEndEvents};
    for (int k = 0; k < 4; ++k)
    {
        for (int32_t i = 0; i < out.counts[k]; ++i)
        {
            lists[k][i].shapeA = ShapeIdOf(w, (int32_t)(keys[k][i] >> 32));
            lists[k][i].shapeB = ShapeIdOf(w, (int32_t)(keys[k][i] & 0xFFFFFFFFu));
        }
        w->eventCounts[k] = out.counts[k];
    }
    for (int32_t i = 0; i < out.counts[4]; ++i)
    {
        const m3gpuHit* g = &out.hits[i];
        m3HitEvent* e = &w->hitEvents[i];
        memset(e, 0, sizeof(*e));
        e->shapeA = ShapeIdOf(w, (int32_t)(g->key >> 32));
        e->shapeB = ShapeIdOf(w, (int32_t)(g->key & 0xFFFFFFFFu));
        e->point.x = g->point[0];
        e->point.y = g->point[1];
        e->point.z = g->point[2];
        e->normal = m3V3(g->normal[0], g->normal[1], g->normal[2]);
        e->approachSpeed = g->approachSpeed;
    }
    w->eventCounts[4] = out.counts[4];
    for (int32_t i = 0; i < out.counts[5]; ++i)
    {
        const m3gpuMove* g = &out.moves[i];
        m3BodyMoveEvent* e = &w->moveEvents[i];
        memset(e, 0, sizeof(*e));
        e->body.index1 = g->body + 1;
        e->body.world0 = w->worldIndex0;
        e->body.generation = w->bodyPool.generations[g->body];
        e->transform.p.x = g->p[0];
        e->transform.p.y = g->p[1];
        e->transform.p.z = g->p[2];
        e->transform.q.x = g->q[0];
        e->transform.q.y = g->q[1];
        e->transform.q.z = g->q[2];
        e->transform.q.w = g->q[3];
        e->fellAsleep = g->fellAsleep != 0;
    }
    w->eventCounts[5] = out.counts[5];
}

static void FreeEvents(m3World* w)
{
    free(w->eventKeys);
    free(w->eventHits);
    free(w->eventMoves);
    free(w->beginEvents);
    free(w->endEvents);
    free(w->sensorBeginEvents);
    free(w->sensorEndEvents);
    free(w->hitEvents);
    free(w->moveEvents);
}

// Every mutation goes through here: build the last step's event streams
// while the device still holds that step's state, read the device state back,
// then mark the groups the mutation touches for upload before the next step.
static void Touch(m3World* w, uint32_t groups)
{
    BuildEvents(w);
    SyncHost(w);
    w->devDirty |= groups;
}

// ---------------------------------------------------------------------------
// Defaults and world lifetime (src/world.c:64-480)
// ---------------------------------------------------------------------------

int m3GetVersion(void) { return 1 * 10000 + 22 * 100 + 0; }

uint64_t m3Hash64(uint64_t h, const void* bytes, int32_t count)
{
    return m3HashBytes(h, bytes, count);
}

m3WorldDef m3DefaultWorldDef(void)
{
    m3WorldDef def;
    memset(&def, 0, sizeof(def));
    def.gravity = m3V3(0.0f, -10.0f, 0.0f);
    def.bodyCapacity = 1024;
    def.shapeCapacity = 2048;
    def.meshCapacity = 4;
    def.jointCapacity = 64;
    def.voxelCapacity = 4;
    def.characterCapacity = 4;
    def.vehicleCapacity = 2;
    def.softBodyCapacity = 2;
    def.workerCount = 1;
    def.contactHertz = M3_CONTACT_HERTZ_DEFAULT;
    def.contactDampingRatio = M3_CONTACT_DAMPING_RATIO_DEFAULT;
    def.contactPushMaxSpeed = M3_CONTACT_PUSH_MAX_SPEED_DEFAULT;
    def.restitutionThreshold = M3_RESTITUTION_THRESHOLD_DEFAULT;
    def.maximumLinearSpeed = M3_MAX_LINEAR_SPEED_DEFAULT;
    def.enableSleeping = 1;
    def.enableContinuous = 1;
    def.hitEventThreshold = M3_HIT_EVENT_THRESHOLD_DEFAULT;
    def.internalValue = M3_WORLD_COOKIE;
    return def;
}

#define ALLOC(field, n, type) (field) = (type*)Zalloc((size_t)(n) * sizeof(type))

m3WorldId m3CreateWorld(const m3WorldDef* def)
{
    m3WorldId nullId = {0, 0};
    if (def == NULL || def->internalValue != M3_WORLD_COOKIE || def->bodyCapacity <= 0 ||
        def->shapeCapacity <= 0 || def->meshCapacity <= 0 || def->jointCapacity <= 0 ||
        def->voxelCapacity <= 0 || def->characterCapacity <= 0 || def->vehicleCapacity <= 0 ||
        def->softBodyCapacity <= 0 || def->workerCount <= 0 ||
        (def->enqueueTask == NULL) != (def->finishTask == NULL) || !FiniteV3(def->gravity) ||
        !FiniteF(def->contactHertz) || def->contactHertz <= 0.0f ||
        !FiniteF(def->contactDampingRatio) || def->contactDampingRatio <= 0.0f ||
        !FiniteF(def->contactPushMaxSpeed) || def->contactPushMaxSpeed <= 0.0f ||
        !FiniteF(def->restitutionThreshold) || def->restitutionThreshold < 0.0f ||
        !FiniteF(def->maximumLinearSpeed) || def->maximumLinearSpeed <= 0.0f ||
        !FiniteF(def->hitEventThreshold) || def->hitEventThreshold < 0.0f)
    {
        return nullId;
    }
    if (def->shapeCapacity > M3_DEVICE_MAX_SHAPES)
    {
        // The device pair list packs two shape slots into 32-bit keys:
        // refuse at creation (the reference's null id for a def it cannot
        // serve) rather than at the first step.
        return nullId;
    }
    int32_t slot = -1;
    for (int32_t i = 0; i < M3_MAX_WORLDS; ++i)
    {
        if (s_worlds[i] == NULL)
        {
            slot = i;
            break;
        }
    }
    if (slot < 0)
    {
        return nullId;
    }
    m3World* w = (m3World*)Zalloc(sizeof(m3World));
    w->gravity = def->gravity;
    w->contactHertz = def->contactHertz;
    w->contactDampingRatio = def->contactDampingRatio;
    w->contactPushMaxSpeed = def->contactPushMaxSpeed;
    w->restitutionThreshold = def->restitutionThreshold;
    w->maximumLinearSpeed = def->maximumLinearSpeed;
    w->maximumAngularSpeed = M3_MAX_ANGULAR_SPEED_DEFAULT;
    w->sleepEnabled = def->enableSleeping != 0 ? 1 : 0;
    w->continuousEnabled = def->enableContinuous != 0 ? 1 : 0;
    w->hitEventThreshold = def->hitEventThreshold;
    w->bodyCapacity = def->bodyCapacity;
    w->shapeCapacity = def->shapeCapacity;
    w->jointCapacity = def->jointCapacity;
    w->meshCapacity = def->meshCapacity;
    w->voxelCapacity = def->voxelCapacity;
    w->characterCapacity = def->characterCapacity;
    w->vehicleCapacity = def->vehicleCapacity;
    w->softBodyCapacity = def->softBodyCapacity;
    w->generation = s_worldGenerations[slot];
    w->worldIndex0 = (uint16_t)slot;

    int32_t cap = def->bodyCapacity;
    w->bodyPool = PoolCreate(cap);
    ALLOC(w->transforms, cap, m3Transform);
    ALLOC(w->linearVelocities, cap, m3Vec3);
    ALLOC(w->angularVelocities, cap, m3Vec3);
    ALLOC(w->invMass, cap, float);
    ALLOC(w->invInertiaLocal, cap, m3Mat3);
    ALLOC(w->inertiaLocal, cap, m3Mat3);
    ALLOC(w->localCenters, cap, m3Vec3);
    ALLOC(w->gravityScales, cap, float);
    ALLOC(w->linearDamping, cap, float);
    ALLOC(w->angularDamping, cap, float);
    ALLOC(w->types, cap, uint8_t);
    ALLOC(w->awake, cap, uint8_t);
    ALLOC(w->sleepTimes, cap, float);
    ALLOC(w->bulletFlags, cap, uint8_t);
    ALLOC(w->minExtents, cap, float);
    ALLOC(w->maxExtents, cap, float);
    ALLOC(w->userData, cap, uint64_t);
    ALLOC(w->bodyForce, cap, m3Vec3);
    ALLOC(w->bodyTorque, cap, m3Vec3);
    ALLOC(w->bodyEnabled, cap, uint8_t);
    ALLOC(w->bodyLocks, cap, uint8_t);
    ALLOC(w->bodySleepThreshold, cap, float);
    ALLOC(w->bodyCanSleep, cap, uint8_t);
    ALLOC(w->bodyHasTarget, cap, uint8_t);
    ALLOC(w->bodyTarget, cap, m3Transform);
    ALLOC(w->bodyShapeHead, cap, int32_t);
    ALLOC(w->bodyJointHead, cap, int32_t);
    for (int32_t i = 0; i < cap; ++i)
    {
        w->bodyShapeHead[i] = -1;
        w->bodyJointHead[i] = -1;
    }

    int32_t shapeCap = def->shapeCapacity;
    w->shapePool = PoolCreate(shapeCap);
    ALLOC(w->shapeBody, shapeCap, int32_t);
    ALLOC(w->shapeType, shapeCap, uint8_t);
    ALLOC(w->shapeGeom, shapeCap, m3ShapeGeom);
    ALLOC(w->shapeDensity, shapeCap, float);
    ALLOC(w->shapeFriction, shapeCap, float);
    ALLOC(w->shapeRestitution, shapeCap, float);
    ALLOC(w->shapeHitEvents, shapeCap, uint8_t);
    ALLOC(w->shapeLocalPos, shapeCap, m3Vec3);
    ALLOC(w->shapeLocalRot, shapeCap, m3Quat);
    ALLOC(w->shapeHasOffset, shapeCap, uint8_t);
    ALLOC(w->shapeSurfaceVel, shapeCap, m3Vec3);
    ALLOC(w->shapePreSolve, shapeCap, uint8_t);
    ALLOC(w->shapeRollingResistance, shapeCap, float);
    ALLOC(w->shapeCategory, shapeCap, uint64_t);
    ALLOC(w->shapeMask, shapeCap, uint64_t);
    ALLOC(w->shapeGroup, shapeCap, int32_t);
    ALLOC(w->shapeUserData, shapeCap, uint64_t);
    ALLOC(w->shapeNext, shapeCap, int32_t);
    ALLOC(w->shapeHullIndex, shapeCap, int32_t);
    ALLOC(w->shapeMeshIndex, shapeCap, int32_t);
    ALLOC(w->shapeSensor, shapeCap, uint8_t);
    ALLOC(w->leafBoxes, (size_t)shapeCap * 6, double);
    for (int32_t i = 0; i < shapeCap; ++i)
    {
        w->shapeBody[i] = -1;
        w->shapeNext[i] = -1;
        w->shapeHullIndex[i] = -1;
        w->shapeMeshIndex[i] = -1;
    }
    if (!m3TreeCreate(&w->tree, 2 * shapeCap))
    {
        fprintf(stderr, "maul3d_amd: out of host memory\n");
        abort();
    }
    ALLOC(w->proxyIds, shapeCap, int32_t);
    ALLOC(w->treeRanks, shapeCap, int32_t);
    w->pairTableLive = 1; // all zero, as the reference's freshly allocated table
    for (int32_t i = 0; i < shapeCap; ++i)
    {
        w->proxyIds[i] = M3_TREE_NULL;
    }
    w->hullPool = PoolCreate(shapeCap);
    ALLOC(w->hullData, shapeCap, m3HullData);
    ALLOC(w->hullRefCounts, shapeCap, int32_t);
    int32_t jointCap = def->jointCapacity;
    w->jointPool = PoolCreate(jointCap);
    ALLOC(w->jointType, jointCap, uint8_t);
    ALLOC(w->jointBodyA, jointCap, int32_t);
    ALLOC(w->jointBodyB, jointCap, int32_t);
    ALLOC(w->jointLocalA, jointCap, m3Vec3);
    ALLOC(w->jointLocalB, jointCap, m3Vec3);
    ALLOC(w->jointCollide, jointCap, uint8_t);
    ALLOC(w->jointImpulse, jointCap, m3Vec3);
    ALLOC(w->jointPerpImpulse, jointCap, m3Vec3);
    ALLOC(w->jointLimitImpulse, jointCap, m3Vec3);
    ALLOC(w->jointAngularImpulse, jointCap, m3Vec3);
    ALLOC(w->jointFrameQA, jointCap, m3Quat);
    ALLOC(w->jointFrameQB, jointCap, m3Quat);
    ALLOC(w->jointFlags, jointCap, uint8_t);
    ALLOC(w->jointMotor, jointCap, m3Vec3);
    ALLOC(w->jointBreak, jointCap, m3Vec3);
    ALLOC(w->jointSpring, jointCap, m3Vec3);
    ALLOC(w->jointTargetScalar, jointCap, float);
    ALLOC(w->jointTargetQ, jointCap, m3Quat);
    ALLOC(w->jointSpringImpulse, jointCap, m3Vec3);
    ALLOC(w->jointLimits, jointCap, m3Vec3);
    ALLOC(w->jointNextA, jointCap, int32_t);
    ALLOC(w->jointNextB, jointCap, int32_t);
    ALLOC(w->jointCompStart, (size_t)jointCap + 1, int32_t);
    ALLOC(w->jointCompList, jointCap, int32_t);
    ALLOC(w->jointScratch, (size_t)jointCap + (size_t)cap, int32_t);
    for (int32_t j = 0; j < jointCap; ++j)
    {
        w->jointBodyA[j] = -1;
        w->jointBodyB[j] = -1;
        w->jointNextA[j] = -1;
        w->jointNextB[j] = -1;
    }

    w->pairCapacity = 8 * shapeCap;
    ALLOC(w->stepVetoKeys, w->pairCapacity, uint64_t);
    ALLOC(w->replayVetoKeys, w->pairCapacity, uint64_t);
    ALLOC(w->pairKeys, w->pairCapacity, uint64_t);
    ALLOC(w->manifolds, w->pairCapacity, m3Manifold);
    ALLOC(w->sleepingPairKeys, w->pairCapacity, uint64_t);
    w->devDirty = M3GPU_FIELDS_ALL;

    s_worlds[slot] = w;
    m3WorldId id = {slot + 1, w->generation};
    return id;
}

void m3DestroyWorld(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL)
    {
        return;
    }
    int32_t slot = w->worldIndex0;
    if (w->dev != NULL)
    {
        m3gpu_world_destroy(w->dev);
    }
    void* arrays[] = {
        w->transforms, w->linearVelocities, w->angularVelocities, w->invMass,
        w->invInertiaLocal, w->inertiaLocal, w->localCenters, w->gravityScales,
        w->linearDamping, w->angularDamping, w->types, w->awake, w->sleepTimes,
        w->bulletFlags, w->minExtents, w->maxExtents, w->userData, w->bodyForce,
        w->bodyTorque, w->bodyEnabled, w->bodyLocks, w->bodySleepThreshold,
        w->bodyCanSleep, w->bodyHasTarget, w->bodyTarget, w->bodyShapeHead,
        w->shapeBody, w->shapeType, w->shapeGeom, w->shapeDensity, w->shapeFriction,
        w->shapeRestitution, w->shapeHitEvents, w->shapeLocalPos, w->shapeLocalRot,
        w->shapeHasOffset, w->shapeSurfaceVel, w->shapePreSolve, w->shapeRollingResistance,
        w->shapeCategory, w->shapeMask, w->shapeGroup, w->shapeUserData, w->shapeNext,
        w->shapeHullIndex, w->shapeMeshIndex, w->shapeSensor, w->leafBoxes, w->hullData,
        w->hullRefCounts, w->pairKeys, w->manifolds, w->sleepingPairKeys,
        w->jointType, w->jointBodyA, w->jointBodyB, w->jointLocalA, w->jointLocalB,
        w->jointCollide, w->jointImpulse, w->jointPerpImpulse, w->jointLimitImpulse,
        w->jointAngularImpulse, w->jointFrameQA, w->jointFrameQB, w->jointFlags, w->jointMotor,
        w->jointBreak, w->jointSpring, w->jointTargetScalar, w->jointTargetQ,
        w->jointSpringImpulse, w->jointLimits, w->jointNextA, w->jointNextB, w->bodyJointHead,
        w->jointCompStart, w->jointCompList, w->jointScratch, w->proxyIds, w->treeRanks,
        w->stepVetoKeys, w->replayVetoKeys,
    };
    for (size_t i = 0; i < sizeof(arrays) / sizeof(arrays[0]); ++i)
    {
        free(arrays[i]);
    }
    FreeEvents(w);
    m3TreeDestroy(&w->tree);
    PoolDestroy(&w->bodyPool);
    PoolDestroy(&w->shapePool);
    PoolDestroy(&w->hullPool);
    PoolDestroy(&w->jointPool);
    free(w);
    s_worlds[slot] = NULL;
    s_worldGenerations[slot] += 1;
}

bool m3World_IsValid(m3WorldId worldId) { return WorldFromId(worldId) != NULL; }

void m3World_SetGravity(m3WorldId worldId, m3Vec3 gravity)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || !FiniteV3(gravity))
    {
        return;
    }
    JournalRecord(w, M3_OP_SET_GRAVITY, &gravity, (int32_t)sizeof(gravity));
    w->gravity = gravity;
}

m3Vec3 m3World_GetGravity(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    return w != NULL ? w->gravity : m3V3(0.0f, 0.0f, 0.0f);
}

void m3World_EnableContinuous(m3WorldId worldId, bool flag)
{
    m3World* w = WorldFromId(worldId);
    if (w != NULL)
    {
        int32_t on = flag ? 1 : 0;
        JournalRecord(w, M3_OP_ENABLE_CONTINUOUS, &on, (int32_t)sizeof(on));
        w->continuousEnabled = flag ? 1 : 0;
    }
}

void m3World_EnableSleeping(m3WorldId worldId, bool flag)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL)
    {
        return;
    }
    int32_t on = flag ? 1 : 0;
    JournalRecord(w, M3_OP_ENABLE_SLEEPING, &on, (int32_t)sizeof(on));
    w->sleepEnabled = flag ? 1 : 0;
    if (!flag)
    {
        // Turning sleep off wakes everyone (the reference toggle).
        Touch(w, M3GPU_FIELDS_DYNAMIC);
        for (int32_t i = 0; i < w->bodyPool.maxIndex; ++i)
        {
            if (w->bodyPool.alive[i] != 0 && w->types[i] == M3_DYNAMIC && w->awake[i] == 0)
            {
                w->awake[i] = 1;
                w->sleepTimes[i] = 0.0f;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Bodies (src/body.c:35-101)
// ---------------------------------------------------------------------------

m3BodyDef m3DefaultBodyDef(void)
{
    m3BodyDef def;
    memset(&def, 0, sizeof(def));
    def.type = m3_staticBody;
    def.rotation = m3IdentityQuat();
    def.gravityScale = 1.0f;
    def.internalValue = M3_BODY_COOKIE;
    return def;
}

static int32_t BodySlot(const m3World* w, m3BodyId id)
{
    int32_t index = id.index1 - 1;
    if (w == NULL || id.world0 != w->worldIndex0 || !PoolValid(&w->bodyPool, index, id.generation))
    {
        return -1;
    }
    return index;
}

static m3World* ResolveBody(m3BodyId id, int32_t* out)
{
    m3World* w = WorldFromIndex0(id.world0);
    int32_t index = w != NULL ? BodySlot(w, id) : -1;
    *out = index;
    return index >= 0 ? w : NULL;
}

m3BodyId m3CreateBody(m3WorldId worldId, const m3BodyDef* def)
{
    m3BodyId nullId = {0, 0, 0};
    m3World* w = WorldFromId(worldId);
    if (w == NULL || def == NULL || def->internalValue != M3_BODY_COOKIE ||
        (def->type != m3_staticBody && def->type != m3_kinematicBody &&
         def->type != m3_dynamicBody))
    {
        return nullId;
    }
    m3real qq = def->rotation.x * def->rotation.x + def->rotation.y * def->rotation.y +
                def->rotation.z * def->rotation.z + def->rotation.w * def->rotation.w;
    if (!FiniteP3(def->position) || !FiniteQ(def->rotation) || !FiniteV3(def->linearVelocity) ||
        !FiniteV3(def->angularVelocity) || !FiniteF(def->gravityScale) ||
        !FiniteF(def->linearDamping) || !FiniteF(def->angularDamping) ||
        def->linearDamping < 0.0f || def->angularDamping < 0.0f || !(qq > 0.98f) ||
        !(qq < 1.02f))
    {
        return nullId;
    }
    Touch(w, M3GPU_FIELDS_ALL);
    int32_t i = PoolAlloc(&w->bodyPool);
    if (i < 0)
    {
        return nullId;
    }
    w->transforms[i].p = def->position;
    w->transforms[i].q = m3NormalizeQuat(def->rotation);
    w->linearVelocities[i] = def->linearVelocity;
    w->angularVelocities[i] = def->angularVelocity;
    w->invMass[i] = def->type == m3_dynamicBody ? 1.0f : 0.0f;
    w->invInertiaLocal[i] = m3ZeroMat3();
    w->inertiaLocal[i] = m3ZeroMat3();
    w->awake[i] = 1;
    w->sleepTimes[i] = 0.0f;
    w->minExtents[i] = 1.0e30f;
    w->maxExtents[i] = 0.0f;
    w->localCenters[i] = m3V3(0.0f, 0.0f, 0.0f);
    w->gravityScales[i] = def->gravityScale;
    w->linearDamping[i] = def->linearDamping;
    w->angularDamping[i] = def->angularDamping;
    w->types[i] = (uint8_t)def->type;
    w->bulletFlags[i] = def->isBullet ? 1 : 0;
    w->userData[i] = def->userData;
    w->bodyForce[i] = m3V3(0.0f, 0.0f, 0.0f);
    w->bodyTorque[i] = m3V3(0.0f, 0.0f, 0.0f);
    w->bodyEnabled[i] = 1;
    w->bodyLocks[i] = 0;
    w->bodySleepThreshold[i] = M3_SLEEP_VELOCITY_DEFAULT;
    w->bodyCanSleep[i] = 1;
    w->bodyHasTarget[i] = 0;
    memset(&w->bodyTarget[i], 0, sizeof(m3Transform));
    w->bodyTarget[i].q.w = 1.0f;
    w->bodyShapeHead[i] = -1;
    m3BodyId id = {i + 1, w->worldIndex0, w->bodyPool.generations[i]};
    if (w->journalActive != 0)
    {
        m3OpCreateBody rec;
        memset(&rec, 0, sizeof(rec));
        rec.def = *def;
        rec.expected = id;
        JournalRecord(w, M3_OP_CREATE_BODY, &rec, (int32_t)sizeof(rec));
    }
    return id;
}

bool m3Body_IsValid(m3BodyId bodyId)
{
    int32_t i;
    return ResolveBody(bodyId, &i) != NULL;
}

m3Pos3 m3Body_GetPosition(m3BodyId bodyId)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    m3Pos3 zero = {0.0, 0.0, 0.0};
    if (w == NULL)
    {
        return zero;
    }
    SyncHost(w);
    return w->transforms[i].p;
}

m3Quat m3Body_GetRotation(m3BodyId bodyId)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    if (w == NULL)
    {
        return m3IdentityQuat();
    }
    SyncHost(w);
    return w->transforms[i].q;
}

m3Vec3 m3Body_GetLinearVelocity(m3BodyId bodyId)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    if (w == NULL)
    {
        return m3V3(0.0f, 0.0f, 0.0f);
    }
    SyncHost(w);
    return w->linearVelocities[i];
}

m3Vec3 m3Body_GetAngularVelocity(m3BodyId bodyId)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    if (w == NULL)
    {
        return m3V3(0.0f, 0.0f, 0.0f);
    }
    SyncHost(w);
    return w->angularVelocities[i];
}

// src/body.c:406-418: a commanded velocity always wakes.
void m3Body_SetLinearVelocity(m3BodyId bodyId, m3Vec3 velocity)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    if (w == NULL || !FiniteV3(velocity))
    {
        return;
    }
    if (w->journalActive != 0)
    {
        m3OpBodyVec rec;
        memset(&rec, 0, sizeof(rec));
        rec.id = bodyId;
        rec.v = velocity;
        JournalRecord(w, M3_OP_SET_LINEAR_VELOCITY, &rec, (int32_t)sizeof(rec));
    }
    Touch(w, M3GPU_FIELDS_DYNAMIC);
    w->awake[i] = 1;
    w->sleepTimes[i] = 0.0f;
    w->linearVelocities[i] = velocity;
}

void m3Body_SetAngularVelocity(m3BodyId bodyId, m3Vec3 velocity)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    if (w == NULL || !FiniteV3(velocity))
    {
        return;
    }
    if (w->journalActive != 0)
    {
        m3OpBodyVec rec;
        memset(&rec, 0, sizeof(rec));
        rec.id = bodyId;
        rec.v = velocity;
        JournalRecord(w, M3_OP_SET_ANGULAR_VELOCITY, &rec, (int32_t)sizeof(rec));
    }
    Touch(w, M3GPU_FIELDS_DYNAMIC);
    w->awake[i] = 1;
    w->sleepTimes[i] = 0.0f;
    w->angularVelocities[i] = velocity;
}

bool m3Body_IsAwake(m3BodyId bodyId)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    if (w == NULL)
    {
        return false;
    }
    SyncHost(w);
    return w->awake[i] != 0;
}

// src/body.c:337-361: forces accumulate for one step and wake a body that
// can respond.
static void ApplyForceLike(m3BodyId bodyId, m3Vec3 value, int torque)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    if (w == NULL || !FiniteV3(value))
    {
        return;
    }
    if (!(w->types[i] == M3_DYNAMIC && w->invMass[i] > 0.0f))
    {
        return;
    }
    if (w->journalActive != 0)
    {
        m3OpBodyVec rec;
        memset(&rec, 0, sizeof(rec));
        rec.id = bodyId;
        rec.v = value;
        JournalRecord(w, torque ? M3_OP_APPLY_TORQUE : M3_OP_APPLY_FORCE, &rec, (int32_t)sizeof(rec));
    }
    Touch(w, M3GPU_FIELDS_DYNAMIC);
    if (torque)
    {
        w->bodyTorque[i] = m3Add3(w->bodyTorque[i], value);
    }
    else
    {
        w->bodyForce[i] = m3Add3(w->bodyForce[i], value);
    }
    if (value.x != 0.0f || value.y != 0.0f || value.z != 0.0f)
    {
        w->awake[i] = 1;
        w->sleepTimes[i] = 0.0f;
    }
}

void m3Body_ApplyForce(m3BodyId bodyId, m3Vec3 force) { ApplyForceLike(bodyId, force, 0); }
void m3Body_ApplyTorque(m3BodyId bodyId, m3Vec3 torque) { ApplyForceLike(bodyId, torque, 1); }

// ---------------------------------------------------------------------------
// Hulls: the canonical box (src/hull.c:14-75); the half-edge adjacency and
// point-cloud hulls are in hull_build.c
// ---------------------------------------------------------------------------

static void BuildBoxHull(m3HullData* out, m3Vec3 he)
{
    memset(out, 0, sizeof(*out));
    m3real hx = he.x;
    m3real hy = he.y;
    m3real hz = he.z;
    // Canonical vertex order: bit0 of the index pattern walks -x,+x around
    // the -z face then the +z face.
    const m3real sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const m3real sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const m3real sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int32_t v = 0; v < 8; ++v)
    {
        out->vertices[v] = m3V3(sx[v] < 0 ? -hx : hx, sy[v] < 0 ? -hy : hy, sz[v] < 0 ? -hz : hz);
    }
    const m3real nrm[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    const m3real off[6] = {hx, hx, hy, hy, hz, hz};
    // Counter-clockwise loops seen from outside: +x, -x, +y, -y, +z, -z.
    const uint8_t loops[6][4] = {{1, 2, 6, 5}, {0, 4, 7, 3}, {2, 3, 7, 6},
                                 {0, 1, 5, 4}, {4, 5, 6, 7}, {0, 3, 2, 1}};
    out->vertexCount = 8;
    out->faceCount = 6;
    out->indexCount = 24;
    for (int32_t f = 0; f < 6; ++f)
    {
        out->faceNormals[f] = m3V3(nrm[f][0], nrm[f][1], nrm[f][2]);
        out->faceOffsets[f] = off[f];
        out->faceVertCounts[f] = 4;
        out->faceVertStart[f] = (uint16_t)(4 * f);
        for (int32_t k = 0; k < 4; ++k)
        {
            out->faceIndices[4 * f + k] = loops[f][k];
        }
    }
    // Closed-form unit-density box: m = 8 hx hy hz, I = m/3 diag(...).
    float m = 8.0f * hx * hy * hz;
    out->unitMass = m;
    out->unitCom = m3V3(0.0f, 0.0f, 0.0f);
    out->unitInertiaCom = m3ZeroMat3();
    out->unitInertiaCom.cx.x = (m / 3.0f) * (hy * hy + hz * hz);
    out->unitInertiaCom.cy.y = (m / 3.0f) * (hx * hx + hz * hz);
    out->unitInertiaCom.cz.z = (m / 3.0f) * (hx * hx + hy * hy);
    m3HostHullHalfEdges(out);
}

// Content dedupe in ascending slot order (src/hull.c:141-163).
static int32_t InternHull(m3World* w, const m3HullData* data)
{
    for (int32_t i = 0; i < w->hullPool.maxIndex; ++i)
    {
        if (w->hullPool.alive[i] != 0 && memcmp(&w->hullData[i], data, sizeof(m3HullData)) == 0)
        {
            w->hullRefCounts[i] += 1;
            return i;
        }
    }
    int32_t index = PoolAlloc(&w->hullPool);
    if (index < 0)
    {
        return -1;
    }
    w->hullData[index] = *data;
    w->hullRefCounts[index] = 1;
    return index;
}

// ---------------------------------------------------------------------------
// Mass properties (src/shape.c:49-301)
// ---------------------------------------------------------------------------

static m3Mat3 InvertSymmetric(m3Mat3 m)
{
    m3real a = m.cx.x;
    m3real b = m.cy.x;
    m3real c = m.cz.x;
    m3real d = m.cy.y;
    m3real e = m.cz.y;
    m3real f = m.cz.z;
    m3real co00 = d * f - e * e;
    m3real co01 = c * e - b * f;
    m3real co02 = b * e - c * d;
    m3real det = a * co00 + b * co01 + c * co02;
    if (!(det > 0.0f))
    {
        return m3ZeroMat3();
    }
    m3real inv = 1.0f / det;
    m3Mat3 r;
    r.cx = m3V3(co00 * inv, co01 * inv, co02 * inv);
    r.cy = m3V3(co01 * inv, (a * f - c * c) * inv, (b * c - a * e) * inv);
    r.cz = m3V3(co02 * inv, (b * c - a * e) * inv, (a * d - b * b) * inv);
    return r;
}

// c' = p + R c, I' = R I R^T for a shape on a compound offset.
static void ComposeMassProps(const m3World* w, int32_t s, m3Vec3* com, m3Mat3* I)
{
    if (w->shapeHasOffset[s] == 0)
    {
        return;
    }
    m3Quat q = w->shapeLocalRot[s];
    *com = m3Add3(w->shapeLocalPos[s], m3RotateVec3(q, *com));
    m3Mat3 r;
    r.cx = m3RotateVec3(q, m3V3(1.0f, 0.0f, 0.0f));
    r.cy = m3RotateVec3(q, m3V3(0.0f, 1.0f, 0.0f));
    r.cz = m3RotateVec3(q, m3V3(0.0f, 0.0f, 1.0f));
    m3Mat3 t;
    t.cx = m3V3(I->cx.x * r.cx.x + I->cy.x * r.cx.y + I->cz.x * r.cx.z,
                I->cx.y * r.cx.x + I->cy.y * r.cx.y + I->cz.y * r.cx.z,
                I->cx.z * r.cx.x + I->cy.z * r.cx.y + I->cz.z * r.cx.z);
    t.cy = m3V3(I->cx.x * r.cy.x + I->cy.x * r.cy.y + I->cz.x * r.cy.z,
                I->cx.y * r.cy.x + I->cy.y * r.cy.y + I->cz.y * r.cy.z,
                I->cx.z * r.cy.x + I->cy.z * r.cy.y + I->cz.z * r.cy.z);
    t.cz = m3V3(I->cx.x * r.cz.x + I->cy.x * r.cz.y + I->cz.x * r.cz.z,
                I->cx.y * r.cz.x + I->cy.y * r.cz.y + I->cz.y * r.cz.z,
                I->cx.z * r.cz.x + I->cy.z * r.cz.y + I->cz.z * r.cz.z);
    I->cx = m3V3(r.cx.x * t.cx.x + r.cy.x * t.cx.y + r.cz.x * t.cx.z,
                 r.cx.y * t.cx.x + r.cy.y * t.cx.y + r.cz.y * t.cx.z,
                 r.cx.z * t.cx.x + r.cy.z * t.cx.y + r.cz.z * t.cx.z);
    I->cy = m3V3(r.cx.x * t.cy.x + r.cy.x * t.cy.y + r.cz.x * t.cy.z,
                 r.cx.y * t.cy.x + r.cy.y * t.cy.y + r.cz.y * t.cy.z,
                 r.cx.z * t.cy.x + r.cy.z * t.cy.y + r.cz.z * t.cy.z);
    I->cz = m3V3(r.cx.x * t.cz.x + r.cy.x * t.cz.y + r.cz.x * t.cz.z,
                 r.cx.y * t.cz.x + r.cy.y * t.cz.y + r.cz.y * t.cz.z,
                 r.cx.z * t.cz.x + r.cy.z * t.cz.y + r.cz.z * t.cz.z);
}

static int ShapeMassProps(const m3World* w, int32_t s, float* mOut, m3Vec3* cOut, m3Mat3* iOut)
{
    uint8_t type = w->shapeType[s];
    if (type == M3_SPHERE)
    {
        float r = w->shapeGeom[s].s;
        float m = w->shapeDensity[s] * (4.0f / 3.0f) * M3_PI * r * r * r;
        float ic = 0.4f * m * r * r;
        *mOut = m;
        *cOut = w->shapeGeom[s].v;
        *iOut = m3ZeroMat3();
        iOut->cx.x = ic;
        iOut->cy.y = ic;
        iOut->cz.z = ic;
        return 1;
    }
    if (type == M3_CAPSULE)
    {
        // Cylinder plus two hemispheres about the segment midpoint:
        // I = Iperp * Id + (Iaxial - Iperp) * (u u^T).
        m3Vec3 p1 = w->shapeGeom[s].v;
        m3Vec3 p2 = w->shapeGeom[s].v2;
        float r = w->shapeGeom[s].s;
        m3Vec3 axis = m3Sub3(p2, p1);
        float length = sqrtf(m3Dot3(axis, axis));
        m3Vec3 u = length > 0.0f ? m3MulSV3(1.0f / length, axis) : m3V3(1.0f, 0.0f, 0.0f);
        float density = w->shapeDensity[s];
        float mCyl = density * M3_PI * r * r * length;
        float mSph = density * (4.0f / 3.0f) * M3_PI * r * r * r;
        float axial = 0.5f * mCyl * r * r + 0.4f * mSph * r * r;
        float perp = mCyl * (length * length / 12.0f + 0.25f * r * r) +
                     mSph * (0.4f * r * r + 0.25f * length * length + 0.375f * length * r);
        *mOut = mCyl + mSph;
        *cOut = m3MulSV3(0.5f, m3Add3(p1, p2));
        float d = axial - perp;
        iOut->cx = m3V3(perp + d * u.x * u.x, d * u.x * u.y, d * u.x * u.z);
        iOut->cy = m3V3(d * u.x * u.y, perp + d * u.y * u.y, d * u.y * u.z);
        iOut->cz = m3V3(d * u.x * u.z, d * u.y * u.z, perp + d * u.z * u.z);
        return 1;
    }
    if (type == M3_HULL)
    {
        const m3HullData* hull = &w->hullData[w->shapeHullIndex[s]];
        float density = w->shapeDensity[s];
        *mOut = density * hull->unitMass;
        *cOut = hull->unitCom;
        m3Mat3 ic = hull->unitInertiaCom;
        ic.cx = m3MulSV3(density, ic.cx);
        ic.cy = m3MulSV3(density, ic.cy);
        ic.cz = m3MulSV3(density, ic.cz);
        *iOut = ic;
        return 1;
    }
    return 0;
}

static void RecomputeMass(m3World* w, int32_t b)
{
    if (w->types[b] != M3_DYNAMIC)
    {
        w->invMass[b] = 0.0f;
        w->invInertiaLocal[b] = m3ZeroMat3();
        w->inertiaLocal[b] = m3ZeroMat3();
        w->localCenters[b] = m3V3(0.0f, 0.0f, 0.0f);
        w->minExtents[b] = 1.0e30f;
        w->maxExtents[b] = 0.0f;
        return;
    }
    float mass = 0.0f;
    m3Vec3 center = m3V3(0.0f, 0.0f, 0.0f);
    for (int32_t s = w->bodyShapeHead[b]; s != -1; s = w->shapeNext[s])
    {
        float m;
        m3Vec3 c;
        m3Mat3 ic;
        if (!ShapeMassProps(w, s, &m, &c, &ic))
        {
            continue;
        }
        ComposeMassProps(w, s, &c, &ic);
        mass += m;
        center = m3Add3(center, m3MulSV3(m, c));
    }
    if (!(mass > 0.0f))
    {
        w->invMass[b] = 1.0f;
        w->invInertiaLocal[b] = m3ZeroMat3();
        w->inertiaLocal[b] = m3ZeroMat3();
        w->localCenters[b] = m3V3(0.0f, 0.0f, 0.0f);
        w->minExtents[b] = 1.0e30f;
        w->maxExtents[b] = 0.0f;
        return;
    }
    center = m3MulSV3(1.0f / mass, center);
    m3Mat3 I = m3ZeroMat3();
    for (int32_t s = w->bodyShapeHead[b]; s != -1; s = w->shapeNext[s])
    {
        float m;
        m3Vec3 c;
        m3Mat3 ic;
        if (!ShapeMassProps(w, s, &m, &c, &ic))
        {
            continue;
        }
        ComposeMassProps(w, s, &c, &ic);
        m3Vec3 d = m3Sub3(c, center);
        float d2 = m3Dot3(d, d);
        // Parallel axis theorem, term by term.
        I.cx.x += ic.cx.x + m * (d2 - d.x * d.x);
        I.cy.y += ic.cy.y + m * (d2 - d.y * d.y);
        I.cz.z += ic.cz.z + m * (d2 - d.z * d.z);
        I.cy.x += ic.cy.x - m * d.x * d.y;
        I.cx.y += ic.cx.y - m * d.x * d.y;
        I.cz.x += ic.cz.x - m * d.x * d.z;
        I.cx.z += ic.cx.z - m * d.x * d.z;
        I.cz.y += ic.cz.y - m * d.y * d.z;
        I.cy.z += ic.cy.z - m * d.y * d.z;
    }
    w->invMass[b] = 1.0f / mass;
    w->inertiaLocal[b] = I;
    w->invInertiaLocal[b] = InvertSymmetric(I);

    float minExtent = 1.0e30f;
    float maxExtent = 0.0f;
    for (int32_t s = w->bodyShapeHead[b]; s != -1; s = w->shapeNext[s])
    {
        uint8_t type = w->shapeType[s];
        if (type == M3_SPHERE)
        {
            float r = w->shapeGeom[s].s;
            m3Vec3 d = m3Sub3(w->shapeGeom[s].v, center);
            minExtent = m3MinF(minExtent, r);
            maxExtent = m3MaxF(maxExtent, sqrtf(m3Dot3(d, d)) + r);
        }
        else if (type == M3_CAPSULE)
        {
            float r = w->shapeGeom[s].s;
            m3Vec3 d1 = m3Sub3(w->shapeGeom[s].v, center);
            m3Vec3 d2 = m3Sub3(w->shapeGeom[s].v2, center);
            minExtent = m3MinF(minExtent, r);
            maxExtent = m3MaxF(maxExtent, sqrtf(m3MaxF(m3Dot3(d1, d1), m3Dot3(d2, d2))) + r);
        }
        else if (type == M3_HULL)
        {
            const m3HullData* hull = &w->hullData[w->shapeHullIndex[s]];
            for (int32_t f = 0; f < hull->faceCount; ++f)
            {
                float dist = hull->faceOffsets[f] - m3Dot3(hull->faceNormals[f], center);
                minExtent = m3MinF(minExtent, dist);
            }
            for (int32_t v = 0; v < hull->vertexCount; ++v)
            {
                m3Vec3 d = m3Sub3(hull->vertices[v], center);
                maxExtent = m3MaxF(maxExtent, sqrtf(m3Dot3(d, d)));
            }
        }
    }
    w->minExtents[b] = minExtent;
    w->maxExtents[b] = maxExtent;
    w->localCenters[b] = center;
}

// ---------------------------------------------------------------------------
// Shape bounds on the host (creation-time proxy box; the device recomputes
// them every step with the same arithmetic).  src/broad_phase.c:24-148.
// ---------------------------------------------------------------------------

static m3Transform ShapeWorldTransform(const m3World* w, int32_t s)
{
    m3Transform xf = w->transforms[w->shapeBody[s]];
    if (w->shapeHasOffset[s] != 0)
    {
        m3Vec3 r = m3RotateVec3(xf.q, w->shapeLocalPos[s]);
        xf.p.x += (double)r.x;
        xf.p.y += (double)r.y;
        xf.p.z += (double)r.z;
        xf.q = m3MulQuat(xf.q, w->shapeLocalRot[s]);
    }
    return xf;
}

static void ShapeBounds(const m3World* w, int32_t s, double box[6])
{
    m3Transform xf = ShapeWorldTransform(w, s);
    uint8_t type = w->shapeType[s];
    if (type == M3_HULL)
    {
        const m3HullData* hull = &w->hullData[w->shapeHullIndex[s]];
        double lo[3] = {1.0e30, 1.0e30, 1.0e30};
        double hi[3] = {-1.0e30, -1.0e30, -1.0e30};
        for (int32_t v = 0; v < hull->vertexCount; ++v)
        {
            m3Vec3 r = m3RotateVec3(xf.q, hull->vertices[v]);
            double p[3] = {xf.p.x + (double)r.x, xf.p.y + (double)r.y, xf.p.z + (double)r.z};
            for (int32_t k = 0; k < 3; ++k)
            {
                lo[k] = p[k] < lo[k] ? p[k] : lo[k];
                hi[k] = p[k] > hi[k] ? p[k] : hi[k];
            }
        }
        for (int32_t k = 0; k < 3; ++k)
        {
            box[k] = lo[k] - (double)M3_AABB_MARGIN;
            box[3 + k] = hi[k] + (double)M3_AABB_MARGIN;
        }
        return;
    }
    if (type == M3_CAPSULE)
    {
        m3Vec3 r1 = m3RotateVec3(xf.q, w->shapeGeom[s].v);
        m3Vec3 r2 = m3RotateVec3(xf.q, w->shapeGeom[s].v2);
        double c1[3] = {xf.p.x + (double)r1.x, xf.p.y + (double)r1.y, xf.p.z + (double)r1.z};
        double c2[3] = {xf.p.x + (double)r2.x, xf.p.y + (double)r2.y, xf.p.z + (double)r2.z};
        double fat = (double)(w->shapeGeom[s].s + M3_AABB_MARGIN);
        for (int32_t k = 0; k < 3; ++k)
        {
            box[k] = (c1[k] < c2[k] ? c1[k] : c2[k]) - fat;
            box[3 + k] = (c1[k] > c2[k] ? c1[k] : c2[k]) + fat;
        }
        return;
    }
    m3Vec3 r = m3RotateVec3(xf.q, w->shapeGeom[s].v);
    double c[3] = {xf.p.x + (double)r.x, xf.p.y + (double)r.y, xf.p.z + (double)r.z};
    double fat = (double)(w->shapeGeom[s].s + M3_AABB_MARGIN);
    for (int32_t k = 0; k < 3; ++k)
    {
        box[k] = c[k] - fat;
        box[3 + k] = c[k] + fat;
    }
}

// ---------------------------------------------------------------------------
// Shapes (src/shape.c:411-625, 716-823, 1334-1344)
// ---------------------------------------------------------------------------

m3ShapeDef m3DefaultShapeDef(void)
{
    m3ShapeDef def;
    memset(&def, 0, sizeof(def));
    def.density = 1.0f;
    def.friction = 0.6f;
    def.restitution = 0.0f;
    def.categoryBits = 1ull;
    def.maskBits = ~0ull;
    def.groupIndex = 0;
    def.localRotation = m3IdentityQuat();
    def.internalValue = M3_SHAPE_COOKIE;
    return def;
}

static m3ShapeId CreateShape(m3BodyId bodyId, const m3ShapeDef* def, uint8_t type,
                             const m3ShapeGeom* geom, const m3HullData* hull)
{
    m3ShapeId nullId = {0, 0, 0};
    m3World* w = WorldFromIndex0(bodyId.world0);
    int32_t b = w != NULL ? BodySlot(w, bodyId) : -1;
    if (b < 0 || def == NULL || def->internalValue != M3_SHAPE_COOKIE)
    {
        return nullId;
    }
    float rotLen2 = def->localRotation.x * def->localRotation.x +
                    def->localRotation.y * def->localRotation.y +
                    def->localRotation.z * def->localRotation.z +
                    def->localRotation.w * def->localRotation.w;
    if (!FiniteV3(def->localPosition) || !FiniteQ(def->localRotation) || rotLen2 < 0.99f ||
        rotLen2 > 1.01f || !FiniteF(def->density) || !(def->density > 0.0f) ||
        !FiniteF(def->friction) || def->friction < 0.0f || !FiniteF(def->restitution) ||
        def->restitution < 0.0f || !FiniteF(def->rollingResistance) ||
        def->rollingResistance < 0.0f)
    {
        return nullId;
    }
    if (hull == NULL)
    {
        // Per-type geometry wall of the plain-geometry door
        // (src/shape.c:433-470): journal replay hands raw record bytes here.
        bool ok = false;
        if (type == M3_SPHERE)
        {
            ok = FiniteV3(geom->v) && FiniteF(geom->s) && geom->s > 0.0f;
        }
        else if (type == M3_PLANE)
        {
            ok = FiniteV3(geom->v) && FiniteF(geom->s) && m3Dot3(geom->v, geom->v) > 1.0e-12f;
        }
        else if (type == M3_CAPSULE)
        {
            ok = FiniteV3(geom->v) && FiniteV3(geom->v2) && FiniteF(geom->s) && geom->s > 0.0f;
        }
        else if (type == M3_HULL)
        {
            ok = FiniteV3(geom->v) && geom->v.x > 0.0f && geom->v.y > 0.0f && geom->v.z > 0.0f;
        }
        if (!ok)
        {
            return nullId;
        }
    }
    Touch(w, M3GPU_FIELDS_ALL);
    int32_t s = PoolAlloc(&w->shapePool);
    if (s < 0)
    {
        return nullId;
    }
    w->shapeBody[s] = b;
    w->shapeType[s] = type;
    w->shapeGeom[s] = *geom;
    w->shapeDensity[s] = def->density;
    w->shapeFriction[s] = def->friction;
    w->shapeRestitution[s] = def->restitution;
    w->shapeRollingResistance[s] = def->rollingResistance;
    w->shapeCategory[s] = def->categoryBits;
    w->shapeMask[s] = def->maskBits;
    w->shapeGroup[s] = def->groupIndex;
    w->shapeUserData[s] = def->userData;
    w->shapeSensor[s] = def->isSensor ? 1 : 0;
    w->shapeHitEvents[s] = def->enableHitEvents ? 1 : 0;
    w->shapePreSolve[s] = def->enablePreSolveEvents ? 1 : 0;
    w->shapeSurfaceVel[s] = m3V3(0.0f, 0.0f, 0.0f);
    w->shapeLocalPos[s] = def->localPosition;
    w->shapeLocalRot[s] = def->localRotation;
    w->shapeHasOffset[s] =
        (def->localPosition.x != 0.0f || def->localPosition.y != 0.0f ||
         def->localPosition.z != 0.0f || def->localRotation.x != 0.0f ||
         def->localRotation.y != 0.0f || def->localRotation.z != 0.0f ||
         def->localRotation.w != 1.0f)
            ? 1
            : 0;
    w->shapeNext[s] = w->bodyShapeHead[b];
    w->bodyShapeHead[b] = s;
    w->shapeHullIndex[s] = -1;
    w->shapeMeshIndex[s] = -1;
    if (type == M3_HULL)
    {
        m3HullData data;
        if (hull != NULL)
        {
            data = *hull;
        }
        else
        {
            BuildBoxHull(&data, geom->v);
        }
        w->shapeHullIndex[s] = InternHull(w, &data);
        if (w->shapeHullIndex[s] < 0)
        {
            w->bodyShapeHead[b] = w->shapeNext[s];
            w->shapeNext[s] = -1;
            w->shapeBody[s] = -1;
            PoolFree(&w->shapePool, s);
            return nullId;
        }
    }
    // Non-plane shapes get a proxy whose stored box starts as the tight
    // (margin-inflated) bounds, like a fresh tree leaf.
    if (type != M3_PLANE)
    {
        ShapeBounds(w, s, &w->leafBoxes[6 * s]);
        // The tree leaf (src/shape.c:592-599); the pool holds two nodes per
        // shape slot, so the insert cannot run out.
        w->proxyIds[s] = m3TreeInsert(&w->tree, &w->leafBoxes[6 * s], &w->leafBoxes[6 * s + 3], s);
    }
    RecomputeMass(w, b);
    m3ShapeId id = {s + 1, w->worldIndex0, w->shapePool.generations[s]};
    if (hull == NULL && w->journalActive != 0)
    {
        // Point-cloud hulls journal their input points instead.
        m3OpCreateShape rec;
        memset(&rec, 0, sizeof(rec));
        rec.def = *def;
        rec.geom = *geom;
        rec.body = bodyId;
        rec.expected = id;
        rec.type = type;
        JournalRecord(w, M3_OP_CREATE_SHAPE, &rec, (int32_t)sizeof(rec));
    }
    return id;
}

m3ShapeId m3CreateSphereShape(m3BodyId bodyId, const m3ShapeDef* def, const m3Sphere* sphere)
{
    m3ShapeId nullId = {0, 0, 0};
    if (sphere == NULL || !(sphere->radius > 0.0f) || !FiniteF(sphere->radius) ||
        !FiniteV3(sphere->center))
    {
        return nullId;
    }
    m3ShapeGeom geom;
    memset(&geom, 0, sizeof(geom));
    geom.v = sphere->center;
    geom.s = sphere->radius;
    return CreateShape(bodyId, def, M3_SPHERE, &geom, NULL);
}

m3ShapeId m3CreatePlaneShape(m3BodyId bodyId, const m3ShapeDef* def, const m3Plane* plane)
{
    m3ShapeId nullId = {0, 0, 0};
    if (plane == NULL || !FiniteV3(plane->normal) || !FiniteF(plane->offset) ||
        !(m3Dot3(plane->normal, plane->normal) > 1.0e-12f))
    {
        return nullId;
    }
    m3World* w = WorldFromIndex0(bodyId.world0);
    int32_t b = w != NULL ? BodySlot(w, bodyId) : -1;
    if (b < 0 || w->types[b] != M3_STATIC)
    {
        return nullId;
    }
    m3ShapeGeom geom;
    memset(&geom, 0, sizeof(geom));
    geom.v = m3Normalize3(plane->normal);
    geom.s = plane->offset;
    return CreateShape(bodyId, def, M3_PLANE, &geom, NULL);
}

m3ShapeId m3CreateCapsuleShape(m3BodyId bodyId, const m3ShapeDef* def, const m3Capsule* capsule)
{
    m3ShapeId nullId = {0, 0, 0};
    if (capsule == NULL || !(capsule->radius > 0.0f) || !FiniteF(capsule->radius) ||
        !FiniteV3(capsule->point1) || !FiniteV3(capsule->point2))
    {
        return nullId;
    }
    m3Vec3 axis = m3Sub3(capsule->point2, capsule->point1);
    if (!(m3Dot3(axis, axis) > 0.0f))
    {
        return nullId;
    }
    m3ShapeGeom geom;
    memset(&geom, 0, sizeof(geom));
    geom.v = capsule->point1;
    geom.s = capsule->radius;
    geom.v2 = capsule->point2;
    return CreateShape(bodyId, def, M3_CAPSULE, &geom, NULL);
}

m3ShapeId m3CreateBoxShape(m3BodyId bodyId, const m3ShapeDef* def, m3Vec3 halfExtents)
{
    m3ShapeId nullId = {0, 0, 0};
    if (!(halfExtents.x > 0.0f) || !(halfExtents.y > 0.0f) || !(halfExtents.z > 0.0f) ||
        !FiniteV3(halfExtents))
    {
        return nullId;
    }
    m3ShapeGeom geom;
    memset(&geom, 0, sizeof(geom));
    geom.v = halfExtents;
    return CreateShape(bodyId, def, M3_HULL, &geom, NULL);
}

// ---------------------------------------------------------------------------
// Step (src/solver.c:4322-4330 wall; the body runs on the device)
// ---------------------------------------------------------------------------

// A convex hull of 4..256 points (src/shape.c:988-1036): QuickHull output,
// interned like every hull; degenerate clouds and clouds past the hull caps
// give the null id.
m3ShapeId m3CreateHullShape(m3BodyId bodyId, const m3ShapeDef* def, const m3Vec3* points,
                            int32_t count)
{
    m3ShapeId nullId = {0, 0, 0};
    m3World* w = WorldFromIndex0(bodyId.world0);
    int32_t b = w != NULL ? BodySlot(w, bodyId) : -1;
    if (b < 0 || def == NULL || def->internalValue != M3_SHAPE_COOKIE || points == NULL ||
        count <= 0)
    {
        return nullId;
    }
    for (int32_t i = 0; i < count; ++i)
    {
        if (!FiniteV3(points[i]))
        {
            return nullId;
        }
    }
    m3HullData data;
    if (!m3HostComputeHull(points, count, &data))
    {
        return nullId;
    }
    m3ShapeGeom geom;
    memset(&geom, 0, sizeof(geom));
    m3ShapeId id = CreateShape(bodyId, def, M3_HULL, &geom, &data);
    if (id.index1 != 0 && w->journalActive != 0)
    {
        m3OpCreateHullShape* rec = count <= M3_HULL_MAX_INPUT
                                       ? (m3OpCreateHullShape*)calloc(1, sizeof(m3OpCreateHullShape))
                                       : NULL;
        if (rec == NULL)
        {
            w->journalOverflow = 1; // cannot record: the journal is invalid
            w->journalActive = 0;
            return id;
        }
        rec->def = *def;
        rec->body = bodyId;
        rec->expected = id;
        rec->count = count;
        memcpy(rec->points, points, (size_t)count * sizeof(m3Vec3));
        JournalRecord(w, M3_OP_CREATE_HULL_SHAPE, rec, (int32_t)sizeof(*rec));
        free(rec);
    }
    return id;
}

// ---------------------------------------------------------------------------
// Joints (src/joint.c:15-520): the spherical joint with its cone and twist
// limits.  Other kinds are refused (m3_nullJointId).
// ---------------------------------------------------------------------------

m3JointDef m3DefaultJointDef(void)
{
    m3JointDef def;
    memset(&def, 0, sizeof(def));
    def.type = m3_sphericalJoint;
    def.localAxisA = m3V3(0.0f, 0.0f, 1.0f);
    def.localAxisB = m3V3(0.0f, 0.0f, 1.0f);
    def.genericMotorAxis = 255;
    def.ratio = 1.0f;
    def.internalValue = M3_JOINT_COOKIE;
    return def;
}

// Quaternion of an orthonormal right-handed column basis, branching on the
// largest diagonal term (src/joint.c:30-75).
static m3Quat QuatFromBasis(m3Vec3 t1, m3Vec3 t2, m3Vec3 axis)
{
    float m00 = t1.x, m01 = t2.x, m02 = axis.x;
    float m10 = t1.y, m11 = t2.y, m12 = axis.y;
    float m20 = t1.z, m21 = t2.z, m22 = axis.z;
    m3Quat q;
    float trace = m00 + m11 + m22;
    if (trace > 0.0f)
    {
        float s = sqrtf(trace + 1.0f) * 2.0f;
        q.w = 0.25f * s;
        q.x = (m21 - m12) / s;
        q.y = (m02 - m20) / s;
        q.z = (m10 - m01) / s;
    }
    else if (m00 > m11 && m00 > m22)
    {
        float s = sqrtf(1.0f + m00 - m11 - m22) * 2.0f;
        q.w = (m21 - m12) / s;
        q.x = 0.25f * s;
        q.y = (m01 + m10) / s;
        q.z = (m02 + m20) / s;
    }
    else if (m11 > m22)
    {
        float s = sqrtf(1.0f + m11 - m00 - m22) * 2.0f;
        q.w = (m02 - m20) / s;
        q.x = (m01 + m10) / s;
        q.y = 0.25f * s;
        q.z = (m12 + m21) / s;
    }
    else
    {
        float s = sqrtf(1.0f + m22 - m00 - m11) * 2.0f;
        q.w = (m10 - m01) / s;
        q.x = (m02 + m20) / s;
        q.y = (m12 + m21) / s;
        q.z = 0.25f * s;
    }
    return m3NormalizeQuat(q);
}

// Frame whose z-axis is the given unit axis; the tangent basis fixes x, y.
static m3Quat QuatFromAxisZ(m3Vec3 axis)
{
    m3Vec3 t1;
    m3Vec3 t2;
    m3TangentBasis(axis, &t1, &t2);
    return QuatFromBasis(t1, t2, axis);
}

static m3World* ResolveJoint(m3JointId id, int32_t* out)
{
    m3World* w = WorldFromIndex0(id.world0);
    int32_t index = id.index1 - 1;
    if (w == NULL || !PoolValid(&w->jointPool, index, id.generation))
    {
        return NULL;
    }
    *out = index;
    return w;
}

m3JointId m3CreateJoint(const m3JointDef* def)
{
    m3JointId none = {0, 0, 0};
    if (def == NULL || def->internalValue != M3_JOINT_COOKIE || def->type != m3_sphericalJoint)
    {
        return none;
    }
    if (!FiniteV3(def->localAnchorA) || !FiniteV3(def->localAnchorB) ||
        !FiniteV3(def->localAxisA) || !FiniteV3(def->localAxisB) || !FiniteF(def->lowerLimit) ||
        !FiniteF(def->upperLimit) || !FiniteF(def->motorSpeed) || !FiniteF(def->maxMotorEffort) ||
        !FiniteF(def->coneAngle) || def->maxMotorEffort < 0.0f ||
        (def->enableLimit && def->lowerLimit > def->upperLimit) ||
        (def->enableCone && def->coneAngle < 0.0f))
    {
        return none;
    }
    m3World* w = WorldFromIndex0(def->bodyA.world0);
    if (w == NULL || def->bodyB.world0 != def->bodyA.world0)
    {
        return none;
    }
    int32_t a = BodySlot(w, def->bodyA);
    int32_t b = BodySlot(w, def->bodyB);
    if (a < 0 || b < 0 || a == b)
    {
        return none;
    }
    if (w->types[a] != (uint8_t)M3_DYNAMIC && w->types[b] != (uint8_t)M3_DYNAMIC)
    {
        return none;
    }
    Touch(w, M3GPU_FIELDS_STATIC | M3GPU_FIELDS_DYNAMIC);
    int32_t index = PoolAlloc(&w->jointPool);
    if (index < 0)
    {
        return none;
    }
    w->jointType[index] = (uint8_t)def->type;
    w->jointBodyA[index] = a;
    w->jointBodyB[index] = b;
    w->jointLocalA[index] = def->localAnchorA;
    w->jointLocalB[index] = def->localAnchorB;
    w->jointCollide[index] = def->collideConnected ? 1 : 0;
    w->jointImpulse[index] = m3V3(0.0f, 0.0f, 0.0f);
    w->jointPerpImpulse[index] = m3V3(0.0f, 0.0f, 0.0f);
    w->jointLimitImpulse[index] = m3V3(0.0f, 0.0f, 0.0f);
    w->jointAngularImpulse[index] = m3V3(0.0f, 0.0f, 0.0f);
    w->jointFrameQA[index] = QuatFromAxisZ(m3Normalize3(def->localAxisA));
    w->jointFrameQB[index] = QuatFromAxisZ(m3Normalize3(def->localAxisB));
    w->jointFlags[index] = (uint8_t)((def->enableLimit ? 1 : 0) | (def->enableMotor ? 2 : 0) |
                                     (def->enableCone ? 4 : 0));
    w->jointMotor[index] = m3V3(def->motorSpeed, def->maxMotorEffort, 0.0f);
    w->jointBreak[index] = m3V3(0.0f, 0.0f, 0.0f);
    w->jointSpring[index] = m3V3(0.0f, 0.0f, 0.0f);
    w->jointTargetScalar[index] = 0.0f;
    w->jointTargetQ[index] = m3Q(0.0f, 0.0f, 0.0f, 1.0f);
    w->jointSpringImpulse[index] = m3V3(0.0f, 0.0f, 0.0f);
    w->jointLimits[index] = m3V3(def->lowerLimit, def->upperLimit, def->coneAngle);
    w->jointNextA[index] = w->bodyJointHead[a];
    w->bodyJointHead[a] = index;
    w->jointNextB[index] = w->bodyJointHead[b];
    w->bodyJointHead[b] = index;
    if (w->types[a] == (uint8_t)M3_DYNAMIC)
    {
        w->awake[a] = 1;
        w->sleepTimes[a] = 0.0f;
    }
    if (w->types[b] == (uint8_t)M3_DYNAMIC)
    {
        w->awake[b] = 1;
        w->sleepTimes[b] = 0.0f;
    }
    m3JointId id = {index + 1, w->worldIndex0, w->jointPool.generations[index]};
    if (w->journalActive != 0)
    {
        m3OpCreateJoint rec;
        memset(&rec, 0, sizeof(rec));
        rec.def = *def;
        rec.expected = id;
        JournalRecord(w, M3_OP_CREATE_JOINT, &rec, (int32_t)sizeof(rec));
    }
    return id;
}

static void UnlinkJoint(m3World* w, int32_t body, int32_t index)
{
    int32_t* cursor = &w->bodyJointHead[body];
    while (*cursor != -1)
    {
        int32_t j = *cursor;
        if (j == index)
        {
            *cursor = w->jointBodyA[j] == body ? w->jointNextA[j] : w->jointNextB[j];
            return;
        }
        cursor = w->jointBodyA[j] == body ? &w->jointNextA[j] : &w->jointNextB[j];
    }
}

void m3DestroyJoint(m3JointId jointId)
{
    int32_t index;
    m3World* w = ResolveJoint(jointId, &index);
    if (w == NULL)
    {
        return;
    }
    JournalRecord(w, M3_OP_DESTROY_JOINT, &jointId, (int32_t)sizeof(jointId));
    Touch(w, M3GPU_FIELDS_STATIC | M3GPU_FIELDS_DYNAMIC);
    int32_t a = w->jointBodyA[index];
    int32_t b = w->jointBodyB[index];
    UnlinkJoint(w, a, index);
    UnlinkJoint(w, b, index);
    if (w->types[a] == (uint8_t)M3_DYNAMIC && w->bodyPool.alive[a] != 0)
    {
        w->awake[a] = 1;
        w->sleepTimes[a] = 0.0f;
    }
    if (w->types[b] == (uint8_t)M3_DYNAMIC && w->bodyPool.alive[b] != 0)
    {
        w->awake[b] = 1;
        w->sleepTimes[b] = 0.0f;
    }
    m3Vec3 zero = m3V3(0.0f, 0.0f, 0.0f);
    m3Quat identity = m3Q(0.0f, 0.0f, 0.0f, 1.0f);
    w->jointType[index] = 0;
    w->jointBodyA[index] = -1;
    w->jointBodyB[index] = -1;
    w->jointLocalA[index] = zero;
    w->jointLocalB[index] = zero;
    w->jointCollide[index] = 0;
    w->jointImpulse[index] = zero;
    w->jointPerpImpulse[index] = zero;
    w->jointLimitImpulse[index] = zero;
    w->jointAngularImpulse[index] = zero;
    w->jointFrameQA[index] = identity;
    w->jointFrameQB[index] = identity;
    w->jointFlags[index] = 0;
    w->jointMotor[index] = zero;
    w->jointBreak[index] = zero;
    w->jointSpring[index] = zero;
    w->jointTargetScalar[index] = 0.0f;
    w->jointTargetQ[index] = identity;
    w->jointSpringImpulse[index] = zero;
    w->jointLimits[index] = zero;
    w->jointNextA[index] = -1;
    w->jointNextB[index] = -1;
    PoolFree(&w->jointPool, index);
}

bool m3Joint_IsValid(m3JointId jointId)
{
    int32_t index;
    return ResolveJoint(jointId, &index) != NULL;
}

// At process exit the device mirrors of live worlds and the library's
// process-wide device state are released before the HIP runtime tears itself
// down: the handler is registered after the first device world exists, and
// exit handlers run in reverse order of registration.
static void ReleaseDeviceAtExit(void)
{
    for (int32_t i = 0; i < M3_MAX_WORLDS; ++i)
    {
        m3World* w = s_worlds[i];
        if (w != NULL && w->dev != NULL)
        {
            m3gpu_world_destroy(w->dev);
            w->dev = NULL;
        }
    }
    m3gpu_shutdown();
}

static int EnsureDevice(m3World* w)
{
    if (w->faulted != 0)
    {
        return 0;
    }
    if (w->dev == NULL)
    {
        m3gpuCaps caps;
        caps.bodyCapacity = w->bodyCapacity;
        caps.shapeCapacity = w->shapeCapacity;
        caps.pairCapacity = w->pairCapacity;
        caps.hullCapacity = w->shapeCapacity;
        caps.jointCapacity = w->jointCapacity;
        w->dev = m3gpu_world_create(&caps);
        if (w->dev == NULL)
        {
            DeviceFail(w, "m3gpu_world_create");
            return 0;
        }
        static int s_exitHandler = 0;
        if (s_exitHandler == 0)
        {
            s_exitHandler = 1;
            atexit(ReleaseDeviceAtExit);
        }
        w->devDirty = M3GPU_FIELDS_ALL;
    }
    if (w->devDirty != 0)
    {
        if (w->devDirty & M3GPU_FIELDS_STATIC)
        {
            BuildJointComponents(w);
        }
        m3gpuHostState h;
        HostView(w, &h);
        if (m3gpu_upload(w->dev, &h, w->devDirty) != 0)
        {
            DeviceFail(w, "m3gpu_upload");
            return 0;
        }
        w->devDirty = 0;
        w->pairsFullQuery = 0; // consumed by the device's next step
    }
    return 1;
}

static void FillParams(const m3World* w, float dt, int32_t substeps, m3gpuStepParams* p)
{
    memset(p, 0, sizeof(*p));
    p->dt = dt;
    p->substeps = substeps;
    p->gravity[0] = w->gravity.x;
    p->gravity[1] = w->gravity.y;
    p->gravity[2] = w->gravity.z;
    p->contactHertz = w->contactHertz;
    p->contactDampingRatio = w->contactDampingRatio;
    p->contactPushMaxSpeed = w->contactPushMaxSpeed;
    p->restitutionThreshold = w->restitutionThreshold;
    p->maximumLinearSpeed = w->maximumLinearSpeed;
    p->maximumAngularSpeed = w->maximumAngularSpeed;
    p->sleepEnabled = w->sleepEnabled;
    p->continuousEnabled = w->continuousEnabled;
}

// ---------------------------------------------------------------------------
// Broadphase tree mirror (m3_tree.h)
// ---------------------------------------------------------------------------

// The reference's canonical rebuild (src/world.c:1140-1198): live proxies in
// shape order with their current leaf boxes.  The host's leaf boxes must be
// current (SyncHost).
static void TreeRebuild(m3World* w)
{
    const int32_t ns = w->shapePool.maxIndex;
    int32_t count = 0;
    for (int32_t s = 0; s < ns; ++s)
    {
        count += w->shapePool.alive[s] != 0 && w->proxyIds[s] >= 0;
    }
    w->treeStale = 0;
    if (count == 0)
    {
        return;
    }
    double (*los)[3] = (double (*)[3])Zalloc((size_t)count * 3 * sizeof(double));
    double (*his)[3] = (double (*)[3])Zalloc((size_t)count * 3 * sizeof(double));
    int32_t* uds = (int32_t*)Zalloc((size_t)count * sizeof(int32_t));
    int32_t* nodes = (int32_t*)Zalloc((size_t)count * sizeof(int32_t));
    int32_t n = 0;
    for (int32_t s = 0; s < ns; ++s)
    {
        if (w->shapePool.alive[s] == 0 || w->proxyIds[s] < 0)
        {
            continue;
        }
        memcpy(los[n], &w->leafBoxes[6 * s], 3 * sizeof(double));
        memcpy(his[n], &w->leafBoxes[6 * s + 3], 3 * sizeof(double));
        uds[n] = s;
        n += 1;
    }
    if (m3TreeRebuild(&w->tree, (const double (*)[3])los, (const double (*)[3])his, uds, n, nodes))
    {
        for (int32_t i = 0; i < n; ++i)
        {
            w->proxyIds[uds[i]] = nodes[i];
        }
    }
    free(los);
    free(his);
    free(uds);
    free(nodes);
}

// Replays one step's leaf moves (src/broad_phase.c:505-539): every shape
// whose leaf box changed escaped its old leaf, and is removed and reinserted
// with the new box, in shape order.  A leaf box changes exactly when the
// shape escaped (the new box contains the tight box the old one did not).
static void TreeReplay(m3World* w, const double* leaves)
{
    const int32_t ns = w->shapePool.maxIndex;
    for (int32_t s = 0; s < ns; ++s)
    {
        const int32_t leaf = w->proxyIds[s];
        if (w->shapePool.alive[s] == 0 || leaf < 0)
        {
            continue;
        }
        const m3TreeNode* node = &w->tree.nodes[leaf];
        const double* box = &leaves[6 * s];
        if (memcmp(node->lo, box, 3 * sizeof(double)) == 0 &&
            memcmp(node->hi, box + 3, 3 * sizeof(double)) == 0)
        {
            continue;
        }
        m3TreeRemove(&w->tree, leaf);
        w->proxyIds[s] = m3TreeInsert(&w->tree, box, box + 3, s);
    }
}

static bool TreeHasLeaves(const m3World* w) { return w->tree.root != M3_TREE_NULL; }

// The mirror is current from here on (rebuilding a stale one).
static bool TreeFresh(m3World* w)
{
    if (w->treeStale != 0)
    {
        if (!SyncHost(w))
        {
            return false;
        }
        TreeRebuild(w);
    }
    return true;
}

static bool AnyPreSolveShape(const m3World* w)
{
    for (int32_t s = 0; s < w->shapePool.maxIndex; ++s)
    {
        if (w->shapePool.alive[s] != 0 && w->shapePreSolve[s] != 0)
        {
            return true;
        }
    }
    return false;
}

// The serial pre-solve pass (src/solver.c:214-248): the callback in pair
// order for every candidate the device gathered; the keys it vetoes (in pair
// order, so ascending) go back to the device and into the journal annex.
static int PreSolvePass(m3World* w, const m3gpuStepParams* p)
{
    const m3gpuPreSolve* cands = NULL;
    int32_t n = 0;
    if (m3gpu_step_presolve(w->dev, &cands, &n) != 0)
    {
        return -1;
    }
    for (int32_t k = 0; k < n; ++k)
    {
        const m3gpuPreSolve* c = &cands[k];
        const int32_t a = (int32_t)(c->key >> 32);
        const int32_t b = (int32_t)(c->key & 0xFFFFFFFFu);
        m3Pos3 point = {c->point[0], c->point[1], c->point[2]};
        m3Vec3 normal = m3V3(c->normal[0], c->normal[1], c->normal[2]);
        if (!w->preSolveFn(ShapeIdOf(w, a), ShapeIdOf(w, b), point, normal, w->preSolveContext))
        {
            w->stepVetoKeys[w->stepVetoCount++] = c->key;
        }
    }
    return m3gpu_step_vetoes(w->dev, p, w->stepVetoKeys, w->stepVetoCount);
}

// One step; false when the step did not run (bad arguments, a faulted world,
// or a world the device step refuses), in which case nothing changed.
static bool StepWorld(m3World* w, float dt, int32_t substeps)
{
    if (w == NULL || !(dt > 0.0f) || substeps < 1 || substeps > M3_MAX_SUBSTEPS)
    {
        return false;
    }
    if (!EnsureDevice(w))
    {
        return false;
    }
    m3gpuStepParams p;
    FillParams(w, dt, substeps, &p);
    const bool treeCcd = m3gpu_step_needs_tree(w->dev, &p) != 0;
    const bool tree = treeCcd || w->treeMirror != 0;
    const bool replayVetoes = w->replayVetoCount > 0;
    // A replayed journal's recorded vetoes own the step; the callback stays
    // silent then (src/solver.c:184-212).
    const bool presolve = !replayVetoes && w->preSolveFn != NULL && AnyPreSolveShape(w);
    w->stepVetoCount = 0;
    int rc;
    if (tree || presolve || replayVetoes)
    {
        // The device stops where the host takes part: after the island
        // wake pass for the pre-solve callback, before the continuous pass
        // for the broadphase tree's leaf moves and the continuous pass's
        // tree order.
        if (tree && !TreeFresh(w))
        {
            return false;
        }
        if (tree && w->pairTableLive != 0 && w->pairTableOnDevice == 0)
        {
            // The device takes over the table from the host arrays.
            m3gpuHostState h;
            HostView(w, &h);
            if (m3gpu_set_pair_table(w->dev, 1) != 0 ||
                m3gpu_upload(w->dev, &h, M3GPU_FIELDS_PAIR_TABLE) != 0)
            {
                DeviceFail(w, "m3gpu_upload");
                return false;
            }
            w->pairTableOnDevice = 1;
        }
        const uint32_t flags = (tree ? M3GPU_STEP_LEAVES : 0u) | (presolve ? M3GPU_STEP_PRESOLVE : 0u);
        rc = m3gpu_step_begin(w->dev, &p, flags, replayVetoes ? w->replayVetoKeys : NULL,
                              replayVetoes ? w->replayVetoCount : 0);
        w->replayVetoCount = 0;
        if (rc == 0 && presolve)
        {
            rc = PreSolvePass(w, &p);
        }
        if (rc == 0)
        {
            const int32_t* ranks = NULL;
            if (tree)
            {
                const double* leaves = m3gpu_step_leaves(w->dev);
                if (leaves == NULL)
                {
                    DeviceFail(w, "m3gpu_step_leaves");
                    return false;
                }
                TreeReplay(w, leaves);
                if (treeCcd)
                {
                    // The reference's query stack holds 64 entries; deeper
                    // trees would drop leaves from its walks.
                    if (m3TreeLeafRanks(&w->tree, w->treeRanks, w->shapePool.maxIndex) > 64)
                    {
                        fprintf(stderr, "maul3d_amd: broadphase tree deeper than the reference's "
                                        "query stack; continuous candidates may differ\n");
                    }
                    ranks = w->treeRanks;
                }
            }
            rc = m3gpu_step_end(w->dev, &p, ranks);
        }
        if (rc == 0 && !tree)
        {
            w->treeStale = TreeHasLeaves(w) ? 1 : 0;
            w->pairTableLive = 0;
            w->pairTableOnDevice = 0;
        }
    }
    else
    {
        if (w->pairTableOnDevice != 0 && m3gpu_set_pair_table(w->dev, 0) != 0)
        {
            DeviceFail(w, "m3gpu_set_pair_table");
            return false;
        }
        rc = m3gpu_step(w->dev, &p, NULL);
        if (rc == 0)
        {
            w->treeStale = TreeHasLeaves(w) ? 1 : 0;
            w->pairTableLive = 0;
            w->pairTableOnDevice = 0;
        }
    }
    if (rc == M3GPU_REFUSED)
    {
        return false; // a world the device step does not support: no-op
    }
    if (rc < 0)
    {
        DeviceFail(w, "m3gpu_step");
        return false;
    }
    w->hostStale = M3GPU_FIELDS_DYNAMIC | M3GPU_FIELDS_PAIRS;
    w->eventsState = 1;
    w->eventThreshold = w->hitEventThreshold;
    // Recorded after the step, as the reference (src/solver.c:4332-4354):
    // the step's vetoes first, so a bare replay repeats them.
    if (w->journalActive != 0)
    {
        if (w->stepVetoCount > 0)
        {
            JournalRecord(w, M3_OP_STEP_VETOES, w->stepVetoKeys,
                          w->stepVetoCount * (int32_t)sizeof(uint64_t));
        }
        m3OpStep rec;
        memset(&rec, 0, sizeof(rec));
        rec.dt = dt;
        rec.substeps = substeps;
        JournalRecord(w, M3_OP_STEP, &rec, (int32_t)sizeof(rec));
    }
    return true;
}

void m3World_Step(m3WorldId worldId, float dt, int32_t substeps)
{
    (void)StepWorld(WorldFromId(worldId), dt, substeps);
}

void m3World_SetPreSolveCallback(m3WorldId worldId, m3PreSolveFn* fn, void* context)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL)
    {
        return;
    }
    w->preSolveFn = fn;
    w->preSolveContext = context;
}

bool m3World_EnableTreeMirror(m3WorldId worldId, bool flag)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || w->faulted != 0)
    {
        return false;
    }
    if (flag && !TreeFresh(w))
    {
        return false;
    }
    w->treeMirror = flag ? 1 : 0;
    return true;
}

bool m3World_IsTreeMirrorCurrent(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    return w != NULL && w->treeStale == 0;
}

void m3World_RebuildBroadphase(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || !SyncHost(w))
    {
        return;
    }
    if (w->journalActive != 0)
    {
        int32_t zero = 0;
        JournalRecord(w, M3_OP_REBUILD_BROADPHASE, &zero, 4);
    }
    // The leaves keep their boxes (the reference preserves fatness), so the
    // device's leaf boxes stay as they are.
    TreeRebuild(w);
}

// The divergence finder's body comparison (src/diff.c): every body slot of
// two worlds of equal capacity, rows of the slots that differ, worst first
// (presence, then position + rotation + velocity error), slot ascending on
// ties.  A read: both worlds are brought up to date from their devices.
static float DiffTotal(const m3BodyDiff* d)
{
    const float presence = (d->onlyInA || d->onlyInB) ? 1.0e30f : 0.0f;
    return presence + d->positionError + d->rotationError + d->velocityError;
}

static int DiffOrder(const void* pa, const void* pb)
{
    const m3BodyDiff* a = (const m3BodyDiff*)pa;
    const m3BodyDiff* b = (const m3BodyDiff*)pb;
    const float ta = DiffTotal(a);
    const float tb = DiffTotal(b);
    if (ta != tb)
    {
        return ta > tb ? -1 : 1;
    }
    return (a->body.index1 > b->body.index1) - (a->body.index1 < b->body.index1);
}

int32_t m3World_DiffReport(m3WorldId worldA, m3WorldId worldB, m3BodyDiff* out, int32_t capacity,
                           int32_t* outCount)
{
    m3World* a = WorldFromId(worldA);
    m3World* b = WorldFromId(worldB);
    if (outCount != NULL)
    {
        *outCount = 0;
    }
    if (a == NULL || b == NULL || out == NULL || capacity <= 0 || outCount == NULL ||
        a->bodyCapacity != b->bodyCapacity || !SyncHost(a) || !SyncHost(b))
    {
        return -1;
    }
    const int32_t maxA = a->bodyPool.maxIndex;
    const int32_t maxB = b->bodyPool.maxIndex;
    const int32_t n = maxA > maxB ? maxA : maxB;
    m3BodyDiff* rows = (m3BodyDiff*)Zalloc((size_t)(n > 0 ? n : 1) * sizeof(m3BodyDiff));
    int32_t found = 0;
    for (int32_t i = 0; i < n; ++i)
    {
        const int inA = i < maxA && a->bodyPool.alive[i] != 0;
        const int inB = i < maxB && b->bodyPool.alive[i] != 0;
        if (!inA && !inB)
        {
            continue;
        }
        m3BodyDiff d;
        memset(&d, 0, sizeof(d));
        d.body.index1 = i + 1;
        d.body.world0 = a->worldIndex0;
        d.body.generation = inA ? a->bodyPool.generations[i] : b->bodyPool.generations[i];
        if (inA != inB)
        {
            d.onlyInA = inA != 0;
            d.onlyInB = inB != 0;
            rows[found++] = d;
            continue;
        }
        const m3Transform* xa = &a->transforms[i];
        const m3Transform* xb = &b->transforms[i];
        const m3Vec3 dp = m3V3((float)(xa->p.x - xb->p.x), (float)(xa->p.y - xb->p.y),
                               (float)(xa->p.z - xb->p.z));
        const float qdot = xa->q.x * xb->q.x + xa->q.y * xb->q.y + xa->q.z * xb->q.z + xa->q.w * xb->q.w;
        const m3Vec3 dv = m3Sub3(a->linearVelocities[i], b->linearVelocities[i]);
        const m3Vec3 dw = m3Sub3(a->angularVelocities[i], b->angularVelocities[i]);
        d.positionError = m3Length3(dp);
        d.rotationError = 1.0f - (qdot < 0.0f ? -qdot : qdot);
        d.velocityError = m3Length3(dv) + m3Length3(dw);
        if (d.positionError != 0.0f || d.rotationError != 0.0f || d.velocityError != 0.0f ||
            a->awake[i] != b->awake[i])
        {
            rows[found++] = d;
        }
    }
    qsort(rows, (size_t)found, sizeof(m3BodyDiff), DiffOrder);
    const int32_t written = found < capacity ? found : capacity;
    memcpy(out, rows, (size_t)written * sizeof(m3BodyDiff));
    *outCount = written;
    free(rows);
    return found;
}

struct m3gpuWorld* m3World_DeviceWorld(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    return (w != NULL && w->faulted == 0) ? w->dev : NULL;
}

const char* m3World_GetDeviceError(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    return (w != NULL && w->faulted != 0) ? w->faultText : NULL;
}

// Batched stepping of many worlds with one call: the rollback and toss
// workloads.  Not in the reference API (its worlds step one by one).
int32_t m3World_StepMany(const m3WorldId* worlds, int32_t count, float dt, int32_t substeps)
{
    if (worlds == NULL || count <= 0 || !(dt > 0.0f) || substeps < 1 ||
        substeps > M3_MAX_SUBSTEPS)
    {
        return -1;
    }
    enum { CHUNK = 1024 };
    m3gpuWorld* devs[CHUNK];
    m3World* hosts[CHUNK];
    m3gpuStepParams params[CHUNK];
    int32_t n = 0;
    int32_t solo = 0;
    for (int32_t k = 0; k < count && n < CHUNK; ++k)
    {
        m3World* w = WorldFromId(worlds[k]);
        if (w == NULL || !EnsureDevice(w))
        {
            continue;
        }
        FillParams(w, dt, substeps, &params[n]);
        if (!m3gpu_step_supported(w->dev, &params[n]) || w->treeMirror != 0)
        {
            // Worlds whose step involves the host tree step on their own
            // (refused worlds do not step at all).
            solo += StepWorld(w, dt, substeps) ? 1 : 0;
            continue;
        }
        devs[n] = w->dev;
        hosts[n] = w;
        n += 1;
    }
    if (n == 0)
    {
        return solo;
    }
    int rc = m3gpu_step_many(devs, params, n);
    if (rc < 0)
    {
        for (int32_t k = 0; k < n; ++k)
        {
            DeviceFail(hosts[k], "m3gpu_step_many");
        }
        return -1;
    }
    m3OpStep rec;
    memset(&rec, 0, sizeof(rec));
    rec.dt = dt;
    rec.substeps = substeps;
    // Only the worlds handed to the device stepped; each records its step
    // (a world listed twice stepped twice and records twice).
    for (int32_t k = 0; k < n; ++k)
    {
        m3World* w = hosts[k];
        w->hostStale = M3GPU_FIELDS_DYNAMIC | M3GPU_FIELDS_PAIRS;
        w->eventsState = 1;
        w->eventThreshold = w->hitEventThreshold;
        if (TreeHasLeaves(w))
        {
            w->treeStale = 1;
        }
        w->pairTableLive = 0;
        w->pairTableOnDevice = 0;
        JournalRecord(w, M3_OP_STEP, &rec, (int32_t)sizeof(rec));
    }
    return n + solo;
}

m3AdjointDef m3DefaultAdjointDef(void)
{
    m3AdjointDef d;
    memset(&d, 0, sizeof(d));
    d.timeStep = 1.0f / 60.0f;
    d.subStepCount = 4;
    d.steps = 120;
    d.checkpointEvery = 10;
    d.precision = 32;
    return d;
}

int32_t m3World_Adjoint(const m3WorldId* worlds, int32_t count, const m3AdjointDef* defs,
                        m3AdjointResult* results)
{
    if (worlds == NULL || defs == NULL || results == NULL || count <= 0)
    {
        return -1;
    }
    enum { CHUNK = 1024 };
    if (count > CHUNK)
    {
        return -1;
    }
    m3gpuWorld* devs[CHUNK];
    m3gpuStepParams params[CHUNK];
    m3gpuAdjointRequest reqs[CHUNK];
    const int32_t precision = defs[0].precision == 64 ? 64 : 32;
    m3gpuHostState* hosts = NULL;
    if (precision == 64)
    {
        hosts = (m3gpuHostState*)calloc((size_t)count, sizeof(m3gpuHostState));
        if (hosts == NULL)
        {
            return -1;
        }
    }
    for (int32_t k = 0; k < count; ++k)
    {
        m3World* w = WorldFromId(worlds[k]);
        const m3AdjointDef* d = &defs[k];
        if (w == NULL || !(d->timeStep > 0.0f) || d->subStepCount < 1 ||
            d->subStepCount > M3_MAX_SUBSTEPS || d->steps < 1 || d->checkpointEvery < 1 ||
            (d->precision == 64) != (precision == 64))
        {
            free(hosts);
            return -1;
        }
        // The run reuses the world's device pair buffers: build the last
        // step's event streams first.
        BuildEvents(w);
        for (int32_t j = 0; j < w->jointPool.maxIndex; ++j)
        {
            if (w->jointPool.alive[j])
            {
                free(hosts);
                return -1; // joints are outside the adjoint's scope
            }
        }
        if (!EnsureDevice(w) || (hosts != NULL && !SyncHost(w)))
        {
            free(hosts);
            return -1;
        }
        devs[k] = w->dev;
        FillParams(w, d->timeStep, d->subStepCount, &params[k]);
        m3gpuAdjointRequest* r = &reqs[k];
        memset(r, 0, sizeof(*r));
        r->steps = d->steps;
        r->checkpointEvery = d->checkpointEvery;
        r->forwardOnly = d->forwardOnly;
        r->target[0] = d->target.x;
        r->target[1] = d->target.y;
        r->target[2] = d->target.z;
        r->force = (const float*)d->force;
        r->offsetKind = d->offsetKind;
        r->offsetBody = d->offsetBody;
        r->offsetComp = d->offsetComp;
        r->offsetValue = d->offsetValue;
        m3AdjointResult* o = &results[k];
        r->loss = &o->loss;
        r->gradLinearVelocity = (float*)o->gradLinearVelocity;
        r->gradAngularVelocity = (float*)o->gradAngularVelocity;
        r->gradMass = o->gradMass;
        r->gradFriction = o->gradFriction;
        r->gradForce = (float*)o->gradForce;
        r->unsupported = &o->unsupported;
        r->ccdPulled = &o->ccdPulled;
        r->precision = precision;
        if (hosts != NULL)
        {
            HostView(w, &hosts[k]);
            r->host = &hosts[k];
        }
        o->unsupported = 0;
        o->ccdPulled = 0;
    }
    int rc = m3gpu_adjoint_many(devs, params, reqs, count);
    free(hosts);
    if (rc < 0)
    {
        for (int32_t k = 0; k < count; ++k)
        {
            DeviceFail(WorldFromId(worlds[k]), "m3gpu_adjoint_many");
        }
    }
    return rc == 0 ? count : -1;
}

// Uploads pending host changes and waits for the world's queued device work
// (timing harnesses call it before and after a timed region).
void m3World_Sync(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL)
    {
        return;
    }
    if (EnsureDevice(w) && m3gpu_sync(w->dev) != 0)
    {
        DeviceFail(w, "m3gpu_sync");
    }
}

// Launch autotuning: times each candidate launch shape on the worlds' own
// current state (snapshotted first and restored before every candidate and
// at the end), keeps the fastest, and leaves the worlds as they were.  The
// shapes never change a step's bits, only its speed.
static double MonotonicSeconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1.0e-9 * (double)ts.tv_nsec;
}

static bool RestoreAll(const m3WorldId* worlds, int32_t count, void* const* snaps,
                       const int32_t* sizes)
{
    for (int32_t k = 0; k < count; ++k)
    {
        if (!m3World_Restore(worlds[k], snaps[k], sizes[k]))
        {
            return false;
        }
    }
    return true;
}

static void SyncAll(const m3WorldId* worlds, int32_t count)
{
    for (int32_t k = 0; k < count; ++k)
    {
        m3World_Sync(worlds[k]);
    }
}

int32_t m3World_Autotune(const m3WorldId* worlds, int32_t count, float timeStep,
                         int32_t subStepCount, int32_t trialSteps, m3AutotuneResult* result)
{
    if (worlds == NULL || result == NULL || count <= 0 || count > 1024 || !(timeStep > 0.0f) ||
        subStepCount < 1 || subStepCount > M3_MAX_SUBSTEPS || trialSteps < 1 || trialSteps > 64)
    {
        return -1;
    }
    for (int32_t k = 0; k < count; ++k)
    {
        m3World* w = WorldFromId(worlds[k]);
        if (w == NULL || !EnsureDevice(w))
        {
            return -1;
        }
    }
    memset(result, 0, sizeof(*result));
    void** snaps = (void**)calloc((size_t)count, sizeof(void*));
    int32_t* sizes = (int32_t*)calloc((size_t)count, sizeof(int32_t));
    // The trials are not part of any session: journals stay silent through
    // them (the worlds are restored afterwards), as during a replay.
    uint8_t* journaling = (uint8_t*)calloc((size_t)count, 1);
    int32_t rc = -1;
    if (snaps == NULL || sizes == NULL || journaling == NULL)
    {
        goto done;
    }
    for (int32_t k = 0; k < count; ++k)
    {
        m3World* w = WorldFromId(worlds[k]);
        journaling[k] = w->journalActive;
        w->journalActive = 0;
    }
    for (int32_t k = 0; k < count; ++k)
    {
        sizes[k] = m3World_SnapshotSize(worlds[k]);
        snaps[k] = sizes[k] > 0 ? malloc((size_t)sizes[k]) : NULL;
        if (snaps[k] == NULL || m3World_Snapshot(worlds[k], snaps[k], sizes[k]) != sizes[k])
        {
            goto done;
        }
    }
    // Candidates: threads per world workgroup for a batch; cooperative
    // workgroups x tail budget for one world.
    const bool batch = count > 1;
    static const int32_t kThreads[] = {64, 128, 256, 512};
    static const int32_t kBlocks[] = {0, 24, 32, 48};
    static const int32_t kTails[] = {512, 1024, 2048};
    int32_t n = 0;
    int32_t best = -1;
    const int32_t total = batch ? 4 : 12;
    for (int32_t c = 0; c < total && n < M3_AUTOTUNE_MAX_CANDIDATES; ++c)
    {
        const int32_t a = batch ? kThreads[c] : kBlocks[c / 3];
        const int32_t b = batch ? 0 : kTails[c % 3];
        if (batch ? m3gpu_set_batch_threads(a) != 0
                  : m3gpu_set_substep_tuning(WorldFromId(worlds[0])->dev, a, b) != 0)
        {
            continue; // beyond this device's limits
        }
        if (!RestoreAll(worlds, count, snaps, sizes))
        {
            goto done;
        }
        // One untimed step (graph capture, first touches), then the trial.
        if (batch)
        {
            m3World_StepMany(worlds, count, timeStep, subStepCount);
        }
        else
        {
            m3World_Step(worlds[0], timeStep, subStepCount);
        }
        SyncAll(worlds, count);
        const double t0 = MonotonicSeconds();
        for (int32_t s = 0; s < trialSteps; ++s)
        {
            if (batch)
            {
                m3World_StepMany(worlds, count, timeStep, subStepCount);
            }
            else
            {
                m3World_Step(worlds[0], timeStep, subStepCount);
            }
        }
        SyncAll(worlds, count);
        const double ms = 1.0e3 * (MonotonicSeconds() - t0) / (double)trialSteps;
        result->candidateA[n] = a;
        result->candidateB[n] = b;
        result->candidateMs[n] = (float)ms;
        if (best < 0 || ms < (double)result->candidateMs[best])
        {
            best = n;
        }
        n += 1;
    }
    if (best < 0)
    {
        goto done;
    }
    if (batch)
    {
        m3gpu_set_batch_threads(result->candidateA[best]);
        result->batchThreads = result->candidateA[best];
    }
    else
    {
        m3gpuWorld* dev = WorldFromId(worlds[0])->dev;
        m3gpu_set_substep_tuning(dev, result->candidateA[best], result->candidateB[best]);
        m3gpu_get_substep_tuning(dev, &result->substepBlocks, &result->tailLimit);
    }
    result->candidates = n;
    if (!RestoreAll(worlds, count, snaps, sizes))
    {
        goto done;
    }
    rc = n;
done:
    if (journaling != NULL)
    {
        for (int32_t k = 0; k < count; ++k)
        {
            m3World* w = WorldFromId(worlds[k]);
            w->journalActive = w->journalActive | journaling[k];
        }
    }
    if (snaps != NULL)
    {
        for (int32_t k = 0; k < count; ++k)
        {
            free(snaps[k]);
        }
    }
    free(snaps);
    free(sizes);
    free(journaling);
    return rc;
}

// Stage profile of the following steps (device event times + counts).
int32_t m3World_ProfileBegin(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || !EnsureDevice(w))
    {
        return -1;
    }
    return m3gpu_profile_begin(w->dev);
}

int32_t m3World_ProfileEnd(m3WorldId worldId, struct m3gpuProfile* out)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || w->dev == NULL)
    {
        return -1;
    }
    return m3gpu_profile_end(w->dev, out);
}

// ---------------------------------------------------------------------------
// Hash (src/snapshot.c:932-1390, the branches live for this build's state)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Event streams (src/world.c:960-1075, src/shape.c:1583-1593)
// ---------------------------------------------------------------------------

static const void* EventList(m3WorldId worldId, int32_t* count, int k)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL)
    {
        if (count != NULL)
        {
            *count = 0;
        }
        return NULL;
    }
    BuildEvents(w);
    if (count != NULL)
    {
        *count = w->eventCounts[k];
    }
    switch (k)
    {
    case 0: return w->beginEvents;
    case 1: return w->endEvents;
    case 2: return w->sensorBeginEvents;
    case 3: return w->sensorEndEvents;
    case 4: return w->hitEvents;
    default: return w->moveEvents;
    }
}

const m3ContactEvent* m3World_ContactBeginEvents(m3WorldId worldId, int32_t* count)
{
    return (const m3ContactEvent*)EventList(worldId, count, 0);
}

const m3ContactEvent* m3World_ContactEndEvents(m3WorldId worldId, int32_t* count)
{
    return (const m3ContactEvent*)EventList(worldId, count, 1);
}

const m3ContactEvent* m3World_SensorBeginEvents(m3WorldId worldId, int32_t* count)
{
    return (const m3ContactEvent*)EventList(worldId, count, 2);
}

const m3ContactEvent* m3World_SensorEndEvents(m3WorldId worldId, int32_t* count)
{
    return (const m3ContactEvent*)EventList(worldId, count, 3);
}

const m3HitEvent* m3World_HitEvents(m3WorldId worldId, int32_t* count)
{
    return (const m3HitEvent*)EventList(worldId, count, 4);
}

// At most one hit per constraint and pairCapacity slots: never drops.
int32_t m3World_HitEventsDropped(m3WorldId worldId)
{
    (void)worldId;
    return 0;
}

const m3BodyMoveEvent* m3World_BodyMoveEvents(m3WorldId worldId, int32_t* count)
{
    return (const m3BodyMoveEvent*)EventList(worldId, count, 5);
}

// Joints never break in this build (no break limits): always empty.
const m3JointBreakEvent* m3World_JointBreakEvents(m3WorldId worldId, int32_t* count)
{
    static m3JointBreakEvent none;
    if (count != NULL)
    {
        *count = 0;
    }
    return WorldFromId(worldId) != NULL ? &none : NULL;
}

void m3World_SetHitEventThreshold(m3WorldId worldId, float value)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || !FiniteF(value) || value < 0.0f)
    {
        return;
    }
    JournalRecord(w, M3_OP_SET_HIT_EVENT_THRESHOLD, &value, (int32_t)sizeof(value));
    w->hitEventThreshold = value; // the last step's streams keep the threshold they ran with
}

static m3World* ResolveShape(m3ShapeId id, int32_t* out)
{
    m3World* w = WorldFromIndex0(id.world0);
    int32_t index = id.index1 - 1;
    if (w == NULL || !PoolValid(&w->shapePool, index, id.generation))
    {
        *out = -1;
        return NULL;
    }
    *out = index;
    return w;
}

void m3Shape_EnableHitEvents(m3ShapeId shapeId, bool flag)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    if (w == NULL)
    {
        return;
    }
    if (w->journalActive != 0)
    {
        m3OpShapeFlag rec;
        memset(&rec, 0, sizeof(rec));
        rec.id = shapeId;
        rec.on = flag ? 1 : 0;
        JournalRecord(w, M3_OP_ENABLE_SHAPE_HIT_EVENTS, &rec, (int32_t)sizeof(rec));
    }
    Touch(w, M3GPU_FIELDS_STATIC);
    w->shapeHitEvents[s] = flag ? 1 : 0;
}

void m3Shape_EnablePreSolve(m3ShapeId shapeId, bool flag)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    if (w == NULL)
    {
        return;
    }
    if (w->journalActive != 0)
    {
        m3OpShapeFlag rec;
        memset(&rec, 0, sizeof(rec));
        rec.id = shapeId;
        rec.on = flag ? 1 : 0;
        JournalRecord(w, M3_OP_ENABLE_SHAPE_PRE_SOLVE, &rec, (int32_t)sizeof(rec));
    }
    Touch(w, M3GPU_FIELDS_STATIC);
    w->shapePreSolve[s] = flag ? 1 : 0;
}

// Shape materials (src/shape.c:1388-1541).  Friction, restitution and
// rolling resistance are read by the next step's prepare; a density change
// rebuilds the body's mass properties when asked.  Each is journaled after
// validation, as the reference records it.
static float* ShapeScalarField(m3World* w, int32_t op)
{
    return op == M3_OP_SET_SHAPE_FRICTION      ? w->shapeFriction
           : op == M3_OP_SET_SHAPE_RESTITUTION ? w->shapeRestitution
                                               : w->shapeRollingResistance;
}

static void SetShapeScalar(m3ShapeId shapeId, int32_t op, float value)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    if (w == NULL || !FiniteF(value) || value < 0.0f)
    {
        return;
    }
    if (w->journalActive != 0)
    {
        m3OpShapeScalar rec;
        memset(&rec, 0, sizeof(rec));
        rec.id = shapeId;
        rec.value = value;
        JournalRecord(w, op, &rec, (int32_t)sizeof(rec));
    }
    Touch(w, M3GPU_FIELDS_STATIC);
    ShapeScalarField(w, op)[s] = value;
}

static float GetShapeScalar(m3ShapeId shapeId, int32_t op)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    return w != NULL ? ShapeScalarField(w, op)[s] : 0.0f;
}

void m3Shape_SetFriction(m3ShapeId shapeId, float friction)
{
    SetShapeScalar(shapeId, M3_OP_SET_SHAPE_FRICTION, friction);
}

float m3Shape_GetFriction(m3ShapeId shapeId) { return GetShapeScalar(shapeId, M3_OP_SET_SHAPE_FRICTION); }

void m3Shape_SetRestitution(m3ShapeId shapeId, float restitution)
{
    SetShapeScalar(shapeId, M3_OP_SET_SHAPE_RESTITUTION, restitution);
}

float m3Shape_GetRestitution(m3ShapeId shapeId) { return GetShapeScalar(shapeId, M3_OP_SET_SHAPE_RESTITUTION); }

void m3Shape_SetRollingResistance(m3ShapeId shapeId, float value)
{
    SetShapeScalar(shapeId, M3_OP_SET_SHAPE_ROLLING, value);
}

float m3Shape_GetRollingResistance(m3ShapeId shapeId) { return GetShapeScalar(shapeId, M3_OP_SET_SHAPE_ROLLING); }

void m3Shape_SetDensity(m3ShapeId shapeId, float density, bool updateBodyMass)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    if (w == NULL || !FiniteF(density) || !(density > 0.0f))
    {
        return;
    }
    if (w->journalActive != 0)
    {
        m3OpShapeDensity rec;
        memset(&rec, 0, sizeof(rec));
        rec.id = shapeId;
        rec.value = density;
        rec.updateMass = updateBodyMass ? 1 : 0;
        JournalRecord(w, M3_OP_SET_SHAPE_DENSITY, &rec, (int32_t)sizeof(rec));
    }
    Touch(w, M3GPU_FIELDS_STATIC);
    w->shapeDensity[s] = density;
    if (updateBodyMass)
    {
        RecomputeMass(w, w->shapeBody[s]);
    }
}

float m3Shape_GetDensity(m3ShapeId shapeId)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    return w != NULL ? w->shapeDensity[s] : 0.0f;
}

bool m3Shape_IsValid(m3ShapeId shapeId)
{
    int32_t s;
    return ResolveShape(shapeId, &s) != NULL;
}

m3BodyId m3Shape_GetBody(m3ShapeId shapeId)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    m3BodyId none = {0, 0, 0};
    if (w == NULL)
    {
        return none;
    }
    const int32_t b = w->shapeBody[s];
    m3BodyId id = {b + 1, w->worldIndex0, w->bodyPool.generations[b]};
    return id;
}

float m3Body_GetMass(m3BodyId bodyId)
{
    int32_t i;
    m3World* w = ResolveBody(bodyId, &i);
    if (w == NULL || w->types[i] != M3_DYNAMIC || !(w->invMass[i] > 0.0f))
    {
        return 0.0f;
    }
    return 1.0f / w->invMass[i];
}

bool m3Shape_IsPreSolveEnabled(m3ShapeId shapeId)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    return w != NULL && w->shapePreSolve[s] != 0;
}

bool m3Shape_AreHitEventsEnabled(m3ShapeId shapeId)
{
    int32_t s;
    m3World* w = ResolveShape(shapeId, &s);
    return w != NULL && w->shapeHitEvents[s] != 0;
}

bool m3World_JournalBegin(m3WorldId worldId, void* buffer, int32_t capacity)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || buffer == NULL || capacity < 8 || w->journalActive != 0)
    {
        return false;
    }
    w->journalBuffer = (uint8_t*)buffer;
    w->journalCapacity = capacity;
    w->journalCursor = 0;
    w->journalActive = 1;
    w->journalOverflow = 0;
    return true;
}

int32_t m3World_JournalEnd(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL)
    {
        return -1;
    }
    int32_t bytes = w->journalOverflow != 0 ? -1 : w->journalCursor;
    w->journalBuffer = NULL;
    w->journalCapacity = 0;
    w->journalCursor = 0;
    w->journalActive = 0;
    w->journalOverflow = 0;
    return bytes;
}

// Booleans read from a journal are normalised to 0/1 (src/world.c:1490-1503).
static void NormalizeBool(void* record, size_t offset)
{
    uint8_t* b = (uint8_t*)record + offset;
    *b = *b != 0 ? 1 : 0;
}

// Applies one journal's records through the public entry points with ids
// retargeted to this world; any record this build cannot apply (an unknown
// or unsupported op, a wrong size, an id that does not come out as
// recorded) fails the whole replay.
static bool JournalApply(m3World* w, m3WorldId worldId, const uint8_t* stream, int32_t size)
{
    int32_t cursor = 0;
    while (cursor < size)
    {
        if (cursor + 8 > size)
        {
            return false;
        }
        int32_t op;
        int32_t bytes;
        memcpy(&op, stream + cursor, 4);
        memcpy(&bytes, stream + cursor + 4, 4);
        cursor += 8;
        if (bytes < 0 || bytes > size - cursor)
        {
            return false;
        }
        const uint8_t* payload = stream + cursor;
        cursor += bytes;
        switch (op)
        {
        case M3_OP_STEP:
        {
            m3OpStep rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            // A hostile record refuses (src/world.c:1635-1642), and so does
            // a step this build's device does not run.
            if (!StepWorld(w, rec.dt, rec.substeps))
            {
                return false;
            }
            break;
        }
        case M3_OP_CREATE_BODY:
        {
            m3OpCreateBody rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            NormalizeBool(&rec.def, offsetof(m3BodyDef, isBullet));
            m3BodyId id = m3CreateBody(worldId, &rec.def);
            if (id.index1 != rec.expected.index1 || id.generation != rec.expected.generation)
            {
                return false;
            }
            break;
        }
        case M3_OP_SET_LINEAR_VELOCITY:
        case M3_OP_SET_ANGULAR_VELOCITY:
        case M3_OP_APPLY_FORCE:
        case M3_OP_APPLY_TORQUE:
        {
            m3OpBodyVec rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            rec.id.world0 = w->worldIndex0;
            if (BodySlot(w, rec.id) < 0)
            {
                return false;
            }
            if (op == M3_OP_SET_LINEAR_VELOCITY)
            {
                m3Body_SetLinearVelocity(rec.id, rec.v);
            }
            else if (op == M3_OP_SET_ANGULAR_VELOCITY)
            {
                m3Body_SetAngularVelocity(rec.id, rec.v);
            }
            else
            {
                ApplyForceLike(rec.id, rec.v, op == M3_OP_APPLY_TORQUE);
            }
            break;
        }
        case M3_OP_REBUILD_BROADPHASE:
        {
            if (bytes != 4)
            {
                return false;
            }
            m3World_RebuildBroadphase(worldId);
            break;
        }
        case M3_OP_CREATE_SHAPE:
        {
            m3OpCreateShape rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            rec.body.world0 = w->worldIndex0;
            NormalizeBool(&rec.def, offsetof(m3ShapeDef, isSensor));
            NormalizeBool(&rec.def, offsetof(m3ShapeDef, enableHitEvents));
            NormalizeBool(&rec.def, offsetof(m3ShapeDef, enablePreSolveEvents));
            if (rec.type != M3_SPHERE && rec.type != M3_PLANE && rec.type != M3_CAPSULE &&
                rec.type != M3_HULL)
            {
                return false;
            }
            m3ShapeId id = CreateShape(rec.body, &rec.def, rec.type, &rec.geom, NULL);
            if (id.index1 != rec.expected.index1 || id.generation != rec.expected.generation)
            {
                return false;
            }
            break;
        }
        case M3_OP_CREATE_HULL_SHAPE:
        {
            m3OpCreateHullShape* rec = (m3OpCreateHullShape*)malloc(sizeof(m3OpCreateHullShape));
            if (rec == NULL || bytes != (int32_t)sizeof(*rec))
            {
                free(rec);
                return false;
            }
            memcpy(rec, payload, sizeof(*rec));
            rec->body.world0 = w->worldIndex0;
            NormalizeBool(&rec->def, offsetof(m3ShapeDef, isSensor));
            NormalizeBool(&rec->def, offsetof(m3ShapeDef, enableHitEvents));
            NormalizeBool(&rec->def, offsetof(m3ShapeDef, enablePreSolveEvents));
            m3ShapeId id = {0, 0, 0};
            if (rec->count >= 1 && rec->count <= M3_HULL_MAX_INPUT)
            {
                id = m3CreateHullShape(rec->body, &rec->def, rec->points, rec->count);
            }
            bool ok = id.index1 == rec->expected.index1 && id.index1 != 0 &&
                      id.generation == rec->expected.generation;
            free(rec);
            if (!ok)
            {
                return false;
            }
            break;
        }
        case M3_OP_CREATE_JOINT:
        {
            m3OpCreateJoint rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            NormalizeBool(&rec.def, offsetof(m3JointDef, enableLimit));
            NormalizeBool(&rec.def, offsetof(m3JointDef, enableMotor));
            NormalizeBool(&rec.def, offsetof(m3JointDef, enableCone));
            NormalizeBool(&rec.def, offsetof(m3JointDef, collideConnected));
            rec.def.bodyA.world0 = w->worldIndex0;
            rec.def.bodyB.world0 = w->worldIndex0;
            m3JointId id = m3CreateJoint(&rec.def);
            if (id.index1 != rec.expected.index1 || id.generation != rec.expected.generation)
            {
                return false;
            }
            break;
        }
        case M3_OP_DESTROY_JOINT:
        {
            m3JointId id;
            if (bytes != (int32_t)sizeof(id))
            {
                return false;
            }
            memcpy(&id, payload, sizeof(id));
            id.world0 = w->worldIndex0;
            int32_t index;
            if (ResolveJoint(id, &index) != w)
            {
                return false;
            }
            m3DestroyJoint(id);
            break;
        }
        case M3_OP_SET_GRAVITY:
        {
            m3Vec3 g;
            if (bytes != (int32_t)sizeof(g))
            {
                return false;
            }
            memcpy(&g, payload, sizeof(g));
            if (!FiniteV3(g))
            {
                return false;
            }
            m3World_SetGravity(worldId, g);
            break;
        }
        case M3_OP_ENABLE_SLEEPING:
        case M3_OP_ENABLE_CONTINUOUS:
        {
            int32_t on;
            if (bytes != (int32_t)sizeof(on))
            {
                return false;
            }
            memcpy(&on, payload, sizeof(on));
            if (op == M3_OP_ENABLE_SLEEPING)
            {
                m3World_EnableSleeping(worldId, on != 0);
            }
            else
            {
                m3World_EnableContinuous(worldId, on != 0);
            }
            break;
        }
        case M3_OP_SET_HIT_EVENT_THRESHOLD:
        {
            float value;
            if (bytes != (int32_t)sizeof(value))
            {
                return false;
            }
            memcpy(&value, payload, sizeof(value));
            if (!FiniteF(value) || value < 0.0f)
            {
                return false;
            }
            m3World_SetHitEventThreshold(worldId, value);
            break;
        }
        case M3_OP_ENABLE_SHAPE_HIT_EVENTS:
        {
            m3OpShapeFlag rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            rec.id.world0 = w->worldIndex0;
            int32_t sIndex;
            if (ResolveShape(rec.id, &sIndex) != w)
            {
                return false;
            }
            m3Shape_EnableHitEvents(rec.id, rec.on != 0);
            break;
        }
        case M3_OP_SET_SHAPE_FRICTION:
        case M3_OP_SET_SHAPE_RESTITUTION:
        case M3_OP_SET_SHAPE_ROLLING:
        {
            m3OpShapeScalar rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            rec.id.world0 = w->worldIndex0;
            int32_t sIndex;
            if (ResolveShape(rec.id, &sIndex) != w || !FiniteF(rec.value) || rec.value < 0.0f)
            {
                return false;
            }
            SetShapeScalar(rec.id, op, rec.value);
            break;
        }
        case M3_OP_SET_SHAPE_DENSITY:
        {
            m3OpShapeDensity rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            rec.id.world0 = w->worldIndex0;
            int32_t sIndex;
            if (ResolveShape(rec.id, &sIndex) != w || !FiniteF(rec.value) || !(rec.value > 0.0f))
            {
                return false;
            }
            m3Shape_SetDensity(rec.id, rec.value, rec.updateMass != 0);
            break;
        }
        case M3_OP_ENABLE_SHAPE_PRE_SOLVE:
        {
            m3OpShapeFlag rec;
            if (bytes != (int32_t)sizeof(rec))
            {
                return false;
            }
            memcpy(&rec, payload, sizeof(rec));
            rec.id.world0 = w->worldIndex0;
            int32_t sIndex;
            if (ResolveShape(rec.id, &sIndex) != w)
            {
                return false;
            }
            m3Shape_EnablePreSolve(rec.id, rec.on != 0);
            break;
        }
        case M3_OP_STEP_VETOES:
        {
            // The next step's recorded vetoes (src/world.c:1613-1621).
            if (bytes <= 0 || (bytes % 8) != 0 || bytes / 8 > w->pairCapacity)
            {
                return false;
            }
            memcpy(w->replayVetoKeys, payload, (size_t)bytes);
            w->replayVetoCount = bytes / 8;
            break;
        }
        default:
            return false; // an op outside this build's surface
        }
    }
    return true;
}

// src/world.c:3187-3220: atomic; a refused journal leaves the world as it was
// (a snapshot taken first is restored).
bool m3World_JournalReplay(m3WorldId worldId, const void* data, int32_t size)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || data == NULL || size < 0)
    {
        return false;
    }
    int32_t snapBytes = m3World_SnapshotSize(worldId);
    void* snap = snapBytes > 0 ? malloc((size_t)snapBytes) : NULL;
    if (snap == NULL || m3World_Snapshot(worldId, snap, snapBytes) != snapBytes)
    {
        free(snap);
        return false;
    }
    // Replay does not record into a journal that is being written.
    const uint8_t active = w->journalActive;
    w->journalActive = 0;
    bool ok = JournalApply(w, worldId, (const uint8_t*)data, size);
    w->journalActive = active;
    if (!ok)
    {
        bool restored = m3World_Restore(worldId, snap, snapBytes);
        (void)restored;
    }
    free(snap);
    return ok;
}

uint64_t m3World_Hash(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || !SyncHost(w))
    {
        return 0; // a faulted world has no current state to hash
    }
    uint64_t h = M3_HASH_INIT;
    h = m3HashBytes(h, &w->stepCount, 8);
    h = m3HashBytes(h, &w->gravity, (int32_t)sizeof(m3Vec3));
    if (w->contactHertz != M3_CONTACT_HERTZ_DEFAULT ||
        w->contactDampingRatio != M3_CONTACT_DAMPING_RATIO_DEFAULT ||
        w->contactPushMaxSpeed != M3_CONTACT_PUSH_MAX_SPEED_DEFAULT ||
        w->restitutionThreshold != M3_RESTITUTION_THRESHOLD_DEFAULT ||
        w->maximumLinearSpeed != M3_MAX_LINEAR_SPEED_DEFAULT || w->sleepEnabled == 0 ||
        w->continuousEnabled == 0 || w->hitEventThreshold != M3_HIT_EVENT_THRESHOLD_DEFAULT)
    {
        h = m3HashBytes(h, &w->contactHertz, 4);
        h = m3HashBytes(h, &w->contactDampingRatio, 4);
        h = m3HashBytes(h, &w->contactPushMaxSpeed, 4);
        h = m3HashBytes(h, &w->restitutionThreshold, 4);
        h = m3HashBytes(h, &w->maximumLinearSpeed, 4);
        h = m3HashBytes(h, &w->sleepEnabled, 1);
        h = m3HashBytes(h, &w->continuousEnabled, 1);
        h = m3HashBytes(h, &w->hitEventThreshold, 4);
    }
    if (w->maximumAngularSpeed != M3_MAX_ANGULAR_SPEED_DEFAULT)
    {
        h = m3HashBytes(h, &w->maximumAngularSpeed, 4);
    }
    for (int32_t i = 0; i < w->bodyPool.maxIndex; ++i)
    {
        uint8_t alive = w->bodyPool.alive[i];
        h = m3HashBytes(h, &alive, 1);
        if (alive == 0)
        {
            continue;
        }
        h = m3HashBytes(h, &w->transforms[i], (int32_t)sizeof(m3Transform));
        h = m3HashBytes(h, &w->linearVelocities[i], (int32_t)sizeof(m3Vec3));
        if (w->bodyForce[i].x != 0.0f || w->bodyForce[i].y != 0.0f || w->bodyForce[i].z != 0.0f ||
            w->bodyTorque[i].x != 0.0f || w->bodyTorque[i].y != 0.0f ||
            w->bodyTorque[i].z != 0.0f)
        {
            h = m3HashBytes(h, &w->bodyForce[i], (int32_t)sizeof(m3Vec3));
            h = m3HashBytes(h, &w->bodyTorque[i], (int32_t)sizeof(m3Vec3));
        }
        if (w->bodyEnabled[i] == 0 || w->bodyLocks[i] != 0 ||
            w->bodySleepThreshold[i] != M3_SLEEP_VELOCITY_DEFAULT || w->bodyCanSleep[i] == 0 ||
            w->bodyHasTarget[i] != 0)
        {
            h = m3HashBytes(h, &w->bodyEnabled[i], 1);
            h = m3HashBytes(h, &w->bodyLocks[i], 1);
            h = m3HashBytes(h, &w->bodySleepThreshold[i], 4);
            h = m3HashBytes(h, &w->bodyCanSleep[i], 1);
            h = m3HashBytes(h, &w->bodyHasTarget[i], 1);
            h = m3HashBytes(h, &w->bodyTarget[i], (int32_t)sizeof(m3Transform));
        }
        h = m3HashBytes(h, &w->angularVelocities[i], (int32_t)sizeof(m3Vec3));
        h = m3HashBytes(h, &w->invMass[i], 4);
        h = m3HashBytes(h, &w->invInertiaLocal[i], (int32_t)sizeof(m3Mat3));
        h = m3HashBytes(h, &w->localCenters[i], (int32_t)sizeof(m3Vec3));
        h = m3HashBytes(h, &w->types[i], 1);
        h = m3HashBytes(h, &w->bulletFlags[i], 1);
        h = m3HashBytes(h, &w->awake[i], 1);
        h = m3HashBytes(h, &w->sleepTimes[i], 4);
    }
    for (int32_t i = 0; i < w->shapePool.maxIndex; ++i)
    {
        uint8_t alive = w->shapePool.alive[i];
        h = m3HashBytes(h, &alive, 1);
        if (alive == 0)
        {
            continue;
        }
        h = m3HashBytes(h, &w->shapeBody[i], 4);
        h = m3HashBytes(h, &w->shapeType[i], 1);
        h = m3HashBytes(h, &w->shapeGeom[i], (int32_t)sizeof(m3ShapeGeom));
        h = m3HashBytes(h, &w->shapeDensity[i], 4);
        h = m3HashBytes(h, &w->shapeFriction[i], 4);
        h = m3HashBytes(h, &w->shapeRestitution[i], 4);
        if (w->shapeHitEvents[i] != 0 || w->shapePreSolve[i] != 0)
        {
            h = m3HashBytes(h, &w->shapeHitEvents[i], 1);
            h = m3HashBytes(h, &w->shapePreSolve[i], 1);
        }
        if (w->shapeSurfaceVel[i].x != 0.0f || w->shapeSurfaceVel[i].y != 0.0f ||
            w->shapeSurfaceVel[i].z != 0.0f)
        {
            h = m3HashBytes(h, &w->shapeSurfaceVel[i], (int32_t)sizeof(m3Vec3));
        }
        if (w->shapeHasOffset[i] != 0)
        {
            h = m3HashBytes(h, &w->shapeLocalPos[i], (int32_t)sizeof(m3Vec3));
            h = m3HashBytes(h, &w->shapeLocalRot[i], (int32_t)sizeof(m3Quat));
        }
        if (w->shapeRollingResistance[i] != 0.0f)
        {
            h = m3HashBytes(h, &w->shapeRollingResistance[i], 4);
        }
        if (w->shapeCategory[i] != 1ull || w->shapeMask[i] != ~0ull || w->shapeGroup[i] != 0)
        {
            h = m3HashBytes(h, &w->shapeCategory[i], 8);
            h = m3HashBytes(h, &w->shapeMask[i], 8);
            h = m3HashBytes(h, &w->shapeGroup[i], 4);
        }
        h = m3HashBytes(h, &w->shapeHullIndex[i], 4);
        h = m3HashBytes(h, &w->shapeMeshIndex[i], 4);
        h = m3HashBytes(h, &w->shapeSensor[i], 1);
    }
    // Character, soft-body, vehicle and voxel pools are empty here, so
    // their blocks contribute nothing.  Joints (src/snapshot.c:1300-1357;
    // the generic and pulley extensions never apply: only spherical joints
    // exist here).
    for (int32_t i = 0; i < w->jointPool.maxIndex; ++i)
    {
        uint8_t alive = w->jointPool.alive[i];
        h = m3HashBytes(h, &alive, 1);
        if (alive == 0)
        {
            continue;
        }
        h = m3HashBytes(h, &w->jointType[i], 1);
        h = m3HashBytes(h, &w->jointBodyA[i], 4);
        h = m3HashBytes(h, &w->jointBodyB[i], 4);
        h = m3HashBytes(h, &w->jointLocalA[i], (int32_t)sizeof(m3Vec3));
        h = m3HashBytes(h, &w->jointLocalB[i], (int32_t)sizeof(m3Vec3));
        h = m3HashBytes(h, &w->jointImpulse[i], (int32_t)sizeof(m3Vec3));
        if (w->jointBreak[i].x != 0.0f || w->jointBreak[i].y != 0.0f)
        {
            h = m3HashBytes(h, &w->jointBreak[i], (int32_t)sizeof(m3Vec3));
        }
        if ((w->jointFlags[i] & 8) != 0 || w->jointTargetScalar[i] != 0.0f ||
            w->jointTargetQ[i].x != 0.0f || w->jointTargetQ[i].y != 0.0f ||
            w->jointTargetQ[i].z != 0.0f || w->jointTargetQ[i].w != 1.0f ||
            w->jointSpringImpulse[i].x != 0.0f || w->jointSpringImpulse[i].y != 0.0f ||
            w->jointSpringImpulse[i].z != 0.0f)
        {
            h = m3HashBytes(h, &w->jointSpring[i], (int32_t)sizeof(m3Vec3));
            h = m3HashBytes(h, &w->jointTargetScalar[i], 4);
            h = m3HashBytes(h, &w->jointTargetQ[i], (int32_t)sizeof(m3Quat));
            h = m3HashBytes(h, &w->jointSpringImpulse[i], (int32_t)sizeof(m3Vec3));
        }
        h = m3HashBytes(h, &w->jointPerpImpulse[i], (int32_t)sizeof(m3Vec3));
        h = m3HashBytes(h, &w->jointLimitImpulse[i], (int32_t)sizeof(m3Vec3));
        h = m3HashBytes(h, &w->jointAngularImpulse[i], (int32_t)sizeof(m3Vec3));
        h = m3HashBytes(h, &w->jointFrameQA[i], (int32_t)sizeof(m3Quat));
        h = m3HashBytes(h, &w->jointFrameQB[i], (int32_t)sizeof(m3Quat));
        h = m3HashBytes(h, &w->jointFlags[i], 1);
    }
    h = m3HashBytes(h, &w->pairCount, 4);
    for (int32_t i = 0; i < w->pairCount; ++i)
    {
        h = m3HashBytes(h, &w->pairKeys[i], 8);
        h = m3HashBytes(h, &w->manifolds[i], (int32_t)sizeof(m3Manifold));
    }
    return h;
}

// ---------------------------------------------------------------------------
// Counters and profile
// ---------------------------------------------------------------------------

static void PullStats(m3World* w)
{
    if (w->dev == NULL)
    {
        return;
    }
    m3gpuStepStats st;
    if (m3gpu_last_stats(w->dev, &st) != 0)
    {
        return;
    }
    w->lastIslandCount = st.islandCount;
    w->lastColorCount = st.colorCount;
    w->lastAwakeCount = st.awakeCount;
    w->lastContactCount = st.constraintCount;
    memset(&w->profile, 0, sizeof(w->profile));
    w->profile.step = st.msStep;
    w->profile.broadphase = st.msBroadphase;
    w->profile.narrowphase = st.msNarrowphase;
    w->profile.prepare = st.msPrepare;
    w->profile.solve = st.msSolve;
    w->profile.continuous = st.msContinuous;
    w->profile.sleep = st.msSleep;
}

m3Counters m3World_GetCounters(m3WorldId worldId)
{
    m3Counters c;
    memset(&c, 0, sizeof(c));
    m3World* w = WorldFromId(worldId);
    if (w == NULL)
    {
        return c;
    }
    SyncHost(w);
    PullStats(w);
    for (int32_t i = 0; i < w->bodyPool.maxIndex; ++i)
    {
        c.bodyCount += w->bodyPool.alive[i];
        if (w->bodyPool.alive[i] != 0 && w->types[i] == M3_DYNAMIC && w->awake[i] != 0)
        {
            c.awakeCount += 1;
        }
    }
    for (int32_t i = 0; i < w->shapePool.maxIndex; ++i)
    {
        c.shapeCount += w->shapePool.alive[i];
    }
    for (int32_t i = 0; i < w->hullPool.maxIndex; ++i)
    {
        c.hullCount += w->hullPool.alive[i];
    }
    for (int32_t i = 0; i < w->pairCount; ++i)
    {
        c.contactCount += w->manifolds[i].pointCount > 0 ? 1 : 0;
    }
    for (int32_t i = 0; i < w->jointPool.maxIndex; ++i)
    {
        c.jointCount += w->jointPool.alive[i];
    }
    c.islandCount = w->lastIslandCount;
    c.colorCount = w->lastColorCount;
    // The reference's tree height (src/world.c:949-950) while the host tree
    // is current; -1 while it is stale (steps taken without the mirror).
    c.treeHeight = w->treeStale != 0 ? -1 : (w->tree.root != M3_TREE_NULL ? m3TreeHeight(&w->tree) : 0);
    c.snapshotBytes = m3World_SnapshotSize(worldId);
    return c;
}

m3Profile m3World_GetProfile(m3WorldId worldId)
{
    m3Profile p;
    memset(&p, 0, sizeof(p));
    m3World* w = WorldFromId(worldId);
    if (w == NULL)
    {
        return p;
    }
    PullStats(w);
    return w->profile;
}

// ---------------------------------------------------------------------------
// Snapshot / restore in the reference's layout (src/snapshot.c:25-929, format
// version 54): a 192-byte header, the persistent arrays as fixed blocks in
// the reference's order, then the hull, mesh and heightfield content sized by
// counts.  A snapshot written here restores in the reference and the other
// way round.  The blocks of subsystems this build does not simulate (voxels,
// characters, vehicles, soft bodies, meshes, heightfields, water, wind, body
// names) are written in the state a world that never used them holds, and a
// restore refuses a snapshot whose such blocks say otherwise.  The tree
// block is the host tree (exact while the mirror runs, else the canonical
// rebuild of the current leaves).  The pair table block holds the whole
// capacity: past pairCount the reference keeps what earlier steps wrote
// there; this build keeps that tail only while the mirror runs from world
// creation (on the device, m3gpu pair table), and writes zeros otherwise.
// ---------------------------------------------------------------------------

#define M3_SNAP_MAGIC 0x4D33534Eu // 'M3SN'
#define M3_SNAP_VERSION 54u
#define M3_SOLVER_REV 21 // src/world_internal.h:47
#define M3_BODY_NAME_CAPACITY 32
#define M3_MAX_WATER_VOLUMES 8
#define M3_VEHICLE_MAX_WHEELS 8
#define M3_DRIVETRAIN_MAX_CURVE 8
#define M3_DRIVETRAIN_MAX_GEARS 6
#define M3_SOFTBODY_MAX_PARTICLES 512
#define M3_SOFTBODY_MAX_EDGES 5120
#define M3_SOFTBODY_MAX_ANCHORS 32
#define M3_SOFTBODY_MAX_TETS 1024
#define M3_VOXEL_CHUNK_BYTES (8 + 512 + 8192 + 4096)

typedef struct m3SnapHeader // src/snapshot.c:84-130
{
    uint32_t magic;
    uint32_t formatVersion;
    uint64_t configHash;
    uint64_t stepCount;
    int32_t bodyCapacity;
    int32_t maxIndex;
    int32_t freeHead;
    int32_t freeCount;
    int32_t retiredCount;
    int32_t shapeCapacity;
    int32_t shapeMaxIndex;
    int32_t shapeFreeHead;
    int32_t shapeFreeCount;
    int32_t shapeRetiredCount;
    int32_t pairCount;
    int32_t treeRoot;
    int32_t treeFreeList;
    int32_t hullMaxIndex;
    int32_t hullFreeHead;
    int32_t hullFreeCount;
    int32_t hullRetiredCount;
    int32_t meshCapacity;
    int32_t meshMaxIndex;
    int32_t meshFreeHead;
    int32_t meshFreeCount;
    int32_t meshRetiredCount;
    int32_t jointCapacity;
    int32_t jointMaxIndex;
    int32_t jointFreeHead;
    int32_t jointFreeCount;
    int32_t jointRetiredCount;
    int32_t voxelCapacity;
    int32_t voxelMaxIndex;
    int32_t voxelFreeHead;
    int32_t voxelFreeCount;
    int32_t voxelRetiredCount;
    int32_t charCapacity;
    int32_t charMaxIndex;
    int32_t charFreeHead;
    int32_t charFreeCount;
    int32_t charRetiredCount;
    int32_t reserved[2];
    m3Vec3 gravity;
} m3SnapHeader;

_Static_assert(sizeof(m3SnapHeader) == 192, "snapshot header layout");

// What the bytes mean: engine version, solver revision, precision and FP
// policy (src/snapshot.c:132-148).
static uint64_t SnapConfigHash(void)
{
    uint64_t h = M3_HASH_INIT;
    const int32_t version = m3GetVersion();
    const int32_t solverRev = M3_SOLVER_REV;
    const int32_t realSize = (int32_t)sizeof(float);
    const int32_t posSize = (int32_t)sizeof(double);
    static const char policy[] = "contract-off;no-fast-math;no-fuse-4wide";
    h = m3HashBytes(h, &version, 4);
    h = m3HashBytes(h, &solverRev, 4);
    h = m3HashBytes(h, &realSize, 4);
    h = m3HashBytes(h, &posSize, 4);
    h = m3HashBytes(h, policy, (int32_t)(sizeof(policy) - 1));
    return h;
}

// One walk serves four passes: measure (no buffers), write (out), verify
// (in, dry: bounds and the fixed contents of unsimulated blocks are
// checked, nothing is written) and read (in).
typedef struct m3Walker
{
    uint8_t* out;
    const uint8_t* in;
    int dry;
    int64_t cursor;
    int64_t limit;
    int failed;
    // payload offsets of the tree nodes and proxy ids (found by walking)
    const void* treeField;
    const void* proxyField;
    int64_t treeAt;
    int64_t proxyAt;
} m3Walker;

static bool WalkRoom(m3Walker* k, int64_t bytes)
{
    if (k->failed)
    {
        return false;
    }
    if (k->limit >= 0 && k->cursor + bytes > k->limit)
    {
        k->failed = 1;
        return false;
    }
    return true;
}

static void Walk(m3Walker* k, void* field, int64_t bytes)
{
    if (bytes <= 0 || !WalkRoom(k, bytes))
    {
        return;
    }
    if (field == k->treeField)
    {
        k->treeAt = k->cursor;
    }
    if (field == k->proxyField)
    {
        k->proxyAt = k->cursor;
    }
    if (k->out != NULL)
    {
        memcpy(k->out + k->cursor, field, (size_t)bytes);
    }
    else if (k->in != NULL && !k->dry)
    {
        memcpy(field, k->in + k->cursor, (size_t)bytes);
    }
    k->cursor += bytes;
}

// A block of `count` copies of a `unit`-byte value (unit at most 16; zero
// when value is null): written as such, and on restore required to be
// exactly that.
static void WalkFixed(m3Walker* k, int64_t count, int32_t unit, const void* value)
{
    const int64_t bytes = count * unit;
    if (bytes <= 0 || !WalkRoom(k, bytes))
    {
        return;
    }
    static const uint8_t zeros[16] = {0};
    const uint8_t* v = value != NULL ? (const uint8_t*)value : zeros;
    if (k->out != NULL)
    {
        for (int64_t i = 0; i < count; ++i)
        {
            memcpy(k->out + k->cursor + i * unit, v, (size_t)unit);
        }
    }
    else if (k->in != NULL)
    {
        for (int64_t i = 0; i < count; ++i)
        {
            if (memcmp(k->in + k->cursor + i * unit, v, (size_t)unit) != 0)
            {
                k->failed = 2; // state this build does not simulate
                return;
            }
        }
    }
    k->cursor += bytes;
}

static const int32_t kMinusOne = -1;

// An id pool's arrays (src/snapshot.c:216-219); the cursors ride the header.
static void WalkPoolArrays(m3Walker* k, m3IdPool* p)
{
    Walk(k, p->generations, (int64_t)p->capacity * 2);
    Walk(k, p->alive, p->capacity);
    Walk(k, p->freeQueue, (int64_t)p->capacity * 4);
}

// An empty pool of a subsystem this build does not simulate.
static void WalkEmptyPool(m3Walker* k, int32_t capacity)
{
    WalkFixed(k, capacity, 2, NULL);
    WalkFixed(k, capacity, 1, NULL);
    WalkFixed(k, capacity, 4, NULL);
}

// The fixed blocks (src/snapshot.c:181-432), in the reference's order.
static void WalkBlocks(m3Walker* k, m3World* w)
{
    const int64_t cap = w->bodyCapacity;
    const int64_t sc = w->shapeCapacity;
    const int64_t jc = w->jointCapacity;
    const int64_t pc = w->pairCapacity;
    const int64_t cc = w->characterCapacity;
    const int64_t vc = w->vehicleCapacity;
    const int64_t sbc = w->softBodyCapacity;
    const int64_t W = M3_VEHICLE_MAX_WHEELS;

    Walk(k, w->transforms, cap * (int64_t)sizeof(m3Transform));
    Walk(k, w->linearVelocities, cap * 12);
    Walk(k, w->angularVelocities, cap * 12);
    Walk(k, w->invMass, cap * 4);
    Walk(k, w->invInertiaLocal, cap * 36);
    Walk(k, w->inertiaLocal, cap * 36);
    Walk(k, w->localCenters, cap * 12);
    Walk(k, w->gravityScales, cap * 4);
    Walk(k, w->linearDamping, cap * 4);
    Walk(k, w->angularDamping, cap * 4);
    Walk(k, w->types, cap);
    Walk(k, w->awake, cap);
    Walk(k, w->sleepTimes, cap * 4);
    Walk(k, w->bulletFlags, cap);
    Walk(k, w->minExtents, cap * 4);
    Walk(k, w->maxExtents, cap * 4);
    Walk(k, w->userData, cap * 8);
    WalkFixed(k, cap * M3_BODY_NAME_CAPACITY, 1, NULL); // debug names
    Walk(k, w->bodyForce, cap * 12);
    Walk(k, w->bodyTorque, cap * 12);
    Walk(k, w->bodyEnabled, cap);
    Walk(k, w->bodyLocks, cap);
    Walk(k, w->bodySleepThreshold, cap * 4);
    Walk(k, w->bodyCanSleep, cap);
    Walk(k, w->bodyHasTarget, cap);
    Walk(k, w->bodyTarget, cap * (int64_t)sizeof(m3Transform));
    Walk(k, &w->contactHertz, 4);
    Walk(k, &w->contactDampingRatio, 4);
    Walk(k, &w->contactPushMaxSpeed, 4);
    Walk(k, &w->restitutionThreshold, 4);
    Walk(k, &w->maximumLinearSpeed, 4);
    Walk(k, &w->maximumAngularSpeed, 4);
    Walk(k, &w->sleepEnabled, 1);
    Walk(k, &w->continuousEnabled, 1);
    Walk(k, &w->hitEventThreshold, 4);
    WalkFixed(k, 1, 12, NULL); // wind direction
    WalkFixed(k, 4, 4, NULL);  // wind speed, gust hertz, gust scale, phase
    WalkPoolArrays(k, &w->bodyPool);
    Walk(k, w->bodyShapeHead, cap * 4);

    Walk(k, w->shapeBody, sc * 4);
    Walk(k, w->shapeType, sc);
    Walk(k, w->shapeGeom, sc * (int64_t)sizeof(m3ShapeGeom));
    Walk(k, w->shapeDensity, sc * 4);
    Walk(k, w->shapeFriction, sc * 4);
    Walk(k, w->shapeRestitution, sc * 4);
    Walk(k, w->shapeUserData, sc * 8);
    Walk(k, w->shapeNext, sc * 4);
    WalkPoolArrays(k, &w->shapePool);

    Walk(k, w->pairKeys, pc * 8);
    Walk(k, w->proxyIds, sc * 4);
    Walk(k, w->tree.nodes, (int64_t)w->tree.capacity * (int64_t)sizeof(m3TreeNode));
    Walk(k, w->shapeHullIndex, sc * 4);
    Walk(k, w->hullRefCounts, sc * 4);
    WalkPoolArrays(k, &w->hullPool);
    Walk(k, w->shapeMeshIndex, sc * 4);
    Walk(k, w->shapeSensor, sc);
    Walk(k, w->shapeRollingResistance, sc * 4);
    Walk(k, w->shapeHitEvents, sc);
    Walk(k, w->shapePreSolve, sc);
    Walk(k, w->shapeSurfaceVel, sc * 12);
    Walk(k, w->shapeLocalPos, sc * 12);
    Walk(k, w->shapeLocalRot, sc * 16);
    Walk(k, w->shapeHasOffset, sc);
    Walk(k, w->shapeCategory, sc * 8);
    Walk(k, w->shapeMask, sc * 8);
    Walk(k, w->shapeGroup, sc * 4);

    // Voxels and heightfields (none on this build).
    WalkFixed(k, (int64_t)w->voxelCapacity * M3_VOXEL_CHUNK_BYTES, 1, NULL);
    WalkFixed(k, w->voxelCapacity, 4, NULL);
    WalkEmptyPool(k, w->voxelCapacity);
    WalkFixed(k, sc, 4, &kMinusOne); // shapeVoxelIndex
    WalkFixed(k, sc, 4, &kMinusOne); // shapeHfIndex
    WalkFixed(k, sc, 4, NULL);       // heightfield ref counts
    WalkEmptyPool(k, (int32_t)sc);

    // Characters (none): charBody is -1, every other block zero.
    WalkFixed(k, cc, 4, &kMinusOne);
    WalkFixed(k, cc * 6, 4, NULL);   // radius, half height, cos slope, snap, skin, step
    WalkFixed(k, cc, 1, NULL);       // grounded
    WalkFixed(k, cc, 12, NULL);      // ground normal
    WalkFixed(k, cc * 3, 4, NULL);   // mass, push max, ground body
    WalkFixed(k, cc, 2, NULL);       // ground generation
    WalkEmptyPool(k, (int32_t)cc);

    // Vehicles (none): the chassis slot is -1, every other block zero.
    WalkFixed(k, vc, 4, &kMinusOne);
    const int64_t vehicleZeroBytes =
        vc * (2 + 4 + 4 + 4 + 4 + 8) + vc * W * (12 + 12 + 4 + 4 + 4 + 4 + 4 + 1 + 4) +
        vc * (1 + 4 + 4 + 4) + vc * W * (4 + 1) + vc * (4 + 4 + 4 + 4) + vc * W * 4 +
        vc * (1 + 4) + vc * M3_DRIVETRAIN_MAX_CURVE * 8 + vc * 4 + vc * M3_DRIVETRAIN_MAX_GEARS * 4 +
        vc * (4 + 4 + 4 + 4) + vc * W * 4 + vc * (4 + 4 + 4 + 1 + 1 + 4 + 4);
    WalkFixed(k, vehicleZeroBytes, 1, NULL);
    WalkEmptyPool(k, (int32_t)vc);

    // Soft bodies (none).
    const int64_t P = M3_SOFTBODY_MAX_PARTICLES;
    const int64_t E = M3_SOFTBODY_MAX_EDGES;
    const int64_t A = M3_SOFTBODY_MAX_ANCHORS;
    const int64_t T = M3_SOFTBODY_MAX_TETS;
    const int64_t softBytes = sbc * (4 + 4 + 4 + 4 + 4 + 2 + 2 + 2 + 4 + 4 + 4) + sbc * T * (2 * 4 + 4) +
                              sbc * P * 24 + sbc * (4 + 4 + 4 + 8) + sbc * P * (24 + 24 + 4 + 12) +
                              sbc * E * (2 + 2 + 4) + sbc * 4 + sbc * A * (4 + 4 + 2 + 12) + sbc * 4 +
                              sbc * A * (4 + 4 + 2 + 4);
    WalkFixed(k, softBytes, 1, NULL);
    WalkEmptyPool(k, (int32_t)sbc);

    // Meshes (none).
    WalkFixed(k, w->meshCapacity, 4, NULL);
    WalkEmptyPool(k, w->meshCapacity);

    // Joints.
    Walk(k, w->jointType, jc);
    Walk(k, w->jointBodyA, jc * 4);
    Walk(k, w->jointBodyB, jc * 4);
    Walk(k, w->jointLocalA, jc * 12);
    Walk(k, w->jointLocalB, jc * 12);
    Walk(k, w->jointCollide, jc);
    Walk(k, w->jointImpulse, jc * 12);
    Walk(k, w->jointPerpImpulse, jc * 12);
    Walk(k, w->jointLimitImpulse, jc * 12);
    Walk(k, w->jointAngularImpulse, jc * 12);
    Walk(k, w->jointFrameQA, jc * 16);
    Walk(k, w->jointFrameQB, jc * 16);
    Walk(k, w->jointFlags, jc);
    Walk(k, w->jointMotor, jc * 12);
    Walk(k, w->jointBreak, jc * 12);
    Walk(k, w->jointSpring, jc * 12);
    Walk(k, w->jointTargetScalar, jc * 4);
    Walk(k, w->jointTargetQ, jc * 16);
    Walk(k, w->jointSpringImpulse, jc * 12);
    Walk(k, w->jointLimits, jc * 12);
    // Generic-joint modes and limits and pulley ground anchors: zero for
    // the spherical joints this build simulates (src/joint.c:240-280).
    WalkFixed(k, jc * (2 + 4 * 12 + 2 * 24), 1, NULL);

    // Water volumes (none).
    WalkFixed(k, (int64_t)M3_MAX_WATER_VOLUMES * (24 + 24 + 4 + 4 + 4 + 12), 1, NULL);
    WalkEmptyPool(k, M3_MAX_WATER_VOLUMES);

    Walk(k, w->jointNextA, jc * 4);
    Walk(k, w->jointNextB, jc * 4);
    Walk(k, w->bodyJointHead, cap * 4);
    WalkPoolArrays(k, &w->jointPool);
    Walk(k, w->manifolds, pc * (int64_t)sizeof(m3Manifold));
}

// Content tails (src/snapshot.c:434-578): every mesh slot's two counts (all
// zero here), every hull slot's four counts and used prefixes, every
// heightfield slot's three words (all zero here).  In the verify pass the
// hull counts come from the buffer and are checked against the caps.
static void WalkTails(m3Walker* k, m3World* w)
{
    WalkFixed(k, (int64_t)w->meshCapacity * 2, 4, NULL);
    for (int32_t h = 0; h < w->shapeCapacity && !k->failed; ++h)
    {
        m3HullData* hull = &w->hullData[h];
        int32_t counts[4] = {hull->vertexCount, hull->faceCount, hull->indexCount, hull->edgeCount};
        if (k->in != NULL)
        {
            if (!WalkRoom(k, 16))
            {
                return;
            }
            memcpy(counts, k->in + k->cursor, 16);
            if (counts[0] < 0 || counts[0] > M3_HULL_MAX_VERTS || counts[1] < 0 ||
                counts[1] > M3_HULL_MAX_FACES || counts[2] < 0 || counts[2] > M3_HULL_MAX_FACE_INDICES ||
                counts[3] < 0 || counts[3] > M3_HULL_MAX_HALF_EDGES ||
                (counts[0] == 0 && (counts[1] | counts[2] | counts[3]) != 0))
            {
                k->failed = 1;
                return;
            }
            if (!k->dry)
            {
                memset(hull, 0, sizeof(*hull));
                hull->vertexCount = counts[0];
                hull->faceCount = counts[1];
                hull->indexCount = counts[2];
                hull->edgeCount = counts[3];
            }
            k->cursor += 16;
        }
        else
        {
            Walk(k, &hull->vertexCount, 4);
            Walk(k, &hull->faceCount, 4);
            Walk(k, &hull->indexCount, 4);
            Walk(k, &hull->edgeCount, 4);
        }
        if (counts[0] > 0)
        {
            Walk(k, &hull->unitMass, 4);
            Walk(k, &hull->unitCom, 12);
            Walk(k, &hull->unitInertiaCom, 36);
            Walk(k, &hull->center, 12);
            Walk(k, hull->vertices, (int64_t)counts[0] * 12);
            Walk(k, hull->faceNormals, (int64_t)counts[1] * 12);
            Walk(k, hull->faceOffsets, (int64_t)counts[1] * 4);
            Walk(k, hull->faceVertCounts, counts[1]);
            Walk(k, hull->faceVertStart, (int64_t)counts[1] * 2);
            Walk(k, hull->faceIndices, counts[2]);
            Walk(k, hull->edges, (int64_t)counts[3] * (int64_t)sizeof(m3HullHalfEdge));
        }
    }
    WalkFixed(k, (int64_t)w->shapeCapacity * 3, 4, NULL);
}

static void PoolToHeader(const m3IdPool* p, int32_t* maxIndex, int32_t* freeHead, int32_t* freeCount,
                         int32_t* retired)
{
    *maxIndex = p->maxIndex;
    *freeHead = p->freeHead;
    *freeCount = p->freeCount;
    *retired = p->retiredCount;
}

static bool PoolHeaderSane(int32_t maxIndex, int32_t freeHead, int32_t freeCount, int32_t retired,
                           int32_t cap) // src/snapshot.c:704-707
{
    return maxIndex >= 0 && maxIndex <= cap && freeHead >= -1 && freeHead < cap && freeCount >= 0 &&
           freeCount <= cap && retired >= 0 && retired <= cap;
}

// The host arrays the snapshot reads, brought up to date: device state,
// the tree (rebuilt when stale), and the pair table (the device's table
// while the mirror keeps one, else the live list and a zero tail).
static bool SnapshotPrepare(m3World* w)
{
    if (w->faulted != 0 || !SyncHost(w) || !TreeFresh(w))
    {
        return false;
    }
    if (w->pairTableLive != 0 && w->pairTableOnDevice != 0)
    {
        m3gpuHostState h;
        HostView(w, &h);
        if (m3gpu_download(w->dev, &h, M3GPU_FIELDS_PAIR_TABLE) != 0)
        {
            DeviceFail(w, "m3gpu_download");
            return false;
        }
    }
    else if (w->pairTableLive == 0)
    {
        const int32_t tail = w->pairCapacity - w->pairCount;
        memset(w->pairKeys + w->pairCount, 0, (size_t)tail * sizeof(uint64_t));
        memset(w->manifolds + w->pairCount, 0, (size_t)tail * sizeof(m3Manifold));
    }
    return true;
}

int32_t m3World_SnapshotSize(m3WorldId worldId)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || w->faulted != 0)
    {
        return -1;
    }
    m3Walker k;
    memset(&k, 0, sizeof(k));
    k.limit = -1;
    WalkBlocks(&k, w);
    WalkTails(&k, w);
    const int64_t total = (int64_t)sizeof(m3SnapHeader) + k.cursor;
    return total > 0x7FFFFFFF ? -1 : (int32_t)total;
}

int32_t m3World_Snapshot(m3WorldId worldId, void* out, int32_t capacity)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || out == NULL || !SnapshotPrepare(w))
    {
        return -1;
    }
    const int32_t size = m3World_SnapshotSize(worldId);
    if (size < 0 || capacity < size)
    {
        return -1;
    }
    m3SnapHeader hdr;
    memset(&hdr, 0, sizeof(hdr));
    hdr.magic = M3_SNAP_MAGIC;
    hdr.formatVersion = M3_SNAP_VERSION;
    hdr.configHash = SnapConfigHash();
    hdr.stepCount = w->stepCount;
    hdr.bodyCapacity = w->bodyCapacity;
    PoolToHeader(&w->bodyPool, &hdr.maxIndex, &hdr.freeHead, &hdr.freeCount, &hdr.retiredCount);
    hdr.shapeCapacity = w->shapeCapacity;
    PoolToHeader(&w->shapePool, &hdr.shapeMaxIndex, &hdr.shapeFreeHead, &hdr.shapeFreeCount,
                 &hdr.shapeRetiredCount);
    hdr.pairCount = w->pairCount;
    hdr.treeRoot = w->tree.root;
    hdr.treeFreeList = w->tree.freeList;
    PoolToHeader(&w->hullPool, &hdr.hullMaxIndex, &hdr.hullFreeHead, &hdr.hullFreeCount,
                 &hdr.hullRetiredCount);
    hdr.meshCapacity = w->meshCapacity;
    hdr.jointCapacity = w->jointCapacity;
    PoolToHeader(&w->jointPool, &hdr.jointMaxIndex, &hdr.jointFreeHead, &hdr.jointFreeCount,
                 &hdr.jointRetiredCount);
    hdr.charCapacity = w->characterCapacity;
    hdr.voxelCapacity = w->voxelCapacity;
    hdr.gravity = w->gravity;
    m3Walker k;
    memset(&k, 0, sizeof(k));
    k.out = (uint8_t*)out + sizeof(hdr);
    k.limit = size - (int64_t)sizeof(hdr);
    WalkBlocks(&k, w);
    WalkTails(&k, w);
    memcpy(out, &hdr, sizeof(hdr));
    return k.failed ? -1 : size;
}

bool m3World_Restore(m3WorldId worldId, const void* data, int32_t size)
{
    m3World* w = WorldFromId(worldId);
    if (w == NULL || w->faulted != 0 || data == NULL || size < (int32_t)sizeof(m3SnapHeader))
    {
        return false;
    }
    m3SnapHeader hdr;
    memcpy(&hdr, data, sizeof(hdr));
    if (hdr.magic != M3_SNAP_MAGIC || hdr.formatVersion != M3_SNAP_VERSION ||
        hdr.configHash != SnapConfigHash() || hdr.bodyCapacity != w->bodyCapacity ||
        hdr.shapeCapacity != w->shapeCapacity || hdr.meshCapacity != w->meshCapacity ||
        hdr.jointCapacity != w->jointCapacity || hdr.voxelCapacity != w->voxelCapacity ||
        hdr.charCapacity != w->characterCapacity)
    {
        return false;
    }
    if (!PoolHeaderSane(hdr.maxIndex, hdr.freeHead, hdr.freeCount, hdr.retiredCount, w->bodyCapacity) ||
        !PoolHeaderSane(hdr.shapeMaxIndex, hdr.shapeFreeHead, hdr.shapeFreeCount, hdr.shapeRetiredCount,
                        w->shapeCapacity) ||
        !PoolHeaderSane(hdr.hullMaxIndex, hdr.hullFreeHead, hdr.hullFreeCount, hdr.hullRetiredCount,
                        w->shapeCapacity) ||
        !PoolHeaderSane(hdr.jointMaxIndex, hdr.jointFreeHead, hdr.jointFreeCount,
                        hdr.jointRetiredCount, w->jointCapacity) ||
        hdr.pairCount < 0 || hdr.pairCount > w->pairCapacity || hdr.treeRoot < -1 ||
        hdr.treeRoot >= w->tree.capacity || hdr.treeFreeList < -1 || hdr.treeFreeList >= w->tree.capacity)
    {
        return false;
    }
    // Pools this build does not simulate must be empty.
    if ((hdr.meshMaxIndex | hdr.meshFreeHead | hdr.meshFreeCount | hdr.meshRetiredCount |
         hdr.voxelMaxIndex | hdr.voxelFreeHead | hdr.voxelFreeCount | hdr.voxelRetiredCount |
         hdr.charMaxIndex | hdr.charFreeHead | hdr.charFreeCount | hdr.charRetiredCount) != 0)
    {
        return false;
    }
    // Verify the whole payload before anything lands (restore is atomic).
    const uint8_t* payload = (const uint8_t*)data + sizeof(hdr);
    m3Walker k;
    memset(&k, 0, sizeof(k));
    k.in = payload;
    k.dry = 1;
    k.limit = size - (int64_t)sizeof(hdr);
    k.treeField = w->tree.nodes;
    k.proxyField = w->proxyIds;
    WalkBlocks(&k, w);
    WalkTails(&k, w);
    if (k.failed || k.cursor != k.limit)
    {
        return false;
    }
    // The tree must be well formed and agree with the proxy ids: every id
    // is -1 or a leaf that names its shape, and every leaf is some shape's.
    {
        m3Tree probe;
        if (!m3TreeCreate(&probe, w->tree.capacity))
        {
            return false;
        }
        memcpy(probe.nodes, payload + k.treeAt, (size_t)probe.capacity * sizeof(m3TreeNode));
        probe.root = hdr.treeRoot;
        probe.freeList = hdr.treeFreeList;
        bool ok = m3TreeValid(&probe);
        int32_t leaves = 0;
        for (int32_t i = 0; ok && i < probe.capacity; ++i)
        {
            leaves += probe.nodes[i].height == 0;
        }
        int32_t named = 0;
        for (int32_t sh = 0; ok && sh < w->shapeCapacity; ++sh)
        {
            int32_t id;
            memcpy(&id, payload + k.proxyAt + 4 * (int64_t)sh, 4);
            if (id == M3_TREE_NULL)
            {
                continue;
            }
            ok = id >= 0 && id < probe.capacity && probe.nodes[id].height == 0 &&
                 probe.nodes[id].userData == sh;
            named += 1;
        }
        ok = ok && named == leaves;
        m3TreeDestroy(&probe);
        if (!ok)
        {
            return false;
        }
    }
    // Validation passed: now write.  Pending device results are dropped,
    // and so are the event streams (src/snapshot.c:878-890).
    w->hostStale = 0;
    w->eventsState = 0;
    memset(w->eventCounts, 0, sizeof(w->eventCounts));
    memset(&k, 0, sizeof(k));
    k.in = payload;
    k.limit = size - (int64_t)sizeof(hdr);
    WalkBlocks(&k, w);
    WalkTails(&k, w);
    w->stepCount = hdr.stepCount;
    w->gravity = hdr.gravity;
    w->bodyPool.maxIndex = hdr.maxIndex;
    w->bodyPool.freeHead = hdr.freeHead;
    w->bodyPool.freeCount = hdr.freeCount;
    w->bodyPool.retiredCount = hdr.retiredCount;
    w->shapePool.maxIndex = hdr.shapeMaxIndex;
    w->shapePool.freeHead = hdr.shapeFreeHead;
    w->shapePool.freeCount = hdr.shapeFreeCount;
    w->shapePool.retiredCount = hdr.shapeRetiredCount;
    w->hullPool.maxIndex = hdr.hullMaxIndex;
    w->hullPool.freeHead = hdr.hullFreeHead;
    w->hullPool.freeCount = hdr.hullFreeCount;
    w->hullPool.retiredCount = hdr.hullRetiredCount;
    w->jointPool.maxIndex = hdr.jointMaxIndex;
    w->jointPool.freeHead = hdr.jointFreeHead;
    w->jointPool.freeCount = hdr.jointFreeCount;
    w->jointPool.retiredCount = hdr.jointRetiredCount;
    w->pairCount = hdr.pairCount;
    w->tree.root = hdr.treeRoot;
    w->tree.freeList = hdr.treeFreeList;
    w->treeStale = 0;
    // The device's leaf boxes are the tree's leaves.
    for (int32_t sh = 0; sh < w->shapeCapacity; ++sh)
    {
        const int32_t id = w->proxyIds[sh];
        if (id >= 0)
        {
            memcpy(&w->leafBoxes[6 * sh], w->tree.nodes[id].lo, 3 * sizeof(double));
            memcpy(&w->leafBoxes[6 * sh + 3], w->tree.nodes[id].hi, 3 * sizeof(double));
        }
    }
    // The frozen-pair buffer is derived from a pair list the restore just
    // replaced: rebuilt by the next step (src/snapshot.c:893-894).
    w->sleepingPairCount = 0;
    w->pairsFullQuery = 1;
    // The pair table holds the snapshot's whole capacity again.
    w->pairTableLive = 1;
    w->pairTableOnDevice = 0;
    w->devDirty = M3GPU_FIELDS_ALL;
    return true;
}
