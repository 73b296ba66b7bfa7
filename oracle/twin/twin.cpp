// Test infrastructure (the gradient checker), not part of the product.
//
// A CPU twin of this build's step: the host world (world.c, the same C
// file the product links) over a CPU implementation of the m3gpu ABI that
// runs the serial step of maul3d_amd/csrc/m3_serial.h, instantiated at the
// scalar type TWIN_REAL.  With float it must reproduce the product (and so
// the reference) bit for bit; with double it is the FP64 forward pass that
// central finite differences of the adjoint's gradients are taken on.  The
// twin also runs the adjoint (m3_adjoint.h) at its precision, so the hand
// written reverse pass can be checked against finite differences of the
// same forward without a GPU.
//
// Built by oracle/twin/Makefile into oracle/_ref/libm3twin32.so and
// libm3twin64.so; only tests load them.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include <algorithm>
#include <vector>

#include "m3gpu.h"
#include "m3_core.h" // the float records of the host world

namespace tw
{
#undef M3_CORE_H
#undef M3_COLLIDE_H
#undef M3_STEP_H
#undef M3_SERIAL_H
#undef M3_ADJOINT_H
#define M3_REAL TWIN_REAL
#include "m3_serial.h"
#include "m3_adjoint.h"
#undef M3_REAL
} // namespace tw

typedef tw::m3real R;

// ---- float records -> twin records -----------------------------------------

static tw::m3Vec3 V(const ::m3Vec3& a) { return tw::m3V3((R)a.x, (R)a.y, (R)a.z); }
static tw::m3Quat Q(const ::m3Quat& a) { return tw::m3Q((R)a.x, (R)a.y, (R)a.z, (R)a.w); }
static tw::m3Mat3 M(const ::m3Mat3& a)
{
    tw::m3Mat3 m;
    m.cx = V(a.cx);
    m.cy = V(a.cy);
    m.cz = V(a.cz);
    return m;
}
static tw::m3Transform X(const ::m3Transform& a)
{
    tw::m3Transform t;
    t.p.x = a.p.x;
    t.p.y = a.p.y;
    t.p.z = a.p.z;
    t.q = Q(a.q);
    return t;
}
static tw::m3ShapeGeom G(const ::m3ShapeGeom& a)
{
    tw::m3ShapeGeom g;
    g.v = V(a.v);
    g.s = (R)a.s;
    g.v2 = V(a.v2);
    g.s2 = (R)a.s2;
    return g;
}
static void H(tw::m3HullData* o, const ::m3HullData& a)
{
    o->vertexCount = a.vertexCount;
    o->faceCount = a.faceCount;
    o->indexCount = a.indexCount;
    o->unitMass = a.unitMass;
    o->unitCom = V(a.unitCom);
    o->unitInertiaCom = M(a.unitInertiaCom);
    for (int k = 0; k < M3_HULL_MAX_VERTS; ++k)
        o->vertices[k] = V(a.vertices[k]);
    for (int k = 0; k < M3_HULL_MAX_FACES; ++k)
    {
        o->faceNormals[k] = V(a.faceNormals[k]);
        o->faceOffsets[k] = (R)a.faceOffsets[k];
        o->faceVertCounts[k] = a.faceVertCounts[k];
        o->faceVertStart[k] = a.faceVertStart[k];
    }
    memcpy(o->faceIndices, a.faceIndices, sizeof(a.faceIndices));
    o->edgeCount = a.edgeCount;
    for (int k = 0; k < M3_HULL_MAX_HALF_EDGES; ++k)
    {
        o->edges[k].twin = a.edges[k].twin;
        o->edges[k].next = a.edges[k].next;
        o->edges[k].origin = a.edges[k].origin;
        o->edges[k].face = a.edges[k].face;
    }
    o->center = V(a.center);
}
static tw::m3Manifold Mf(const ::m3Manifold& a)
{
    tw::m3Manifold m;
    memset(&m, 0, sizeof(m));
    m.normal = V(a.normal);
    m.pointCount = a.pointCount;
    m.frictionImpulse = V(a.frictionImpulse);
    m.twistImpulse = (R)a.twistImpulse;
    m.rollingImpulse = V(a.rollingImpulse);
    for (int k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        m.points[k].anchorA = V(a.points[k].anchorA);
        m.points[k].anchorB = V(a.points[k].anchorB);
        m.points[k].separation = (R)a.points[k].separation;
        m.points[k].normalImpulse = (R)a.points[k].normalImpulse;
        m.points[k].id = a.points[k].id;
        m.points[k].flags = a.points[k].flags;
    }
    return m;
}

// ---- twin records -> float records (rounded; the float host world keeps
// its own layout) -------------------------------------------------------------

static ::m3Vec3 Vb(const tw::m3Vec3& a)
{
    ::m3Vec3 v = {(float)a.x, (float)a.y, (float)a.z};
    return v;
}
static ::m3Quat Qb(const tw::m3Quat& a)
{
    ::m3Quat q = {(float)a.x, (float)a.y, (float)a.z, (float)a.w};
    return q;
}
static ::m3Transform Xb(const tw::m3Transform& a)
{
    ::m3Transform t;
    t.p.x = a.p.x;
    t.p.y = a.p.y;
    t.p.z = a.p.z;
    t.q = Qb(a.q);
    return t;
}
static ::m3Manifold Mb(const tw::m3Manifold& a)
{
    ::m3Manifold m;
    memset(&m, 0, sizeof(m));
    m.normal = Vb(a.normal);
    m.pointCount = a.pointCount;
    m.frictionImpulse = Vb(a.frictionImpulse);
    m.twistImpulse = (float)a.twistImpulse;
    m.rollingImpulse = Vb(a.rollingImpulse);
    for (int k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        m.points[k].anchorA = Vb(a.points[k].anchorA);
        m.points[k].anchorB = Vb(a.points[k].anchorB);
        m.points[k].separation = (float)a.points[k].separation;
        m.points[k].normalImpulse = (float)a.points[k].normalImpulse;
        m.points[k].id = a.points[k].id;
        m.points[k].flags = a.points[k].flags;
    }
    return m;
}

template <typename T> static T* A(int64_t n) { return (T*)calloc(n > 0 ? n : 1, sizeof(T)); }

struct m3gpuWorld
{
    m3gpuCaps caps;
    tw::m3View v;
    std::vector<uint64_t> keysBuf[2];
    std::vector<tw::m3Manifold> mansBuf[2];
    int cur;
    std::vector<int32_t> root;
    std::vector<uint32_t> masks;
    int jointsPresent;
    int targetsPresent;
    int ccdNeedsTree; // continuous sweeps meet tree shapes (static non-plane shapes or bullets)
    int begun;        // m3gpu_step_begin ran; m3gpu_step_end finishes the step
    int beginStalled;
    int pairTable;    // keep the pair table in the reference's layout
    std::vector<m3gpuPreSolve> preCands;
    std::vector<uint8_t> vetoFlags;
    std::vector<uint64_t> tableKeys;
    std::vector<::m3Manifold> tableMans;
    tw::m3HitRec* hitRecBuf;
};

static char s_err[256];
extern "C" int m3gpu_device_count(void) { return 1; }
extern "C" const char* m3gpu_last_error(void) { return s_err; }

extern "C" m3gpuWorld* m3gpu_world_create(const m3gpuCaps* c)
{
    m3gpuWorld* w = new m3gpuWorld();
    w->caps = *c;
    tw::m3View& v = w->v;
    memset(&v, 0, sizeof(v));
    int nb = c->bodyCapacity, ns = c->shapeCapacity, np = c->pairCapacity;
    v.pairCapacity = np;
    v.emitCapacity = 2 * np;
    v.xf = A<tw::m3Transform>(nb);
    v.lv = A<tw::m3Vec3>(nb);
    v.av = A<tw::m3Vec3>(nb);
    v.invMass = A<R>(nb);
    v.invI = A<tw::m3Mat3>(nb);
    v.inertia = A<tw::m3Mat3>(nb);
    v.lc = A<tw::m3Vec3>(nb);
    v.gravityScale = A<R>(nb);
    v.linDamp = A<R>(nb);
    v.angDamp = A<R>(nb);
    v.bodyAlive = A<uint8_t>(nb);
    v.type = A<uint8_t>(nb);
    v.awake = A<uint8_t>(nb);
    v.sleepTime = A<R>(nb);
    v.bullet = A<uint8_t>(nb);
    v.minExt = A<R>(nb);
    v.maxExt = A<R>(nb);
    v.force = A<tw::m3Vec3>(nb);
    v.torque = A<tw::m3Vec3>(nb);
    v.enabled = A<uint8_t>(nb);
    v.locks = A<uint8_t>(nb);
    v.sleepThreshold = A<R>(nb);
    v.canSleep = A<uint8_t>(nb);
    v.hasTarget = A<uint8_t>(nb);
    v.target = A<tw::m3Transform>(nb);
    v.shapeHead = A<int32_t>(nb);
    v.shapeAlive = A<uint8_t>(ns);
    v.shapeBody = A<int32_t>(ns);
    v.shapeType = A<uint8_t>(ns);
    v.geom = A<tw::m3ShapeGeom>(ns);
    v.friction = A<R>(ns);
    v.restitution = A<R>(ns);
    v.rolling = A<R>(ns);
    v.sensor = A<uint8_t>(ns);
    v.hitEvents = A<uint8_t>(ns);
    w->hitRecBuf = A<tw::m3HitRec>(np);
    v.hasOffset = A<uint8_t>(ns);
    v.localPos = A<tw::m3Vec3>(ns);
    v.localRot = A<tw::m3Quat>(ns);
    v.surfaceVel = A<tw::m3Vec3>(ns);
    v.category = A<uint64_t>(ns);
    v.mask = A<uint64_t>(ns);
    v.group = A<int32_t>(ns);
    v.shapeNext = A<int32_t>(ns);
    v.hullIndex = A<int32_t>(ns);
    v.leaf = A<double>(6 * (int64_t)ns);
    v.fresh = A<double>(6 * (int64_t)ns);
    v.planes = A<int32_t>(ns);
    v.shapeRank = A<int32_t>(ns);
    v.ccdStatics = A<int32_t>(ns);
    v.preSolve = A<uint8_t>(ns);
    v.hulls = A<tw::m3HullData>(c->hullCapacity);
    for (int k = 0; k < 2; ++k)
    {
        w->keysBuf[k].resize(np);
        w->mansBuf[k].resize(np);
    }
    v.sleepKeys = A<uint64_t>(np);
    v.emit = A<uint32_t>(v.emitCapacity);
    v.cons = A<tw::m3Constraint>(np);
    v.consPair = A<int32_t>(np);
    v.colors = A<uint8_t>(np);
    v.colorLists = A<int32_t>(np);
    v.slotOf = A<int32_t>(np);
    v.movers = A<int32_t>(nb);
    v.deltaPos = A<tw::m3Vec3>(nb);
    v.deltaRot = A<tw::m3Quat>(nb);
    v.com0 = A<tw::m3Pos3>(nb);
    v.rot0 = A<tw::m3Quat>(nb);
    v.parent = A<int32_t>(nb);
    v.forced = A<uint8_t>(nb);
    v.flag = A<uint8_t>(nb);
    v.pending = A<uint8_t>(nb);
    v.sc = A<tw::m3Scalars>(1);
    w->root.resize(nb);
    w->masks.resize(nb);
    w->cur = 0;
    v.keys = w->keysBuf[0].data();
    v.manifolds = w->mansBuf[0].data();
    v.oldKeys = w->keysBuf[1].data();
    v.oldManifolds = w->mansBuf[1].data();
    return w;
}

extern "C" void m3gpu_world_destroy(m3gpuWorld* w) { delete w; }
extern "C" void m3gpu_shutdown(void) {}

extern "C" int m3gpu_upload(m3gpuWorld* w, const m3gpuHostState* h, uint32_t f)
{
    tw::m3View& v = w->v;
    v.maxBody = h->maxBody;
    v.maxShape = h->maxShape;
    int nb = h->maxBody, ns = h->maxShape;
    if (f & M3GPU_FIELDS_STATIC)
    {
        const ::m3Mat3* invI = (const ::m3Mat3*)h->invInertiaLocal;
        const ::m3Mat3* inertia = (const ::m3Mat3*)h->inertiaLocal;
        const ::m3Vec3* lc = (const ::m3Vec3*)h->localCenters;
        const ::m3Transform* target = (const ::m3Transform*)h->bodyTarget;
        for (int b = 0; b < nb; ++b)
        {
            v.invMass[b] = (R)h->invMass[b];
            v.invI[b] = M(invI[b]);
            v.inertia[b] = M(inertia[b]);
            v.lc[b] = V(lc[b]);
            v.gravityScale[b] = (R)h->gravityScales[b];
            v.linDamp[b] = (R)h->linearDamping[b];
            v.angDamp[b] = (R)h->angularDamping[b];
            v.bodyAlive[b] = h->bodyAlive[b];
            v.type[b] = h->types[b];
            v.bullet[b] = h->bulletFlags[b];
            v.minExt[b] = (R)h->minExtents[b];
            v.maxExt[b] = (R)h->maxExtents[b];
            v.enabled[b] = h->bodyEnabled[b];
            v.locks[b] = h->bodyLocks[b];
            v.sleepThreshold[b] = (R)h->bodySleepThreshold[b];
            v.canSleep[b] = h->bodyCanSleep[b];
            v.target[b] = X(target[b]);
            v.shapeHead[b] = h->bodyShapeHead[b];
        }
        const ::m3ShapeGeom* geom = (const ::m3ShapeGeom*)h->shapeGeom;
        const ::m3Vec3* lp = (const ::m3Vec3*)h->shapeLocalPos;
        const ::m3Quat* lr = (const ::m3Quat*)h->shapeLocalRot;
        const ::m3Vec3* sv = (const ::m3Vec3*)h->shapeSurfaceVel;
        int pc = 0;
        bool anyHit = false;
        for (int s = 0; s < ns; ++s)
        {
            v.shapeAlive[s] = h->shapeAlive[s];
            v.shapeBody[s] = h->shapeBody[s];
            v.shapeType[s] = h->shapeType[s];
            v.geom[s] = G(geom[s]);
            v.friction[s] = (R)h->shapeFriction[s];
            v.restitution[s] = (R)h->shapeRestitution[s];
            v.rolling[s] = (R)h->shapeRollingResistance[s];
            v.sensor[s] = h->shapeSensor[s];
            v.hitEvents[s] = h->shapeHitEvents[s];
            v.preSolve[s] = h->shapePreSolve[s];
            anyHit = anyHit || (h->shapeAlive[s] != 0 && h->shapeHitEvents[s] != 0);
            v.hasOffset[s] = h->shapeHasOffset[s];
            v.localPos[s] = V(lp[s]);
            v.localRot[s] = Q(lr[s]);
            v.surfaceVel[s] = V(sv[s]);
            v.category[s] = h->shapeCategory[s];
            v.mask[s] = h->shapeMask[s];
            v.group[s] = h->shapeGroup[s];
            v.shapeNext[s] = h->shapeNext[s];
            v.hullIndex[s] = h->shapeHullIndex[s];
            if (h->shapeAlive[s] && h->shapeType[s] == M3_PLANE)
            {
                v.planes[pc++] = s;
            }
        }
        v.planeCount = pc;
        v.hitRecs = anyHit ? w->hitRecBuf : nullptr;
        const ::m3HullData* hulls = (const ::m3HullData*)h->hullData;
        for (int k = 0; k < h->maxHull; ++k)
        {
            H((tw::m3HullData*)&v.hulls[k], hulls[k]);
        }
        w->jointsPresent = 0;
        for (int j = 0; j < h->maxJoint; ++j)
        {
            if (h->jointAlive[j])
            {
                w->jointsPresent = 1;
            }
        }
    }
    if (f & M3GPU_FIELDS_DYNAMIC)
    {
        const ::m3Transform* xf = (const ::m3Transform*)h->transforms;
        const ::m3Vec3* lv = (const ::m3Vec3*)h->linearVelocities;
        const ::m3Vec3* av = (const ::m3Vec3*)h->angularVelocities;
        const ::m3Vec3* fo = (const ::m3Vec3*)h->bodyForce;
        const ::m3Vec3* to = (const ::m3Vec3*)h->bodyTorque;
        w->targetsPresent = 0;
        w->ccdNeedsTree = 0;
        v.ccdStaticCount = 0;
        for (int s2 = 0; s2 < ns; ++s2)
        {
            if (h->shapeAlive[s2] != 0 && h->shapeType[s2] != M3_PLANE &&
                h->types[h->shapeBody[s2]] == M3_STATIC)
            {
                v.ccdStatics[v.ccdStaticCount++] = s2;
                w->ccdNeedsTree |= h->shapeSensor[s2] == 0;
            }
        }
        for (int b = 0; b < nb; ++b)
        {
            if (h->bodyAlive[b] != 0 && h->bulletFlags[b] != 0)
            {
                w->ccdNeedsTree = 1;
            }
        }
        for (int b = 0; b < nb; ++b)
        {
            v.xf[b] = X(xf[b]);
            v.lv[b] = V(lv[b]);
            v.av[b] = V(av[b]);
            v.awake[b] = h->awake[b];
            v.sleepTime[b] = (R)h->sleepTimes[b];
            v.force[b] = V(fo[b]);
            v.torque[b] = V(to[b]);
            v.hasTarget[b] = h->bodyHasTarget[b];
            w->targetsPresent |= h->bodyHasTarget[b] != 0;
        }
        v.sc->stepCount = *h->stepCount;
    }
    if (f & M3GPU_FIELDS_PAIRS)
    {
        int pc = *h->pairCount, zc = *h->sleepingPairCount;
        const ::m3Manifold* mans = (const ::m3Manifold*)h->manifolds;
        for (int i = 0; i < pc; ++i)
        {
            v.keys[i] = h->pairKeys[i];
            v.manifolds[i] = Mf(mans[i]);
        }
        v.sc->pairCount = pc;
        memcpy(v.sleepKeys, h->sleepingPairKeys, (size_t)zc * sizeof(uint64_t));
        v.sc->sleepCount = zc;
        memcpy(v.leaf, h->leafBoxes, (size_t)6 * ns * sizeof(double));
        v.sc->fullQuery = *h->pairsFullQuery != 0;
    }
    if (f & M3GPU_FIELDS_PAIR_TABLE)
    {
        const ::m3Manifold* mans = (const ::m3Manifold*)h->manifolds;
        w->tableKeys.assign(h->pairKeys, h->pairKeys + v.pairCapacity);
        w->tableMans.assign(mans, mans + v.pairCapacity);
    }
    return 0;
}

extern "C" int m3gpu_set_pair_table(m3gpuWorld* w, int32_t on)
{
    w->pairTable = on != 0;
    if (on && w->tableKeys.empty())
    {
        w->tableKeys.assign(w->v.pairCapacity, 0);
        w->tableMans.assign(w->v.pairCapacity, ::m3Manifold());
    }
    return 0;
}

extern "C" int m3gpu_download(m3gpuWorld* w, m3gpuHostState* h, uint32_t f)
{
    tw::m3View& v = w->v;
    int nb = v.maxBody, ns = v.maxShape;
    if (f & M3GPU_FIELDS_DYNAMIC)
    {
        ::m3Transform* xf = (::m3Transform*)h->transforms;
        ::m3Vec3* lv = (::m3Vec3*)h->linearVelocities;
        ::m3Vec3* av = (::m3Vec3*)h->angularVelocities;
        ::m3Vec3* fo = (::m3Vec3*)h->bodyForce;
        ::m3Vec3* to = (::m3Vec3*)h->bodyTorque;
        for (int b = 0; b < nb; ++b)
        {
            xf[b] = Xb(v.xf[b]);
            lv[b] = Vb(v.lv[b]);
            av[b] = Vb(v.av[b]);
            h->awake[b] = v.awake[b];
            h->sleepTimes[b] = (float)v.sleepTime[b];
            fo[b] = Vb(v.force[b]);
            to[b] = Vb(v.torque[b]);
            h->bodyHasTarget[b] = v.hasTarget[b];
        }
        *h->stepCount = v.sc->stepCount;
    }
    if (f & M3GPU_FIELDS_PAIRS)
    {
        *h->pairCount = v.sc->pairCount;
        *h->sleepingPairCount = v.sc->sleepCount;
        ::m3Manifold* mans = (::m3Manifold*)h->manifolds;
        for (int i = 0; i < v.sc->pairCount; ++i)
        {
            h->pairKeys[i] = v.keys[i];
            mans[i] = Mb(v.manifolds[i]);
        }
        memcpy(h->sleepingPairKeys, v.sleepKeys, (size_t)v.sc->sleepCount * sizeof(uint64_t));
        memcpy(h->leafBoxes, v.leaf, (size_t)6 * ns * sizeof(double));
        *h->frozenDirty = 0;
    }
    if (f & M3GPU_FIELDS_PAIR_TABLE)
    {
        if (w->tableKeys.empty())
        {
            snprintf(s_err, sizeof(s_err), "twin: no pair table");
            return -1;
        }
        memcpy(h->pairKeys, w->tableKeys.data(), w->tableKeys.size() * sizeof(uint64_t));
        memcpy(h->manifolds, w->tableMans.data(), w->tableMans.size() * sizeof(::m3Manifold));
    }
    return 0;
}

static tw::m3SerialParams Params(const m3gpuStepParams* p)
{
    tw::m3SerialParams s;
    s.dt = (R)p->dt;
    s.substeps = p->substeps;
    s.gravity = tw::m3V3((R)p->gravity[0], (R)p->gravity[1], (R)p->gravity[2]);
    s.contactHertz = (R)p->contactHertz;
    s.contactDampingRatio = (R)p->contactDampingRatio;
    s.contactPushMaxSpeed = (R)p->contactPushMaxSpeed;
    s.restitutionThreshold = (R)p->restitutionThreshold;
    s.maximumLinearSpeed = (R)p->maximumLinearSpeed;
    s.maximumAngularSpeed = (R)p->maximumAngularSpeed;
    s.sleepEnabled = p->sleepEnabled;
    s.continuousEnabled = p->continuousEnabled;
    return s;
}

// Through the island wake pass; nonzero on a stall (1) or a failure (-1).
static int StepA(m3gpuWorld* w, const m3gpuStepParams* p, tw::m3Tape* tape)
{
    if (w->jointsPresent || w->targetsPresent)
    {
        snprintf(s_err, sizeof(s_err), "twin: joints and kinematic targets are not supported");
        return -1;
    }
    w->cur ^= 1;
    tw::m3View& v = w->v;
    v.keys = w->keysBuf[w->cur].data();
    v.manifolds = w->mansBuf[w->cur].data();
    v.oldKeys = w->keysBuf[w->cur ^ 1].data();
    v.oldManifolds = w->mansBuf[w->cur ^ 1].data();
    v.veto = nullptr;
    tw::m3SerialScratch s;
    s.root = w->root.data();
    s.masks = w->masks.data();
    tw::m3SerialParams sp = Params(p);
    v.sc->stall = 0;
    if (tw::m3SerialStepA(&v, &s, &sp, tape) != 0)
    {
        w->cur ^= 1;
        return 1;
    }
    return 0;
}

// From the constraint list to the impulse store, `keys` (ascending) vetoed.
static void StepB(m3gpuWorld* w, const m3gpuStepParams* p, tw::m3Tape* tape, const uint64_t* keys, int32_t n)
{
    tw::m3View& v = w->v;
    v.veto = nullptr;
    if (n > 0)
    {
        w->vetoFlags.assign(v.pairCapacity, 0);
        for (int32_t i = 0; i < v.sc->pairCount; ++i)
        {
            w->vetoFlags[i] = std::binary_search(keys, keys + n, v.keys[i]) ? 1 : 0;
        }
        v.veto = w->vetoFlags.data();
    }
    tw::m3SerialScratch s;
    s.root = w->root.data();
    s.masks = w->masks.data();
    tw::m3SerialParams sp = Params(p);
    tw::m3SerialStepB(&v, &s, &sp, tape);
    v.veto = nullptr;
}

static int StepBegin(m3gpuWorld* w, const m3gpuStepParams* p, tw::m3Tape* tape)
{
    const int rc = StepA(w, p, tape);
    if (rc != 0)
    {
        return rc;
    }
    StepB(w, p, tape, nullptr, 0);
    return 0;
}

static void StepEnd(m3gpuWorld* w, const m3gpuStepParams* p, tw::m3Tape* tape, const int32_t* ranks)
{
    tw::m3View& v = w->v;
    v.treeCcd = ranks != nullptr ? 1 : 0;
    if (ranks != nullptr)
    {
        memcpy(v.shapeRank, ranks, (size_t)v.maxShape * sizeof(int32_t));
    }
    tw::m3SerialScratch s;
    s.root = w->root.data();
    s.masks = w->masks.data();
    tw::m3SerialParams sp = Params(p);
    tw::m3SerialStepEnd(&v, &s, &sp, tape);
    if (w->pairTable)
    {
        for (int32_t i = 0; i < v.sc->pairCount; ++i)
        {
            w->tableKeys[i] = v.keys[i];
            w->tableMans[i] = Mb(v.manifolds[i]);
        }
    }
}

static int Step(m3gpuWorld* w, const m3gpuStepParams* p, tw::m3Tape* tape)
{
    const int rc = StepBegin(w, p, tape);
    if (rc != 0)
    {
        return rc;
    }
    StepEnd(w, p, tape, nullptr);
    return 0;
}

extern "C" int m3gpu_step_needs_tree(const m3gpuWorld* w, const m3gpuStepParams* p)
{
    return w != nullptr && p != nullptr && p->continuousEnabled && w->ccdNeedsTree;
}

extern "C" int m3gpu_step_supported(const m3gpuWorld* w, const m3gpuStepParams* p)
{
    return w != nullptr && p != nullptr && !m3gpu_step_needs_tree(w, p);
}

extern "C" int m3gpu_step_begin(m3gpuWorld* w, const m3gpuStepParams* p, uint32_t flags,
                                const uint64_t* vetoKeys, int32_t vetoCount)
{
    const int rc = StepA(w, p, nullptr);
    if (rc < 0)
    {
        return -1;
    }
    w->begun = 1;
    w->beginStalled = rc != 0;
    w->preCands.clear();
    if (w->beginStalled)
    {
        return 0;
    }
    if (flags & M3GPU_STEP_PRESOLVE)
    {
        tw::m3View& v = w->v;
        for (int32_t i = 0; i < v.sc->pairCount; ++i)
        {
            double point[3];
            tw::m3Vec3 normal;
            if (tw::m3PreSolveCandidate(&v, i, point, &normal))
            {
                m3gpuPreSolve c;
                memset(&c, 0, sizeof(c));
                c.key = v.keys[i];
                c.pair = i;
                c.point[0] = point[0];
                c.point[1] = point[1];
                c.point[2] = point[2];
                c.normal[0] = (float)normal.x;
                c.normal[1] = (float)normal.y;
                c.normal[2] = (float)normal.z;
                w->preCands.push_back(c);
            }
        }
        return 0;
    }
    StepB(w, p, nullptr, vetoKeys, vetoCount);
    return 0;
}

extern "C" int m3gpu_step_presolve(m3gpuWorld* w, const m3gpuPreSolve** candidates, int32_t* count)
{
    *candidates = w->preCands.data();
    *count = (int32_t)w->preCands.size();
    return 0;
}

extern "C" int m3gpu_step_vetoes(m3gpuWorld* w, const m3gpuStepParams* p, const uint64_t* vetoKeys,
                                 int32_t vetoCount)
{
    if (!w->beginStalled)
    {
        StepB(w, p, nullptr, vetoKeys, vetoCount);
    }
    return 0;
}

extern "C" const double* m3gpu_step_leaves(m3gpuWorld* w) { return w->v.leaf; }

extern "C" int m3gpu_step_end(m3gpuWorld* w, const m3gpuStepParams* p, const int32_t* ranks)
{
    if (!w->begun)
    {
        snprintf(s_err, sizeof(s_err), "twin: m3gpu_step_end without m3gpu_step_begin");
        return -1;
    }
    w->begun = 0;
    if (!w->beginStalled)
    {
        StepEnd(w, p, nullptr, ranks);
    }
    return 0;
}

extern "C" int m3gpu_step(m3gpuWorld* w, const m3gpuStepParams* p, m3gpuStepStats* st)
{
    if (!m3gpu_step_supported(w, p))
    {
        snprintf(s_err, sizeof(s_err), "twin: the continuous pass needs the tree ranks");
        return M3GPU_REFUSED;
    }
    int rc = Step(w, p, nullptr);
    if (rc < 0)
    {
        return -1;
    }
    if (st != nullptr)
    {
        memset(st, 0, sizeof(*st));
        st->pairCount = w->v.sc->pairCount;
        st->constraintCount = w->v.sc->constraintCount;
        st->colorCount = w->v.sc->colorCount;
        st->islandCount = w->v.sc->islandCount;
    }
    return rc;
}

extern "C" int m3gpu_step_many(m3gpuWorld* const* ws, const m3gpuStepParams* ps, int32_t n)
{
    for (int32_t k = 0; k < n; ++k)
    {
        if (m3gpu_step(ws[k], &ps[k], nullptr) < 0)
        {
            return -1;
        }
    }
    return 0;
}

// The event streams of the last step, by the device's rule (every begin
// event a new pair touching now and not before, every end event an old pair
// touching before and not now; hits in pair order; moves in mover order).
static int32_t TwinFindKey(const uint64_t* keys, int32_t n, uint64_t key)
{
    const uint64_t* it = std::lower_bound(keys, keys + n, key);
    return (it != keys + n && *it == key) ? (int32_t)(it - keys) : -1;
}

static int TwinPairCode(const tw::m3View& v, uint64_t key)
{
    const int32_t a = (int32_t)(key >> 32);
    const int32_t b = (int32_t)(key & 0xFFFFFFFFu);
    if (v.shapeAlive[a] == 0 || v.shapeAlive[b] == 0)
    {
        return 0;
    }
    return (v.sensor[a] != 0 || v.sensor[b] != 0) ? 2 : 1;
}

extern "C" int m3gpu_events(m3gpuWorld* w, float hitThreshold, m3gpuEvents* out)
{
    const tw::m3View& v = w->v;
    const tw::m3Scalars* sc = v.sc;
    memset(out->counts, 0, sizeof(out->counts));
    if (sc->stall != 0)
    {
        out->counts[6] = 1;
        return 0;
    }
    const int32_t np = sc->pairCount;
    const int32_t no = sc->oldPairCount;
    for (int32_t i = 0; i < np; ++i)
    {
        const uint64_t key = v.keys[i];
        const int32_t j = TwinFindKey(v.oldKeys, no, key);
        const bool touchNew = v.manifolds[i].pointCount > 0;
        const bool touchOld = j >= 0 && v.oldManifolds[j].pointCount > 0;
        const int code = (touchNew && !touchOld) ? TwinPairCode(v, key) : 0;
        if (code == 1)
        {
            out->beginKeys[out->counts[0]++] = key;
        }
        else if (code == 2)
        {
            out->sensorBeginKeys[out->counts[2]++] = key;
        }
        if (v.hitRecs != nullptr && v.hitRecs[i].flag != 0 &&
            v.hitRecs[i].relativeVelocity < -(R)hitThreshold)
        {
            const tw::m3HitRec& r = v.hitRecs[i];
            m3gpuHit& hit = out->hits[out->counts[4]++];
            hit.key = key;
            hit.point[0] = r.point[0];
            hit.point[1] = r.point[1];
            hit.point[2] = r.point[2];
            hit.normal[0] = (float)r.normal.x;
            hit.normal[1] = (float)r.normal.y;
            hit.normal[2] = (float)r.normal.z;
            hit.approachSpeed = (float)-r.relativeVelocity;
        }
    }
    for (int32_t i = 0; i < no; ++i)
    {
        const uint64_t key = v.oldKeys[i];
        const int32_t j = TwinFindKey(v.keys, np, key);
        const bool touchOld = v.oldManifolds[i].pointCount > 0;
        const bool touchNew = j >= 0 && v.manifolds[j].pointCount > 0;
        const int code = (touchOld && !touchNew) ? TwinPairCode(v, key) : 0;
        if (code == 1)
        {
            out->endKeys[out->counts[1]++] = key;
        }
        else if (code == 2)
        {
            out->sensorEndKeys[out->counts[3]++] = key;
        }
    }
    for (int32_t m = 0; m < sc->moverCount; ++m)
    {
        const int32_t b = v.movers[m];
        m3gpuMove& mv = out->moves[m];
        memset(&mv, 0, sizeof(mv));
        mv.body = b;
        mv.fellAsleep = v.type[b] == M3_DYNAMIC && v.awake[b] == 0;
        mv.p[0] = v.xf[b].p.x;
        mv.p[1] = v.xf[b].p.y;
        mv.p[2] = v.xf[b].p.z;
        mv.q[0] = (float)v.xf[b].q.x;
        mv.q[1] = (float)v.xf[b].q.y;
        mv.q[2] = (float)v.xf[b].q.z;
        mv.q[3] = (float)v.xf[b].q.w;
    }
    out->counts[5] = sc->moverCount;
    return 0;
}

extern "C" int m3gpu_sync(m3gpuWorld*) { return 0; }

extern "C" int m3gpu_last_stats(m3gpuWorld* w, m3gpuStepStats* s)
{
    memset(s, 0, sizeof(*s));
    s->pairCount = w->v.sc->pairCount;
    s->constraintCount = w->v.sc->constraintCount;
    s->colorCount = w->v.sc->colorCount;
    s->islandCount = w->v.sc->islandCount;
    return 0;
}

extern "C" int m3gpu_profile_begin(m3gpuWorld*) { return 0; }
extern "C" int m3gpu_profile_end(m3gpuWorld*, m3gpuProfile* p)
{
    memset(p, 0, sizeof(*p));
    return 0;
}

// ---- trajectory gradients ---------------------------------------------------

namespace
{
struct Checkpoint
{
    int cur;
    std::vector<tw::m3Transform> xf;
    std::vector<tw::m3Vec3> lv, av, force, torque;
    std::vector<uint8_t> awake, pending;
    std::vector<R> sleepTime;
    std::vector<uint64_t> keys, sleepKeys;
    std::vector<tw::m3Manifold> mans;
    std::vector<double> leaf;
    int32_t pairCount, sleepCount, fullQuery;
    uint64_t stepCount;
};

void Save(m3gpuWorld* w, Checkpoint* c)
{
    tw::m3View& v = w->v;
    int nb = v.maxBody, ns = v.maxShape;
    c->cur = w->cur;
    c->xf.assign(v.xf, v.xf + nb);
    c->lv.assign(v.lv, v.lv + nb);
    c->av.assign(v.av, v.av + nb);
    c->force.assign(v.force, v.force + nb);
    c->torque.assign(v.torque, v.torque + nb);
    c->awake.assign(v.awake, v.awake + nb);
    c->pending.assign(v.pending, v.pending + nb);
    c->sleepTime.assign(v.sleepTime, v.sleepTime + nb);
    c->pairCount = v.sc->pairCount;
    c->keys.assign(w->keysBuf[w->cur].begin(), w->keysBuf[w->cur].begin() + c->pairCount);
    c->mans.assign(w->mansBuf[w->cur].begin(), w->mansBuf[w->cur].begin() + c->pairCount);
    c->sleepCount = v.sc->sleepCount;
    c->sleepKeys.assign(v.sleepKeys, v.sleepKeys + c->sleepCount);
    c->leaf.assign(v.leaf, v.leaf + 6 * (size_t)ns);
    c->fullQuery = v.sc->fullQuery;
    c->stepCount = v.sc->stepCount;
}

void Load(m3gpuWorld* w, const Checkpoint* c)
{
    tw::m3View& v = w->v;
    w->cur = c->cur;
    std::copy(c->xf.begin(), c->xf.end(), v.xf);
    std::copy(c->lv.begin(), c->lv.end(), v.lv);
    std::copy(c->av.begin(), c->av.end(), v.av);
    std::copy(c->force.begin(), c->force.end(), v.force);
    std::copy(c->torque.begin(), c->torque.end(), v.torque);
    std::copy(c->awake.begin(), c->awake.end(), v.awake);
    std::copy(c->pending.begin(), c->pending.end(), v.pending);
    std::copy(c->sleepTime.begin(), c->sleepTime.end(), v.sleepTime);
    v.sc->pairCount = c->pairCount;
    std::copy(c->keys.begin(), c->keys.end(), w->keysBuf[w->cur].begin());
    std::copy(c->mans.begin(), c->mans.end(), w->mansBuf[w->cur].begin());
    v.keys = w->keysBuf[w->cur].data();
    v.manifolds = w->mansBuf[w->cur].data();
    v.oldKeys = w->keysBuf[w->cur ^ 1].data();
    v.oldManifolds = w->mansBuf[w->cur ^ 1].data();
    v.sc->sleepCount = c->sleepCount;
    std::copy(c->sleepKeys.begin(), c->sleepKeys.end(), v.sleepKeys);
    std::copy(c->leaf.begin(), c->leaf.end(), v.leaf);
    v.sc->fullQuery = c->fullQuery;
    v.sc->stepCount = c->stepCount;
}

// Tape storage of one step.
struct TapeStore
{
    std::vector<tw::m3Transform> xf0, sxf, xfEnd;
    std::vector<tw::m3Vec3> lv0, av0, force0, slv, sav, sdp, ilv, iav, plv, pav, rdp;
    std::vector<tw::m3Quat> sdq, rdq;
    std::vector<uint8_t> slept, ccd, mover;
    std::vector<uint64_t> keys;
    std::vector<tw::m3Manifold> man;
    std::vector<int32_t> carryPair, consPair, counts, colorStarts;
    std::vector<int8_t> carryPoint;
    std::vector<tw::m3TapePass> pass, fin;
    tw::m3Tape t;

    void Init(int nb, int np, int substeps)
    {
        size_t sb = (size_t)substeps * nb;
        xf0.resize(nb);
        lv0.resize(nb);
        av0.resize(nb);
        force0.resize(nb);
        xfEnd.resize(nb);
        slept.resize(nb);
        ccd.resize(nb);
        mover.resize(nb);
        keys.resize(np);
        man.resize(np);
        carryPair.resize(np);
        carryPoint.resize(4 * (size_t)np);
        consPair.resize(np);
        counts.resize(4);
        colorStarts.resize(M3_GRAPH_COLORS + 2);
        sxf.resize(sb);
        slv.resize(sb);
        sav.resize(sb);
        sdp.resize(sb);
        sdq.resize(sb);
        ilv.resize(sb);
        iav.resize(sb);
        plv.resize(sb);
        pav.resize(sb);
        rdp.resize(sb);
        rdq.resize(sb);
        pass.resize((size_t)substeps * 2 * np);
        fin.resize(np);
        t.bodyCap = nb;
        t.pairCap = np;
        t.substeps = substeps;
        t.xf0 = xf0.data();
        t.lv0 = lv0.data();
        t.av0 = av0.data();
        t.force0 = force0.data();
        t.xfEnd = xfEnd.data();
        t.slept = slept.data();
        t.ccd = ccd.data();
        t.mover = mover.data();
        t.keys = keys.data();
        t.man = man.data();
        t.carryPair = carryPair.data();
        t.carryPoint = carryPoint.data();
        t.consPair = consPair.data();
        t.counts = counts.data();
        t.colorStarts = colorStarts.data();
        t.sxf = sxf.data();
        t.slv = slv.data();
        t.sav = sav.data();
        t.sdp = sdp.data();
        t.sdq = sdq.data();
        t.ilv = ilv.data();
        t.iav = iav.data();
        t.plv = plv.data();
        t.pav = pav.data();
        t.rdp = rdp.data();
        t.rdq = rdq.data();
        t.pass = pass.data();
        t.fin = fin.data();
    }
};

int FirstShape(const tw::m3View& v, int b) { return v.shapeHead[b]; }

void ApplyForces(m3gpuWorld* w, const std::vector<tw::m3Vec3>& F)
{
    tw::m3View& v = w->v;
    for (int b = 0; b < v.maxBody; ++b)
    {
        if (v.bodyAlive[b] && v.type[b] == M3_DYNAMIC)
        {
            v.force[b] = F[b];
        }
    }
}

// Mass-weighted centre of mass of the dynamic bodies (double).
void Com(const tw::m3View& v, double out[3], double* totalMass)
{
    double M = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int b = 0; b < v.maxBody; ++b)
    {
        if (!v.bodyAlive[b] || v.type[b] != M3_DYNAMIC || v.invMass[b] == 0.0f)
        {
            continue;
        }
        double m = 1.0 / (double)v.invMass[b];
        tw::m3Vec3 r = tw::m3RotateVec3(v.xf[b].q, v.lc[b]);
        cx += m * (v.xf[b].p.x + (double)r.x);
        cy += m * (v.xf[b].p.y + (double)r.y);
        cz += m * (v.xf[b].p.z + (double)r.z);
        M += m;
    }
    out[0] = cx / M;
    out[1] = cy / M;
    out[2] = cz / M;
    *totalMass = M;
}

int Adjoint(m3gpuWorld* w, const m3gpuStepParams* p, const m3gpuAdjointRequest* rq)
{
    tw::m3View& v = w->v;
    const int nb = v.maxBody;
    const int np = v.pairCapacity;
    const int K = rq->checkpointEvery > 0 ? rq->checkpointEvery : 10;
    const int steps = rq->steps;
    // Originals to restore afterwards.
    Checkpoint origin;
    Save(w, &origin);
    std::vector<R> invMass0(v.invMass, v.invMass + nb);
    std::vector<tw::m3Mat3> invI0(v.invI, v.invI + nb), inertia0(v.inertia, v.inertia + nb);
    std::vector<R> friction0(v.friction, v.friction + v.maxShape);
    std::vector<tw::m3Vec3> F(nb, tw::m3V3(0, 0, 0));
    if (rq->force != nullptr)
    {
        for (int b = 0; b < nb; ++b)
        {
            F[b] = tw::m3V3((R)rq->force[3 * b], (R)rq->force[3 * b + 1], (R)rq->force[3 * b + 2]);
        }
    }
    // Parameter offset.
    if (rq->offsetKind != 0)
    {
        int b = rq->offsetBody, k = rq->offsetComp;
        R d = (R)rq->offsetValue;
        R* comp = nullptr;
        if (rq->offsetKind == 1)
            comp = k == 0 ? &v.lv[b].x : (k == 1 ? &v.lv[b].y : &v.lv[b].z);
        if (rq->offsetKind == 2)
            comp = k == 0 ? &v.av[b].x : (k == 1 ? &v.av[b].y : &v.av[b].z);
        if (rq->offsetKind == 5)
            comp = k == 0 ? &F[b].x : (k == 1 ? &F[b].y : &F[b].z);
        if (comp != nullptr)
        {
            *comp += d;
        }
        if (rq->offsetKind == 3)
        {
            R m = (R)1.0 / v.invMass[b];
            R m2 = m + d;
            v.invMass[b] = (R)1.0 / m2;
            R si = m / m2, sf = m2 / m;
            tw::m3Mat3& a = v.invI[b];
            a.cx = tw::m3MulSV3(si, a.cx);
            a.cy = tw::m3MulSV3(si, a.cy);
            a.cz = tw::m3MulSV3(si, a.cz);
            tw::m3Mat3& c = v.inertia[b];
            c.cx = tw::m3MulSV3(sf, c.cx);
            c.cy = tw::m3MulSV3(sf, c.cy);
            c.cz = tw::m3MulSV3(sf, c.cz);
        }
        if (rq->offsetKind == 4)
        {
            v.friction[FirstShape(v, b)] += d;
        }
    }
    // Forward with checkpoints.
    int segs = (steps + K - 1) / K;
    std::vector<Checkpoint> cps(segs);
    int rc = 0;
    for (int s = 0; s < steps && rc == 0; ++s)
    {
        if (s % K == 0)
        {
            Save(w, &cps[s / K]);
        }
        ApplyForces(w, F);
        rc = Step(w, p, nullptr);
    }
    double C[3], M;
    Com(v, C, &M);
    double L = 0.0;
    for (int k = 0; k < 3; ++k)
    {
        L += (C[k] - (double)rq->target[k]) * (C[k] - (double)rq->target[k]);
    }
    if (rq->loss != nullptr)
    {
        *rq->loss = L;
    }
    if (rc == 0 && !rq->forwardOnly)
    {
        std::vector<tw::m3BodyBar> body(nb);
        std::vector<tw::m3ParamBar> param(nb);
        std::vector<tw::m3ConBar> cons(np);
        std::vector<tw::m3ManBar> store(np), storePrev(np);
        std::vector<tw::m3Constraint> prep(np);
        std::vector<tw::m3Vec3> ccdP(nb);
        std::vector<tw::m3ConContrib> contrib(np);
        std::vector<tw::m3Quat> ccdQ(nb);
        memset(body.data(), 0, sizeof(tw::m3BodyBar) * nb);
        memset(param.data(), 0, sizeof(tw::m3ParamBar) * nb);
        memset(store.data(), 0, sizeof(tw::m3ManBar) * np);
        R Cb[3];
        for (int k = 0; k < 3; ++k)
        {
            Cb[k] = (R)(2.0 * (C[k] - (double)rq->target[k]));
        }
        std::vector<R> comBar(nb, (R)0);
        for (int b = 0; b < nb; ++b)
        {
            if (!v.bodyAlive[b] || v.type[b] != M3_DYNAMIC)
                continue;
            R m = (R)1.0 / v.invMass[b];
            tw::m3Vec3 cbar = tw::m3V3(Cb[0] * m / (R)M, Cb[1] * m / (R)M, Cb[2] * m / (R)M);
            body[b].p = cbar;
            tw::m3AdjRotate(v.xf[b].q, v.lc[b], cbar, &body[b].q);
            tw::m3Vec3 r = tw::m3RotateVec3(v.xf[b].q, v.lc[b]);
            double cx = v.xf[b].p.x + (double)r.x - C[0];
            double cy = v.xf[b].p.y + (double)r.y - C[1];
            double cz = v.xf[b].p.z + (double)r.z - C[2];
            comBar[b] = (R)((Cb[0] * cx + Cb[1] * cy + Cb[2] * cz) / M);
        }
        tw::m3AdjWorld aw;
        aw.body = body.data();
        aw.param = param.data();
        aw.cons = cons.data();
        aw.store = store.data();
        aw.storePrev = storePrev.data();
        aw.prep = prep.data();
        aw.ccdP = ccdP.data();
        aw.contrib = contrib.data();
        aw.ccdQ = ccdQ.data();
        aw.unsupported = 0;
        aw.ccdPulled = 0;
        tw::m3SerialParams sp = Params(p);
        std::vector<TapeStore> tapes(K);
        for (int k = 0; k < K; ++k)
        {
            tapes[k].Init(nb, np, p->substeps);
        }
        for (int seg = segs - 1; seg >= 0; --seg)
        {
            Load(w, &cps[seg]);
            int s0 = seg * K;
            int s1 = s0 + K < steps ? s0 + K : steps;
            for (int s = s0; s < s1; ++s)
            {
                ApplyForces(w, F);
                Step(w, p, &tapes[s - s0].t);
            }
            for (int s = s1 - 1; s >= s0; --s)
            {
                tw::m3AdjStep(&v, &sp, &tapes[s - s0].t, &aw);
                std::swap(aw.store, aw.storePrev);
            }
        }
        for (int b = 0; b < nb; ++b)
        {
            bool dyn = v.bodyAlive[b] && v.type[b] == M3_DYNAMIC;
            R m = dyn ? (R)1.0 / v.invMass[b] : (R)0;
            R gm = dyn ? param[b].invMass * (-v.invMass[b] * v.invMass[b]) + param[b].inertia / m -
                             param[b].invI / m + comBar[b]
                       : (R)0;
            if (rq->gradLinearVelocity)
            {
                rq->gradLinearVelocity[3 * b] = (float)body[b].v.x;
                rq->gradLinearVelocity[3 * b + 1] = (float)body[b].v.y;
                rq->gradLinearVelocity[3 * b + 2] = (float)body[b].v.z;
            }
            if (rq->gradAngularVelocity)
            {
                rq->gradAngularVelocity[3 * b] = (float)body[b].w.x;
                rq->gradAngularVelocity[3 * b + 1] = (float)body[b].w.y;
                rq->gradAngularVelocity[3 * b + 2] = (float)body[b].w.z;
            }
            if (rq->gradMass)
                rq->gradMass[b] = (float)gm;
            if (rq->gradFriction)
                rq->gradFriction[b] = (float)param[b].friction;
            if (rq->gradForce)
            {
                rq->gradForce[3 * b] = (float)param[b].force.x;
                rq->gradForce[3 * b + 1] = (float)param[b].force.y;
                rq->gradForce[3 * b + 2] = (float)param[b].force.z;
            }
        }
        if (rq->unsupported)
            *rq->unsupported = aw.unsupported;
        if (rq->ccdPulled)
            *rq->ccdPulled = aw.ccdPulled;
    }
    // Restore.
    Load(w, &origin);
    std::copy(invMass0.begin(), invMass0.end(), v.invMass);
    std::copy(invI0.begin(), invI0.end(), v.invI);
    std::copy(inertia0.begin(), inertia0.end(), v.inertia);
    std::copy(friction0.begin(), friction0.end(), v.friction);
    return rc;
}
} // namespace

extern "C" int m3gpu_adjoint_many(m3gpuWorld* const* ws, const m3gpuStepParams* ps,
                                  const m3gpuAdjointRequest* rqs, int32_t count)
{
    int rc = 0;
    for (int32_t k = 0; k < count; ++k)
    {
        rc |= Adjoint(ws[k], &ps[k], &rqs[k]);
    }
    return rc;
}

// The FP64 state itself, for checks that read it without rounding.
extern "C" int m3twin_body_state(m3gpuWorld* w, int32_t b, double out[13])
{
    tw::m3View& v = w->v;
    out[0] = v.xf[b].p.x;
    out[1] = v.xf[b].p.y;
    out[2] = v.xf[b].p.z;
    out[3] = (double)v.xf[b].q.x;
    out[4] = (double)v.xf[b].q.y;
    out[5] = (double)v.xf[b].q.z;
    out[6] = (double)v.xf[b].q.w;
    out[7] = (double)v.lv[b].x;
    out[8] = (double)v.lv[b].y;
    out[9] = (double)v.lv[b].z;
    out[10] = (double)v.av[b].x;
    out[11] = (double)v.av[b].y;
    out[12] = (double)v.av[b].z;
    return 0;
}

// Launch tuning has no meaning for the serial CPU twin: accepted, ignored.
extern "C" int m3gpu_set_substep_tuning(m3gpuWorld* w, int32_t blocks, int32_t tailLimit)
{
    (void)w;
    return blocks >= 0 && tailLimit >= 0 ? 0 : -1;
}

extern "C" int m3gpu_get_substep_tuning(const m3gpuWorld* w, int32_t* blocks, int32_t* tailLimit)
{
    (void)w;
    if (blocks != nullptr)
    {
        *blocks = 1;
    }
    if (tailLimit != nullptr)
    {
        *tailLimit = 0;
    }
    return 0;
}

extern "C" int m3gpu_set_batch_threads(int32_t threads)
{
    return threads >= 0 ? 0 : -1;
}

extern "C" int32_t m3gpu_get_batch_threads(void)
{
    return 1;
}
