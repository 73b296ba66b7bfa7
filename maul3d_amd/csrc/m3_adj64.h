// Trajectory gradients in double precision (m3gpu_adjoint_many, precision
// 64).  The serial step and its reverse pass (m3_serial.h, m3_adjoint.h)
// are instantiated at double in namespace m3f64.  The host converts a
// world's current host arrays into a double-precision image of its view,
// with room for the checkpoints, the tapes of one segment and the reverse
// pass's bars (Adj64Carve); one workgroup per world then runs Adj64Body:
// the checkpointed forward (lane 0 steps, the lanes copy), each segment's
// taped replay and reverse (lanes over constraints, bodies and pairs), and
// the gradients.  The world itself is not touched.
//
// Everything here is host-and-device code over plain memory, so the same
// image and driver also run on one CPU thread (lanes = 1, no barrier).
// Include after m3_adjoint.h (the float records) and m3gpu.h.
#ifndef M3_ADJ64_H
#define M3_ADJ64_H

#include <stdint.h>
#include <string.h>

#include "m3gpu.h"

namespace m3f64
{
#undef M3_CORE_H
#undef M3_COLLIDE_H
#undef M3_STEP_H
#undef M3_SERIAL_H
#undef M3_ADJOINT_H
#define M3_REAL double
#include "m3_serial.h"
#include "m3_adjoint.h"
#undef M3_REAL
} // namespace m3f64

struct m3Adj64World
{
    m3f64::m3View v;
    m3f64::m3View tv; // the reverse pass's view of the tape being reversed
    uint64_t* keysBuf[2];
    m3f64::m3Manifold* mansBuf[2];
    int32_t cur;
    int32_t steps;
    int32_t every;
    int32_t segs;
    int32_t forwardOnly;
    int32_t failed; // a step stalled: no loss, no gradients
    m3f64::m3SerialScratch scratch;
    m3f64::m3SerialParams sp;
    double target[3];
    m3f64::m3Vec3* force; // [maxBody] applied every step
    uint8_t* ckpt;        // [segs] checkpoints
    int64_t ckptStride;
    m3f64::m3Tape* tapes; // [every]
    m3f64::m3AdjWorld aw;
    double* comBar; // [maxBody]
    double loss;
    float* grads; // [maxBody * 11]
};

M3_HD int64_t Adj64Align(int64_t x) { return (x + 15) & ~(int64_t)15; }

template <typename T> M3_HD void Adj64Copy(T* dst, const T* src, int64_t n, int lane, int lanes)
{
    for (int64_t i = lane; i < n; i += lanes)
    {
        dst[i] = src[i];
    }
}

// What the lanes of one world share (LDS on the device).
struct m3Adj64Shared
{
    int failed;
    double C[3];
    double M;
};

M3_HD int64_t Adj64CheckpointBytes(int32_t nb, int32_t ns, int32_t np)
{
    int64_t b = 32;
    b += Adj64Align((int64_t)nb * sizeof(m3f64::m3Transform));
    b += 4 * Adj64Align((int64_t)nb * sizeof(m3f64::m3Vec3));
    b += 2 * Adj64Align((int64_t)nb);
    b += Adj64Align((int64_t)nb * sizeof(double));
    b += 2 * Adj64Align((int64_t)np * sizeof(uint64_t));
    b += Adj64Align((int64_t)np * sizeof(m3f64::m3Manifold));
    b += Adj64Align((int64_t)ns * 6 * sizeof(double));
    return b;
}

// Saves (save != 0) or loads the world's mutable state at checkpoint
// `slot`; every lane calls it between barriers.
template <typename Sync>
M3_HD void Adj64Checkpoint(m3Adj64World* c, int32_t slot, int save, int lane, int lanes, Sync sync)
{
    m3f64::m3View& v = c->v;
    const int32_t nb = v.maxBody;
    const int32_t ns = v.maxShape;
    const int32_t np = v.pairCapacity;
    uint8_t* p = c->ckpt + (int64_t)slot * c->ckptStride;
    int32_t* hdr = (int32_t*)p;
    uint64_t* stepCount = (uint64_t*)(p + 16);
    uint8_t* q = p + 32;
    m3f64::m3Transform* xf = (m3f64::m3Transform*)q;
    q += Adj64Align((int64_t)nb * sizeof(m3f64::m3Transform));
    m3f64::m3Vec3* lv = (m3f64::m3Vec3*)q;
    q += Adj64Align((int64_t)nb * sizeof(m3f64::m3Vec3));
    m3f64::m3Vec3* av = (m3f64::m3Vec3*)q;
    q += Adj64Align((int64_t)nb * sizeof(m3f64::m3Vec3));
    m3f64::m3Vec3* force = (m3f64::m3Vec3*)q;
    q += Adj64Align((int64_t)nb * sizeof(m3f64::m3Vec3));
    m3f64::m3Vec3* torque = (m3f64::m3Vec3*)q;
    q += Adj64Align((int64_t)nb * sizeof(m3f64::m3Vec3));
    uint8_t* awake = q;
    q += Adj64Align((int64_t)nb);
    uint8_t* pending = q;
    q += Adj64Align((int64_t)nb);
    double* sleepTime = (double*)q;
    q += Adj64Align((int64_t)nb * sizeof(double));
    uint64_t* keys = (uint64_t*)q;
    q += Adj64Align((int64_t)np * sizeof(uint64_t));
    m3f64::m3Manifold* mans = (m3f64::m3Manifold*)q;
    q += Adj64Align((int64_t)np * sizeof(m3f64::m3Manifold));
    uint64_t* sleepKeys = (uint64_t*)q;
    q += Adj64Align((int64_t)np * sizeof(uint64_t));
    double* leaf = (double*)q;
    if (save)
    {
        const int32_t cur = c->cur;
        const int32_t pairs = v.sc->pairCount;
        const int32_t sleeping = v.sc->sleepCount;
        if (lane == 0)
        {
            hdr[0] = cur;
            hdr[1] = pairs;
            hdr[2] = sleeping;
            hdr[3] = v.sc->fullQuery;
            *stepCount = v.sc->stepCount;
        }
        Adj64Copy(xf, v.xf, nb, lane, lanes);
        Adj64Copy(lv, v.lv, nb, lane, lanes);
        Adj64Copy(av, v.av, nb, lane, lanes);
        Adj64Copy(force, v.force, nb, lane, lanes);
        Adj64Copy(torque, v.torque, nb, lane, lanes);
        Adj64Copy(awake, v.awake, nb, lane, lanes);
        Adj64Copy(pending, v.pending, nb, lane, lanes);
        Adj64Copy(sleepTime, v.sleepTime, nb, lane, lanes);
        Adj64Copy(keys, c->keysBuf[cur], pairs, lane, lanes);
        Adj64Copy(mans, c->mansBuf[cur], pairs, lane, lanes);
        Adj64Copy(sleepKeys, v.sleepKeys, sleeping, lane, lanes);
        Adj64Copy(leaf, v.leaf, 6 * (int64_t)ns, lane, lanes);
        return;
    }
    const int32_t cur = hdr[0];
    const int32_t pairs = hdr[1];
    const int32_t sleeping = hdr[2];
    Adj64Copy(v.xf, xf, nb, lane, lanes);
    Adj64Copy(v.lv, lv, nb, lane, lanes);
    Adj64Copy(v.av, av, nb, lane, lanes);
    Adj64Copy(v.force, force, nb, lane, lanes);
    Adj64Copy(v.torque, torque, nb, lane, lanes);
    Adj64Copy(v.awake, awake, nb, lane, lanes);
    Adj64Copy(v.pending, pending, nb, lane, lanes);
    Adj64Copy(v.sleepTime, sleepTime, nb, lane, lanes);
    Adj64Copy(c->keysBuf[cur], keys, pairs, lane, lanes);
    Adj64Copy(c->mansBuf[cur], mans, pairs, lane, lanes);
    Adj64Copy(v.sleepKeys, sleepKeys, sleeping, lane, lanes);
    Adj64Copy(v.leaf, leaf, 6 * (int64_t)ns, lane, lanes);
    sync();
    if (lane == 0)
    {
        c->cur = cur;
        v.sc->pairCount = pairs;
        v.sc->sleepCount = sleeping;
        v.sc->fullQuery = hdr[3];
        v.sc->stepCount = *stepCount;
        v.keys = c->keysBuf[cur];
        v.manifolds = c->mansBuf[cur];
        v.oldKeys = c->keysBuf[cur ^ 1];
        v.oldManifolds = c->mansBuf[cur ^ 1];
    }
}

// One step of the world on the calling lane, with the applied forces set
// first; nonzero on a stall.
M3_HD int Adj64Step(m3Adj64World* c, m3f64::m3Tape* tape)
{
    m3f64::m3View& v = c->v;
    for (int32_t b = 0; b < v.maxBody; ++b)
    {
        if (v.bodyAlive[b] != 0 && v.type[b] == M3_DYNAMIC)
        {
            v.force[b] = c->force[b];
        }
    }
    c->cur ^= 1;
    v.keys = c->keysBuf[c->cur];
    v.manifolds = c->mansBuf[c->cur];
    v.oldKeys = c->keysBuf[c->cur ^ 1];
    v.oldManifolds = c->mansBuf[c->cur ^ 1];
    v.veto = nullptr;
    v.sc->stall = 0;
    if (m3f64::m3SerialStepA(&v, &c->scratch, &c->sp, tape) != 0)
    {
        c->cur ^= 1;
        return 1;
    }
    m3f64::m3SerialStepB(&v, &c->scratch, &c->sp, tape);
    v.treeCcd = 0;
    m3f64::m3SerialStepEnd(&v, &c->scratch, &c->sp, tape);
    return 0;
}

// The whole run of one world; every lane of its workgroup calls it.
template <typename Sync>
M3_HD void Adj64Body(m3Adj64World* c, int lane, int lanes, Sync sync, m3Adj64Shared* sh)
{
    m3f64::m3View& v = c->v;
    const int32_t K = c->every;
    int& failed = sh->failed;
    if (lane == 0)
    {
        failed = 0;
    }
    sync();
    // Forward with a checkpoint at the start of every segment.
    for (int32_t s = 0; s < c->steps && failed == 0; ++s)
    {
        if (s % K == 0)
        {
            Adj64Checkpoint(c, s / K, 1, lane, lanes, sync);
            sync();
        }
        if (lane == 0 && Adj64Step(c, nullptr) != 0)
        {
            failed = 1;
        }
        sync();
    }
    if (failed != 0)
    {
        if (lane == 0)
        {
            c->failed = 1;
        }
        return;
    }
    // Loss: squared distance of the dynamic bodies' mass-weighted centre.
    double* C = sh->C;
    double& M = sh->M;
    if (lane == 0)
    {
        double m0 = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
        for (int32_t b = 0; b < v.maxBody; ++b)
        {
            if (!v.bodyAlive[b] || v.type[b] != M3_DYNAMIC || v.invMass[b] == 0.0)
            {
                continue;
            }
            const double m = 1.0 / v.invMass[b];
            const m3f64::m3Vec3 r = m3f64::m3RotateVec3(v.xf[b].q, v.lc[b]);
            cx += m * (v.xf[b].p.x + r.x);
            cy += m * (v.xf[b].p.y + r.y);
            cz += m * (v.xf[b].p.z + r.z);
            m0 += m;
        }
        C[0] = cx / m0;
        C[1] = cy / m0;
        C[2] = cz / m0;
        M = m0;
        double L = 0.0;
        for (int i = 0; i < 3; ++i)
        {
            L += (C[i] - c->target[i]) * (C[i] - c->target[i]);
        }
        c->loss = L;
        c->aw.unsupported = 0;
        c->aw.ccdPulled = 0;
    }
    sync();
    if (c->forwardOnly)
    {
        return;
    }
    // Seeds: the loss's bars on the final state.
    double Cb[3];
    for (int i = 0; i < 3; ++i)
    {
        Cb[i] = 2.0 * (C[i] - c->target[i]);
    }
    for (int32_t b = lane; b < v.maxBody; b += lanes)
    {
        memset(&c->aw.body[b], 0, sizeof(m3f64::m3BodyBar));
        memset(&c->aw.param[b], 0, sizeof(m3f64::m3ParamBar));
        c->comBar[b] = 0.0;
        if (!v.bodyAlive[b] || v.type[b] != M3_DYNAMIC)
        {
            continue;
        }
        const double m = 1.0 / v.invMass[b];
        const m3f64::m3Vec3 cbar = m3f64::m3V3(Cb[0] * m / M, Cb[1] * m / M, Cb[2] * m / M);
        c->aw.body[b].p = cbar;
        m3f64::m3AdjRotate(v.xf[b].q, v.lc[b], cbar, &c->aw.body[b].q);
        const m3f64::m3Vec3 r = m3f64::m3RotateVec3(v.xf[b].q, v.lc[b]);
        const double dx = v.xf[b].p.x + r.x - C[0];
        const double dy = v.xf[b].p.y + r.y - C[1];
        const double dz = v.xf[b].p.z + r.z - C[2];
        c->comBar[b] = (Cb[0] * dx + Cb[1] * dy + Cb[2] * dz) / M;
    }
    for (int32_t i = lane; i < v.pairCapacity; i += lanes)
    {
        memset(&c->aw.store[i], 0, sizeof(m3f64::m3ManBar));
    }
    sync();
    // Each lane keeps its own counters; the store/storePrev swap is the
    // same on every lane.
    m3f64::m3AdjWorld a = c->aw;
    for (int32_t seg = c->segs - 1; seg >= 0 && failed == 0; --seg)
    {
        Adj64Checkpoint(c, seg, 0, lane, lanes, sync);
        sync();
        const int32_t s0 = seg * K;
        const int32_t s1 = s0 + K < c->steps ? s0 + K : c->steps;
        for (int32_t s = s0; s < s1 && failed == 0; ++s)
        {
            if (lane == 0 && Adj64Step(c, &c->tapes[s - s0]) != 0)
            {
                failed = 1;
            }
            sync();
        }
        for (int32_t s = s1 - 1; s >= s0 && failed == 0; --s)
        {
            const m3f64::m3Tape* t = &c->tapes[s - s0];
            if (lane == 0)
            {
                c->tv = v;
                m3f64::m3AdjTapeView(&c->tv, t);
            }
            sync();
            m3f64::m3AdjStepPar(&c->tv, &c->sp, t, &a, lane, lanes, sync);
            sync();
            m3f64::m3ManBar* sw = a.store;
            a.store = a.storePrev;
            a.storePrev = sw;
        }
    }
    if (failed != 0)
    {
        if (lane == 0)
        {
            c->failed = 1;
        }
        return;
    }
#if defined(__HIP_DEVICE_COMPILE__)
    if (a.unsupported != 0)
    {
        atomicAdd(&c->aw.unsupported, a.unsupported);
    }
    if (a.ccdPulled != 0)
    {
        atomicAdd(&c->aw.ccdPulled, a.ccdPulled);
    }
#else
    c->aw.unsupported += a.unsupported;
    c->aw.ccdPulled += a.ccdPulled;
#endif
    for (int32_t b = lane; b < v.maxBody; b += lanes)
    {
        float* g = &c->grads[11 * b];
        const m3f64::m3BodyBar* bb = &c->aw.body[b];
        const m3f64::m3ParamBar* pb = &c->aw.param[b];
        const int dyn = v.bodyAlive[b] && v.type[b] == M3_DYNAMIC;
        const double im = v.invMass[b];
        const double m = dyn ? 1.0 / im : 0.0;
        g[0] = (float)bb->v.x;
        g[1] = (float)bb->v.y;
        g[2] = (float)bb->v.z;
        g[3] = (float)bb->w.x;
        g[4] = (float)bb->w.y;
        g[5] = (float)bb->w.z;
        g[6] = dyn ? (float)(pb->invMass * (-im * im) + pb->inertia / m - pb->invI / m + c->comBar[b]) : 0.0f;
        g[7] = (float)pb->friction;
        g[8] = (float)pb->force.x;
        g[9] = (float)pb->force.y;
        g[10] = (float)pb->force.z;
    }
}

// ---- the host image ----------------------------------------------------------

static inline m3f64::m3Vec3 F64V(const m3Vec3& a) { return m3f64::m3V3(a.x, a.y, a.z); }
static inline m3f64::m3Quat F64Q(const m3Quat& a) { return m3f64::m3Q(a.x, a.y, a.z, a.w); }
static inline m3f64::m3Mat3 F64M(const m3Mat3& a)
{
    m3f64::m3Mat3 m;
    m.cx = F64V(a.cx);
    m.cy = F64V(a.cy);
    m.cz = F64V(a.cz);
    return m;
}
static inline m3f64::m3Transform F64X(const m3Transform& a)
{
    m3f64::m3Transform t;
    t.p.x = a.p.x;
    t.p.y = a.p.y;
    t.p.z = a.p.z;
    t.q = F64Q(a.q);
    return t;
}
static inline m3f64::m3ShapeGeom F64G(const m3ShapeGeom& a)
{
    m3f64::m3ShapeGeom g;
    g.v = F64V(a.v);
    g.s = a.s;
    g.v2 = F64V(a.v2);
    g.s2 = a.s2;
    return g;
}
static inline void F64H(m3f64::m3HullData* o, const m3HullData& a)
{
    memset(o, 0, sizeof(*o));
    o->vertexCount = a.vertexCount;
    o->faceCount = a.faceCount;
    o->indexCount = a.indexCount;
    o->unitMass = a.unitMass;
    o->unitCom = F64V(a.unitCom);
    o->unitInertiaCom = F64M(a.unitInertiaCom);
    for (int k = 0; k < M3_HULL_MAX_VERTS; ++k)
    {
        o->vertices[k] = F64V(a.vertices[k]);
    }
    for (int k = 0; k < M3_HULL_MAX_FACES; ++k)
    {
        o->faceNormals[k] = F64V(a.faceNormals[k]);
        o->faceOffsets[k] = a.faceOffsets[k];
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
    o->center = F64V(a.center);
}
static inline m3f64::m3Manifold F64Mf(const m3Manifold& a)
{
    m3f64::m3Manifold m;
    memset(&m, 0, sizeof(m));
    m.normal = F64V(a.normal);
    m.pointCount = a.pointCount;
    m.frictionImpulse = F64V(a.frictionImpulse);
    m.twistImpulse = a.twistImpulse;
    m.rollingImpulse = F64V(a.rollingImpulse);
    for (int k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        m.points[k].anchorA = F64V(a.points[k].anchorA);
        m.points[k].anchorB = F64V(a.points[k].anchorB);
        m.points[k].separation = a.points[k].separation;
        m.points[k].normalImpulse = a.points[k].normalImpulse;
        m.points[k].id = a.points[k].id;
        m.points[k].flags = a.points[k].flags;
    }
    return m;
}

// Carves one world's arrays out of the image: the sizing pass (host null)
// only advances the offset; the filling pass (same sequence) also returns
// each array's place in the host image.
struct m3Adj64Carver
{
    uint8_t* host; // image start, or null in the sizing pass
    uint8_t* dev;  // the image's device address
    int64_t off;
    template <typename T> T* take(int64_t n, T** hostOut = nullptr)
    {
        const int64_t at = off;
        off += Adj64Align((n > 0 ? n : 1) * (int64_t)sizeof(T));
        if (hostOut != nullptr)
        {
            *hostOut = host != nullptr ? (T*)(host + at) : nullptr;
        }
        return (T*)(dev + at);
    }
};

// The world's double-precision image: view arrays filled from the host
// state (with the request's parameter offset applied), pair buffers,
// scratch, checkpoints, tapes and bars; `c` receives the device pointers.
static inline void Adj64Carve(m3Adj64Carver& B, int32_t pairCapacity, const m3gpuStepParams* p,
                              const m3gpuAdjointRequest* r, m3Adj64World* c)
{
    const m3gpuHostState* h = r->host;
    const int32_t nb = h->maxBody;
    const int32_t ns = h->maxShape;
    const int32_t nh = h->maxHull > 0 ? h->maxHull : 1;
    const int32_t np = pairCapacity;
    const int32_t K = r->checkpointEvery > 0 ? r->checkpointEvery : 10;
    const int32_t segs = (r->steps + K - 1) / K;
    const bool fill = B.host != nullptr;
    memset(c, 0, sizeof(*c));
    m3f64::m3View& v = c->v;
    v.maxBody = nb;
    v.maxShape = ns;
    v.pairCapacity = np;
    v.emitCapacity = 2 * np;
    // bodies
    m3f64::m3Transform* xf;
    m3f64::m3Vec3 *lv, *av, *lc, *force, *torque;
    double *invMass, *gravityScale, *linDamp, *angDamp, *sleepTime, *minExt, *maxExt, *sleepThreshold;
    m3f64::m3Mat3 *invI, *inertia;
    uint8_t *bodyAlive, *type, *awake, *bullet, *enabled, *locks, *canSleep, *hasTarget;
    m3f64::m3Transform* target;
    int32_t* shapeHead;
    v.xf = B.take(nb, &xf);
    v.lv = B.take(nb, &lv);
    v.av = B.take(nb, &av);
    v.invMass = B.take(nb, &invMass);
    v.invI = B.take(nb, &invI);
    v.inertia = B.take(nb, &inertia);
    v.lc = B.take(nb, &lc);
    v.gravityScale = B.take(nb, &gravityScale);
    v.linDamp = B.take(nb, &linDamp);
    v.angDamp = B.take(nb, &angDamp);
    v.bodyAlive = B.take(nb, &bodyAlive);
    v.type = B.take(nb, &type);
    v.awake = B.take(nb, &awake);
    v.sleepTime = B.take(nb, &sleepTime);
    v.bullet = B.take(nb, &bullet);
    v.minExt = B.take(nb, &minExt);
    v.maxExt = B.take(nb, &maxExt);
    v.force = B.take(nb, &force);
    v.torque = B.take(nb, &torque);
    v.enabled = B.take(nb, &enabled);
    v.locks = B.take(nb, &locks);
    v.sleepThreshold = B.take(nb, &sleepThreshold);
    v.canSleep = B.take(nb, &canSleep);
    v.hasTarget = B.take(nb, &hasTarget);
    v.target = B.take(nb, &target);
    v.shapeHead = B.take(nb, &shapeHead);
    if (fill)
    {
        const m3Transform* hxf = (const m3Transform*)h->transforms;
        const m3Vec3* hlv = (const m3Vec3*)h->linearVelocities;
        const m3Vec3* hav = (const m3Vec3*)h->angularVelocities;
        const m3Mat3* hinvI = (const m3Mat3*)h->invInertiaLocal;
        const m3Mat3* hinertia = (const m3Mat3*)h->inertiaLocal;
        const m3Vec3* hlc = (const m3Vec3*)h->localCenters;
        const m3Vec3* hfo = (const m3Vec3*)h->bodyForce;
        const m3Vec3* hto = (const m3Vec3*)h->bodyTorque;
        const m3Transform* htarget = (const m3Transform*)h->bodyTarget;
        for (int32_t b = 0; b < nb; ++b)
        {
            xf[b] = F64X(hxf[b]);
            lv[b] = F64V(hlv[b]);
            av[b] = F64V(hav[b]);
            invMass[b] = h->invMass[b];
            invI[b] = F64M(hinvI[b]);
            inertia[b] = F64M(hinertia[b]);
            lc[b] = F64V(hlc[b]);
            gravityScale[b] = h->gravityScales[b];
            linDamp[b] = h->linearDamping[b];
            angDamp[b] = h->angularDamping[b];
            bodyAlive[b] = h->bodyAlive[b];
            type[b] = h->types[b];
            awake[b] = h->awake[b];
            sleepTime[b] = h->sleepTimes[b];
            bullet[b] = h->bulletFlags[b];
            minExt[b] = h->minExtents[b];
            maxExt[b] = h->maxExtents[b];
            force[b] = F64V(hfo[b]);
            torque[b] = F64V(hto[b]);
            enabled[b] = h->bodyEnabled[b];
            locks[b] = h->bodyLocks[b];
            sleepThreshold[b] = h->bodySleepThreshold[b];
            canSleep[b] = h->bodyCanSleep[b];
            hasTarget[b] = h->bodyHasTarget[b];
            target[b] = F64X(htarget[b]);
            shapeHead[b] = h->bodyShapeHead[b];
        }
    }
    // shapes
    uint8_t *shapeAlive, *shapeType, *sensor, *hitEvents, *preSolve, *hasOffset;
    int32_t *shapeBody, *group, *shapeNext, *hullIndex, *planes, *shapeRank, *ccdStatics;
    m3f64::m3ShapeGeom* geom;
    double *friction, *restitution, *rolling, *leaf, *fresh;
    m3f64::m3Vec3 *localPos, *surfaceVel;
    m3f64::m3Quat* localRot;
    uint64_t *category, *mask;
    m3f64::m3HullData* hulls;
    v.shapeAlive = B.take(ns, &shapeAlive);
    v.shapeBody = B.take(ns, &shapeBody);
    v.shapeType = B.take(ns, &shapeType);
    v.geom = B.take(ns, &geom);
    v.friction = B.take(ns, &friction);
    v.restitution = B.take(ns, &restitution);
    v.rolling = B.take(ns, &rolling);
    v.sensor = B.take(ns, &sensor);
    v.hitEvents = B.take(ns, &hitEvents);
    v.preSolve = B.take(ns, &preSolve);
    v.hasOffset = B.take(ns, &hasOffset);
    v.localPos = B.take(ns, &localPos);
    v.localRot = B.take(ns, &localRot);
    v.surfaceVel = B.take(ns, &surfaceVel);
    v.category = B.take(ns, &category);
    v.mask = B.take(ns, &mask);
    v.group = B.take(ns, &group);
    v.shapeNext = B.take(ns, &shapeNext);
    v.hullIndex = B.take(ns, &hullIndex);
    v.leaf = B.take(6 * (int64_t)ns, &leaf);
    v.fresh = B.take(6 * (int64_t)ns, &fresh);
    v.planes = B.take(ns, &planes);
    v.shapeRank = B.take(ns, &shapeRank);
    v.ccdStatics = B.take(ns, &ccdStatics);
    v.hulls = B.take(nh, &hulls);
    if (fill)
    {
        const m3ShapeGeom* hgeom = (const m3ShapeGeom*)h->shapeGeom;
        const m3Vec3* hlp = (const m3Vec3*)h->shapeLocalPos;
        const m3Quat* hlr = (const m3Quat*)h->shapeLocalRot;
        const m3Vec3* hsv = (const m3Vec3*)h->shapeSurfaceVel;
        int32_t pc = 0;
        int32_t sc = 0;
        for (int32_t s = 0; s < ns; ++s)
        {
            shapeAlive[s] = h->shapeAlive[s];
            shapeBody[s] = h->shapeBody[s];
            shapeType[s] = h->shapeType[s];
            geom[s] = F64G(hgeom[s]);
            friction[s] = h->shapeFriction[s];
            restitution[s] = h->shapeRestitution[s];
            rolling[s] = h->shapeRollingResistance[s];
            sensor[s] = h->shapeSensor[s];
            hitEvents[s] = h->shapeHitEvents[s];
            preSolve[s] = h->shapePreSolve[s];
            hasOffset[s] = h->shapeHasOffset[s];
            localPos[s] = F64V(hlp[s]);
            localRot[s] = F64Q(hlr[s]);
            surfaceVel[s] = F64V(hsv[s]);
            category[s] = h->shapeCategory[s];
            mask[s] = h->shapeMask[s];
            group[s] = h->shapeGroup[s];
            shapeNext[s] = h->shapeNext[s];
            hullIndex[s] = h->shapeHullIndex[s];
            shapeRank[s] = -1;
            if (h->shapeAlive[s] != 0 && h->shapeType[s] == M3_PLANE)
            {
                planes[pc++] = s;
            }
            else if (h->shapeAlive[s] != 0 && h->types[h->shapeBody[s]] == M3_STATIC)
            {
                ccdStatics[sc++] = s;
            }
        }
        memcpy(leaf, h->leafBoxes, (size_t)6 * ns * sizeof(double));
        v.planeCount = pc;
        v.ccdStaticCount = sc;
        const m3HullData* hh = (const m3HullData*)h->hullData;
        for (int32_t k = 0; k < h->maxHull; ++k)
        {
            F64H(&hulls[k], hh[k]);
        }
    }
    // pairs: the current list in buffer 0 (cur = 0); the first step swaps
    uint64_t* keys0;
    m3f64::m3Manifold* mans0;
    c->keysBuf[0] = B.take(np, &keys0);
    c->keysBuf[1] = B.take<uint64_t>(np);
    c->mansBuf[0] = B.take(np, &mans0);
    c->mansBuf[1] = B.take<m3f64::m3Manifold>(np);
    c->cur = 0;
    v.keys = c->keysBuf[0];
    v.manifolds = c->mansBuf[0];
    v.oldKeys = c->keysBuf[1];
    v.oldManifolds = c->mansBuf[1];
    uint64_t* sleepKeys;
    m3f64::m3Scalars* scal;
    v.sleepKeys = B.take(np, &sleepKeys);
    v.emit = B.take<uint32_t>(2 * (int64_t)np);
    v.cons = B.take<m3f64::m3Constraint>(np);
    v.consPair = B.take<int32_t>(np);
    v.colors = B.take<uint8_t>(np);
    v.colorLists = B.take<int32_t>(np);
    v.slotOf = B.take<int32_t>(np);
    v.movers = B.take<int32_t>(nb);
    v.deltaPos = B.take<m3f64::m3Vec3>(nb);
    v.deltaRot = B.take<m3f64::m3Quat>(nb);
    v.com0 = B.take<m3f64::m3Pos3>(nb);
    v.rot0 = B.take<m3f64::m3Quat>(nb);
    v.parent = B.take<int32_t>(nb);
    v.forced = B.take<uint8_t>(nb);
    v.flag = B.take<uint8_t>(nb);
    v.pending = B.take<uint8_t>(nb);
    v.sc = B.take(1, &scal);
    v.hitRecs = nullptr;
    c->scratch.root = B.take<int32_t>(nb);
    c->scratch.masks = B.take<uint32_t>(nb);
    if (fill)
    {
        const int32_t pairs = *h->pairCount;
        const int32_t sleeping = *h->sleepingPairCount;
        const m3Manifold* hm = (const m3Manifold*)h->manifolds;
        for (int32_t i = 0; i < pairs; ++i)
        {
            keys0[i] = h->pairKeys[i];
            mans0[i] = F64Mf(hm[i]);
        }
        memcpy(sleepKeys, h->sleepingPairKeys, (size_t)sleeping * sizeof(uint64_t));
        memset(scal, 0, sizeof(*scal));
        scal->pairCount = pairs;
        scal->sleepCount = sleeping;
        scal->fullQuery = *h->pairsFullQuery != 0;
        scal->stepCount = *h->stepCount;
    }
    // run parameters and the applied forces
    m3f64::m3SerialParams& sp = c->sp;
    sp.dt = p->dt;
    sp.substeps = p->substeps;
    sp.gravity = m3f64::m3V3(p->gravity[0], p->gravity[1], p->gravity[2]);
    sp.contactHertz = p->contactHertz;
    sp.contactDampingRatio = p->contactDampingRatio;
    sp.contactPushMaxSpeed = p->contactPushMaxSpeed;
    sp.restitutionThreshold = p->restitutionThreshold;
    sp.maximumLinearSpeed = p->maximumLinearSpeed;
    sp.maximumAngularSpeed = p->maximumAngularSpeed;
    sp.sleepEnabled = p->sleepEnabled;
    sp.continuousEnabled = p->continuousEnabled;
    c->steps = r->steps;
    c->every = K;
    c->segs = segs;
    c->forwardOnly = r->forwardOnly;
    for (int i = 0; i < 3; ++i)
    {
        c->target[i] = r->target[i];
    }
    m3f64::m3Vec3* F;
    c->force = B.take(nb, &F);
    if (fill)
    {
        for (int32_t b = 0; b < nb; ++b)
        {
            F[b] = r->force != nullptr ? m3f64::m3V3(r->force[3 * b], r->force[3 * b + 1], r->force[3 * b + 2])
                                       : m3f64::m3V3(0.0, 0.0, 0.0);
        }
        // The request's parameter offset (central differences).
        const int32_t ob = r->offsetBody;
        const int32_t oc = r->offsetComp;
        const double d = r->offsetValue;
        if (r->offsetKind != 0 && ob >= 0 && ob < nb)
        {
            m3f64::m3Vec3* vec = r->offsetKind == 1 ? &lv[ob] : r->offsetKind == 2 ? &av[ob] : r->offsetKind == 5 ? &F[ob] : nullptr;
            if (vec != nullptr)
            {
                double* comp = oc == 0 ? &vec->x : oc == 1 ? &vec->y : &vec->z;
                *comp += d;
            }
            if (r->offsetKind == 3 && invMass[ob] > 0.0)
            {
                const double m = 1.0 / invMass[ob];
                const double m2 = m + d;
                invMass[ob] = 1.0 / m2;
                const double si = m / m2;
                const double sf = m2 / m;
                invI[ob].cx = m3f64::m3MulSV3(si, invI[ob].cx);
                invI[ob].cy = m3f64::m3MulSV3(si, invI[ob].cy);
                invI[ob].cz = m3f64::m3MulSV3(si, invI[ob].cz);
                inertia[ob].cx = m3f64::m3MulSV3(sf, inertia[ob].cx);
                inertia[ob].cy = m3f64::m3MulSV3(sf, inertia[ob].cy);
                inertia[ob].cz = m3f64::m3MulSV3(sf, inertia[ob].cz);
            }
            if (r->offsetKind == 4 && shapeHead[ob] >= 0)
            {
                friction[shapeHead[ob]] += d;
            }
        }
    }
    // checkpoints, the tapes of one segment, the reverse pass's bars
    c->ckptStride = Adj64CheckpointBytes(nb, ns, np);
    c->ckpt = B.take<uint8_t>(c->ckptStride * segs);
    m3f64::m3Tape* tapes;
    c->tapes = B.take(K, &tapes);
    const int32_t sub = p->substeps;
    const int64_t sb = (int64_t)sub * nb;
    for (int32_t k = 0; k < K; ++k)
    {
        m3f64::m3Tape t;
        memset(&t, 0, sizeof(t));
        t.bodyCap = nb;
        t.pairCap = np;
        t.substeps = sub;
        t.xf0 = B.take<m3f64::m3Transform>(nb);
        t.lv0 = B.take<m3f64::m3Vec3>(nb);
        t.av0 = B.take<m3f64::m3Vec3>(nb);
        t.force0 = B.take<m3f64::m3Vec3>(nb);
        t.xfEnd = B.take<m3f64::m3Transform>(nb);
        t.slept = B.take<uint8_t>(nb);
        t.ccd = B.take<uint8_t>(nb);
        t.mover = B.take<uint8_t>(nb);
        t.keys = B.take<uint64_t>(np);
        t.man = B.take<m3f64::m3Manifold>(np);
        t.carryPair = B.take<int32_t>(np);
        t.carryPoint = B.take<int8_t>(4 * (int64_t)np);
        t.consPair = B.take<int32_t>(np);
        t.counts = B.take<int32_t>(4);
        t.colorStarts = B.take<int32_t>(M3_GRAPH_COLORS + 2);
        t.sxf = B.take<m3f64::m3Transform>(sb);
        t.slv = B.take<m3f64::m3Vec3>(sb);
        t.sav = B.take<m3f64::m3Vec3>(sb);
        t.sdp = B.take<m3f64::m3Vec3>(sb);
        t.sdq = B.take<m3f64::m3Quat>(sb);
        t.ilv = B.take<m3f64::m3Vec3>(sb);
        t.iav = B.take<m3f64::m3Vec3>(sb);
        t.plv = B.take<m3f64::m3Vec3>(sb);
        t.pav = B.take<m3f64::m3Vec3>(sb);
        t.rdp = B.take<m3f64::m3Vec3>(sb);
        t.rdq = B.take<m3f64::m3Quat>(sb);
        t.pass = B.take<m3f64::m3TapePass>((int64_t)sub * 2 * np);
        t.fin = B.take<m3f64::m3TapePass>(np);
        if (fill)
        {
            tapes[k] = t;
        }
    }
    m3f64::m3AdjWorld& aw = c->aw;
    aw.body = B.take<m3f64::m3BodyBar>(nb);
    aw.param = B.take<m3f64::m3ParamBar>(nb);
    aw.cons = B.take<m3f64::m3ConBar>(np);
    aw.store = B.take<m3f64::m3ManBar>(np);
    aw.storePrev = B.take<m3f64::m3ManBar>(np);
    aw.prep = B.take<m3f64::m3Constraint>(np);
    aw.contrib = B.take<m3f64::m3ConContrib>(np);
    aw.ccdP = B.take<m3f64::m3Vec3>(nb);
    aw.ccdQ = B.take<m3f64::m3Quat>(nb);
    c->comBar = B.take<double>(nb);
    c->grads = B.take<float>(11 * (int64_t)nb);
}


#endif // M3_ADJ64_H
