// Per-item stage bodies of the device step.  Each function handles one
// shape, pair, constraint or body and is called by exactly one GPU thread;
// the orchestration (ordering, compaction, graph-color scheduling) lives in
// m3_device.hip.  Arithmetic follows the reference stage it names, in the
// same operation order, so the device reproduces the CPU engine's bits.
#ifndef M3_STEP_H
#define M3_STEP_H

#include "m3_collide.h"

// ---------------------------------------------------------------------------
// Device-resident world: structure-of-arrays in HBM.  Body and shape records
// keep the reference byte layouts so uploads and downloads are plain copies.
// ---------------------------------------------------------------------------

typedef struct m3Softness
{
    m3real biasRate;
    m3real massScale;
    m3real impulseScale;
} m3Softness;

typedef struct m3CPoint
{
    m3Vec3 rA;
    m3Vec3 rB;
    m3real baseSeparation;
    m3real normalMass;
    m3real leverArm;
    m3real relativeVelocity;
    m3real normalImpulse;
    m3real totalNormalImpulse;
} m3CPoint;

// One contact constraint (the prepared form of one touching manifold).
typedef struct m3Constraint
{
    int32_t bodyA;
    int32_t bodyB;
    int32_t manifoldIndex;
    int32_t pointCount;
    m3Vec3 normal;
    m3Vec3 t1;
    m3Vec3 t2;
    m3Vec3 originA;
    m3Vec3 originB;
    m3real frictionK11;
    m3real frictionK12;
    m3real frictionK22;
    m3real frictionImpulse1;
    m3real frictionImpulse2;
    m3real twistMass;
    m3real twistImpulse;
    m3real tangentVelocity1;
    m3real tangentVelocity2;
    m3real friction;
    m3real restitution;
    m3real invMassA;
    m3real invMassB;
    m3Mat3 invIA;
    m3Mat3 invIB;
    m3Softness softness;
    m3real rollingResistance;
    m3Vec3 rollingImpulse;
    m3Mat3 rollingK;
    uint8_t dynA; // body A is dynamic (impulses land on it)
    uint8_t dynB;
    uint8_t pad[2];
    m3CPoint points[M3_MANIFOLD_MAX_POINTS];
} m3Constraint;

// One prepared spherical joint (the fields of the reference's
// m3JointConstraint that the spherical rows read, src/solver.c:584-639).
typedef struct m3JointCon
{
    int32_t bodyA;
    int32_t bodyB;
    uint8_t flags; // bit0 twist limit, bit1 motor, bit2 cone, bit3 drive spring
    uint8_t pad[3];
    m3Vec3 rA; // COM-relative anchors, world-rotated at prepare
    m3Vec3 rB;
    m3Vec3 deltaCenter; // cB - cA at prepare
    m3real invMassA;
    m3real invMassB;
    m3Mat3 invIA;
    m3Mat3 invIB;
    m3Softness softness;
    m3Vec3 impulse;
    m3Quat frameQA;
    m3Quat frameQB;
    m3Vec3 swingAxis;
    m3Vec3 twistJacobian;
    m3real swingMass;
    m3real twistMass;
    m3real coneAngle;
    m3real lowerLimit;
    m3real upperLimit;
    m3real lowerImpulse;
    m3real upperImpulse;
    m3real swingImpulse;
    m3Vec3 springImpulseV;
} m3JointCon;

// Step-wide scalars that live on the device so no stage needs a host
// round trip.
typedef struct m3Scalars
{
    int32_t pairCount;
    int32_t oldPairCount;
    int32_t sleepCount;
    int32_t emitCount;
    int32_t uniqueCount;
    int32_t constraintCount;
    int32_t moverCount;
    int32_t stall;
    int32_t anyRestitution;
    int32_t islandCount;
    int32_t awakeCount;
    int32_t colorCount;
    int32_t discoverCount;
    int32_t pointCount; // contact points over all constraints
    int32_t colorStarts[M3_GRAPH_COLORS + 2];
    int32_t fault;     // sticky: a grid barrier timed out or the coloring stalled
    int32_t fullQuery; // the next step rebuilds the frozen-pair buffer
    int32_t jointActive; // joints prepared this step
    int32_t satCount;    // wide hull pairs listed for the wavefront SAT queries
    int32_t flowVeto;    // a constraint does not fit the barrier-free substep passes
    int32_t flowKinematic; // a contact touches a kinematic body, whose position deltas move
    uint64_t stepCount;
} m3Scalars;

// A step's hit candidate for one pair (src/solver.c:2991-3026): the
// fastest-approaching point among those whose accumulated normal impulse is
// positive, kept whatever the world's hit threshold; the threshold is applied
// when the events are read.  flag is cleared by the narrowphase and set by
// the impulse store.
typedef struct m3HitRec
{
    double point[3]; // world point: body A's position + rotated local center + rA
    m3Vec3 normal;
    m3real relativeVelocity;
    int32_t flag;
    int32_t pad;
} m3HitRec;

typedef struct m3View
{
    int32_t maxBody;
    int32_t maxShape;
    int32_t pairCapacity;
    int32_t emitCapacity;
    int32_t planeCount;
    // bodies
    m3Transform* xf;
    m3Vec3* lv;
    m3Vec3* av;
    m3real* invMass;
    m3Mat3* invI;
    m3Mat3* inertia;
    m3Vec3* lc;
    m3real* gravityScale;
    m3real* linDamp;
    m3real* angDamp;
    uint8_t* bodyAlive;
    uint8_t* type;
    uint8_t* awake;
    m3real* sleepTime;
    uint8_t* bullet;
    m3real* minExt;
    m3real* maxExt;
    m3Vec3* force;
    m3Vec3* torque;
    uint8_t* enabled;
    uint8_t* locks;
    m3real* sleepThreshold;
    uint8_t* canSleep;
    uint8_t* hasTarget;
    m3Transform* target;
    int32_t* shapeHead;
    // shapes
    uint8_t* shapeAlive;
    int32_t* shapeBody;
    uint8_t* shapeType;
    m3ShapeGeom* geom;
    m3real* friction;
    m3real* restitution;
    m3real* rolling;
    uint8_t* sensor;
    uint8_t* hitEvents; // per shape: hit events opted in
    uint8_t* hasOffset;
    m3Vec3* localPos;
    m3Quat* localRot;
    m3Vec3* surfaceVel;
    uint64_t* category;
    uint64_t* mask;
    int32_t* group;
    int32_t* shapeNext;
    int32_t* hullIndex;
    double* leaf;   // 6 per shape
    double* fresh;  // 6 per shape, this step's tight bounds
    int32_t* planes; // alive plane shapes, ascending
    const m3HullData* hulls;
    // pairs
    uint64_t* keys;
    m3Manifold* manifolds;
    const uint64_t* oldKeys;
    const m3Manifold* oldManifolds;
    uint64_t* sleepKeys;
    uint32_t* emit;
    m3SatQuery* satQueries; // per pair: SAT queries of wide hull pairs, evaluated ahead
    int32_t satBoxes;       // box pairs count as wide (one large world)
    int32_t* satList;       // pair indices of the wide hull pairs (any order)
    // solver scratch
    m3Constraint* cons;   // prepared constraints in color-slot order
    int32_t* consPair;    // pair index of each constraint (constraint order)
    int32_t* consBodies;  // awake dynamic bodies of each constraint (-1 = none), 2 per constraint
    uint8_t* colors;      // color of each constraint (constraint order)
    uint8_t* slotColors;  // color of each slot (sorted)
    int32_t* colorLists;  // constraint index of each slot
    int32_t* slotOf;      // slot of each constraint
    int32_t* slotTable;   // per body and color < 16: slot << 1 | side
    uint32_t* slotMask;   // per body: bit c = a slot in color c; bit 16 = overflow slots
    int32_t* movers;
    m3Vec3* deltaPos;
    m3Quat* deltaRot;
    m3Pos3* com0;
    m3Quat* rot0;
    int32_t* parent;
    uint8_t* forced;
    uint8_t* flag;
    uint8_t* pending; // bodies whose frozen pairs the next sweep records
    m3Scalars* sc;
    m3HitRec* hitRecs; // per pair of the current list, or null (no hit records)
    // joints: slot arrays (the reference's world arrays), the joint
    // components (joints linked through non-static bodies, each a run of
    // ascending slots), and this step's prepared records
    int32_t maxJoint;
    int32_t jointCompCount;
    uint8_t* jointAlive;
    uint8_t* jointType;
    int32_t* jointBodyA;
    int32_t* jointBodyB;
    m3Vec3* jointLocalA;
    m3Vec3* jointLocalB;
    uint8_t* jointCollide;
    m3Quat* jointFrameQA;
    m3Quat* jointFrameQB;
    uint8_t* jointFlags;
    m3Vec3* jointLimits;
    m3Vec3* jointImpulse;
    m3Vec3* jointPerpImpulse;
    m3Vec3* jointLimitImpulse;
    m3Vec3* jointAngularImpulse;
    m3Vec3* jointSpringImpulse;
    int32_t* jointNextA;
    int32_t* jointNextB;
    int32_t* bodyJointHead;
    int32_t* jointCompStart;
    int32_t* jointCompList;
    m3JointCon* jcons;
    uint8_t* jointActiveFlag;
    // continuous collision against tree shapes: each shape's depth-first
    // rank in the broadphase tree (-1 when not in the tree), valid for the
    // step when treeCcd is set; the static tree shapes, ascending
    int32_t* shapeRank;
    int32_t treeCcd;
    int32_t ccdStaticCount;
    int32_t* ccdStatics;
    // pre-solve: per shape opt-in (m3Shape_EnablePreSolve) and, for the
    // step that carries vetoes, one flag per pair of the current list
    uint8_t* preSolve;
    const uint8_t* veto;
} m3View;

// ---------------------------------------------------------------------------
// Shared helpers
// ---------------------------------------------------------------------------

// src/world_internal.h:1069-1083
M3_HD m3Transform m3ShapeXf(const m3View* v, int32_t s)
{
    m3Transform xf = v->xf[v->shapeBody[s]];
    if (v->hasOffset[s] != 0)
    {
        m3Vec3 r = m3RotateVec3(xf.q, v->localPos[s]);
        xf.p.x += (double)r.x;
        xf.p.y += (double)r.y;
        xf.p.z += (double)r.z;
        xf.q = m3MulQuat(xf.q, v->localRot[s]);
    }
    return xf;
}

// src/broad_phase.c:24-148: the margin-inflated tight box of one shape.
M3_HD void m3FreshBox(const m3View* v, int32_t s, double box[6])
{
    m3Transform xf = m3ShapeXf(v, s);
    uint8_t type = v->shapeType[s];
    if (type == M3_HULL)
    {
        const m3HullData* hull = &v->hulls[v->hullIndex[s]];
        double lo0 = 1.0e30, lo1 = 1.0e30, lo2 = 1.0e30;
        double hi0 = -1.0e30, hi1 = -1.0e30, hi2 = -1.0e30;
        for (int32_t k = 0; k < hull->vertexCount; ++k)
        {
            m3Vec3 r = m3RotateVec3(xf.q, hull->vertices[k]);
            double p0 = xf.p.x + (double)r.x;
            double p1 = xf.p.y + (double)r.y;
            double p2 = xf.p.z + (double)r.z;
            lo0 = p0 < lo0 ? p0 : lo0;
            hi0 = p0 > hi0 ? p0 : hi0;
            lo1 = p1 < lo1 ? p1 : lo1;
            hi1 = p1 > hi1 ? p1 : hi1;
            lo2 = p2 < lo2 ? p2 : lo2;
            hi2 = p2 > hi2 ? p2 : hi2;
        }
        box[0] = lo0 - (double)M3_AABB_MARGIN;
        box[1] = lo1 - (double)M3_AABB_MARGIN;
        box[2] = lo2 - (double)M3_AABB_MARGIN;
        box[3] = hi0 + (double)M3_AABB_MARGIN;
        box[4] = hi1 + (double)M3_AABB_MARGIN;
        box[5] = hi2 + (double)M3_AABB_MARGIN;
        return;
    }
    if (type == M3_CAPSULE)
    {
        m3Vec3 r1 = m3RotateVec3(xf.q, v->geom[s].v);
        m3Vec3 r2 = m3RotateVec3(xf.q, v->geom[s].v2);
        double c1[3] = {xf.p.x + (double)r1.x, xf.p.y + (double)r1.y, xf.p.z + (double)r1.z};
        double c2[3] = {xf.p.x + (double)r2.x, xf.p.y + (double)r2.y, xf.p.z + (double)r2.z};
        double fat = (double)(v->geom[s].s + M3_AABB_MARGIN);
        for (int32_t k = 0; k < 3; ++k)
        {
            box[k] = (c1[k] < c2[k] ? c1[k] : c2[k]) - fat;
            box[3 + k] = (c1[k] > c2[k] ? c1[k] : c2[k]) + fat;
        }
        return;
    }
    m3Vec3 r = m3RotateVec3(xf.q, v->geom[s].v);
    double cx = xf.p.x + (double)r.x;
    double cy = xf.p.y + (double)r.y;
    double cz = xf.p.z + (double)r.z;
    double fat = (double)(v->geom[s].s + M3_AABB_MARGIN);
    box[0] = cx - fat;
    box[1] = cy - fat;
    box[2] = cz - fat;
    box[3] = cx + fat;
    box[4] = cy + fat;
    box[5] = cz + fat;
}

M3_HD int m3BoxOverlap(const double* a, const double* b)
{
    return a[0] <= b[3] && b[0] <= a[3] && a[1] <= b[4] && b[1] <= a[4] && a[2] <= b[5] &&
           b[2] <= a[5];
}

// src/broad_phase.c:182-232 (joint lists are empty on this path).
M3_HD int m3PairAllowed(const m3View* v, int32_t i, int32_t j)
{
    int32_t bi = v->shapeBody[i];
    int32_t bj = v->shapeBody[j];
    if (bi == bj)
    {
        return 0;
    }
    if (v->enabled[bi] == 0 || v->enabled[bj] == 0)
    {
        return 0;
    }
    if (v->type[bi] == M3_STATIC && v->type[bj] == M3_STATIC)
    {
        return 0;
    }
    if (v->sensor[i] != 0 && v->sensor[j] != 0)
    {
        return 0;
    }
    int32_t gi = v->group[i];
    if (gi != 0 && gi == v->group[j])
    {
        if (gi < 0)
        {
            return 0;
        }
    }
    else if (!((v->category[i] & v->mask[j]) != 0 && (v->category[j] & v->mask[i]) != 0))
    {
        return 0;
    }
    // Jointed bodies skip contact unless the joint allows it
    // (src/broad_phase.c:221-231).
    if (v->bodyJointHead != 0)
    {
        for (int32_t jt = v->bodyJointHead[bi]; jt != -1;
             jt = v->jointBodyA[jt] == bi ? v->jointNextA[jt] : v->jointNextB[jt])
        {
            int32_t other = v->jointBodyA[jt] == bi ? v->jointBodyB[jt] : v->jointBodyA[jt];
            if (other == bj && v->jointCollide[jt] == 0)
            {
                return 0;
            }
        }
    }
    return 1;
}

// src/broad_phase.c:238-251
M3_HD int m3PairHot(const m3View* v, int32_t i, int32_t j)
{
    int32_t bi = v->shapeBody[i];
    int32_t bj = v->shapeBody[j];
    if (v->type[bi] != M3_STATIC && v->awake[bi] != 0)
    {
        return 1;
    }
    return v->type[bj] != M3_STATIC && v->awake[bj] != 0;
}

// A shape runs its own tree query when its body is static or awake.
M3_HD int m3ShapeQueries(const m3View* v, int32_t s)
{
    int32_t b = v->shapeBody[s];
    return v->type[b] == M3_STATIC || v->awake[b] != 0;
}

// Would the reference's tree pass emit the pair (a, b), whose fresh boxes
// overlap?  The query runs from the querying side (from the lower index
// when both query) and hits the other shape when that shape's stored proxy
// box overlaps the querier's fresh box, then re-tests fresh boxes
// (src/broad_phase.c:285-349, src/dynamic_tree.c:367-406).  A querying
// shape's proxy is refreshed to contain its fresh box before the queries
// run, so for two queriers the proxy test is implied by the fresh overlap;
// a non-querying shape's proxy is never refreshed, so its stored box is
// tested as it stands.  This lets the test run before the refresh.
M3_HD int m3TreePairEmits(const m3View* v, int32_t a, int32_t b)
{
    int qa = m3ShapeQueries(v, a);
    int qb = m3ShapeQueries(v, b);
    if (!qa && !qb)
    {
        return 0;
    }
    if (!m3PairAllowed(v, a, b) || !m3PairHot(v, a, b))
    {
        return 0;
    }
    if (qa && qb)
    {
        return 1;
    }
    int32_t self = qa ? a : b;
    int32_t other = qa ? b : a;
    return m3BoxOverlap(&v->leaf[6 * other], &v->fresh[6 * self]);
}

// Frozen-pair discovery of `self` (a shape of a body that froze): the tree
// hit needs the other proxy's stored box to overlap self's fresh box, plus
// the pair filter (src/broad_phase.c:351-383); the fresh overlap is the
// caller's.
M3_HD int m3DiscoverHit(const m3View* v, int32_t self, int32_t other)
{
    return m3BoxOverlap(&v->leaf[6 * other], &v->fresh[6 * self]) &&
           m3PairAllowed(v, self, other);
}

// Compact 32-bit pair key: low shape in the high half.
M3_HD uint32_t m3CompactKey(int32_t i, int32_t j)
{
    return i < j ? (((uint32_t)i << 16) | (uint32_t)j) : (((uint32_t)j << 16) | (uint32_t)i);
}

M3_HD uint64_t m3ExpandKey(uint32_t k)
{
    return ((uint64_t)(k >> 16) << 32) | (uint64_t)(k & 0xFFFFu);
}

// ---------------------------------------------------------------------------
// Narrowphase (src/manifold.c:2357-2920)
// ---------------------------------------------------------------------------

M3_HD m3Vec3 m3FromCom(const m3View* v, int32_t body, double px, double py, double pz)
{
    const m3Transform xf = v->xf[body];
    m3Vec3 rlc = m3RotateVec3(xf.q, v->lc[body]);
    return m3V3((m3real)(px - xf.p.x - (double)rlc.x), (m3real)(py - xf.p.y - (double)rlc.y),
                (m3real)(pz - xf.p.z - (double)rlc.z));
}

M3_HD m3Proxy m3MakeProxy(const m3View* v, int32_t s, m3Vec3 scratch[2])
{
    m3Proxy proxy;
    uint8_t type = v->shapeType[s];
    if (type == M3_HULL)
    {
        const m3HullData* hull = &v->hulls[v->hullIndex[s]];
        proxy.points = hull->vertices;
        proxy.count = hull->vertexCount;
        proxy.radius = 0.0f;
        return proxy;
    }
    if (type == M3_CAPSULE)
    {
        scratch[0] = v->geom[s].v;
        scratch[1] = v->geom[s].v2;
        proxy.points = scratch;
        proxy.count = 2;
        proxy.radius = v->geom[s].s;
        return proxy;
    }
    scratch[0] = v->geom[s].v;
    proxy.points = scratch;
    proxy.count = 1;
    proxy.radius = v->geom[s].s;
    return proxy;
}

M3_HD void m3SwapAnchors(m3Manifold* m, int32_t count)
{
    m->normal = m3Neg3(m->normal);
    for (int32_t k = 0; k < count; ++k)
    {
        m3Vec3 tmp = m->points[k].anchorA;
        m->points[k].anchorA = m->points[k].anchorB;
        m->points[k].anchorB = tmp;
    }
}

M3_HD m3Manifold m3CollidePlaneHull(const m3View* v, int32_t shapeA, int32_t planeShape,
                                    int32_t hullShape)
{
    m3Manifold fresh = m3ZeroManifold();
    const m3HullData* hull = &v->hulls[v->hullIndex[hullShape]];
    int32_t planeBody = v->shapeBody[planeShape];
    int32_t hullBody = v->shapeBody[hullShape];
    m3Transform xf = m3ShapeXf(v, hullShape);
    m3Vec3 n = v->geom[planeShape].v;
    m3real offset = v->geom[planeShape].s;
    // Deepest four below the margin (ties to the lower vertex), kept in
    // ascending vertex order: a running top-4 is the same selection as the
    // reference's repeated min scan over all candidates.
    int32_t kIdx[M3_MANIFOLD_MAX_POINTS];
    m3real kSep[M3_MANIFOLD_MAX_POINTS];
    double kW[M3_MANIFOLD_MAX_POINTS][3];
    int32_t kc = 0;
    for (int32_t vi = 0; vi < hull->vertexCount; ++vi)
    {
        m3Vec3 r = m3RotateVec3(xf.q, hull->vertices[vi]);
        double wx = xf.p.x + (double)r.x;
        double wy = xf.p.y + (double)r.y;
        double wz = xf.p.z + (double)r.z;
        double sep = (double)n.x * wx + (double)n.y * wy + (double)n.z * wz - (double)offset;
        m3real fs = (m3real)sep;
        if (!(fs < M3_SPECULATIVE_DISTANCE))
        {
            continue;
        }
        // Insert if among the four deepest so far: candidates arrive in
        // ascending index, so a strictly smaller separation wins and ties
        // keep the earlier (lower) index.
        int32_t pos = kc;
        while (pos > 0 && fs < kSep[pos - 1])
        {
            pos -= 1;
        }
        if (pos >= M3_MANIFOLD_MAX_POINTS)
        {
            continue;
        }
        int32_t last = kc < M3_MANIFOLD_MAX_POINTS ? kc : M3_MANIFOLD_MAX_POINTS - 1;
        for (int32_t q = last; q > pos; --q)
        {
            kIdx[q] = kIdx[q - 1];
            kSep[q] = kSep[q - 1];
            kW[q][0] = kW[q - 1][0];
            kW[q][1] = kW[q - 1][1];
            kW[q][2] = kW[q - 1][2];
        }
        kIdx[pos] = vi;
        kSep[pos] = fs;
        kW[pos][0] = wx;
        kW[pos][1] = wy;
        kW[pos][2] = wz;
        if (kc < M3_MANIFOLD_MAX_POINTS)
        {
            kc += 1;
        }
    }
    // Ascending vertex index among the kept.
    for (int32_t a = 0; a < kc; ++a)
    {
        for (int32_t b = a + 1; b < kc; ++b)
        {
            if (kIdx[b] < kIdx[a])
            {
                int32_t ti = kIdx[a];
                kIdx[a] = kIdx[b];
                kIdx[b] = ti;
                m3real ts = kSep[a];
                kSep[a] = kSep[b];
                kSep[b] = ts;
                for (int32_t c = 0; c < 3; ++c)
                {
                    double tw = kW[a][c];
                    kW[a][c] = kW[b][c];
                    kW[b][c] = tw;
                }
            }
        }
    }
    fresh.normal = n;
    fresh.pointCount = kc;
    for (int32_t k = 0; k < kc; ++k)
    {
        double sepD = (double)kSep[k];
        double px = kW[k][0] - (double)n.x * sepD;
        double py = kW[k][1] - (double)n.y * sepD;
        double pz = kW[k][2] - (double)n.z * sepD;
        fresh.points[k].anchorA = m3FromCom(v, planeBody, px, py, pz);
        fresh.points[k].anchorB = m3FromCom(v, hullBody, kW[k][0], kW[k][1], kW[k][2]);
        fresh.points[k].separation = kSep[k];
        fresh.points[k].id = (uint16_t)kIdx[k];
    }
    if (planeShape != shapeA)
    {
        m3SwapAnchors(&fresh, kc);
    }
    return fresh;
}

// B's pose in A's frame, as the hull pair is evaluated (src/manifold.c:2526-2536).
M3_HD void m3HullPairFrame(const m3View* v, int32_t shapeA, int32_t shapeB, m3Transform* xfA,
                           m3Quat* qRel, m3Vec3* pRel)
{
    *xfA = m3ShapeXf(v, shapeA);
    m3Transform xfB = m3ShapeXf(v, shapeB);
    *qRel = m3MulQuat(m3Conj(xfA->q), xfB.q);
    m3Vec3 dp = m3V3((m3real)(xfB.p.x - xfA->p.x), (m3real)(xfB.p.y - xfA->p.y),
                     (m3real)(xfB.p.z - xfA->p.z));
    *pRel = m3InvRotateVec3(xfA->q, dp);
}

M3_HD m3Manifold m3CollidePairHulls(const m3View* v, int32_t shapeA, int32_t shapeB,
                                    const m3SatQuery* sat)
{
    int32_t bodyA = v->shapeBody[shapeA];
    int32_t bodyB = v->shapeBody[shapeB];
    m3Transform xfA;
    m3Quat qRel;
    m3Vec3 pRel;
    m3HullPairFrame(v, shapeA, shapeB, &xfA, &qRel, &pRel);
    m3Manifold fresh = m3CollideHullsQ(&v->hulls[v->hullIndex[shapeA]],
                                       &v->hulls[v->hullIndex[shapeB]], qRel, pRel, sat);
    if (fresh.pointCount > 0)
    {
        fresh.normal = m3RotateVec3(xfA.q, fresh.normal);
        for (int32_t k = 0; k < fresh.pointCount; ++k)
        {
            m3Vec3 rA = m3RotateVec3(xfA.q, fresh.points[k].anchorA);
            m3Vec3 rB = m3RotateVec3(xfA.q, fresh.points[k].anchorB);
            fresh.points[k].anchorA = m3FromCom(v, bodyA, xfA.p.x + (double)rA.x,
                                                xfA.p.y + (double)rA.y, xfA.p.z + (double)rA.z);
            fresh.points[k].anchorB = m3FromCom(v, bodyB, xfA.p.x + (double)rB.x,
                                                xfA.p.y + (double)rB.y, xfA.p.z + (double)rB.z);
        }
    }
    return fresh;
}

M3_HD m3Manifold m3CollidePlaneCapsule(const m3View* v, int32_t shapeA, int32_t planeShape,
                                       int32_t capShape)
{
    m3Manifold fresh = m3ZeroManifold();
    int32_t planeBody = v->shapeBody[planeShape];
    int32_t capBody = v->shapeBody[capShape];
    m3Transform xf = m3ShapeXf(v, capShape);
    m3Vec3 n = v->geom[planeShape].v;
    m3real offset = v->geom[planeShape].s;
    m3real radius = v->geom[capShape].s;
    m3Vec3 caps[2] = {v->geom[capShape].v, v->geom[capShape].v2};
    int32_t count = 0;
    for (int32_t k = 0; k < 2; ++k)
    {
        m3Vec3 r = m3RotateVec3(xf.q, caps[k]);
        double wx = xf.p.x + (double)r.x;
        double wy = xf.p.y + (double)r.y;
        double wz = xf.p.z + (double)r.z;
        double centerDist = (double)n.x * wx + (double)n.y * wy + (double)n.z * wz - (double)offset;
        m3real sep = (m3real)centerDist - radius;
        if (sep > M3_SPECULATIVE_DISTANCE)
        {
            continue;
        }
        double px = wx - (double)n.x * centerDist;
        double py = wy - (double)n.y * centerDist;
        double pz = wz - (double)n.z * centerDist;
        fresh.points[count].anchorA = m3FromCom(v, planeBody, px, py, pz);
        fresh.points[count].anchorB =
            m3FromCom(v, capBody, wx - (double)(n.x * radius), wy - (double)(n.y * radius),
                      wz - (double)(n.z * radius));
        fresh.points[count].separation = sep;
        fresh.points[count].id = (uint16_t)k;
        count += 1;
    }
    fresh.normal = n;
    fresh.pointCount = count;
    if (count > 0 && planeShape != shapeA)
    {
        m3SwapAnchors(&fresh, count);
    }
    return fresh;
}

// The capsule's segment in the hull shape's frame; returns the hull shape's
// world transform.
M3_HD m3Transform m3CapsuleInHull(const m3View* v, int32_t hullShape, int32_t capShape, m3Vec3* s1,
                                  m3Vec3* s2)
{
    m3Transform xfH = m3ShapeXf(v, hullShape);
    m3Transform xfC = m3ShapeXf(v, capShape);
    m3Quat qRel = m3MulQuat(m3Conj(xfH.q), xfC.q);
    m3Vec3 dp = m3V3((m3real)(xfC.p.x - xfH.p.x), (m3real)(xfC.p.y - xfH.p.y),
                     (m3real)(xfC.p.z - xfH.p.z));
    m3Vec3 pRel = m3InvRotateVec3(xfH.q, dp);
    *s1 = m3Add3(m3RotateVec3(qRel, v->geom[capShape].v), pRel);
    *s2 = m3Add3(m3RotateVec3(qRel, v->geom[capShape].v2), pRel);
    return xfH;
}

// `sat` as for m3CollideSegmentHullQ (the queries of the segment against
// the hull, evaluated ahead), or null.
M3_HD m3Manifold m3CollideCapsuleHull(const m3View* v, int32_t shapeA, int32_t hullShape,
                                      int32_t capShape, const m3SatQuery* sat)
{
    m3Manifold fresh = m3ZeroManifold();
    int32_t hullBody = v->shapeBody[hullShape];
    int32_t capBody = v->shapeBody[capShape];
    m3Vec3 s1;
    m3Vec3 s2;
    m3Transform xfH = m3CapsuleInHull(v, hullShape, capShape, &s1, &s2);
    m3Manifold local =
        m3CollideSegmentHullQ(&v->hulls[v->hullIndex[hullShape]], s1, s2, v->geom[capShape].s, sat);
    if (local.pointCount > 0)
    {
        m3Vec3 nWorld = m3RotateVec3(xfH.q, local.normal);
        fresh.normal = hullShape == shapeA ? nWorld : m3Neg3(nWorld);
        fresh.pointCount = local.pointCount;
        for (int32_t k = 0; k < local.pointCount; ++k)
        {
            m3Vec3 rH = m3RotateVec3(xfH.q, local.points[k].anchorA);
            m3Vec3 rC = m3RotateVec3(xfH.q, local.points[k].anchorB);
            m3Vec3 aHull = m3FromCom(v, hullBody, xfH.p.x + (double)rH.x, xfH.p.y + (double)rH.y,
                                     xfH.p.z + (double)rH.z);
            m3Vec3 aCap = m3FromCom(v, capBody, xfH.p.x + (double)rC.x, xfH.p.y + (double)rC.y,
                                    xfH.p.z + (double)rC.z);
            fresh.points[k].anchorA = hullShape == shapeA ? aHull : aCap;
            fresh.points[k].anchorB = hullShape == shapeA ? aCap : aHull;
            fresh.points[k].separation = local.points[k].separation;
            fresh.points[k].id = local.points[k].id;
        }
    }
    return fresh;
}

// Hull-sphere, capsule-sphere and capsule-capsule: GJK on the cores, radii
// analytic, one point (src/manifold.c:2666-2808).
M3_HD m3Manifold m3CollideGjkPair(const m3View* v, int32_t shapeA, int32_t shapeB, int hullPair)
{
    m3Manifold fresh = m3ZeroManifold();
    uint8_t typeA = v->shapeType[shapeA];
    int32_t bodyA = v->shapeBody[shapeA];
    int32_t bodyB = v->shapeBody[shapeB];
    m3Transform xfA = m3ShapeXf(v, shapeA);
    m3Transform xfB = m3ShapeXf(v, shapeB);
    m3Quat q = m3MulQuat(m3Conj(xfA.q), xfB.q);
    m3Vec3 dp = m3V3((m3real)(xfB.p.x - xfA.p.x), (m3real)(xfB.p.y - xfA.p.y),
                     (m3real)(xfB.p.z - xfA.p.z));
    m3Vec3 p = m3InvRotateVec3(xfA.q, dp);
    m3Vec3 pointsA[2];
    m3Vec3 pointsB[2];
    m3Proxy proxyA = m3MakeProxy(v, shapeA, pointsA);
    m3Proxy proxyB = m3MakeProxy(v, shapeB, pointsB);
    m3GjkOut out = m3ShapeDistance(&proxyA, &proxyB, q, p);
    m3real rA = proxyA.radius;
    m3real rB = proxyB.radius;
    m3real sep = out.distance - rA - rB;
    if (!(sep <= M3_SPECULATIVE_DISTANCE))
    {
        return fresh;
    }
    m3Vec3 nLocal;
    m3Vec3 pALocal;
    m3Vec3 pBLocal;
    if (out.distance > 0.0f)
    {
        nLocal = out.normal;
        pALocal = m3Add3(out.pointA, m3MulSV3(rA, nLocal));
        pBLocal = m3Sub3(out.pointB, m3MulSV3(rB, nLocal));
    }
    else if (hullPair)
    {
        int hullIsA = typeA == M3_HULL;
        const m3HullData* hull = &v->hulls[v->hullIndex[hullIsA ? shapeA : shapeB]];
        const m3Proxy* round = hullIsA ? &proxyB : &proxyA;
        m3Vec3 core = hullIsA ? m3Add3(m3RotateVec3(q, round->points[0]), p)
                              : m3InvRotateVec3(q, m3Sub3(round->points[0], p));
        m3Vec3 nHull;
        m3real coreSep;
        m3Vec3 onHull;
        m3DeepPointInHull(hull, core, &nHull, &coreSep, &onHull);
        m3real rRound = round->radius;
        sep = coreSep - rRound;
        if (hullIsA)
        {
            nLocal = nHull;
            pALocal = onHull;
            pBLocal = m3Sub3(core, m3MulSV3(rRound, nHull));
        }
        else
        {
            m3Vec3 nA = m3RotateVec3(q, nHull);
            nLocal = m3Neg3(nA);
            pALocal = m3Sub3(m3Add3(m3RotateVec3(q, core), p), m3MulSV3(rRound, nA));
            pBLocal = m3Add3(m3RotateVec3(q, onHull), p);
        }
    }
    else
    {
        m3Vec3 axis;
        m3Vec3 cA;
        m3Vec3 cB;
        if (proxyA.count == 2 && proxyB.count == 2)
        {
            m3Vec3 dirA = m3Sub3(pointsA[1], pointsA[0]);
            m3Vec3 dirB = m3RotateVec3(q, m3Sub3(pointsB[1], pointsB[0]));
            axis = m3Cross3(dirA, dirB);
            cA = m3MulSV3(0.5f, m3Add3(pointsA[0], pointsA[1]));
            cB = m3Add3(m3RotateVec3(q, m3MulSV3(0.5f, m3Add3(pointsB[0], pointsB[1]))), p);
            if (m3Dot3(axis, axis) < 1.0e-10f)
            {
                axis = m3Sub3(cB, cA);
            }
        }
        else
        {
            m3Vec3 dir = proxyA.count == 2 ? m3Sub3(pointsA[1], pointsA[0])
                                           : m3RotateVec3(q, m3Sub3(pointsB[1], pointsB[0]));
            m3Vec3 t1;
            m3Vec3 t2;
            m3TangentBasis(m3Normalize3(dir), &t1, &t2);
            axis = t1;
            cA = proxyA.count == 1 ? pointsA[0] : m3MulSV3(0.5f, m3Add3(pointsA[0], pointsA[1]));
            cB = proxyB.count == 1
                     ? m3Add3(m3RotateVec3(q, pointsB[0]), p)
                     : m3Add3(m3RotateVec3(q, m3MulSV3(0.5f, m3Add3(pointsB[0], pointsB[1]))), p);
        }
        nLocal = m3Normalize3(axis);
        if (m3Dot3(nLocal, m3Sub3(cB, cA)) < 0.0f)
        {
            nLocal = m3Neg3(nLocal);
        }
        m3Vec3 mid = m3MulSV3(0.5f, m3Add3(cA, cB));
        pALocal = mid;
        pBLocal = mid;
        sep = -(rA + rB);
    }
    fresh.normal = m3RotateVec3(xfA.q, nLocal);
    m3Vec3 wA = m3RotateVec3(xfA.q, pALocal);
    m3Vec3 wB = m3RotateVec3(xfA.q, pBLocal);
    fresh.points[0].anchorA =
        m3FromCom(v, bodyA, xfA.p.x + (double)wA.x, xfA.p.y + (double)wA.y, xfA.p.z + (double)wA.z);
    fresh.points[0].anchorB =
        m3FromCom(v, bodyB, xfA.p.x + (double)wB.x, xfA.p.y + (double)wB.y, xfA.p.z + (double)wB.z);
    fresh.points[0].separation = sep;
    fresh.points[0].id = 0;
    fresh.pointCount = 1;
    return fresh;
}

M3_HD void m3SphereCenter(const m3View* v, int32_t s, double* cx, double* cy, double* cz)
{
    m3Transform xf = m3ShapeXf(v, s);
    m3Vec3 r = m3RotateVec3(xf.q, v->geom[s].v);
    *cx = xf.p.x + (double)r.x;
    *cy = xf.p.y + (double)r.y;
    *cz = xf.p.z + (double)r.z;
}

M3_HD m3Manifold m3CollidePlaneSpherePair(const m3View* v, int32_t shapeA, int32_t shapeB)
{
    uint8_t typeA = v->shapeType[shapeA];
    int32_t planeShape = typeA == M3_PLANE ? shapeA : shapeB;
    int32_t sphereShape = planeShape == shapeA ? shapeB : shapeA;
    m3Vec3 n = v->geom[planeShape].v;
    m3real offset = v->geom[planeShape].s;
    double cx, cy, cz;
    m3SphereCenter(v, sphereShape, &cx, &cy, &cz);
    double distD = (double)n.x * cx + (double)n.y * cy + (double)n.z * cz - (double)offset;
    m3Manifold fresh = m3CollidePlaneSphere(n, (m3real)distD, v->geom[sphereShape].s);
    if (fresh.pointCount > 0)
    {
        int32_t planeBody = v->shapeBody[planeShape];
        int32_t ballBody = v->shapeBody[sphereShape];
        double px = cx - (double)(fresh.normal.x * (m3real)distD);
        double py = cy - (double)(fresh.normal.y * (m3real)distD);
        double pz = cz - (double)(fresh.normal.z * (m3real)distD);
        fresh.points[0].anchorA = m3FromCom(v, planeBody, px, py, pz);
        fresh.points[0].anchorB =
            m3Add3(m3FromCom(v, ballBody, cx, cy, cz), fresh.points[0].anchorB);
        if (planeShape != shapeA)
        {
            m3Vec3 tmp = fresh.points[0].anchorA;
            fresh.points[0].anchorA = fresh.points[0].anchorB;
            fresh.points[0].anchorB = tmp;
            fresh.normal = m3Neg3(fresh.normal);
        }
    }
    return fresh;
}

M3_HD m3Manifold m3CollideSpherePair(const m3View* v, int32_t shapeA, int32_t shapeB)
{
    double ax, ay, az, bx, by, bz;
    m3SphereCenter(v, shapeA, &ax, &ay, &az);
    m3SphereCenter(v, shapeB, &bx, &by, &bz);
    m3Vec3 d = m3V3((m3real)(bx - ax), (m3real)(by - ay), (m3real)(bz - az));
    m3Manifold fresh = m3CollideSpheres(d, v->geom[shapeA].s, v->geom[shapeB].s);
    if (fresh.pointCount > 0)
    {
        fresh.points[0].anchorA =
            m3Add3(m3FromCom(v, v->shapeBody[shapeA], ax, ay, az), fresh.points[0].anchorA);
        fresh.points[0].anchorB =
            m3Add3(m3FromCom(v, v->shapeBody[shapeB], bx, by, bz), fresh.points[0].anchorB);
    }
    return fresh;
}

// One pair: fresh manifold by shape family, then the warm-start carry by
// feature id from the previous step's sorted list.
// Index of key in the ascending, duplicate-free keys[0, n), or -1.  A few
// interpolation probes come first (pair keys spread evenly enough that they
// land on or beside the key), then bisection of the range that remains.
// Every probe keeps keys[lo] < key < keys[hi], so the answer is the same as
// plain bisection's; only the number of dependent loads differs.
M3_HD int32_t m3FindKey(const uint64_t* keys, int32_t n, uint64_t key)
{
    if (n <= 0)
    {
        return -1;
    }
    int32_t lo = 0;
    int32_t hi = n - 1;
    uint64_t klo = keys[lo];
    uint64_t khi = keys[hi];
    if (key == klo)
    {
        return lo;
    }
    if (key == khi)
    {
        return hi;
    }
    if (key < klo || key > khi)
    {
        return -1;
    }
    for (int probe = 0; probe < 3 && hi - lo > 1; ++probe)
    {
        const double f = (double)(key - klo) / (double)(khi - klo);
        int32_t mid = lo + (int32_t)(f * (double)(hi - lo));
        mid = mid <= lo ? lo + 1 : (mid >= hi ? hi - 1 : mid);
        const uint64_t k = keys[mid];
        if (k == key)
        {
            return mid;
        }
        if (k < key)
        {
            lo = mid;
            klo = k;
        }
        else
        {
            hi = mid;
            khi = k;
        }
    }
    // keys[lo] < key < keys[hi]: bisect the open interval.
    lo += 1;
    hi -= 1;
    while (lo <= hi)
    {
        const int32_t mid = (lo + hi) / 2;
        const uint64_t k = keys[mid];
        if (k == key)
        {
            return mid;
        }
        if (k < key)
        {
            lo = mid + 1;
        }
        else
        {
            hi = mid - 1;
        }
    }
    return -1;
}

// Returns the new manifold's point count.
M3_HD int32_t m3NarrowPair(const m3View* v, int32_t i, int32_t oldCount)
{
    uint64_t key = v->keys[i];
    int32_t shapeA = (int32_t)(key >> 32);
    int32_t shapeB = (int32_t)(key & 0xFFFFFFFFu);
    uint8_t typeA = v->shapeType[shapeA];
    uint8_t typeB = v->shapeType[shapeB];
    int planePair = typeA == M3_PLANE || typeB == M3_PLANE;
    int hullPair = typeA == M3_HULL || typeB == M3_HULL;
    int capsulePair = typeA == M3_CAPSULE || typeB == M3_CAPSULE;
    m3Manifold fresh;
    if (planePair && hullPair)
    {
        int32_t planeShape = typeA == M3_PLANE ? shapeA : shapeB;
        int32_t hullShape = planeShape == shapeA ? shapeB : shapeA;
        fresh = m3CollidePlaneHull(v, shapeA, planeShape, hullShape);
    }
    else if (typeA == M3_HULL && typeB == M3_HULL)
    {
        const m3HullData* hullA = &v->hulls[v->hullIndex[shapeA]];
        const m3HullData* hullB = &v->hulls[v->hullIndex[shapeB]];
        const m3SatQuery* sat =
            v->satQueries != NULL && m3WideHullPair(hullA, hullB, v->satBoxes) ? &v->satQueries[i] : NULL;
        fresh = m3CollidePairHulls(v, shapeA, shapeB, sat);
    }
    else if (planePair && capsulePair)
    {
        int32_t planeShape = typeA == M3_PLANE ? shapeA : shapeB;
        int32_t capShape = planeShape == shapeA ? shapeB : shapeA;
        fresh = m3CollidePlaneCapsule(v, shapeA, planeShape, capShape);
    }
    else if (hullPair && capsulePair)
    {
        int32_t hullShape = typeA == M3_HULL ? shapeA : shapeB;
        int32_t capShape = hullShape == shapeA ? shapeB : shapeA;
        const m3SatQuery* sat =
            v->satQueries != NULL && m3WideSegmentHull(&v->hulls[v->hullIndex[hullShape]]) ? &v->satQueries[i]
                                                                                           : NULL;
        fresh = m3CollideCapsuleHull(v, shapeA, hullShape, capShape, sat);
    }
    else if (hullPair || capsulePair)
    {
        fresh = m3CollideGjkPair(v, shapeA, shapeB, hullPair);
    }
    else if (planePair)
    {
        fresh = m3CollidePlaneSpherePair(v, shapeA, shapeB);
    }
    else
    {
        fresh = m3CollideSpherePair(v, shapeA, shapeB);
    }

    const int32_t mid = fresh.pointCount > 0 ? m3FindKey(v->oldKeys, oldCount, key) : -1;
    if (mid >= 0)
    {
        const m3Manifold* prev = &v->oldManifolds[mid];
        fresh.frictionImpulse = prev->frictionImpulse;
        fresh.twistImpulse = prev->twistImpulse;
        fresh.rollingImpulse = prev->rollingImpulse;
        int32_t prevCount = prev->pointCount;
        for (int32_t k2 = 0; k2 < fresh.pointCount; ++k2)
        {
            for (int32_t o = 0; o < prevCount; ++o)
            {
                if (prev->points[o].id == fresh.points[k2].id)
                {
                    fresh.points[k2].normalImpulse = prev->points[o].normalImpulse;
                    fresh.points[k2].flags |= 1;
                    break;
                }
            }
        }
    }
    v->manifolds[i] = fresh;
    return fresh.pointCount;
}

// ---------------------------------------------------------------------------
// Contact solver (src/solver.c:30-600)
// ---------------------------------------------------------------------------

M3_HD m3Softness m3MakeSoft(m3real hertz, m3real zeta, m3real h)
{
    m3Softness s;
    if (hertz == 0.0f)
    {
        s.biasRate = 0.0f;
        s.massScale = 0.0f;
        s.impulseScale = 0.0f;
        return s;
    }
    m3real omega = 2.0f * M3_PI * hertz;
    m3real a1 = 2.0f * zeta + h * omega;
    m3real a2 = h * omega * a1;
    m3real a3 = 1.0f / (1.0f + a2);
    s.biasRate = omega / a1;
    s.massScale = a2 * a3;
    s.impulseScale = a3;
    return s;
}

// I_w^-1 = R I_l^-1 R^T built column by column on the world basis.
M3_HD m3Mat3 m3WorldInvInertia(const m3View* v, int32_t body)
{
    if (v->type[body] != M3_DYNAMIC)
    {
        return m3ZeroMat3();
    }
    m3Quat q = v->xf[body].q;
    m3Mat3 il = v->invI[body];
    m3Mat3 r;
    r.cx = m3RotateVec3(q, m3MulMV3(il, m3InvRotateVec3(q, m3V3(1.0f, 0.0f, 0.0f))));
    r.cy = m3RotateVec3(q, m3MulMV3(il, m3InvRotateVec3(q, m3V3(0.0f, 1.0f, 0.0f))));
    r.cz = m3RotateVec3(q, m3MulMV3(il, m3InvRotateVec3(q, m3V3(0.0f, 0.0f, 1.0f))));
    return r;
}

M3_HD m3real m3EffectiveMass(const m3Constraint* c, m3Vec3 rA, m3Vec3 rB, m3Vec3 dir)
{
    m3Vec3 arm1 = m3Cross3(rA, dir);
    m3Vec3 arm2 = m3Cross3(rB, dir);
    m3real k = c->invMassA + c->invMassB + m3Dot3(arm1, m3MulMV3(c->invIA, arm1)) +
               m3Dot3(arm2, m3MulMV3(c->invIB, arm2));
    return k > 0.0f ? 1.0f / k : 0.0f;
}

M3_HD m3Vec3 m3VelocityAt(const m3View* v, int32_t body, m3Vec3 arm)
{
    return m3Add3(v->lv[body], m3Cross3(v->av[body], arm));
}

// Does pair i become a constraint this step (src/solver.c:156-186)?
M3_HD int m3PairIsConstraint(const m3View* v, int32_t i)
{
    const m3Manifold* m = &v->manifolds[i];
    if (m->pointCount == 0)
    {
        return 0;
    }
    uint64_t key = v->keys[i];
    int32_t shapeA = (int32_t)(key >> 32);
    int32_t shapeB = (int32_t)(key & 0xFFFFFFFFu);
    int32_t bodyA = v->shapeBody[shapeA];
    int32_t bodyB = v->shapeBody[shapeB];
    int awakeDynA = v->type[bodyA] == M3_DYNAMIC && v->awake[bodyA] != 0;
    int awakeDynB = v->type[bodyB] == M3_DYNAMIC && v->awake[bodyB] != 0;
    if (!awakeDynA && !awakeDynB)
    {
        return 0;
    }
    if (v->sensor[shapeA] != 0 || v->sensor[shapeB] != 0)
    {
        return 0;
    }
    if (v->veto != NULL && v->veto[i] != 0)
    {
        return 0; // vetoed for this step (src/solver.c:184-248)
    }
    return 1;
}

// The pre-solve callback's inputs for pair i (src/solver.c:214-236): pairs
// that would become constraints and where either shape opted in.  The point
// is the deepest manifold point (ties to the lower index) in world space from
// body B's centre of mass; the normal points from A to B.  Returns 0 when
// the pair is no candidate.
M3_HD int m3PreSolveCandidate(const m3View* v, int32_t i, double point[3], m3Vec3* normal)
{
    const m3Manifold* m = &v->manifolds[i];
    if (m->pointCount == 0)
    {
        return 0;
    }
    const uint64_t key = v->keys[i];
    const int32_t shapeA = (int32_t)(key >> 32);
    const int32_t shapeB = (int32_t)(key & 0xFFFFFFFFu);
    const int32_t bodyA = v->shapeBody[shapeA];
    const int32_t bodyB = v->shapeBody[shapeB];
    const int awakeDynA = v->type[bodyA] == M3_DYNAMIC && v->awake[bodyA] != 0;
    const int awakeDynB = v->type[bodyB] == M3_DYNAMIC && v->awake[bodyB] != 0;
    if ((!awakeDynA && !awakeDynB) || v->sensor[shapeA] != 0 || v->sensor[shapeB] != 0 ||
        (v->preSolve[shapeA] == 0 && v->preSolve[shapeB] == 0))
    {
        return 0;
    }
    int32_t deep = 0;
    for (int32_t k = 1; k < m->pointCount; ++k)
    {
        if (m->points[k].separation < m->points[deep].separation)
        {
            deep = k;
        }
    }
    const m3Transform xf = v->xf[bodyB];
    const m3Vec3 lc = m3RotateVec3(xf.q, v->lc[bodyB]);
    point[0] = xf.p.x + (double)lc.x + (double)m->points[deep].anchorB.x;
    point[1] = xf.p.y + (double)lc.y + (double)m->points[deep].anchorB.y;
    point[2] = xf.p.z + (double)lc.z + (double)m->points[deep].anchorB.z;
    *normal = m->normal;
    return 1;
}

// src/solver.c:249-403 for one constraint.
M3_HD void m3PrepareContact(const m3View* v, int32_t pairIndex, m3Constraint* c, m3Softness soft,
                            m3Softness staticSoft)
{
    const m3Manifold* manifold = &v->manifolds[pairIndex];
    uint64_t key = v->keys[pairIndex];
    int32_t shapeA = (int32_t)(key >> 32);
    int32_t shapeB = (int32_t)(key & 0xFFFFFFFFu);
    int32_t bodyA = v->shapeBody[shapeA];
    int32_t bodyB = v->shapeBody[shapeB];
    memset(c, 0, sizeof(*c));
    c->bodyA = bodyA;
    c->bodyB = bodyB;
    c->manifoldIndex = pairIndex;
    c->pointCount = manifold->pointCount;
    c->normal = manifold->normal;
    m3TangentBasis(c->normal, &c->t1, &c->t2);
    c->dynA = v->type[bodyA] == M3_DYNAMIC ? 1 : 0;
    c->dynB = v->type[bodyB] == M3_DYNAMIC ? 1 : 0;
    c->invMassA = c->dynA ? v->invMass[bodyA] : 0.0f;
    c->invMassB = c->dynB ? v->invMass[bodyB] : 0.0f;
    c->invIA = m3WorldInvInertia(v, bodyA);
    c->invIB = m3WorldInvInertia(v, bodyB);
    m3real fricA = v->friction[shapeA];
    m3real fricB = v->friction[shapeB];
    m3real restA = v->restitution[shapeA];
    m3real restB = v->restitution[shapeB];
    m3real rollA = v->rolling[shapeA];
    m3real rollB = v->rolling[shapeB];
    m3Vec3 surfA = v->surfaceVel[shapeA];
    m3Vec3 surfB = v->surfaceVel[shapeB];
    c->friction = m3Sqrt(fricA * fricB);
    c->restitution = m3MaxF(restA, restB);
    c->rollingResistance = m3MaxF(rollA, rollB) * m3MaxF(v->maxExt[bodyA], v->maxExt[bodyB]);
    c->rollingImpulse = manifold->rollingImpulse;
    if (c->rollingResistance > 0.0f)
    {
        m3Mat3 sum = c->invIA;
        sum.cx = m3Add3(sum.cx, c->invIB.cx);
        sum.cy = m3Add3(sum.cy, c->invIB.cy);
        sum.cz = m3Add3(sum.cz, c->invIB.cz);
        c->rollingK = sum;
    }
    c->softness = (c->invMassA == 0.0f || c->invMassB == 0.0f) ? staticSoft : soft;
    for (int32_t k = 0; k < c->pointCount; ++k)
    {
        m3CPoint* cp = &c->points[k];
        cp->rA = manifold->points[k].anchorA;
        cp->rB = manifold->points[k].anchorB;
        cp->baseSeparation = manifold->points[k].separation -
                             (m3Dot3(cp->rB, c->normal) - m3Dot3(cp->rA, c->normal));
        cp->normalMass = m3EffectiveMass(c, cp->rA, cp->rB, c->normal);
        cp->relativeVelocity =
            m3Dot3(m3Sub3(m3VelocityAt(v, bodyB, cp->rB), m3VelocityAt(v, bodyA, cp->rA)),
                   c->normal);
        cp->normalImpulse = manifold->points[k].normalImpulse;
        cp->totalNormalImpulse = 0.0f;
    }
    m3Vec3 originA = m3V3(0.0f, 0.0f, 0.0f);
    m3Vec3 originB = m3V3(0.0f, 0.0f, 0.0f);
    for (int32_t k = 0; k < c->pointCount; ++k)
    {
        originA = m3Add3(originA, c->points[k].rA);
        originB = m3Add3(originB, c->points[k].rB);
    }
    m3real invCount = 1.0f / (m3real)c->pointCount;
    c->originA = m3MulSV3(invCount, originA);
    c->originB = m3MulSV3(invCount, originB);
    for (int32_t k = 0; k < c->pointCount; ++k)
    {
        c->points[k].leverArm = m3Length3(m3Sub3(c->points[k].rA, c->originA));
    }
    m3Vec3 rtA1 = m3Cross3(c->originA, c->t1);
    m3Vec3 rtA2 = m3Cross3(c->originA, c->t2);
    m3Vec3 rtB1 = m3Cross3(c->originB, c->t1);
    m3Vec3 rtB2 = m3Cross3(c->originB, c->t2);
    m3real k11 = c->invMassA + c->invMassB + m3Dot3(rtA1, m3MulMV3(c->invIA, rtA1)) +
                 m3Dot3(rtB1, m3MulMV3(c->invIB, rtB1));
    m3real k22 = c->invMassA + c->invMassB + m3Dot3(rtA2, m3MulMV3(c->invIA, rtA2)) +
                 m3Dot3(rtB2, m3MulMV3(c->invIB, rtB2));
    m3real k12 = m3Dot3(rtA1, m3MulMV3(c->invIA, rtA2)) + m3Dot3(rtB1, m3MulMV3(c->invIB, rtB2));
    m3real det = k11 * k22 - k12 * k12;
    if (det > 0.0f)
    {
        m3real invDet = 1.0f / det;
        c->frictionK11 = k22 * invDet;
        c->frictionK22 = k11 * invDet;
        c->frictionK12 = -k12 * invDet;
    }
    else
    {
        c->frictionK11 = 0.0f;
        c->frictionK22 = 0.0f;
        c->frictionK12 = 0.0f;
    }
    c->frictionImpulse1 = m3Dot3(manifold->frictionImpulse, c->t1);
    c->frictionImpulse2 = m3Dot3(manifold->frictionImpulse, c->t2);
    m3Vec3 surf = m3Sub3(surfA, surfB);
    c->tangentVelocity1 = m3Dot3(surf, c->t1);
    c->tangentVelocity2 = m3Dot3(surf, c->t2);
    m3Vec3 iSumN = m3Add3(m3MulMV3(c->invIA, c->normal), m3MulMV3(c->invIB, c->normal));
    m3real twistK = m3Dot3(c->normal, iSumN);
    c->twistMass = twistK > 0.0f ? 1.0f / twistK : 0.0f;
    c->twistImpulse = manifold->twistImpulse;
}

M3_HD void m3ApplyImpulse(const m3View* v, const m3Constraint* c, m3Vec3 impulse, m3Vec3 rsA,
                          m3Vec3 rsB)
{
    if (c->dynA)
    {
        int32_t a = c->bodyA;
        v->lv[a] = m3Sub3(v->lv[a], m3MulSV3(c->invMassA, impulse));
        v->av[a] = m3Sub3(v->av[a], m3MulMV3(c->invIA, m3Cross3(rsA, impulse)));
    }
    if (c->dynB)
    {
        int32_t b = c->bodyB;
        v->lv[b] = m3Add3(v->lv[b], m3MulSV3(c->invMassB, impulse));
        v->av[b] = m3Add3(v->av[b], m3MulMV3(c->invIB, m3Cross3(rsB, impulse)));
    }
}

// src/solver.c:408-429
M3_HD void m3WarmStartOne(const m3View* v, m3Constraint* c)
{
    for (int32_t k = 0; k < c->pointCount; ++k)
    {
        m3CPoint* cp = &c->points[k];
        m3ApplyImpulse(v, c, m3MulSV3(cp->normalImpulse, c->normal), cp->rA, cp->rB);
    }
    m3Vec3 f = m3Add3(m3MulSV3(c->frictionImpulse1, c->t1), m3MulSV3(c->frictionImpulse2, c->t2));
    m3ApplyImpulse(v, c, f, c->originA, c->originB);
    m3Vec3 tw = m3Add3(m3MulSV3(c->twistImpulse, c->normal), c->rollingImpulse);
    if (c->dynA)
    {
        v->av[c->bodyA] = m3Sub3(v->av[c->bodyA], m3MulMV3(c->invIA, tw));
    }
    if (c->dynB)
    {
        v->av[c->bodyB] = m3Add3(v->av[c->bodyB], m3MulMV3(c->invIB, tw));
    }
}

// The warm-start updates of constraint c applied to one of its bodies only
// (side 0 = A, 1 = B), on that body's velocity held in registers.
// WarmStartOne's impulses depend on stored accumulators, never on
// velocities, so replaying each body's constraints in color order with this
// function performs exactly the operations the colored pass performs on
// that body, in the same order.  All fields are loaded before the first use.
M3_HD void m3WarmStartSide(const m3Constraint* c, int side, m3Vec3* linIo, m3Vec3* angIo)
{
    const int32_t pointCount = c->pointCount;
    const m3Vec3 normal = c->normal;
    m3Vec3 arm[M3_MANIFOLD_MAX_POINTS];
    m3real imp[M3_MANIFOLD_MAX_POINTS];
    M3_UNROLL
    for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        arm[k] = side == 0 ? c->points[k].rA : c->points[k].rB;
        imp[k] = c->points[k].normalImpulse;
    }
    const m3real invMass = side == 0 ? c->invMassA : c->invMassB;
    const m3Mat3 invI = side == 0 ? c->invIA : c->invIB;
    const m3Vec3 t1 = c->t1;
    const m3Vec3 t2 = c->t2;
    const m3real f1 = c->frictionImpulse1;
    const m3real f2 = c->frictionImpulse2;
    const m3Vec3 origin = side == 0 ? c->originA : c->originB;
    const m3real twist = c->twistImpulse;
    const m3Vec3 rolling = c->rollingImpulse;
    m3Vec3 lin = *linIo;
    m3Vec3 ang = *angIo;
    // Slots up to the largest point count among the lanes running together
    // are evaluated and those past a lane's own count discarded (the bound
    // is uniform across the wavefront; the loads were issued above).
#if defined(__HIP_DEVICE_COMPILE__)
    const int32_t slots = __ballot(pointCount > 2) != 0 ? (__ballot(pointCount > 3) != 0 ? 4 : 3)
                                                        : (__ballot(pointCount > 1) != 0 ? 2 : 1);
#else
    const int32_t slots = M3_MANIFOLD_MAX_POINTS;
#endif
    M3_UNROLL
    for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        if (k >= slots)
        {
            break;
        }
        m3Vec3 impulse = m3MulSV3(imp[k], normal);
        m3Vec3 nl;
        m3Vec3 na;
        if (side == 0)
        {
            nl = m3Sub3(lin, m3MulSV3(invMass, impulse));
            na = m3Sub3(ang, m3MulMV3(invI, m3Cross3(arm[k], impulse)));
        }
        else
        {
            nl = m3Add3(lin, m3MulSV3(invMass, impulse));
            na = m3Add3(ang, m3MulMV3(invI, m3Cross3(arm[k], impulse)));
        }
        if (k < pointCount)
        {
            lin = nl;
            ang = na;
        }
    }
    m3Vec3 f = m3Add3(m3MulSV3(f1, t1), m3MulSV3(f2, t2));
    if (side == 0)
    {
        lin = m3Sub3(lin, m3MulSV3(invMass, f));
        ang = m3Sub3(ang, m3MulMV3(invI, m3Cross3(origin, f)));
    }
    else
    {
        lin = m3Add3(lin, m3MulSV3(invMass, f));
        ang = m3Add3(ang, m3MulMV3(invI, m3Cross3(origin, f)));
    }
    m3Vec3 tw = m3Add3(m3MulSV3(twist, normal), rolling);
    if (side == 0)
    {
        ang = m3Sub3(ang, m3MulMV3(invI, tw));
    }
    else
    {
        ang = m3Add3(ang, m3MulMV3(invI, tw));
    }
    *linIo = lin;
    *angIo = ang;
}

// The warm-start updates that constraint c applies to one body (side 0 =
// A, subtracted; 1 = B, added), each already multiplied out: per live
// point the linear and angular change of its normal impulse, then the
// friction pair's, then the twist and rolling row's angular change.  They
// depend only on the constraint's stored accumulators and geometry, so any
// thread may evaluate them ahead; a body applying them one by one in color
// order performs the operations of m3WarmStartSide in the same order.
typedef struct m3WarmAdd
{
    m3Vec3 lin[M3_MANIFOLD_MAX_POINTS];
    m3Vec3 ang[M3_MANIFOLD_MAX_POINTS];
    m3Vec3 linF;
    m3Vec3 angF;
    m3Vec3 angT;
    int32_t pointCount;
    int32_t pad[3];
} m3WarmAdd;

// m3WarmStartSide's addends for both sides from the constraint's geometry
// and accumulators (normal impulses imp[], friction f1, f2, twist, rolling).
M3_HD void m3WarmAddends(int32_t pointCount, m3Vec3 normal, const m3Vec3* rA, const m3Vec3* rB,
                         const m3real* imp, m3real invMassA, m3real invMassB, const m3Mat3* invIA,
                         const m3Mat3* invIB, m3Vec3 t1, m3Vec3 t2, m3real f1, m3real f2,
                         m3Vec3 originA, m3Vec3 originB, m3real twist, m3Vec3 rolling, int dynA,
                         int dynB, m3WarmAdd* outA, m3WarmAdd* outB)
{
    const m3Vec3 f = m3Add3(m3MulSV3(f1, t1), m3MulSV3(f2, t2));
    const m3Vec3 tw = m3Add3(m3MulSV3(twist, normal), rolling);
    for (int side = 0; side < 2; ++side)
    {
        if (side == 0 ? !dynA : !dynB)
        {
            continue;
        }
        m3WarmAdd* o = side == 0 ? outA : outB;
        const m3real invMass = side == 0 ? invMassA : invMassB;
        const m3Mat3* invI = side == 0 ? invIA : invIB;
        const m3Vec3* arm = side == 0 ? rA : rB;
        M3_UNROLL
        for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
        {
            if (k < pointCount)
            {
                const m3Vec3 impulse = m3MulSV3(imp[k], normal);
                o->lin[k] = m3MulSV3(invMass, impulse);
                o->ang[k] = m3MulMV3(*invI, m3Cross3(arm[k], impulse));
            }
        }
        o->linF = m3MulSV3(invMass, f);
        o->angF = m3MulMV3(*invI, m3Cross3(side == 0 ? originA : originB, f));
        o->angT = m3MulMV3(*invI, tw);
        o->pointCount = pointCount;
    }
}

// Applies one side's addends (m3WarmAddends) in m3WarmStartSide's order.
M3_HD void m3WarmApply(const m3WarmAdd* a, int side, m3Vec3* linIo, m3Vec3* angIo)
{
    m3Vec3 lin = *linIo;
    m3Vec3 ang = *angIo;
    const int32_t pointCount = a->pointCount;
    M3_UNROLL
    for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        if (k < pointCount)
        {
            lin = side == 0 ? m3Sub3(lin, a->lin[k]) : m3Add3(lin, a->lin[k]);
            ang = side == 0 ? m3Sub3(ang, a->ang[k]) : m3Add3(ang, a->ang[k]);
        }
    }
    lin = side == 0 ? m3Sub3(lin, a->linF) : m3Add3(lin, a->linF);
    ang = side == 0 ? m3Sub3(ang, a->angF) : m3Add3(ang, a->angF);
    ang = side == 0 ? m3Sub3(ang, a->angT) : m3Add3(ang, a->angT);
    *linIo = lin;
    *angIo = ang;
}

// Warm start of every constraint on dynamic body b in the order the
// colored pass reaches them: one slot per color 0..15 from the body's slot
// table, then the overflow color's slots [ovStart, ovEnd) in list order.
// The body's velocity stays in registers throughout, and the next slot's
// entry is fetched while the current constraint is applied.
M3_HD void m3WarmStartBody(const m3View* v, int32_t b, int32_t ovStart, int32_t ovEnd)
{
    const uint32_t mask = v->slotMask[b];
    const int32_t* row = v->slotTable + (int64_t)b * M3_GRAPH_COLORS;
    m3Vec3 lin = v->lv[b];
    m3Vec3 ang = v->av[b];
    uint32_t colors = mask & 0xFFFFu;
    int32_t next = colors != 0u ? row[__builtin_ctz(colors)] : 0;
    while (colors != 0u)
    {
        const int32_t e = next;
        colors &= colors - 1u;
        if (colors != 0u)
        {
            next = row[__builtin_ctz(colors)];
        }
        m3WarmStartSide(&v->cons[e >> 1], e & 1, &lin, &ang);
    }
    if ((mask & 0x10000u) != 0u)
    {
        for (int32_t k = ovStart; k < ovEnd; ++k)
        {
            const m3Constraint* c = &v->cons[k];
            if (c->dynA && c->bodyA == b)
            {
                m3WarmStartSide(c, 0, &lin, &ang);
            }
            if (c->dynB && c->bodyB == b)
            {
                m3WarmStartSide(c, 1, &lin, &ang);
            }
        }
    }
    v->lv[b] = lin;
    v->av[b] = ang;
}

// Cramer's rule; singular returns zero (src/solver.c:2909-2926).
M3_HD m3Vec3 m3Solve3(const m3Mat3* J, m3Vec3 b)
{
    m3Vec3 cxy = m3Cross3(J->cy, J->cz);
    m3real det = m3Dot3(J->cx, cxy);
    if (det == 0.0f)
    {
        return m3V3(0.0f, 0.0f, 0.0f);
    }
    m3real inv = 1.0f / det;
    m3Vec3 x;
    x.x = inv * m3Dot3(b, cxy);
    x.y = inv * m3Dot3(J->cx, m3Cross3(b, J->cz));
    x.z = inv * m3Dot3(J->cx, m3Cross3(J->cy, b));
    return x;
}

// ---------------------------------------------------------------------------
// Spherical joints (src/solver.c:577-811, 1044-1066, 1195-1337, 2657-2730)
// ---------------------------------------------------------------------------

M3_HD m3real m3TwistAngle(m3Quat q)
{
    m3real twist = q.w < 0.0f ? m3Atan2(-q.z, -q.w) : m3Atan2(q.z, q.w);
    return 2.0f * twist;
}

M3_HD m3real m3SwingAngle(m3Quat q)
{
    m3real x = m3Sqrt(q.z * q.z + q.w * q.w);
    m3real y = m3Sqrt(q.x * q.x + q.y * q.y);
    return 2.0f * m3Atan2(y, x);
}

// Joint j prepared at the step's start transforms (src/solver.c:699-811,
// the spherical branch).
M3_HD void m3PrepareJoint(const m3View* v, int32_t j, m3JointCon* c, m3real h)
{
    m3JointCon z;
    memset(&z, 0, sizeof(z));
    *c = z;
    int32_t a = v->jointBodyA[j];
    int32_t b = v->jointBodyB[j];
    c->bodyA = a;
    c->bodyB = b;
    m3Transform xfA = v->xf[a];
    m3Transform xfB = v->xf[b];
    c->rA = m3RotateVec3(xfA.q, m3Sub3(v->jointLocalA[j], v->lc[a]));
    c->rB = m3RotateVec3(xfB.q, m3Sub3(v->jointLocalB[j], v->lc[b]));
    m3Vec3 rlcA = m3RotateVec3(xfA.q, v->lc[a]);
    m3Vec3 rlcB = m3RotateVec3(xfB.q, v->lc[b]);
    c->deltaCenter = m3V3((m3real)(xfB.p.x + (double)rlcB.x - xfA.p.x - (double)rlcA.x),
                          (m3real)(xfB.p.y + (double)rlcB.y - xfA.p.y - (double)rlcA.y),
                          (m3real)(xfB.p.z + (double)rlcB.z - xfA.p.z - (double)rlcA.z));
    c->invMassA = v->type[a] == M3_DYNAMIC ? v->invMass[a] : 0.0f;
    c->invMassB = v->type[b] == M3_DYNAMIC ? v->invMass[b] : 0.0f;
    c->invIA = m3WorldInvInertia(v, a);
    c->invIB = m3WorldInvInertia(v, b);
    c->softness = m3MakeSoft(60.0f, 2.0f, h);
    c->impulse = v->jointImpulse[j];
    c->flags = v->jointFlags[j];
    c->springImpulseV = v->jointSpringImpulse[j];
    if ((c->flags & (1 | 4 | 8)) != 0)
    {
        c->frameQA = m3MulQuat(xfA.q, v->jointFrameQA[j]);
        c->frameQB = m3MulQuat(xfB.q, v->jointFrameQB[j]);
        m3Vec3 coneAxis = m3RotateVec3(c->frameQA, m3V3(0.0f, 0.0f, 1.0f));
        m3Vec3 twistAxis = m3RotateVec3(c->frameQB, m3V3(0.0f, 0.0f, 1.0f));
        m3Vec3 swing = m3Cross3(coneAxis, twistAxis);
        m3real swingLen2 = m3Dot3(swing, swing);
        if (swingLen2 > 1.0e-12f)
        {
            swing = m3MulSV3(1.0f / m3Sqrt(swingLen2), swing);
        }
        else
        {
            m3Vec3 t2;
            m3TangentBasis(coneAxis, &swing, &t2);
        }
        c->swingAxis = swing;
        m3Vec3 sumSwing = m3Add3(m3MulMV3(c->invIA, swing), m3MulMV3(c->invIB, swing));
        m3real kSwing = m3Dot3(swing, sumSwing);
        c->swingMass = kSwing > 0.0f ? 1.0f / kSwing : 0.0f;
        m3Quat conjA = m3Q(-c->frameQA.x, -c->frameQA.y, -c->frameQA.z, c->frameQA.w);
        m3Quat relQ = m3MulQuat(conjA, c->frameQB);
        m3real denom = relQ.z * relQ.z + relQ.w * relQ.w;
        m3real tanHalf =
            denom > 1.0e-12f ? m3Sqrt((relQ.x * relQ.x + relQ.y * relQ.y) / denom) : 0.0f;
        m3Vec3 perp = m3Cross3(swing, coneAxis);
        c->twistJacobian = m3Add3(coneAxis, m3MulSV3(tanHalf, perp));
        m3Vec3 sumTwist =
            m3Add3(m3MulMV3(c->invIA, c->twistJacobian), m3MulMV3(c->invIB, c->twistJacobian));
        m3real kTwist = m3Dot3(c->twistJacobian, sumTwist);
        c->twistMass = kTwist > 0.0f ? 1.0f / kTwist : 0.0f;
        c->coneAngle = v->jointLimits[j].z;
        c->lowerLimit = v->jointLimits[j].x;
        c->upperLimit = v->jointLimits[j].y;
        c->lowerImpulse = v->jointLimitImpulse[j].x;
        c->upperImpulse = v->jointLimitImpulse[j].y;
        c->swingImpulse = v->jointLimitImpulse[j].z;
    }
}

M3_HD void m3WarmStartJoint(const m3View* v, const m3JointCon* c)
{
    m3Vec3 rA = m3RotateVec3(v->deltaRot[c->bodyA], c->rA);
    m3Vec3 rB = m3RotateVec3(v->deltaRot[c->bodyB], c->rB);
    m3Vec3 angularImpulse = m3V3(0.0f, 0.0f, 0.0f);
    m3Vec3 linearExtra = m3V3(0.0f, 0.0f, 0.0f);
    if ((c->flags & (1 | 4 | 8)) != 0)
    {
        angularImpulse = m3MulSV3(-c->swingImpulse, c->swingAxis);
        angularImpulse =
            m3Add3(angularImpulse, m3MulSV3(c->lowerImpulse - c->upperImpulse, c->twistJacobian));
        if ((c->flags & 8) != 0)
        {
            angularImpulse = m3Add3(angularImpulse, c->springImpulseV);
        }
    }
    m3Vec3 totalLinear = m3Add3(c->impulse, linearExtra);
    v->lv[c->bodyA] = m3Sub3(v->lv[c->bodyA], m3MulSV3(c->invMassA, totalLinear));
    v->av[c->bodyA] = m3Sub3(v->av[c->bodyA],
                             m3MulMV3(c->invIA, m3Add3(m3Cross3(rA, totalLinear), angularImpulse)));
    v->lv[c->bodyB] = m3Add3(v->lv[c->bodyB], m3MulSV3(c->invMassB, totalLinear));
    v->av[c->bodyB] = m3Add3(v->av[c->bodyB],
                             m3MulMV3(c->invIB, m3Add3(m3Cross3(rB, totalLinear), angularImpulse)));
}

// One limit row: impulse delta along jac, accumulated and clamped at 0.
M3_HD m3real m3JointLimitDelta(const m3JointCon* c, m3real cc, m3real cdot, m3real mass,
                               m3real* acc, m3real invH, int useBias)
{
    m3real bias = 0.0f;
    m3real massScale = 1.0f;
    m3real impulseScale = 0.0f;
    if (cc > 0.0f)
    {
        bias = cc * invH;
    }
    else if (useBias)
    {
        bias = c->softness.biasRate * cc;
        massScale = c->softness.massScale;
        impulseScale = c->softness.impulseScale;
    }
    m3real old = *acc;
    m3real delta = -massScale * mass * (cdot + bias) - impulseScale * old;
    *acc = m3MaxF(old + delta, 0.0f);
    return *acc - old;
}

// Twist range, cone, then the point constraint (src/solver.c:1211-1337,
// 2657-2700); the drive spring (flag 8) has no setter in this build.
M3_HD void m3SolveJoint(const m3View* v, m3JointCon* c, m3real invH, int useBias)
{
    m3Vec3 rA = m3RotateVec3(v->deltaRot[c->bodyA], c->rA);
    m3Vec3 rB = m3RotateVec3(v->deltaRot[c->bodyB], c->rB);
    m3Vec3 vA = v->lv[c->bodyA];
    m3Vec3 wA = v->av[c->bodyA];
    m3Vec3 vB = v->lv[c->bodyB];
    m3Vec3 wB = v->av[c->bodyB];
    if ((c->flags & (1 | 4 | 8)) != 0)
    {
        m3Quat quatA = m3MulQuat(v->deltaRot[c->bodyA], c->frameQA);
        m3Quat quatB = m3MulQuat(v->deltaRot[c->bodyB], c->frameQB);
        m3Quat conjA = m3Q(-quatA.x, -quatA.y, -quatA.z, quatA.w);
        m3Quat relQ = m3MulQuat(conjA, quatB);
        if ((c->flags & 1) != 0)
        {
            m3real angle = m3TwistAngle(relQ);
            m3Vec3 jac = c->twistJacobian;
            m3real delta = m3JointLimitDelta(c, angle - c->lowerLimit, m3Dot3(m3Sub3(wB, wA), jac),
                                             c->twistMass, &c->lowerImpulse, invH, useBias);
            wA = m3Sub3(wA, m3MulSV3(delta, m3MulMV3(c->invIA, jac)));
            wB = m3Add3(wB, m3MulSV3(delta, m3MulMV3(c->invIB, jac)));
            delta = m3JointLimitDelta(c, c->upperLimit - angle, m3Dot3(m3Sub3(wA, wB), jac),
                                      c->twistMass, &c->upperImpulse, invH, useBias);
            wA = m3Add3(wA, m3MulSV3(delta, m3MulMV3(c->invIA, jac)));
            wB = m3Sub3(wB, m3MulSV3(delta, m3MulMV3(c->invIB, jac)));
        }
        if ((c->flags & 4) != 0)
        {
            m3real swingAngle = m3SwingAngle(relQ);
            m3Vec3 axis = c->swingAxis;
            m3real delta = m3JointLimitDelta(c, c->coneAngle - swingAngle,
                                             m3Dot3(m3Sub3(wA, wB), axis), c->swingMass,
                                             &c->swingImpulse, invH, useBias);
            wA = m3Add3(wA, m3MulSV3(delta, m3MulMV3(c->invIA, axis)));
            wB = m3Sub3(wB, m3MulSV3(delta, m3MulMV3(c->invIB, axis)));
        }
        v->av[c->bodyA] = wA;
        v->av[c->bodyB] = wB;
    }
    m3Vec3 cdot = m3Sub3(m3Add3(vB, m3Cross3(wB, rB)), m3Add3(vA, m3Cross3(wA, rA)));
    m3Vec3 bias = m3V3(0.0f, 0.0f, 0.0f);
    m3real massScale = 1.0f;
    m3real impulseScale = 0.0f;
    if (useBias)
    {
        m3Vec3 separation = m3Add3(
            m3Add3(m3Sub3(v->deltaPos[c->bodyB], v->deltaPos[c->bodyA]), m3Sub3(rB, rA)),
            c->deltaCenter);
        bias = m3MulSV3(c->softness.biasRate, separation);
        massScale = c->softness.massScale;
        impulseScale = c->softness.impulseScale;
    }
    m3Mat3 k;
    for (int32_t axisIndex = 0; axisIndex < 3; ++axisIndex)
    {
        m3Vec3 e = m3V3(axisIndex == 0 ? 1.0f : 0.0f, axisIndex == 1 ? 1.0f : 0.0f,
                        axisIndex == 2 ? 1.0f : 0.0f);
        m3Vec3 tA = m3Cross3(rA, m3MulMV3(c->invIA, m3Cross3(rA, e)));
        m3Vec3 tB = m3Cross3(rB, m3MulMV3(c->invIB, m3Cross3(rB, e)));
        m3Vec3 col = m3Sub3(m3MulSV3(c->invMassA + c->invMassB, e), m3Add3(tA, tB));
        if (axisIndex == 0)
        {
            k.cx = col;
        }
        else if (axisIndex == 1)
        {
            k.cy = col;
        }
        else
        {
            k.cz = col;
        }
    }
    m3Vec3 sol = m3Solve3(&k, m3Add3(cdot, bias));
    m3Vec3 impulse = m3Sub3(m3MulSV3(-massScale, sol), m3MulSV3(impulseScale, c->impulse));
    c->impulse = m3Add3(c->impulse, impulse);
    v->lv[c->bodyA] = m3Sub3(vA, m3MulSV3(c->invMassA, impulse));
    v->av[c->bodyA] = m3Sub3(wA, m3MulMV3(c->invIA, m3Cross3(rA, impulse)));
    v->lv[c->bodyB] = m3Add3(vB, m3MulSV3(c->invMassB, impulse));
    v->av[c->bodyB] = m3Add3(wB, m3MulMV3(c->invIB, m3Cross3(rB, impulse)));
}

// src/solver.c:2704-2730 for a spherical joint.
M3_HD void m3StoreJoint(const m3View* v, const m3JointCon* c, int32_t j)
{
    v->jointImpulse[j] = c->impulse;
    v->jointPerpImpulse[j] = m3V3(0.0f, 0.0f, 0.0f);
    v->jointLimitImpulse[j] = m3V3(c->lowerImpulse, c->upperImpulse, c->swingImpulse);
    v->jointAngularImpulse[j] = m3V3(0.0f, 0.0f, 0.0f);
    v->jointSpringImpulse[j] = c->springImpulseV;
}

// Every joint of component k, serially in ascending slot order (the
// reference's order restricted to joints that share non-static bodies).
M3_HD void m3JointComponentPass(const m3View* v, int32_t k, int mode, m3real invH)
{
    for (int32_t e = v->jointCompStart[k]; e < v->jointCompStart[k + 1]; ++e)
    {
        int32_t j = v->jointCompList[e];
        if (v->jointActiveFlag[j] == 0)
        {
            continue;
        }
        if (mode == 0)
        {
            m3WarmStartJoint(v, &v->jcons[j]);
        }
        else
        {
            m3SolveJoint(v, &v->jcons[j], invH, mode == 1);
        }
    }
}

// The velocities of a constraint's two bodies, held in registers while the
// constraint is solved.  The two bodies of a contact constraint are always
// distinct, so these locals go through exactly the values the body arrays
// would hold between the individual updates.
typedef struct m3PairVelocity
{
    m3Vec3 vA;
    m3Vec3 wA;
    m3Vec3 vB;
    m3Vec3 wB;
} m3PairVelocity;

// The fields of one contact constraint that a solve or relax pass reads,
// gathered into registers.  Loading them apart from the solve lets a
// caller issue the loads early (before the barrier that precedes the
// constraint's color); the fields a constraint's own solve writes belong to
// that solve alone, so a copy taken after the constraint's previous solve
// stays current.
typedef struct m3ContactIn
{
    int32_t a;
    int32_t b;
    int32_t pointCount;
    uint8_t dynA;
    uint8_t dynB;
    m3Vec3 normal;
    m3Vec3 prA[M3_MANIFOLD_MAX_POINTS];
    m3Vec3 prB[M3_MANIFOLD_MAX_POINTS];
    m3real pBase[M3_MANIFOLD_MAX_POINTS];
    m3real pMass[M3_MANIFOLD_MAX_POINTS];
    m3real pLever[M3_MANIFOLD_MAX_POINTS];
    m3real pImpulse[M3_MANIFOLD_MAX_POINTS];
    m3real pTotal[M3_MANIFOLD_MAX_POINTS];
    m3real invMassA;
    m3real invMassB;
    m3Mat3 invIA;
    m3Mat3 invIB;
    m3Softness soft;
    m3real friction;
    // Relax pass only: twist, rolling and friction rows.
    m3Vec3 t1;
    m3Vec3 t2;
    m3Vec3 originA;
    m3Vec3 originB;
    m3real k11;
    m3real k12;
    m3real k22;
    m3real oldF1;
    m3real oldF2;
    m3real tv1;
    m3real tv2;
    m3real twistMass;
    m3real oldTwist;
    m3real rolling;
} m3ContactIn;

// All loads are issued together; with a constant useBias the relax-only
// fields of a solve pass are not loaded at all.
M3_HD void m3LoadContact(const m3Constraint* c, int useBias, m3ContactIn* in)
{
    in->a = c->bodyA;
    in->b = c->bodyB;
    in->pointCount = c->pointCount;
    in->normal = c->normal;
    M3_UNROLL
    for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        const m3CPoint* cp = &c->points[k];
        in->prA[k] = cp->rA;
        in->prB[k] = cp->rB;
        in->pBase[k] = cp->baseSeparation;
        in->pMass[k] = cp->normalMass;
        in->pLever[k] = cp->leverArm;
        in->pImpulse[k] = cp->normalImpulse;
        in->pTotal[k] = cp->totalNormalImpulse;
    }
    in->dynA = c->dynA;
    in->dynB = c->dynB;
    in->invMassA = c->invMassA;
    in->invMassB = c->invMassB;
    in->invIA = c->invIA;
    in->invIB = c->invIB;
    in->soft = c->softness;
    in->friction = c->friction;
    if (!useBias)
    {
        in->t1 = c->t1;
        in->t2 = c->t2;
        in->originA = c->originA;
        in->originB = c->originB;
        in->k11 = c->frictionK11;
        in->k12 = c->frictionK12;
        in->k22 = c->frictionK22;
        in->oldF1 = c->frictionImpulse1;
        in->oldF2 = c->frictionImpulse2;
        in->tv1 = c->tangentVelocity1;
        in->tv2 = c->tangentVelocity2;
        in->twistMass = c->twistMass;
        in->oldTwist = c->twistImpulse;
        in->rolling = c->rollingResistance;
    }
}

// The per-point terms of a contact solve that depend only on the bodies'
// position deltas (src/solver.c:446-470: current separation, then the
// speculative bias or the soft bias, mass scale and impulse scale).  The
// deltas do not change inside a solve or relax pass, so a caller may
// evaluate these before the velocities it waits for are ready.
typedef struct m3ContactBias
{
    m3real bias[M3_MANIFOLD_MAX_POINTS];
    m3real massScale[M3_MANIFOLD_MAX_POINTS];
    m3real impulseScale[M3_MANIFOLD_MAX_POINTS];
} m3ContactBias;

M3_HD void m3ContactBiasTerms(const m3ContactIn* in, m3Vec3 dpA, m3Quat dqA, m3Vec3 dpB, m3Quat dqB,
                              m3real invH, m3real pushMax, int useBias, int32_t slots,
                              m3ContactBias* out)
{
    M3_UNROLL
    for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        if (k >= slots)
        {
            break;
        }
        m3Vec3 rsA = m3RotateVec3(dqA, in->prA[k]);
        m3Vec3 rsB = m3RotateVec3(dqB, in->prB[k]);
        m3Vec3 ds = m3Add3(m3Sub3(dpB, dpA), m3Sub3(rsB, rsA));
        m3real sep = in->pBase[k] + m3Dot3(ds, in->normal);
        m3real bias = 0.0f;
        m3real massScale = 1.0f;
        m3real impulseScale = 0.0f;
        if (sep > 0.0f)
        {
            bias = sep * invH;
        }
        else if (useBias)
        {
            bias = m3MaxF(in->soft.biasRate * sep, -pushMax);
            massScale = in->soft.massScale;
            impulseScale = in->soft.impulseScale;
        }
        out->bias[k] = bias;
        out->massScale[k] = massScale;
        out->impulseScale[k] = impulseScale;
    }
}

// The velocity part of a contact solve (src/solver.c:471-581): the normal
// points in order, then (relax pass) twist, rolling resistance and
// friction, on the pair's velocities held in registers.  Writes the
// constraint's accumulators; returns whether the rolling row ran (it
// writes the angular velocity of both sides).
typedef struct m3ContactAccum
{
    m3real normal[M3_MANIFOLD_MAX_POINTS];
    m3real f1;
    m3real f2;
    m3real twist;
} m3ContactAccum;

M3_HD int m3ContactImpulses(m3Constraint* c, const m3ContactIn* in, const m3ContactBias* cb,
                            int useBias, int32_t slots, m3PairVelocity* sv, m3ContactAccum* acc = nullptr)
{
    const int32_t pointCount = in->pointCount;
    const m3Vec3 normal = in->normal;
    const uint8_t dynA = in->dynA;
    const uint8_t dynB = in->dynB;
    const m3real invMassA = in->invMassA;
    const m3real invMassB = in->invMassB;
    const m3Mat3 invIA = in->invIA;
    const m3Mat3 invIB = in->invIB;
    const m3real friction = in->friction;
    m3PairVelocity s = *sv;
    m3real passNormal = 0.0f;
    m3real twistLimit = 0.0f;
    M3_UNROLL
    for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        if (k >= slots)
        {
            break;
        }
        const int live = k < pointCount;
        const m3Vec3 rA = in->prA[k];
        const m3Vec3 rB = in->prB[k];
        m3Vec3 velB = m3Add3(s.vB, m3Cross3(s.wB, rB));
        m3Vec3 velA = m3Add3(s.vA, m3Cross3(s.wA, rA));
        m3real vn = m3Dot3(m3Sub3(velB, velA), normal);
        const m3real oldImpulse = in->pImpulse[k];
        m3real impulse = -in->pMass[k] * cb->massScale[k] * (vn + cb->bias[k]) -
                         cb->impulseScale[k] * oldImpulse;
        m3real newImpulse = m3MaxF(oldImpulse + impulse, 0.0f);
        impulse = newImpulse - oldImpulse;
        if (live)
        {
            m3CPoint* cp = &c->points[k];
            cp->normalImpulse = newImpulse;
            cp->totalNormalImpulse = in->pTotal[k] + newImpulse;
            passNormal += newImpulse;
            twistLimit += in->pLever[k] * newImpulse;
        }
        if (acc != nullptr)
        {
            acc->normal[k] = newImpulse;
        }
        const m3Vec3 P = m3MulSV3(impulse, normal);
        const m3Vec3 nvA = m3Sub3(s.vA, m3MulSV3(invMassA, P));
        const m3Vec3 nwA = m3Sub3(s.wA, m3MulMV3(invIA, m3Cross3(rA, P)));
        const m3Vec3 nvB = m3Add3(s.vB, m3MulSV3(invMassB, P));
        const m3Vec3 nwB = m3Add3(s.wB, m3MulMV3(invIB, m3Cross3(rB, P)));
        if (live && dynA)
        {
            s.vA = nvA;
            s.wA = nwA;
        }
        if (live && dynB)
        {
            s.vB = nvB;
            s.wB = nwB;
        }
    }
    int rolled = 0;
    if (!useBias)
    {
        const m3Vec3 t1 = in->t1;
        const m3Vec3 t2 = in->t2;
        const m3Vec3 originA = in->originA;
        const m3Vec3 originB = in->originB;
        const m3real k11 = in->k11;
        const m3real k12 = in->k12;
        const m3real k22 = in->k22;
        const m3real oldF1 = in->oldF1;
        const m3real oldF2 = in->oldF2;
        const m3real tv1 = in->tv1;
        const m3real tv2 = in->tv2;
        const m3real twistMass = in->twistMass;
        const m3real oldTwist = in->oldTwist;
        const m3real rolling = in->rolling;
        {
            m3Vec3 wRel = m3Sub3(s.wB, s.wA);
            m3real twistSpeed = m3Dot3(normal, wRel);
            m3real maxTwist = friction * twistLimit;
            m3real delta = -twistMass * twistSpeed;
            m3real newTwist = oldTwist + delta;
            if (newTwist < -maxTwist)
            {
                newTwist = -maxTwist;
            }
            else if (newTwist > maxTwist)
            {
                newTwist = maxTwist;
            }
            delta = newTwist - oldTwist;
            c->twistImpulse = newTwist;
            if (acc != nullptr)
            {
                acc->twist = newTwist;
            }
            m3Vec3 tw = m3MulSV3(delta, normal);
            if (dynA)
            {
                s.wA = m3Sub3(s.wA, m3MulMV3(invIA, tw));
            }
            if (dynB)
            {
                s.wB = m3Add3(s.wB, m3MulMV3(invIB, tw));
            }
        }
        if (rolling > 0.0f)
        {
            m3Vec3 wRel = m3Sub3(s.wB, s.wA);
            m3Vec3 delta = m3MulSV3(-1.0f, m3Solve3(&c->rollingK, wRel));
            m3Vec3 accum = m3Add3(c->rollingImpulse, delta);
            m3real maxRoll = rolling * passNormal;
            m3real mag2 = m3Dot3(accum, accum);
            if (mag2 > maxRoll * maxRoll && mag2 > 0.0f)
            {
                accum = m3MulSV3(maxRoll / m3Sqrt(mag2), accum);
            }
            delta = m3Sub3(accum, c->rollingImpulse);
            c->rollingImpulse = accum;
            // The reference writes both sides unconditionally here; a body
            // without inverse inertia receives a zero change.
            s.wA = m3Sub3(s.wA, m3MulMV3(invIA, delta));
            s.wB = m3Add3(s.wB, m3MulMV3(invIB, delta));
            rolled = 1;
        }
        {
            m3Vec3 velB = m3Add3(s.vB, m3Cross3(s.wB, originB));
            m3Vec3 velA = m3Add3(s.vA, m3Cross3(s.wA, originA));
            m3Vec3 vrel = m3Sub3(velB, velA);
            m3real vt1 = m3Dot3(vrel, t1) - tv1;
            m3real vt2 = m3Dot3(vrel, t2) - tv2;
            m3real f1 = oldF1 - (k11 * vt1 + k12 * vt2);
            m3real f2 = oldF2 - (k12 * vt1 + k22 * vt2);
            m3real maxFriction = friction * passNormal;
            m3real mag2 = f1 * f1 + f2 * f2;
            if (mag2 > maxFriction * maxFriction)
            {
                m3real mag = m3Sqrt(mag2);
                m3real scale = mag > 0.0f ? maxFriction / mag : 0.0f;
                f1 *= scale;
                f2 *= scale;
            }
            m3real d1 = f1 - oldF1;
            m3real d2 = f2 - oldF2;
            c->frictionImpulse1 = f1;
            c->frictionImpulse2 = f2;
            if (acc != nullptr)
            {
                acc->f1 = f1;
                acc->f2 = f2;
            }
            const m3Vec3 P = m3Add3(m3MulSV3(d1, t1), m3MulSV3(d2, t2));
            if (dynA)
            {
                s.vA = m3Sub3(s.vA, m3MulSV3(invMassA, P));
                s.wA = m3Sub3(s.wA, m3MulMV3(invIA, m3Cross3(originA, P)));
            }
            if (dynB)
            {
                s.vB = m3Add3(s.vB, m3MulSV3(invMassB, P));
                s.wB = m3Add3(s.wB, m3MulMV3(invIB, m3Cross3(originB, P)));
            }
        }
    }
    *sv = s;
    return rolled;
}

// The number of manifold slots a solve evaluates: every slot up to the
// largest point count among the lanes running this solve together.  The
// results of slots past a lane's own pointCount are discarded, so no load
// waits behind a branch on the point count and the bound is uniform across
// the wavefront; slots past every lane's count would be discarded anyway,
// so skipping them changes no result.
M3_HD int32_t m3ContactSlots(int32_t pointCount)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(pointCount > 2) != 0 ? (__ballot(pointCount > 3) != 0 ? 4 : 3)
                                         : (__ballot(pointCount > 1) != 0 ? 2 : 1);
#else
    (void)pointCount;
    return M3_MANIFOLD_MAX_POINTS;
#endif
}

// src/solver.c:430-581 on a gathered constraint (m3LoadContact of c with
// the same useBias).  A constraint solve is a chain of short dependent
// steps, so memory round trips dominate it: the body velocities are read
// once, updated in registers in the reference's operation order, and
// written back once: linear and angular velocity of each dynamic side, and
// the angular velocity of a non-dynamic side that the rolling-resistance
// row writes.
M3_HD void m3SolveContactIn(const m3View* v, m3Constraint* c, const m3ContactIn* in, m3real invH,
                            m3real pushMax, int useBias)
{
    const int32_t a = in->a;
    const int32_t b = in->b;
    m3PairVelocity s;
    s.vA = v->lv[a];
    s.wA = v->av[a];
    s.vB = v->lv[b];
    s.wB = v->av[b];
    const int32_t slots = m3ContactSlots(in->pointCount);
    m3ContactBias cb;
    m3ContactBiasTerms(in, v->deltaPos[a], v->deltaRot[a], v->deltaPos[b], v->deltaRot[b], invH,
                       pushMax, useBias, slots, &cb);
    const int rolled = m3ContactImpulses(c, in, &cb, useBias, slots, &s);
    if (in->dynA)
    {
        v->lv[a] = s.vA;
        v->av[a] = s.wA;
    }
    else if (rolled)
    {
        v->av[a] = s.wA;
    }
    if (in->dynB)
    {
        v->lv[b] = s.vB;
        v->av[b] = s.wB;
    }
    else if (rolled)
    {
        v->av[b] = s.wB;
    }
}

// src/solver.c:430-581 for one constraint.
M3_HD void m3SolveContact(const m3View* v, m3Constraint* c, m3real invH, m3real pushMax,
                          int useBias)
{
    m3ContactIn in;
    m3LoadContact(c, useBias, &in);
    m3SolveContactIn(v, c, &in, invH, pushMax, useBias);
}

// src/solver.c:2879-2907 for one constraint.
M3_HD void m3RestitutionOne(const m3View* v, m3Constraint* c, m3real threshold)
{
    if (c->restitution == 0.0f)
    {
        return;
    }
    for (int32_t k = 0; k < c->pointCount; ++k)
    {
        m3CPoint* cp = &c->points[k];
        if (cp->relativeVelocity > -threshold || cp->totalNormalImpulse == 0.0f)
        {
            continue;
        }
        m3real vn = m3Dot3(
            m3Sub3(m3VelocityAt(v, c->bodyB, cp->rB), m3VelocityAt(v, c->bodyA, cp->rA)), c->normal);
        m3real impulse = -cp->normalMass * (vn + c->restitution * cp->relativeVelocity);
        m3real newImpulse = m3MaxF(cp->normalImpulse + impulse, 0.0f);
        impulse = newImpulse - cp->normalImpulse;
        cp->normalImpulse = newImpulse;
        m3ApplyImpulse(v, c, m3MulSV3(impulse, c->normal), cp->rA, cp->rB);
    }
}

// src/solver.c:2979-3048.  The hit event's point selection (:2991-3026)
// is recorded per pair when a hit record array is present: among the points
// with a positive accumulated impulse, the first with the lowest relative
// normal velocity.  When that velocity is below minus the threshold it is
// exactly the point the reference's loop selects (the reference's candidate
// set is this set cut at the threshold, and its minimum is this minimum);
// otherwise the reference selects none.
M3_HD void m3StoreOne(const m3View* v, const m3Constraint* c)
{
    m3Manifold* m = &v->manifolds[c->manifoldIndex];
    for (int32_t k = 0; k < c->pointCount; ++k)
    {
        m->points[k].normalImpulse = c->points[k].normalImpulse;
    }
    if (v->hitRecs != nullptr)
    {
        const uint64_t key = v->keys[c->manifoldIndex];
        const int32_t shapeA = (int32_t)(key >> 32);
        const int32_t shapeB = (int32_t)(key & 0xFFFFFFFFu);
        if (v->hitEvents[shapeA] != 0 || v->hitEvents[shapeB] != 0)
        {
            int32_t best = -1;
            for (int32_t k = 0; k < c->pointCount; ++k)
            {
                const m3real rv = c->points[k].relativeVelocity;
                if (rv == rv && c->points[k].totalNormalImpulse > 0.0f &&
                    (best < 0 || rv < c->points[best].relativeVelocity))
                {
                    best = k;
                }
            }
            if (best >= 0)
            {
                m3HitRec* r = &v->hitRecs[c->manifoldIndex];
                const m3Transform xa = v->xf[c->bodyA];
                const m3Vec3 lcA = m3RotateVec3(xa.q, v->lc[c->bodyA]);
                const m3Vec3 rA = c->points[best].rA;
                r->point[0] = xa.p.x + (double)lcA.x + (double)rA.x;
                r->point[1] = xa.p.y + (double)lcA.y + (double)rA.y;
                r->point[2] = xa.p.z + (double)lcA.z + (double)rA.z;
                r->normal = c->normal;
                r->relativeVelocity = c->points[best].relativeVelocity;
                r->flag = 1;
            }
        }
    }
    m->frictionImpulse =
        m3Add3(m3MulSV3(c->frictionImpulse1, c->t1), m3MulSV3(c->frictionImpulse2, c->t2));
    m->twistImpulse = c->twistImpulse;
    m->rollingImpulse = c->rollingImpulse;
}

// ---------------------------------------------------------------------------
// Integration (src/solver.c:2927-2977, 4087-4199)
// ---------------------------------------------------------------------------

// One implicit Newton step of the gyroscopic term in body coordinates;
// exactly isotropic tensors skip it.
M3_HD m3Vec3 m3GyroscopicOmega(const m3View* v, int32_t body, m3Vec3 w, m3real h)
{
    const m3Mat3 inertia = v->inertia[body];
    const m3real i00 = inertia.cx.x;
    const m3real i01 = inertia.cy.x;
    const m3real i02 = inertia.cz.x;
    const m3real i11 = inertia.cy.y;
    const m3real i12 = inertia.cz.y;
    const m3real i22 = inertia.cz.z;
    if (i01 == 0.0f && i02 == 0.0f && i12 == 0.0f && i00 == i11 && i11 == i22)
    {
        return w;
    }
    m3Quat q = v->xf[body].q;
    m3Vec3 omega2 = m3InvRotateVec3(q, w);
    const m3real w1 = omega2.x;
    const m3real w2 = omega2.y;
    const m3real w3 = omega2.z;
    const m3real Iw1 = i00 * w1 + i01 * w2 + i02 * w3;
    const m3real Iw2 = i01 * w1 + i11 * w2 + i12 * w3;
    const m3real Iw3 = i02 * w1 + i12 * w2 + i22 * w3;
    m3Vec3 b = m3V3(h * (w2 * Iw3 - w3 * Iw2), h * (w3 * Iw1 - w1 * Iw3), h * (w1 * Iw2 - w2 * Iw1));
    m3Mat3 J;
    J.cx = m3V3(i00 + h * (w2 * i02 - w3 * i01), i01 + h * (w3 * i00 - w1 * i02 - Iw3),
                i02 + h * (w1 * i01 - w2 * i00 + Iw2));
    J.cy = m3V3(i01 + h * (w2 * i12 - w3 * i11 + Iw3), i11 + h * (w3 * i01 - w1 * i12),
                i12 + h * (w1 * i11 - w2 * i01 - Iw1));
    J.cz = m3V3(i02 + h * (w2 * i22 - w3 * i12 - Iw2), i12 + h * (w3 * i02 - w1 * i22 + Iw1),
                i22 + h * (w1 * i12 - w2 * i02));
    omega2 = m3Sub3(omega2, m3Solve3(&J, b));
    return m3RotateVec3(q, omega2);
}

M3_HD void m3IntegrateVelocity(const m3View* v, int32_t i, m3real h, m3Vec3 gravity,
                               m3real maxLinear, m3real maxAngular)
{
    if (v->type[i] != M3_DYNAMIC)
    {
        return;
    }
    m3Vec3 lin = v->lv[i];
    m3Vec3 ang = v->av[i];
    lin = m3Add3(lin, m3MulSV3(h * v->gravityScale[i], gravity));
    lin = m3Add3(lin, m3MulSV3(h * v->invMass[i], v->force[i]));
    m3Vec3 tq = v->torque[i];
    if (tq.x != 0.0f || tq.y != 0.0f || tq.z != 0.0f)
    {
        ang = m3Add3(ang, m3MulSV3(h, m3MulMV3(m3WorldInvInertia(v, i), tq)));
    }
    lin = m3MulSV3(1.0f / (1.0f + h * v->linDamp[i]), lin);
    ang = m3MulSV3(1.0f / (1.0f + h * v->angDamp[i]), ang);
    ang = m3GyroscopicOmega(v, i, ang, h);
    m3real v2 = m3Dot3(lin, lin);
    if (v2 > maxLinear * maxLinear)
    {
        lin = m3MulSV3(maxLinear / m3Sqrt(v2), lin);
    }
    if ((v->locks[i] & 0x40u) == 0)
    {
        m3real w2 = m3Dot3(ang, ang);
        if (w2 > maxAngular * maxAngular)
        {
            ang = m3MulSV3(maxAngular / m3Sqrt(w2), ang);
        }
    }
    v->lv[i] = lin;
    v->av[i] = ang;
}

M3_HD void m3IntegratePosition(const m3View* v, int32_t i, m3real h)
{
    m3Vec3 lin = v->lv[i];
    m3Vec3 ang = v->av[i];
    uint8_t locks = v->locks[i];
    if (locks != 0)
    {
        if (locks & 1u)
            lin.x = 0.0f;
        if (locks & 2u)
            lin.y = 0.0f;
        if (locks & 4u)
            lin.z = 0.0f;
        if (locks & 8u)
            ang.x = 0.0f;
        if (locks & 16u)
            ang.y = 0.0f;
        if (locks & 32u)
            ang.z = 0.0f;
        v->lv[i] = lin;
        v->av[i] = ang;
    }
    m3Vec3 lc = v->lc[i];
    m3Transform xf = v->xf[i];
    m3Vec3 rlcOld = m3RotateVec3(xf.q, lc);
    double cx = xf.p.x + (double)rlcOld.x + (double)(h * lin.x);
    double cy = xf.p.y + (double)rlcOld.y + (double)(h * lin.y);
    double cz = xf.p.z + (double)rlcOld.z + (double)(h * lin.z);
    m3Vec3 dw = m3MulSV3(h, ang);
    xf.q = m3IntegrateRotation(xf.q, dw);
    m3Vec3 rlcNew = m3RotateVec3(xf.q, lc);
    xf.p.x = cx - (double)rlcNew.x;
    xf.p.y = cy - (double)rlcNew.y;
    xf.p.z = cz - (double)rlcNew.z;
    v->xf[i] = xf;
    v->deltaPos[i] = m3Add3(v->deltaPos[i], m3MulSV3(h, lin));
    v->deltaRot[i] = m3IntegrateRotation(v->deltaRot[i], dw);
}

// Kinematic servo: velocities that land the body on its target this step
// (src/solver.c:3955-3981).
M3_HD void m3KinematicServo(const m3View* v, int32_t i, m3real dt)
{
    m3real servoInvDt = 1.0f / dt;
    const m3Transform now = v->xf[i];
    const m3Transform want = v->target[i];
    v->lv[i] = m3V3((m3real)(want.p.x - now.p.x) * servoInvDt, (m3real)(want.p.y - now.p.y) * servoInvDt,
                    (m3real)(want.p.z - now.p.z) * servoInvDt);
    m3Quat dq = m3MulQuat(want.q, m3Q(-now.q.x, -now.q.y, -now.q.z, now.q.w));
    if (dq.w < 0.0f)
    {
        dq = m3Q(-dq.x, -dq.y, -dq.z, -dq.w);
    }
    v->av[i] = m3MulSV3(2.0f * servoInvDt, m3V3(dq.x, dq.y, dq.z));
    v->hasTarget[i] = 0;
}

// ---------------------------------------------------------------------------
// Continuous collision against planes (src/solver.c:3065-3079, 3280-3430;
// sweep transform src/distance.c:712-732)
// ---------------------------------------------------------------------------

typedef struct m3Sweep
{
    m3Vec3 localCenter;
    m3Vec3 c1;
    m3Vec3 c2;
    m3Quat q1;
    m3Quat q2;
} m3Sweep;

M3_HD m3Transform m3SweepTransform(const m3Sweep* sweep, m3real time)
{
    m3Transform transform;
    m3Quat q2 = sweep->q2;
    if (sweep->q1.x * q2.x + sweep->q1.y * q2.y + sweep->q1.z * q2.z + sweep->q1.w * q2.w < 0.0f)
    {
        q2 = m3Q(-q2.x, -q2.y, -q2.z, -q2.w);
    }
    m3Quat q = m3Q(sweep->q1.x + time * (q2.x - sweep->q1.x), sweep->q1.y + time * (q2.y - sweep->q1.y),
                   sweep->q1.z + time * (q2.z - sweep->q1.z), sweep->q1.w + time * (q2.w - sweep->q1.w));
    transform.q = m3NormalizeQuat(q);
    m3Vec3 c = m3Add3(sweep->c1, m3MulSV3(time, m3Sub3(sweep->c2, sweep->c1)));
    m3Vec3 r = m3RotateVec3(transform.q, sweep->localCenter);
    transform.p.x = (double)(c.x - r.x);
    transform.p.y = (double)(c.y - r.y);
    transform.p.z = (double)(c.z - r.z);
    return transform;
}

M3_HD m3Sweep m3RelativeSweep(const m3View* v, int32_t body, m3Pos3 base)
{
    m3Sweep sweep;
    sweep.localCenter = v->lc[body];
    sweep.c1 = m3V3((m3real)(v->com0[body].x - base.x), (m3real)(v->com0[body].y - base.y),
                    (m3real)(v->com0[body].z - base.z));
    m3Vec3 rlc = m3RotateVec3(v->xf[body].q, v->lc[body]);
    sweep.c2 = m3V3((m3real)(v->xf[body].p.x + (double)rlc.x - base.x),
                    (m3real)(v->xf[body].p.y + (double)rlc.y - base.y),
                    (m3real)(v->xf[body].p.z + (double)rlc.z - base.z));
    sweep.q1 = v->rot0[body];
    sweep.q2 = v->xf[body].q;
    return sweep;
}

// ---------------------------------------------------------------------------
// Time of impact between two swept proxies: conservative advancement with
// separation functions, the reference's kernel (src/distance.c:734-1215).
// Both sweeps are re-centred so A starts at the origin; the state reached
// (hit, separated, overlapped, failed) and the fraction follow the same
// branches, so the fraction's bits equal the reference's.
// ---------------------------------------------------------------------------

enum { M3_TOI_UNKNOWN = 0, M3_TOI_FAILED = 1, M3_TOI_OVERLAPPED = 2, M3_TOI_HIT = 3, M3_TOI_SEPARATED = 4 };
enum { M3_SEP_VERTICES = 0, M3_SEP_EDGES = 1, M3_SEP_FACE_A = 2, M3_SEP_FACE_B = 3 };

typedef struct m3SepFn
{
    const m3Proxy* proxyA;
    const m3Proxy* proxyB;
    const m3Sweep* sweepA;
    const m3Sweep* sweepB;
    m3Vec3 witness1;
    m3Vec3 witness2;
    int32_t type;
} m3SepFn;

M3_HD m3Vec3 m3ToiPoint(const m3Transform* xf, m3Vec3 p) // src/distance.c:736-740
{
    m3Vec3 r = m3RotateVec3(xf->q, p);
    return m3V3((m3real)xf->p.x + r.x, (m3real)xf->p.y + r.y, (m3real)xf->p.z + r.z);
}

M3_HD int32_t m3ToiUnique(int32_t count, const int32_t idx[3]) // src/distance.c:742-761
{
    if (count == 1)
    {
        return 1;
    }
    if (count == 2)
    {
        return idx[0] != idx[1] ? 2 : 1;
    }
    if (idx[0] != idx[1] && idx[0] != idx[2] && idx[1] != idx[2])
    {
        return 3;
    }
    if (idx[0] == idx[1] && idx[0] == idx[2])
    {
        return 1;
    }
    return 2;
}

M3_HD int32_t m3ToiSupport(const m3Proxy* proxy, m3Vec3 d) // src/distance.c:763-776
{
    int32_t best = 0;
    m3real bestDot = m3Dot3(proxy->points[0], d);
    for (int32_t i = 1; i < proxy->count; ++i)
    {
        const m3real dot = m3Dot3(proxy->points[i], d);
        if (dot > bestDot)
        {
            bestDot = dot;
            best = i;
        }
    }
    return best;
}

// Does the edge cross product flip across the sweep (src/distance.c:780-789)?
M3_HD bool m3ToiFastEdges(const m3Sweep* sweepA, const m3Sweep* sweepB, m3Vec3 edgeLocalA,
                          m3Vec3 edgeLocalB, m3Vec3 axis0)
{
    m3Transform xfA = m3SweepTransform(sweepA, 1.0f);
    m3Transform xfB = m3SweepTransform(sweepB, 1.0f);
    m3Vec3 axis = m3Cross3(m3RotateVec3(xfA.q, edgeLocalA), m3RotateVec3(xfB.q, edgeLocalB));
    return m3Dot3(axis, axis0) < 0.0f;
}

// The separation function of the cached simplex at t1 (src/distance.c:809-929).
M3_HD m3SepFn m3ToiMakeSep(const m3SimplexCache* cache, const m3Proxy* proxyA, const m3Sweep* sweepA,
                           const m3Proxy* proxyB, const m3Sweep* sweepB, m3Vec3 worldNormal, m3real t1)
{
    m3SepFn f;
    f.proxyA = proxyA;
    f.proxyB = proxyB;
    f.sweepA = sweepA;
    f.sweepB = sweepB;
    f.witness1 = worldNormal;
    f.witness2 = m3V3(0.0f, 0.0f, 0.0f);
    f.type = M3_SEP_VERTICES;
    const int32_t ia3[3] = {cache->indexA[0], cache->indexA[1], cache->indexA[2]};
    const int32_t ib3[3] = {cache->indexB[0], cache->indexB[1], cache->indexB[2]};
    const int32_t count = cache->count < 3 ? (int32_t)cache->count : 3;
    const int32_t uniqueA = m3ToiUnique(count, ia3);
    const int32_t uniqueB = m3ToiUnique(count, ib3);
    m3Transform xfA1 = m3SweepTransform(sweepA, t1);
    m3Transform xfB1 = m3SweepTransform(sweepB, t1);
    const m3Vec3 deltaP = m3V3((m3real)(xfB1.p.x - xfA1.p.x), (m3real)(xfB1.p.y - xfA1.p.y),
                               (m3real)(xfB1.p.z - xfA1.p.z));
    if (count == 2 || (count == 3 && uniqueA <= 2 && uniqueB <= 2))
    {
        int32_t ea[2];
        int32_t eb[2];
        if (count == 2 && uniqueA == 2 && uniqueB == 2)
        {
            ea[0] = ia3[0];
            ea[1] = ia3[1];
            eb[0] = ib3[0];
            eb[1] = ib3[1];
        }
        else if (count == 3 && uniqueA == 2 && uniqueB == 2)
        {
            ea[0] = ia3[0];
            ea[1] = ia3[0] != ia3[1] ? ia3[1] : ia3[2];
            eb[0] = ib3[0];
            eb[1] = ib3[0] != ib3[1] ? ib3[1] : ib3[2];
        }
        else
        {
            return f; // vertex against edge: the world axis
        }
        const m3Vec3 vA1 = proxyA->points[ea[0]];
        m3Vec3 edgeLocalA = m3Normalize3(m3Sub3(proxyA->points[ea[1]], vA1));
        const m3Vec3 edgeA = m3RotateVec3(xfA1.q, edgeLocalA);
        const m3Vec3 vB1 = proxyB->points[eb[0]];
        m3Vec3 edgeLocalB = m3Normalize3(m3Sub3(proxyB->points[eb[1]], vB1));
        const m3Vec3 edgeB = m3RotateVec3(xfB1.q, edgeLocalB);
        m3Vec3 axis = m3Cross3(edgeA, edgeB);
        const m3real len2 = m3Dot3(axis, axis);
        const m3real tolerance2 = count == 2 ? 0.05f * 0.05f : 0.005f * 0.005f;
        if (len2 < tolerance2)
        {
            return f; // near parallel: the world axis
        }
        const m3Vec3 delta =
            m3Add3(m3Sub3(m3RotateVec3(xfB1.q, vB1), m3RotateVec3(xfA1.q, vA1)), deltaP);
        if (m3Dot3(delta, axis) < 0.0f)
        {
            axis = m3Neg3(axis);
            edgeLocalB = m3Neg3(edgeLocalB);
        }
        if (m3ToiFastEdges(sweepA, sweepB, edgeLocalA, edgeLocalB, axis))
        {
            f.witness1 = m3Normalize3(axis); // unsafe: a fixed world axis
            return f;
        }
        f.type = M3_SEP_EDGES;
        f.witness1 = edgeLocalA;
        f.witness2 = edgeLocalB;
        return f;
    }
    if (count == 3 && uniqueA == 3)
    {
        const m3Vec3 vA1 = proxyA->points[ia3[0]];
        const m3Vec3 vA2 = proxyA->points[ia3[1]];
        const m3Vec3 vA3 = proxyA->points[ia3[2]];
        m3Vec3 axisLocal = m3Normalize3(m3Cross3(m3Sub3(vA2, vA1), m3Sub3(vA3, vA1)));
        const m3Vec3 axisA = m3RotateVec3(xfA1.q, axisLocal);
        const m3Vec3 pointLocalA = m3MulSV3(1.0f / 3.0f, m3Add3(m3Add3(vA1, vA2), vA3));
        const m3Vec3 pointLocalB = proxyB->points[ib3[0]];
        const m3Vec3 delta = m3Add3(
            m3Sub3(m3RotateVec3(xfB1.q, pointLocalB), m3RotateVec3(xfA1.q, pointLocalA)), deltaP);
        if (m3Dot3(delta, axisA) < 0.0f)
        {
            axisLocal = m3Neg3(axisLocal);
        }
        f.type = M3_SEP_FACE_A;
        f.witness1 = axisLocal;
        f.witness2 = pointLocalA;
        return f;
    }
    if (count == 3 && uniqueB == 3)
    {
        const m3Vec3 vB1 = proxyB->points[ib3[0]];
        const m3Vec3 vB2 = proxyB->points[ib3[1]];
        const m3Vec3 vB3 = proxyB->points[ib3[2]];
        m3Vec3 axisLocal = m3Normalize3(m3Cross3(m3Sub3(vB2, vB1), m3Sub3(vB3, vB1)));
        const m3Vec3 axisB = m3RotateVec3(xfB1.q, axisLocal);
        const m3Vec3 pointLocalA = proxyA->points[ia3[0]];
        const m3Vec3 pointLocalB = m3MulSV3(1.0f / 3.0f, m3Add3(m3Add3(vB1, vB2), vB3));
        const m3Vec3 delta = m3Sub3(
            m3Sub3(m3RotateVec3(xfA1.q, pointLocalA), m3RotateVec3(xfB1.q, pointLocalB)), deltaP);
        if (m3Dot3(delta, axisB) < 0.0f)
        {
            axisLocal = m3Neg3(axisLocal);
        }
        f.type = M3_SEP_FACE_B;
        f.witness1 = axisLocal;
        f.witness2 = pointLocalB;
        return f;
    }
    return f; // one vertex or a mixed simplex: the world axis
}

// The deepest points along the function's axis at time t
// (src/distance.c:931-984).
M3_HD m3real m3ToiMinSep(const m3SepFn* f, int32_t* indexA, int32_t* indexB, m3real t)
{
    m3Transform xfA = m3SweepTransform(f->sweepA, t);
    m3Transform xfB = m3SweepTransform(f->sweepB, t);
    const m3Vec3 deltaP =
        m3V3((m3real)(xfB.p.x - xfA.p.x), (m3real)(xfB.p.y - xfA.p.y), (m3real)(xfB.p.z - xfA.p.z));
    switch (f->type)
    {
    case M3_SEP_VERTICES:
    {
        const m3Vec3 axis = f->witness1;
        *indexA = m3ToiSupport(f->proxyA, m3InvRotateVec3(xfA.q, axis));
        *indexB = m3ToiSupport(f->proxyB, m3InvRotateVec3(xfB.q, m3Neg3(axis)));
        const m3Vec3 delta = m3Add3(m3Sub3(m3RotateVec3(xfB.q, f->proxyB->points[*indexB]),
                                           m3RotateVec3(xfA.q, f->proxyA->points[*indexA])),
                                    deltaP);
        return m3Dot3(delta, axis);
    }
    case M3_SEP_EDGES:
    {
        const m3Vec3 edgeA = m3RotateVec3(xfA.q, f->witness1);
        const m3Vec3 edgeB = m3RotateVec3(xfB.q, f->witness2);
        const m3Vec3 axis = m3Normalize3(m3Cross3(edgeA, edgeB));
        *indexA = m3ToiSupport(f->proxyA, m3InvRotateVec3(xfA.q, axis));
        *indexB = m3ToiSupport(f->proxyB, m3InvRotateVec3(xfB.q, m3Neg3(axis)));
        const m3Vec3 delta = m3Add3(m3Sub3(m3RotateVec3(xfB.q, f->proxyB->points[*indexB]),
                                           m3RotateVec3(xfA.q, f->proxyA->points[*indexA])),
                                    deltaP);
        return m3Dot3(delta, axis);
    }
    case M3_SEP_FACE_A:
    {
        const m3Vec3 normal = m3RotateVec3(xfA.q, f->witness1);
        *indexA = -1;
        const m3Vec3 pointA = m3ToiPoint(&xfA, f->witness2);
        *indexB = m3ToiSupport(f->proxyB, m3InvRotateVec3(xfB.q, m3Neg3(normal)));
        const m3Vec3 pointB = m3ToiPoint(&xfB, f->proxyB->points[*indexB]);
        return m3Dot3(m3Sub3(pointB, pointA), normal);
    }
    default:
    {
        const m3Vec3 normal = m3RotateVec3(xfB.q, f->witness1);
        *indexA = m3ToiSupport(f->proxyA, m3InvRotateVec3(xfA.q, m3Neg3(normal)));
        const m3Vec3 pointA = m3ToiPoint(&xfA, f->proxyA->points[*indexA]);
        *indexB = -1;
        const m3Vec3 pointB = m3ToiPoint(&xfB, f->witness2);
        return m3Dot3(m3Sub3(pointA, pointB), normal);
    }
    }
}

// The separation of a fixed point pair at time t (src/distance.c:986-1027).
M3_HD m3real m3ToiEvalSep(const m3SepFn* f, int32_t indexA, int32_t indexB, m3real t)
{
    m3Transform xfA = m3SweepTransform(f->sweepA, t);
    m3Transform xfB = m3SweepTransform(f->sweepB, t);
    switch (f->type)
    {
    case M3_SEP_VERTICES:
    {
        const m3Vec3 pointA = m3ToiPoint(&xfA, f->proxyA->points[indexA]);
        const m3Vec3 pointB = m3ToiPoint(&xfB, f->proxyB->points[indexB]);
        return m3Dot3(m3Sub3(pointB, pointA), f->witness1);
    }
    case M3_SEP_EDGES:
    {
        const m3Vec3 edgeA = m3RotateVec3(xfA.q, f->witness1);
        const m3Vec3 edgeB = m3RotateVec3(xfB.q, f->witness2);
        const m3Vec3 axis = m3Normalize3(m3Cross3(edgeA, edgeB));
        const m3Vec3 pointA = m3ToiPoint(&xfA, f->proxyA->points[indexA]);
        const m3Vec3 pointB = m3ToiPoint(&xfB, f->proxyB->points[indexB]);
        return m3Dot3(m3Sub3(pointB, pointA), axis);
    }
    case M3_SEP_FACE_A:
    {
        const m3Vec3 axis = m3RotateVec3(xfA.q, f->witness1);
        const m3Vec3 pointA = m3ToiPoint(&xfA, f->witness2);
        const m3Vec3 pointB = m3ToiPoint(&xfB, f->proxyB->points[indexB]);
        return m3Dot3(m3Sub3(pointB, pointA), axis);
    }
    default:
    {
        const m3Vec3 axis = m3RotateVec3(xfB.q, f->witness1);
        const m3Vec3 pointA = m3ToiPoint(&xfA, f->proxyA->points[indexA]);
        const m3Vec3 pointB = m3ToiPoint(&xfB, f->witness2);
        return m3Dot3(m3Sub3(pointA, pointB), axis);
    }
    }
}

// An edge function whose root search failed becomes a fixed world axis
// (src/distance.c:1029-1037).
M3_HD void m3ToiFixAxis(m3SepFn* f, m3real t)
{
    m3Transform xfA = m3SweepTransform(f->sweepA, t);
    m3Transform xfB = m3SweepTransform(f->sweepB, t);
    const m3Vec3 edgeA = m3RotateVec3(xfA.q, f->witness1);
    const m3Vec3 edgeB = m3RotateVec3(xfB.q, f->witness2);
    f->witness1 = m3Normalize3(m3Cross3(edgeA, edgeB));
    f->witness2 = m3V3(0.0f, 0.0f, 0.0f);
    f->type = M3_SEP_VERTICES;
}

// m3TimeOfImpact (src/distance.c:1039-1215); returns the state and writes the
// fraction.
M3_HD int32_t m3TimeOfImpact(const m3Proxy* proxyA, m3Sweep sweepA, const m3Proxy* proxyB, m3Sweep sweepB,
                             m3real tMax, m3real* fractionOut)
{
    int32_t state = M3_TOI_UNKNOWN;
    m3real fraction = tMax;
    const m3Vec3 origin = sweepA.c1;
    sweepA.c1 = m3Sub3(sweepA.c1, origin);
    sweepA.c2 = m3Sub3(sweepA.c2, origin);
    sweepB.c1 = m3Sub3(sweepB.c1, origin);
    sweepB.c2 = m3Sub3(sweepB.c2, origin);
    const m3real linearSlop = M3_LINEAR_SLOP;
    const m3real totalRadius = proxyA->radius + proxyB->radius;
    const m3real target = m3MaxF(linearSlop, totalRadius - linearSlop);
    const m3real tolerance = 0.25f * linearSlop;
    const int32_t maxPushBack = proxyA->count + proxyB->count;
    m3real t1 = 0.0f;
    int32_t distanceIterations = 0;
    m3SimplexCache cache;
    cache.count = 0;
    cache.metric = 0.0f;
    for (;;)
    {
        m3Transform xfA = m3SweepTransform(&sweepA, t1);
        m3Transform xfB = m3SweepTransform(&sweepB, t1);
        const m3Quat conjA = m3Q(-xfA.q.x, -xfA.q.y, -xfA.q.z, xfA.q.w);
        const m3Quat q = m3MulQuat(conjA, xfB.q);
        const m3Vec3 dp = m3V3((m3real)(xfB.p.x - xfA.p.x), (m3real)(xfB.p.y - xfA.p.y),
                               (m3real)(xfB.p.z - xfA.p.z));
        const m3Vec3 p = m3InvRotateVec3(xfA.q, dp);
        const m3GjkOut d = m3ShapeDistanceCached(proxyA, proxyB, q, p, &cache);
        distanceIterations += 1;
        if (d.distance <= 0.0f)
        {
            state = M3_TOI_OVERLAPPED;
            fraction = 0.0f;
            break;
        }
        if (d.distance <= target + tolerance)
        {
            state = M3_TOI_HIT;
            fraction = t1;
            break;
        }
        if (distanceIterations == 25)
        {
            state = M3_TOI_FAILED;
            fraction = t1;
            break;
        }
        const m3Vec3 worldNormal = m3RotateVec3(xfA.q, d.normal);
        m3SepFn f = m3ToiMakeSep(&cache, proxyA, &sweepA, proxyB, &sweepB, worldNormal, t1);
        bool done = false;
        m3real t2 = tMax;
        int32_t pushBack = 0;
        for (;;)
        {
            int32_t indexA;
            int32_t indexB;
            m3real s2 = m3ToiMinSep(&f, &indexA, &indexB, t2);
            if (s2 - target > tolerance)
            {
                state = M3_TOI_SEPARATED;
                fraction = tMax;
                done = true;
                break;
            }
            if (s2 >= target - tolerance)
            {
                t1 = t2;
                break;
            }
            m3real s1 = m3ToiEvalSep(&f, indexA, indexB, t1);
            if (s1 < target - tolerance)
            {
                state = M3_TOI_FAILED;
                fraction = t1;
                done = true;
                break;
            }
            if (s1 <= target + tolerance)
            {
                state = M3_TOI_HIT;
                fraction = t1;
                done = true;
                break;
            }
            // False position alternating with bisection.
            int32_t rootIterations = 0;
            m3real a1 = t1;
            m3real a2 = t2;
            for (;;)
            {
                m3real t;
                if (rootIterations & 1)
                {
                    t = a1 + (target - s1) * (a2 - a1) / (s2 - s1);
                }
                else
                {
                    t = 0.5f * (a1 + a2);
                }
                rootIterations += 1;
                const m3real sep = m3ToiEvalSep(&f, indexA, indexB, t);
                const m3real err = sep - target;
                if ((err < 0.0f ? -err : err) <= tolerance)
                {
                    t2 = t;
                    break;
                }
                if (sep > target)
                {
                    a1 = t;
                    s1 = sep;
                }
                else
                {
                    a2 = t;
                    s2 = sep;
                }
                if (rootIterations == 50)
                {
                    break;
                }
            }
            if (rootIterations == 49 && f.type == M3_SEP_EDGES)
            {
                t2 = tMax;
                m3ToiFixAxis(&f, t1);
            }
            pushBack += 1;
            if (pushBack == maxPushBack)
            {
                break;
            }
        }
        if (done)
        {
            break;
        }
    }
    *fractionOut = fraction;
    return state;
}

M3_HD m3real m3ContinuousVersusPlane(const m3View* v, int32_t fastBody, int32_t fastShape,
                                     const m3Sweep* sweep, m3Pos3 base, int32_t planeShape,
                                     m3real fraction)
{
    m3Vec3 n = v->geom[planeShape].v;
    m3real offset = v->geom[planeShape].s -
                    (m3real)((double)n.x * base.x + (double)n.y * base.y + (double)n.z * base.z);
    m3Vec3 scratch[2];
    m3Proxy proxy = m3MakeProxy(v, fastShape, scratch);
    const m3real linearSlop = M3_LINEAR_SLOP;
    m3real target = m3MaxF(linearSlop, proxy.radius - linearSlop);
    m3real tolerance = 0.25f * linearSlop;
    m3Vec3 travel = m3Sub3(sweep->c2, sweep->c1);
    m3Quat dq = m3MulQuat(sweep->q2, m3Q(-sweep->q1.x, -sweep->q1.y, -sweep->q1.z, sweep->q1.w));
    m3real arc = 2.0f * m3Sqrt(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z) * v->maxExt[fastBody];
    m3real rate = m3Sqrt(m3Dot3(travel, travel)) + arc;
    if (!(rate > 0.0f))
    {
        return fraction;
    }
    m3real t = 0.0f;
    for (int32_t iter = 0; iter < 25; ++iter)
    {
        m3Transform xf = m3SweepTransform(sweep, t);
        m3real minD = 3.4e38f;
        for (int32_t k = 0; k < proxy.count; ++k)
        {
            m3Vec3 r = m3RotateVec3(xf.q, proxy.points[k]);
            m3real d = n.x * ((m3real)xf.p.x + r.x) + n.y * ((m3real)xf.p.y + r.y) +
                       n.z * ((m3real)xf.p.z + r.z);
            minD = m3MinF(minD, d);
        }
        m3real sep = minD - offset;
        if (sep <= 0.0f)
        {
            return fraction;
        }
        if (sep <= target + tolerance)
        {
            if (t > 0.0f && t < fraction)
            {
                fraction = t;
            }
            return fraction;
        }
        t += (sep - target) / rate;
        if (t >= fraction)
        {
            return fraction;
        }
    }
    return fraction;
}

// Poses the continuous pass sees for its targets.  The reference sweeps the
// bodies one by one in slot order and pulls each back at once
// (src/solver.c:3349-3430), so a bullet sweeping against a dynamic target
// of a lower slot sees that target's pulled-back pose.  The serial form
// applies pull-backs as it goes (poses = null: the current transforms);
// the parallel form computes the non-bullet bodies first into `pulled` /
// `hit` and gives bullets the poses of lower slots from there.
typedef struct m3CcdPoses
{
    const m3Transform* pulled;
    const uint8_t* hit;
} m3CcdPoses;

M3_HD m3Transform m3CcdPose(const m3View* v, const m3CcdPoses* poses, int32_t body, int32_t fastBody)
{
    if (poses != NULL && body < fastBody && poses->hit[body] != 0)
    {
        return poses->pulled[body];
    }
    return v->xf[body];
}

// MakeRelativeSweep for a body at pose `xf` (src/solver.c:3065-3079).
M3_HD m3Sweep m3RelativeSweepAt(const m3View* v, int32_t body, m3Transform xf, m3Pos3 base)
{
    m3Sweep sweep;
    sweep.localCenter = v->lc[body];
    sweep.c1 = m3V3((m3real)(v->com0[body].x - base.x), (m3real)(v->com0[body].y - base.y),
                    (m3real)(v->com0[body].z - base.z));
    m3Vec3 rlc = m3RotateVec3(xf.q, v->lc[body]);
    sweep.c2 = m3V3((m3real)(xf.p.x + (double)rlc.x - base.x), (m3real)(xf.p.y + (double)rlc.y - base.y),
                    (m3real)(xf.p.z + (double)rlc.z - base.z));
    sweep.q1 = v->rot0[body];
    sweep.q2 = xf.q;
    return sweep;
}

// The next tree candidate of a continuous query: among the shapes whose tree
// leaf overlaps [lo, hi], the one of smallest depth-first rank above
// `after` (-1 when none).  The reference meets its candidates in exactly
// this order (a depth-first walk visits the leaves it reaches in the full
// walk's order).  Non-bullet fast bodies only ever hit static targets, so
// their scan runs over the static tree shapes only.
M3_HD int32_t m3CcdNextCandidate(const m3View* v, const double box[6], int32_t after, int allShapes,
                                 int32_t* rankOut)
{
    int32_t best = -1;
    int32_t bestRank = 0x7FFFFFFF;
    const int32_t n = allShapes ? v->maxShape : v->ccdStaticCount;
    for (int32_t k = 0; k < n; ++k)
    {
        const int32_t c = allShapes ? k : v->ccdStatics[k];
        const int32_t r = v->shapeRank[c];
        if (r <= after || r >= bestRank)
        {
            continue;
        }
        if (!m3BoxOverlap(&v->leaf[6 * c], box))
        {
            continue;
        }
        best = c;
        bestRank = r;
    }
    *rankOut = bestRank;
    return best;
}

// One tree candidate of ContinuousQueryCallback (src/solver.c:3081-3278):
// the filters in the reference's order, then the time of impact with the
// fraction threaded through.
M3_HD m3real m3CcdVersusShape(const m3View* v, const m3CcdPoses* poses, int32_t fastBody, int32_t fastShape,
                              const m3Sweep* fastSweep, m3Pos3 base, int32_t shape, m3real fraction)
{
    if (shape == fastShape)
    {
        return fraction;
    }
    const int32_t body = v->shapeBody[shape];
    if (body == fastBody || v->enabled[body] == 0 || v->bullet[body] != 0 || v->sensor[shape] != 0)
    {
        return fraction;
    }
    if (v->type[body] != M3_STATIC && v->bullet[fastBody] == 0)
    {
        return fraction; // only bullets sweep against dynamic and kinematic bodies
    }
    const int32_t gi = v->group[shape];
    if (gi != 0 && gi == v->group[fastShape])
    {
        if (gi < 0)
        {
            return fraction;
        }
    }
    else if (!((v->category[shape] & v->mask[fastShape]) != 0 &&
               (v->category[fastShape] & v->mask[shape]) != 0))
    {
        return fraction;
    }
    m3Vec3 scratchA[2];
    m3Vec3 scratchB[2];
    const m3Proxy proxyA = m3MakeProxy(v, shape, scratchA);
    const m3Proxy proxyB = m3MakeProxy(v, fastShape, scratchB);
    const m3Sweep sweepA = m3RelativeSweepAt(v, body, m3CcdPose(v, poses, body, fastBody), base);
    m3real t = fraction;
    const int32_t state = m3TimeOfImpact(&proxyA, sweepA, &proxyB, *fastSweep, fraction, &t);
    if (state == M3_TOI_HIT && 0.0f < t && t < fraction)
    {
        return t;
    }
    return fraction;
}

// The fast test, the tree and plane sweeps of one dynamic body
// (src/solver.c:3349-3430).  Returns 1 and the pulled-back pose when the
// body is pulled back; nothing is written to the view.  Tree candidates are
// swept only when the step carries the tree's leaf ranks (v->treeCcd).
M3_HD int m3ContinuousSweep(const m3View* v, const m3CcdPoses* poses, int32_t i, m3Transform* pulled)
{
    if (v->bodyAlive[i] == 0 || v->type[i] != M3_DYNAMIC)
    {
        return 0;
    }
    m3Transform xfi = v->xf[i];
    m3Vec3 rlc = m3RotateVec3(xfi.q, v->lc[i]);
    double cx = xfi.p.x + (double)rlc.x;
    double cy = xfi.p.y + (double)rlc.y;
    double cz = xfi.p.z + (double)rlc.z;
    m3Pos3 c0 = v->com0[i];
    m3Vec3 dc = m3V3((m3real)(cx - c0.x), (m3real)(cy - c0.y), (m3real)(cz - c0.z));
    m3Quat q0 = v->rot0[i];
    m3Quat dq = m3MulQuat(xfi.q, m3Q(-q0.x, -q0.y, -q0.z, q0.w));
    m3real arc = 2.0f * m3Sqrt(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z) * v->maxExt[i];
    m3real maxMotion = m3Sqrt(m3Dot3(dc, dc)) + arc;
    if (!(maxMotion > 0.5f * v->minExt[i]))
    {
        return 0;
    }
    m3Pos3 base = c0;
    m3real fraction = 1.0f;
    m3Sweep sweep = m3RelativeSweepAt(v, i, xfi, base);
    // Swept candidate box: both centre-of-mass endpoints padded by the
    // body's largest extent.
    const double pad = (double)(v->maxExt[i] + M3_AABB_MARGIN);
    double box[6];
    box[0] = (c0.x < cx ? c0.x : cx) - pad;
    box[1] = (c0.y < cy ? c0.y : cy) - pad;
    box[2] = (c0.z < cz ? c0.z : cz) - pad;
    box[3] = (c0.x > cx ? c0.x : cx) + pad;
    box[4] = (c0.y > cy ? c0.y : cy) + pad;
    box[5] = (c0.z > cz ? c0.z : cz) + pad;
    const int allShapes = v->bullet[i] != 0;
    for (int32_t s = v->shapeHead[i]; s != -1; s = v->shapeNext[s])
    {
        if (v->sensor[s] != 0)
        {
            continue;
        }
        if (v->treeCcd != 0)
        {
            int32_t rank = -1;
            for (;;)
            {
                const int32_t c = m3CcdNextCandidate(v, box, rank, allShapes, &rank);
                if (c < 0)
                {
                    break;
                }
                fraction = m3CcdVersusShape(v, poses, i, s, &sweep, base, c, fraction);
            }
        }
        for (int32_t pi = 0; pi < v->planeCount; ++pi)
        {
            fraction = m3ContinuousVersusPlane(v, i, s, &sweep, base, v->planes[pi], fraction);
        }
    }
    if (fraction < 1.0f)
    {
        m3Transform xf = m3SweepTransform(&sweep, fraction);
        pulled->q = xf.q;
        pulled->p.x = base.x + xf.p.x;
        pulled->p.y = base.y + xf.p.y;
        pulled->p.z = base.z + xf.p.z;
        return 1;
    }
    return 0;
}

// The serial form: sweep and apply in place (bodies in slot order give the
// reference's sequence).
M3_HD int m3ContinuousBody(const m3View* v, int32_t i)
{
    m3Transform xf;
    if (m3ContinuousSweep(v, NULL, i, &xf))
    {
        v->xf[i] = xf;
        return 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Sleep (src/solver.c:3582-3625): the per-body timer update.
// ---------------------------------------------------------------------------

M3_HD void m3SleepTimer(const m3View* v, int32_t i, m3real dt)
{
    m3real invDt = dt > 0.0f ? 1.0f / dt : 0.0f;
    m3Vec3 lin = v->lv[i];
    m3Vec3 ang = v->av[i];
    m3real velocity = m3Sqrt(m3Dot3(lin, lin)) + m3Sqrt(m3Dot3(ang, ang)) * v->maxExt[i];
    m3Transform xf = v->xf[i];
    m3Vec3 rlc = m3RotateVec3(xf.q, v->lc[i]);
    m3Pos3 c0 = v->com0[i];
    m3Vec3 dc = m3V3((m3real)(xf.p.x + (double)rlc.x - c0.x), (m3real)(xf.p.y + (double)rlc.y - c0.y),
                     (m3real)(xf.p.z + (double)rlc.z - c0.z));
    m3Quat q0 = v->rot0[i];
    m3Quat dq = m3MulQuat(xf.q, m3Q(-q0.x, -q0.y, -q0.z, q0.w));
    m3real motion = m3Sqrt(m3Dot3(dc, dc)) +
                    2.0f * m3Sqrt(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z) * v->maxExt[i];
    m3real sleepVelocity = m3MaxF(velocity, 0.5f * invDt * motion);
    if (v->canSleep[i] == 0)
    {
        sleepVelocity = 3.4e38f;
    }
    if (sleepVelocity < v->sleepThreshold[i])
    {
        v->sleepTime[i] += dt;
    }
    else
    {
        v->sleepTime[i] = 0.0f;
    }
}

#endif // M3_STEP_H
