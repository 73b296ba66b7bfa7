// One world's whole step on one thread, over the same per-item functions
// the parallel kernels use (m3_step.h), in the reference's serial order
// (src/solver.c:3662-4320).  It serves small worlds whose step must be
// replayed with a record of its intermediate values: the adjoint replays
// checkpointed segments with it, one thread per world, and the gradient
// checker instantiates it at double precision (M3_REAL).  Joints and
// kinematic targets are outside its scope (the caller refuses such worlds).
//
// Pair list: every shape pair is visited once in ascending compact-key
// order and emitted by the same tree, plane and frozen-buffer rules as the
// sweep (m3TreePairEmits is symmetric in its shapes), so the list comes out
// sorted and unique without a sort.
#ifndef M3_SERIAL_H
#define M3_SERIAL_H

#include "m3_step.h"

// Per-world scratch of the serial step.
typedef struct m3SerialScratch
{
    int32_t* root;   // maxBody: island root of each body this step
    uint32_t* masks; // maxBody: color masks
} m3SerialScratch;

// Intermediate values of one step, written when a tape is given.  Arrays
// are per step of a segment; `step` selects the slice.
typedef struct m3TapePass
{
    m3Vec3 vA, wA, vB, wB; // body velocities before the constraint is solved
    m3real lambda[M3_MANIFOLD_MAX_POINTS];
    m3real f1, f2, twist; // accumulators before the constraint is solved
} m3TapePass;

typedef struct m3Tape
{
    int32_t bodyCap;
    int32_t pairCap;
    int32_t substeps;
    // per step
    m3Transform* xf0;  // [bodyCap] pose at step start
    m3Vec3* lv0;       // [bodyCap]
    m3Vec3* av0;       // [bodyCap]
    m3Vec3* force0;    // [bodyCap] applied force of the step
    m3Transform* xfEnd; // [bodyCap] pose after the substeps (continuous pass input)
    uint8_t* slept;    // [bodyCap] frozen by this step's sleep pass
    uint8_t* ccd;      // [bodyCap] pulled back by this step's continuous pass
    uint8_t* mover;    // [bodyCap] solved this step
    uint64_t* keys;    // [pairCap] this step's pair list
    m3Manifold* man;   // [pairCap] manifolds after narrowphase and the carry
    int32_t* carryPair;   // [pairCap] previous-step pair the carry read, -1 none
    int8_t* carryPoint;   // [pairCap * 4] previous-step point each point matched, -1 none
    int32_t* consPair;    // [pairCap] pair of each constraint, slot order
    int32_t* counts;      // [4]: pairs, constraints, movers, color starts follow
    int32_t* colorStarts; // [M3_GRAPH_COLORS + 2]
    // per substep
    m3Transform* sxf; // [substeps * bodyCap] at substep start
    m3Vec3* slv;      // [substeps * bodyCap] at substep start
    m3Vec3* sav;
    m3Vec3* sdp;      // [substeps * bodyCap] accumulated position change
    m3Quat* sdq;
    m3Vec3* ilv;      // [substeps * bodyCap] after integrate-velocities
    m3Vec3* iav;
    m3Vec3* plv;      // [substeps * bodyCap] after the solve pass (integrate-positions input)
    m3Vec3* pav;
    m3Vec3* rdp;      // [substeps * bodyCap] after integrate-positions (relax pass input)
    m3Quat* rdq;
    m3TapePass* pass; // [substeps * 2 * pairCap] pre-solve records, slot order
    m3TapePass* fin;  // [pairCap] accumulators after the last pass, slot order
} m3Tape;

M3_HD int32_t m3SerialFind(const int32_t* parent, int32_t i)
{
    while (parent[i] != i)
    {
        i = parent[i];
    }
    return i;
}

M3_HD void m3SerialUnion(int32_t* parent, int32_t a, int32_t b)
{
    int32_t ra = m3SerialFind(parent, a);
    int32_t rb = m3SerialFind(parent, b);
    if (ra != rb)
    {
        if (ra < rb)
        {
            parent[rb] = ra;
        }
        else
        {
            parent[ra] = rb;
        }
    }
}

// Is the compact key in the (sorted) frozen buffer?
M3_HD int m3SerialFrozenHas(const uint64_t* keys, int32_t count, uint64_t key)
{
    int32_t lo = 0;
    int32_t hi = count - 1;
    while (lo <= hi)
    {
        int32_t mid = (lo + hi) / 2;
        if (keys[mid] == key)
        {
            return 1;
        }
        if (keys[mid] < key)
        {
            lo = mid + 1;
        }
        else
        {
            hi = mid - 1;
        }
    }
    return 0;
}

// Scalar parameters of the step at the scalar type.
typedef struct m3SerialParams
{
    m3real dt;
    int32_t substeps;
    m3Vec3 gravity;
    m3real contactHertz;
    m3real contactDampingRatio;
    m3real contactPushMaxSpeed;
    m3real restitutionThreshold;
    m3real maximumLinearSpeed;
    m3real maximumAngularSpeed;
    int32_t sleepEnabled;
    int32_t continuousEnabled;
} m3SerialParams;

// The step.  The caller has pointed v->keys/manifolds at this step's buffers
// and v->oldKeys/oldManifolds at the previous step's.  Returns 0, or -1 on a
// pair-capacity overflow (the world is then left unadvanced, as a stall).
// The frozen buffer v->sleepKeys holds sorted keys; discovery appends in the
// reference are merged the same way here: a pair discovered this step joins
// the frozen set before the merge.
// The step through the island wake pass (pair list, narrowphase, islands):
// where the pre-solve callback's candidates are read.  Nonzero on a capacity
// stall (the world did not advance).
M3_HD int m3SerialStepA(const m3View* v, m3SerialScratch* s, const m3SerialParams* p, m3Tape* tape)
{
    m3Scalars* sc = v->sc;
    int32_t nb = v->maxBody;
    int32_t ns = v->maxShape;
    int32_t oldPairCount = sc->pairCount;
    for (int32_t a = 0; a < ns; ++a)
    {
        if (v->shapeAlive[a] && v->shapeType[a] != M3_PLANE)
        {
            m3FreshBox(v, a, &v->fresh[6 * a]);
        }
    }
    if (sc->fullQuery)
    {
        sc->sleepCount = 0;
        for (int32_t b = 0; b < nb; ++b)
        {
            v->pending[b] = (v->bodyAlive[b] && v->type[b] != M3_STATIC && v->awake[b] == 0) ? 1 : 0;
        }
        sc->fullQuery = 0;
    }
    // Pairs in ascending key order.
    int32_t frozenCount = sc->sleepCount;
    int32_t count = 0;
    for (int32_t a = 0; a < ns; ++a)
    {
        if (!v->shapeAlive[a])
        {
            continue;
        }
        for (int32_t b = a + 1; b < ns; ++b)
        {
            if (!v->shapeAlive[b])
            {
                continue;
            }
            int planeA = v->shapeType[a] == M3_PLANE;
            int planeB = v->shapeType[b] == M3_PLANE;
            if (planeA && planeB)
            {
                continue;
            }
            uint64_t key = ((uint64_t)a << 32) | (uint64_t)b;
            int emit = 0;
            int discovered = 0;
            if (planeA || planeB)
            {
                int32_t pl = planeA ? a : b;
                int32_t sh = planeA ? b : a;
                if (m3PairAllowed(v, pl, sh))
                {
                    discovered = v->pending[v->shapeBody[sh]] != 0;
                    emit = m3PairHot(v, pl, sh);
                }
            }
            else if (m3BoxOverlap(&v->fresh[6 * a], &v->fresh[6 * b]))
            {
                discovered = (v->pending[v->shapeBody[a]] && m3DiscoverHit(v, a, b)) ||
                             (v->pending[v->shapeBody[b]] && m3DiscoverHit(v, b, a));
                emit = m3TreePairEmits(v, a, b);
            }
            if (!emit && !m3PairHot(v, a, b) &&
                (discovered || m3SerialFrozenHas(v->sleepKeys, frozenCount, key)))
            {
                emit = 1;
            }
            if (emit)
            {
                if (count >= v->pairCapacity)
                {
                    sc->stall = 1;
                    return -1;
                }
                v->keys[count++] = key;
            }
        }
    }
    // Leaf refresh of awake shapes (after the pair rules read the leaves).
    for (int32_t a = 0; a < ns; ++a)
    {
        if (!v->shapeAlive[a] || v->shapeType[a] == M3_PLANE || v->awake[v->shapeBody[a]] == 0)
        {
            continue;
        }
        const double* f = &v->fresh[6 * a];
        double* leaf = &v->leaf[6 * a];
        if (leaf[0] <= f[0] && leaf[1] <= f[1] && leaf[2] <= f[2] && f[3] <= leaf[3] &&
            f[4] <= leaf[4] && f[5] <= leaf[5])
        {
            continue;
        }
        for (int32_t k = 0; k < 3; ++k)
        {
            leaf[k] = f[k] - (double)M3_AABB_MARGIN;
            leaf[3 + k] = f[3 + k] + (double)M3_AABB_MARGIN;
        }
    }
    sc->oldPairCount = oldPairCount;
    sc->pairCount = count;
    sc->sleepCount = 0;
    for (int32_t i = 0; i < count; ++i)
    {
        uint64_t key = v->keys[i];
        if (!m3PairHot(v, (int32_t)(key >> 32), (int32_t)(key & 0xFFFFFFFFu)))
        {
            v->sleepKeys[sc->sleepCount++] = key;
        }
    }
    if (tape != NULL)
    {
        for (int32_t b = 0; b < nb; ++b)
        {
            tape->xf0[b] = v->xf[b];
            tape->lv0[b] = v->lv[b];
            tape->av0[b] = v->av[b];
            tape->force0[b] = v->force[b];
            tape->slept[b] = 0;
            tape->ccd[b] = 0;
            tape->mover[b] = 0;
        }
    }
    for (int32_t i = 0; i < count; ++i)
    {
        if (v->hitRecs != nullptr)
        {
            v->hitRecs[i].flag = 0;
        }
        m3NarrowPair(v, i, oldPairCount);
    }
    if (tape != NULL)
    {
        // Which previous pair and points the carry read (the same search as
        // m3NarrowPair's).
        for (int32_t i = 0; i < count; ++i)
        {
            const m3Manifold* m = &v->manifolds[i];
            tape->keys[i] = v->keys[i];
            tape->man[i] = *m;
            tape->carryPair[i] = -1;
            for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
            {
                tape->carryPoint[4 * i + k] = -1;
            }
            if (m->pointCount == 0 || oldPairCount == 0)
            {
                continue;
            }
            int32_t lo = 0;
            int32_t hi = oldPairCount - 1;
            while (lo <= hi)
            {
                int32_t mid = (lo + hi) / 2;
                uint64_t k2 = v->oldKeys[mid];
                if (k2 == v->keys[i])
                {
                    tape->carryPair[i] = mid;
                    const m3Manifold* prev = &v->oldManifolds[mid];
                    for (int32_t k = 0; k < m->pointCount; ++k)
                    {
                        for (int32_t o = 0; o < prev->pointCount; ++o)
                        {
                            if (prev->points[o].id == m->points[k].id)
                            {
                                tape->carryPoint[4 * i + k] = (int8_t)o;
                                break;
                            }
                        }
                    }
                    break;
                }
                if (k2 < v->keys[i])
                {
                    lo = mid + 1;
                }
                else
                {
                    hi = mid - 1;
                }
            }
        }
    }
    for (int32_t i = 0; i < nb; ++i)
    {
        if (v->bodyAlive[i])
        {
            m3Transform xf = v->xf[i];
            m3Vec3 rlc = m3RotateVec3(xf.q, v->lc[i]);
            v->com0[i].x = xf.p.x + (double)rlc.x;
            v->com0[i].y = xf.p.y + (double)rlc.y;
            v->com0[i].z = xf.p.z + (double)rlc.z;
            v->rot0[i] = xf.q;
        }
        v->parent[i] = i;
        v->forced[i] = 0;
        v->flag[i] = 0;
        v->pending[i] = 0;
        v->deltaPos[i] = m3V3(0.0f, 0.0f, 0.0f);
        v->deltaRot[i] = m3IdentityQuat();
    }
    // Islands and wake.
    for (int32_t i = 0; i < count; ++i)
    {
        if (v->manifolds[i].pointCount == 0)
        {
            continue;
        }
        uint64_t key = v->keys[i];
        int32_t sA = (int32_t)(key >> 32);
        int32_t sB = (int32_t)(key & 0xFFFFFFFFu);
        if (v->sensor[sA] || v->sensor[sB])
        {
            continue;
        }
        int32_t bA = v->shapeBody[sA];
        int32_t bB = v->shapeBody[sB];
        int dynA = v->type[bA] == M3_DYNAMIC;
        int dynB = v->type[bB] == M3_DYNAMIC;
        if (dynA && dynB)
        {
            m3SerialUnion(v->parent, bA, bB);
        }
        else if (dynA || dynB)
        {
            int32_t kin = dynA ? bB : bA;
            int32_t dyn = dynA ? bA : bB;
            if (v->type[kin] == M3_KINEMATIC)
            {
                m3Vec3 l = v->lv[kin];
                m3Vec3 an = v->av[kin];
                if (m3Dot3(l, l) > 0.0f || m3Dot3(an, an) > 0.0f)
                {
                    v->forced[dyn] = 1;
                }
            }
        }
    }
    for (int32_t i = 0; i < nb; ++i)
    {
        s->root[i] = m3SerialFind(v->parent, i);
    }
    for (int32_t i = 0; i < nb; ++i)
    {
        if (v->bodyAlive[i] && v->type[i] == M3_DYNAMIC && (v->awake[i] || v->forced[i]))
        {
            v->flag[s->root[i]] = 1;
        }
    }
    for (int32_t i = 0; i < nb; ++i)
    {
        if (v->bodyAlive[i] && v->type[i] == M3_DYNAMIC && !v->awake[i] && v->flag[s->root[i]])
        {
            v->awake[i] = 1;
            v->sleepTime[i] = 0.0f;
        }
    }
    return 0;
}

// From the constraint list (pairs vetoed through v->veto left out) to the
// impulse store: everything before the continuous pass.
M3_HD void m3SerialStepB(const m3View* v, m3SerialScratch* s, const m3SerialParams* p, m3Tape* tape)
{
    m3Scalars* sc = v->sc;
    const int32_t nb = v->maxBody;
    const int32_t count = sc->pairCount;
    // Constraints, coloring, prepare in color-slot order.
    m3real h = p->dt / (m3real)p->substeps;
    m3real invH = h > 0.0f ? 1.0f / h : 0.0f;
    m3Softness soft = m3MakeSoft(p->contactHertz, p->contactDampingRatio, h);
    m3Softness staticSoft = m3MakeSoft(2.0f * p->contactHertz, p->contactDampingRatio, h);
    int32_t nc = 0;
    for (int32_t i = 0; i < count; ++i)
    {
        if (m3PairIsConstraint(v, i))
        {
            v->consPair[nc++] = i;
        }
    }
    sc->constraintCount = nc;
    for (int32_t b = 0; b < nb; ++b)
    {
        s->masks[b] = 0u;
    }
    int32_t counts[M3_GRAPH_COLORS + 1];
    for (int32_t k = 0; k <= M3_GRAPH_COLORS; ++k)
    {
        counts[k] = 0;
    }
    for (int32_t c = 0; c < nc; ++c)
    {
        uint64_t key = v->keys[v->consPair[c]];
        int32_t a = v->shapeBody[(int32_t)(key >> 32)];
        int32_t b = v->shapeBody[(int32_t)(key & 0xFFFFFFFFu)];
        int32_t ca = (v->type[a] == M3_DYNAMIC && v->awake[a]) ? a : -1;
        int32_t cb = (v->type[b] == M3_DYNAMIC && v->awake[b]) ? b : -1;
        uint32_t mask = (ca >= 0 ? s->masks[ca] : 0u) | (cb >= 0 ? s->masks[cb] : 0u);
        int32_t color = M3_GRAPH_COLORS;
        for (int32_t bit = 0; bit < M3_GRAPH_COLORS; ++bit)
        {
            if (!(mask & (1u << bit)))
            {
                color = bit;
                break;
            }
        }
        v->colors[c] = (uint8_t)color;
        if (color < M3_GRAPH_COLORS)
        {
            if (ca >= 0)
            {
                s->masks[ca] |= 1u << color;
            }
            if (cb >= 0)
            {
                s->masks[cb] |= 1u << color;
            }
        }
        counts[color] += 1;
    }
    int32_t cursor = 0;
    int32_t used = 0;
    for (int32_t k = 0; k <= M3_GRAPH_COLORS; ++k)
    {
        sc->colorStarts[k] = cursor;
        cursor += counts[k];
        used += counts[k] > 0;
    }
    sc->colorStarts[M3_GRAPH_COLORS + 1] = cursor;
    sc->colorCount = used;
    int32_t fill[M3_GRAPH_COLORS + 1];
    for (int32_t k = 0; k <= M3_GRAPH_COLORS; ++k)
    {
        fill[k] = sc->colorStarts[k];
    }
    for (int32_t c = 0; c < nc; ++c)
    {
        int32_t k = fill[v->colors[c]]++;
        v->colorLists[k] = c;
    }
    for (int32_t k = 0; k < nc; ++k)
    {
        int32_t c = v->colorLists[k];
        m3PrepareContact(v, v->consPair[c], &v->cons[k], soft, staticSoft);
        v->slotOf[c] = k;
    }
    if (tape != NULL)
    {
        tape->counts[0] = count;
        tape->counts[1] = nc;
        for (int32_t k = 0; k < nc; ++k)
        {
            tape->consPair[k] = v->consPair[v->colorLists[k]];
        }
        for (int32_t k = 0; k < M3_GRAPH_COLORS + 2; ++k)
        {
            tape->colorStarts[k] = sc->colorStarts[k];
        }
    }
    int32_t nm = 0;
    for (int32_t i = 0; i < nb; ++i)
    {
        if (!v->bodyAlive[i] || !v->enabled[i])
        {
            continue;
        }
        uint8_t t = v->type[i];
        if (t == M3_KINEMATIC || (t == M3_DYNAMIC && v->awake[i]))
        {
            v->movers[nm++] = i;
            if (tape != NULL)
            {
                tape->mover[i] = 1;
            }
        }
    }
    sc->moverCount = nm;
    if (tape != NULL)
    {
        tape->counts[2] = nm;
    }
    for (int32_t sub = 0; sub < p->substeps; ++sub)
    {
        int32_t base = sub * (tape != NULL ? tape->bodyCap : 0);
        if (tape != NULL)
        {
            for (int32_t b = 0; b < nb; ++b)
            {
                tape->sxf[base + b] = v->xf[b];
                tape->slv[base + b] = v->lv[b];
                tape->sav[base + b] = v->av[b];
                tape->sdp[base + b] = v->deltaPos[b];
                tape->sdq[base + b] = v->deltaRot[b];
            }
        }
        for (int32_t m = 0; m < nm; ++m)
        {
            m3IntegrateVelocity(v, v->movers[m], h, p->gravity, p->maximumLinearSpeed,
                                p->maximumAngularSpeed);
        }
        if (tape != NULL)
        {
            for (int32_t b = 0; b < nb; ++b)
            {
                tape->ilv[base + b] = v->lv[b];
                tape->iav[base + b] = v->av[b];
            }
        }
        for (int32_t k = 0; k < nc; ++k)
        {
            m3WarmStartOne(v, &v->cons[k]);
        }
        for (int32_t mode = 1; mode <= 2; ++mode)
        {
            if (mode == 2)
            {
                if (tape != NULL)
                {
                    for (int32_t b = 0; b < nb; ++b)
                    {
                        tape->plv[base + b] = v->lv[b];
                        tape->pav[base + b] = v->av[b];
                    }
                }
                for (int32_t m = 0; m < nm; ++m)
                {
                    m3IntegratePosition(v, v->movers[m], h);
                }
                if (tape != NULL)
                {
                    for (int32_t b = 0; b < nb; ++b)
                    {
                        tape->rdp[base + b] = v->deltaPos[b];
                        tape->rdq[base + b] = v->deltaRot[b];
                    }
                }
            }
            for (int32_t k = 0; k < nc; ++k)
            {
                m3Constraint* c = &v->cons[k];
                if (tape != NULL)
                {
                    m3TapePass* r =
                        &tape->pass[((int64_t)sub * 2 + (mode - 1)) * tape->pairCap + k];
                    r->vA = v->lv[c->bodyA];
                    r->wA = v->av[c->bodyA];
                    r->vB = v->lv[c->bodyB];
                    r->wB = v->av[c->bodyB];
                    for (int32_t q = 0; q < M3_MANIFOLD_MAX_POINTS; ++q)
                    {
                        r->lambda[q] = q < c->pointCount ? c->points[q].normalImpulse : 0.0f;
                    }
                    r->f1 = c->frictionImpulse1;
                    r->f2 = c->frictionImpulse2;
                    r->twist = c->twistImpulse;
                }
                m3SolveContact(v, c, invH, p->contactPushMaxSpeed, mode == 1);
            }
        }
    }
    if (tape != NULL)
    {
        for (int32_t k = 0; k < nc; ++k)
        {
            const m3Constraint* c = &v->cons[k];
            m3TapePass* r = &tape->fin[k];
            r->vA = v->lv[c->bodyA];
            r->wA = v->av[c->bodyA];
            r->vB = v->lv[c->bodyB];
            r->wB = v->av[c->bodyB];
            for (int32_t q = 0; q < M3_MANIFOLD_MAX_POINTS; ++q)
            {
                r->lambda[q] = q < c->pointCount ? c->points[q].normalImpulse : 0.0f;
            }
            r->f1 = c->frictionImpulse1;
            r->f2 = c->frictionImpulse2;
            r->twist = c->twistImpulse;
        }
    }
    for (int32_t m = 0; m < nm; ++m)
    {
        v->force[v->movers[m]] = m3V3(0.0f, 0.0f, 0.0f);
        v->torque[v->movers[m]] = m3V3(0.0f, 0.0f, 0.0f);
    }
    for (int32_t c = 0; c < nc; ++c)
    {
        m3RestitutionOne(v, &v->cons[v->slotOf[c]], p->restitutionThreshold);
    }
    for (int32_t k = 0; k < nc; ++k)
    {
        m3StoreOne(v, &v->cons[k]);
    }
    if (tape != NULL)
    {
        for (int32_t b = 0; b < nb; ++b)
        {
            tape->xfEnd[b] = v->xf[b];
        }
    }
}

// The step up to the continuous pass (m3SerialStepEnd runs the rest);
// nonzero on a capacity stall (the world did not advance).
M3_HD int m3SerialStepBegin(const m3View* v, m3SerialScratch* s, const m3SerialParams* p, m3Tape* tape)
{
    if (m3SerialStepA(v, s, p, tape) != 0)
    {
        return 1;
    }
    m3SerialStepB(v, s, p, tape);
    return 0;
}

// The continuous pass (tree candidates when v->treeCcd is set), islands
// and sleep, and the step count.
M3_HD void m3SerialStepEnd(const m3View* v, m3SerialScratch* s, const m3SerialParams* p, m3Tape* tape)
{
    m3Scalars* sc = v->sc;
    const int32_t nb = v->maxBody;
    if (p->continuousEnabled)
    {
        for (int32_t i = 0; i < nb; ++i)
        {
            int pulled = m3ContinuousBody(v, i);
            if (tape != NULL)
            {
                tape->ccd[i] = (uint8_t)pulled;
            }
        }
    }
    for (int32_t i = 0; i < nb; ++i)
    {
        v->flag[i] = 1;
    }
    sc->islandCount = 0;
    for (int32_t i = 0; i < nb; ++i)
    {
        if (!v->bodyAlive[i] || v->type[i] != M3_DYNAMIC || !v->awake[i])
        {
            continue;
        }
        if (s->root[i] == i)
        {
            sc->islandCount += 1;
        }
        if (p->sleepEnabled)
        {
            m3SleepTimer(v, i, p->dt);
        }
    }
    if (p->sleepEnabled)
    {
        for (int32_t i = 0; i < nb; ++i)
        {
            if (v->bodyAlive[i] && v->type[i] == M3_DYNAMIC && v->awake[i] && v->sleepTime[i] < 0.5f)
            {
                v->flag[s->root[i]] = 0;
            }
        }
        for (int32_t i = 0; i < nb; ++i)
        {
            if (!v->bodyAlive[i] || v->type[i] != M3_DYNAMIC || !v->awake[i])
            {
                continue;
            }
            if (v->flag[s->root[i]])
            {
                v->awake[i] = 0;
                v->lv[i] = m3V3(0.0f, 0.0f, 0.0f);
                v->av[i] = m3V3(0.0f, 0.0f, 0.0f);
                v->pending[i] = 1;
                if (tape != NULL)
                {
                    tape->slept[i] = 1;
                }
            }
        }
    }
    sc->stepCount += 1;
}

M3_HD int m3SerialStep(const m3View* v, m3SerialScratch* s, const m3SerialParams* p, m3Tape* tape)
{
    if (m3SerialStepBegin(v, s, p, tape) != 0)
    {
        return 1;
    }
    m3SerialStepEnd(v, s, p, tape);
    return 0;
}

#endif // M3_SERIAL_H
