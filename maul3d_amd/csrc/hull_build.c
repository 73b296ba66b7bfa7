// Point-cloud hulls for m3CreateHullShape: QuickHull with face merging and
// the unit-density mass integral, restated from src/quickhull.c.  The hull
// feeds the world hash byte for byte (vertex order, face order and loops,
// normals, offsets, mass integrals), so every step below follows the
// reference's order: the tolerance, the four-point seed, per-face conflict
// lists with a cached farthest point, the iterative horizon walk, the cone,
// the three merge passes (flipped, concave, coplanar), orphan and face
// resolution, cleanup, emission and the half-edge build.  Comparisons are
// strict improvements, so ties keep the earliest candidate.
#include "m3_hull.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { QH_VISIBLE = 0, QH_DELETED = 1 };

typedef struct QhNode
{
    struct QhNode* prev; // NULL while the node is on no list
    struct QhNode* next;
} QhNode;

typedef struct QhFace QhFace;

typedef struct QhVert
{
    QhNode node; // first: list nodes convert back to vertices
    QhFace* conflict;
    m3Vec3 p;
    int32_t out; // index in the emitted hull
    int32_t used;
} QhVert;

typedef struct QhEdge
{
    struct QhEdge* prev; // ring of the owning face
    struct QhEdge* next;
    QhVert* origin;
    QhFace* face;
    struct QhEdge* twin;
    struct QhEdge* freeNext;
} QhEdge;

struct QhFace
{
    QhNode node; // first: list nodes convert back to faces
    QhEdge* edge;
    int32_t mark;
    float area;
    m3Vec3 normal;
    float offset;
    m3Vec3 centroid;
    float farthestDist;
    QhNode conflicts;
    QhVert* farthest;
    int32_t flipped;
    QhFace* freeNext;
};

typedef struct QhFrame
{
    QhFace* face;
    QhEdge* first;
    QhEdge* edge;
    int32_t started;
} QhFrame;

#define QH_MAX_VERTS 64
#define QH_MAX_FACES 124
#define QH_MAX_INDICES 372
#define QH_VERT_CAP (M3_HULL_MAX_INPUT + 4)
#define QH_EDGE_CAP (24 * QH_MAX_VERTS - 48)
#define QH_FACE_CAP (5 * QH_MAX_VERTS - 10)
#define QH_RING_CAP (3 * QH_MAX_VERTS - 6)
#define QH_MERGE_CAP (2 * QH_MAX_VERTS - 4)

typedef struct QhWork
{
    float tolerance;
    float minRadius;
    float minOutside;
    m3Vec3 inside;
    QhNode orphans;
    QhNode verts;
    QhNode faces;
    int32_t vertUsed;
    int32_t edgeUsed;
    int32_t faceUsed;
    QhEdge* edgeFree;
    QhFace* faceFree;
    int32_t horizonCount;
    int32_t coneCount;
    int32_t mergedCount;
    int32_t outVerts;
    int32_t outHalfEdges;
    int32_t outFaces;
    int failed; // a pool ran out: the cloud is refused
    QhVert vertPool[QH_VERT_CAP];
    QhEdge edgePool[QH_EDGE_CAP];
    QhFace facePool[QH_FACE_CAP];
    QhEdge* horizon[QH_RING_CAP];
    QhFace* cone[QH_RING_CAP];
    QhFace* merged[QH_MERGE_CAP];
    QhFrame stack[QH_MERGE_CAP];
    m3Vec3 local[M3_HULL_MAX_INPUT];
} QhWork;

static void ListInit(QhNode* head)
{
    head->prev = head;
    head->next = head;
}

static int ListHas(const QhNode* n) { return n->prev != NULL; }

static void ListUnlink(QhNode* n)
{
    n->prev->next = n->next;
    n->next->prev = n->prev;
    n->prev = NULL;
    n->next = NULL;
}

static void ListAppend(QhNode* head, QhNode* n)
{
    n->prev = head->prev;
    n->next = head;
    head->prev->next = n;
    head->prev = n;
}

static float Above(const QhFace* f, m3Vec3 p) { return m3Dot3(f->normal, p) - f->offset; }

static QhVert* VertNew(QhWork* w, m3Vec3 p)
{
    if (w->vertUsed >= QH_VERT_CAP)
    {
        w->failed = 1;
        w->vertUsed = QH_VERT_CAP - 1;
    }
    QhVert* v = &w->vertPool[w->vertUsed++];
    v->node.prev = NULL;
    v->node.next = NULL;
    v->conflict = NULL;
    v->p = p;
    v->out = -1;
    v->used = 0;
    return v;
}

static QhEdge* EdgeNew(QhWork* w)
{
    if (w->edgeFree != NULL)
    {
        QhEdge* e = w->edgeFree;
        w->edgeFree = e->freeNext;
        return e;
    }
    if (w->edgeUsed >= QH_EDGE_CAP)
    {
        w->failed = 1;
        w->edgeUsed = QH_EDGE_CAP - 1;
    }
    return &w->edgePool[w->edgeUsed++];
}

static void EdgeFree(QhWork* w, QhEdge* e)
{
    e->freeNext = w->edgeFree;
    w->edgeFree = e;
}

// Triangle a, b, c with its plane; the three edges are created in loop order.
static QhFace* FaceNew(QhWork* w, QhVert* a, QhVert* b, QhVert* c)
{
    QhFace* f;
    if (w->faceFree != NULL)
    {
        f = w->faceFree;
        w->faceFree = f->freeNext;
    }
    else
    {
        if (w->faceUsed >= QH_FACE_CAP)
        {
            w->failed = 1;
            w->faceUsed = QH_FACE_CAP - 1;
        }
        f = &w->facePool[w->faceUsed++];
    }
    f->node.prev = NULL;
    f->node.next = NULL;
    f->farthest = NULL;
    f->farthestDist = 0.0f;
    QhEdge* e1 = EdgeNew(w);
    QhEdge* e2 = EdgeNew(w);
    QhEdge* e3 = EdgeNew(w);
    m3Vec3 n = m3Cross3(m3Sub3(b->p, a->p), m3Sub3(c->p, a->p));
    float len = sqrtf(m3Dot3(n, n));
    n = len > 0.0f ? m3MulSV3(1.0f / len, n) : m3V3(0.0f, 1.0f, 0.0f);
    f->edge = e1;
    f->mark = QH_VISIBLE;
    f->area = 0.5f * len;
    f->centroid = m3MulSV3(1.0f / 3.0f, m3Add3(a->p, m3Add3(b->p, c->p)));
    f->normal = n;
    f->offset = m3Dot3(n, a->p);
    f->flipped = Above(f, w->inside) > 0.0f;
    ListInit(&f->conflicts);
    QhEdge* ring[3] = {e1, e2, e3};
    QhVert* org[3] = {a, b, c};
    for (int k = 0; k < 3; ++k)
    {
        ring[k]->prev = ring[(k + 2) % 3];
        ring[k]->next = ring[(k + 1) % 3];
        ring[k]->origin = org[k];
        ring[k]->face = f;
        ring[k]->twin = NULL;
    }
    return f;
}

static void FaceFree(QhWork* w, QhFace* f)
{
    if (ListHas(&f->node))
    {
        ListUnlink(&f->node);
    }
    f->edge = NULL;
    f->freeNext = w->faceFree;
    w->faceFree = f;
}

static int EdgeConvex(const QhEdge* e, float tol) { return Above(e->face, e->twin->face->centroid) < -tol; }

static int EdgeConcave(const QhEdge* e, float tol) { return Above(e->face, e->twin->face->centroid) > tol; }

static int32_t FaceDegree(const QhFace* f)
{
    int32_t n = 0;
    const QhEdge* e = f->edge;
    do
    {
        n += 1;
        e = e->next;
    } while (e != f->edge);
    return n;
}

static QhEdge* RingAt(const QhFace* f, int32_t k)
{
    QhEdge* e = f->edge;
    while (k-- > 0)
    {
        e = e->next;
    }
    return e;
}

static void Twin(QhEdge* a, QhEdge* b)
{
    a->twin = b;
    b->twin = a;
}

// A ring the merges left inconsistent (an edge without twin, or one that
// does not close): such clouds are refused instead of followed.
static int RingBroken(const QhFace* f)
{
    const QhEdge* e = f->edge;
    int32_t guard = 0;
    if (e == NULL)
    {
        return 1;
    }
    do
    {
        if (e->twin == NULL || e->origin == NULL || e->twin->origin == NULL || e->next == NULL ||
            ++guard > QH_EDGE_CAP)
        {
            return 1;
        }
        e = e->next;
    } while (e != f->edge);
    return 0;
}

// Newell normal and centroid of an n-gon, relative to its first vertex.
static void FacePlane(QhWork* w, QhFace* f)
{
    if (RingBroken(f))
    {
        w->failed = 1;
        return;
    }
    int32_t count = 0;
    m3Vec3 centroid = m3V3(0.0f, 0.0f, 0.0f);
    m3Vec3 n = m3V3(0.0f, 0.0f, 0.0f);
    QhEdge* e = f->edge;
    m3Vec3 o = e->origin->p;
    do
    {
        m3Vec3 p = m3Sub3(e->origin->p, o);
        m3Vec3 q = m3Sub3(e->twin->origin->p, o);
        count += 1;
        centroid = m3Add3(centroid, p);
        n.x += (p.y - q.y) * (p.z + q.z);
        n.y += (p.z - q.z) * (p.x + q.x);
        n.z += (p.x - q.x) * (p.y + q.y);
        e = e->next;
    } while (e != f->edge);
    centroid = m3Add3(m3MulSV3(1.0f / (float)count, centroid), o);
    float len = sqrtf(m3Dot3(n, n));
    n = len > 0.0f ? m3MulSV3(1.0f / len, n) : m3V3(0.0f, 1.0f, 0.0f);
    f->centroid = centroid;
    f->normal = n;
    f->offset = m3Dot3(n, centroid);
    f->area = 0.5f * len;
}

static void Tolerance(QhWork* w, int32_t n, const m3Vec3* p)
{
    m3Vec3 lo = p[0];
    m3Vec3 hi = p[0];
    for (int32_t i = 1; i < n; ++i)
    {
        lo.x = m3MinF(lo.x, p[i].x);
        lo.y = m3MinF(lo.y, p[i].y);
        lo.z = m3MinF(lo.z, p[i].z);
        hi.x = m3MaxF(hi.x, p[i].x);
        hi.y = m3MaxF(hi.y, p[i].y);
        hi.z = m3MaxF(hi.z, p[i].z);
    }
    m3Vec3 m;
    m.x = m3MaxF(lo.x < 0.0f ? -lo.x : lo.x, hi.x < 0.0f ? -hi.x : hi.x);
    m.y = m3MaxF(lo.y < 0.0f ? -lo.y : lo.y, hi.y < 0.0f ? -hi.y : hi.y);
    m.z = m3MaxF(lo.z < 0.0f ? -lo.z : lo.z, hi.z < 0.0f ? -hi.z : hi.z);
    float sum = m.x + m.y + m.z;
    float largest = m3MaxF(m.x, m3MaxF(m.y, m.z));
    float reach = m3MinF(1.7320508f * largest, sum);
    w->tolerance = (3.0f * reach * 1.01f + largest) * FLT_EPSILON;
    w->minRadius = 4.0f * w->tolerance;
    w->minOutside = 2.0f * w->minRadius;
}

// The widest cardinal extent: its lowest and highest points (first found
// wins a tie on either end).
static void SeedAxis(int32_t* lo, int32_t* hi, float tol, int32_t n, const m3Vec3* p)
{
    *lo = -1;
    *hi = -1;
    float minV[3] = {p[0].x, p[0].y, p[0].z};
    float maxV[3] = {p[0].x, p[0].y, p[0].z};
    int32_t minI[3] = {0, 0, 0};
    int32_t maxI[3] = {0, 0, 0};
    for (int32_t i = 1; i < n; ++i)
    {
        float c[3] = {p[i].x, p[i].y, p[i].z};
        for (int a = 0; a < 3; ++a)
        {
            if (c[a] < minV[a])
            {
                minV[a] = c[a];
                minI[a] = i;
            }
            else if (c[a] > maxV[a])
            {
                maxV[a] = c[a];
                maxI[a] = i;
            }
        }
    }
    float span[3] = {maxV[0] - minV[0], maxV[1] - minV[1], maxV[2] - minV[2]};
    int a = 0;
    if (span[1] > span[a])
    {
        a = 1;
    }
    if (span[2] > span[a])
    {
        a = 2;
    }
    if (span[a] > 2.0f * tol)
    {
        *lo = minI[a];
        *hi = maxI[a];
    }
}

static int32_t SeedLine(int32_t i1, int32_t i2, float tol, int32_t n, const m3Vec3* p)
{
    m3Vec3 a = p[i1];
    m3Vec3 ab = m3Sub3(p[i2], a);
    float ab2 = m3Dot3(ab, ab);
    if (!(ab2 > 0.0f))
    {
        return -1;
    }
    float inv = 1.0f / ab2;
    float best = 4.0f * tol * tol;
    int32_t pick = -1;
    for (int32_t i = 0; i < n; ++i)
    {
        if (i == i1 || i == i2)
        {
            continue;
        }
        m3Vec3 c = m3Cross3(m3Sub3(p[i], a), ab);
        float d2 = m3Dot3(c, c) * inv;
        if (d2 > best)
        {
            best = d2;
            pick = i;
        }
    }
    return pick;
}

static int32_t SeedPlane(int32_t i1, int32_t i2, int32_t i3, float tol, int32_t n, const m3Vec3* p)
{
    m3Vec3 a = p[i1];
    m3Vec3 nrm = m3Cross3(m3Sub3(p[i2], a), m3Sub3(p[i3], a));
    float len = sqrtf(m3Dot3(nrm, nrm));
    if (!(len > 0.0f))
    {
        return -1;
    }
    nrm = m3MulSV3(1.0f / len, nrm);
    float off = m3Dot3(nrm, a);
    float best = 2.0f * tol;
    int32_t pick = -1;
    for (int32_t i = 0; i < n; ++i)
    {
        if (i == i1 || i == i2 || i == i3)
        {
            continue;
        }
        float d = m3Dot3(nrm, p[i]) - off;
        d = d < 0.0f ? -d : d;
        if (d > best)
        {
            best = d;
            pick = i;
        }
    }
    return pick;
}

// Files v under face f when it lies farther out than minOutside.
static void AssignConflict(QhFace* f, QhVert* v, float dist)
{
    v->conflict = f;
    ListAppend(&f->conflicts, &v->node);
    if (dist > f->farthestDist)
    {
        f->farthestDist = dist;
        f->farthest = v;
    }
}

static int Seed(QhWork* w, int32_t n, const m3Vec3* p)
{
    int32_t i1;
    int32_t i2;
    SeedAxis(&i1, &i2, w->tolerance, n, p);
    if (i1 < 0 || i2 < 0)
    {
        return 0;
    }
    int32_t i3 = SeedLine(i1, i2, w->tolerance, n, p);
    if (i3 < 0)
    {
        return 0;
    }
    int32_t i4 = SeedPlane(i1, i2, i3, w->tolerance, n, p);
    if (i4 < 0)
    {
        return 0;
    }
    m3Vec3 d1 = m3Sub3(p[i1], p[i4]);
    m3Vec3 d2 = m3Sub3(p[i2], p[i4]);
    m3Vec3 d3 = m3Sub3(p[i3], p[i4]);
    if (m3Dot3(d1, m3Cross3(d2, d3)) < 0.0f)
    {
        int32_t t = i2;
        i2 = i3;
        i3 = t;
    }
    w->inside = m3MulSV3(0.25f, m3Add3(m3Add3(p[i1], p[i2]), m3Add3(p[i3], p[i4])));
    QhVert* v1 = VertNew(w, p[i1]);
    ListAppend(&w->verts, &v1->node);
    QhVert* v2 = VertNew(w, p[i2]);
    ListAppend(&w->verts, &v2->node);
    QhVert* v3 = VertNew(w, p[i3]);
    ListAppend(&w->verts, &v3->node);
    QhVert* v4 = VertNew(w, p[i4]);
    ListAppend(&w->verts, &v4->node);
    QhFace* f1 = FaceNew(w, v1, v2, v3);
    ListAppend(&w->faces, &f1->node);
    QhFace* f2 = FaceNew(w, v4, v2, v1);
    ListAppend(&w->faces, &f2->node);
    QhFace* f3 = FaceNew(w, v4, v3, v2);
    ListAppend(&w->faces, &f3->node);
    QhFace* f4 = FaceNew(w, v4, v1, v3);
    ListAppend(&w->faces, &f4->node);
    Twin(RingAt(f1, 0), RingAt(f2, 1));
    Twin(RingAt(f1, 1), RingAt(f3, 1));
    Twin(RingAt(f1, 2), RingAt(f4, 1));
    Twin(RingAt(f2, 0), RingAt(f3, 2));
    Twin(RingAt(f3, 0), RingAt(f4, 2));
    Twin(RingAt(f4, 0), RingAt(f2, 2));
    for (int32_t i = 0; i < n; ++i)
    {
        if (i == i1 || i == i2 || i == i3 || i == i4)
        {
            continue;
        }
        float best = w->minOutside;
        QhFace* bestFace = NULL;
        for (QhNode* it = w->faces.next; it != &w->faces; it = it->next)
        {
            QhFace* f = (QhFace*)it;
            float d = Above(f, p[i]);
            if (d > best)
            {
                best = d;
                bestFace = f;
            }
        }
        if (bestFace != NULL)
        {
            AssignConflict(bestFace, VertNew(w, p[i]), best);
        }
    }
    return 1;
}

// Re-find a face's farthest conflict after its plane moved.
static void Refarthest(QhFace* f, float minOutside)
{
    QhVert* best = NULL;
    float bestDist = minOutside;
    for (QhNode* it = f->conflicts.next; it != &f->conflicts; it = it->next)
    {
        QhVert* v = (QhVert*)it;
        float d = Above(f, v->p);
        if (d > bestDist)
        {
            bestDist = d;
            best = v;
        }
    }
    f->farthest = best;
    f->farthestDist = bestDist;
}

static QhVert* NextApex(const QhWork* w)
{
    QhVert* best = NULL;
    float bestDist = w->minOutside;
    for (const QhNode* it = w->faces.next; it != &w->faces; it = it->next)
    {
        const QhFace* f = (const QhFace*)it;
        if (f->farthest != NULL && f->farthestDist > bestDist)
        {
            bestDist = f->farthestDist;
            best = f->farthest;
        }
    }
    return best;
}

static void Orphan(QhWork* w, QhFace* f)
{
    QhNode* it = f->conflicts.next;
    while (it != &f->conflicts)
    {
        QhVert* v = (QhVert*)it;
        it = it->next;
        v->conflict = NULL;
        ListUnlink(&v->node);
        ListAppend(&w->orphans, &v->node);
    }
}

static void VisitFace(QhWork* w, QhFace* f, QhEdge* entry, QhFrame* fr)
{
    f->mark = QH_DELETED;
    Orphan(w, f);
    fr->face = f;
    fr->started = 0;
    fr->first = entry != NULL ? entry : f->edge;
    fr->edge = entry != NULL ? entry->next : f->edge;
}

// Depth-first walk over the faces the apex sees; the horizon is the ring
// of edges whose far face it does not see, in walk order.
static void Horizon(QhWork* w, QhVert* apex, QhFace* seed)
{
    int32_t top = 0;
    VisitFace(w, seed, NULL, &w->stack[top++]);
    while (top > 0)
    {
        QhFrame* fr = &w->stack[top - 1];
        if (fr->started && fr->edge == fr->first)
        {
            top -= 1;
            continue;
        }
        fr->started = 1;
        QhEdge* e = fr->edge;
        QhEdge* t = e->twin;
        fr->edge = e->next;
        if (t->face->mark != QH_VISIBLE)
        {
            continue;
        }
        if (Above(t->face, apex->p) > w->minRadius)
        {
            if (top >= QH_MERGE_CAP)
            {
                w->failed = 1;
                return;
            }
            VisitFace(w, t->face, t, &w->stack[top++]);
        }
        else
        {
            if (w->horizonCount >= QH_RING_CAP)
            {
                w->failed = 1;
                return;
            }
            w->horizon[w->horizonCount++] = e;
        }
    }
}

static void Cone(QhWork* w, QhVert* apex)
{
    for (int32_t i = 0; i < w->horizonCount; ++i)
    {
        QhEdge* e = w->horizon[i];
        QhFace* f = FaceNew(w, apex, e->origin, e->twin->origin);
        w->cone[w->coneCount++] = f;
        Twin(RingAt(f, 1), e->twin);
    }
    QhFace* prev = w->cone[w->coneCount - 1];
    for (int32_t i = 0; i < w->coneCount; ++i)
    {
        QhFace* f = w->cone[i];
        Twin(RingAt(prev, 2), RingAt(f, 0));
        prev = f;
    }
}

static void FreeRun(QhWork* w, QhEdge* from, QhEdge* to)
{
    QhEdge* e = from;
    while (e != to)
    {
        QhEdge* next = e->next;
        EdgeFree(w, e);
        e = next;
    }
}

// Joins `a` to `b` in a ring.  When both face the same neighbour the
// shared vertex would dangle: a triangular neighbour dissolves, a larger
// one loses the redundant edge.
static void Splice(QhWork* w, QhEdge* a, QhEdge* b)
{
    if (a->twin->face != b->twin->face)
    {
        a->next = b;
        b->prev = a;
        return;
    }
    if (b->face->edge == b)
    {
        b->face->edge = a;
    }
    QhEdge* keep;
    if (FaceDegree(a->twin->face) == 3)
    {
        QhEdge* dead0 = a->twin;
        QhEdge* dead1 = b->twin;
        QhEdge* dead2 = b->twin->prev;
        keep = dead2->twin;
        QhFace* gone = a->twin->face;
        gone->mark = QH_DELETED;
        w->merged[w->mergedCount++] = gone;
        a->next = b->next;
        a->next->prev = a;
        Twin(a, keep);
        ListUnlink(&b->origin->node);
        EdgeFree(w, dead0);
        EdgeFree(w, dead1);
        EdgeFree(w, dead2);
    }
    else
    {
        keep = b->twin;
        if (keep->face->edge == a->twin)
        {
            keep->face->edge = keep;
        }
        keep->next = a->twin->next;
        keep->next->prev = keep;
        EdgeFree(w, a->twin);
        a->next = b->next;
        a->next->prev = a;
        Twin(a, keep);
        ListUnlink(&b->origin->node);
    }
    FacePlane(w, keep->face);
    if (w->failed)
    {
        return;
    }
    Refarthest(keep->face, w->minOutside);
}

// Conflicts of the absorbed faces move to f or become orphans.
static void Absorb(QhWork* w, QhFace* f)
{
    for (int32_t i = 0; i < w->mergedCount; ++i)
    {
        QhNode* head = &w->merged[i]->conflicts;
        QhNode* it = head->next;
        while (it != head)
        {
            QhVert* v = (QhVert*)it;
            it = it->next;
            ListUnlink(&v->node);
            float d = Above(f, v->p);
            if (d > w->minOutside)
            {
                AssignConflict(f, v, d);
            }
            else
            {
                ListAppend(&w->orphans, &v->node);
                v->conflict = NULL;
            }
        }
        FaceFree(w, w->merged[i]);
    }
}

// Merges the face across e into e's face.
static void Merge(QhWork* w, QhEdge* e)
{
    if (w->failed || RingBroken(e->face) || RingBroken(e->twin->face))
    {
        w->failed = 1;
        return;
    }
    QhFace* f = e->face;
    QhEdge* t = e->twin;
    QhEdge* ePrev = e->prev;
    QhEdge* eNext = e->next;
    QhEdge* tPrev = t->prev;
    QhEdge* tNext = t->next;
    while (ePrev->twin->face == t->face)
    {
        ePrev = ePrev->prev;
        tNext = tNext->next;
    }
    while (eNext->twin->face == t->face)
    {
        eNext = eNext->next;
        tPrev = tPrev->prev;
    }
    f->edge = ePrev;
    w->mergedCount = 0;
    w->merged[w->mergedCount++] = t->face;
    t->face->mark = QH_DELETED;
    t->face->edge = NULL;
    for (QhEdge* a = tNext; a != tPrev->next; a = a->next)
    {
        a->face = f;
    }
    FreeRun(w, ePrev->next, eNext);
    FreeRun(w, tPrev->next, tNext);
    Splice(w, ePrev, tNext);
    if (w->failed)
    {
        return;
    }
    Splice(w, tPrev, eNext);
    if (w->failed)
    {
        return;
    }
    FacePlane(w, f);
    if (w->failed)
    {
        return;
    }
    Refarthest(f, w->minOutside);
    Absorb(w, f);
}

static int MergeConcave(QhWork* w, QhFace* f)
{
    if (w->failed || RingBroken(f))
    {
        w->failed = 1;
        return 0;
    }
    QhEdge* e = f->edge;
    do
    {
        if (EdgeConcave(e, w->minRadius) || EdgeConcave(e->twin, w->minRadius))
        {
            Merge(w, e);
            return 1;
        }
        e = e->next;
    } while (e != f->edge);
    return 0;
}

static int MergeCoplanar(QhWork* w, QhFace* f)
{
    if (w->failed || RingBroken(f))
    {
        w->failed = 1;
        return 0;
    }
    QhEdge* e = f->edge;
    do
    {
        if (!EdgeConvex(e, w->minRadius) || !EdgeConvex(e->twin, w->minRadius))
        {
            Merge(w, e);
            return 1;
        }
        e = e->next;
    } while (e != f->edge);
    return 0;
}

static void MergeCone(QhWork* w)
{
    // A cone face that sees the interior point merges into its largest
    // neighbour.
    for (int32_t i = 0; i < w->coneCount; ++i)
    {
        QhFace* f = w->cone[i];
        if (f->mark != QH_VISIBLE || !f->flipped)
        {
            continue;
        }
        f->flipped = 0;
        float bestArea = 0.0f;
        QhEdge* best = NULL;
        QhEdge* e = f->edge;
        do
        {
            float a = e->twin->face->area;
            if (a > bestArea)
            {
                bestArea = a;
                best = e;
            }
            e = e->next;
        } while (e != f->edge);
        if (best == NULL)
        {
            w->failed = 1;
            return;
        }
        Merge(w, best);
        if (w->failed)
        {
            return;
        }
    }
    for (int32_t i = 0; i < w->coneCount; ++i)
    {
        if (w->cone[i]->mark == QH_VISIBLE)
        {
            while (MergeConcave(w, w->cone[i]))
            {
            }
        }
    }
    for (int32_t i = 0; i < w->coneCount; ++i)
    {
        if (w->cone[i]->mark == QH_VISIBLE)
        {
            while (MergeCoplanar(w, w->cone[i]))
            {
            }
        }
    }
}

// Orphans go to the cone face they are farthest outside, or are dropped.
static void Rehome(QhWork* w)
{
    QhNode* it = w->orphans.next;
    while (it != &w->orphans)
    {
        QhVert* v = (QhVert*)it;
        it = it->next;
        ListUnlink(&v->node);
        float best = w->minOutside;
        QhFace* bestFace = NULL;
        for (int32_t i = 0; i < w->coneCount; ++i)
        {
            if (w->cone[i]->mark == QH_VISIBLE)
            {
                float d = Above(w->cone[i], v->p);
                if (d > best)
                {
                    best = d;
                    bestFace = w->cone[i];
                }
            }
        }
        if (bestFace != NULL)
        {
            AssignConflict(bestFace, v, best);
        }
    }
}

// Deleted faces leave the face list; surviving cone faces join its end.
static void Refile(QhWork* w)
{
    QhNode* it = w->faces.next;
    while (it != &w->faces)
    {
        QhFace* f = (QhFace*)it;
        it = it->next;
        if (f->mark == QH_DELETED && ListHas(&f->node))
        {
            ListUnlink(&f->node);
        }
    }
    for (int32_t i = 0; i < w->coneCount; ++i)
    {
        if (w->cone[i]->mark != QH_DELETED)
        {
            ListAppend(&w->faces, &w->cone[i]->node);
        }
    }
}

static void AddApex(QhWork* w, QhVert* apex)
{
    QhFace* f = apex->conflict;
    apex->conflict = NULL;
    ListUnlink(&apex->node);
    ListAppend(&w->verts, &apex->node);
    w->horizonCount = 0;
    Horizon(w, apex, f);
    if (w->failed || w->horizonCount == 0)
    {
        w->failed = 1;
        return;
    }
    w->coneCount = 0;
    Cone(w, apex);
    MergeCone(w);
    if (w->failed)
    {
        return;
    }
    Rehome(w);
    Refile(w);
}

// Drops vertices no face uses; counts what is left.
static void Tidy(QhWork* w)
{
    int32_t faces = 0;
    int32_t halfEdges = 0;
    for (QhNode* it = w->faces.next; it != &w->faces; it = it->next)
    {
        QhFace* f = (QhFace*)it;
        if (RingBroken(f))
        {
            w->failed = 1;
            return;
        }
        QhEdge* e = f->edge;
        do
        {
            e->origin->used = 1;
            e = e->next;
            halfEdges += 1;
        } while (e != f->edge);
        faces += 1;
    }
    int32_t verts = 0;
    QhNode* it = w->verts.next;
    while (it != &w->verts)
    {
        QhVert* v = (QhVert*)it;
        it = it->next;
        if (!v->used)
        {
            ListUnlink(&v->node);
        }
        else
        {
            verts += 1;
        }
    }
    w->outVerts = verts;
    w->outHalfEdges = halfEdges;
    w->outFaces = faces;
}

// Unit-density volume, centroid and central inertia: each face fans into
// triangles from its first vertex, each triangle with the hull's first
// vertex forms a tetrahedron (src/quickhull.c:1041-1108).
static int MassProperties(m3HullData* h)
{
    m3Vec3 o = h->vertices[0];
    float volume = 0.0f;
    m3Vec3 center = m3V3(0.0f, 0.0f, 0.0f);
    float xx = 0.0f;
    float yy = 0.0f;
    float zz = 0.0f;
    float xy = 0.0f;
    float xz = 0.0f;
    float yz = 0.0f;
    for (int32_t f = 0; f < h->faceCount; ++f)
    {
        int32_t n = h->faceVertCounts[f];
        int32_t s = h->faceVertStart[f];
        m3Vec3 v1 = m3Sub3(h->vertices[h->faceIndices[s]], o);
        for (int32_t k = 1; k + 1 < n; ++k)
        {
            m3Vec3 v2 = m3Sub3(h->vertices[h->faceIndices[s + k]], o);
            m3Vec3 v3 = m3Sub3(h->vertices[h->faceIndices[s + k + 1]], o);
            float det = m3Dot3(v1, m3Cross3(v2, v3));
            volume += det;
            m3Vec3 v4 = m3Add3(v1, m3Add3(v2, v3));
            center = m3Add3(center, m3MulSV3(det, v4));
            xx += det * (v1.x * v1.x + v2.x * v2.x + v3.x * v3.x + v4.x * v4.x);
            yy += det * (v1.y * v1.y + v2.y * v2.y + v3.y * v3.y + v4.y * v4.y);
            zz += det * (v1.z * v1.z + v2.z * v2.z + v3.z * v3.z + v4.z * v4.z);
            xy += det * (v1.x * v1.y + v2.x * v2.y + v3.x * v3.y + v4.x * v4.y);
            xz += det * (v1.x * v1.z + v2.x * v2.z + v3.x * v3.z + v4.x * v4.z);
            yz += det * (v1.y * v1.z + v2.y * v2.z + v3.y * v3.z + v4.y * v4.z);
        }
    }
    if (!(volume > 0.0f))
    {
        return 0;
    }
    m3Vec3 c = m3MulSV3(0.25f / volume, center);
    float mass = volume / 6.0f;
    m3Mat3 I;
    I.cx = m3V3(yy + zz, -xy, -xz);
    I.cy = m3V3(-xy, xx + zz, -yz);
    I.cz = m3V3(-xz, -yz, xx + yy);
    float scale = 1.0f / 120.0f;
    I.cx = m3MulSV3(scale, I.cx);
    I.cy = m3MulSV3(scale, I.cy);
    I.cz = m3MulSV3(scale, I.cz);
    float c2 = m3Dot3(c, c);
    I.cx.x -= mass * (c2 - c.x * c.x);
    I.cy.y -= mass * (c2 - c.y * c.y);
    I.cz.z -= mass * (c2 - c.z * c.z);
    I.cy.x += mass * c.x * c.y;
    I.cx.y += mass * c.x * c.y;
    I.cz.x += mass * c.x * c.z;
    I.cx.z += mass * c.x * c.z;
    I.cz.y += mass * c.y * c.z;
    I.cy.z += mass * c.y * c.z;
    h->unitMass = mass;
    h->unitCom = m3Add3(c, o);
    h->unitInertiaCom = I;
    return 1;
}

// Half edges from the face loops: each undirected edge takes the next
// twin pair (2k, 2k+1) when first met, its reverse takes the other slot;
// then next pointers follow the loops and the vertex centroid is stored
// (src/hull.c:77-137).
void m3HostHullHalfEdges(m3HullData* hull)
{
    int32_t count = 0;
    int32_t dirKey[M3_HULL_MAX_HALF_EDGES];
    int32_t dirSlot[M3_HULL_MAX_HALF_EDGES];
    int32_t dirCount = 0;
    int32_t faceEdgeSlot[M3_HULL_MAX_FACE_INDICES];
    for (int32_t f = 0; f < hull->faceCount; ++f)
    {
        int32_t n = hull->faceVertCounts[f];
        int32_t start = hull->faceVertStart[f];
        for (int32_t k = 0; k < n; ++k)
        {
            int32_t a = hull->faceIndices[start + k];
            int32_t b = hull->faceIndices[start + (k + 1) % n];
            int32_t reverseKey = b * 256 + a;
            int32_t slot = -1;
            for (int32_t d = 0; d < dirCount; ++d)
            {
                if (dirKey[d] == reverseKey)
                {
                    slot = dirSlot[d] ^ 1;
                    break;
                }
            }
            if (slot < 0)
            {
                slot = count;
                count += 2;
                dirKey[dirCount] = a * 256 + b;
                dirSlot[dirCount] = slot;
                dirCount += 1;
            }
            hull->edges[slot].origin = (uint8_t)a;
            hull->edges[slot].twin = (uint16_t)(slot ^ 1);
            hull->edges[slot].face = (uint8_t)f;
            faceEdgeSlot[start + k] = slot;
        }
    }
    for (int32_t f = 0; f < hull->faceCount; ++f)
    {
        int32_t n = hull->faceVertCounts[f];
        int32_t start = hull->faceVertStart[f];
        for (int32_t k = 0; k < n; ++k)
        {
            hull->edges[faceEdgeSlot[start + k]].next = (uint16_t)faceEdgeSlot[start + (k + 1) % n];
        }
    }
    hull->edgeCount = count;
    m3Vec3 c = m3V3(0.0f, 0.0f, 0.0f);
    for (int32_t v = 0; v < hull->vertexCount; ++v)
    {
        c = m3Add3(c, hull->vertices[v]);
    }
    hull->center = m3MulSV3(1.0f / (m3real)hull->vertexCount, c);
}

int m3HostComputeHull(const m3Vec3* points, int32_t count, m3HullData* out)
{
    if (points == NULL || out == NULL || count < 4 || count > M3_HULL_MAX_INPUT)
    {
        return 0;
    }
    for (int32_t i = 0; i < count; ++i)
    {
        const m3Vec3 q = points[i];
        if (!(q.x == q.x) || !(q.y == q.y) || !(q.z == q.z) || q.x > 1.0e18f || q.x < -1.0e18f ||
            q.y > 1.0e18f || q.y < -1.0e18f || q.z > 1.0e18f || q.z < -1.0e18f)
        {
            return 0;
        }
    }
    QhWork* w = (QhWork*)calloc(1, sizeof(QhWork));
    if (w == NULL)
    {
        return 0;
    }
    ListInit(&w->orphans);
    ListInit(&w->verts);
    ListInit(&w->faces);
    // Work relative to the first point to limit round-off.
    m3Vec3 origin = points[0];
    for (int32_t i = 0; i < count; ++i)
    {
        w->local[i] = m3Sub3(points[i], origin);
    }
    Tolerance(w, count, w->local);
    int ok = Seed(w, count, w->local) && !w->failed;
    int32_t budget = QH_MAX_VERTS - 4;
    QhVert* apex = ok ? NextApex(w) : NULL;
    while (ok && apex != NULL && budget > 0)
    {
        AddApex(w, apex);
        if (w->failed)
        {
            ok = 0;
            break;
        }
        apex = NextApex(w);
        budget -= 1;
    }
    if (ok)
    {
        Tidy(w);
    }
    if (ok && !w->failed)
    {
        int32_t v = w->outVerts;
        int32_t e = w->outHalfEdges / 2;
        int32_t f = w->outFaces;
        ok = v - e + f == 2 && f >= 4 && v <= QH_MAX_VERTS && f <= QH_MAX_FACES &&
             w->outHalfEdges <= QH_MAX_INDICES;
    }
    else
    {
        ok = 0;
    }
    if (!ok)
    {
        free(w);
        return 0;
    }
    memset(out, 0, sizeof(*out));
    int32_t nv = 0;
    for (QhNode* it = w->verts.next; it != &w->verts; it = it->next)
    {
        QhVert* v = (QhVert*)it;
        v->out = nv;
        out->vertices[nv] = m3Add3(v->p, origin);
        nv += 1;
    }
    int32_t nf = 0;
    int32_t ni = 0;
    for (QhNode* it = w->faces.next; it != &w->faces; it = it->next)
    {
        QhFace* f = (QhFace*)it;
        out->faceNormals[nf] = f->normal;
        out->faceOffsets[nf] = f->offset + m3Dot3(f->normal, origin);
        out->faceVertStart[nf] = (uint16_t)ni;
        int32_t n = 0;
        QhEdge* e = f->edge;
        do
        {
            out->faceIndices[ni++] = (uint8_t)e->origin->out;
            n += 1;
            e = e->next;
        } while (e != f->edge);
        out->faceVertCounts[nf] = (uint8_t)n;
        nf += 1;
    }
    out->vertexCount = nv;
    out->faceCount = nf;
    out->indexCount = ni;
    free(w);
    if (!MassProperties(out))
    {
        memset(out, 0, sizeof(*out));
        return 0;
    }
    m3HostHullHalfEdges(out);
    return 1;
}
