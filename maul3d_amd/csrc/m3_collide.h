// Narrowphase for the device: per-pair manifold generation in the
// reference's arithmetic order (bit-exact with the CPU engine under
// -ffp-contract=off).  Each function names the reference routine whose
// operation sequence it reproduces.  All of them are pure functions of their
// arguments, so one GPU thread evaluates one pair.
#ifndef M3_COLLIDE_H
#define M3_COLLIDE_H

#include "m3_core.h"

#include <float.h>

// Clipped incident polygon bound: the incident face has at most
// M3_HULL_MAX_VERTS corners and each side plane adds at most one vertex.
#define M3_CLIP_MAX (2 * M3_HULL_MAX_VERTS)

M3_HD m3Manifold m3ZeroManifold(void)
{
    m3Manifold m;
    memset(&m, 0, sizeof(m));
    return m;
}

// src/manifold.c:41-62
M3_HD m3Manifold m3CollideSpheres(m3Vec3 d, m3real radiusA, m3real radiusB)
{
    m3Manifold m = m3ZeroManifold();
    m3real distance = m3Length3(d);
    m3real separation = distance - radiusA - radiusB;
    if (separation > M3_SPECULATIVE_DISTANCE)
    {
        return m;
    }
    m3Vec3 normal = m3Normalize3(d);
    m.normal = normal;
    m.pointCount = 1;
    m.points[0].anchorA = m3MulSV3(radiusA, normal);
    m.points[0].anchorB = m3MulSV3(-radiusB, normal);
    m.points[0].separation = separation;
    m.points[0].id = 0;
    return m;
}

// src/manifold.c:64-85
M3_HD m3Manifold m3CollidePlaneSphere(m3Vec3 planeNormal, m3real dist, m3real radius)
{
    m3Manifold m = m3ZeroManifold();
    m3real separation = dist - radius;
    if (separation > M3_SPECULATIVE_DISTANCE)
    {
        return m;
    }
    m.normal = planeNormal;
    m.pointCount = 1;
    m.points[0].anchorB = m3MulSV3(-radius, planeNormal);
    m.points[0].separation = separation;
    m.points[0].id = 0;
    return m;
}

// ---- hull versus hull SAT (src/manifold.c:96-437) --------------------------

typedef struct m3FaceQuery
{
    m3real separation;
    int32_t faceIndex;
} m3FaceQuery;

// Deepest support of `other` (moved into ref's frame by q,p or its inverse)
// against every face plane of `ref`, with the reference's early exits.
M3_HD m3FaceQuery m3QueryFaces(const m3HullData* ref, const m3HullData* other, m3Quat q, m3Vec3 p,
                               int refIsA)
{
    m3FaceQuery query;
    query.separation = -3.4e38f;
    query.faceIndex = 0;
    m3Vec3 w[M3_HULL_MAX_VERTS];
    int32_t nv = other->vertexCount;
    for (int32_t v = 0; v < nv; ++v)
    {
        w[v] = refIsA ? m3Add3(m3RotateVec3(q, other->vertices[v]), p)
                      : m3InvRotateVec3(q, m3Sub3(other->vertices[v], p));
    }
    for (int32_t f = 0; f < ref->faceCount; ++f)
    {
        m3Vec3 n = ref->faceNormals[f];
        m3real off = ref->faceOffsets[f];
        m3real best = 3.4e38f;
        for (int32_t v = 0; v < nv; ++v)
        {
            m3real d = m3Dot3(n, w[v]) - off;
            best = m3MinF(best, d);
            if (best <= query.separation)
            {
                break;
            }
        }
        if (best > query.separation)
        {
            query.separation = best;
            query.faceIndex = f;
            if (best > M3_SPECULATIVE_DISTANCE)
            {
                return query;
            }
        }
    }
    return query;
}

typedef struct m3EdgeQuery
{
    m3real separation;
    int32_t indexA;
    int32_t indexB;
    m3Vec3 axis;
} m3EdgeQuery;

// Gauss-map pruned edge pairs (src/manifold.c:150-220).
M3_HD m3EdgeQuery m3QueryEdges(const m3HullData* hullA, const m3HullData* hullB, m3Quat q, m3Vec3 p)
{
    m3EdgeQuery query;
    query.separation = -3.4e38f;
    query.indexA = -1;
    query.indexB = -1;
    query.axis = m3V3(0.0f, 1.0f, 0.0f);
    for (int32_t ib = 0; ib < hullB->edgeCount; ib += 2)
    {
        m3HullHalfEdge edgeB = hullB->edges[ib];
        m3HullHalfEdge twinB = hullB->edges[ib + 1];
        m3Vec3 qB = m3Add3(m3RotateVec3(q, hullB->vertices[twinB.origin]), p);
        m3Vec3 pB = m3Add3(m3RotateVec3(q, hullB->vertices[edgeB.origin]), p);
        m3Vec3 eB = m3Sub3(qB, pB);
        m3Vec3 uB = m3RotateVec3(q, hullB->faceNormals[edgeB.face]);
        m3Vec3 vB = m3RotateVec3(q, hullB->faceNormals[twinB.face]);
        for (int32_t ia = 0; ia < hullA->edgeCount; ia += 2)
        {
            m3HullHalfEdge edgeA = hullA->edges[ia];
            m3HullHalfEdge twinA = hullA->edges[ia + 1];
            m3Vec3 pA = hullA->vertices[edgeA.origin];
            m3Vec3 qA = hullA->vertices[twinA.origin];
            m3Vec3 eA = m3Sub3(qA, pA);
            m3Vec3 uA = hullA->faceNormals[edgeA.face];
            m3Vec3 vA = hullA->faceNormals[twinA.face];
            m3real cba = m3Dot3(uB, eA);
            m3real dba = m3Dot3(vB, eA);
            m3real adc = -m3Dot3(uA, eB);
            m3real bdc = -m3Dot3(vA, eB);
            if (!(cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f))
            {
                continue;
            }
            m3Vec3 axis = m3Cross3(eA, eB);
            m3real len = m3Length3(axis);
            if (len < 1.0e-6f)
            {
                continue;
            }
            axis = m3MulSV3(1.0f / len, axis);
            if (m3Dot3(axis, m3Sub3(pA, hullA->center)) < 0.0f)
            {
                axis = m3Neg3(axis);
            }
            m3real separation = m3Dot3(axis, m3Sub3(pB, pA));
            if (separation > query.separation)
            {
                query.separation = separation;
                query.indexA = ia;
                query.indexB = ib;
                query.axis = axis;
                if (separation > M3_SPECULATIVE_DISTANCE)
                {
                    return query;
                }
            }
        }
    }
    return query;
}

// Closest points between two segments (src/manifold.c:222-238).
M3_HD void m3SegmentClosest(m3Vec3 p1, m3Vec3 d1, m3Vec3 p2, m3Vec3 d2, m3Vec3* c1, m3Vec3* c2)
{
    m3Vec3 r = m3Sub3(p1, p2);
    m3real a = m3Dot3(d1, d1);
    m3real e = m3Dot3(d2, d2);
    m3real f = m3Dot3(d2, r);
    m3real c = m3Dot3(d1, r);
    m3real b = m3Dot3(d1, d2);
    m3real denom = a * e - b * b;
    m3real s = denom > 1.0e-9f ? m3ClampF((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
    m3real t = e > 1.0e-9f ? m3ClampF((b * s + f) / e, 0.0f, 1.0f) : 0.0f;
    s = a > 1.0e-9f ? m3ClampF((b * t - c) / a, 0.0f, 1.0f) : 0.0f;
    *c1 = m3Add3(p1, m3MulSV3(s, d1));
    *c2 = m3Add3(p2, m3MulSV3(t, d2));
}

// The three separating-axis queries of a hull pair, computed ahead of the
// manifold by a wavefront when the pair is wide (m3WideHullPair).
typedef struct m3SatQuery
{
    m3FaceQuery faceA;
    m3FaceQuery faceB;
    m3EdgeQuery edge;
} m3SatQuery;

// Pairs whose edge query is larger than box against box (12 x 12 edge
// pairs), or at least that large when `boxes` is set, take the wavefront
// query: its cost grows with the product of the edge counts (up to 186 x 186
// for two 64-vertex hulls).  One large world sends box pairs there too; a
// batch of small worlds keeps them one thread per pair.
M3_HD int m3WideHullPair(const m3HullData* hullA, const m3HullData* hullB, int boxes)
{
    int32_t product = (hullA->edgeCount >> 1) * (hullB->edgeCount >> 1);
    return boxes ? product >= 144 : product > 144;
}

// Hull versus hull in A's frame; B arrives through (q, p).  Face queries
// both ways, the edge query, then an edge contact or a Sutherland-Hodgman
// clip of the incident face with the deepest-four reduction.  `sat`, when
// given, holds the three queries already evaluated (they are pure
// functions of the hulls and (q, p), so evaluating them ahead and all three
// changes nothing but the work done).
M3_HD m3Manifold m3CollideHullsQ(const m3HullData* hullA, const m3HullData* hullB, m3Quat q,
                                 m3Vec3 p, const m3SatQuery* sat)
{
    m3Manifold manifold = m3ZeroManifold();
    m3FaceQuery faceA = sat != NULL ? sat->faceA : m3QueryFaces(hullA, hullB, q, p, 1);
    if (faceA.separation > M3_SPECULATIVE_DISTANCE)
    {
        return manifold;
    }
    m3FaceQuery faceB = sat != NULL ? sat->faceB : m3QueryFaces(hullB, hullA, q, p, 0);
    if (faceB.separation > M3_SPECULATIVE_DISTANCE)
    {
        return manifold;
    }
    m3EdgeQuery edge = sat != NULL ? sat->edge : m3QueryEdges(hullA, hullB, q, p);
    if (edge.indexA >= 0 && edge.separation > M3_SPECULATIVE_DISTANCE)
    {
        return manifold;
    }
    const m3real linearSlop = M3_LINEAR_SLOP;
    m3real maxFace = m3MaxF(faceA.separation, faceB.separation);
    if (edge.indexA >= 0 && edge.separation > maxFace + 0.1f * linearSlop)
    {
        m3HullHalfEdge eA = hullA->edges[edge.indexA];
        m3HullHalfEdge tA = hullA->edges[edge.indexA + 1];
        m3HullHalfEdge eB = hullB->edges[edge.indexB];
        m3HullHalfEdge tB = hullB->edges[edge.indexB + 1];
        m3Vec3 pA = hullA->vertices[eA.origin];
        m3Vec3 dA = m3Sub3(hullA->vertices[tA.origin], pA);
        m3Vec3 pB = m3Add3(m3RotateVec3(q, hullB->vertices[eB.origin]), p);
        m3Vec3 dB = m3Sub3(m3Add3(m3RotateVec3(q, hullB->vertices[tB.origin]), p), pB);
        m3Vec3 cA;
        m3Vec3 cB;
        m3SegmentClosest(pA, dA, pB, dB, &cA, &cB);
        manifold.normal = edge.axis;
        manifold.pointCount = 1;
        manifold.points[0].anchorA = cA;
        manifold.points[0].anchorB = cB;
        manifold.points[0].separation = edge.separation;
        manifold.points[0].id = (uint16_t)(0x8000u | (((uint32_t)edge.indexA >> 1) << 7) |
                                           ((uint32_t)edge.indexB >> 1));
        return manifold;
    }

    int refIsA = faceB.separation <= faceA.separation + 0.1f * linearSlop ? 1 : 0;
    const m3HullData* ref = refIsA ? hullA : hullB;
    const m3HullData* inc = refIsA ? hullB : hullA;
    int32_t refFace = refIsA ? faceA.faceIndex : faceB.faceIndex;
    m3Vec3 refN = ref->faceNormals[refFace];
    m3real refOff = ref->faceOffsets[refFace];

    m3Vec3 refNInc = refIsA ? m3InvRotateVec3(q, refN) : m3RotateVec3(q, refN);
    int32_t incFace = 0;
    m3real minDot = 3.4e38f;
    for (int32_t f = 0; f < inc->faceCount; ++f)
    {
        m3real d = m3Dot3(refNInc, inc->faceNormals[f]);
        if (d < minDot)
        {
            minDot = d;
            incFace = f;
        }
    }

    m3Vec3 poly[M3_CLIP_MAX];
    uint16_t polyId[M3_CLIP_MAX];
    m3Vec3 outPoly[M3_CLIP_MAX];
    uint16_t outId[M3_CLIP_MAX];
    int32_t polyCount = inc->faceVertCounts[incFace];
    for (int32_t k = 0; k < polyCount; ++k)
    {
        uint8_t vi = inc->faceIndices[inc->faceVertStart[incFace] + k];
        m3Vec3 v = inc->vertices[vi];
        poly[k] = refIsA ? m3Add3(m3RotateVec3(q, v), p) : m3InvRotateVec3(q, m3Sub3(v, p));
        polyId[k] = vi;
    }

    int32_t n = ref->faceVertCounts[refFace];
    int32_t start = ref->faceVertStart[refFace];
    for (int32_t e = 0; e < n && polyCount > 0; ++e)
    {
        m3Vec3 v1 = ref->vertices[ref->faceIndices[start + e]];
        m3Vec3 v2 = ref->vertices[ref->faceIndices[start + (e + 1) % n]];
        m3Vec3 tangent = m3Normalize3(m3Sub3(v2, v1));
        m3Vec3 sideN = m3Cross3(tangent, refN);
        m3real sideOff = m3Dot3(sideN, v1);
        int32_t outCount = 0;
        for (int32_t k = 0; k < polyCount; ++k)
        {
            m3Vec3 cur = poly[k];
            m3Vec3 nxt = poly[(k + 1) % polyCount];
            m3real dc = m3Dot3(sideN, cur) - sideOff;
            m3real dn = m3Dot3(sideN, nxt) - sideOff;
            if (dc <= 0.0f)
            {
                outPoly[outCount] = cur;
                outId[outCount] = polyId[k];
                outCount += 1;
            }
            if (dc * dn < 0.0f)
            {
                m3real t = dc / (dc - dn);
                outPoly[outCount] = m3Add3(m3MulSV3(1.0f - t, cur), m3MulSV3(t, nxt));
                outId[outCount] = (uint16_t)(0x4000u | ((uint32_t)e << 8) | polyId[k]);
                outCount += 1;
            }
        }
        for (int32_t k = 0; k < outCount; ++k)
        {
            poly[k] = outPoly[k];
            polyId[k] = outId[k];
        }
        polyCount = outCount;
    }

    // Keep points below the margin: the four deepest (ties to the lower
    // id), emitted in ascending id order.
    int32_t candIdx[M3_CLIP_MAX];
    m3real candSep[M3_CLIP_MAX];
    int32_t candCount = 0;
    for (int32_t k = 0; k < polyCount; ++k)
    {
        m3real sep = m3Dot3(refN, poly[k]) - refOff;
        if (sep < M3_SPECULATIVE_DISTANCE)
        {
            candIdx[candCount] = k;
            candSep[candCount] = sep;
            candCount += 1;
        }
    }
    int32_t want = candCount < M3_MANIFOLD_MAX_POINTS ? candCount : M3_MANIFOLD_MAX_POINTS;
    int32_t kept[M3_MANIFOLD_MAX_POINTS];
    // `used` as a bitmask over candidates (at most M3_CLIP_MAX of them).
    uint32_t used[M3_CLIP_MAX / 32];
    for (int32_t k = 0; k < M3_CLIP_MAX / 32; ++k)
    {
        used[k] = 0u;
    }
    for (int32_t k = 0; k < want; ++k)
    {
        int32_t best = -1;
        for (int32_t c = 0; c < candCount; ++c)
        {
            if ((used[c >> 5] >> (c & 31)) & 1u)
            {
                continue;
            }
            if (best < 0 || candSep[c] < candSep[best] ||
                (candSep[c] == candSep[best] && polyId[candIdx[c]] < polyId[candIdx[best]]))
            {
                best = c;
            }
        }
        used[best >> 5] |= 1u << (best & 31);
        kept[k] = best;
    }
    for (int32_t a = 0; a < want; ++a)
    {
        for (int32_t b = a + 1; b < want; ++b)
        {
            if (polyId[candIdx[kept[b]]] < polyId[candIdx[kept[a]]])
            {
                int32_t tmp = kept[a];
                kept[a] = kept[b];
                kept[b] = tmp;
            }
        }
    }
    m3Vec3 outN = refIsA ? refN : m3Neg3(m3RotateVec3(q, refN));
    manifold.normal = outN;
    manifold.pointCount = want;
    for (int32_t k = 0; k < want; ++k)
    {
        int32_t c = kept[k];
        m3Vec3 onInc = poly[candIdx[c]];
        m3real sep = candSep[c];
        m3Vec3 onRef = m3Sub3(onInc, m3MulSV3(sep, refN));
        m3Vec3 wInc = refIsA ? onInc : m3Add3(m3RotateVec3(q, onInc), p);
        m3Vec3 wRef = refIsA ? onRef : m3Add3(m3RotateVec3(q, onRef), p);
        manifold.points[k].anchorA = refIsA ? wRef : wInc;
        manifold.points[k].anchorB = refIsA ? wInc : wRef;
        manifold.points[k].separation = sep;
        manifold.points[k].id = (uint16_t)(((uint32_t)refIsA << 15) | ((uint32_t)refFace << 8) |
                                           (polyId[candIdx[c]] & 0xFFu));
    }
    return manifold;
}

M3_HD m3Manifold m3CollideHulls(const m3HullData* hullA, const m3HullData* hullB, m3Quat q, m3Vec3 p)
{
    return m3CollideHullsQ(hullA, hullB, q, p, NULL);
}

// ---- point core inside a hull (src/manifold.c:503-520) ---------------------
M3_HD void m3DeepPointInHull(const m3HullData* hull, m3Vec3 q, m3Vec3* normalOut,
                             m3real* coreSepOut, m3Vec3* onHullOut)
{
    int32_t best = 0;
    m3real bestD = -3.4e38f;
    for (int32_t f = 0; f < hull->faceCount; ++f)
    {
        m3real d = m3Dot3(hull->faceNormals[f], q) - hull->faceOffsets[f];
        if (d > bestD)
        {
            bestD = d;
            best = f;
        }
    }
    *normalOut = hull->faceNormals[best];
    *coreSepOut = bestD;
    *onHullOut = m3Sub3(q, m3MulSV3(bestD, hull->faceNormals[best]));
}

// ---- capsule segment versus hull (src/manifold.c:534-714) ------------------

// Separation of the segment against hull face f: the deeper endpoint.
M3_HD m3real m3SegmentFaceSep(const m3HullData* hull, int32_t f, m3Vec3 p1, m3Vec3 p2)
{
    m3real d1 = m3Dot3(hull->faceNormals[f], p1) - hull->faceOffsets[f];
    m3real d2 = m3Dot3(hull->faceNormals[f], p2) - hull->faceOffsets[f];
    return m3MinF(d1, d2);
}

// The axis of hull edge e (even half-edge index) crossed with the segment,
// oriented away from the hull's centre; 0 when the two are parallel.
M3_HD int m3SegmentEdgeAxis(const m3HullData* hull, int32_t e, m3Vec3 segDir, m3Vec3* axisOut)
{
    m3Vec3 a = hull->vertices[hull->edges[e].origin];
    m3Vec3 b = hull->vertices[hull->edges[e + 1].origin];
    m3Vec3 axis = m3Cross3(m3Sub3(b, a), segDir);
    m3real len2 = m3Dot3(axis, axis);
    if (len2 < 1.0e-10f)
    {
        return 0;
    }
    axis = m3MulSV3(1.0f / m3Sqrt(len2), axis);
    if (m3Dot3(axis, m3Sub3(a, hull->center)) < 0.0f)
    {
        axis = m3Neg3(axis);
    }
    *axisOut = axis;
    return 1;
}

// Hulls whose edge query against a segment (edges x vertices projections)
// is larger than a box's (12 x 8) take the wavefront query on the device.
M3_HD int m3WideSegmentHull(const m3HullData* hull)
{
    return (hull->edgeCount >> 1) * hull->vertexCount > 96;
}

// `sat`, when given, holds the face query (faceA: the deepest face and its
// separation) and the edge query (edge: indexA the half-edge, -1 when no
// edge axis, its separation and axis) already evaluated; they are pure
// functions of the hull and the segment, so taking them from there changes
// nothing but the work done.
M3_HD m3Manifold m3CollideSegmentHullQ(const m3HullData* hull, m3Vec3 p1, m3Vec3 p2, m3real radius,
                                       const m3SatQuery* sat)
{
    const m3real linearSlop = M3_LINEAR_SLOP;
    m3Manifold manifold = m3ZeroManifold();

    int32_t bestFace = 0;
    m3real bestFaceSep = -3.4e38f;
    m3Vec3 segDir = m3Sub3(p2, p1);
    m3Vec3 bestEdgeAxis = m3V3(0.0f, 1.0f, 0.0f);
    m3real bestEdgeSep = -3.4e38f;
    int32_t bestEdge = -1;
    if (sat != NULL)
    {
        bestFace = sat->faceA.faceIndex;
        bestFaceSep = sat->faceA.separation;
        bestEdge = sat->edge.indexA;
        bestEdgeSep = sat->edge.separation;
        bestEdgeAxis = sat->edge.axis;
    }
    else
    {
        for (int32_t f = 0; f < hull->faceCount; ++f)
        {
            m3real sep = m3SegmentFaceSep(hull, f, p1, p2);
            if (sep > bestFaceSep)
            {
                bestFaceSep = sep;
                bestFace = f;
            }
        }
        for (int32_t e = 0; e < hull->edgeCount; e += 2)
        {
            m3Vec3 axis;
            if (!m3SegmentEdgeAxis(hull, e, segDir, &axis))
            {
                continue;
            }
            m3real hullMax = -3.4e38f;
            for (int32_t v = 0; v < hull->vertexCount; ++v)
            {
                m3real proj = m3Dot3(axis, hull->vertices[v]);
                if (proj > hullMax)
                {
                    hullMax = proj;
                }
            }
            m3real sep = m3MinF(m3Dot3(axis, p1), m3Dot3(axis, p2)) - hullMax;
            if (sep > bestEdgeSep)
            {
                bestEdgeSep = sep;
                bestEdgeAxis = axis;
                bestEdge = e;
            }
        }
    }

    int edgeWins = bestEdge >= 0 && bestEdgeSep > bestFaceSep + 0.1f * linearSlop;
    m3real coreSep = edgeWins ? bestEdgeSep : bestFaceSep;
    if (coreSep - radius > M3_SPECULATIVE_DISTANCE)
    {
        return manifold;
    }
    if (edgeWins)
    {
        m3Vec3 a = hull->vertices[hull->edges[bestEdge].origin];
        m3Vec3 b = hull->vertices[hull->edges[bestEdge + 1].origin];
        m3Vec3 cSeg;
        m3Vec3 cEdge;
        m3SegmentClosest(p1, segDir, a, m3Sub3(b, a), &cSeg, &cEdge);
        manifold.normal = bestEdgeAxis;
        manifold.pointCount = 1;
        manifold.points[0].anchorA = cEdge;
        manifold.points[0].anchorB = m3Sub3(cSeg, m3MulSV3(radius, bestEdgeAxis));
        manifold.points[0].separation = bestEdgeSep - radius;
        manifold.points[0].id = (uint16_t)(0x8000u | (uint32_t)bestEdge);
        return manifold;
    }

    m3Vec3 n = hull->faceNormals[bestFace];
    m3real off = hull->faceOffsets[bestFace];
    int32_t count = hull->faceVertCounts[bestFace];
    int32_t startIdx = hull->faceVertStart[bestFace];
    m3Vec3 centroid = m3V3(0.0f, 0.0f, 0.0f);
    for (int32_t k = 0; k < count; ++k)
    {
        centroid = m3Add3(centroid, hull->vertices[hull->faceIndices[startIdx + k]]);
    }
    centroid = m3MulSV3(1.0f / (m3real)count, centroid);

    m3real t0 = 0.0f;
    m3real t1 = 1.0f;
    int emptySlab = 0;
    for (int32_t k = 0; k < count; ++k)
    {
        m3Vec3 a = hull->vertices[hull->faceIndices[startIdx + k]];
        m3Vec3 b = hull->vertices[hull->faceIndices[startIdx + (k + 1) % count]];
        m3Vec3 sideN = m3Cross3(m3Sub3(b, a), n);
        if (m3Dot3(sideN, m3Sub3(centroid, a)) < 0.0f)
        {
            sideN = m3Neg3(sideN);
        }
        m3real c0 = m3Dot3(sideN, m3Sub3(p1, a));
        m3real cd = m3Dot3(sideN, segDir);
        if (cd > -1.0e-9f && cd < 1.0e-9f)
        {
            if (c0 < 0.0f)
            {
                emptySlab = 1;
                break;
            }
            continue;
        }
        m3real tc = -c0 / cd;
        if (cd > 0.0f)
        {
            t0 = m3MaxF(t0, tc);
        }
        else
        {
            t1 = m3MinF(t1, tc);
        }
    }
    if (emptySlab || t0 > t1)
    {
        m3real tm = 0.5f * (t0 + t1);
        tm = m3MaxF(0.0f, m3MinF(1.0f, tm));
        t0 = tm;
        t1 = tm;
    }
    int32_t emitted = 0;
    for (int32_t k = 0; k < 2; ++k)
    {
        if (k == 1 && !(t1 > t0))
        {
            break;
        }
        m3real t = k == 0 ? t0 : t1;
        m3Vec3 pt = m3Add3(p1, m3MulSV3(t, segDir));
        m3real d = m3Dot3(n, pt) - off;
        m3real sepK = d - radius;
        if (sepK > M3_SPECULATIVE_DISTANCE)
        {
            continue;
        }
        manifold.points[emitted].anchorA = m3Sub3(pt, m3MulSV3(d, n));
        manifold.points[emitted].anchorB = m3Sub3(pt, m3MulSV3(radius, n));
        manifold.points[emitted].separation = sepK;
        manifold.points[emitted].id = (uint16_t)k;
        emitted += 1;
    }
    manifold.normal = n;
    manifold.pointCount = emitted;
    return manifold;
}

M3_HD m3Manifold m3CollideSegmentHull(const m3HullData* hull, m3Vec3 p1, m3Vec3 p2, m3real radius)
{
    return m3CollideSegmentHullQ(hull, p1, p2, radius, NULL);
}

// ---- GJK distance (src/distance.c:1-710) -----------------------------------

#define M3_MAX_GJK_ITERATIONS 32

typedef struct m3Proxy
{
    const m3Vec3* points;
    int32_t count;
    m3real radius;
} m3Proxy;

typedef struct m3SimplexVertex
{
    m3Vec3 wA;
    m3Vec3 wB;
    m3Vec3 w;
    m3real a;
    int32_t indexA;
    int32_t indexB;
} m3SimplexVertex;

typedef struct m3Simplex
{
    m3SimplexVertex v[4];
    int32_t count;
} m3Simplex;


typedef struct m3GjkOut
{
    m3Vec3 pointA;
    m3Vec3 pointB;
    m3Vec3 normal;
    m3real distance;
} m3GjkOut;

M3_HD m3real m3ScalarTriple(m3Vec3 a, m3Vec3 b, m3Vec3 c) { return m3Dot3(a, m3Cross3(b, c)); }

M3_HD m3Vec3 m3Blend2(m3real a, m3Vec3 wa, m3real b, m3Vec3 wb)
{
    return m3V3(a * wa.x + b * wb.x, a * wa.y + b * wb.y, a * wa.z + b * wb.z);
}

M3_HD m3Vec3 m3Blend3(m3real a, m3Vec3 wa, m3real b, m3Vec3 wb, m3real c, m3Vec3 wc)
{
    return m3V3(a * wa.x + b * wb.x + c * wc.x, a * wa.y + b * wb.y + c * wc.y,
                a * wa.z + b * wb.z + c * wc.z);
}

M3_HD int32_t m3ProxySupport(const m3Proxy* proxy, m3Vec3 axis)
{
    m3Vec3 origin = proxy->points[0];
    int32_t maxIndex = 0;
    m3real maxProjection = 0.0f;
    for (int32_t i = 1; i < proxy->count; ++i)
    {
        m3real projection = m3Dot3(axis, m3Sub3(proxy->points[i], origin));
        if (projection > maxProjection)
        {
            maxIndex = i;
            maxProjection = projection;
        }
    }
    return maxIndex;
}

M3_HD void m3BaryEdge(m3real out[3], m3Vec3 a, m3Vec3 b)
{
    m3Vec3 ab = m3Sub3(b, a);
    out[0] = m3Dot3(b, ab);
    out[1] = -m3Dot3(a, ab);
    out[2] = m3Dot3(ab, ab);
}

M3_HD void m3BaryTri(m3real out[4], m3Vec3 a, m3Vec3 b, m3Vec3 c)
{
    m3Vec3 ab = m3Sub3(b, a);
    m3Vec3 ac = m3Sub3(c, a);
    m3Vec3 n = m3Cross3(ab, ac);
    out[0] = m3Dot3(m3Cross3(b, c), n);
    out[1] = m3Dot3(m3Cross3(c, a), n);
    out[2] = m3Dot3(m3Cross3(a, b), n);
    out[3] = m3Dot3(n, n);
}

M3_HD void m3BaryTet(m3real out[5], m3Vec3 a, m3Vec3 b, m3Vec3 c, m3Vec3 d)
{
    m3Vec3 ab = m3Sub3(b, a);
    m3Vec3 ac = m3Sub3(c, a);
    m3Vec3 ad = m3Sub3(d, a);
    m3real divisor = m3ScalarTriple(ab, ac, ad);
    m3real sign = divisor < 0.0f ? -1.0f : 1.0f;
    out[0] = sign * m3ScalarTriple(b, c, d);
    out[1] = sign * m3ScalarTriple(a, d, c);
    out[2] = sign * m3ScalarTriple(a, b, d);
    out[3] = sign * m3ScalarTriple(a, c, b);
    out[4] = sign * divisor;
}

// Sets a two-vertex edge simplex (x, y) with weights from a bary triple.
M3_HD bool m3SetEdge(m3Simplex* s, m3SimplexVertex x, m3SimplexVertex y, const m3real w[3])
{
    s->count = 2;
    s->v[0] = x;
    s->v[1] = y;
    if (w[2] <= 0.0f)
    {
        return false;
    }
    s->v[0].a = w[0] / w[2];
    s->v[1].a = w[1] / w[2];
    return true;
}

M3_HD bool m3SetTri(m3Simplex* s, m3SimplexVertex x, m3SimplexVertex y, m3SimplexVertex z,
                    const m3real w[4])
{
    s->count = 3;
    s->v[0] = x;
    s->v[1] = y;
    s->v[2] = z;
    if (w[3] <= 0.0f)
    {
        return false;
    }
    s->v[0].a = w[0] / w[3];
    s->v[1].a = w[1] / w[3];
    s->v[2].a = w[2] / w[3];
    return true;
}

M3_HD void m3SetVertex(m3Simplex* s, m3SimplexVertex x)
{
    s->count = 1;
    s->v[0] = x;
    s->v[0].a = 1.0f;
}

M3_HD bool m3SolveSimplex2(m3Simplex* s)
{
    m3Vec3 a = s->v[0].w;
    m3Vec3 b = s->v[1].w;
    m3Vec3 ab = m3Sub3(b, a);
    m3real divisor = m3Dot3(ab, ab);
    m3real u = m3Dot3(b, ab);
    m3real v = -m3Dot3(a, ab);
    if (v <= 0.0f)
    {
        s->count = 1;
        s->v[0].a = 1.0f;
        return true;
    }
    if (u <= 0.0f)
    {
        m3SetVertex(s, s->v[1]);
        return true;
    }
    if (divisor <= 0.0f)
    {
        return false;
    }
    m3real denominator = 1.0f / divisor;
    s->v[0].a = denominator * u;
    s->v[1].a = denominator * v;
    return true;
}

M3_HD bool m3SolveSimplex3(m3Simplex* s)
{
    m3SimplexVertex v1 = s->v[0];
    m3SimplexVertex v2 = s->v[1];
    m3SimplexVertex v3 = s->v[2];
    m3real wAB[3];
    m3real wBC[3];
    m3real wCA[3];
    m3BaryEdge(wAB, v1.w, v2.w);
    m3BaryEdge(wBC, v2.w, v3.w);
    m3BaryEdge(wCA, v3.w, v1.w);
    if (wAB[1] <= 0.0f && wCA[0] <= 0.0f)
    {
        m3SetVertex(s, v1);
        return true;
    }
    if (wBC[1] <= 0.0f && wAB[0] <= 0.0f)
    {
        m3SetVertex(s, v2);
        return true;
    }
    if (wCA[1] <= 0.0f && wBC[0] <= 0.0f)
    {
        m3SetVertex(s, v3);
        return true;
    }
    m3real wABC[4];
    m3BaryTri(wABC, v1.w, v2.w, v3.w);
    if (wABC[2] <= 0.0f && wAB[0] > 0.0f && wAB[1] > 0.0f)
    {
        return m3SetEdge(s, v1, v2, wAB);
    }
    if (wABC[0] <= 0.0f && wBC[0] > 0.0f && wBC[1] > 0.0f)
    {
        return m3SetEdge(s, v2, v3, wBC);
    }
    if (wABC[1] <= 0.0f && wCA[0] > 0.0f && wCA[1] > 0.0f)
    {
        return m3SetEdge(s, v3, v1, wCA);
    }
    if (wABC[3] <= 0.0f)
    {
        return false;
    }
    s->v[0].a = wABC[0] / wABC[3];
    s->v[1].a = wABC[1] / wABC[3];
    s->v[2].a = wABC[2] / wABC[3];
    return true;
}

M3_HD bool m3SolveSimplex4(m3Simplex* s)
{
    m3SimplexVertex vA = s->v[0];
    m3SimplexVertex vB = s->v[1];
    m3SimplexVertex vC = s->v[2];
    m3SimplexVertex vD = s->v[3];
    m3real wAB[3], wAC[3], wAD[3], wBC[3], wCD[3], wDB[3];
    m3BaryEdge(wAB, vA.w, vB.w);
    m3BaryEdge(wAC, vA.w, vC.w);
    m3BaryEdge(wAD, vA.w, vD.w);
    m3BaryEdge(wBC, vB.w, vC.w);
    m3BaryEdge(wCD, vC.w, vD.w);
    m3BaryEdge(wDB, vD.w, vB.w);
    if (wAB[1] <= 0.0f && wAC[1] <= 0.0f && wAD[1] <= 0.0f)
    {
        m3SetVertex(s, vA);
        return true;
    }
    if (wAB[0] <= 0.0f && wDB[0] <= 0.0f && wBC[1] <= 0.0f)
    {
        m3SetVertex(s, vB);
        return true;
    }
    if (wAC[0] <= 0.0f && wBC[0] <= 0.0f && wCD[1] <= 0.0f)
    {
        m3SetVertex(s, vC);
        return true;
    }
    if (wAD[0] <= 0.0f && wCD[0] <= 0.0f && wDB[1] <= 0.0f)
    {
        m3SetVertex(s, vD);
        return true;
    }
    m3real wACB[4], wABD[4], wADC[4], wBCD[4];
    m3BaryTri(wACB, vA.w, vC.w, vB.w);
    m3BaryTri(wABD, vA.w, vB.w, vD.w);
    m3BaryTri(wADC, vA.w, vD.w, vC.w);
    m3BaryTri(wBCD, vB.w, vC.w, vD.w);
    if (wABD[2] <= 0.0f && wACB[1] <= 0.0f && wAB[0] > 0.0f && wAB[1] > 0.0f)
    {
        return m3SetEdge(s, vA, vB, wAB);
    }
    if (wACB[2] <= 0.0f && wADC[1] <= 0.0f && wAC[0] > 0.0f && wAC[1] > 0.0f)
    {
        return m3SetEdge(s, vA, vC, wAC);
    }
    if (wADC[2] <= 0.0f && wABD[1] <= 0.0f && wAD[0] > 0.0f && wAD[1] > 0.0f)
    {
        return m3SetEdge(s, vA, vD, wAD);
    }
    if (wACB[0] <= 0.0f && wBCD[2] <= 0.0f && wBC[0] > 0.0f && wBC[1] > 0.0f)
    {
        return m3SetEdge(s, vB, vC, wBC);
    }
    if (wADC[0] <= 0.0f && wBCD[0] <= 0.0f && wCD[0] > 0.0f && wCD[1] > 0.0f)
    {
        return m3SetEdge(s, vC, vD, wCD);
    }
    if (wABD[0] <= 0.0f && wBCD[1] <= 0.0f && wDB[0] > 0.0f && wDB[1] > 0.0f)
    {
        return m3SetEdge(s, vD, vB, wDB);
    }
    m3real wABCD[5];
    m3BaryTet(wABCD, vA.w, vB.w, vC.w, vD.w);
    if (wABCD[3] < 0.0f && wACB[0] > 0.0f && wACB[1] > 0.0f && wACB[2] > 0.0f)
    {
        return m3SetTri(s, vA, vC, vB, wACB);
    }
    if (wABCD[2] < 0.0f && wABD[0] > 0.0f && wABD[1] > 0.0f && wABD[2] > 0.0f)
    {
        return m3SetTri(s, vA, vB, vD, wABD);
    }
    if (wABCD[1] < 0.0f && wADC[0] > 0.0f && wADC[1] > 0.0f && wADC[2] > 0.0f)
    {
        return m3SetTri(s, vA, vD, vC, wADC);
    }
    if (wABCD[0] < 0.0f && wBCD[0] > 0.0f && wBCD[1] > 0.0f && wBCD[2] > 0.0f)
    {
        return m3SetTri(s, vB, vC, vD, wBCD);
    }
    if (wABCD[4] <= 0.0f)
    {
        return false;
    }
    s->v[0].a = wABCD[0] / wABCD[4];
    s->v[1].a = wABCD[1] / wABCD[4];
    s->v[2].a = wABCD[2] / wABCD[4];
    s->v[3].a = wABCD[3] / wABCD[4];
    return true;
}

M3_HD void m3WitnessPoints(const m3Simplex* s, m3Vec3* pointA, m3Vec3* pointB)
{
    switch (s->count)
    {
    case 1:
        *pointA = s->v[0].wA;
        *pointB = s->v[0].wB;
        break;
    case 2:
        *pointA = m3Blend2(s->v[0].a, s->v[0].wA, s->v[1].a, s->v[1].wA);
        *pointB = m3Blend2(s->v[0].a, s->v[0].wB, s->v[1].a, s->v[1].wB);
        break;
    case 3:
        *pointA = m3Blend3(s->v[0].a, s->v[0].wA, s->v[1].a, s->v[1].wA, s->v[2].a, s->v[2].wA);
        *pointB = m3Blend3(s->v[0].a, s->v[0].wB, s->v[1].a, s->v[1].wB, s->v[2].a, s->v[2].wB);
        break;
    case 4:
    {
        m3Vec3 sum = m3Add3(m3Blend2(s->v[0].a, s->v[0].wA, s->v[1].a, s->v[1].wA),
                            m3Blend2(s->v[2].a, s->v[2].wA, s->v[3].a, s->v[3].wA));
        *pointA = sum;
        *pointB = sum;
        break;
    }
    default:
        *pointA = m3V3(0.0f, 0.0f, 0.0f);
        *pointB = m3V3(0.0f, 0.0f, 0.0f);
        break;
    }
}

// The simplex cache a time-of-impact search carries between its distance
// queries (src/world_internal.h:1310-1316, written by src/distance.c:123-132).
typedef struct m3SimplexCache
{
    m3real metric;
    uint16_t count;
    uint8_t indexA[4];
    uint8_t indexB[4];
} m3SimplexCache;

M3_HD m3real m3SimplexMetric(const m3Simplex* s) // src/distance.c:103-121
{
    switch (s->count)
    {
    case 1:
        return 0.0f;
    case 2:
        return m3Length3(m3Sub3(s->v[1].w, s->v[0].w));
    case 3:
        return 0.5f * m3Length3(m3Cross3(m3Sub3(s->v[1].w, s->v[0].w), m3Sub3(s->v[2].w, s->v[0].w)));
    case 4:
        return m3ScalarTriple(m3Sub3(s->v[1].w, s->v[0].w), m3Sub3(s->v[2].w, s->v[0].w),
                              m3Sub3(s->v[3].w, s->v[0].w)) /
               6.0f;
    default:
        return 0.0f;
    }
}

// GJK on the cores (radii not applied), seeded from and writing back a
// simplex cache (src/distance.c:505-706).  The overlap exits return before
// the cache is written, as there.
M3_HD m3GjkOut m3ShapeDistanceCached(const m3Proxy* proxyA, const m3Proxy* proxyB, m3Quat q, m3Vec3 p,
                                     m3SimplexCache* cache)
{
    m3GjkOut out;
    out.pointA = m3V3(0.0f, 0.0f, 0.0f);
    out.pointB = out.pointA;
    out.normal = out.pointA;
    out.distance = 0.0f;

    m3Simplex simplex;
    memset(&simplex, 0, sizeof(simplex));
    simplex.count = cache->count <= 4 ? (int32_t)cache->count : 0;
    for (int32_t i = 0; i < simplex.count; ++i)
    {
        const int32_t indexA = cache->indexA[i];
        const int32_t indexB = cache->indexB[i];
        if (indexA >= proxyA->count || indexB >= proxyB->count)
        {
            simplex.count = 0;
            break;
        }
        simplex.v[i].indexA = indexA;
        simplex.v[i].indexB = indexB;
        simplex.v[i].wA = proxyA->points[indexA];
        simplex.v[i].wB = m3Add3(m3RotateVec3(q, proxyB->points[indexB]), p);
        simplex.v[i].w = m3Sub3(simplex.v[i].wB, simplex.v[i].wA);
        simplex.v[i].a = 0.0f;
    }
    if (simplex.count > 0)
    {
        // Flush a cache whose simplex changed size substantially.
        const m3real metric1 = cache->metric;
        const m3real metric2 = m3SimplexMetric(&simplex);
        if (2.0f * metric1 < metric2 || metric2 < 0.5f * metric1 || metric2 < FLT_EPSILON)
        {
            simplex.count = 0;
        }
    }
    if (simplex.count == 0)
    {
        simplex.count = 1;
        simplex.v[0].indexA = 0;
        simplex.v[0].indexB = 0;
        simplex.v[0].wA = proxyA->points[0];
        simplex.v[0].wB = m3Add3(m3RotateVec3(q, proxyB->points[0]), p);
        simplex.v[0].w = m3Sub3(simplex.v[0].wB, simplex.v[0].wA);
        simplex.v[0].a = 0.0f;
    }

    m3Simplex backup;
    memset(&backup, 0, sizeof(backup));
    m3real distanceSq = FLT_MAX;
    m3Vec3 normal = m3V3(0.0f, 0.0f, 0.0f);

    for (int32_t iteration = 0; iteration < M3_MAX_GJK_ITERATIONS; ++iteration)
    {
        bool solved = false;
        switch (simplex.count)
        {
        case 1:
            simplex.v[0].a = 1.0f;
            solved = true;
            break;
        case 2:
            solved = m3SolveSimplex2(&simplex);
            break;
        case 3:
            solved = m3SolveSimplex3(&simplex);
            break;
        case 4:
            solved = m3SolveSimplex4(&simplex);
            break;
        default:
            break;
        }
        if (!solved)
        {
            if (backup.count == 0)
            {
                break;
            }
            simplex = backup;
            break;
        }
        if (simplex.count == 4)
        {
            m3WitnessPoints(&simplex, &out.pointA, &out.pointB);
            return out;
        }
        m3real oldDistanceSq = distanceSq;
        m3Vec3 closest;
        switch (simplex.count)
        {
        case 1:
            closest = simplex.v[0].w;
            break;
        case 2:
            closest = m3Blend2(simplex.v[0].a, simplex.v[0].w, simplex.v[1].a, simplex.v[1].w);
            break;
        default:
            closest = m3Blend3(simplex.v[0].a, simplex.v[0].w, simplex.v[1].a, simplex.v[1].w,
                               simplex.v[2].a, simplex.v[2].w);
            break;
        }
        distanceSq = m3Dot3(closest, closest);
        if (distanceSq >= oldDistanceSq)
        {
            if (backup.count != 0)
            {
                simplex = backup;
            }
            break;
        }
        m3Vec3 search;
        switch (simplex.count)
        {
        case 1:
            search = m3Neg3(simplex.v[0].w);
            break;
        case 2:
        {
            m3Vec3 ab = m3Sub3(simplex.v[1].w, simplex.v[0].w);
            search = m3Cross3(m3Cross3(ab, m3Neg3(simplex.v[0].w)), ab);
            break;
        }
        default:
        {
            m3Vec3 ab = m3Sub3(simplex.v[1].w, simplex.v[0].w);
            m3Vec3 ac = m3Sub3(simplex.v[2].w, simplex.v[0].w);
            m3Vec3 n = m3Cross3(ab, ac);
            search = m3Dot3(n, simplex.v[0].w) < 0.0f ? n : m3Neg3(n);
            break;
        }
        }
        if (m3Dot3(search, search) < 1000.0f * FLT_MIN)
        {
            m3WitnessPoints(&simplex, &out.pointA, &out.pointB);
            return out;
        }
        normal = m3Neg3(search);
        int32_t indexA = m3ProxySupport(proxyA, m3Neg3(search));
        m3Vec3 supportA = proxyA->points[indexA];
        int32_t indexB = m3ProxySupport(proxyB, m3InvRotateVec3(q, search));
        m3Vec3 supportB = m3Add3(m3RotateVec3(q, proxyB->points[indexB]), p);
        backup = simplex;
        bool duplicate = false;
        for (int32_t i = 0; i < simplex.count; ++i)
        {
            if (simplex.v[i].indexA == indexA && simplex.v[i].indexB == indexB)
            {
                duplicate = true;
                break;
            }
        }
        if (duplicate)
        {
            break;
        }
        int32_t c = simplex.count;
        simplex.v[c].indexA = indexA;
        simplex.v[c].indexB = indexB;
        simplex.v[c].wA = supportA;
        simplex.v[c].wB = supportB;
        simplex.v[c].w = m3Sub3(supportB, supportA);
        simplex.count += 1;
    }
    normal = m3Normalize3(normal);
    m3Vec3 pointA;
    m3Vec3 pointB;
    m3WitnessPoints(&simplex, &pointA, &pointB);
    cache->metric = m3SimplexMetric(&simplex);
    cache->count = (uint16_t)simplex.count;
    for (int32_t i = 0; i < simplex.count; ++i)
    {
        cache->indexA[i] = (uint8_t)simplex.v[i].indexA;
        cache->indexB[i] = (uint8_t)simplex.v[i].indexB;
    }
    out.pointA = pointA;
    out.pointB = pointB;
    out.distance = m3Length3(m3Sub3(pointA, pointB));
    out.normal = normal;
    return out;
}

// Cold-start GJK on the cores (radii not applied): the form the narrowphase
// uses (empty cache, useRadii false; src/distance.c:505-706).
M3_HD m3GjkOut m3ShapeDistance(const m3Proxy* proxyA, const m3Proxy* proxyB, m3Quat q, m3Vec3 p)
{
    m3SimplexCache cache;
    cache.metric = 0.0f;
    cache.count = 0;
    return m3ShapeDistanceCached(proxyA, proxyB, q, p, &cache);
}

#endif // M3_COLLIDE_H
