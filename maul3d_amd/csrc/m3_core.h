// Shared state layouts and pinned-order math for the host world (C17) and
// the gfx950 kernels (HIP C++).  Every function is plain IEEE arithmetic in a
// fixed evaluation order; both sides are compiled with -ffp-contract=off and
// without fast-math, so host and device produce the same bits.
//
// Layouts are the reference's padding-free persistent records, because the
// world hash and the snapshot read them byte for byte:
//   m3Vec3/m3Pos3/m3Quat/m3Transform/m3Mat3   include/maul3d/math.h:24-52
//   m3ShapeGeom                              src/world_internal.h:457-467
//   m3ManifoldPoint / m3Manifold             src/world_internal.h:471-508
//   m3HullHalfEdge / m3HullData              src/world_internal.h:219-248
// The operation order of each helper follows include/maul3d/math.h:62-226.
#ifndef M3_CORE_H
#define M3_CORE_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__)
#define M3_HD __host__ __device__ static inline
#else
#define M3_HD static inline
#endif

// Full unrolling of a loop with a constant trip count (so that arrays it
// indexes stay in registers).
#if defined(__clang__)
#define M3_UNROLL _Pragma("unroll")
#else
#define M3_UNROLL
#endif

#if defined(__cplusplus) && !defined(M3_REAL)
extern "C" {
#endif

// The scalar type of every local quantity: float for the engine.  A C++
// translation unit may define M3_REAL (e.g. double) and include these
// headers inside a namespace of its own to instantiate the same operation
// sequences at another precision.
#ifdef M3_REAL
typedef M3_REAL m3real;
#else
typedef float m3real;
#endif

#define M3_PI 3.14159265358979323846f
#define M3_AABB_MARGIN (4.0f * 0.005f)
#define M3_SPECULATIVE_DISTANCE M3_AABB_MARGIN
#define M3_LINEAR_SLOP 0.005f
#define M3_MANIFOLD_MAX_POINTS 4
#define M3_GRAPH_COLORS 16
#define M3_MAX_SUBSTEPS 256

#define M3_HULL_MAX_VERTS 64
#define M3_HULL_MAX_FACES 124
#define M3_HULL_MAX_FACE_INDICES 372
#define M3_HULL_MAX_HALF_EDGES 372

enum { M3_STATIC = 0, M3_KINEMATIC = 1, M3_DYNAMIC = 2 };
enum { M3_SPHERE = 0, M3_PLANE = 1, M3_HULL = 2, M3_CAPSULE = 3, M3_MESH = 4, M3_VOXEL = 5,
       M3_HEIGHTFIELD = 6 };

typedef struct m3Vec3 { m3real x, y, z; } m3Vec3;
typedef struct m3Pos3 { double x, y, z; } m3Pos3;
typedef struct m3Quat { m3real x, y, z, w; } m3Quat;
typedef struct m3Transform { m3Pos3 p; m3Quat q; } m3Transform;
typedef struct m3Mat3 { m3Vec3 cx, cy, cz; } m3Mat3;

typedef struct m3ShapeGeom
{
    m3Vec3 v;  // sphere center | plane normal | box half extents | capsule p1
    m3real s;  // radius | offset | unused | radius
    m3Vec3 v2; // capsule p2
    m3real s2;
} m3ShapeGeom;

typedef struct m3ManifoldPoint
{
    m3Vec3 anchorA;
    m3Vec3 anchorB;
    m3real separation;
    m3real normalImpulse;
    uint16_t id;
    uint16_t flags;
} m3ManifoldPoint;

typedef struct m3Manifold
{
    m3Vec3 normal;
    int32_t pointCount;
    m3Vec3 frictionImpulse;
    m3real twistImpulse;
    m3Vec3 rollingImpulse;
    m3ManifoldPoint points[M3_MANIFOLD_MAX_POINTS];
} m3Manifold;

typedef struct m3HullHalfEdge
{
    uint16_t twin;
    uint16_t next;
    uint8_t origin;
    uint8_t face;
} m3HullHalfEdge;

typedef struct m3HullData
{
    int32_t vertexCount;
    int32_t faceCount;
    int32_t indexCount;
    float unitMass;
    m3Vec3 unitCom;
    m3Mat3 unitInertiaCom;
    m3Vec3 vertices[M3_HULL_MAX_VERTS];
    m3Vec3 faceNormals[M3_HULL_MAX_FACES];
    m3real faceOffsets[M3_HULL_MAX_FACES];
    uint8_t faceVertCounts[M3_HULL_MAX_FACES];
    uint16_t faceVertStart[M3_HULL_MAX_FACES];
    uint8_t faceIndices[M3_HULL_MAX_FACE_INDICES];
    int32_t edgeCount;
    m3HullHalfEdge edges[M3_HULL_MAX_HALF_EDGES];
    m3Vec3 center;
} m3HullData;

#if defined(M3_REAL)
#elif defined(__cplusplus)
static_assert(sizeof(m3Transform) == 40, "transform layout");
static_assert(sizeof(m3ShapeGeom) == 32, "geom layout");
static_assert(sizeof(m3ManifoldPoint) == 36, "manifold point layout");
static_assert(sizeof(m3Manifold) == 188, "manifold layout");
static_assert(sizeof(m3HullData) == 5808, "hull layout");
#else
_Static_assert(sizeof(m3Transform) == 40, "transform layout");
_Static_assert(sizeof(m3ShapeGeom) == 32, "geom layout");
_Static_assert(sizeof(m3ManifoldPoint) == 36, "manifold point layout");
_Static_assert(sizeof(m3Manifold) == 188, "manifold layout");
_Static_assert(sizeof(m3HullData) == 5808, "hull layout");
#endif

// ---- constructors ---------------------------------------------------------

M3_HD m3Vec3 m3V3(m3real x, m3real y, m3real z)
{
    m3Vec3 v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
}

M3_HD m3Quat m3Q(m3real x, m3real y, m3real z, m3real w)
{
    m3Quat q;
    q.x = x;
    q.y = y;
    q.z = z;
    q.w = w;
    return q;
}

M3_HD m3Mat3 m3ZeroMat3(void)
{
    m3Mat3 m;
    m.cx = m3V3(0.0f, 0.0f, 0.0f);
    m.cy = m3V3(0.0f, 0.0f, 0.0f);
    m.cz = m3V3(0.0f, 0.0f, 0.0f);
    return m;
}

M3_HD m3Quat m3IdentityQuat(void) { return m3Q(0.0f, 0.0f, 0.0f, 1.0f); }

// ---- scalar ---------------------------------------------------------------

M3_HD m3real m3MinF(m3real a, m3real b) { return a < b ? a : b; }
M3_HD m3real m3MaxF(m3real a, m3real b) { return a > b ? a : b; }
M3_HD m3real m3ClampF(m3real a, m3real lo, m3real hi) { return m3MaxF(lo, m3MinF(a, hi)); }
M3_HD m3real m3AbsF(m3real a) { return a < 0.0f ? -a : a; }

// ---- vector ---------------------------------------------------------------

M3_HD m3Vec3 m3MulMV3(m3Mat3 m, m3Vec3 v)
{
    return m3V3(m.cx.x * v.x + m.cy.x * v.y + m.cz.x * v.z,
                m.cx.y * v.x + m.cy.y * v.y + m.cz.y * v.z,
                m.cx.z * v.x + m.cy.z * v.y + m.cz.z * v.z);
}

M3_HD m3Vec3 m3Add3(m3Vec3 a, m3Vec3 b) { return m3V3(a.x + b.x, a.y + b.y, a.z + b.z); }
M3_HD m3Vec3 m3Sub3(m3Vec3 a, m3Vec3 b) { return m3V3(a.x - b.x, a.y - b.y, a.z - b.z); }
M3_HD m3Vec3 m3MulSV3(m3real s, m3Vec3 v) { return m3V3(s * v.x, s * v.y, s * v.z); }
M3_HD m3Vec3 m3Neg3(m3Vec3 v) { return m3V3(-v.x, -v.y, -v.z); }
M3_HD m3real m3Dot3(m3Vec3 a, m3Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

M3_HD m3Vec3 m3Cross3(m3Vec3 a, m3Vec3 b)
{
    return m3V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

#ifdef M3_REAL
M3_HD m3real m3Sqrt(m3real x) { return sqrt(x); }
#else
M3_HD m3real m3Sqrt(m3real x) { return sqrtf(x); }
#endif

M3_HD m3real m3Length3(m3Vec3 v) { return m3Sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

// Degenerate input falls back to +y.
M3_HD m3Vec3 m3Normalize3(m3Vec3 v)
{
    m3real length = m3Length3(v);
    if (length < 1.19209290e-7f)
    {
        return m3V3(0.0f, 1.0f, 0.0f);
    }
    m3real inv = 1.0f / length;
    return m3V3(inv * v.x, inv * v.y, inv * v.z);
}

// ---- quaternion -----------------------------------------------------------

M3_HD m3Quat m3MulQuat(m3Quat a, m3Quat b)
{
    m3Quat q;
    q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    q.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    q.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return q;
}

M3_HD m3Quat m3Conj(m3Quat q) { return m3Q(-q.x, -q.y, -q.z, q.w); }

// Degenerate input falls back to identity.
M3_HD m3Quat m3NormalizeQuat(m3Quat q)
{
    m3real mag = m3Sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (mag < 1.19209290e-7f)
    {
        return m3IdentityQuat();
    }
    m3real inv = 1.0f / mag;
    return m3Q(inv * q.x, inv * q.y, inv * q.z, inv * q.w);
}

// v' = v + w t + u x t with t = 2 (u x v), u the vector part.
M3_HD m3Vec3 m3RotateVec3(m3Quat q, m3Vec3 v)
{
    m3Vec3 u = m3V3(q.x, q.y, q.z);
    m3Vec3 t = m3MulSV3(2.0f, m3Cross3(u, v));
    return m3Add3(m3Add3(v, m3MulSV3(q.w, t)), m3Cross3(u, t));
}

M3_HD m3Vec3 m3InvRotateVec3(m3Quat q, m3Vec3 v) { return m3RotateVec3(m3Conj(q), v); }

// q2 = normalize(q + 0.5 (dr, 0) q).
M3_HD m3Quat m3IntegrateRotation(m3Quat q, m3Vec3 dr)
{
    m3real ax = 0.5f * dr.x;
    m3real ay = 0.5f * dr.y;
    m3real az = 0.5f * dr.z;
    m3Quat qd;
    qd.w = -(ax * q.x + ay * q.y + az * q.z);
    qd.x = ax * q.w + ay * q.z - az * q.y;
    qd.y = -ax * q.z + ay * q.w + az * q.x;
    qd.z = ax * q.y - ay * q.x + az * q.w;
    return m3NormalizeQuat(m3Q(q.x + qd.x, q.y + qd.y, q.z + qd.z, q.w + qd.w));
}

// Minimax atan2 (src/math_functions.c:14-49 constants).
M3_HD m3real m3Atan2(m3real y, m3real x)
{
    if (x == 0.0f && y == 0.0f)
    {
        return 0.0f;
    }
    m3real ax = m3AbsF(x);
    m3real ay = m3AbsF(y);
    m3real mx = m3MaxF(ay, ax);
    m3real mn = m3MinF(ay, ax);
    m3real a = mn / mx;
    m3real s = a * a;
    m3real c = s * a;
    m3real q = s * s;
    m3real r = 0.024840285f * q + 0.18681418f;
    m3real t = -0.094097948f * q - 0.33213072f;
    r = r * s + t;
    r = r * c + a;
    if (ay > ax)
    {
        r = 1.57079637f - r;
    }
    if (x < 0.0f)
    {
        r = 3.14159274f - r;
    }
    if (y < 0.0f)
    {
        r = -r;
    }
    return r;
}

// FNV-1a 64 (src/core.c:46-54).
#define M3_HASH_INIT 0xCBF29CE484222325ull
M3_HD uint64_t m3HashBytes(uint64_t h, const void* bytes, int32_t count)
{
    const uint8_t* p = (const uint8_t*)bytes;
    for (int32_t i = 0; i < count; ++i)
    {
        h = (h ^ p[i]) * 0x100000001B3ull;
    }
    return h;
}

// Deterministic tangent basis: the axis with the smallest |n_i| (ties x<y<z).
M3_HD void m3TangentBasis(m3Vec3 normal, m3Vec3* t1, m3Vec3* t2)
{
    m3real ax = m3AbsF(normal.x);
    m3real ay = m3AbsF(normal.y);
    m3real az = m3AbsF(normal.z);
    m3Vec3 axis;
    if (ax <= ay && ax <= az)
    {
        axis = m3V3(1.0f, 0.0f, 0.0f);
    }
    else if (ay <= az)
    {
        axis = m3V3(0.0f, 1.0f, 0.0f);
    }
    else
    {
        axis = m3V3(0.0f, 0.0f, 1.0f);
    }
    *t1 = m3Normalize3(m3Cross3(normal, axis));
    *t2 = m3Cross3(normal, *t1);
}

#if defined(__cplusplus) && !defined(M3_REAL)
}
#endif

#endif // M3_CORE_H
