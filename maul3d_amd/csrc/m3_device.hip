// gfx950 step for the soft-step contact solver: the m3gpu_* C ABI
// (include/m3gpu.h) over a device-resident, structure-of-arrays world.
//
// One step is a fixed sequence of kernels on the world's stream with no
// host round trip: every count the later stages need (pairs, constraints,
// movers, colors) stays in device memory (m3Scalars) and kernels read it.
// Stage order and semantics follow m3StepInternal (src/solver.c:3662-4320).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "m3gpu.h"
#include "m3_step.h"

#define M3_SENTINEL32 0xFFFFFFFFu
#define M3_SENTINEL64 0xFFFFFFFFFFFFFFFFull

static char s_lastError[512];

static int Fail(const char* what, hipError_t e)
{
    snprintf(s_lastError, sizeof(s_lastError), "%s: %s", what, hipGetErrorString(e));
    return -1;
}

#define HIP_TRY(expr)                                                                              \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
        {                                                                                          \
            return Fail(#expr, e_);                                                                \
        }                                                                                          \
    } while (0)

struct m3gpuWorld
{
    m3gpuCaps caps;
    hipStream_t stream;
    m3View view;          // device pointers, counts filled at upload
    uint64_t* keysBuf[2]; // current pair list and the previous step's
    m3Manifold* mansBuf[2];
    int cur;
    int fullQuery; // rebuild the frozen-pair buffer before the next step
    int ccdUnsupported;
    // scratch for the broadphase sorts and compactions
    uint64_t* sapKeysIn;
    uint64_t* sapKeysOut;
    int32_t* sapValsIn;
    int32_t* sapValsOut;
    float4* sapXY;
    float2* sapZ;
    int4* sapAttr;
    uint32_t* emitSorted;
    uint32_t* emitUnique;
    uint8_t* pairFlags;
    uint8_t* bodyFlags;
    int32_t* root;
    uint16_t* bodyMasks;
    int32_t* bodyOwner;
    int32_t* iota;
    void* cubTemp;
    size_t cubBytes;
    m3Scalars* hostScalars; // pinned mirror for stats
    struct m3Profiler* prof;
    int coopBlocks;    // co-resident blocks of the substep kernel
    int coopThreads;   // threads per block of the substep kernel
    uint32_t* barrier; // substep grid barrier: arrival count, generation
};

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------

static inline unsigned Blocks(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b > 0 ? b : 1);
}

__device__ static inline uint64_t SortableDouble(double d)
{
    uint64_t u = (uint64_t)__double_as_longlong(d);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

__device__ static inline int32_t LoadRelaxed(const int32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ static inline void Emit(const m3View& v, uint32_t key)
{
    int32_t slot = atomicAdd(&v.sc->emitCount, 1);
    if (slot < v.emitCapacity)
    {
        v.emit[slot] = key;
    }
    else
    {
        v.sc->stall = 1;
    }
}

// Per-lane emission buffer: one counter atomic per M3_EMIT_BUF keys instead
// of one per key (tens of thousands of pairs contend on one counter).
#define M3_EMIT_BUF 16
struct m3EmitBuf
{
    uint32_t key[M3_EMIT_BUF];
    int32_t n;
};

__device__ static inline void EmitFlush(const m3View& v, m3EmitBuf& b)
{
    if (b.n == 0)
    {
        return;
    }
    int32_t base = atomicAdd(&v.sc->emitCount, b.n);
    for (int32_t k = 0; k < b.n; ++k)
    {
        if (base + k < v.emitCapacity)
        {
            v.emit[base + k] = b.key[k];
        }
        else
        {
            v.sc->stall = 1;
        }
    }
    b.n = 0;
}

__device__ static inline void EmitBuffered(const m3View& v, m3EmitBuf& b, uint32_t key)
{
    b.key[b.n++] = key;
    if (b.n == M3_EMIT_BUF)
    {
        EmitFlush(v, b);
    }
}

// One key per lane at most: a single atomic per wavefront.
__device__ static inline void EmitWave(const m3View& v, bool valid, uint32_t key)
{
    uint64_t mask = __ballot(valid);
    if (mask == 0)
    {
        return;
    }
    int32_t lane = __lane_id();
    int32_t leader = __ffsll((unsigned long long)mask) - 1;
    int32_t base = 0;
    if (lane == leader)
    {
        base = atomicAdd(&v.sc->emitCount, __popcll(mask));
    }
    base = __shfl(base, leader);
    if (valid)
    {
        int32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
        if (base + rank < v.emitCapacity)
        {
            v.emit[base + rank] = key;
        }
        else
        {
            v.sc->stall = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Broadphase (src/broad_phase.c:391-637)
// ---------------------------------------------------------------------------

__global__ void k_begin(m3View v)
{
    m3Scalars* sc = v.sc;
    sc->oldPairCount = sc->pairCount;
    sc->emitCount = 0;
    sc->stall = 0;
    sc->anyRestitution = 0;
    sc->islandCount = 0;
    sc->awakeCount = 0;
    sc->discoverCount = 0;
    sc->pointCount = 0;
}

__global__ void k_fill32(uint32_t* p, int32_t n, uint32_t value)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        p[i] = value;
    }
}

// Fresh bounds of every alive non-plane shape and its sweep key.
__global__ void k_fresh(m3View v, uint64_t* sapKeys, int32_t* sapVals)
{
    int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.maxShape)
    {
        return;
    }
    uint64_t key = M3_SENTINEL64;
    if (v.shapeAlive[s] != 0 && v.shapeType[s] != M3_PLANE)
    {
        double box[6];
        m3FreshBox(&v, s, box);
        for (int k = 0; k < 6; ++k)
        {
            v.fresh[6 * s + k] = box[k];
        }
        key = SortableDouble(box[0]);
        if (key == M3_SENTINEL64)
        {
            key -= 1;
        }
    }
    if (sapKeys != nullptr)
    {
        sapKeys[s] = key;
        sapVals[s] = s;
    }
}

__global__ void k_reset_sleep(m3View v) { v.sc->sleepCount = 0; }

// After a restore every sleeping body rediscovers its frozen pairs.
__global__ void k_pending_all(m3View v)
{
    int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= v.maxBody)
    {
        return;
    }
    v.pending[b] = (v.bodyAlive[b] != 0 && v.type[b] != M3_STATIC && v.awake[b] == 0) ? 1 : 0;
}

__device__ static inline void AppendFrozen(const m3View& v, int32_t i, int32_t j)
{
    int32_t slot = atomicAdd(&v.sc->sleepCount, 1);
    if (slot < v.pairCapacity)
    {
        v.sleepKeys[slot] = m3ExpandKey(m3CompactKey(i, j));
    }
}

// A proxy moves only when its fresh box escapes the stored box; it is then
// re-stored inflated by the margin (src/broad_phase.c:505-539).
__global__ void k_refresh(m3View v)
{
    int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.maxShape || v.shapeAlive[s] == 0 || v.shapeType[s] == M3_PLANE ||
        v.awake[v.shapeBody[s]] == 0)
    {
        return;
    }
    const double* f = &v.fresh[6 * s];
    double* leaf = &v.leaf[6 * s];
    if (leaf[0] <= f[0] && leaf[1] <= f[1] && leaf[2] <= f[2] && f[3] <= leaf[3] &&
        f[4] <= leaf[4] && f[5] <= leaf[5])
    {
        return;
    }
    for (int k = 0; k < 3; ++k)
    {
        leaf[k] = f[k] - (double)M3_AABB_MARGIN;
        leaf[3 + k] = f[3 + k] + (double)M3_AABB_MARGIN;
    }
}

// Outward-rounded float copies of the sorted fresh boxes: a float test that
// fails implies the exact double test fails, so the sweep filters on floats
// and confirms survivors in double.
__device__ static inline float RoundDown(double d)
{
    float f = (float)d;
    return (double)f > d ? nextafterf(f, -INFINITY) : f;
}

__device__ static inline float RoundUp(double d)
{
    float f = (float)d;
    return (double)f < d ? nextafterf(f, INFINITY) : f;
}

// Per-shape pair facts packed at gather time so the sweep's filters are
// register operations (src/broad_phase.c:182-251 read these per pair).
enum
{
    SA_QUERIES = 1,  // static or awake: runs its own tree query
    SA_HOT = 2,      // awake and not static
    SA_STATIC = 4,
    SA_ENABLED = 8,
    SA_SENSOR = 16,
    SA_PENDING = 32, // its body froze last step
    SA_DEFAULT_FILTER = 64,
};

__global__ void k_sap_gather(m3View v, const uint64_t* keys, const int32_t* vals, float4* xy,
                             float2* zz, int4* attr)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxShape)
    {
        return;
    }
    if (keys[i] == M3_SENTINEL64)
    {
        xy[i] = make_float4(INFINITY, -INFINITY, INFINITY, -INFINITY);
        zz[i] = make_float2(INFINITY, -INFINITY);
        attr[i] = make_int4(-1, -1, 0, 0);
        return;
    }
    int32_t s = vals[i];
    const double* f = &v.fresh[6 * s];
    xy[i] = make_float4(RoundDown(f[0]), RoundUp(f[3]), RoundDown(f[1]), RoundUp(f[4]));
    zz[i] = make_float2(RoundDown(f[2]), RoundUp(f[5]));
    int32_t b = v.shapeBody[s];
    int32_t fl = 0;
    bool isStatic = v.type[b] == M3_STATIC;
    fl |= (isStatic || v.awake[b] != 0) ? SA_QUERIES : 0;
    fl |= (!isStatic && v.awake[b] != 0) ? SA_HOT : 0;
    fl |= isStatic ? SA_STATIC : 0;
    fl |= v.enabled[b] != 0 ? SA_ENABLED : 0;
    fl |= v.sensor[s] != 0 ? SA_SENSOR : 0;
    fl |= v.pending[b] != 0 ? SA_PENDING : 0;
    fl |= (v.category[s] == 1ull && v.mask[s] == ~0ull && v.group[s] == 0) ? SA_DEFAULT_FILTER : 0;
    attr[i] = make_int4(s, b, fl, 0);
}

// m3PairAllowed from packed facts; shapes with custom filters take the
// full test.
__device__ static inline bool AllowedPacked(const m3View& v, int4 a, int4 b)
{
    if (a.y == b.y || !(a.z & SA_ENABLED) || !(b.z & SA_ENABLED) ||
        ((a.z & SA_STATIC) && (b.z & SA_STATIC)) || ((a.z & SA_SENSOR) && (b.z & SA_SENSOR)))
    {
        return false;
    }
    if ((a.z & SA_DEFAULT_FILTER) && (b.z & SA_DEFAULT_FILTER))
    {
        return true;
    }
    return m3PairAllowed(&v, a.x, b.x) != 0;
}

// Sweep along x over the sorted fresh boxes: every overlapping pair is
// visited once, emitted when the tree pass would emit it (m3TreePairEmits),
// and recorded in the frozen buffer when one side belongs to a body that
// froze last step (m3DiscoverHit).  Runs before the proxy refresh: the
// frozen discovery reads last step's stored boxes, as the reference's
// end-of-step discovery does.
__global__ void k_sap(m3View v, const float4* xy, const float2* zz, const int4* attr)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxShape)
    {
        return;
    }
    const int4 me = attr[i];
    if (me.x < 0)
    {
        return;
    }
    const float4 mxy = xy[i];
    const float2 mz = zz[i];
    const double* a = &v.fresh[6 * me.x];
    m3EmitBuf buf;
    buf.n = 0;
    for (int32_t j = i + 1; j < v.maxShape; ++j)
    {
        float4 oxy = xy[j];
        if (oxy.x > mxy.y) // sentinels have +inf lower x: the sweep ends
        {
            break;
        }
        if (oxy.z > mxy.w || mxy.z > oxy.w)
        {
            continue;
        }
        float2 oz = zz[j];
        if (oz.x > mz.y || mz.x > oz.y)
        {
            continue;
        }
        const int4 ot = attr[j];
        const double* b = &v.fresh[6 * ot.x];
        if (!m3BoxOverlap(a, b))
        {
            continue;
        }
        bool pend = ((me.z | ot.z) & SA_PENDING) != 0;
        bool hot = ((me.z | ot.z) & SA_HOT) != 0;
        bool queries = ((me.z | ot.z) & SA_QUERIES) != 0;
        if (!pend && !(hot && queries))
        {
            continue;
        }
        if (!AllowedPacked(v, me, ot))
        {
            continue;
        }
        if ((me.z & SA_PENDING) && m3BoxOverlap(&v.leaf[6 * ot.x], a))
        {
            AppendFrozen(v, me.x, ot.x);
        }
        if ((ot.z & SA_PENDING) && m3BoxOverlap(&v.leaf[6 * me.x], b))
        {
            AppendFrozen(v, ot.x, me.x);
        }
        if (!hot || !queries)
        {
            continue;
        }
        bool qa = (me.z & SA_QUERIES) != 0;
        bool qb = (ot.z & SA_QUERIES) != 0;
        bool emit = (qa && qb) || (qa ? m3BoxOverlap(&v.leaf[6 * ot.x], a)
                                      : m3BoxOverlap(&v.leaf[6 * me.x], b));
        if (emit)
        {
            EmitBuffered(v, buf, m3CompactKey(me.x, ot.x));
        }
    }
    EmitFlush(v, buf);
}

// Planes pair with every allowed hot shape (src/broad_phase.c:541-564).
__global__ void k_planes(m3View v)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool hot = false;
    uint32_t key = 0;
    if (idx < (int64_t)v.maxShape * v.planeCount)
    {
        int32_t s = (int32_t)(idx % v.maxShape);
        int32_t p = v.planes[idx / v.maxShape];
        if (v.shapeAlive[s] != 0 && v.shapeType[s] != M3_PLANE && m3PairAllowed(&v, p, s))
        {
            if (v.pending[v.shapeBody[s]])
            {
                AppendFrozen(v, p, s); // planes pair unconditionally with a frozen body
            }
            hot = m3PairHot(&v, p, s) != 0;
            key = m3CompactKey(p, s);
        }
    }
    EmitWave(v, hot, key);
}

// Still-cold frozen pairs ride along (src/broad_phase.c:607-632).
__global__ void k_frozen(m3View v)
{
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t n = v.sc->sleepCount < v.pairCapacity ? v.sc->sleepCount : v.pairCapacity;
    bool cold = false;
    uint32_t key = 0;
    if (k < n)
    {
        uint64_t full = v.sleepKeys[k];
        int32_t i = (int32_t)(full >> 32);
        int32_t j = (int32_t)(full & 0xFFFFFFFFu);
        cold = v.shapeAlive[i] != 0 && v.shapeAlive[j] != 0 && !m3PairHot(&v, i, j);
        key = m3CompactKey(i, j);
    }
    EmitWave(v, cold, key);
}

// The unique sorted keys become this step's pair list; the cold ones are
// the frozen buffer carried to the next step.
__global__ void k_pair_count(m3View v, const uint32_t* unique)
{
    m3Scalars* sc = v.sc;
    int32_t n = sc->uniqueCount;
    if (n > 0 && unique[n - 1] == M3_SENTINEL32)
    {
        n -= 1;
    }
    if (sc->emitCount > v.emitCapacity || n > v.pairCapacity)
    {
        sc->stall = 1;
    }
    if (sc->stall)
    {
        sc->pairCount = sc->oldPairCount;
        return;
    }
    sc->pairCount = n;
}

__global__ void k_expand(m3View v, const uint32_t* unique, uint8_t* coldFlags)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.pairCapacity)
    {
        return;
    }
    uint8_t cold = 0;
    if (v.sc->stall == 0 && i < v.sc->pairCount)
    {
        uint64_t key = m3ExpandKey(unique[i]);
        v.keys[i] = key;
        cold = m3PairHot(&v, (int32_t)(key >> 32), (int32_t)(key & 0xFFFFFFFFu)) ? 0 : 1;
    }
    coldFlags[i] = cold;
}

// On a capacity stall the previous pair list stays the world's list.
__global__ void k_stall_restore(m3View v)
{
    if (v.sc->stall == 0)
    {
        return;
    }
    int32_t n = v.sc->oldPairCount;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        v.keys[i] = v.oldKeys[i];
        v.manifolds[i] = v.oldManifolds[i];
    }
}

// ---------------------------------------------------------------------------
// Narrowphase (src/manifold.c:2357-2944)
// ---------------------------------------------------------------------------

__global__ void k_narrow(m3View v)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || i >= v.sc->pairCount)
    {
        return;
    }
    m3NarrowPair(&v, i, v.sc->oldPairCount);
}

// ---------------------------------------------------------------------------
// Sweep capture and islands (src/solver.c:3437-3580, 3851-3881)
// ---------------------------------------------------------------------------

__global__ void k_capture(m3View v)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0)
    {
        return;
    }
    if (v.bodyAlive[i] != 0)
    {
        m3Transform xf = v.xf[i];
        m3Vec3 rlc = m3RotateVec3(xf.q, v.lc[i]);
        m3Pos3 c;
        c.x = xf.p.x + (double)rlc.x;
        c.y = xf.p.y + (double)rlc.y;
        c.z = xf.p.z + (double)rlc.z;
        v.com0[i] = c;
        v.rot0[i] = xf.q;
    }
    v.parent[i] = i;
    v.forced[i] = 0;
    v.flag[i] = 0;
    v.pending[i] = 0; // the sweep has recorded last step's freezes
    v.deltaPos[i] = m3V3(0.0f, 0.0f, 0.0f);
    v.deltaRot[i] = m3IdentityQuat();
}

__device__ static int32_t Find(const int32_t* parent, int32_t i)
{
    while (true)
    {
        int32_t p = LoadRelaxed(&parent[i]);
        if (p == i)
        {
            return i;
        }
        i = p;
    }
}

// Lock-free union that always hooks the larger root under the smaller, so
// each component's root is its lowest body index: the reference's
// "lower root wins" labels.
__device__ static void Unite(int32_t* parent, int32_t a, int32_t b)
{
    while (true)
    {
        a = Find(parent, a);
        b = Find(parent, b);
        if (a == b)
        {
            return;
        }
        int32_t lo = a < b ? a : b;
        int32_t hi = a < b ? b : a;
        if (atomicCAS(&parent[hi], hi, lo) == hi)
        {
            return;
        }
    }
}

__global__ void k_union(m3View v)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || i >= v.sc->pairCount || v.manifolds[i].pointCount == 0)
    {
        return;
    }
    uint64_t key = v.keys[i];
    int32_t sA = (int32_t)(key >> 32);
    int32_t sB = (int32_t)(key & 0xFFFFFFFFu);
    if (v.sensor[sA] != 0 || v.sensor[sB] != 0)
    {
        return;
    }
    int32_t bA = v.shapeBody[sA];
    int32_t bB = v.shapeBody[sB];
    int dynA = v.type[bA] == M3_DYNAMIC;
    int dynB = v.type[bB] == M3_DYNAMIC;
    if (dynA && dynB)
    {
        Unite(v.parent, bA, bB);
    }
    else if (dynA || dynB)
    {
        int32_t kin = dynA ? bB : bA;
        int32_t dyn = dynA ? bA : bB;
        if (v.type[kin] == M3_KINEMATIC)
        {
            m3Vec3 lin = v.lv[kin];
            m3Vec3 ang = v.av[kin];
            if (m3Dot3(lin, lin) > 0.0f || m3Dot3(ang, ang) > 0.0f)
            {
                v.forced[dyn] = 1;
            }
        }
    }
}

__global__ void k_root(m3View v, int32_t* root)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0)
    {
        return;
    }
    root[i] = Find(v.parent, i);
    v.slotMask[i] = 0u;
}

__global__ void k_wake_mark(m3View v, const int32_t* root)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0 || v.bodyAlive[i] == 0 || v.type[i] != M3_DYNAMIC)
    {
        return;
    }
    if (v.awake[i] != 0 || v.forced[i] != 0)
    {
        v.flag[root[i]] = 1;
    }
}

__global__ void k_wake_apply(m3View v, const int32_t* root)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0 || v.bodyAlive[i] == 0 || v.type[i] != M3_DYNAMIC)
    {
        return;
    }
    if (v.awake[i] == 0 && v.flag[root[i]] != 0)
    {
        v.awake[i] = 1;
        v.sleepTime[i] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Prepare and coloring (src/solver.c:156-404, 2741-2822)
// ---------------------------------------------------------------------------

__global__ void k_constraint_flags(m3View v, uint8_t* flags)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.pairCapacity)
    {
        return;
    }
    flags[i] = (v.sc->stall == 0 && i < v.sc->pairCount && m3PairIsConstraint(&v, i)) ? 1 : 0;
}

// The awake dynamic bodies of each constraint, the only bodies coloring
// tracks (src/solver.c:2741-2822).
__global__ void k_cons_bodies(m3View v)
{
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || c >= v.sc->constraintCount)
    {
        return;
    }
    uint64_t key = v.keys[v.consPair[c]];
    int32_t a = v.shapeBody[(int32_t)(key >> 32)];
    int32_t b = v.shapeBody[(int32_t)(key & 0xFFFFFFFFu)];
    v.consBodies[2 * c] = (v.type[a] == M3_DYNAMIC && v.awake[a] != 0) ? a : -1;
    v.consBodies[2 * c + 1] = (v.type[b] == M3_DYNAMIC && v.awake[b] != 0) ? b : -1;
}

__device__ static inline void RegisterSlot(const m3View& v, int32_t body, int32_t slot, int side,
                                           int32_t color)
{
    if (color < M3_GRAPH_COLORS)
    {
        v.slotTable[(int64_t)body * M3_GRAPH_COLORS + color] = (slot << 1) | side;
        atomicOr(&v.slotMask[body], 1u << color);
    }
    else
    {
        atomicOr(&v.slotMask[body], 0x10000u);
    }
}

// Prepares constraints straight into color-slot order (the substep passes
// then read slot k without an index hop) and records each dynamic body's
// slots.  A dynamic body in a constraint is awake: it touches an awake
// dynamic body or a static/kinematic one, and touching dynamic bodies share
// an island, which the wake pass woke whole.  So coloring gave it at most
// one slot per color.
__global__ void k_prepare(m3View v, float hertz, float zeta, float h)
{
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || k >= v.sc->constraintCount)
    {
        return;
    }
    m3Softness soft = m3MakeSoft(hertz, zeta, h);
    m3Softness staticSoft = m3MakeSoft(2.0f * hertz, zeta, h);
    int32_t c = v.colorLists[k];
    m3Constraint* con = &v.cons[k];
    m3PrepareContact(&v, v.consPair[c], con, soft, staticSoft);
    v.slotOf[c] = k;
    int32_t color = v.slotColors[k];
    if (con->dynA)
    {
        RegisterSlot(v, con->bodyA, k, 0, color);
    }
    if (con->dynB)
    {
        RegisterSlot(v, con->bodyB, k, 1, color);
    }
    atomicAdd(&v.sc->pointCount, con->pointCount);
    if (con->restitution > 0.0f)
    {
        v.sc->anyRestitution = 1;
    }
}

// Greedy first-fit coloring in constraint order (src/solver.c:2741-2822),
// reproduced exactly by one workgroup.  Constraints are taken 1024 at a
// time; inside a chunk, a constraint colors in the first round in which it
// is the lowest pending constraint on each of its awake dynamic bodies
// (an owner table claimed with atomic min), so every earlier constraint that
// shares a body has already updated that body's mask.  Body masks and
// owners live in LDS for worlds up to M3_COLOR_LDS_BODIES bodies.
#define M3_COLOR_THREADS 1024
#define M3_COLOR_LDS_BODIES 16384

template <bool kLds>
__global__ void __launch_bounds__(M3_COLOR_THREADS)
    k_color(m3View v, uint16_t* gMasks, int32_t* gOwner)
{
    __shared__ uint16_t sMask[kLds ? M3_COLOR_LDS_BODIES : 1];
    __shared__ int32_t sOwner[kLds ? M3_COLOR_LDS_BODIES : 1];
    __shared__ int32_t counts[M3_GRAPH_COLORS + 1];
    uint16_t* masks = kLds ? sMask : gMasks;
    int32_t* owner = kLds ? sOwner : gOwner;
    const int32_t t = threadIdx.x;
    m3Scalars* sc = v.sc;
    const int32_t n = sc->stall != 0 ? 0 : sc->constraintCount;
    for (int32_t i = t; i < v.maxBody; i += M3_COLOR_THREADS)
    {
        masks[i] = 0;
        owner[i] = 0x7FFFFFFF;
    }
    // Unused slots sort after every color in the list build.
    for (int32_t c = n + t; c < v.pairCapacity; c += M3_COLOR_THREADS)
    {
        v.colors[c] = 0xFF;
    }
    if (t <= M3_GRAPH_COLORS)
    {
        counts[t] = 0;
    }
    __syncthreads();
    for (int32_t base = 0; base < n; base += M3_COLOR_THREADS)
    {
        int32_t c = base + t;
        int32_t A = -1;
        int32_t B = -1;
        int pend = c < n;
        if (pend)
        {
            A = v.consBodies[2 * c];
            B = v.consBodies[2 * c + 1];
        }
        while (true)
        {
            if (pend)
            {
                if (A >= 0)
                {
                    atomicMin(&owner[A], t);
                }
                if (B >= 0)
                {
                    atomicMin(&owner[B], t);
                }
            }
            if (!__syncthreads_or(pend))
            {
                break;
            }
            int ready = pend && (A < 0 || owner[A] == t) && (B < 0 || owner[B] == t);
            if (ready)
            {
                uint32_t mask = (A >= 0 ? (uint32_t)masks[A] : 0u) | (B >= 0 ? (uint32_t)masks[B] : 0u);
                int32_t color = M3_GRAPH_COLORS;
                for (int32_t bit = 0; bit < M3_GRAPH_COLORS; ++bit)
                {
                    if ((mask & (1u << bit)) == 0u)
                    {
                        color = bit;
                        break;
                    }
                }
                if (color < M3_GRAPH_COLORS)
                {
                    if (A >= 0)
                    {
                        masks[A] = (uint16_t)(masks[A] | (1u << color));
                    }
                    if (B >= 0)
                    {
                        masks[B] = (uint16_t)(masks[B] | (1u << color));
                    }
                }
                v.colors[c] = (uint8_t)color;
                atomicAdd(&counts[color], 1);
            }
            __syncthreads();
            if (ready)
            {
                pend = 0;
                if (A >= 0)
                {
                    owner[A] = 0x7FFFFFFF;
                }
                if (B >= 0)
                {
                    owner[B] = 0x7FFFFFFF;
                }
            }
            __syncthreads();
        }
    }
    if (t == 0)
    {
        int32_t cursor = 0;
        int32_t used = 0;
        for (int32_t k = 0; k <= M3_GRAPH_COLORS; ++k)
        {
            sc->colorStarts[k] = cursor;
            cursor += counts[k];
            used += counts[k] > 0 ? 1 : 0;
        }
        sc->colorStarts[M3_GRAPH_COLORS + 1] = cursor;
        sc->colorCount = used;
    }
}

__global__ void k_iota(int32_t* p, int32_t n)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        p[i] = i;
    }
}

// ---------------------------------------------------------------------------
// Movers and the substep loop (src/solver.c:3938-4203)
// ---------------------------------------------------------------------------

__global__ void k_mover_flags(m3View v, uint8_t* flags, float dt)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody)
    {
        return;
    }
    uint8_t f = 0;
    if (v.sc->stall == 0 && v.bodyAlive[i] != 0 && v.enabled[i] != 0)
    {
        uint8_t type = v.type[i];
        if (type == M3_KINEMATIC || (type == M3_DYNAMIC && v.awake[i] != 0))
        {
            f = 1;
        }
        if (type == M3_KINEMATIC && v.hasTarget[i] != 0)
        {
            m3KinematicServo(&v, i, dt);
        }
    }
    flags[i] = f;
}

struct m3SubstepParams
{
    int32_t substeps;
    float h;
    float invH;
    float push;
    float maxLinear;
    float maxAngular;
    m3Vec3 gravity;
};

// Grid-wide barrier of the substep kernel.  The fences are those of a
// cooperative-groups grid sync (agent-scope release before arriving,
// acquire once the generation flips); arrival is one counter atomic per
// block.  All blocks are co-resident (cooperative launch).  The wait is
// bounded: a barrier that times out marks the world faulted (the host
// reports it at the next synchronization) and later barriers of a faulted
// world pass at once, so every block still runs to completion.
#define M3_BARRIER_SPINS (1u << 19)

__device__ static inline void GridBarrier(uint32_t* bar, int32_t* fault)
{
    __syncthreads();
    if (threadIdx.x == 0 &&
        __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    {
        uint32_t gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t prev =
            __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == gridDim.x - 1u)
        {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        else
        {
            // Relaxed polling reads the coherent copy without invalidating
            // this XCD's L2 on every poll; one acquire fence follows.
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen)
            {
                __builtin_amdgcn_s_sleep(1);
                if (++spins == M3_BARRIER_SPINS ||
                    __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
                {
                    __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
}

// One graph color of one pass over the grid (mode 1 solve with bias,
// 2 relax); the overflow color runs serially on one thread.
__device__ static inline void ColorPhase(const m3View& v, int32_t start, int32_t end, int mode,
                                         const m3SubstepParams& sp, int32_t gid, int32_t gsz,
                                         bool serial)
{
    if (serial)
    {
        if (gid == 0)
        {
            for (int32_t k = start; k < end; ++k)
            {
                m3SolveContact(&v, &v.cons[k], sp.invH, sp.push, mode == 1);
            }
        }
        return;
    }
    for (int32_t k = start + gid; k < end; k += gsz)
    {
        m3SolveContact(&v, &v.cons[k], sp.invH, sp.push, mode == 1);
    }
}

#define M3_SUBSTEP_THREADS 256

// The whole substep loop of src/solver.c:4085-4203 in one cooperative
// launch, with a grid barrier after each phase:
//   integrate velocities and warm start (one phase: a warm-start impulse
//   depends only on stored accumulators, so each body replays its own
//   constraints in color order, the exact operation sequence of the colored
//   pass on that body), solve by color, integrate positions, relax by color.
// Every branch depends only on counts written before the launch, so all
// blocks reach the same barriers.
__global__ void __launch_bounds__(M3_SUBSTEP_THREADS)
    k_substeps(m3View v, m3SubstepParams sp, uint32_t* bar)
{
    if (v.sc->stall != 0)
    {
        return;
    }
    int32_t starts[M3_GRAPH_COLORS + 2];
    for (int32_t k = 0; k < M3_GRAPH_COLORS + 2; ++k)
    {
        starts[k] = v.sc->colorStarts[k];
    }
    const int32_t movers = v.sc->moverCount;
    const int32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t gsz = gridDim.x * blockDim.x;
    for (int32_t sub = 0; sub < sp.substeps; ++sub)
    {
        for (int32_t m = gid; m < movers; m += gsz)
        {
            int32_t b = v.movers[m];
            m3IntegrateVelocity(&v, b, sp.h, sp.gravity, sp.maxLinear, sp.maxAngular);
            m3WarmStartBody(&v, b, starts[M3_GRAPH_COLORS], starts[M3_GRAPH_COLORS + 1]);
        }
        GridBarrier(bar, &v.sc->fault);
        for (int mode = 1; mode < 3; ++mode)
        {
            if (mode == 2)
            {
                for (int32_t m = gid; m < movers; m += gsz)
                {
                    m3IntegratePosition(&v, v.movers[m], sp.h);
                }
                GridBarrier(bar, &v.sc->fault);
            }
            for (int32_t color = 0; color <= M3_GRAPH_COLORS; ++color)
            {
                int32_t start = starts[color];
                int32_t end = starts[color + 1];
                if (end == start)
                {
                    continue;
                }
                ColorPhase(v, start, end, mode, sp, gid, gsz, color == M3_GRAPH_COLORS);
                GridBarrier(bar, &v.sc->fault);
            }
        }
    }
}

__global__ void k_clear_forces(m3View v)
{
    int32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || m >= v.sc->moverCount)
    {
        return;
    }
    int32_t i = v.movers[m];
    v.force[i] = m3V3(0.0f, 0.0f, 0.0f);
    v.torque[i] = m3V3(0.0f, 0.0f, 0.0f);
}

__global__ void k_restitution(m3View v, float threshold)
{
    if (v.sc->stall != 0 || v.sc->anyRestitution == 0)
    {
        return;
    }
    int32_t n = v.sc->constraintCount;
    for (int32_t c = 0; c < n; ++c)
    {
        m3RestitutionOne(&v, &v.cons[v.slotOf[c]], threshold);
    }
}

__global__ void k_store(m3View v)
{
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || c >= v.sc->constraintCount)
    {
        return;
    }
    m3StoreOne(&v, &v.cons[c]);
}

// ---------------------------------------------------------------------------
// Continuous collision and sleep (src/solver.c:3349-3430, 3582-3660)
// ---------------------------------------------------------------------------

__global__ void k_ccd(m3View v)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0)
    {
        return;
    }
    m3ContinuousBody(&v, i);
}

__global__ void k_sleep_timers(m3View v, const int32_t* root, float dt, int sleepEnabled)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0)
    {
        return;
    }
    v.flag[i] = 1; // island readiness, cleared below by any slow member
    if (v.bodyAlive[i] == 0 || v.type[i] != M3_DYNAMIC || v.awake[i] == 0)
    {
        return;
    }
    if (root[i] == i)
    {
        atomicAdd(&v.sc->islandCount, 1);
    }
    if (sleepEnabled)
    {
        m3SleepTimer(&v, i, dt);
    }
}

__global__ void k_sleep_ready(m3View v, const int32_t* root)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0 || v.bodyAlive[i] == 0 || v.type[i] != M3_DYNAMIC ||
        v.awake[i] == 0)
    {
        return;
    }
    if (v.sleepTime[i] < 0.5f)
    {
        v.flag[root[i]] = 0;
    }
}

__global__ void k_sleep_freeze(m3View v, const int32_t* root)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0 || v.bodyAlive[i] == 0 || v.type[i] != M3_DYNAMIC ||
        v.awake[i] == 0)
    {
        return;
    }
    if (v.flag[root[i]] != 0)
    {
        v.awake[i] = 0;
        v.lv[i] = m3V3(0.0f, 0.0f, 0.0f);
        v.av[i] = m3V3(0.0f, 0.0f, 0.0f);
        v.pending[i] = 1; // frozen pairs are recorded by the next sweep
        atomicAdd(&v.sc->discoverCount, 1);
    }
}

__global__ void k_end(m3View v)
{
    m3Scalars* sc = v.sc;
    if (sc->stall == 0)
    {
        sc->stepCount += 1;
    }
    if (sc->sleepCount > v.pairCapacity)
    {
        sc->sleepCount = v.pairCapacity;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

template <typename T>
static int DevAlloc(T** p, int64_t count)
{
    int64_t bytes = count > 0 ? count * (int64_t)sizeof(T) : (int64_t)sizeof(T);
    HIP_TRY(hipMalloc((void**)p, (size_t)bytes));
    HIP_TRY(hipMemset(*p, 0, (size_t)bytes));
    return 0;
}

extern "C" int m3gpu_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
    {
        return 0;
    }
    return n;
}

extern "C" const char* m3gpu_last_error(void) { return s_lastError; }

static int CreateImpl(m3gpuWorld* w)
{
    const m3gpuCaps& c = w->caps;
    m3View& v = w->view;
    memset(&v, 0, sizeof(v));
    v.pairCapacity = c.pairCapacity;
    v.emitCapacity = 2 * c.pairCapacity;
    int32_t nb = c.bodyCapacity;
    int32_t ns = c.shapeCapacity;
    int32_t np = c.pairCapacity;
    HIP_TRY(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    if (DevAlloc(&v.xf, nb) || DevAlloc(&v.lv, nb) || DevAlloc(&v.av, nb) ||
        DevAlloc(&v.invMass, nb) || DevAlloc(&v.invI, nb) || DevAlloc(&v.inertia, nb) ||
        DevAlloc(&v.lc, nb) || DevAlloc(&v.gravityScale, nb) || DevAlloc(&v.linDamp, nb) ||
        DevAlloc(&v.angDamp, nb) || DevAlloc(&v.bodyAlive, nb) || DevAlloc(&v.type, nb) ||
        DevAlloc(&v.awake, nb) || DevAlloc(&v.sleepTime, nb) || DevAlloc(&v.bullet, nb) ||
        DevAlloc(&v.minExt, nb) || DevAlloc(&v.maxExt, nb) || DevAlloc(&v.force, nb) ||
        DevAlloc(&v.torque, nb) || DevAlloc(&v.enabled, nb) || DevAlloc(&v.locks, nb) ||
        DevAlloc(&v.sleepThreshold, nb) || DevAlloc(&v.canSleep, nb) ||
        DevAlloc(&v.hasTarget, nb) || DevAlloc(&v.target, nb) || DevAlloc(&v.shapeHead, nb))
    {
        return -1;
    }
    if (DevAlloc(&v.shapeAlive, ns) || DevAlloc(&v.shapeBody, ns) || DevAlloc(&v.shapeType, ns) ||
        DevAlloc(&v.geom, ns) || DevAlloc(&v.friction, ns) || DevAlloc(&v.restitution, ns) ||
        DevAlloc(&v.rolling, ns) || DevAlloc(&v.sensor, ns) || DevAlloc(&v.hasOffset, ns) ||
        DevAlloc(&v.localPos, ns) || DevAlloc(&v.localRot, ns) || DevAlloc(&v.surfaceVel, ns) ||
        DevAlloc(&v.category, ns) || DevAlloc(&v.mask, ns) || DevAlloc(&v.group, ns) ||
        DevAlloc(&v.shapeNext, ns) || DevAlloc(&v.hullIndex, ns) || DevAlloc(&v.leaf, 6 * (int64_t)ns) ||
        DevAlloc(&v.fresh, 6 * (int64_t)ns) || DevAlloc(&v.planes, ns))
    {
        return -1;
    }
    m3HullData* hulls = nullptr;
    if (DevAlloc(&hulls, c.hullCapacity))
    {
        return -1;
    }
    v.hulls = hulls;
    for (int k = 0; k < 2; ++k)
    {
        if (DevAlloc(&w->keysBuf[k], np) || DevAlloc(&w->mansBuf[k], np))
        {
            return -1;
        }
    }
    if (DevAlloc(&v.sleepKeys, np) || DevAlloc(&v.emit, v.emitCapacity) ||
        DevAlloc(&w->emitSorted, v.emitCapacity) || DevAlloc(&w->emitUnique, v.emitCapacity) ||
        DevAlloc(&v.cons, np) || DevAlloc(&v.consPair, np) || DevAlloc(&v.colors, np) ||
        DevAlloc(&v.colorLists, np) || DevAlloc(&v.movers, nb) || DevAlloc(&v.deltaPos, nb) ||
        DevAlloc(&v.deltaRot, nb) || DevAlloc(&v.com0, nb) || DevAlloc(&v.rot0, nb) ||
        DevAlloc(&v.parent, nb) || DevAlloc(&v.forced, nb) || DevAlloc(&v.flag, nb) ||
        DevAlloc(&v.pending, nb) ||
        DevAlloc(&v.sc, 1) || DevAlloc(&w->sapKeysIn, ns) || DevAlloc(&w->sapKeysOut, ns) ||
        DevAlloc(&w->sapValsIn, ns) || DevAlloc(&w->sapValsOut, ns) ||
        DevAlloc(&w->sapXY, ns) || DevAlloc(&w->sapZ, ns) || DevAlloc(&w->sapAttr, ns) ||
        DevAlloc(&w->pairFlags, np) || DevAlloc(&w->bodyFlags, nb) || DevAlloc(&w->root, nb) ||
        DevAlloc(&w->bodyMasks, nb) || DevAlloc(&w->bodyOwner, nb) ||
        DevAlloc(&w->iota, np) || DevAlloc(&v.slotColors, np) ||
        DevAlloc(&v.consBodies, 2 * (int64_t)np) || DevAlloc(&v.slotOf, np) ||
        DevAlloc(&v.slotTable, (int64_t)M3_GRAPH_COLORS * nb) || DevAlloc(&v.slotMask, nb) ||
        DevAlloc(&w->barrier, 4))
    {
        return -1;
    }
    HIP_TRY(hipHostMalloc((void**)&w->hostScalars, sizeof(m3Scalars), hipHostMallocDefault));
    memset(w->hostScalars, 0, sizeof(m3Scalars));

    // One temp buffer sized for the largest primitive call.
    size_t need = 0;
    size_t bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, w->sapKeysIn, w->sapKeysOut,
                                               w->sapValsIn, w->sapValsOut, ns, 0, 64, w->stream));
    need = bytes > need ? bytes : need;
    HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, v.emit, w->emitSorted,
                                              v.emitCapacity, 0, 32, w->stream));
    need = bytes > need ? bytes : need;
    HIP_TRY(hipcub::DeviceSelect::Unique(nullptr, bytes, w->emitSorted, w->emitUnique,
                                         &v.sc->uniqueCount, v.emitCapacity, w->stream));
    need = bytes > need ? bytes : need;
    HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, bytes, v.keys, w->pairFlags, v.sleepKeys,
                                          &v.sc->sleepCount, np, w->stream));
    need = bytes > need ? bytes : need;
    hipcub::CountingInputIterator<int32_t> counting(0);
    HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, bytes, counting, w->pairFlags, v.consPair,
                                          &v.sc->constraintCount, np, w->stream));
    need = bytes > need ? bytes : need;
    HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, bytes, counting, w->bodyFlags, v.movers,
                                          &v.sc->moverCount, nb, w->stream));
    need = bytes > need ? bytes : need;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, v.colors, v.slotColors, w->iota,
                                               v.colorLists, np, 0, 5, w->stream));
    need = bytes > need ? bytes : need;
    hipLaunchKernelGGL(k_iota, dim3(Blocks(np, 256)), dim3(256), 0, w->stream, w->iota, np);
    w->cubBytes = need > 0 ? need : 1;
    {
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, dev));
        if (!prop.cooperativeLaunch)
        {
            snprintf(s_lastError, sizeof(s_lastError), "device lacks cooperative launch");
            return -1;
        }
        int perCu = 0;
        w->coopThreads = M3_SUBSTEP_THREADS;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_substeps, w->coopThreads, 0));
        int maxBlocks = perCu * prop.multiProcessorCount;
        const char* env = getenv("MAUL3D_AMD_SUBSTEP_BLOCKS");
        int want = env != nullptr ? atoi(env) : 32;
        w->coopBlocks = want < 1 ? 1 : (want > maxBlocks ? maxBlocks : want);
    }
    HIP_TRY(hipMalloc(&w->cubTemp, w->cubBytes));
    w->cur = 0;
    v.keys = w->keysBuf[0];
    v.manifolds = w->mansBuf[0];
    v.oldKeys = w->keysBuf[1];
    v.oldManifolds = w->mansBuf[1];
    HIP_TRY(hipStreamSynchronize(w->stream));
    return 0;
}

extern "C" m3gpuWorld* m3gpu_world_create(const m3gpuCaps* caps)
{
    if (caps == nullptr || caps->bodyCapacity <= 0 || caps->shapeCapacity <= 0 ||
        caps->pairCapacity <= 0 || caps->shapeCapacity > 65535)
    {
        snprintf(s_lastError, sizeof(s_lastError),
                 "m3gpu_world_create: capacities must be positive and shapeCapacity <= 65535");
        return nullptr;
    }
    m3gpuWorld* w = (m3gpuWorld*)calloc(1, sizeof(m3gpuWorld));
    if (w == nullptr)
    {
        return nullptr;
    }
    w->caps = *caps;
    if (CreateImpl(w) != 0)
    {
        free(w);
        return nullptr;
    }
    return w;
}

static void ProfFree(m3gpuWorld* w);

extern "C" void m3gpu_world_destroy(m3gpuWorld* w)
{
    if (w == nullptr)
    {
        return;
    }
    (void)hipStreamSynchronize(w->stream);
    m3View& v = w->view;
    void* ptrs[] = {v.xf, v.lv, v.av, v.invMass, v.invI, v.inertia, v.lc, v.gravityScale,
                    v.linDamp, v.angDamp, v.bodyAlive, v.type, v.awake, v.sleepTime, v.bullet,
                    v.minExt, v.maxExt, v.force, v.torque, v.enabled, v.locks, v.sleepThreshold,
                    v.canSleep, v.hasTarget, v.target, v.shapeHead, v.shapeAlive, v.shapeBody,
                    v.shapeType, v.geom, v.friction, v.restitution, v.rolling, v.sensor,
                    v.hasOffset, v.localPos, v.localRot, v.surfaceVel, v.category, v.mask,
                    v.group, v.shapeNext, v.hullIndex, v.leaf, v.fresh, v.planes,
                    (void*)v.hulls, w->keysBuf[0], w->keysBuf[1], w->mansBuf[0], w->mansBuf[1],
                    v.sleepKeys, v.emit, w->emitSorted, w->emitUnique, v.cons, v.consPair,
                    v.colors, v.colorLists, v.movers, v.deltaPos, v.deltaRot, v.com0, v.rot0,
                    v.parent, v.forced, v.flag, v.pending, v.sc, w->sapKeysIn, w->sapKeysOut, w->sapValsIn,
                    w->sapValsOut, w->sapXY, w->sapZ, w->sapAttr, w->pairFlags, w->bodyFlags, w->root,
                    w->bodyMasks, w->bodyOwner, w->iota, v.slotColors, v.consBodies, v.slotOf,
                    v.slotTable, v.slotMask, w->barrier, w->cubTemp};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i)
    {
        if (ptrs[i] != nullptr)
        {
            (void)hipFree(ptrs[i]);
        }
    }
    (void)hipHostFree(w->hostScalars);
    ProfFree(w);
    (void)hipStreamDestroy(w->stream);
    free(w);
}

static int Copy(m3gpuWorld* w, void* dst, const void* src, int64_t bytes, hipMemcpyKind kind)
{
    if (bytes <= 0)
    {
        return 0;
    }
    HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, kind, w->stream));
    return 0;
}

#define UP(dev, host, count, type) \
    if (Copy(w, (dev), (host), (int64_t)(count) * (int64_t)sizeof(type), hipMemcpyHostToDevice)) return -1
#define DOWN(host, dev, count, type) \
    if (Copy(w, (host), (dev), (int64_t)(count) * (int64_t)sizeof(type), hipMemcpyDeviceToHost)) return -1

extern "C" int m3gpu_upload(m3gpuWorld* w, const m3gpuHostState* h, uint32_t fields)
{
    if (w == nullptr || h == nullptr)
    {
        return -1;
    }
    if (h->maxBody > w->caps.bodyCapacity || h->maxShape > w->caps.shapeCapacity ||
        h->maxHull > w->caps.hullCapacity)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_upload: counts exceed capacities");
        return -1;
    }
    m3View& v = w->view;
    v.maxBody = h->maxBody;
    v.maxShape = h->maxShape;
    int32_t nb = h->maxBody;
    int32_t ns = h->maxShape;
    if (fields & M3GPU_FIELDS_STATIC)
    {
        UP(v.invMass, h->invMass, nb, float);
        UP(v.invI, h->invInertiaLocal, nb, m3Mat3);
        UP(v.inertia, h->inertiaLocal, nb, m3Mat3);
        UP(v.lc, h->localCenters, nb, m3Vec3);
        UP(v.gravityScale, h->gravityScales, nb, float);
        UP(v.linDamp, h->linearDamping, nb, float);
        UP(v.angDamp, h->angularDamping, nb, float);
        UP(v.bodyAlive, h->bodyAlive, nb, uint8_t);
        UP(v.type, h->types, nb, uint8_t);
        UP(v.bullet, h->bulletFlags, nb, uint8_t);
        UP(v.minExt, h->minExtents, nb, float);
        UP(v.maxExt, h->maxExtents, nb, float);
        UP(v.enabled, h->bodyEnabled, nb, uint8_t);
        UP(v.locks, h->bodyLocks, nb, uint8_t);
        UP(v.sleepThreshold, h->bodySleepThreshold, nb, float);
        UP(v.canSleep, h->bodyCanSleep, nb, uint8_t);
        UP(v.target, h->bodyTarget, nb, m3Transform);
        UP(v.shapeHead, h->bodyShapeHead, nb, int32_t);
        UP(v.shapeAlive, h->shapeAlive, ns, uint8_t);
        UP(v.shapeBody, h->shapeBody, ns, int32_t);
        UP(v.shapeType, h->shapeType, ns, uint8_t);
        UP(v.geom, h->shapeGeom, ns, m3ShapeGeom);
        UP(v.friction, h->shapeFriction, ns, float);
        UP(v.restitution, h->shapeRestitution, ns, float);
        UP(v.rolling, h->shapeRollingResistance, ns, float);
        UP(v.sensor, h->shapeSensor, ns, uint8_t);
        UP(v.hasOffset, h->shapeHasOffset, ns, uint8_t);
        UP(v.localPos, h->shapeLocalPos, ns, m3Vec3);
        UP(v.localRot, h->shapeLocalRot, ns, m3Quat);
        UP(v.surfaceVel, h->shapeSurfaceVel, ns, m3Vec3);
        UP(v.category, h->shapeCategory, ns, uint64_t);
        UP(v.mask, h->shapeMask, ns, uint64_t);
        UP(v.group, h->shapeGroup, ns, int32_t);
        UP(v.shapeNext, h->shapeNext, ns, int32_t);
        UP(v.hullIndex, h->shapeHullIndex, ns, int32_t);
        UP((void*)v.hulls, h->hullData, h->maxHull, m3HullData);
        // The plane list and the unsupported-target census are host facts.
        const uint8_t* alive = h->shapeAlive;
        const uint8_t* stype = h->shapeType;
        int32_t* planes = (int32_t*)malloc((size_t)(ns > 0 ? ns : 1) * sizeof(int32_t));
        int32_t pc = 0;
        int unsupported = 0;
        for (int32_t s = 0; s < ns; ++s)
        {
            if (alive[s] == 0)
            {
                continue;
            }
            if (stype[s] == M3_PLANE)
            {
                planes[pc++] = s;
            }
            else if (stype[s] != M3_SPHERE && stype[s] != M3_HULL && stype[s] != M3_CAPSULE)
            {
                unsupported = 1;
            }
            else if (h->types[h->shapeBody[s]] == M3_STATIC)
            {
                unsupported = 1; // static tree targets for continuous sweeps
            }
        }
        for (int32_t b = 0; b < nb; ++b)
        {
            if (h->bodyAlive[b] != 0 && h->bulletFlags[b] != 0)
            {
                unsupported = 1;
            }
        }
        v.planeCount = pc;
        w->ccdUnsupported = unsupported;
        int rc = Copy(w, v.planes, planes, (int64_t)pc * 4, hipMemcpyHostToDevice);
        HIP_TRY(hipStreamSynchronize(w->stream));
        free(planes);
        if (rc)
        {
            return -1;
        }
    }
    if (fields & M3GPU_FIELDS_DYNAMIC)
    {
        UP(v.xf, h->transforms, nb, m3Transform);
        UP(v.lv, h->linearVelocities, nb, m3Vec3);
        UP(v.av, h->angularVelocities, nb, m3Vec3);
        UP(v.awake, h->awake, nb, uint8_t);
        UP(v.sleepTime, h->sleepTimes, nb, float);
        UP(v.force, h->bodyForce, nb, m3Vec3);
        UP(v.torque, h->bodyTorque, nb, m3Vec3);
        UP(v.hasTarget, h->bodyHasTarget, nb, uint8_t);
        UP(&v.sc->stepCount, h->stepCount, 1, uint64_t);
    }
    if (fields & M3GPU_FIELDS_PAIRS)
    {
        int32_t pc = *h->pairCount;
        int32_t zc = *h->sleepingPairCount;
        UP(v.keys, h->pairKeys, pc, uint64_t);
        UP(v.manifolds, h->manifolds, pc, m3Manifold);
        UP(&v.sc->pairCount, h->pairCount, 1, int32_t);
        UP(v.sleepKeys, h->sleepingPairKeys, zc, uint64_t);
        UP(&v.sc->sleepCount, h->sleepingPairCount, 1, int32_t);
        UP(v.leaf, h->leafBoxes, 6 * (int64_t)ns, double);
        w->fullQuery = *h->pairsFullQuery != 0;
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    return 0;
}

// A world whose substep barrier timed out holds no valid state.
static int CheckFault(m3gpuWorld* w)
{
    HIP_TRY(hipMemcpyAsync(&w->hostScalars->fault, &w->view.sc->fault, sizeof(int32_t),
                           hipMemcpyDeviceToHost, w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    if (w->hostScalars->fault != 0)
    {
        snprintf(s_lastError, sizeof(s_lastError), "substep grid barrier timed out");
        return -1;
    }
    return 0;
}

extern "C" int m3gpu_download(m3gpuWorld* w, m3gpuHostState* h, uint32_t fields)
{
    if (w == nullptr || h == nullptr)
    {
        return -1;
    }
    m3View& v = w->view;
    int32_t nb = v.maxBody;
    int32_t ns = v.maxShape;
    if (fields & M3GPU_FIELDS_DYNAMIC)
    {
        DOWN(h->transforms, v.xf, nb, m3Transform);
        DOWN(h->linearVelocities, v.lv, nb, m3Vec3);
        DOWN(h->angularVelocities, v.av, nb, m3Vec3);
        DOWN(h->awake, v.awake, nb, uint8_t);
        DOWN(h->sleepTimes, v.sleepTime, nb, float);
        DOWN(h->bodyForce, v.force, nb, m3Vec3);
        DOWN(h->bodyTorque, v.torque, nb, m3Vec3);
        DOWN(h->bodyHasTarget, v.hasTarget, nb, uint8_t);
        DOWN(h->stepCount, &v.sc->stepCount, 1, uint64_t);
    }
    if (fields & M3GPU_FIELDS_PAIRS)
    {
        DOWN(h->pairCount, &v.sc->pairCount, 1, int32_t);
        DOWN(h->sleepingPairCount, &v.sc->sleepCount, 1, int32_t);
        HIP_TRY(hipStreamSynchronize(w->stream));
        int32_t pc = *h->pairCount;
        int32_t zc = *h->sleepingPairCount;
        DOWN(h->pairKeys, v.keys, pc, uint64_t);
        DOWN(h->manifolds, v.manifolds, pc, m3Manifold);
        DOWN(h->sleepingPairKeys, v.sleepKeys, zc, uint64_t);
        DOWN(h->leafBoxes, v.leaf, 6 * (int64_t)ns, double);
        *h->frozenDirty = 0;
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    return CheckFault(w);
}

// Stage names for the profile (event-timed on the world's stream).
enum m3Stage
{
    ST_BEGIN, ST_FILL, ST_FRESH, ST_DISCOVER_ALL, ST_REFRESH, ST_SAP_SORT, ST_SAP_GATHER, ST_SAP, ST_PLANES,
    ST_FROZEN, ST_EMIT_SORT, ST_UNIQUE, ST_PAIR_COUNT, ST_EXPAND, ST_STALL_RESTORE,
    ST_SLEEP_SELECT, ST_NARROW, ST_CAPTURE, ST_UNION, ST_ROOT, ST_WAKE_MARK, ST_WAKE_APPLY,
    ST_CONS_FLAGS, ST_CONS_SELECT, ST_CONS_BODIES, ST_COLOR, ST_COLOR_LISTS, ST_PREPARE,
    ST_MOVER_FLAGS, ST_MOVER_SELECT, ST_SUBSTEPS, ST_CLEAR_FORCES, ST_RESTITUTION, ST_STORE, ST_CCD,
    ST_SLEEP_TIMERS, ST_SLEEP_READY, ST_SLEEP_FREEZE, ST_END,
    ST_COUNT
};

static const char* const kStageNames[] = {
    "begin", "fill_emit", "fresh_bounds", "discover_all", "refresh_proxies", "sap_sort", "sap_gather", "sap_sweep",
    "plane_pairs", "frozen_merge", "pair_sort", "pair_unique", "pair_count", "pair_expand",
    "stall_restore", "frozen_select", "narrowphase", "capture", "island_union", "island_root",
    "wake_mark", "wake_apply", "constraint_flags", "constraint_select", "constraint_bodies",
    "coloring", "color_lists", "prepare", "mover_flags", "mover_select", "substeps", "clear_forces",
    "restitution", "store_impulses", "continuous", "sleep_timers", "sleep_ready", "sleep_freeze",
    "end"};
static_assert(sizeof(kStageNames) / sizeof(kStageNames[0]) == ST_COUNT, "stage names out of sync");

struct m3Profiler
{
    int on;
    int pending; // events recorded in the current step
    hipEvent_t ev[2 * 1024];
    int stage[1024];
    double ms[ST_COUNT];
    int64_t launches[ST_COUNT];
    int64_t steps;
    // per-step counts summed over profiled steps (for the byte models)
    int64_t sumPairs, sumOldPairs, sumConstraints, sumPoints, sumMovers, sumSubsteps;
    int64_t sumShapes, sumBodies, sumOverflow;
};

static m3Profiler* Prof(m3gpuWorld* w);

// Records a launch between two events when the world is being profiled.
#define STAGE(id, launch)                                                                          \
    do                                                                                             \
    {                                                                                              \
        m3Profiler* pr_ = Prof(w);                                                                 \
        int slot_ = -1;                                                                            \
        if (pr_ != nullptr && pr_->pending < 1024)                                                 \
        {                                                                                          \
            slot_ = pr_->pending++;                                                                \
            pr_->stage[slot_] = (id);                                                              \
            HIP_TRY(hipEventRecord(pr_->ev[2 * slot_], st));                                       \
        }                                                                                          \
        launch;                                                                                    \
        if (slot_ >= 0)                                                                            \
        {                                                                                          \
            HIP_TRY(hipEventRecord(pr_->ev[2 * slot_ + 1], st));                                   \
        }                                                                                          \
    } while (0)

#define KL(...) hipLaunchKernelGGL(__VA_ARGS__)

// Enqueue one step.  Launch geometry only depends on host-known capacities
// and high-water marks; every data-dependent count is read on the device.
static int EnqueueStep(m3gpuWorld* w, const m3gpuStepParams* p)
{
    hipStream_t st = w->stream;
    // Swap the pair buffers: last step's list becomes the carry source.
    w->cur ^= 1;
    m3View& v = w->view;
    v.keys = w->keysBuf[w->cur];
    v.manifolds = w->mansBuf[w->cur];
    v.oldKeys = w->keysBuf[w->cur ^ 1];
    v.oldManifolds = w->mansBuf[w->cur ^ 1];
    const int T = 256;
    int32_t nb = v.maxBody;
    int32_t ns = v.maxShape;
    int32_t np = v.pairCapacity;
    int32_t ne = v.emitCapacity;
    size_t tb = w->cubBytes;
    void* ct = w->cubTemp;

    STAGE(ST_BEGIN, KL(k_begin, dim3(1), dim3(1), 0, st, v));
    STAGE(ST_FILL, KL(k_fill32, dim3(Blocks(ne, T)), dim3(T), 0, st, v.emit, ne, M3_SENTINEL32));
    // Broadphase.
    STAGE(ST_FRESH, KL(k_fresh, dim3(Blocks(ns, T)), dim3(T), 0, st, v, w->sapKeysIn, w->sapValsIn));
    if (w->fullQuery)
    {
        // After a restore the frozen buffer is rebuilt from every sleeping
        // body (src/broad_phase.c:432-448).
        KL(k_reset_sleep, dim3(1), dim3(1), 0, st, v);
        STAGE(ST_DISCOVER_ALL, KL(k_pending_all, dim3(Blocks(nb, T)), dim3(T), 0, st, v));
        w->fullQuery = 0;
    }
    STAGE(ST_SAP_SORT, HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
                           ct, tb, w->sapKeysIn, w->sapKeysOut, w->sapValsIn, w->sapValsOut, ns, 0,
                           64, st)));
    STAGE(ST_SAP_GATHER, KL(k_sap_gather, dim3(Blocks(ns, T)), dim3(T), 0, st, v, w->sapKeysOut,
                            w->sapValsOut, w->sapXY, w->sapZ, w->sapAttr));
    STAGE(ST_SAP, KL(k_sap, dim3(Blocks(ns, 64)), dim3(64), 0, st, v, w->sapXY, w->sapZ,
                     w->sapAttr));
    if (v.planeCount > 0)
    {
        STAGE(ST_PLANES,
              KL(k_planes, dim3(Blocks((int64_t)ns * v.planeCount, T)), dim3(T), 0, st, v));
    }
    STAGE(ST_REFRESH, KL(k_refresh, dim3(Blocks(ns, T)), dim3(T), 0, st, v));
    STAGE(ST_FROZEN, KL(k_frozen, dim3(Blocks(np, T)), dim3(T), 0, st, v));
    STAGE(ST_EMIT_SORT, HIP_TRY(hipcub::DeviceRadixSort::SortKeys(ct, tb, v.emit, w->emitSorted,
                                                                   ne, 0, 32, st)));
    STAGE(ST_UNIQUE, HIP_TRY(hipcub::DeviceSelect::Unique(ct, tb, w->emitSorted, w->emitUnique,
                                                          &v.sc->uniqueCount, ne, st)));
    STAGE(ST_PAIR_COUNT, KL(k_pair_count, dim3(1), dim3(1), 0, st, v, w->emitUnique));
    STAGE(ST_EXPAND,
          KL(k_expand, dim3(Blocks(np, T)), dim3(T), 0, st, v, w->emitUnique, w->pairFlags));
    STAGE(ST_STALL_RESTORE, KL(k_stall_restore, dim3(Blocks(np, T)), dim3(T), 0, st, v));
    STAGE(ST_SLEEP_SELECT, HIP_TRY(hipcub::DeviceSelect::Flagged(ct, tb, v.keys, w->pairFlags,
                                                                 v.sleepKeys, &v.sc->sleepCount,
                                                                 np, st)));
    // Narrowphase.
    STAGE(ST_NARROW, KL(k_narrow, dim3(Blocks(np, 64)), dim3(64), 0, st, v));
    // Sweeps, islands, wake.
    STAGE(ST_CAPTURE, KL(k_capture, dim3(Blocks(nb, T)), dim3(T), 0, st, v));
    STAGE(ST_UNION, KL(k_union, dim3(Blocks(np, T)), dim3(T), 0, st, v));
    STAGE(ST_ROOT, KL(k_root, dim3(Blocks(nb, T)), dim3(T), 0, st, v, w->root));
    STAGE(ST_WAKE_MARK, KL(k_wake_mark, dim3(Blocks(nb, T)), dim3(T), 0, st, v, w->root));
    STAGE(ST_WAKE_APPLY, KL(k_wake_apply, dim3(Blocks(nb, T)), dim3(T), 0, st, v, w->root));
    // Prepare and color.
    float h = p->dt / (float)p->substeps;
    float invH = h > 0.0f ? 1.0f / h : 0.0f;
    STAGE(ST_CONS_FLAGS,
          KL(k_constraint_flags, dim3(Blocks(np, T)), dim3(T), 0, st, v, w->pairFlags));
    hipcub::CountingInputIterator<int32_t> counting(0);
    STAGE(ST_CONS_SELECT, HIP_TRY(hipcub::DeviceSelect::Flagged(ct, tb, counting, w->pairFlags,
                                                                v.consPair,
                                                                &v.sc->constraintCount, np, st)));
    STAGE(ST_CONS_BODIES, KL(k_cons_bodies, dim3(Blocks(np, T)), dim3(T), 0, st, v));
    if (nb <= M3_COLOR_LDS_BODIES)
    {
        STAGE(ST_COLOR, KL(k_color<true>, dim3(1), dim3(M3_COLOR_THREADS), 0, st, v, w->bodyMasks,
                           w->bodyOwner));
    }
    else
    {
        STAGE(ST_COLOR, KL(k_color<false>, dim3(1), dim3(M3_COLOR_THREADS), 0, st, v,
                           w->bodyMasks, w->bodyOwner));
    }
    // Stable radix sort on the color byte: each color's list ascends.
    STAGE(ST_COLOR_LISTS, HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
                              ct, tb, v.colors, v.slotColors, w->iota, v.colorLists, np, 0, 5,
                              st)));
    STAGE(ST_PREPARE, KL(k_prepare, dim3(Blocks(np, 64)), dim3(64), 0, st, v, p->contactHertz,
                         p->contactDampingRatio, h));
    STAGE(ST_MOVER_FLAGS,
          KL(k_mover_flags, dim3(Blocks(nb, T)), dim3(T), 0, st, v, w->bodyFlags, p->dt));
    STAGE(ST_MOVER_SELECT, HIP_TRY(hipcub::DeviceSelect::Flagged(ct, tb, counting, w->bodyFlags,
                                                                 v.movers, &v.sc->moverCount, nb,
                                                                 st)));
    // Substeps: one cooperative launch.
    m3SubstepParams sp;
    sp.substeps = p->substeps;
    sp.h = h;
    sp.invH = invH;
    sp.push = p->contactPushMaxSpeed;
    sp.maxLinear = p->maximumLinearSpeed;
    sp.maxAngular = p->maximumAngularSpeed;
    sp.gravity = m3V3(p->gravity[0], p->gravity[1], p->gravity[2]);
    {
        void* args[] = {(void*)&v, (void*)&sp, (void*)&w->barrier};
        STAGE(ST_SUBSTEPS, HIP_TRY(hipLaunchCooperativeKernel((const void*)k_substeps,
                                                              dim3(w->coopBlocks), dim3(w->coopThreads), args,
                                                              0, st)));
    }
    STAGE(ST_CLEAR_FORCES, KL(k_clear_forces, dim3(Blocks(nb, T)), dim3(T), 0, st, v));
    STAGE(ST_RESTITUTION, KL(k_restitution, dim3(1), dim3(1), 0, st, v, p->restitutionThreshold));
    STAGE(ST_STORE, KL(k_store, dim3(Blocks(np, T)), dim3(T), 0, st, v));
    if (p->continuousEnabled)
    {
        STAGE(ST_CCD, KL(k_ccd, dim3(Blocks(nb, 64)), dim3(64), 0, st, v));
    }
    STAGE(ST_SLEEP_TIMERS, KL(k_sleep_timers, dim3(Blocks(nb, T)), dim3(T), 0, st, v, w->root,
                              p->dt, p->sleepEnabled));
    if (p->sleepEnabled)
    {
        STAGE(ST_SLEEP_READY, KL(k_sleep_ready, dim3(Blocks(nb, T)), dim3(T), 0, st, v, w->root));
        STAGE(ST_SLEEP_FREEZE, KL(k_sleep_freeze, dim3(Blocks(nb, T)), dim3(T), 0, st, v, w->root));
    }
    STAGE(ST_END, KL(k_end, dim3(1), dim3(1), 0, st, v));
    HIP_TRY(hipGetLastError());
    return 0;
}

// Waits for the profiled step, folds its event times and counts in.
static int ProfileCollect(m3gpuWorld* w, int32_t substeps)
{
    m3Profiler* pr = Prof(w);
    if (pr == nullptr)
    {
        return 0;
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    for (int k = 0; k < pr->pending; ++k)
    {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, pr->ev[2 * k], pr->ev[2 * k + 1]));
        pr->ms[pr->stage[k]] += ms;
        pr->launches[pr->stage[k]] += 1;
    }
    pr->pending = 0;
    HIP_TRY(hipMemcpy(w->hostScalars, w->view.sc, sizeof(m3Scalars), hipMemcpyDeviceToHost));
    const m3Scalars* s = w->hostScalars;
    pr->steps += 1;
    pr->sumPairs += s->pairCount;
    pr->sumOldPairs += s->oldPairCount;
    pr->sumConstraints += s->constraintCount;
    pr->sumPoints += s->pointCount;
    pr->sumMovers += s->moverCount;
    pr->sumSubsteps += substeps;
    pr->sumShapes += w->view.maxShape;
    pr->sumBodies += w->view.maxBody;
    pr->sumOverflow += s->colorStarts[M3_GRAPH_COLORS + 1] - s->colorStarts[M3_GRAPH_COLORS];
    return 0;
}

extern "C" int m3gpu_step(m3gpuWorld* w, const m3gpuStepParams* p, m3gpuStepStats* stats)
{
    if (w == nullptr || p == nullptr || !(p->dt > 0.0f) || p->substeps < 1 ||
        p->substeps > M3_MAX_SUBSTEPS)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_step: bad arguments");
        return -1;
    }
    if (p->continuousEnabled && w->ccdUnsupported)
    {
        snprintf(s_lastError, sizeof(s_lastError),
                 "m3gpu_step: continuous collision against static non-plane shapes or bullets "
                 "is not implemented on the device path");
        return -1;
    }
    if (EnqueueStep(w, p) != 0)
    {
        return -1;
    }
    if (ProfileCollect(w, p->substeps) != 0)
    {
        return -1;
    }
    if (stats != nullptr)
    {
        return m3gpu_last_stats(w, stats) != 0 ? -1 : (w->hostScalars->stall ? M3GPU_STALL : 0);
    }
    return 0;
}

extern "C" int m3gpu_step_many(m3gpuWorld* const* worlds, const m3gpuStepParams* params,
                               int32_t count)
{
    for (int32_t k = 0; k < count; ++k)
    {
        if (m3gpu_step(worlds[k], &params[k], nullptr) < 0)
        {
            return -1;
        }
    }
    return 0;
}

extern "C" int m3gpu_sync(m3gpuWorld* w)
{
    if (w == nullptr)
    {
        return -1;
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    return CheckFault(w);
}

extern "C" int m3gpu_last_stats(m3gpuWorld* w, m3gpuStepStats* stats)
{
    if (w == nullptr || stats == nullptr)
    {
        return -1;
    }
    HIP_TRY(hipMemcpyAsync(w->hostScalars, w->view.sc, sizeof(m3Scalars), hipMemcpyDeviceToHost,
                           w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    const m3Scalars* s = w->hostScalars;
    memset(stats, 0, sizeof(*stats));
    stats->pairCount = s->pairCount;
    stats->constraintCount = s->constraintCount;
    stats->colorCount = s->colorCount;
    stats->islandCount = s->islandCount;
    stats->awakeCount = s->moverCount;
    return 0;
}

static m3Profiler* Prof(m3gpuWorld* w) { return (w->prof != nullptr && w->prof->on) ? w->prof : nullptr; }

extern "C" int m3gpu_profile_begin(m3gpuWorld* w)
{
    if (w == nullptr)
    {
        return -1;
    }
    if (w->prof == nullptr)
    {
        w->prof = (m3Profiler*)calloc(1, sizeof(m3Profiler));
        if (w->prof == nullptr)
        {
            return -1;
        }
        for (int k = 0; k < 2 * 1024; ++k)
        {
            HIP_TRY(hipEventCreate(&w->prof->ev[k]));
        }
    }
    m3Profiler* pr = w->prof;
    memset(pr->ms, 0, sizeof(pr->ms));
    memset(pr->launches, 0, sizeof(pr->launches));
    pr->steps = pr->sumPairs = pr->sumOldPairs = pr->sumConstraints = pr->sumPoints = 0;
    pr->sumMovers = pr->sumSubsteps = pr->sumShapes = pr->sumBodies = pr->sumOverflow = 0;
    pr->pending = 0;
    pr->on = 1;
    return 0;
}

extern "C" int m3gpu_profile_end(m3gpuWorld* w, m3gpuProfile* out)
{
    if (w == nullptr || out == nullptr || w->prof == nullptr)
    {
        return -1;
    }
    m3Profiler* pr = w->prof;
    pr->on = 0;
    memset(out, 0, sizeof(*out));
    int n = 0;
    for (int k = 0; k < ST_COUNT && n < M3GPU_MAX_STAGES; ++k)
    {
        if (pr->launches[k] == 0)
        {
            continue;
        }
        strncpy(out->stage[n].name, kStageNames[k], sizeof(out->stage[n].name) - 1);
        out->stage[n].launches = pr->launches[k];
        out->stage[n].ms = pr->ms[k];
        n += 1;
    }
    out->stageCount = n;
    out->steps = pr->steps;
    out->sumPairs = pr->sumPairs;
    out->sumOldPairs = pr->sumOldPairs;
    out->sumConstraints = pr->sumConstraints;
    out->sumPoints = pr->sumPoints;
    out->sumMovers = pr->sumMovers;
    out->sumSubsteps = pr->sumSubsteps;
    out->sumShapes = pr->sumShapes;
    out->sumBodies = pr->sumBodies;
    out->sumOverflow = pr->sumOverflow;
    return 0;
}

static void ProfFree(m3gpuWorld* w)
{
    if (w->prof == nullptr)
    {
        return;
    }
    for (int k = 0; k < 2 * 1024; ++k)
    {
        (void)hipEventDestroy(w->prof->ev[k]);
    }
    free(w->prof);
    w->prof = nullptr;
}
