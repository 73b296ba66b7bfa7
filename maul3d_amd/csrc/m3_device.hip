// gfx950 step for the soft-step contact solver: the m3gpu_* C ABI
// (include/m3gpu.h) over a device-resident, structure-of-arrays world.
//
// One step is a fixed sequence of kernels on the world's stream with no
// host round trip: every count the later stages need (pairs, constraints,
// movers, colors) stays in device memory (m3Scalars) and kernels read it.
// Stage order and semantics follow m3StepInternal (src/solver.c:3662-4320).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <type_traits>
#include <vector>

#include "m3gpu.h"
#include "m3_adjoint.h"
#include "m3_adj64.h"

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


// What a kernel sees of one world: the device view plus the step's scratch.
// Kernels take an array of these and work on entry blockIdx.y, so one
// launch serves one world or a batch of independent worlds alike.
// A shape's sweep key: the sortable bits of its lower x bound rounded down
// to float (the value the sweep compares) over its index; shapes out of
// the sweep (dead, planes) carry an all-ones high word and sort last.
#define M3_SAP_OUT 0xFFFFFFFFu

struct m3Dev
{
    m3View v;
    uint64_t* sapKeysIn;  // per shape: its sweep key (k_open); k_sap_sort leaves its sorted runs here
    uint64_t* sapKeysOut; // the keys sorted
    float4* sapXY;
    float2* sapZ;
    int4* sapAttr;
    uint32_t* emitUnique;
    // Pair-key grouping by lower shape (k_emit_pairs): per shape the count,
    // start, fill cursor, unique count and unique start of its group, the
    // groups too large for one thread and their count; per key its group.
    int32_t* grp;
    uint32_t* grpKeys;
    int32_t* grpOf;
    uint8_t* pairFlags;
    uint8_t* bodyFlags;
    int32_t* blockCounts; // set flags per 256-thread workgroup of the last flag pass
    int32_t* root;
    uint16_t* bodyMasks;
    int32_t* bodyOwner;
    uint32_t* barrier; // substep grid barrier: arrival count, generation
    // Barrier-free passes of the single-world substep kernel: per body six
    // versioned velocity words, per slot the constraint's ranks on its two
    // bodies (see SubstepCells).
    uint64_t* flowRec;
    uint32_t* flowRank;
    m3WarmAdd* warmAdd; // per slot and side: the next warm start's addends (barrier-free form)
    uint64_t* deltaRec; // per body: versioned position and rotation deltas
    // Cells of the barrier-free form: per mover its cell and index there,
    // per cell its movers and counts, per slot the solving cell.
    uint8_t* cellOf;
    uint16_t* cellLocal;
    int32_t* cellBodies;
    int32_t* cellCount;
    int32_t* cellConsCount;
    int32_t* cellCons; // per cell: its slots (any order)
    uint8_t* slotOwner;
    int2* slotBodies; // per slot: its awake dynamic bodies (or -1), for the cell owners
    uint32_t* cellDone; // per cell: the substeps whose relax pass (and addends) it has finished
    m3Tape* tape;      // intermediate-value record of this step (adjoint replays), or null
    // Pairs binned by collision routine: 8 counts, then one bin of
    // pairCapacity entries per routine.
    int32_t* narrowBins;
    // Continuous pass with bullets: the pulled-back pose of each body and
    // whether it was pulled back, before they are applied.
    m3Transform* ccdXf;
    uint8_t* ccdHit;
    int32_t ccdBullets; // some bullet body exists (the three-pass form)
};

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
    int ccdNeedsTree; // continuous sweeps meet tree shapes (static non-plane targets or bullets)
    int anyRestitution; // some live shape has restitution (else the restitution pass is not launched)
    // pre-solve: candidate records and their count, the vetoed keys and
    // per-pair flags of the step, pinned host staging
    m3gpuPreSolve* preCands;
    int32_t* preCount;
    uint64_t* vetoKeys;
    uint8_t* vetoFlags;
    m3gpuPreSolve* preStage;
    uint64_t* vetoStage;
    hipEvent_t preGathered;
    // A pinned stage is rewritten only after the copy that reads it has run:
    // the host may enqueue the next step before the last one started.
    hipEvent_t vetoCopied;
    int vetoCopyPending;
    // the pair table in the reference's layout (m3gpu_set_pair_table)
    uint64_t* tableKeys;
    m3Manifold* tableMans;
    int pairTable;
    int32_t* rankStage; // pinned staging of the tree ranks a step carries
    double* leafStage;  // pinned copy of the step's leaf boxes (m3gpu_step_leaves)
    hipEvent_t leavesCopied;
    hipEvent_t ranksCopied;
    int rankCopyPending;
    m3Dev dev;      // host copy of what kernels see (scratch pointers, view)
    m3Dev* devDev;  // device copies: [parity] = pair buffers for a step with cur = parity
    int devDirty;   // the device copies need refreshing
    int64_t layout; // bumped whenever the device copies change
    int streamBusy; // work enqueued on the world's stream since it was last idle
    int batched;    // the world's latest device work ran in the batch stream
    int64_t stamp;  // duplicate detection in batched calls
    m3Scalars* hostScalars; // pinned mirror for stats
    struct m3Profiler* prof;
    int coopBlocks;  // co-resident blocks of the substep kernel (maximum)
    int coopCompact; // co-resident 256-thread blocks of k_islands, k_constraints and k_movers
    int coopFixed;   // coopBlocks set from the environment or tuning: launch exactly that many
    int coopDefault; // coopBlocks as created (restored when tuning is cleared)
    int coopMax;     // co-resident limit of the cooperative launch
    int coopEnvFixed; // MAUL3D_AMD_SUBSTEP_BLOCKS was set at creation
    int tailLimit;   // tuned budget of the workgroup-0 tail (0: M3_TAIL_CONSTRAINTS)
    int flow;        // barrier-free solve and relax passes (MAUL3D_AMD_FLOW=0 turns them off)
    int cellsX;      // cells of the barrier-free passes: slabs along x
    int cellsZ;      // and cells per slab along z
    int coopThreads; // threads per block of the substep kernel
    int wideHulls;   // some hull is larger than a box (wide hull pairs possible)
    void* cubTemp;       // temp storage of the sweep-key radix sort
    size_t cubBytes;
    m3HitRec* hitRecBuf; // per pair; in the view only while some shape has hit events
    struct m3EventBufs* events; // event classification and output (allocated at first read)
    int boxHulls;    // some hull is at least box-sized
    int smallHulls; // every hull within SatLdsBox (k_sat's small LDS layout)
};

// The entry is read through the constant address space: its fields load as
// invariant scalars and the pointers in it are known to address global
// memory (as for a kernel argument), so accesses through them compile to
// global rather than flat memory operations.
// (The host compilation pass of kernel bodies takes the plain read.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) m3Dev m3ConstDev;
#define M3_WORLD_ENTRY (((m3ConstDev*)ds)[blockIdx.y])
#else
#define M3_WORLD_ENTRY (ds[blockIdx.y])
#endif

#define M3_WORLD                                                                                   \
    const m3Dev d = M3_WORLD_ENTRY;                                                                \
    const m3View& v = d.v;                                                                         \
    (void)d

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------

// Rank of a constraint side that is not an awake dynamic body (barrier-free
// substep passes).
#define M3_FLOW_NONE 0xFFu
#define M3_FLOW_STALE 0xFFFFFFFEu // a record version no pass writes
// Per body, the position and rotation deltas integrate-positions left
// (seven versioned words, one unused): the passes that follow read them
// there, since no grid barrier separates them from the writer.
#define M3_DELTA_WORDS 8
// Solve and relax passes of one step the versions cover (two per substep;
// more substeps take the colored path).
#define M3_FLOW_PASSES 16

static inline unsigned Blocks(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b > 0 ? b : 1);
}

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

__device__ static inline void StepBegin(const m3Dev& d)
{
    const m3View& v = d.v;
    m3Scalars* sc = v.sc;
    sc->oldPairCount = sc->pairCount;
    sc->emitCount = 0;
    sc->stall = 0;
    sc->anyRestitution = 0;
    sc->islandCount = 0;
    sc->awakeCount = 0;
    sc->discoverCount = 0;
    sc->pointCount = 0;
    sc->jointActive = 0;
    sc->satCount = 0;
    sc->flowVeto = 0;
    sc->flowKinematic = 0;
    if (d.narrowBins != nullptr)
    {
        for (int c = 0; c < 8; ++c)
        {
            d.narrowBins[c] = 0;
        }
    }
}

__device__ static inline void FreshShape(const m3Dev& d, int32_t s);

// The step's opening: thread 0 resets the step's scalars, every thread
// fills its slot of the emission buffer with the sentinel, and each shape's
// fresh box and sweep key are computed (none of these reads another).
__global__ void k_open(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
    {
        StepBegin(d);
    }
    if (i < v.maxShape)
    {
        FreshShape(d, i);
    }
    if (i < v.emitCapacity)
    {
        v.emit[i] = M3_SENTINEL32;
    }
    // The pair-key groups' counters start at zero (k_emit_pairs).
    const int32_t stride = v.maxShape + 1;
    if (i < v.maxShape)
    {
        d.grp[i] = 0;
        d.grp[2 * stride + i] = 0;
    }
    if (i == 0)
    {
        d.grp[6 * stride] = 0;
    }
}

// ---------------------------------------------------------------------------
// Ordered compaction.  A flag pass leaves the number of set flags of each
// 256-thread workgroup in blockCounts (FlagCount); the matching write pass,
// launched with the same grid, gives each workgroup its output offset by
// summing the counts of the workgroups before it and writes its flagged
// items in index order (FlagRank), the last workgroup the total.
// ---------------------------------------------------------------------------

#define M3_COMPACT_THREADS 256
// Workgroups of a cooperative compaction (k_islands, k_constraints,
// k_movers): the grid barrier's arrivals serialise on one counter.
#define M3_COOP_COMPACT_BLOCKS 64

// Every thread of the workgroup calls it.
__device__ static inline void FlagCount(int flag, int32_t* blockCounts, int32_t chunk)
{
    __shared__ int32_t waveCount[M3_COMPACT_THREADS / 64];
    const uint64_t mask = __ballot(flag);
    const int32_t lane = threadIdx.x & 63;
    const int32_t wave = threadIdx.x >> 6;
    if (lane == 0)
    {
        waveCount[wave] = __popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int32_t n = 0;
        for (int32_t k = 0; k < M3_COMPACT_THREADS / 64; ++k)
        {
            n += waveCount[k];
        }
        blockCounts[chunk] = n;
    }
}

// This thread's output position if its flag is set (every thread calls it);
// *total receives the count over all workgroups.
__device__ static inline int32_t FlagRank(int flag, const int32_t* blockCounts, int32_t chunk, int32_t* total)
{
    __shared__ int32_t partial[M3_COMPACT_THREADS];
    __shared__ int32_t waveCount[M3_COMPACT_THREADS / 64];
    const int32_t t = threadIdx.x;
    int32_t sum = 0;
    for (int32_t k = t; k < chunk; k += M3_COMPACT_THREADS)
    {
        sum += blockCounts[k];
    }
    partial[t] = sum;
    const uint64_t mask = __ballot(flag);
    const int32_t lane = t & 63;
    const int32_t wave = t >> 6;
    if (lane == 0)
    {
        waveCount[wave] = __popcll(mask);
    }
    __syncthreads();
    for (int32_t half = M3_COMPACT_THREADS / 2; half > 0; half >>= 1)
    {
        if (t < half)
        {
            partial[t] += partial[t + half];
        }
        __syncthreads();
    }
    int32_t pos = partial[0];
    for (int32_t k = 0; k < wave; ++k)
    {
        pos += waveCount[k];
    }
    pos += __popcll(mask & ((1ull << lane) - 1ull));
    int32_t mine = 0;
    for (int32_t k = 0; k < M3_COMPACT_THREADS / 64; ++k)
    {
        mine += waveCount[k];
    }
    *total = partial[0] + mine;
    return pos;
}

// Fresh bounds of every alive non-plane shape and its sweep key.
__device__ static inline void FreshShape(const m3Dev& d, int32_t s)
{
    const m3View& v = d.v;
    uint32_t u = M3_SAP_OUT;
    if (v.shapeAlive[s] != 0 && v.shapeType[s] != M3_PLANE)
    {
        double box[6];
        m3FreshBox(&v, s, box);
        for (int k = 0; k < 6; ++k)
        {
            v.fresh[6 * s + k] = box[k];
        }
        u = __float_as_uint(RoundDown(box[0]));
        u ^= (u & 0x80000000u) != 0u ? 0xFFFFFFFFu : 0x80000000u;
        u = u < M3_SAP_OUT ? u : M3_SAP_OUT - 1u;
    }
    d.sapKeysIn[s] = ((uint64_t)u << 32) | (uint32_t)s;
}

__global__ void k_reset_sleep(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    if (v.sc->fullQuery != 0)
    {
        v.sc->sleepCount = 0;
    }
}

// After a restore every sleeping body rediscovers its frozen pairs.
__global__ void k_pending_all(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= v.maxBody || v.sc->fullQuery == 0)
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

// Outward-rounded float copies of the sorted fresh boxes: a float test that
// fails implies the exact double test fails, so the sweep filters on floats
// and confirms survivors in double.

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

__global__ void k_sap_gather(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    const uint64_t* sorted = d.sapKeysOut;
    float4* xy = d.sapXY;
    float2* zz = d.sapZ;
    int4* attr = d.sapAttr;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxShape)
    {
        return;
    }
    if ((uint32_t)(sorted[i] >> 32) == M3_SAP_OUT)
    {
        xy[i] = make_float4(INFINITY, -INFINITY, INFINITY, -INFINITY);
        zz[i] = make_float2(INFINITY, -INFINITY);
        attr[i] = make_int4(-1, -1, 0, 0);
        return;
    }
    int32_t s = (int32_t)(sorted[i] & 0xFFFFFFFFu);
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
    fl |= (v.category[s] == 1ull && v.mask[s] == ~0ull && v.group[s] == 0 &&
           v.bodyJointHead[b] == -1)
              ? SA_DEFAULT_FILTER
              : 0;
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
// The workgroup's shapes and the sorted entries after them are staged in
// LDS: a shape's candidates are the entries that follow it in the x order,
// so neighbouring shapes scan largely the same window.
#define M3_SAP_THREADS 64
#define M3_SAP_WINDOW 1536

// The sweep.  Each lane first collects, from the x-sorted entries in LDS,
// the candidates whose float sweep boxes overlap its shape's (cheap, no
// global loads), then the lanes test their candidates in step: the full
// double-precision tests and filters load from global memory, and doing
// them as they turn up would make every iteration of the wavefront's scan
// wait on some lane's loads.
#define M3_SAP_CAND 32
// Lanes per shape: the lanes of a shape scan interleaved entries of its
// sweep range (entry i+1+k, i+1+k+M3_SAP_SPLIT, ...), which shortens the
// serial scan and puts more wavefronts on the chip.  Each overlapping pair
// is still found by exactly one lane.
#define M3_SAP_SPLIT 4
#define M3_SAP_SHAPES (M3_SAP_THREADS / M3_SAP_SPLIT)

__global__ void __launch_bounds__(M3_SAP_THREADS) k_sap(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    __shared__ float4 sxy[M3_SAP_WINDOW];
    __shared__ float2 szz[M3_SAP_WINDOW];
    __shared__ int32_t cand[M3_SAP_CAND][M3_SAP_THREADS];
    const float4* xy = d.sapXY;
    const float2* zz = d.sapZ;
    const int4* attr = d.sapAttr;
    const int32_t base = blockIdx.x * M3_SAP_SHAPES;
    const int32_t ns = v.maxShape;
    const int32_t wend = base + M3_SAP_WINDOW < ns ? base + M3_SAP_WINDOW : ns;
    for (int32_t j = base + threadIdx.x; j < wend; j += blockDim.x)
    {
        sxy[j - base] = xy[j];
        szz[j - base] = zz[j];
    }
    __syncthreads();
    const int lane = threadIdx.x;
    const int32_t i = base + lane / M3_SAP_SPLIT;
    int4 me = make_int4(-1, 0, 0, 0);
    if (i < ns)
    {
        me = attr[i];
    }
    bool done = !(i < ns && me.x >= 0);
    float4 mxy = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float2 mz = make_float2(0.0f, 0.0f);
    const double* a = nullptr;
    if (!done)
    {
        mxy = sxy[i - base];
        mz = szz[i - base];
        a = &v.fresh[6 * me.x];
    }
    int32_t j = i + 1 + lane % M3_SAP_SPLIT;
    m3EmitBuf buf;
    buf.n = 0;
    while (true)
    {
        int32_t nc = 0;
        if (!done)
        {
            for (; j < wend && nc < M3_SAP_CAND; j += M3_SAP_SPLIT)
            {
                float4 oxy = sxy[j - base];
                if (oxy.x > mxy.y) // sentinels have +inf lower x: the sweep ends
                {
                    done = true;
                    break;
                }
                if (oxy.z > mxy.w || mxy.z > oxy.w)
                {
                    continue;
                }
                float2 oz = szz[j - base];
                if (oz.x > mz.y || mz.x > oz.y)
                {
                    continue;
                }
                cand[nc++][lane] = j;
            }
            if (!done && j >= wend)
            {
                for (; j < ns && nc < M3_SAP_CAND; j += M3_SAP_SPLIT)
                {
                    float4 oxy = xy[j];
                    if (oxy.x > mxy.y)
                    {
                        done = true;
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
                    cand[nc++][lane] = j;
                }
            }
            if (j >= ns)
            {
                done = true;
            }
        }
        for (int32_t k = 0; k < nc; ++k)
        {
            const int32_t jj = cand[k][lane];
            const int4 ot = attr[jj];
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
        if (__all(done))
        {
            break;
        }
    }
    EmitFlush(v, buf);
}

// Planes pair with every allowed hot shape (src/broad_phase.c:541-564).

// Still-cold frozen pairs ride along (src/broad_phase.c:607-632).
__global__ void k_frozen(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
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
// the frozen buffer carried to the next step.  On a capacity stall the
// previous pair list stays the world's list.  Every thread derives the
// count and the stall from scalars that no thread of this pass changes
// (thread 0 only publishes them), then expands its key or restores its old
// pair, flags the cold keys and counts them per workgroup.

// The cold keys of the pair list, in order, become the frozen buffer.
__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_compact_frozen(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int flag = i < v.pairCapacity && d.pairFlags[i] != 0;
    int32_t total;
    const int32_t pos = FlagRank(flag, d.blockCounts, blockIdx.x, &total);
    if (flag)
    {
        v.sleepKeys[pos] = v.keys[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    {
        v.sc->sleepCount = total;
    }
}

// ---------------------------------------------------------------------------
// Narrowphase (src/manifold.c:2357-2944)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Wide hull pairs: the separating-axis queries one wavefront per pair
// ---------------------------------------------------------------------------
//
// A hull pair's face queries cost faces x vertices and its edge query
// (edges of A) x (edges of B) tests, up to 186 x 186 for two 64-vertex hulls:
// milliseconds for one thread.  Pairs above box-against-box size are listed
// first, then each listed pair's three queries are evaluated by a 64-lane
// workgroup: lanes split the faces and the edge pairs, and a reduction
// picks what the serial scan with its early exits would return:
//   * the first candidate (scan order) whose separation exceeds the
//     speculative distance, where the scan returns at once, else
//   * the first candidate attaining the largest separation (updates are
//     strict), else the scan's initial value.
// A face's inner scan may stop early in the serial form only when its value
// could no longer be selected, so the full minimum over the vertices selects
// identically.  Every candidate value comes from the same operations as the
// serial scan, so the selected separation and axis are bit-identical.

__device__ static inline void RefreshShape(const m3View& v, int32_t s)
{
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

// Shape-plane pairs (idx = plane slot * maxShape + shape) and, for
// idx < maxShape, the shape's proxy refresh: neither reads the other's
// output.
__global__ void k_planes(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
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
    if (idx < v.maxShape)
    {
        RefreshShape(v, (int32_t)idx);
    }
}


#define M3_HULL_MAX_EDGES (M3_HULL_MAX_HALF_EDGES / 2)

template <int kVerts, int kEdges>
struct SatLdsT
{
    m3Vec3 w[kVerts]; // the other hull's vertices in the reference hull's frame
    m3Vec3 aP[kEdges]; // A's edges: origin, direction, the two face normals
    m3Vec3 aE[kEdges];
    m3Vec3 aU[kEdges];
    m3Vec3 aV[kEdges];
    m3Vec3 bP[kEdges]; // B's edges, moved into A's frame
    m3Vec3 bE[kEdges];
    m3Vec3 bU[kEdges];
    m3Vec3 bV[kEdges];
};
// Any hull (18.6 KB: eight workgroups per CU), and worlds whose hulls all
// have at most 8 vertices and 12 edges, boxes (1.2 KB).
typedef SatLdsT<M3_HULL_MAX_VERTS, M3_HULL_MAX_EDGES> SatLds;
typedef SatLdsT<8, 12> SatLdsBox;
#define M3_SAT_BOX_VERTS 8
#define M3_SAT_BOX_HALF_EDGES 24

// (separation, index) of the serial scan's choice among this lane's and the
// other lanes' candidates: first exceeding index if any, else first maximum.
struct SatPick
{
    int32_t hit;  // smallest candidate index above the speculative distance
    m3real hitSep;
    int32_t best; // first candidate index of the largest separation
    m3real bestSep;
};

__device__ inline void SatOffer(SatPick& s, int32_t idx, m3real sep)
{
    // Candidates arrive in ascending index per lane.
    if (sep > M3_SPECULATIVE_DISTANCE && s.hit == INT32_MAX)
    {
        s.hit = idx;
        s.hitSep = sep;
    }
    if (sep > s.bestSep)
    {
        s.bestSep = sep;
        s.best = idx;
    }
}

__device__ inline SatPick SatReduce(SatPick s)
{
    for (int off = 32; off > 0; off >>= 1)
    {
        int32_t oHit = __shfl_xor(s.hit, off, 64);
        m3real oHitSep = __shfl_xor(s.hitSep, off, 64);
        int32_t oBest = __shfl_xor(s.best, off, 64);
        m3real oBestSep = __shfl_xor(s.bestSep, off, 64);
        if (oHit < s.hit)
        {
            s.hit = oHit;
            s.hitSep = oHitSep;
        }
        if (oBestSep > s.bestSep || (oBestSep == s.bestSep && oBest < s.best))
        {
            s.bestSep = oBestSep;
            s.best = oBest;
        }
    }
    return s;
}

__device__ inline SatPick SatPickInit()
{
    SatPick s;
    s.hit = INT32_MAX;
    s.hitSep = 0.0f;
    s.best = INT32_MAX;
    s.bestSep = -3.4e38f;
    return s;
}

// m3QueryFaces over the lanes (same values, same choice).
template <class Lds>
__device__ m3FaceQuery WaveQueryFaces(const m3HullData* ref, const m3HullData* other, m3Quat q,
                                      m3Vec3 p, int refIsA, Lds& L, int lane)
{
    int32_t nv = other->vertexCount;
    for (int32_t k = lane; k < nv; k += 64)
    {
        L.w[k] = refIsA ? m3Add3(m3RotateVec3(q, other->vertices[k]), p)
                        : m3InvRotateVec3(q, m3Sub3(other->vertices[k], p));
    }
    __syncthreads();
    SatPick s = SatPickInit();
    int32_t nf = ref->faceCount;
    for (int32_t f = lane; f < nf; f += 64)
    {
        m3Vec3 n = ref->faceNormals[f];
        m3real off = ref->faceOffsets[f];
        m3real best = 3.4e38f;
        for (int32_t k = 0; k < nv; ++k)
        {
            best = m3MinF(best, m3Dot3(n, L.w[k]) - off);
        }
        SatOffer(s, f, best);
    }
    s = SatReduce(s);
    __syncthreads();
    m3FaceQuery out;
    if (s.hit != INT32_MAX)
    {
        out.separation = s.hitSep;
        out.faceIndex = s.hit;
    }
    else if (s.best != INT32_MAX)
    {
        out.separation = s.bestSep;
        out.faceIndex = s.best;
    }
    else
    {
        out.separation = -3.4e38f;
        out.faceIndex = 0;
    }
    return out;
}

// Separation of edge pair (ja, jb) as m3QueryEdges computes it; false when
// the Gauss-map test or the degenerate-axis test rejects the pair.
template <class Lds>
__device__ inline bool SatEdgePair(const Lds& L, int32_t ja, int32_t jb, m3Vec3 centerA,
                                   m3real* sepOut, m3Vec3* axisOut)
{
    m3Vec3 eA = L.aE[ja];
    m3Vec3 eB = L.bE[jb];
    m3real cba = m3Dot3(L.bU[jb], eA);
    m3real dba = m3Dot3(L.bV[jb], eA);
    m3real adc = -m3Dot3(L.aU[ja], eB);
    m3real bdc = -m3Dot3(L.aV[ja], eB);
    if (!(cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f))
    {
        return false;
    }
    m3Vec3 axis = m3Cross3(eA, eB);
    m3real len = m3Length3(axis);
    if (len < 1.0e-6f)
    {
        return false;
    }
    axis = m3MulSV3(1.0f / len, axis);
    m3Vec3 pA = L.aP[ja];
    if (m3Dot3(axis, m3Sub3(pA, centerA)) < 0.0f)
    {
        axis = m3Neg3(axis);
    }
    *sepOut = m3Dot3(axis, m3Sub3(L.bP[jb], pA));
    *axisOut = axis;
    return true;
}

// m3QueryEdges over the lanes: candidates in the serial scan order
// t = (B edge) * (A edges) + (A edge).
template <class Lds>
__device__ m3EdgeQuery WaveQueryEdges(const m3HullData* hullA, const m3HullData* hullB, m3Quat q,
                                      m3Vec3 p, Lds& L, int lane)
{
    int32_t nA = hullA->edgeCount >> 1;
    int32_t nB = hullB->edgeCount >> 1;
    for (int32_t k = lane; k < nA; k += 64)
    {
        m3HullHalfEdge edgeA = hullA->edges[2 * k];
        m3HullHalfEdge twinA = hullA->edges[2 * k + 1];
        m3Vec3 pA = hullA->vertices[edgeA.origin];
        L.aP[k] = pA;
        L.aE[k] = m3Sub3(hullA->vertices[twinA.origin], pA);
        L.aU[k] = hullA->faceNormals[edgeA.face];
        L.aV[k] = hullA->faceNormals[twinA.face];
    }
    for (int32_t k = lane; k < nB; k += 64)
    {
        m3HullHalfEdge edgeB = hullB->edges[2 * k];
        m3HullHalfEdge twinB = hullB->edges[2 * k + 1];
        m3Vec3 qB = m3Add3(m3RotateVec3(q, hullB->vertices[twinB.origin]), p);
        m3Vec3 pB = m3Add3(m3RotateVec3(q, hullB->vertices[edgeB.origin]), p);
        L.bP[k] = pB;
        L.bE[k] = m3Sub3(qB, pB);
        L.bU[k] = m3RotateVec3(q, hullB->faceNormals[edgeB.face]);
        L.bV[k] = m3RotateVec3(q, hullB->faceNormals[twinB.face]);
    }
    __syncthreads();
    m3Vec3 centerA = hullA->center;
    SatPick s = SatPickInit();
    int32_t total = nA * nB;
    int32_t jb = lane / nA;
    int32_t ja = lane - jb * nA;
    for (int32_t t = lane; t < total; t += 64)
    {
        m3real sep;
        m3Vec3 axis;
        if (SatEdgePair(L, ja, jb, centerA, &sep, &axis))
        {
            SatOffer(s, t, sep);
        }
        ja += 64;
        while (ja >= nA)
        {
            ja -= nA;
            jb += 1;
        }
    }
    s = SatReduce(s);
    m3EdgeQuery out;
    out.separation = -3.4e38f;
    out.indexA = -1;
    out.indexB = -1;
    out.axis = m3V3(0.0f, 1.0f, 0.0f);
    int32_t pick = s.hit != INT32_MAX ? s.hit : s.best;
    if (pick != INT32_MAX)
    {
        int32_t pb = pick / nA;
        int32_t pa = pick - pb * nA;
        m3real sep;
        m3Vec3 axis;
        SatEdgePair(L, pa, pb, centerA, &sep, &axis);
        out.separation = sep;
        out.indexA = 2 * pa;
        out.indexB = 2 * pb;
        out.axis = axis;
    }
    __syncthreads();
    return out;
}

// The face and edge queries of a capsule segment (p1, p2 in the hull's
// frame) against the hull over the lanes: each lane evaluates its share of
// the faces and edges with the serial scan's expressions (m3SegmentFaceSep,
// m3SegmentEdgeAxis and the vertex maximum), and the reduction keeps the
// serial scan's choice, the first index of the largest separation (a lane
// offers its candidates in ascending index, with the same strict test), so
// the queries equal m3CollideSegmentHullQ's own loops bit for bit.
template <class Lds>
__device__ m3SatQuery WaveSegmentHull(const m3HullData* hull, m3Vec3 p1, m3Vec3 p2, Lds& L, int lane)
{
    const int32_t nv = hull->vertexCount;
    for (int32_t k = lane; k < nv; k += 64)
    {
        L.w[k] = hull->vertices[k];
    }
    __syncthreads();
    int32_t face = 0;
    m3real faceSep = -3.4e38f;
    for (int32_t f = lane; f < hull->faceCount; f += 64)
    {
        const m3real sep = m3SegmentFaceSep(hull, f, p1, p2);
        if (sep > faceSep)
        {
            faceSep = sep;
            face = f;
        }
    }
    const m3Vec3 segDir = m3Sub3(p2, p1);
    int32_t edge = -1;
    m3real edgeSep = -3.4e38f;
    m3Vec3 edgeAxis = m3V3(0.0f, 1.0f, 0.0f);
    for (int32_t e = 2 * lane; e < hull->edgeCount; e += 128)
    {
        m3Vec3 axis;
        if (!m3SegmentEdgeAxis(hull, e, segDir, &axis))
        {
            continue;
        }
        m3real hullMax = -3.4e38f;
        for (int32_t k = 0; k < nv; ++k)
        {
            const m3real proj = m3Dot3(axis, L.w[k]);
            if (proj > hullMax)
            {
                hullMax = proj;
            }
        }
        const m3real sep = m3MinF(m3Dot3(axis, p1), m3Dot3(axis, p2)) - hullMax;
        if (sep > edgeSep)
        {
            edgeSep = sep;
            edgeAxis = axis;
            edge = e;
        }
    }
    for (int off = 32; off > 0; off >>= 1)
    {
        const int32_t oFace = __shfl_xor(face, off, 64);
        const m3real oFaceSep = __shfl_xor(faceSep, off, 64);
        if (oFaceSep > faceSep || (oFaceSep == faceSep && oFace < face))
        {
            faceSep = oFaceSep;
            face = oFace;
        }
        const int32_t oEdge = __shfl_xor(edge, off, 64);
        const m3real oEdgeSep = __shfl_xor(edgeSep, off, 64);
        const m3real ox = __shfl_xor(edgeAxis.x, off, 64);
        const m3real oy = __shfl_xor(edgeAxis.y, off, 64);
        const m3real oz = __shfl_xor(edgeAxis.z, off, 64);
        if (oEdgeSep > edgeSep || (oEdgeSep == edgeSep && oEdge < edge))
        {
            edgeSep = oEdgeSep;
            edge = oEdge;
            edgeAxis = m3V3(ox, oy, oz);
        }
    }
    __syncthreads();
    m3SatQuery out;
    out.faceA.faceIndex = face;
    out.faceA.separation = faceSep;
    out.faceB.faceIndex = 0;
    out.faceB.separation = 0.0f;
    out.edge.indexA = edge;
    out.edge.indexB = -1;
    out.edge.separation = edgeSep;
    out.edge.axis = edgeAxis;
    return out;
}

// One 64-lane workgroup per listed pair, grid-stride over the list.
template <class Lds>
__global__ __launch_bounds__(64) void k_sat(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    __shared__ Lds L;
    if (v.sc->stall != 0)
    {
        return;
    }
    int lane = threadIdx.x;
    int32_t count = v.sc->satCount;
    for (int32_t k = blockIdx.x; k < count; k += gridDim.x)
    {
        int32_t i = v.satList[k];
        uint64_t key = v.keys[i];
        int32_t shapeA = (int32_t)(key >> 32);
        int32_t shapeB = (int32_t)(key & 0xFFFFFFFFu);
        if (v.shapeType[shapeA] == M3_CAPSULE || v.shapeType[shapeB] == M3_CAPSULE)
        {
            const int32_t hullShape = v.shapeType[shapeA] == M3_HULL ? shapeA : shapeB;
            const int32_t capShape = hullShape == shapeA ? shapeB : shapeA;
            m3Vec3 s1;
            m3Vec3 s2;
            m3CapsuleInHull(&v, hullShape, capShape, &s1, &s2);
            const m3SatQuery out = WaveSegmentHull(&v.hulls[v.hullIndex[hullShape]], s1, s2, L, lane);
            if (lane == 0)
            {
                v.satQueries[i] = out;
            }
            continue;
        }
        const m3HullData* hullA = &v.hulls[v.hullIndex[shapeA]];
        const m3HullData* hullB = &v.hulls[v.hullIndex[shapeB]];
        m3Transform xfA;
        m3Quat q;
        m3Vec3 p;
        m3HullPairFrame(&v, shapeA, shapeB, &xfA, &q, &p);
        m3SatQuery out;
        out.faceA = WaveQueryFaces(hullA, hullB, q, p, 1, L, lane);
        out.faceB.separation = 0.0f;
        out.faceB.faceIndex = 0;
        out.edge.separation = 0.0f;
        out.edge.indexA = -1;
        out.edge.indexB = -1;
        out.edge.axis = m3V3(0.0f, 1.0f, 0.0f);
        // The manifold reads the later queries only when the earlier ones
        // found no separation.
        if (!(out.faceA.separation > M3_SPECULATIVE_DISTANCE))
        {
            out.faceB = WaveQueryFaces(hullB, hullA, q, p, 0, L, lane);
            if (!(out.faceB.separation > M3_SPECULATIVE_DISTANCE))
            {
                out.edge = WaveQueryEdges(hullA, hullB, q, p, L, lane);
            }
        }
        if (lane == 0)
        {
            v.satQueries[i] = out;
        }
    }
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

// Island coupling of touching pair i (src/solver.c:3468-3580): two dynamic
// bodies join; a dynamic body touching a moving kinematic one is forced
// awake.
__device__ static inline void UnionTouching(const m3View& v, int32_t i)
{
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

// Collision routine a pair dispatches to in m3NarrowPair (0 plane-hull,
// 1 hull-hull, 2 plane-capsule, 3 capsule-hull, 4 GJK, 5 plane-sphere,
// 6 sphere-sphere).
__device__ static inline int NarrowClass(const m3View& v, int32_t i)
{
    const uint64_t key = v.keys[i];
    const uint8_t typeA = v.shapeType[(int32_t)(key >> 32)];
    const uint8_t typeB = v.shapeType[(int32_t)(key & 0xFFFFFFFFu)];
    const int planePair = typeA == M3_PLANE || typeB == M3_PLANE;
    const int hullPair = typeA == M3_HULL || typeB == M3_HULL;
    const int capsulePair = typeA == M3_CAPSULE || typeB == M3_CAPSULE;
    if (planePair && hullPair)
    {
        return 0;
    }
    if (typeA == M3_HULL && typeB == M3_HULL)
    {
        return 1;
    }
    if (planePair && capsulePair)
    {
        return 2;
    }
    if (hullPair && capsulePair)
    {
        return 3;
    }
    if (hullPair || capsulePair)
    {
        return 4;
    }
    return planePair ? 5 : 6;
}

// Step-start body state and union-find reset of body i.
__device__ static inline void CaptureBody(const m3View& v, int32_t i)
{
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

// The pairs are first binned by collision routine (k_pairs_settle; the
// order inside a bin is immaterial: every pair writes its own manifold
// slot), then each wavefront of k_narrow_binned takes 64 pairs of one bin,
// so no wavefront runs more than one routine (a wavefront mixing routines
// runs them one after another).  Batched worlds bin per world (grid.y).
// The same launch captures every body's step-start state; the narrowphase
// that follows unites bodies over the reset union-find.
#define M3_NARROW_CLASSES 7

// Pair i listed for the wavefront SAT queries when it is a wide hull pair.
__device__ static inline void SatList(const m3View& v, int32_t i, uint64_t key)
{
    int32_t shapeA = (int32_t)(key >> 32);
    int32_t shapeB = (int32_t)(key & 0xFFFFFFFFu);
    const uint8_t typeA = v.shapeType[shapeA];
    const uint8_t typeB = v.shapeType[shapeB];
    int wide = 0;
    if (typeA == M3_HULL && typeB == M3_HULL)
    {
        wide = m3WideHullPair(&v.hulls[v.hullIndex[shapeA]], &v.hulls[v.hullIndex[shapeB]], v.satBoxes);
    }
    else if ((typeA == M3_HULL && typeB == M3_CAPSULE) || (typeA == M3_CAPSULE && typeB == M3_HULL))
    {
        wide = m3WideSegmentHull(&v.hulls[v.hullIndex[typeA == M3_HULL ? shapeA : shapeB]]);
    }
    if (wide)
    {
        v.satList[atomicAdd(&v.sc->satCount, 1)] = i;
    }
}

// The step's pair list from the unique keys (or, on a stall, the old list
// restored) and each pair's sleep flag (FlagCount for k_compact_frozen);
// then, with the key in hand, the wide hull pairs listed for the SAT
// queries (`sat`) and each pair binned by collision routine; and each
// body's step-start state and union-find reset.  The grid covers
// max(pairCapacity, maxBody).
__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_pairs_settle(const m3Dev* __restrict__ ds, int sat)
{
    M3_WORLD;
    __shared__ int32_t cnt[8];
    __shared__ int32_t base[8];
    const uint32_t* unique = d.emitUnique;
    m3Scalars* sc = v.sc;
    int32_t n = sc->uniqueCount;
    if (n > 0 && unique[n - 1] == M3_SENTINEL32)
    {
        n -= 1;
    }
    const int stall = sc->stall != 0 || sc->emitCount > v.emitCapacity || n > v.pairCapacity;
    const int32_t oldCount = sc->oldPairCount;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
    {
        if (stall)
        {
            sc->stall = 1;
        }
        sc->pairCount = stall ? oldCount : n;
    }
    uint8_t cold = 0;
    if (i < v.pairCapacity)
    {
        if (stall)
        {
            if (i < oldCount)
            {
                v.keys[i] = v.oldKeys[i];
                v.manifolds[i] = v.oldManifolds[i];
            }
        }
        else if (i < n)
        {
            const uint64_t key = m3ExpandKey(unique[i]);
            v.keys[i] = key;
            cold = m3PairHot(&v, (int32_t)(key >> 32), (int32_t)(key & 0xFFFFFFFFu)) ? 0 : 1;
        }
        d.pairFlags[i] = cold;
    }
    FlagCount(cold, d.blockCounts, blockIdx.x);
    if (i < v.maxBody && !stall)
    {
        CaptureBody(v, i);
    }
    const bool live = !stall && i < n;
    if (sat != 0 && live)
    {
        SatList(v, i, v.keys[i]);
    }
    if (threadIdx.x < 8)
    {
        cnt[threadIdx.x] = 0;
    }
    __syncthreads();
    const int cls = live ? NarrowClass(v, i) : M3_NARROW_CLASSES;
    int32_t rank = 0;
    if (cls < M3_NARROW_CLASSES)
    {
        rank = atomicAdd(&cnt[cls], 1);
    }
    __syncthreads();
    if (threadIdx.x < M3_NARROW_CLASSES && cnt[threadIdx.x] > 0)
    {
        base[threadIdx.x] = atomicAdd(&d.narrowBins[threadIdx.x], cnt[threadIdx.x]);
    }
    __syncthreads();
    if (cls < M3_NARROW_CLASSES)
    {
        d.narrowBins[8 + (int64_t)cls * v.pairCapacity + base[cls] + rank] = i;
    }
}


// Grid: ceil(pairCapacity / 64) + M3_NARROW_CLASSES workgroups of 64, enough
// for every bin rounded up to whole wavefronts.  Workgroup b walks the bins'
// wavefront counts to find its bin and its 64 entries.
__global__ void __launch_bounds__(64) k_narrow_binned(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    if (v.sc->stall != 0)
    {
        return;
    }
    int32_t b = blockIdx.x;
    int cls = -1;
    int32_t count = 0;
    for (int c = 0; c < M3_NARROW_CLASSES; ++c)
    {
        count = d.narrowBins[c];
        const int32_t waves = (count + 63) / 64;
        if (b < waves)
        {
            cls = c;
            break;
        }
        b -= waves;
    }
    const int32_t k = b * 64 + (int32_t)threadIdx.x;
    if (cls < 0 || k >= count)
    {
        return;
    }
    const int32_t i = d.narrowBins[8 + (int64_t)cls * v.pairCapacity + k];
    if (v.hitRecs != nullptr)
    {
        v.hitRecs[i].flag = 0;
    }
    // The island union of a touching pair follows its manifold here (the
    // bodies' union-find entries were reset by k_pairs_settle before this launch).
    if (m3NarrowPair(&v, i, v.sc->oldPairCount) > 0)
    {
        UnionTouching(v, i);
    }
}

// ---------------------------------------------------------------------------
// Sweep capture and islands (src/solver.c:3437-3580, 3851-3881)
// ---------------------------------------------------------------------------

// Joints couple islands like touching contacts; a moving kinematic
// partner is a wake source through a joint too (src/solver.c:3520-3547).
__global__ void k_union_joints(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || j >= v.maxJoint || v.jointAlive[j] == 0)
    {
        return;
    }
    int32_t bA = v.jointBodyA[j];
    int32_t bB = v.jointBodyB[j];
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

// Body i's island root; an awake or forced dynamic body marks its island
// for waking (the mark needs only the body's own root).
__device__ static inline void RootMark(const m3Dev& d, int32_t i)
{
    const m3View& v = d.v;
    const int32_t r = Find(v.parent, i);
    d.root[i] = r;
    v.slotMask[i] = 0u;
    if (v.bodyAlive[i] != 0 && v.type[i] == M3_DYNAMIC && (v.awake[i] != 0 || v.forced[i] != 0))
    {
        v.flag[r] = 1;
    }
}

// Every dynamic body of a marked island wakes.
__device__ static inline void WakeApply(const m3Dev& d, int32_t i)
{
    const m3View& v = d.v;
    if (v.bodyAlive[i] != 0 && v.type[i] == M3_DYNAMIC && v.awake[i] == 0 && v.flag[d.root[i]] != 0)
    {
        v.awake[i] = 1;
        v.sleepTime[i] = 0.0f;
    }
}

__global__ void k_root(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0)
    {
        return;
    }
    RootMark(d, i);
}

__global__ void k_wake_apply(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0)
    {
        return;
    }
    WakeApply(d, i);
}

// ---------------------------------------------------------------------------
// Prepare and coloring (src/solver.c:156-404, 2741-2822)
// ---------------------------------------------------------------------------

// Pre-solve candidates of this step (src/solver.c:214-236), appended in any
// order with their pair index (the host puts them in pair order).
__global__ void k_presolve_gather(const m3Dev* __restrict__ ds, m3gpuPreSolve* out, int32_t* count)
{
    M3_WORLD;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || i >= v.sc->pairCount)
    {
        return;
    }
    double point[3];
    m3Vec3 normal;
    if (!m3PreSolveCandidate(&v, i, point, &normal))
    {
        return;
    }
    const int32_t k = atomicAdd(count, 1);
    m3gpuPreSolve r;
    r.key = v.keys[i];
    r.pair = i;
    r.pad = 0;
    r.point[0] = point[0];
    r.point[1] = point[1];
    r.point[2] = point[2];
    r.normal[0] = normal.x;
    r.normal[1] = normal.y;
    r.normal[2] = normal.z;
    r.pad2 = 0.0f;
    out[k] = r;
}

// One flag per pair of the current list: its key is among the step's
// vetoed keys (ascending).
__global__ void k_veto_mark(const m3Dev* __restrict__ ds, const uint64_t* keys, int32_t n, uint8_t* flags)
{
    M3_WORLD;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.pairCapacity)
    {
        return;
    }
    uint8_t hit = 0;
    if (i < v.sc->pairCount)
    {
        const uint64_t key = v.keys[i];
        int32_t lo = 0;
        int32_t hi = n - 1;
        while (lo <= hi)
        {
            const int32_t mid = (lo + hi) >> 1;
            if (keys[mid] == key)
            {
                hit = 1;
                break;
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
    }
    flags[i] = hit;
}

__device__ static inline uint8_t ConstraintFlag(const m3Dev& d, int32_t i)
{
    const m3View& v = d.v;
    uint8_t f = 0;
    if (i < v.pairCapacity)
    {
        f = (v.sc->stall == 0 && i < v.sc->pairCount && m3PairIsConstraint(&v, i)) ? 1 : 0;
        d.pairFlags[i] = f;
    }
    return f;
}

__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_constraint_flags(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    const uint8_t f = ConstraintFlag(d, blockIdx.x * blockDim.x + threadIdx.x);
    FlagCount(f, d.blockCounts, blockIdx.x);
}

// The constraint list in pair order, with the awake dynamic bodies of each
// constraint, the only bodies coloring tracks (src/solver.c:2741-2822).
// Chunk `chunk` of the constraint list (constraint c: its pair and its
// awake dynamic bodies, or -1); the last chunk's thread 0 writes the count.
__device__ static inline void CompactConstraints(const m3Dev& d, int32_t chunk, int32_t chunks)
{
    const m3View& v = d.v;
    const int32_t i = chunk * M3_COMPACT_THREADS + threadIdx.x;
    const int flag = i < v.pairCapacity && d.pairFlags[i] != 0;
    int32_t total;
    const int32_t c = FlagRank(flag, d.blockCounts, chunk, &total);
    if (flag)
    {
        v.consPair[c] = i;
        const uint64_t key = v.keys[i];
        const int32_t a = v.shapeBody[(int32_t)(key >> 32)];
        const int32_t b = v.shapeBody[(int32_t)(key & 0xFFFFFFFFu)];
        v.consBodies[2 * c] = (v.type[a] == M3_DYNAMIC && v.awake[a] != 0) ? a : -1;
        v.consBodies[2 * c + 1] = (v.type[b] == M3_DYNAMIC && v.awake[b] != 0) ? b : -1;
    }
    if (chunk == chunks - 1 && threadIdx.x == 0)
    {
        v.sc->constraintCount = total;
    }
}

__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_compact_constraints(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    CompactConstraints(d, blockIdx.x, gridDim.x);
}

// The awake dynamic bodies of each constraint, the only bodies coloring
// tracks (src/solver.c:2741-2822).
__global__ void k_cons_bodies(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
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
__global__ void k_prepare(const m3Dev* __restrict__ ds, float hertz, float zeta, float h, int flowPrep)
{
    M3_WORLD;
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
    if (flowPrep)
    {
        // The barrier-free passes track awake dynamic bodies only, and the
        // rolling row writes a non-dynamic side's angular velocity.
        const int2 ab = make_int2(v.consBodies[2 * c], v.consBodies[2 * c + 1]);
        d.slotBodies[k] = ab;
        // The constraint's rank on each awake dynamic body: the colors
        // below its own in which the body has a constraint (the final masks
        // k_color left).
        const uint32_t lower = (1u << color) - 1u;
        const uint32_t rankA = ab.x >= 0 ? (uint32_t)__popc((uint32_t)d.bodyMasks[ab.x] & lower) : M3_FLOW_NONE;
        const uint32_t rankB = ab.y >= 0 ? (uint32_t)__popc((uint32_t)d.bodyMasks[ab.y] & lower) : M3_FLOW_NONE;
        d.flowRank[k] = rankA | (rankB << 8);
        const bool awakeA = ab.x >= 0;
        const bool awakeB = ab.y >= 0;
        if ((con->dynA != 0) != awakeA || (con->dynB != 0) != awakeB ||
            (con->rollingResistance > 0.0f && (!con->dynA || !con->dynB)))
        {
            v.sc->flowVeto = 1;
        }
        // A kinematic side's position deltas move during the substeps
        // without a solve count of their own.
        if ((!con->dynA && v.type[con->bodyA] == M3_KINEMATIC) ||
            (!con->dynB && v.type[con->bodyB] == M3_KINEMATIC))
        {
            v.sc->flowKinematic = 1;
        }
        m3Vec3 rA[M3_MANIFOLD_MAX_POINTS];
        m3Vec3 rB[M3_MANIFOLD_MAX_POINTS];
        m3real imp[M3_MANIFOLD_MAX_POINTS];
        for (int32_t q = 0; q < M3_MANIFOLD_MAX_POINTS; ++q)
        {
            rA[q] = con->points[q].rA;
            rB[q] = con->points[q].rB;
            imp[q] = con->points[q].normalImpulse;
        }
        m3WarmAddends(con->pointCount, con->normal, rA, rB, imp, con->invMassA, con->invMassB, &con->invIA,
                      &con->invIB, con->t1, con->t2, con->frictionImpulse1, con->frictionImpulse2,
                      con->originA, con->originB, con->twistImpulse, con->rollingImpulse, con->dynA,
                      con->dynB, &d.warmAdd[2 * k], &d.warmAdd[2 * k + 1]);
    }
}

// Joints with an awake dynamic side are prepared in place by slot
// (src/solver.c:699-811); filter joints have no rows.
__global__ void k_prepare_joints(const m3Dev* __restrict__ ds, float h)
{
    M3_WORLD;
    int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || j >= v.maxJoint)
    {
        return;
    }
    uint8_t active = 0;
    if (v.jointAlive[j] != 0 && v.jointType[j] != 7)
    {
        int32_t a = v.jointBodyA[j];
        int32_t b = v.jointBodyB[j];
        if ((v.type[a] == M3_DYNAMIC && v.awake[a] != 0) ||
            (v.type[b] == M3_DYNAMIC && v.awake[b] != 0))
        {
            active = 1;
            m3PrepareJoint(&v, j, &v.jcons[j], h);
            atomicAdd(&v.sc->jointActive, 1);
        }
    }
    v.jointActiveFlag[j] = active;
}

__global__ void k_store_joints(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (v.sc->stall != 0 || j >= v.maxJoint || v.jointActiveFlag[j] == 0)
    {
        return;
    }
    m3StoreJoint(&v, &v.jcons[j], j);
}

// Greedy first-fit coloring in constraint order (src/solver.c:2741-2822),
// reproduced exactly by one workgroup.  A constraint colors in the first
// round in which it is the lowest pending constraint on each of its awake
// dynamic bodies (an owner table claimed with atomic min), so every earlier
// constraint that shares a body has already updated that body's mask.  Body masks and
// owners live in LDS for worlds up to M3_COLOR_LDS_BODIES bodies.
#define M3_COLOR_THREADS 1024
#define M3_COLOR_LDS_BODIES 16384

// Round barrier of the coloring loop: with the tables in LDS it orders LDS
// only, so global loads and stores stay in flight across it.
template <bool kLds>
__device__ static inline void ColorBarrier()
{
    if (kLds)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
    else
    {
        __syncthreads();
    }
}

// Inclusive sum within each row of 16 lanes (DPP row shifts).
__device__ static inline uint32_t RowInclusiveSum(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false); // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false); // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false); // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false); // row_shr:8
    return x;
}

static_assert(M3_COLOR_THREADS / 64 == 16, "k_color scans one DPP row of wave counts per color");
static_assert((M3_GRAPH_COLORS + 1) * 16 <= M3_COLOR_THREADS, "k_color's color scan needs a row per color");

template <bool kLds>
__global__ void __launch_bounds__(M3_COLOR_THREADS)
    k_color(const m3Dev* __restrict__ ds, int flowPrep)
{
    M3_WORLD;
    uint16_t* gMasks = d.bodyMasks;
    int32_t* gOwner = d.bodyOwner;
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
    __shared__ int32_t cursor;
    __shared__ int32_t anyPending[2];
    if (t == 0)
    {
        cursor = M3_COLOR_THREADS;
        anyPending[0] = 0;
        anyPending[1] = 0;
    }
    __syncthreads();
    // A sliding window of one pending constraint per participant: a
    // participant whose constraint colors takes the next unclaimed index.
    // Taken indices exceed every constraint already in the window, so the
    // lowest-pending rule still sees every earlier constraint on a body.
    // With the tables in LDS the round barriers order LDS only, and the
    // load of a newly taken constraint's bodies overlaps the barrier.
    const int32_t lane = t & 63;
    int32_t cur = t < n ? t : -1;
    int32_t A = -1;
    int32_t B = -1;
    if (cur >= 0)
    {
        A = v.consBodies[2 * cur];
        B = v.consBodies[2 * cur + 1];
    }
    for (int32_t round = 0;; ++round)
    {
        if (cur >= 0)
        {
            if (A >= 0)
            {
                atomicMin(&owner[A], cur);
            }
            if (B >= 0)
            {
                atomicMin(&owner[B], cur);
            }
            anyPending[round & 1] = 1;
        }
        ColorBarrier<kLds>();
        if (anyPending[round & 1] == 0)
        {
            break;
        }
        // Every round colors at least the lowest pending constraint, so
        // more than n rounds would mean a broken table: stop and report.
        if (round > n)
        {
            if (t == 0)
            {
                __hip_atomic_store(&sc->fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            break;
        }
        const int ready = cur >= 0 && (A < 0 || owner[A] == cur) && (B < 0 || owner[B] == cur);
        if (ready)
        {
            uint32_t mask = (A >= 0 ? (uint32_t)masks[A] : 0u) | (B >= 0 ? (uint32_t)masks[B] : 0u);
            // Lowest color free on both bodies (the serial scan's first
            // clear bit), M3_GRAPH_COLORS when all are taken.
            const uint32_t freeBits = ~mask & ((1u << M3_GRAPH_COLORS) - 1u);
            const int32_t color = freeBits != 0u ? (int32_t)__builtin_ctz(freeBits) : M3_GRAPH_COLORS;
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
            v.colors[cur] = (uint8_t)color;
            atomicAdd(&counts[color], 1);
            // Only a body's owner matches its entry, so clearing it here
            // cannot make another participant of this round ready.
            if (A >= 0)
            {
                owner[A] = 0x7FFFFFFF;
            }
            if (B >= 0)
            {
                owner[B] = 0x7FFFFFFF;
            }
        }
        // Participants that colored take the next indices (one LDS atomic
        // per wave).
        const uint64_t took = __ballot(ready);
        int32_t base = 0;
        if (lane == 0 && took != 0ull)
        {
            base = atomicAdd(&cursor, (int32_t)__popcll(took));
        }
        base = __shfl(base, 0);
        if (ready)
        {
            const int32_t id = base + (int32_t)__popcll(took & ((1ull << lane) - 1ull));
            cur = id < n ? id : -1;
            A = -1;
            B = -1;
            if (cur >= 0)
            {
                A = v.consBodies[2 * cur];
                B = v.consBodies[2 * cur + 1];
            }
        }
        // The other round's flag was last read before this round's first
        // barrier; it is next set after the barrier below.
        if (t == 0)
        {
            anyPending[(round + 1) & 1] = 0;
        }
        ColorBarrier<kLds>();
    }
    // The list build reads every participant's colors from global memory.
    __syncthreads();
    if (kLds && flowPrep)
    {
        // The final body masks, for each constraint's ranks (k_prepare).
        for (int32_t i = t; i < v.maxBody; i += M3_COLOR_THREADS)
        {
            gMasks[i] = sMask[i];
        }
    }
    __shared__ int32_t starts[M3_GRAPH_COLORS + 2];
    if (t == 0)
    {
        int32_t cursor = 0;
        int32_t used = 0;
        for (int32_t k = 0; k <= M3_GRAPH_COLORS; ++k)
        {
            starts[k] = cursor;
            sc->colorStarts[k] = cursor;
            cursor += counts[k];
            used += counts[k] > 0 ? 1 : 0;
        }
        sc->colorStarts[M3_GRAPH_COLORS + 1] = cursor;
        sc->colorCount = used;
    }
    // Color lists: each color's slots hold its constraints in ascending
    // order (a stable counting sort; in-wave ranks from ballots).
    __shared__ int32_t waveBase[M3_GRAPH_COLORS + 1][M3_COLOR_THREADS / 64];
    __shared__ int32_t running[M3_GRAPH_COLORS + 1];
    if (t <= M3_GRAPH_COLORS)
    {
        running[t] = 0;
    }
    __syncthreads();
    const int32_t wave = t >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    // Eight chunks' colors are loaded at a time, ahead of their passes.
    int32_t ahead[8];
    for (int32_t base = 0; base < n; base += M3_COLOR_THREADS)
    {
        const int32_t q = (base / M3_COLOR_THREADS) & 7;
        if (q == 0)
        {
            M3_UNROLL
            for (int32_t j = 0; j < 8; ++j)
            {
                const int32_t cj = base + j * M3_COLOR_THREADS + t;
                ahead[j] = cj < n ? (int32_t)v.colors[cj] : -1;
            }
        }
        int32_t c = base + t;
        int32_t color = -1;
        M3_UNROLL
        for (int32_t j = 0; j < 8; ++j)
        {
            color = j == q ? ahead[j] : color;
        }
        int32_t rank = 0;
        for (int32_t k = 0; k <= M3_GRAPH_COLORS; ++k)
        {
            uint64_t m = __ballot(color == k);
            if (color == k)
            {
                rank = __popcll(m & below);
            }
            if (lane == 0)
            {
                waveBase[k][wave] = __popcll(m);
            }
        }
        __syncthreads();
        // Each color's 16 wave counts are one DPP row: an exclusive scan
        // across the row, on top of the color's running count.
        if (t < (M3_GRAPH_COLORS + 1) * 16)
        {
            const int32_t k = t >> 4;
            const int32_t w = t & 15;
            const uint32_t x = (uint32_t)waveBase[k][w];
            const uint32_t incl = RowInclusiveSum(x);
            const int32_t run = running[k];
            waveBase[k][w] = run + (int32_t)(incl - x);
            if (w == 15)
            {
                running[k] = run + (int32_t)incl;
            }
        }
        __syncthreads();
        if (color >= 0)
        {
            int32_t slot = starts[color] + waveBase[color][wave] + rank;
            v.colorLists[slot] = c;
            v.slotColors[slot] = (uint8_t)color;
        }
        __syncthreads();
    }
}


// ---------------------------------------------------------------------------
// Movers and the substep loop (src/solver.c:3938-4203)
// ---------------------------------------------------------------------------

// Body i's mover flag (kinematic, or dynamic and awake); a kinematic body
// with a target gets its servo velocities.
__device__ static inline uint8_t MoverFlag(const m3Dev& d, int32_t i, float dt)
{
    const m3View& v = d.v;
    uint8_t* flags = d.bodyFlags;
    uint8_t f = 0;
    if (i < v.maxBody && v.sc->stall == 0 && v.bodyAlive[i] != 0 && v.enabled[i] != 0)
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
    if (i < v.maxBody)
    {
        flags[i] = f;
        if (v.sc->stall == 0)
        {
            v.flag[i] = 1; // island readiness for the sleep pass, cleared by any slow member
        }
    }
    if (f != 0 && d.flowRec != nullptr)
    {
        // The barrier-free passes' records of a mover: no pass writes this
        // version, so no word left by an earlier step (or never written)
        // can pass for one of this step's.
        for (int w = 0; w < 6; ++w)
        {
            d.flowRec[6 * (int64_t)i + w] = (uint64_t)M3_FLOW_STALE << 32;
        }
        for (int w = 0; w < M3_DELTA_WORDS; ++w)
        {
            d.deltaRec[M3_DELTA_WORDS * (int64_t)i + w] = (uint64_t)M3_FLOW_STALE << 32;
        }
    }
    return f;
}

__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_mover_flags(const m3Dev* __restrict__ ds, float dt)
{
    M3_WORLD;
    const uint8_t f = MoverFlag(d, blockIdx.x * blockDim.x + threadIdx.x, dt);
    FlagCount(f, d.blockCounts, blockIdx.x);
}

// The movers (solved bodies) in slot order.
__device__ static inline void CompactMovers(const m3Dev& d, int32_t chunk, int32_t chunks)
{
    const m3View& v = d.v;
    const int32_t i = chunk * M3_COMPACT_THREADS + threadIdx.x;
    const int flag = i < v.maxBody && d.bodyFlags[i] != 0;
    int32_t total;
    const int32_t pos = FlagRank(flag, d.blockCounts, chunk, &total);
    if (flag)
    {
        v.movers[pos] = i;
    }
    if (chunk == chunks - 1 && threadIdx.x == 0)
    {
        v.sc->moverCount = total;
    }
}

__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_compact_movers(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    CompactMovers(d, blockIdx.x, gridDim.x);
}

struct m3SubstepParams
{
    int32_t substeps;
    int32_t tailLimit; // constraint budget of the trailing colors solved in workgroup 0
    int32_t flow;      // barrier-free passes allowed (single-world kernel)
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
    // Every wavefront waits for its own stores first: the workgroup barrier
    // alone does not (on gfx950 it is a bare s_barrier), and the arriving
    // thread's release waits only for its own wavefront's.  Once they are
    // acknowledged, device-scope stores are at the coherence point and
    // plain stores are in this XCD's L2, which that release writes back.
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
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

// The velocities and accumulators a constraint solve starts from.
__device__ static inline void TapeSolve(const m3View& v, const m3Constraint* c, m3TapePass* r)
{
    r->vA = v.lv[c->bodyA];
    r->wA = v.av[c->bodyA];
    r->vB = v.lv[c->bodyB];
    r->wB = v.av[c->bodyB];
    for (int32_t q = 0; q < M3_MANIFOLD_MAX_POINTS; ++q)
    {
        r->lambda[q] = q < c->pointCount ? c->points[q].normalImpulse : 0.0f;
    }
    r->f1 = c->frictionImpulse1;
    r->f2 = c->frictionImpulse2;
    r->twist = c->twistImpulse;
}

// The contact constraint a participant solves first in its next color
// phase, gathered before the barrier that precedes that phase: its loads
// then overlap the barrier instead of following it.  A constraint's slot
// maps to the same participant in every pass, and the fields gathered are
// read-only in the substep loop or written only by that participant's own
// solve of the constraint, so the copy is current when it is used.
struct m3Prefetch
{
    int32_t k; // constraint slot, or -1
    m3ContactIn in;
};

template <int kMode>
__device__ static inline void GatherFirst(const m3View& v, int32_t k, m3Prefetch* pf)
{
    pf->k = k;
    if (k >= 0)
    {
        m3LoadContact(&v.cons[k], kMode == 1, &pf->in);
    }
}

template <int kMode>
__device__ static inline void ColorRun(const m3View& v, int32_t start, int32_t end,
                                       const m3SubstepParams& sp, int32_t gid, int32_t gsz,
                                       m3TapePass* rec, const m3Prefetch* pf)
{
    int32_t k = start + gid;
    if (k < end && pf->k == k)
    {
        if (rec != nullptr)
        {
            TapeSolve(v, &v.cons[k], &rec[k]);
        }
        m3SolveContactIn(&v, &v.cons[k], &pf->in, sp.invH, sp.push, kMode == 1);
        k += gsz;
    }
    for (; k < end; k += gsz)
    {
        if (rec != nullptr)
        {
            TapeSolve(v, &v.cons[k], &rec[k]);
        }
        m3SolveContact(&v, &v.cons[k], sp.invH, sp.push, kMode == 1);
    }
}

// One graph color of one pass (mode 1 solve with bias, 2 relax) over the
// participants; the overflow color runs serially on one thread.  Each
// pass has a loop of its own so that every inlined solve sees a constant
// mode.
__device__ static inline void ColorPhase(const m3View& v, int32_t start, int32_t end, int mode,
                                         const m3SubstepParams& sp, int32_t gid, int32_t gsz,
                                         bool serial, m3TapePass* rec, const m3Prefetch* pf)
{
    if (serial)
    {
        if (gid == 0)
        {
            for (int32_t k = start; k < end; ++k)
            {
                if (rec != nullptr)
                {
                    TapeSolve(v, &v.cons[k], &rec[k]);
                }
                m3SolveContact(&v, &v.cons[k], sp.invH, sp.push, mode == 1);
            }
        }
        return;
    }
    if (mode == 1)
    {
        ColorRun<1>(v, start, end, sp, gid, gsz, rec, pf);
    }
    else
    {
        ColorRun<2>(v, start, end, sp, gid, gsz, rec, pf);
    }
}

#define M3_SUBSTEP_THREADS 256 // M3_CELL_THREADS
#ifndef M3_BLOCK_GATHER
#define M3_BLOCK_GATHER 0
#endif

// The whole substep loop of src/solver.c:4085-4203, with a barrier after
// each phase:
//   integrate velocities and warm start (one phase: a warm-start impulse
//   depends only on stored accumulators, so each body replays its own
//   constraints in color order, the exact operation sequence of the colored
//   pass on that body), solve by color, integrate positions, relax by color.
// Every branch depends only on counts written before the loop starts, so
// all participants reach the same barriers.
struct m3GridSync
{
    uint32_t* bar;
    int32_t* fault;
    __device__ void operator()() const { GridBarrier(bar, fault); }
};

struct m3BlockSync
{
    __device__ void operator()() const { __syncthreads(); }
};

// ---------------------------------------------------------------------------
// Cooperative forms of the island pass and the two flag-and-compact pairs of
// the single-world step: one launch each, a grid barrier between the
// passes.  Chunk c (items 256 c .. 256 c + 255) belongs to workgroup
// c % gridDim.x, so any co-resident grid computes the same result.
// ---------------------------------------------------------------------------

__device__ static inline int32_t Chunks(int32_t n)
{
    return n > 0 ? (n + M3_COMPACT_THREADS - 1) / M3_COMPACT_THREADS : 1;
}

// Roots and wake marks, then the wake (src/solver.c:3532-3580).
__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_islands(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    const m3GridSync sync = {d.barrier, &v.sc->fault};
    const bool live = v.sc->stall == 0;
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; live && i < v.maxBody; i += stride)
    {
        RootMark(d, i);
    }
    sync();
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; live && i < v.maxBody; i += stride)
    {
        WakeApply(d, i);
    }
}

// The step's constraints: flags, then the ordered compaction, over the
// live pairs only (the pair count is the previous launch's).
__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_constraints(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    const m3GridSync sync = {d.barrier, &v.sc->fault};
    const int32_t live = v.sc->pairCount < v.pairCapacity ? v.sc->pairCount : v.pairCapacity;
    const int32_t chunks = Chunks(live);
    for (int32_t c = blockIdx.x; c < chunks; c += gridDim.x)
    {
        const uint8_t f = ConstraintFlag(d, c * M3_COMPACT_THREADS + threadIdx.x);
        FlagCount(f, d.blockCounts, c);
        __syncthreads();
    }
    sync();
    for (int32_t c = blockIdx.x; c < chunks; c += gridDim.x)
    {
        CompactConstraints(d, c, chunks);
        __syncthreads();
    }
}

// The step's movers: flags (and kinematic servos), then the compaction.
__global__ void __launch_bounds__(M3_COMPACT_THREADS) k_movers(const m3Dev* __restrict__ ds, float dt)
{
    M3_WORLD;
    const m3GridSync sync = {d.barrier, &v.sc->fault};
    const int32_t chunks = Chunks(v.maxBody);
    for (int32_t c = blockIdx.x; c < chunks; c += gridDim.x)
    {
        const uint8_t f = MoverFlag(d, c * M3_COMPACT_THREADS + threadIdx.x, dt);
        FlagCount(f, d.blockCounts, c);
        __syncthreads();
    }
    sync();
    for (int32_t c = blockIdx.x; c < chunks; c += gridDim.x)
    {
        CompactMovers(d, c, chunks);
        __syncthreads();
    }
}

// On a grid, the last colors of a pass are small (the greedy coloring's
// sizes fall off quickly): the trailing colors holding at most
// M3_TAIL_CONSTRAINTS constraints in all run in workgroup 0 alone, color
// after color with workgroup barriers, and the grid meets once after them.
#define M3_TAIL_CONSTRAINTS 1024

// The small trailing colors of a pass in workgroup 0, without a barrier
// per color: a constraint waits only until every earlier tail constraint
// on each of its dynamic bodies has been solved.  Each dynamic body keeps
// a count of its solved tail constraints in LDS; a constraint's rank on a
// body is the number of tail colors below its own in which the body has a
// constraint (its slot mask), so it runs when both counts equal its ranks
// and each body still sees its constraints in color order, the operation
// sequence of the colored pass.  The lowest pending slot is always ready,
// so the wait ends; it is bounded all the same, and a wait that gives up
// marks the world faulted (reported by the host) instead of hanging.
#define M3_TAIL_LDS_BODIES 32768
#define M3_TAIL_SPINS (1u << 20)

template <int kMode>
__device__ static inline void TailFlow(const m3View& v, const m3SubstepParams& sp, int32_t tb,
                                       int32_t te, uint32_t tailColors, uint8_t* done)
{
    volatile uint8_t* vd = done;
    const int32_t t = threadIdx.x;
    for (int32_t i = t; i < v.maxBody; i += blockDim.x)
    {
        done[i] = 0;
    }
    __syncthreads();
    for (int32_t s0 = tb; s0 < te; s0 += blockDim.x)
    {
        const int32_t slot = s0 + t;
        int pending = slot < te;
        m3ContactIn in;
        int32_t rA = -1;
        int32_t rB = -1;
        if (pending)
        {
            m3LoadContact(&v.cons[slot], kMode == 1, &in);
            const uint32_t below = tailColors & ((1u << v.slotColors[slot]) - 1u);
            if (in.dynA)
            {
                rA = __popc(v.slotMask[in.a] & below);
            }
            if (in.dynB)
            {
                rB = __popc(v.slotMask[in.b] & below);
            }
        }
        uint32_t spins = 0;
        while (__any(pending))
        {
            if (pending && (rA < 0 || vd[in.a] == rA) && (rB < 0 || vd[in.b] == rB))
            {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                m3SolveContactIn(&v, &v.cons[slot], &in, sp.invH, sp.push, kMode == 1);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (rA >= 0)
                {
                    vd[in.a] = (uint8_t)(rA + 1);
                }
                if (rB >= 0)
                {
                    vd[in.b] = (uint8_t)(rB + 1);
                }
                pending = 0;
            }
            if (++spins == M3_TAIL_SPINS)
            {
                if (pending)
                {
                    __hip_atomic_store(&v.sc->fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                break;
            }
        }
    }
}

template <typename Sync, bool kGridTail = false>
__device__ static inline void SubstepLoop(const m3View& v, const m3SubstepParams& sp, int32_t gid,
                                          int32_t gsz, Sync sync, m3Tape* tape = nullptr)
{
    const bool gridTail = kGridTail;
    const int32_t nb = v.maxBody;
    int32_t starts[M3_GRAPH_COLORS + 2];
    for (int32_t k = 0; k < M3_GRAPH_COLORS + 2; ++k)
    {
        starts[k] = v.sc->colorStarts[k];
    }
    const int32_t movers = v.sc->moverCount;
    // First color of the workgroup-0 tail (M3_GRAPH_COLORS + 1: none).
    int32_t tailStart = M3_GRAPH_COLORS + 1;
    if (gridTail)
    {
        int32_t rem = starts[M3_GRAPH_COLORS + 1] - starts[M3_GRAPH_COLORS];
        if (rem <= sp.tailLimit)
        {
            tailStart = M3_GRAPH_COLORS;
            for (int32_t c = M3_GRAPH_COLORS - 1; c >= 0; --c)
            {
                int32_t size = starts[c + 1] - starts[c];
                if (rem + size > sp.tailLimit)
                {
                    break;
                }
                rem += size;
                tailStart = c;
            }
            if (rem == 0)
            {
                tailStart = M3_GRAPH_COLORS + 1;
            }
        }
    }
    // Joints run before the contacts in each pass, serially within a joint
    // component (src/solver.c:4149-4201); with joints the warm-start gather
    // waits for the joint warm start.
    const bool joints = v.sc->jointActive > 0;
    const int32_t comps = v.jointCompCount;
    // This participant's first slot in a color of a pass (-1: none, or the
    // serial overflow color), and the next non-empty color after one.
    auto firstSlot = [&](int32_t color) -> int32_t {
        if (color < 0 || color >= M3_GRAPH_COLORS)
        {
            return -1;
        }
        int32_t k = starts[color];
        if (color < tailStart)
        {
            k += gid;
        }
        else
        {
            // The barrier-free tail loads its constraints itself.
            if (blockIdx.x != 0 || (kGridTail && nb <= M3_TAIL_LDS_BODIES))
            {
                return -1;
            }
            k += (int32_t)threadIdx.x;
        }
        return k < starts[color + 1] ? k : -1;
    };
    auto nextColor = [&](int32_t color) -> int32_t {
        for (int32_t c = color + 1; c <= M3_GRAPH_COLORS; ++c)
        {
            if (starts[c + 1] > starts[c])
            {
                return c;
            }
        }
        return -1;
    };
    const int32_t firstColor = nextColor(-1);
    // Gathering pays where a grid barrier separates the phases; under a
    // workgroup barrier the loads would only lengthen register lifetimes.
    constexpr bool kGather = M3_BLOCK_GATHER || !std::is_same<Sync, m3BlockSync>::value;
    m3Prefetch pf;
    pf.k = -1;
    auto gather = [&](int mode, int32_t color) {
        if (!kGather)
        {
            return;
        }
        const int32_t k = firstSlot(color);
        if (mode == 1)
        {
            GatherFirst<1>(v, k, &pf);
        }
        else
        {
            GatherFirst<2>(v, k, &pf);
        }
    };
    for (int32_t sub = 0; sub < sp.substeps; ++sub)
    {
        if (tape != nullptr)
        {
            const int32_t base = sub * tape->bodyCap;
            for (int32_t b = gid; b < nb; b += gsz)
            {
                tape->sxf[base + b] = v.xf[b];
                tape->slv[base + b] = v.lv[b];
                tape->sav[base + b] = v.av[b];
                tape->sdp[base + b] = v.deltaPos[b];
                tape->sdq[base + b] = v.deltaRot[b];
            }
            sync();
        }
        for (int32_t m = gid; m < movers; m += gsz)
        {
            int32_t b = v.movers[m];
            m3IntegrateVelocity(&v, b, sp.h, sp.gravity, sp.maxLinear, sp.maxAngular);
            if (!joints)
            {
                m3WarmStartBody(&v, b, starts[M3_GRAPH_COLORS], starts[M3_GRAPH_COLORS + 1]);
            }
        }
        if (!joints)
        {
            gather(1, firstColor);
        }
        sync();
        if (joints)
        {
            for (int32_t k = gid; k < comps; k += gsz)
            {
                m3JointComponentPass(&v, k, 0, sp.invH);
            }
            sync();
            for (int32_t m = gid; m < movers; m += gsz)
            {
                m3WarmStartBody(&v, v.movers[m], starts[M3_GRAPH_COLORS],
                                starts[M3_GRAPH_COLORS + 1]);
            }
            sync();
        }
        for (int mode = 1; mode < 3; ++mode)
        {
            if (mode == 2)
            {
                if (tape != nullptr)
                {
                    const int32_t base = sub * tape->bodyCap;
                    for (int32_t b = gid; b < nb; b += gsz)
                    {
                        tape->plv[base + b] = v.lv[b];
                        tape->pav[base + b] = v.av[b];
                    }
                    sync();
                }
                for (int32_t m = gid; m < movers; m += gsz)
                {
                    m3IntegratePosition(&v, v.movers[m], sp.h);
                }
                if (!joints)
                {
                    gather(2, firstColor);
                }
                sync();
                if (tape != nullptr)
                {
                    const int32_t base = sub * tape->bodyCap;
                    for (int32_t b = gid; b < nb; b += gsz)
                    {
                        tape->rdp[base + b] = v.deltaPos[b];
                        tape->rdq[base + b] = v.deltaRot[b];
                    }
                }
            }
            if (joints)
            {
                for (int32_t k = gid; k < comps; k += gsz)
                {
                    m3JointComponentPass(&v, k, mode, sp.invH);
                }
                gather(mode, firstColor);
                sync();
            }
            m3TapePass* rec =
                tape != nullptr ? tape->pass + ((int64_t)sub * 2 + (mode - 1)) * tape->pairCap : nullptr;
            for (int32_t color = 0; color < tailStart && color <= M3_GRAPH_COLORS; ++color)
            {
                int32_t start = starts[color];
                int32_t end = starts[color + 1];
                if (end == start)
                {
                    continue;
                }
                ColorPhase(v, start, end, mode, sp, gid, gsz, color == M3_GRAPH_COLORS, rec, &pf);
                gather(mode, nextColor(color));
                sync();
            }
            if (kGridTail && tailStart <= M3_GRAPH_COLORS && nb <= M3_TAIL_LDS_BODIES)
            {
                if (blockIdx.x == 0)
                {
                    const int32_t tb = starts[tailStart];
                    const int32_t te = starts[M3_GRAPH_COLORS];
                    const uint32_t tailColors =
                        ((1u << M3_GRAPH_COLORS) - 1u) & ~((1u << tailStart) - 1u);
                    __shared__ uint8_t tailDone[M3_TAIL_LDS_BODIES];
                    if (mode == 1)
                    {
                        TailFlow<1>(v, sp, tb, te, tailColors, tailDone);
                    }
                    else
                    {
                        TailFlow<2>(v, sp, tb, te, tailColors, tailDone);
                    }
                    __syncthreads();
                    ColorPhase(v, starts[M3_GRAPH_COLORS], starts[M3_GRAPH_COLORS + 1], mode, sp,
                               (int32_t)threadIdx.x, (int32_t)blockDim.x, true, rec, &pf);
                }
                sync();
            }
            else if (tailStart <= M3_GRAPH_COLORS)
            {
                if (blockIdx.x == 0)
                {
                    for (int32_t color = tailStart; color <= M3_GRAPH_COLORS; ++color)
                    {
                        int32_t start = starts[color];
                        int32_t end = starts[color + 1];
                        if (end == start)
                        {
                            continue;
                        }
                        ColorPhase(v, start, end, mode, sp, (int32_t)threadIdx.x, (int32_t)blockDim.x,
                                   color == M3_GRAPH_COLORS, rec, &pf);
                        gather(mode, nextColor(color));
                        __syncthreads();
                    }
                }
                sync();
            }
        }
    }
    if (tape != nullptr)
    {
        const int32_t nc = v.sc->constraintCount;
        for (int32_t k = gid; k < nc; k += gsz)
        {
            TapeSolve(v, &v.cons[k], &tape->fin[k]);
        }
        for (int32_t b = gid; b < nb; b += gsz)
        {
            tape->xfEnd[b] = v.xf[b];
        }
    }
}

// ---------------------------------------------------------------------------
// Barrier-free solve and relax passes over cells (single world)
//
// Inside a pass, a contact constraint only needs each of its dynamic bodies
// to have received exactly the constraints of lower colors on that body:
// the colored pass gives each body its constraints in color order, and
// solves on different bodies commute.  So instead of a grid barrier per
// color, a constraint waits for its two bodies alone: a body's version is
// pass << 8 | the number of its constraints solved in the pass, and a
// constraint of rank r on a body (the number of colors below its own in
// which the body has a constraint, from the coloring) waits for version
// pass << 8 | r and leaves r + 1.
//
// Where a body's velocity travels decides the cost of that wait, so the
// island is cut into cells, one workgroup each (k_cells: a quantile grid
// over the movers' x and z).  A body lives in its cell's LDS; a constraint
// belongs to the cell of one of its dynamic bodies (k_cells).  When
// the previous constraint on a body was solved in the same cell, the wait
// is on an LDS version.  Otherwise the body's velocity comes from its six
// global words: each holds one float and the version, written and read
// with agent-scope atomics, so a reader on any XCD sees each word whole,
// and the version in the word says which write it holds (no fence between
// data and flag).  One wavefront of each workgroup (the mail wavefront)
// polls those words for the cell's pending imports and copies each into an
// LDS mailbox, so solver wavefronts only ever wait on LDS.  Every solve
// writes the global words of its dynamic bodies, so after a pass they hold
// the final velocities.
//
// Every participant takes its cell's constraints in slot order, every
// mail lane polls all its imports each round, and all workgroups are
// co-resident (a grid within the occupancy limit), so the lowest unsolved
// slot is always ready and the passes end.  Waits are bounded all the
// same: one that gives up marks the world faulted, and every wait ends once
// it sees the fault.  The relax pass takes the position deltas that
// integrate-positions left from versioned per-body records.  Between
// substeps no grid barrier: each cell publishes its finished substep
// (release) and the next substep of a cell waits (acquire) only for its
// neighbours, the cells whose plain stores it reads (warm-start addends,
// position deltas).
// ---------------------------------------------------------------------------
// A wait gives up after this much wall time (s_memrealtime runs at
// 100 MHz): far beyond any step, so only a broken wait reaches it.
#define M3_FLOW_TICKS 200000000ull
#define M3_CELL_SOLVERS 192 // solver threads of a cell workgroup
#define M3_CELL_MAIL 64     // the mail wavefront
#define M3_CELL_THREADS (M3_CELL_SOLVERS + M3_CELL_MAIL)

__device__ static inline uint64_t FlowWord(float x, uint32_t ver)
{
    return ((uint64_t)ver << 32) | (uint64_t)__float_as_uint(x);
}

__device__ static inline void FlowStore(uint64_t* rec, int32_t b, m3Vec3 lin, m3Vec3 ang, uint32_t ver)
{
    uint64_t* r = rec + 6 * (int64_t)b;
    __hip_atomic_store(&r[0], FlowWord(lin.x, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&r[1], FlowWord(lin.y, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&r[2], FlowWord(lin.z, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&r[3], FlowWord(ang.x, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&r[4], FlowWord(ang.y, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&r[5], FlowWord(ang.z, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reads a body's six words.  Returns the version all six carry, or
// 0xFFFFFFFF when they disagree (a write in progress).
__device__ static inline uint32_t FlowLoad(const uint64_t* rec, int32_t b, float* x)
{
    const uint64_t* r = rec + 6 * (int64_t)b;
    uint64_t w[6];
    M3_UNROLL
    for (int j = 0; j < 6; ++j)
    {
        w[j] = __hip_atomic_load(&r[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint32_t ver = (uint32_t)(w[5] >> 32);
    bool same = true;
    M3_UNROLL
    for (int j = 0; j < 6; ++j)
    {
        same = same && (uint32_t)(w[j] >> 32) == ver;
        x[j] = __uint_as_float((uint32_t)w[j]);
    }
    return same ? ver : 0xFFFFFFFFu;
}

// A pass whose position deltas have not moved since the step began reads
// them directly.
#define M3_DELTA_FIXED 0xFFFFFFFFu

__device__ static inline void DeltaStore(uint64_t* rec, int32_t b, m3Vec3 dp, m3Quat dq, uint32_t ver)
{
    uint64_t* r = rec + M3_DELTA_WORDS * (int64_t)b;
    const float x[7] = {dp.x, dp.y, dp.z, dq.x, dq.y, dq.z, dq.w};
    M3_UNROLL
    for (int j = 0; j < 7; ++j)
    {
        __hip_atomic_store(&r[j], FlowWord(x[j], ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// True when all seven words carry version ver; then dp and dq hold them.
__device__ static inline bool DeltaLoad(const uint64_t* rec, int32_t b, uint32_t ver, m3Vec3* dp, m3Quat* dq)
{
    const uint64_t* r = rec + M3_DELTA_WORDS * (int64_t)b;
    uint64_t w[7];
    M3_UNROLL
    for (int j = 0; j < 7; ++j)
    {
        w[j] = __hip_atomic_load(&r[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    bool same = true;
    M3_UNROLL
    for (int j = 0; j < 7; ++j)
    {
        same = same && (uint32_t)(w[j] >> 32) == ver;
    }
    *dp = m3V3(__uint_as_float((uint32_t)w[0]), __uint_as_float((uint32_t)w[1]), __uint_as_float((uint32_t)w[2]));
    dq->x = __uint_as_float((uint32_t)w[3]);
    dq->y = __uint_as_float((uint32_t)w[4]);
    dq->z = __uint_as_float((uint32_t)w[5]);
    dq->w = __uint_as_float((uint32_t)w[6]);
    return same;
}


// Quantile cells over the movers: gx slabs of equal mover count along x by
// gz of equal count along z.  Any partition gives the same results (the
// per-body operation order does not depend on it); a compact one keeps
// most consecutive constraints on a body in one cell.
#define M3_CELL_MAX 256
#define M3_CELL_BODIES 1024 // movers homed in one cell
#define M3_CELL_CONS 1024   // constraints owned by one cell
#define M3_CELL_BINS 1024   // histogram bins per axis (one per thread of k_cells)
#define M3_CELL_GX_MAX 16
#define M3_CELL_CACHE 12288 // movers whose x, z k_cells keeps in LDS
#define M3_CELL_LDS_BODIES 32768 // bodies whose cell k_cells keeps in LDS

__device__ static inline int32_t CellBin(float x, float lo, float scale)
{
    float f = (x - lo) * scale;
    if (!(f >= 0.0f))
    {
        f = 0.0f;
    }
    if (f > (float)(M3_CELL_BINS - 1))
    {
        f = (float)(M3_CELL_BINS - 1);
    }
    return (int32_t)f;
}

__device__ static inline float WaveMin(float x)
{
    for (int o = 32; o > 0; o >>= 1)
    {
        x = fminf(x, __shfl_xor(x, o));
    }
    return x;
}

__device__ static inline float WaveMax(float x)
{
    for (int o = 32; o > 0; o >>= 1)
    {
        x = fmaxf(x, __shfl_xor(x, o));
    }
    return x;
}

// Exclusive prefix sum over one value per thread of a 1024-thread block.
__device__ static inline int32_t BlockExclusiveSum(int32_t x, int32_t* waveTotals)
{
    const int32_t lane = threadIdx.x & 63;
    const int32_t wave = threadIdx.x >> 6;
    int32_t inc = x;
    for (int o = 1; o < 64; o <<= 1)
    {
        const int32_t y = __shfl_up(inc, o);
        if (lane >= o)
        {
            inc += y;
        }
    }
    if (lane == 63)
    {
        waveTotals[wave] = inc;
    }
    __syncthreads();
    int32_t before = 0;
    for (int32_t w = 0; w < wave; ++w)
    {
        before += waveTotals[w];
    }
    __syncthreads();
    return before + inc - x;
}

__global__ void __launch_bounds__(M3_CELL_BINS) k_cells(const m3Dev* __restrict__ ds, int32_t gx, int32_t gz)
{
    M3_WORLD;
    __shared__ float red[4][M3_CELL_BINS / 64];
    __shared__ float box[4];
    __shared__ int32_t hx[M3_CELL_BINS];
    __shared__ int32_t hz[M3_CELL_BINS];
    __shared__ uint8_t xOf[M3_CELL_BINS];
    __shared__ uint8_t zOf[M3_CELL_BINS];
    __shared__ int32_t totals[M3_CELL_BINS / 64];
    __shared__ int32_t cnt[M3_CELL_MAX];
    __shared__ int32_t consCnt[M3_CELL_MAX];
    __shared__ float2 pos[M3_CELL_CACHE];
    __shared__ uint8_t cellOfL[M3_CELL_LDS_BODIES];
    const int32_t t = threadIdx.x;
    const int32_t cells = gx * gz;
    const int32_t n = v.sc->stall != 0 ? 0 : v.sc->moverCount;
    if (t < M3_CELL_MAX)
    {
        cnt[t] = 0;
        consCnt[t] = 0;
        d.cellDone[t] = 0u;
    }
    hx[t] = 0;
    hz[t] = 0;
    // Bodies that are not movers this step belong to no cell.
    for (int32_t b = t; b < v.maxBody; b += blockDim.x)
    {
        d.cellOf[b] = 0xFF;
        if (b < M3_CELL_LDS_BODIES)
        {
            cellOfL[b] = 0xFF;
        }
    }
    auto at = [&](int32_t m) -> float2 {
        if (m < M3_CELL_CACHE)
        {
            return pos[m];
        }
        const m3Pos3 p = v.xf[v.movers[m]].p;
        return make_float2((float)p.x, (float)p.z);
    };
    float x0 = INFINITY, x1 = -INFINITY, z0 = INFINITY, z1 = -INFINITY;
    // The cached movers' loads are issued together (index, then position).
    constexpr int32_t kPer = M3_CELL_CACHE / M3_CELL_BINS;
    int32_t mb[kPer];
    M3_UNROLL
    for (int32_t j = 0; j < kPer; ++j)
    {
        const int32_t m = t + j * M3_CELL_BINS;
        mb[j] = m < n ? v.movers[m] : -1;
    }
    M3_UNROLL
    for (int32_t j = 0; j < kPer; ++j)
    {
        if (mb[j] >= 0)
        {
            const m3Pos3 p = v.xf[mb[j]].p;
            const float2 q = make_float2((float)p.x, (float)p.z);
            pos[t + j * M3_CELL_BINS] = q;
            x0 = fminf(x0, q.x);
            x1 = fmaxf(x1, q.x);
            z0 = fminf(z0, q.y);
            z1 = fmaxf(z1, q.y);
        }
    }
    for (int32_t m = M3_CELL_CACHE + t; m < n; m += blockDim.x)
    {
        const m3Pos3 p = v.xf[v.movers[m]].p;
        x0 = fminf(x0, (float)p.x);
        x1 = fmaxf(x1, (float)p.x);
        z0 = fminf(z0, (float)p.z);
        z1 = fmaxf(z1, (float)p.z);
    }
    x0 = WaveMin(x0);
    x1 = WaveMax(x1);
    z0 = WaveMin(z0);
    z1 = WaveMax(z1);
    if ((t & 63) == 0)
    {
        red[0][t >> 6] = x0;
        red[1][t >> 6] = x1;
        red[2][t >> 6] = z0;
        red[3][t >> 6] = z1;
    }
    __syncthreads();
    if (t == 0)
    {
        float a = red[0][0], b = red[1][0], c = red[2][0], e = red[3][0];
        for (int32_t w = 1; w < M3_CELL_BINS / 64; ++w)
        {
            a = fminf(a, red[0][w]);
            b = fmaxf(b, red[1][w]);
            c = fminf(c, red[2][w]);
            e = fmaxf(e, red[3][w]);
        }
        box[0] = a;
        box[1] = b > a ? (float)M3_CELL_BINS / (b - a) : 0.0f;
        box[2] = c;
        box[3] = e > c ? (float)M3_CELL_BINS / (e - c) : 0.0f;
    }
    __syncthreads();
    for (int32_t m = t; m < n; m += blockDim.x)
    {
        const float2 q = at(m);
        atomicAdd(&hx[CellBin(q.x, box[0], box[1])], 1);
        atomicAdd(&hz[CellBin(q.y, box[2], box[3])], 1);
    }
    __syncthreads();
    // Bin t's slab (row) is the quantile of its middle mover.
    const int32_t ex = BlockExclusiveSum(hx[t], totals);
    const int32_t ez = BlockExclusiveSum(hz[t], totals);
    const int64_t nn = n > 0 ? n : 1;
    const int32_t sx = (int32_t)(((int64_t)ex + hx[t] / 2) * gx / nn);
    const int32_t sz = (int32_t)(((int64_t)ez + hz[t] / 2) * gz / nn);
    xOf[t] = (uint8_t)(sx < gx ? sx : gx - 1);
    zOf[t] = (uint8_t)(sz < gz ? sz : gz - 1);
    __syncthreads();
    // The cached movers' bodies are still in registers (mb); the rest are
    // loaded again.
    for (int32_t m = t; m < n; m += blockDim.x)
    {
        const float2 q = at(m);
        int32_t b = 0;
        if (m < M3_CELL_CACHE)
        {
            M3_UNROLL
            for (int32_t j = 0; j < kPer; ++j)
            {
                b = t + j * M3_CELL_BINS == m ? mb[j] : b;
            }
        }
        else
        {
            b = v.movers[m];
        }
        const int32_t cell = xOf[CellBin(q.x, box[0], box[1])] * gz + zOf[CellBin(q.y, box[2], box[3])];
        const int32_t local = atomicAdd(&cnt[cell], 1);
        d.cellOf[b] = (uint8_t)cell;
        if (b < M3_CELL_LDS_BODIES)
        {
            cellOfL[b] = (uint8_t)cell;
        }
        d.cellLocal[b] = (uint16_t)(local < M3_CELL_BODIES ? local : 0);
        if (local < M3_CELL_BODIES)
        {
            d.cellBodies[cell * M3_CELL_BODIES + local] = b;
        }
    }
    // Bodies past the LDS copy are read back from global memory: every
    // wavefront's stores complete before the barrier, and the reads are
    // device-coherent.
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (t < cells)
    {
        d.cellCount[t] = cnt[t] < M3_CELL_BODIES ? cnt[t] : M3_CELL_BODIES;
        if (cnt[t] > M3_CELL_BODIES)
        {
            v.sc->flowVeto = 1;
        }
    }
    // The cell that solves each slot: the cell of its body A when A is
    // dynamic, else of body B.  A dynamic side outside every cell, or a
    // cell over its constraint capacity, sends the step to the colored
    // form.  A cell's list is in no particular order (the substep kernel
    // sorts it).
    auto cellOfBody = [&](int32_t b) -> uint32_t {
        if (b < 0)
        {
            return 0u;
        }
        if (b < M3_CELL_LDS_BODIES)
        {
            return cellOfL[b];
        }
        return __hip_atomic_load(&d.cellOf[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    const int32_t nc = v.sc->stall != 0 ? 0 : v.sc->colorStarts[M3_GRAPH_COLORS];
    bool veto = false;
    for (int32_t k0 = t; k0 < nc; k0 += 8 * M3_CELL_BINS)
    {
        int2 ab[8];
        M3_UNROLL
        for (int q = 0; q < 8; ++q)
        {
            const int32_t k = k0 + q * M3_CELL_BINS;
            ab[q] = k < nc ? d.slotBodies[k] : make_int2(-1, -1);
        }
        M3_UNROLL
        for (int q = 0; q < 8; ++q)
        {
            const int32_t k = k0 + q * M3_CELL_BINS;
            if (k >= nc)
            {
                continue;
            }
            const uint32_t ca = cellOfBody(ab[q].x);
            const uint32_t cb = cellOfBody(ab[q].y);
            const uint32_t owner = ab[q].x >= 0 ? ca : cb;
            d.slotOwner[k] = (uint8_t)owner;
            if ((ab[q].x >= 0 && ca >= (uint32_t)cells) || (ab[q].y >= 0 && cb >= (uint32_t)cells))
            {
                veto = true;
                continue;
            }
            const int32_t at = atomicAdd(&consCnt[owner], 1);
            if (at >= M3_CELL_CONS)
            {
                veto = true;
                continue;
            }
            d.cellCons[owner * M3_CELL_CONS + at] = k;
        }
    }
    if (veto)
    {
        v.sc->flowVeto = 1;
    }
    __syncthreads();
    if (t < cells)
    {
        d.cellConsCount[t] = consCnt[t];
    }
}

// Per-cell LDS of the substep kernel.
struct m3CellLds
{
    float state[M3_CELL_BODIES][6]; // velocities of the cell's bodies
    uint32_t ver[M3_CELL_BODIES];
    uint8_t deg[M3_CELL_BODIES]; // constraints on each of the cell's bodies
    float mail[2 * M3_CELL_CONS][6]; // imported velocities
    uint32_t mailVer[2 * M3_CELL_CONS];
    int32_t mailBody[2 * M3_CELL_CONS];
    int32_t slot[M3_CELL_CONS]; // the cell's constraints, ascending
    uint16_t loc[M3_CELL_CONS][2]; // per side: where it reads (LDS body or mailbox index)
    uint16_t own[M3_CELL_CONS][2]; // per side: the LDS body it writes (home sides)
    uint8_t flags[M3_CELL_CONS];   // per side: bit 0 / 2 read LDS body, bit 1 / 3 write LDS body
    uint32_t nbr[M3_CELL_MAX / 32]; // cells whose finished substeps this one waits for
    int32_t count;
    int32_t imports;
};

// Sorts the cell's slots ascending (bitonic, in LDS; count <= M3_CELL_CONS).
__device__ static inline void CellSort(m3CellLds& L, int32_t count)
{
    int32_t n = 1;
    while (n < count)
    {
        n <<= 1;
    }
    for (int32_t i = count + (int32_t)threadIdx.x; i < n; i += blockDim.x)
    {
        L.slot[i] = 0x7FFFFFFF;
    }
    __syncthreads();
    for (int32_t size = 2; size <= n; size <<= 1)
    {
        for (int32_t stride = size >> 1; stride > 0; stride >>= 1)
        {
            for (int32_t i = threadIdx.x; i < n; i += blockDim.x)
            {
                const int32_t j = i ^ stride;
                if (j > i)
                {
                    const int32_t a = L.slot[i];
                    const int32_t b = L.slot[j];
                    const bool up = (i & size) == 0;
                    if ((a > b) == up)
                    {
                        L.slot[i] = b;
                        L.slot[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Warm start of awake dynamic body b from its constraints' addends, in
// color order, four constraints' loads at a time.
__device__ static inline void FlowWarmStart(const m3View& v, const m3WarmAdd* warm, int32_t b,
                                            m3Vec3* lin, m3Vec3* ang)
{
    uint32_t colors = v.slotMask[b] & 0xFFFFu;
    const int32_t* row = v.slotTable + (int64_t)b * M3_GRAPH_COLORS;
    while (colors != 0u)
    {
        int32_t e[4];
        int32_t n = 0;
        M3_UNROLL
        for (int q = 0; q < 4; ++q)
        {
            if (colors != 0u)
            {
                e[q] = row[__builtin_ctz(colors)];
                colors &= colors - 1u;
                n = q + 1;
            }
            else
            {
                e[q] = -1;
            }
        }
        m3WarmAdd a[4];
        M3_UNROLL
        for (int q = 0; q < 4; ++q)
        {
            if (q < n)
            {
                a[q] = warm[e[q]]; // slot << 1 | side indexes the pair of records
            }
        }
        M3_UNROLL
        for (int q = 0; q < 4; ++q)
        {
            if (q < n)
            {
                m3WarmApply(&a[q], e[q] & 1, lin, ang);
            }
        }
    }
}

__device__ static inline void CellPut(m3CellLds& L, int32_t l, m3Vec3 lin, m3Vec3 ang, uint32_t ver)
{
    L.state[l][0] = lin.x;
    L.state[l][1] = lin.y;
    L.state[l][2] = lin.z;
    L.state[l][3] = ang.x;
    L.state[l][4] = ang.y;
    L.state[l][5] = ang.z;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __hip_atomic_store(&L.ver[l], ver, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The warm-start addends of a prepared or solved constraint (both sides).
__device__ static inline void CellAddends(const m3Constraint* con, m3WarmAdd* out)
{
    m3Vec3 rA[M3_MANIFOLD_MAX_POINTS];
    m3Vec3 rB[M3_MANIFOLD_MAX_POINTS];
    m3real imp[M3_MANIFOLD_MAX_POINTS];
    M3_UNROLL
    for (int32_t q = 0; q < M3_MANIFOLD_MAX_POINTS; ++q)
    {
        rA[q] = con->points[q].rA;
        rB[q] = con->points[q].rB;
        imp[q] = con->points[q].normalImpulse;
    }
    m3WarmAddends(con->pointCount, con->normal, rA, rB, imp, con->invMassA, con->invMassB, &con->invIA,
                  &con->invIB, con->t1, con->t2, con->frictionImpulse1, con->frictionImpulse2, con->originA,
                  con->originB, con->twistImpulse, con->rollingImpulse, con->dynA, con->dynB, &out[0],
                  &out[1]);
}

// Waits are bounded in wall time: every 256 rounds a waiting participant
// checks the clock (from the first check of this wait) and the world's
// fault flag; one that gives up marks the world faulted.  `since` is 0 at
// the start of each wait.
__device__ static inline bool FlowGiveUp(uint32_t* spins, uint64_t* since, int32_t* fault)
{
    if ((++*spins & 255u) != 0u)
    {
        return false;
    }
    const uint64_t now = __builtin_amdgcn_s_memrealtime();
    if (*since == 0ull)
    {
        *since = now;
    }
    if (now - *since > M3_FLOW_TICKS ||
        __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    {
        __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
    return false;
}

// One pass (kMode 1 solve, 2 relax) of one cell.  Solver threads take the
// cell's constraints in slot order; the mail wavefront fills the
// mailboxes.  Dynamic sides read their position deltas from the delta
// records at version deltaVer (M3_DELTA_FIXED: directly).  A relax pass
// that precedes another substep (`addends`) also leaves each constraint's
// warm-start addends for that substep.
template <int kMode>
__device__ static inline void CellPass(const m3View& v, const m3SubstepParams& sp, m3CellLds& L,
                                       uint64_t* rec, const uint32_t* rank, const uint64_t* drec,
                                       m3WarmAdd* warm, uint32_t deltaVer, bool addends, uint32_t pass,
                                       const m3ContactIn* first)
{
    const uint32_t base = pass << 8;
    int32_t* fault = &v.sc->fault;
    const int32_t count = L.count;
    const int32_t t = threadIdx.x;
    if (t >= M3_CELL_SOLVERS)
    {
        // Mail: each lane polls its imports (import q of item i waits for
        // the version the item's rank names) until all are in.
        const int32_t imports = L.imports;
        uint32_t spins = 0;
        uint64_t since = 0;
        uint64_t left = 0; // bit j: import t - M3_CELL_SOLVERS + 64 j still pending
        for (int32_t q = t - M3_CELL_SOLVERS, j = 0; q < imports; q += M3_CELL_MAIL, ++j)
        {
            left |= 1ull << j;
        }
        while (__any(left != 0ull))
        {
            bool got = false;
            for (uint64_t m = left; m != 0ull;)
            {
                int32_t js[4];
                M3_UNROLL
                for (int g = 0; g < 4; ++g)
                {
                    js[g] = -1;
                    if (m != 0ull)
                    {
                        js[g] = (int32_t)__builtin_ctzll(m);
                        m &= m - 1ull;
                    }
                }
                float x[4][6];
                uint32_t got4[4];
                M3_UNROLL
                for (int g = 0; g < 4; ++g)
                {
                    got4[g] = 0xFFFFFFFFu;
                    if (js[g] >= 0)
                    {
                        got4[g] = FlowLoad(rec, L.mailBody[t - M3_CELL_SOLVERS + M3_CELL_MAIL * js[g]], x[g]);
                    }
                }
                M3_UNROLL
                for (int g = 0; g < 4; ++g)
                {
                    if (js[g] < 0)
                    {
                        continue;
                    }
                    const int32_t q = t - M3_CELL_SOLVERS + M3_CELL_MAIL * js[g];
                    const uint32_t want = L.mailVer[q] & 0x7FFFFFFFu;
                    if (got4[g] == want)
                    {
                        for (int c = 0; c < 6; ++c)
                        {
                            L.mail[q][c] = x[g][c];
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                        __hip_atomic_store(&L.mailVer[q], want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        left &= ~(1ull << js[g]);
                        got = true;
                    }
                }
            }
            if (!got)
            {
                if (left != 0ull && FlowGiveUp(&spins, &since, fault))
                {
                    left = 0ull;
                }
            }
            else
            {
                spins = 0;
                since = 0;
            }
        }
        return;
    }
    int32_t i = t;
    bool have = i < count;
    int32_t k = 0;
    m3ContactIn in;
    m3ContactBias cb;
    m3Vec3 sLinA, sAngA, sLinB, sAngB;
    uint32_t wantA = 0, wantB = 0;
    uint32_t fl = 0;
    int32_t locA = 0, locB = 0, ownA = 0, ownB = 0;
    bool biased = false;
    // The bias terms, once both sides' deltas are in.
    auto bias = [&]() -> bool {
        m3Vec3 dpA, dpB;
        m3Quat dqA, dqB;
        if (deltaVer == M3_DELTA_FIXED || !in.dynA)
        {
            dpA = v.deltaPos[in.a];
            dqA = v.deltaRot[in.a];
        }
        else if (!DeltaLoad(drec, in.a, deltaVer, &dpA, &dqA))
        {
            return false;
        }
        if (deltaVer == M3_DELTA_FIXED || !in.dynB)
        {
            dpB = v.deltaPos[in.b];
            dqB = v.deltaRot[in.b];
        }
        else if (!DeltaLoad(drec, in.b, deltaVer, &dpB, &dqB))
        {
            return false;
        }
        M3_UNROLL
        for (int32_t q = 0; q < M3_MANIFOLD_MAX_POINTS; ++q)
        {
            cb.bias[q] = 0.0f;
            cb.massScale[q] = 1.0f;
            cb.impulseScale[q] = 0.0f;
        }
        const int32_t slots =
            in.pointCount < M3_MANIFOLD_MAX_POINTS ? in.pointCount : M3_MANIFOLD_MAX_POINTS;
        m3ContactBiasTerms(&in, dpA, dqA, dpB, dqB, sp.invH, sp.push, kMode == 1, slots, &cb);
        return true;
    };
    auto take = [&](bool loaded) {
        k = L.slot[i];
        if (!loaded)
        {
            m3LoadContact(&v.cons[k], kMode == 1, &in);
        }
        const uint32_t rk = rank[k];
        wantA = base | (rk & 0xFFu);
        wantB = base | ((rk >> 8) & 0xFFu);
        fl = L.flags[i];
        locA = L.loc[i][0];
        locB = L.loc[i][1];
        ownA = L.own[i][0];
        ownB = L.own[i][1];
        // Non-dynamic sides keep their velocities for the whole pass.
        sLinA = v.lv[in.a];
        sAngA = v.av[in.a];
        sLinB = v.lv[in.b];
        sAngB = v.av[in.b];
        biased = bias();
    };
    if (have)
    {
        in = *first;
        take(true);
    }
    uint32_t spins = 0;
    uint64_t since = 0;
    while (__any(have))
    {
        bool progressed = false;
        if (have && !biased)
        {
            biased = bias();
        }
        if (have)
        {
            // A side reads its cell body (bit 0 / 2) or its mailbox.
            bool ready = biased;
            m3PairVelocity s;
            s.vA = sLinA;
            s.wA = sAngA;
            s.vB = sLinB;
            s.wB = sAngB;
            for (int side = 0; side < 2; ++side)
            {
                const bool dyn = side == 0 ? in.dynA != 0 : in.dynB != 0;
                if (!dyn || !ready)
                {
                    continue;
                }
                const uint32_t want = side == 0 ? wantA : wantB;
                const int32_t l = side == 0 ? locA : locB;
                const bool body = (fl >> (2 * side)) & 1u;
                const uint32_t seen =
                    __hip_atomic_load(body ? &L.ver[l] : &L.mailVer[l], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
                if (seen != want)
                {
                    ready = false;
                    continue;
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                const float* x = body ? L.state[l] : L.mail[l];
                const m3Vec3 lin = m3V3(x[0], x[1], x[2]);
                const m3Vec3 ang = m3V3(x[3], x[4], x[5]);
                if (side == 0)
                {
                    s.vA = lin;
                    s.wA = ang;
                }
                else
                {
                    s.vB = lin;
                    s.wB = ang;
                }
            }
            if (ready)
            {
                const int32_t slots = m3ContactSlots(in.pointCount);
                m3ContactImpulses(&v.cons[k], &in, &cb, kMode == 1, slots, &s);
                // Global words first: once the LDS version lets another
                // wavefront of this cell write the body's next version,
                // this one's stores to the same words are already issued
                // ahead of it, so the next version cannot be overwritten by
                // this older one landing late.
                if (in.dynA)
                {
                    FlowStore(rec, in.a, s.vA, s.wA, wantA + 1u);
                }
                if (in.dynB)
                {
                    FlowStore(rec, in.b, s.vB, s.wB, wantB + 1u);
                }
                if (in.dynA && (fl & 2u))
                {
                    CellPut(L, ownA, s.vA, s.wA, wantA + 1u);
                }
                if (in.dynB && (fl & 8u))
                {
                    CellPut(L, ownB, s.vB, s.wB, wantB + 1u);
                }
                progressed = true;
                spins = 0;
                since = 0;
                i += M3_CELL_SOLVERS;
                have = i < count;
                if (have)
                {
                    take(false);
                }
            }
            else if (FlowGiveUp(&spins, &since, fault))
            {
                have = false;
            }
        }
        if (!__any(progressed))
        {
            __builtin_amdgcn_s_sleep(1);
        }
    }
    // The next substep's warm-start addends, from the accumulators this
    // pass left (off the pass's dependency chains).
    if (kMode == 2 && addends)
    {
        for (int32_t j = t; j < count; j += M3_CELL_SOLVERS)
        {
            const int32_t kk = L.slot[j];
            CellAddends(&v.cons[kk], &warm[2 * kk]);
        }
    }
}

// Cell prologue: the cell's constraints in slot order, each side's place
// (cell body or mailbox), and the imports with the versions they wait for
// (bit 31 set: not yet in).  A side reads its cell body when the body is
// homed here and the body's previous operation in the pass (the previous
// constraint in color order, or the phase operation for rank 0) also ran
// here; every other dynamic side is an import.
__device__ static inline void CellSetup(const m3View& v, const m3Dev& d, m3CellLds& L, int32_t cell)
{
    const int32_t t = threadIdx.x;
    const int32_t nHome = d.cellCount[cell];
    for (int32_t j = t; j < nHome; j += blockDim.x)
    {
        L.deg[j] = (uint8_t)__popc(v.slotMask[d.cellBodies[cell * M3_CELL_BODIES + j]] & 0xFFFFu);
    }
    const int32_t total = d.cellConsCount[cell];
    const int32_t count = total < M3_CELL_CONS ? total : M3_CELL_CONS;
    if (t == 0)
    {
        L.count = count;
        L.imports = 0;
    }
    if (t < M3_CELL_MAX / 32)
    {
        L.nbr[t] = 0u;
    }
    for (int32_t i = t; i < count; i += blockDim.x)
    {
        L.slot[i] = d.cellCons[cell * M3_CELL_CONS + i];
    }
    __syncthreads();
    // Neighbours: the owners of the constraints on this cell's bodies (their
    // warm-start addends) and, below, the home cells of the bodies of this
    // cell's constraints (their position deltas).
    for (int32_t j = t; j < nHome; j += blockDim.x)
    {
        const int32_t b = d.cellBodies[cell * M3_CELL_BODIES + j];
        for (uint32_t m = v.slotMask[b] & 0xFFFFu; m != 0u; m &= m - 1u)
        {
            const int32_t slot = v.slotTable[(int64_t)b * M3_GRAPH_COLORS + __builtin_ctz(m)] >> 1;
            const uint32_t o = d.slotOwner[slot];
            if (o != (uint32_t)cell && o < M3_CELL_MAX)
            {
                atomicOr(&L.nbr[o >> 5], 1u << (o & 31u));
            }
        }
    }
    CellSort(L, count);
    for (int32_t i = t; i < count; i += blockDim.x)
    {
        const int32_t k = L.slot[i];
        const m3Constraint* c = &v.cons[k];
        const uint32_t lower = (1u << v.slotColors[k]) - 1u;
        uint32_t fl = 0;
        for (int side = 0; side < 2; ++side)
        {
            const bool dyn = side == 0 ? c->dynA != 0 : c->dynB != 0;
            if (!dyn)
            {
                L.loc[i][side] = 0;
                L.own[i][side] = 0;
                continue;
            }
            const int32_t b = side == 0 ? c->bodyA : c->bodyB;
            const bool home = d.cellOf[b] == (uint8_t)cell;
            if (!home && d.cellOf[b] < M3_CELL_MAX)
            {
                const uint32_t hc = d.cellOf[b];
                atomicOr(&L.nbr[hc >> 5], 1u << (hc & 31u));
            }
            const uint32_t below = v.slotMask[b] & lower;
            bool prevHere = below == 0u;
            if (!prevHere)
            {
                const int32_t prevColor = 31 - __clz(below);
                const int32_t prev = v.slotTable[(int64_t)b * M3_GRAPH_COLORS + prevColor] >> 1;
                prevHere = d.slotOwner[prev] == (uint8_t)cell;
            }
            L.own[i][side] = home ? d.cellLocal[b] : 0;
            if (home)
            {
                fl |= 2u << (2 * side);
            }
            if (home && prevHere)
            {
                fl |= 1u << (2 * side);
                L.loc[i][side] = d.cellLocal[b];
            }
            else
            {
                const int32_t q = atomicAdd(&L.imports, 1);
                L.loc[i][side] = (uint16_t)q;
                L.mailBody[q] = b;
                // The wanted version's pass is filled in by the pass.
                L.mailVer[q] = 0x80000000u | (d.flowRank[k] >> (8 * side) & 0xFFu);
            }
        }
        L.flags[i] = (uint8_t)fl;
    }
    __syncthreads();
}

// A home body's velocities after pass `pass`: from LDS when the body's last
// constraint of the pass ran in this cell (its LDS version then counts
// all of them), else from its global words once all six carry that last
// constraint's version.  An acknowledged device-scope store from another
// XCD can still be on its way after the grid barrier; waiting for the
// final version also means no older store to these words can land after
// the body phase rewrites them (every solve saw its predecessor's version
// in each word before writing).
__device__ static inline void CellFinal(const m3CellLds& L, const uint64_t* rec, int32_t j, int32_t b,
                                        uint32_t pass, m3Vec3* lin, m3Vec3* ang, int32_t* fault)
{
    float x[6];
    const uint32_t want = (pass << 8) | L.deg[j];
    if (L.ver[j] == want)
    {
        for (int c = 0; c < 6; ++c)
        {
            x[c] = L.state[j][c];
        }
    }
    else
    {
        uint32_t spins = 0;
        uint64_t since = 0;
        while (FlowLoad(rec, b, x) != want && !FlowGiveUp(&spins, &since, fault))
        {
        }
    }
    *lin = m3V3(x[0], x[1], x[2]);
    *ang = m3V3(x[3], x[4], x[5]);
}

// The first constraint of a solver thread, loaded before the barrier that
// opens a pass (its fields are fixed or written only by this thread).
template <int kMode>
__device__ static inline void CellFirst(const m3View& v, const m3CellLds& L, m3ContactIn* in)
{
    const int32_t t = threadIdx.x;
    if (t < M3_CELL_SOLVERS && t < L.count)
    {
        m3LoadContact(&v.cons[L.slot[t]], kMode == 1, in);
    }
}

// Arms the mailboxes for a pass: each waits for version pass << 8 | rank.
__device__ static inline void CellArm(m3CellLds& L, uint32_t pass)
{
    for (int32_t q = threadIdx.x; q < L.imports; q += blockDim.x)
    {
        L.mailVer[q] = 0x80000000u | (pass << 8) | (L.mailVer[q] & 0xFFu);
    }
}

// After its relax pass (and the addends it leaves) a cell publishes the
// substep as done: every wavefront's stores are acknowledged, then one
// agent-scope release makes them visible to any XCD that acquires.
__device__ static inline void CellSignal(uint32_t* done, int32_t cell, uint32_t substeps)
{
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&done[cell], substeps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Before substep `sub` a cell waits until each neighbour has finished
// substep sub - 1 (lanes of wavefront 0 poll four cells each), then
// acquires, so the neighbours' addends and position deltas are visible.
__device__ static inline void CellWait(const m3CellLds& L, const uint32_t* done, uint32_t sub, int32_t* fault)
{
    if (threadIdx.x < 64)
    {
        for (int32_t c = (int32_t)threadIdx.x; c < M3_CELL_MAX; c += 64)
        {
            if ((L.nbr[c >> 5] >> (c & 31)) & 1u)
            {
                uint32_t spins = 0;
                uint64_t since = 0;
                while (__hip_atomic_load(&done[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < sub &&
                       !FlowGiveUp(&spins, &since, fault))
                {
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// The substep loop of src/solver.c:4085-4203 over cells.  Body phases run
// in each body's home cell.
__device__ static inline void SubstepCells(const m3View& v, const m3Dev& d, const m3SubstepParams& sp,
                                           m3GridSync sync, m3CellLds& L)
{
    const int32_t cell = blockIdx.x;
    const int32_t t = threadIdx.x;
    const int32_t nHome = d.cellCount[cell];
    const int32_t* home = d.cellBodies + cell * M3_CELL_BODIES;
    uint64_t* rec = d.flowRec;
    int32_t* fault = &v.sc->fault;
    // A kinematic body has no solve count, and contacts read its deltas
    // directly: a contact on one keeps the grid barriers around
    // integrate-positions.
    const bool kinematic = v.sc->flowKinematic != 0;
    CellSetup(v, d, L, cell);
    for (int32_t sub = 0; sub < sp.substeps; ++sub)
    {
        const uint32_t solvePass = 2u * (uint32_t)sub;
        if (sub > 0 && !kinematic)
        {
            CellWait(L, d.cellDone, (uint32_t)sub, fault);
        }
        for (int32_t j = t; j < nHome; j += blockDim.x)
        {
            const int32_t b = home[j];
            if (sub > 0)
            {
                m3Vec3 lin, ang;
                CellFinal(L, rec, j, b, solvePass - 1u, &lin, &ang, fault);
                v.lv[b] = lin;
                v.av[b] = ang;
            }
            m3IntegrateVelocity(&v, b, sp.h, sp.gravity, sp.maxLinear, sp.maxAngular);
            m3Vec3 lin = v.lv[b];
            m3Vec3 ang = v.av[b];
            if (v.type[b] == M3_DYNAMIC)
            {
                FlowWarmStart(v, d.warmAdd, b, &lin, &ang);
            }
            v.lv[b] = lin;
            v.av[b] = ang;
            FlowStore(rec, b, lin, ang, solvePass << 8); // global before LDS, as in CellPass
            CellPut(L, j, lin, ang, solvePass << 8);
        }
        // No grid barrier opens a pass: a constraint waits for its bodies'
        // versions.  The solve pass reads the deltas directly (the last
        // grid barrier followed their writes); the relax pass reads them
        // from the delta records.  The workgroup barriers keep the
        // mailboxes armed until the whole cell is through a pass.
        CellArm(L, solvePass);
        {
            m3ContactIn first;
            CellFirst<1>(v, L, &first);
            __syncthreads();
            CellPass<1>(v, sp, L, rec, d.flowRank, d.deltaRec, d.warmAdd, M3_DELTA_FIXED, false, solvePass,
                        &first);
        }
        if (kinematic)
        {
            sync();
        }
        else
        {
            __syncthreads();
        }
        // A dynamic body waits for its last solve, after which no solve
        // anywhere reads its deltas.
        for (int32_t j = t; j < nHome; j += blockDim.x)
        {
            const int32_t b = home[j];
            m3Vec3 lin, ang;
            CellFinal(L, rec, j, b, solvePass, &lin, &ang, fault);
            v.lv[b] = lin;
            v.av[b] = ang;
            m3IntegratePosition(&v, b, sp.h);
            DeltaStore(d.deltaRec, b, v.deltaPos[b], v.deltaRot[b], (solvePass + 1u) << 8);
            FlowStore(rec, b, v.lv[b], v.av[b], (solvePass + 1u) << 8);
            CellPut(L, j, v.lv[b], v.av[b], (solvePass + 1u) << 8);
        }
        CellArm(L, solvePass + 1u);
        {
            m3ContactIn first;
            CellFirst<2>(v, L, &first);
            if (kinematic)
            {
                sync();
            }
            else
            {
                __syncthreads();
            }
            CellPass<2>(v, sp, L, rec, d.flowRank, d.deltaRec, d.warmAdd, (solvePass + 1u) << 8,
                        sub + 1 < sp.substeps, solvePass + 1u, &first);
        }
        // The next warm start reads the addends the relax pass left, and the
        // next solve pass the position deltas: the neighbours that wrote
        // them (owners of this cell's bodies' constraints, homes of its
        // constraints' bodies) are waited for at the top of the loop.  A
        // contact on a kinematic body keeps the grid barrier.
        if (kinematic)
        {
            sync();
        }
        else
        {
            CellSignal(d.cellDone, cell, (uint32_t)sub + 1u);
        }
    }
    for (int32_t j = t; j < nHome; j += blockDim.x)
    {
        const int32_t b = home[j];
        m3Vec3 lin, ang;
        CellFinal(L, rec, j, b, 2u * (uint32_t)sp.substeps - 1u, &lin, &ang, fault);
        v.lv[b] = lin;
        v.av[b] = ang;
    }
}

// One world over a cooperative grid: over cells when every constraint of
// the step fits them (see SubstepCells), else color by color.
__global__ void __launch_bounds__(M3_CELL_THREADS)
    k_substeps(const m3Dev* __restrict__ ds, m3SubstepParams sp)
{
    M3_WORLD;
    if (v.sc->stall != 0)
    {
        return;
    }
    __shared__ m3CellLds cellLds;
    const int32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t gsz = gridDim.x * blockDim.x;
    const m3GridSync sync{d.barrier, &v.sc->fault};
    // Every input of the choice was written by earlier launches, so all
    // participants take the same branch.
    const bool cells = sp.flow != 0 && d.flowRec != nullptr && 2 * sp.substeps <= M3_FLOW_PASSES &&
                       v.sc->jointActive == 0 && v.sc->flowVeto == 0 &&
                       v.sc->colorStarts[M3_GRAPH_COLORS] == v.sc->colorStarts[M3_GRAPH_COLORS + 1];
    if (cells)
    {
        SubstepCells(v, d, sp, sync, cellLds);
    }
    else
    {
        SubstepLoop<m3GridSync, true>(v, sp, gid, gsz, sync);
    }
}

// A batch: worlds are independent, so each workgroup runs its own world's
// whole substep loop with workgroup barriers.
#define M3_BLOCK_MAX_THREADS 512

__global__ void __launch_bounds__(M3_BLOCK_MAX_THREADS)
    k_substeps_block(const m3Dev* __restrict__ ds, m3SubstepParams sp)
{
    M3_WORLD;
    if (v.sc->stall != 0)
    {
        return;
    }
    SubstepLoop(v, sp, (int32_t)threadIdx.x, (int32_t)blockDim.x, m3BlockSync{}, d.tape);
}

__global__ void k_restitution(const m3Dev* __restrict__ ds, float threshold)
{
    M3_WORLD;
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

// The impulse store, and the applied forces of every mover cleared (the
// grid covers the larger of the pair and body capacities).

// ---------------------------------------------------------------------------
// Continuous collision and sleep (src/solver.c:3349-3430, 3582-3660)
// ---------------------------------------------------------------------------

// Continuous collision (src/solver.c:3349-3430).  Without bullets a fast
// body sweeps only against static shapes and planes, which the pass never
// moves, so bodies run in parallel and apply their own pull-back.  With
// bullets, k_finish computes the other bodies into ccdXf / ccdHit,
// k_ccd_bullets the bullets (each sees the targets of lower slots pulled
// back, as the reference's slot-order loop does; bullets never sweep against
// bullets), and k_ccd_apply applies both.
// Island count and sleep timer of body i; a body still too young to sleep
// clears its island's readiness flag (raised by the mover pass).
__device__ static inline void SleepTimer(const m3Dev& d, int32_t i, float dt, int sleepEnabled)
{
    const m3View& v = d.v;
    if (v.bodyAlive[i] == 0 || v.type[i] != M3_DYNAMIC || v.awake[i] == 0)
    {
        return;
    }
    const int32_t r = d.root[i];
    if (r == i)
    {
        atomicAdd(&v.sc->islandCount, 1);
    }
    if (sleepEnabled)
    {
        m3SleepTimer(&v, i, dt);
        if (v.sleepTime[i] < 0.5f)
        {
            v.flag[r] = 0;
        }
    }
}

// The step's closing work, one item per thread: constraint c's
// accumulators back into its manifold and mover c's force cleared
// (src/solver.c:4210-4240) in the workgroups below `bodyBase`; from
// `bodyBase` on, body i's continuous pass when `ccd` is set and, with
// `timers`, body i's sleep timer after its own continuous pass (without
// bullets a body's pass moves only that body).  The two halves touch
// disjoint data, and separate wavefronts run their load chains side by side.
__global__ void __launch_bounds__(64) k_finish(const m3Dev* __restrict__ ds, int ccd, int timers, float dt,
                                               int sleepEnabled, int32_t bodyBase)
{
    M3_WORLD;
    if (v.sc->stall != 0)
    {
        return;
    }
    if ((int32_t)blockIdx.x < bodyBase)
    {
        const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c < v.sc->moverCount)
        {
            const int32_t b = v.movers[c];
            v.force[b] = m3V3(0.0f, 0.0f, 0.0f);
            v.torque[b] = m3V3(0.0f, 0.0f, 0.0f);
        }
        if (c < v.sc->constraintCount)
        {
            m3StoreOne(&v, &v.cons[c]);
        }
        return;
    }
    const int32_t i = ((int32_t)blockIdx.x - bodyBase) * blockDim.x + threadIdx.x;
    if (i >= v.maxBody)
    {
        return;
    }
    if (ccd != 0)
    {
        if (d.ccdBullets == 0)
        {
            int pulled = m3ContinuousBody(&v, i);
            if (d.tape != nullptr)
            {
                d.tape->ccd[i] = (uint8_t)pulled;
            }
        }
        else
        {
            d.ccdHit[i] = 0;
            m3Transform xf;
            if (v.bullet[i] == 0 && m3ContinuousSweep(&v, nullptr, i, &xf))
            {
                d.ccdXf[i] = xf;
                d.ccdHit[i] = 1;
            }
        }
    }
    if (timers != 0)
    {
        SleepTimer(d, i, dt, sleepEnabled);
    }
}

__global__ void k_ccd_bullets(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0 || v.bullet[i] == 0)
    {
        return;
    }
    const m3CcdPoses poses = {d.ccdXf, d.ccdHit};
    m3Transform xf;
    if (m3ContinuousSweep(&v, &poses, i, &xf))
    {
        d.ccdXf[i] = xf;
        d.ccdHit[i] = 1;
    }
}

__global__ void k_ccd_apply(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0)
    {
        return;
    }
    if (d.ccdHit[i] != 0)
    {
        v.xf[i] = d.ccdXf[i];
    }
}

// The step's pair list over the first pairCount entries of the table; the
// rest keeps what earlier steps left (the reference's single pair array).
__global__ void k_pair_table(const m3Dev* __restrict__ ds, uint64_t* tableKeys, m3Manifold* tableMans)
{
    M3_WORLD;
    if (v.sc->stall != 0)
    {
        return;
    }
    const int32_t n = v.sc->pairCount;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        tableKeys[i] = v.keys[i];
        tableMans[i] = v.manifolds[i];
    }
}

// The sleep timers in a launch of their own (worlds with bullets, whose
// pulled-back poses land after every body's continuous pass).
__global__ void k_sleep_timers(const m3Dev* __restrict__ ds, float dt, int sleepEnabled)
{
    M3_WORLD;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.maxBody || v.sc->stall != 0)
    {
        return;
    }
    SleepTimer(d, i, dt, sleepEnabled);
}

// The step's close: thread 0 counts the step and ends the full query;
// with sleep on, every body of a ready island freezes.
__global__ void k_close(const m3Dev* __restrict__ ds, int sleepEnabled)
{
    M3_WORLD;
    const int32_t* root = d.root;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
    {
        m3Scalars* sc = v.sc;
        if (sc->stall == 0)
        {
            sc->stepCount += 1;
        }
        sc->fullQuery = 0;
        if (sc->sleepCount > v.pairCapacity)
        {
            sc->sleepCount = v.pairCapacity;
        }
    }
    if (sleepEnabled == 0 || i >= v.maxBody || v.sc->stall != 0 || v.bodyAlive[i] == 0 ||
        v.type[i] != M3_DYNAMIC || v.awake[i] == 0)
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
        if (d.tape != nullptr)
        {
            d.tape->slept[i] = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Event streams of the last step (src/solver.c:3732-3815 contact and sensor
// begin/end, :2991-3026 hits, :4295-4306 body moves), built when a host reads
// them.  The reference walks the old and new pair lists (both ascending)
// together; every begin event is a new pair (touching now, not touching or
// absent before) and every end event an old pair (touching before, not
// touching or absent now), so each stream is one list filtered in its own
// order.  Hits keep constraint order, which is pair order.
// ---------------------------------------------------------------------------

struct m3EventBufs
{
    uint8_t* codeNew;   // per new pair: 1 begin, 2 sensor begin, | 4 hit
    uint8_t* codeOld;   // per old pair: 1 end, 2 sensor end
    int32_t* counts;    // begin, end, sensor begin, sensor end, hits, moves, stalled
    uint64_t* lists;    // 4 x pairCapacity keys: begin, end, sensor begin, sensor end
    m3gpuHit* hits;     // pairCapacity
    m3gpuMove* moves;   // bodyCapacity
    int32_t* hostCounts; // pinned
};

// The index of key in keys[0..n) (ascending), or -1.
__device__ static inline int32_t FindKey(const uint64_t* keys, int32_t n, uint64_t key)
{
    int32_t lo = 0;
    int32_t hi = n;
    while (lo < hi)
    {
        const int32_t mid = (lo + hi) >> 1;
        const uint64_t k = keys[mid];
        if (k < key)
        {
            lo = mid + 1;
        }
        else if (k > key)
        {
            hi = mid;
        }
        else
        {
            return mid;
        }
    }
    return -1;
}

__device__ static inline uint8_t PairEventCode(const m3View& v, uint64_t key)
{
    const int32_t sA = (int32_t)(key >> 32);
    const int32_t sB = (int32_t)(key & 0xFFFFFFFFu);
    // A pair whose shape died emits nothing (its id would be stale).
    if (v.shapeAlive[sA] == 0 || v.shapeAlive[sB] == 0)
    {
        return 0;
    }
    return (v.sensor[sA] != 0 || v.sensor[sB] != 0) ? 2 : 1;
}

__global__ void k_event_codes(const m3Dev* __restrict__ ds, m3EventBufs e, float hitThreshold)
{
    M3_WORLD;
    const m3Scalars* sc = v.sc;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (sc->stall != 0)
    {
        return;
    }
    const int32_t np = sc->pairCount;
    const int32_t no = sc->oldPairCount;
    if (i < np)
    {
        const uint64_t key = v.keys[i];
        const int touchNew = v.manifolds[i].pointCount > 0;
        const int32_t j = FindKey(v.oldKeys, no, key);
        const int touchOld = j >= 0 && v.oldManifolds[j].pointCount > 0;
        uint8_t code = (touchNew && !touchOld) ? PairEventCode(v, key) : 0;
        if (v.hitRecs != nullptr && v.hitRecs[i].flag != 0 &&
            v.hitRecs[i].relativeVelocity < -hitThreshold)
        {
            code |= 4;
        }
        e.codeNew[i] = code;
    }
    if (i < no)
    {
        const uint64_t key = v.oldKeys[i];
        const int touchOld = v.oldManifolds[i].pointCount > 0;
        const int32_t j = FindKey(v.keys, np, key);
        const int touchNew = j >= 0 && v.manifolds[j].pointCount > 0;
        e.codeOld[i] = (touchOld && !touchNew) ? PairEventCode(v, key) : 0;
    }
    if (i < sc->moverCount)
    {
        const int32_t b = v.movers[i];
        m3gpuMove mv;
        mv.body = b;
        mv.fellAsleep = v.type[b] == M3_DYNAMIC && v.awake[b] == 0;
        mv.pad[0] = 0;
        mv.pad[1] = 0;
        mv.pad[2] = 0;
        const m3Transform xf = v.xf[b];
        mv.p[0] = xf.p.x;
        mv.p[1] = xf.p.y;
        mv.p[2] = xf.p.z;
        mv.q[0] = xf.q.x;
        mv.q[1] = xf.q.y;
        mv.q[2] = xf.q.z;
        mv.q[3] = xf.q.w;
        e.moves[i] = mv;
    }
}

#define M3_EVENT_THREADS 1024
#define M3_EVENT_LISTS 5

// One workgroup: the five filtered lists in order, by ballot ranks within a
// wavefront and per-wavefront offsets in LDS.
__global__ void __launch_bounds__(M3_EVENT_THREADS) k_event_compact(const m3Dev* __restrict__ ds, m3EventBufs e)
{
    M3_WORLD;
    const m3Scalars* sc = v.sc;
    constexpr int W = M3_EVENT_THREADS / 64;
    __shared__ int32_t waveCount[M3_EVENT_LISTS][W];
    __shared__ int32_t total[M3_EVENT_LISTS];
    const int t = (int)threadIdx.x;
    const int wave = t >> 6;
    const int lane = t & 63;
    const int stalled = sc->stall != 0;
    if (t < M3_EVENT_LISTS)
    {
        total[t] = 0;
    }
    __syncthreads();
    const int32_t np = stalled ? 0 : sc->pairCount;
    const int32_t no = stalled ? 0 : sc->oldPairCount;
    const int32_t n = np > no ? np : no;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int32_t base = 0; base < n; base += M3_EVENT_THREADS)
    {
        const int32_t i = base + t;
        const uint8_t cn = i < np ? e.codeNew[i] : 0;
        const uint8_t co = i < no ? e.codeOld[i] : 0;
        bool pred[M3_EVENT_LISTS];
        pred[0] = (cn & 3) == 1;
        pred[1] = (co & 3) == 1;
        pred[2] = (cn & 3) == 2;
        pred[3] = (co & 3) == 2;
        pred[4] = (cn & 4) != 0;
        uint64_t mask[M3_EVENT_LISTS];
        for (int k = 0; k < M3_EVENT_LISTS; ++k)
        {
            mask[k] = __ballot(pred[k]);
            if (lane == 0)
            {
                waveCount[k][wave] = __popcll(mask[k]);
            }
        }
        __syncthreads();
        for (int k = 0; k < M3_EVENT_LISTS; ++k)
        {
            if (!pred[k])
            {
                continue;
            }
            int32_t pos = total[k];
            for (int w2 = 0; w2 < wave; ++w2)
            {
                pos += waveCount[k][w2];
            }
            pos += __popcll(mask[k] & below);
            if (k == 0 || k == 2)
            {
                e.lists[(int64_t)k * v.pairCapacity + pos] = v.keys[i];
            }
            else if (k == 1 || k == 3)
            {
                e.lists[(int64_t)k * v.pairCapacity + pos] = v.oldKeys[i];
            }
            else
            {
                const m3HitRec r = v.hitRecs[i];
                m3gpuHit hit;
                hit.key = v.keys[i];
                hit.point[0] = r.point[0];
                hit.point[1] = r.point[1];
                hit.point[2] = r.point[2];
                hit.normal[0] = r.normal.x;
                hit.normal[1] = r.normal.y;
                hit.normal[2] = r.normal.z;
                hit.approachSpeed = -r.relativeVelocity;
                e.hits[pos] = hit;
            }
        }
        __syncthreads();
        if (t < M3_EVENT_LISTS)
        {
            int32_t sum = 0;
            for (int w2 = 0; w2 < W; ++w2)
            {
                sum += waveCount[t][w2];
            }
            total[t] += sum;
        }
        __syncthreads();
    }
    if (t < M3_EVENT_LISTS)
    {
        e.counts[t] = total[t];
    }
    if (t == 0)
    {
        e.counts[5] = stalled ? 0 : sc->moverCount;
        e.counts[6] = stalled;
    }
}

// ---------------------------------------------------------------------------
// Batches of small worlds: one workgroup per world replaces each
// device-wide sort and compaction, with the world's data staged in LDS.
// ---------------------------------------------------------------------------

#define M3_BATCH_THREADS 1024
#define M3_BATCH_EMIT_MAX 32768  // emitted pair keys a world sorts in LDS
#define M3_BATCH_SHAPES_MAX 4096 // shapes a world sweep-sorts in LDS

// Exclusive prefix sum over the block (blockDim.x a multiple of 64, at most
// 1024); *total receives the block sum.  Every thread must call it.
__device__ static inline int32_t BlockExclusiveScan(int32_t x, int32_t* total)
{
    __shared__ int32_t waveSum[M3_BATCH_THREADS / 64 + 1];
    const int32_t lane = threadIdx.x & 63;
    const int32_t wave = threadIdx.x >> 6;
    const int32_t waves = blockDim.x >> 6;
    int32_t incl = x;
    for (int32_t off = 1; off < 64; off <<= 1)
    {
        int32_t y = __shfl_up(incl, off);
        if (lane >= off)
        {
            incl += y;
        }
    }
    if (lane == 63)
    {
        waveSum[wave] = incl;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int32_t acc = 0;
        for (int32_t w = 0; w < waves; ++w)
        {
            int32_t t = waveSum[w];
            waveSum[w] = acc;
            acc += t;
        }
        waveSum[M3_BATCH_THREADS / 64] = acc;
    }
    __syncthreads();
    int32_t result = waveSum[wave] + incl - x;
    *total = waveSum[M3_BATCH_THREADS / 64];
    __syncthreads();
    return result;
}

// Bitonic sort of a[0, n) in LDS, n a power of two.
template <typename K>
__device__ static inline void BitonicSortLds(K* a, int32_t n)
{
    for (int32_t k = 2; k <= n; k <<= 1)
    {
        for (int32_t j = k >> 1; j > 0; j >>= 1)
        {
            for (int32_t i = threadIdx.x; i < n; i += blockDim.x)
            {
                int32_t l = i ^ j;
                if (l > i)
                {
                    K x = a[i];
                    K y = a[l];
                    bool ascending = (i & k) == 0;
                    if ((y < x) == ascending)
                    {
                        a[i] = y;
                        a[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ static inline int32_t NextPow2(int32_t n)
{
    int32_t p = 1;
    while (p < n)
    {
        p <<= 1;
    }
    return p;
}


__global__ void __launch_bounds__(M3_BATCH_THREADS) k_sap_sort_block(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    __shared__ uint64_t keys[M3_BATCH_SHAPES_MAX];
    const int32_t n = v.maxShape;
    const int32_t n2 = NextPow2(n);
    for (int32_t i = threadIdx.x; i < n2; i += blockDim.x)
    {
        keys[i] = i < n ? d.sapKeysIn[i] : M3_SENTINEL64;
    }
    __syncthreads();
    BitonicSortLds(keys, n2);
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x)
    {
        d.sapKeysOut[i] = keys[i];
    }
}

// The sweep keys of one world, sorted in one launch of 1024-thread
// workgroups.  Every key is distinct (its low word is the shape).
//   1. Workgroup r < runs sorts run r (keys [r R, r R + R), R at most
//      M3_RSORT_RUN) in LDS and writes it back in place: a stable LSD radix
//      sort of the high words, 4 bits a pass, each thread holding `per`
//      consecutive keys.  A thread counts its digits in registers (16
//      counters of 16 bits packed in four 64-bit words), DPP scans those
//      counters across each wavefront and wavefront 0 across the
//      wavefronts, and a key's place is its digit's base + the counts of
//      that digit in earlier threads + its rank among the thread's own, so
//      equal digits keep their order.  The keys arrive in shape order, so
//      the high words' stable order is the full keys' order.
//   2. After the grid barrier every workgroup loads the sorted runs into
//      LDS, and each key goes to its place in its own run plus, for every
//      other run, the number of that run's keys below it (binary search).
// Worlds over M3_RSORT_MAX shapes take the library sort.
#define M3_RSORT_THREADS 1024
#define M3_RSORT_ITEMS 2
#define M3_RSORT_RUN (M3_RSORT_THREADS * M3_RSORT_ITEMS)
#define M3_RSORT_MAX 14336
#define M3_RSORT_BITS 4

// Inclusive sum over a wavefront by DPP: shifts within each row of 16
// lanes, then each row's total broadcast into the rows above it.
__device__ static inline uint32_t WaveInclusiveSum(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false); // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false); // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false); // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false); // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false); // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false); // row_bcast:31
    return x;
}

// The same over packed 16-bit counters (no field exceeds 65535, so the
// 32-bit sums carry nothing between fields).
__device__ static inline uint64_t WaveInclusiveSum64(uint64_t x)
{
    const uint32_t lo = WaveInclusiveSum((uint32_t)x);
    const uint32_t hi = WaveInclusiveSum((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// 16-bit field d (0-15) of four packed words.
__device__ static inline uint32_t RsortField(const uint64_t* p, uint32_t d)
{
    const uint32_t q = d >> 2;
    const uint64_t w = q == 0 ? p[0] : (q == 1 ? p[1] : (q == 2 ? p[2] : p[3]));
    return (uint32_t)(w >> (16 * (d & 3))) & 0xFFFFu;
}

// Stable sort of keys[0, per * 1024) by their high words, in LDS.
__device__ static inline void RunSort(uint64_t* keys, int32_t per, uint64_t (*waveTot)[4], uint64_t* grand)
{
    const int32_t t = threadIdx.x;
    const int32_t lane = t & 63;
    const int32_t wave = t >> 6;
    uint64_t k[M3_RSORT_ITEMS];
    M3_UNROLL
    for (int32_t j = 0; j < M3_RSORT_ITEMS; ++j)
    {
        k[j] = j < per ? keys[t * per + j] : 0ull;
    }
    const uint64_t ones = 0x0001000100010001ull;
    for (int32_t bit = 32; bit < 64; bit += M3_RSORT_BITS)
    {
        uint64_t c[4] = {0ull, 0ull, 0ull, 0ull};
        uint32_t rank[M3_RSORT_ITEMS];
        M3_UNROLL
        for (int32_t j = 0; j < M3_RSORT_ITEMS; ++j)
        {
            if (j < per)
            {
                const uint32_t dg = (uint32_t)(k[j] >> bit) & 15u;
                rank[j] = RsortField(c, dg);
                const uint64_t inc = 1ull << (16 * (dg & 3));
                const uint32_t q = dg >> 2;
                c[0] += q == 0 ? inc : 0ull;
                c[1] += q == 1 ? inc : 0ull;
                c[2] += q == 2 ? inc : 0ull;
                c[3] += q == 3 ? inc : 0ull;
            }
        }
        uint64_t sc[4];
        M3_UNROLL
        for (int r = 0; r < 4; ++r)
        {
            sc[r] = WaveInclusiveSum64(c[r]);
        }
        if (lane == 63)
        {
            M3_UNROLL
            for (int r = 0; r < 4; ++r)
            {
                waveTot[wave][r] = sc[r];
            }
        }
        __syncthreads();
        if (wave == 0)
        {
            uint64_t x[4];
            uint64_t in[4];
            M3_UNROLL
            for (int r = 0; r < 4; ++r)
            {
                x[r] = lane < M3_RSORT_THREADS / 64 ? waveTot[lane][r] : 0ull;
                in[r] = WaveInclusiveSum64(x[r]);
            }
            if (lane < M3_RSORT_THREADS / 64)
            {
                M3_UNROLL
                for (int r = 0; r < 4; ++r)
                {
                    waveTot[lane][r] = in[r] - x[r];
                }
            }
            if (lane == M3_RSORT_THREADS / 64 - 1)
            {
                M3_UNROLL
                for (int r = 0; r < 4; ++r)
                {
                    grand[r] = in[r];
                }
            }
        }
        __syncthreads();
        // Each field of `at`: the digit's base (keys of lower digits) + the
        // digit's keys in earlier threads.
        uint64_t at[4];
        uint64_t carry = 0ull;
        M3_UNROLL
        for (int r = 0; r < 4; ++r)
        {
            const uint64_t g = grand[r];
            const uint64_t before = (g << 16) + (g << 32) + (g << 48); // earlier fields of this word
            at[r] = before + carry * ones + waveTot[wave][r] + sc[r] - c[r];
            carry += (g * ones) >> 48; // the word's total
        }
        M3_UNROLL
        for (int32_t j = 0; j < M3_RSORT_ITEMS; ++j)
        {
            if (j < per)
            {
                const uint32_t dg = (uint32_t)(k[j] >> bit) & 15u;
                keys[RsortField(at, dg) + rank[j]] = k[j];
            }
        }
        __syncthreads();
        M3_UNROLL
        for (int32_t j = 0; j < M3_RSORT_ITEMS; ++j)
        {
            if (j < per)
            {
                k[j] = keys[t * per + j];
            }
        }
    }
}

// Keys of a sorted run below `key`.
__device__ static inline int32_t RunBelow(const uint64_t* run, int32_t len, uint64_t key)
{
    int32_t lo = 0;
    int32_t hi = len;
    while (lo < hi)
    {
        const int32_t mid = (lo + hi) >> 1;
        if (run[mid] < key)
        {
            lo = mid + 1;
        }
        else
        {
            hi = mid;
        }
    }
    return lo;
}

__global__ void __launch_bounds__(M3_RSORT_THREADS) k_sap_sort(const m3Dev* __restrict__ ds, int32_t runs)
{
    M3_WORLD;
    __shared__ uint64_t keys[M3_RSORT_MAX];
    __shared__ uint64_t waveTot[M3_RSORT_THREADS / 64][4];
    __shared__ uint64_t grand[4];
    const m3GridSync sync = {d.barrier, &v.sc->fault};
    const int32_t t = threadIdx.x;
    const int32_t n = v.maxShape;
    const int32_t runLen = (n + runs - 1) / runs;
    uint64_t* buf = d.sapKeysIn;
    if ((int32_t)blockIdx.x < runs)
    {
        const int32_t base = blockIdx.x * runLen;
        const int32_t len = n - base < runLen ? (n - base > 0 ? n - base : 0) : runLen;
        const int32_t per = (len + M3_RSORT_THREADS - 1) / M3_RSORT_THREADS;
        const int32_t m = per * M3_RSORT_THREADS;
        // Padding keys sort after every real key and are never written out.
        M3_UNROLL
        for (int32_t j = 0; j < M3_RSORT_ITEMS; ++j)
        {
            const int32_t i = t + j * M3_RSORT_THREADS;
            if (i < m)
            {
                keys[i] = i < len ? buf[base + i] : ~0ull;
            }
        }
        __syncthreads();
        RunSort(keys, per, waveTot, grand);
        M3_UNROLL
        for (int32_t j = 0; j < M3_RSORT_ITEMS; ++j)
        {
            const int32_t i = t + j * M3_RSORT_THREADS;
            if (i < len)
            {
                buf[base + i] = keys[i];
            }
        }
    }
    sync();
    for (int32_t i = t; i < n; i += M3_RSORT_THREADS)
    {
        keys[i] = buf[i];
    }
    __syncthreads();
    const int32_t i = blockIdx.x * M3_RSORT_THREADS + t;
    if (i < n)
    {
        const int32_t own = i / runLen;
        const uint64_t key = keys[i];
        int32_t at = i - own * runLen;
        for (int32_t r = 0; r < runs; ++r)
        {
            if (r != own)
            {
                const int32_t base = r * runLen;
                const int32_t len = n - base < runLen ? (n - base > 0 ? n - base : 0) : runLen;
                at += RunBelow(keys + base, len, key);
            }
        }
        d.sapKeysOut[at] = key;
    }
}

#define M3_SORT_THREADS 1024
#define M3_SORT_GRID 32 // workgroups of the pair-key launch

// ---------------------------------------------------------------------------
// The pair list's candidate keys sorted and made unique (src/broad_phase.c:636
// sorts the unique keys) in one cooperative launch.  A compact key is
// lower shape << 16 | higher shape, so keys group by their lower shape:
// count per group, scan, scatter; then each group is ordered and its
// duplicates dropped (up to M3_GROUP_SMALL keys in one thread's registers,
// larger groups by a workgroup through an LDS bitmap over the higher
// shape); a second scan places the groups' unique keys.
// ---------------------------------------------------------------------------
#define M3_GROUP_SMALL 16
#define M3_GROUP_BITMAP_WORDS 2048 // higher shapes < 65536

// Exclusive scan of in[0, n) into out[0, n] (out[n] = total) by one
// workgroup of M3_SORT_THREADS; groups above `big` keys are listed.  Each
// thread takes a contiguous run of the entries (its loads issued eight at
// a time), one workgroup scan places the runs, and each thread writes its
// run's prefixes.
__device__ static inline void GroupScan(const int32_t* in, int32_t* out, int32_t n, int32_t big,
                                        int32_t* list, int32_t* listCount)
{
    const int32_t per = (n + (int32_t)blockDim.x - 1) / (int32_t)blockDim.x;
    const int32_t b = (int32_t)threadIdx.x * per;
    int32_t sum = 0;
    for (int32_t j = 0; j < per; j += 8)
    {
        int32_t x[8];
        M3_UNROLL
        for (int32_t q = 0; q < 8; ++q)
        {
            const int32_t i = b + j + q;
            x[q] = (j + q < per && i < n) ? in[i] : 0;
        }
        M3_UNROLL
        for (int32_t q = 0; q < 8; ++q)
        {
            sum += x[q];
        }
    }
    int32_t total = 0;
    int32_t pos = BlockExclusiveScan(sum, &total);
    for (int32_t j = 0; j < per; j += 8)
    {
        int32_t x[8];
        M3_UNROLL
        for (int32_t q = 0; q < 8; ++q)
        {
            const int32_t i = b + j + q;
            x[q] = (j + q < per && i < n) ? in[i] : 0;
        }
        M3_UNROLL
        for (int32_t q = 0; q < 8; ++q)
        {
            const int32_t i = b + j + q;
            if (j + q < per && i < n)
            {
                out[i] = pos;
                pos += x[q];
                if (big >= 0 && x[q] > big)
                {
                    list[atomicAdd(listCount, 1)] = i;
                }
            }
        }
    }
    if (threadIdx.x == 0)
    {
        out[n] = total;
    }
}

__global__ void __launch_bounds__(M3_SORT_THREADS) k_emit_pairs(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    __shared__ uint32_t bits[M3_GROUP_BITMAP_WORDS];
    const m3GridSync sync{d.barrier, &v.sc->fault};
    const int32_t ns = v.maxShape;
    const int32_t stride = ns + 1;
    int32_t* cnt = d.grp;
    int32_t* start = d.grp + stride;
    int32_t* fill = d.grp + 2 * stride;
    int32_t* ucnt = d.grp + 3 * stride;
    int32_t* ustart = d.grp + 4 * stride;
    int32_t* large = d.grp + 5 * stride;
    int32_t* largeCount = d.grp + 6 * stride;
    int32_t n = v.sc->emitCount;
    n = n < 0 ? 0 : (n < v.emitCapacity ? n : v.emitCapacity);
    const int32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t gsz = gridDim.x * blockDim.x;
    // Emission writes each wavefront's keys together, so runs of one lower
    // shape (a plane's pairs) share a wavefront: their lanes add once.
    const int32_t lane = threadIdx.x & 63;
    for (int32_t i0 = gid - lane; i0 < n; i0 += gsz)
    {
        const int32_t i = i0 + lane;
        const uint32_t lo = i < n ? v.emit[i] >> 16 : 0xFFFFFFFFu;
        const bool live = lo < (uint32_t)ns;
        const uint32_t lead = (uint32_t)__shfl((int32_t)lo, 0);
        const uint64_t same = __ballot(live && lo == lead);
        if (live && lo != lead)
        {
            atomicAdd(&cnt[lo], 1);
        }
        if (lane == 0 && same != 0ull)
        {
            atomicAdd(&cnt[lead], (int32_t)__popcll(same));
        }
    }
    sync();
    if (blockIdx.x == 0)
    {
        GroupScan(cnt, start, ns, M3_GROUP_SMALL, large, largeCount);
    }
    sync();
    for (int32_t i0 = gid - lane; i0 < n; i0 += gsz)
    {
        const int32_t i = i0 + lane;
        const uint32_t key = i < n ? v.emit[i] : M3_SENTINEL32;
        const uint32_t lo = key >> 16;
        const bool live = lo < (uint32_t)ns;
        const uint32_t lead = (uint32_t)__shfl((int32_t)lo, 0);
        const uint64_t same = __ballot(live && lo == lead);
        int32_t base = 0;
        if (lane == 0 && same != 0ull)
        {
            base = atomicAdd(&fill[lead], (int32_t)__popcll(same));
        }
        base = __shfl(base, 0);
        if (live)
        {
            const int32_t at = lo == lead ? start[lo] + base + (int32_t)__popcll(same & ((1ull << lane) - 1ull))
                                          : start[lo] + atomicAdd(&fill[lo], 1);
            d.grpKeys[at] = key;
            d.grpOf[at] = (int32_t)lo;
        }
    }
    sync();
    // Large groups: a workgroup each, through a bitmap over the higher shape.
    const int32_t words = (ns + 31) >> 5;
    const int32_t nLarge = *largeCount;
    for (int32_t g = blockIdx.x; g < nLarge; g += gridDim.x)
    {
        const int32_t lo = large[g];
        const int32_t s0 = start[lo];
        const int32_t c = cnt[lo];
        for (int32_t w = threadIdx.x; w < words; w += blockDim.x)
        {
            bits[w] = 0u;
        }
        __syncthreads();
        for (int32_t j = threadIdx.x; j < c; j += blockDim.x)
        {
            const uint32_t hi = d.grpKeys[s0 + j] & 0xFFFFu;
            atomicOr(&bits[hi >> 5], 1u << (hi & 31u));
        }
        __syncthreads();
        int32_t carry = 0;
        for (int32_t w0 = 0; w0 < words; w0 += blockDim.x)
        {
            const int32_t w = w0 + threadIdx.x;
            const uint32_t m = w < words ? bits[w] : 0u;
            int32_t total = 0;
            int32_t pos = carry + BlockExclusiveScan(__popc(m), &total);
            for (uint32_t r = m; r != 0u; r &= r - 1u)
            {
                const uint32_t hi = ((uint32_t)w << 5) | (uint32_t)__builtin_ctz(r);
                d.grpKeys[s0 + pos++] = ((uint32_t)lo << 16) | hi;
            }
            carry += total;
        }
        if (threadIdx.x == 0)
        {
            ucnt[lo] = carry;
        }
        __syncthreads();
    }
    // Small groups: one thread each, in registers.
    for (int32_t lo = gid; lo < ns; lo += gsz)
    {
        const int32_t c = cnt[lo];
        if (c > M3_GROUP_SMALL)
        {
            continue;
        }
        const int32_t s0 = start[lo];
        uint32_t k[M3_GROUP_SMALL];
        M3_UNROLL
        for (int32_t j = 0; j < M3_GROUP_SMALL; ++j)
        {
            k[j] = j < c ? d.grpKeys[s0 + j] : M3_SENTINEL32;
        }
        // Bitonic network over the 16 registers (ascending).
        M3_UNROLL
        for (int32_t size = 2; size <= M3_GROUP_SMALL; size <<= 1)
        {
            M3_UNROLL
            for (int32_t half = size >> 1; half > 0; half >>= 1)
            {
                M3_UNROLL
                for (int32_t a = 0; a < M3_GROUP_SMALL; ++a)
                {
                    const int32_t b = a ^ half;
                    if (b > a)
                    {
                        const bool up = (a & size) == 0;
                        const uint32_t x = k[a];
                        const uint32_t y = k[b];
                        const bool swap = up ? x > y : x < y;
                        k[a] = swap ? y : x;
                        k[b] = swap ? x : y;
                    }
                }
            }
        }
        int32_t u = 0;
        M3_UNROLL
        for (int32_t j = 0; j < M3_GROUP_SMALL; ++j)
        {
            if (j < c && (j == 0 || k[j] != k[j - 1]))
            {
                d.grpKeys[s0 + u] = k[j];
                ++u;
            }
        }
        ucnt[lo] = u;
    }
    sync();
    if (blockIdx.x == 0)
    {
        GroupScan(ucnt, ustart, ns, -1, nullptr, nullptr);
    }
    sync();
    const int32_t placed = start[ns];
    for (int32_t p = gid; p < placed; p += gsz)
    {
        const int32_t lo = d.grpOf[p];
        const int32_t j = p - start[lo];
        if (j < ucnt[lo])
        {
            d.emitUnique[ustart[lo] + j] = d.grpKeys[p];
        }
    }
    if (gid == 0)
    {
        v.sc->uniqueCount = ustart[ns];
    }
}

// The emitted keys sorted and made unique (the pair list candidates).
__global__ void __launch_bounds__(M3_BATCH_THREADS) k_emit_sort_block(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    __shared__ uint32_t keys[M3_BATCH_EMIT_MAX];
    int32_t n = v.sc->emitCount;
    n = n < v.emitCapacity ? n : v.emitCapacity;
    const int32_t n2 = NextPow2(n);
    for (int32_t i = threadIdx.x; i < n2; i += blockDim.x)
    {
        keys[i] = i < n ? v.emit[i] : M3_SENTINEL32;
    }
    __syncthreads();
    BitonicSortLds(keys, n2);
    int32_t written = 0;
    for (int32_t base = 0; base < n; base += blockDim.x)
    {
        int32_t i = base + threadIdx.x;
        int32_t first = (i < n && (i == 0 || keys[i] != keys[i - 1])) ? 1 : 0;
        int32_t total = 0;
        int32_t pos = BlockExclusiveScan(first, &total);
        if (first)
        {
            d.emitUnique[written + pos] = keys[i];
        }
        written += total;
    }
    if (threadIdx.x == 0)
    {
        v.sc->uniqueCount = written;
    }
}

// Stable compactions: 0 = cold pair keys into the frozen buffer,
// 1 = constraint pair indices, 2 = moving bodies.
__global__ void __launch_bounds__(M3_BATCH_THREADS)
    k_select_block(const m3Dev* __restrict__ ds, int mode)
{
    M3_WORLD;
    const int32_t n = mode == 2 ? v.maxBody : v.pairCapacity;
    const uint8_t* flags = mode == 2 ? d.bodyFlags : d.pairFlags;
    int32_t written = 0;
    for (int32_t base = 0; base < n; base += blockDim.x)
    {
        int32_t i = base + threadIdx.x;
        int32_t f = (i < n && flags[i] != 0) ? 1 : 0;
        int32_t total = 0;
        int32_t pos = BlockExclusiveScan(f, &total);
        if (f)
        {
            if (mode == 0)
            {
                v.sleepKeys[written + pos] = v.keys[i];
            }
            else if (mode == 1)
            {
                v.consPair[written + pos] = i;
            }
            else
            {
                v.movers[written + pos] = i;
            }
        }
        written += total;
    }
    if (threadIdx.x == 0)
    {
        int32_t* out = mode == 0 ? &v.sc->sleepCount
                                 : (mode == 1 ? &v.sc->constraintCount : &v.sc->moverCount);
        *out = written;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static int JoinBatch(m3gpuWorld* w);
static void BatchForget(m3gpuWorld* w);

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

static int RefreshDevCopies(m3gpuWorld* w);

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
        DevAlloc(&v.fresh, 6 * (int64_t)ns) || DevAlloc(&v.planes, ns) || DevAlloc(&v.hitEvents, ns) ||
        DevAlloc(&v.preSolve, ns))
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
        DevAlloc(&w->dev.emitUnique, v.emitCapacity) ||
        DevAlloc(&v.satQueries, np) || DevAlloc(&v.satList, np) ||
        DevAlloc(&v.cons, np) || DevAlloc(&v.consPair, np) || DevAlloc(&v.colors, np) ||
        DevAlloc(&v.colorLists, np) || DevAlloc(&v.movers, nb) || DevAlloc(&v.deltaPos, nb) ||
        DevAlloc(&v.deltaRot, nb) || DevAlloc(&v.com0, nb) || DevAlloc(&v.rot0, nb) ||
        DevAlloc(&v.parent, nb) || DevAlloc(&v.forced, nb) || DevAlloc(&v.flag, nb) ||
        DevAlloc(&v.pending, nb) ||
        DevAlloc(&v.sc, 1) || DevAlloc(&w->dev.sapKeysIn, ns) || DevAlloc(&w->dev.sapKeysOut, ns) ||
        DevAlloc(&w->dev.grp, 6 * ((int64_t)ns + 1) + 1) || DevAlloc(&w->dev.grpKeys, 2 * (int64_t)np) ||
        DevAlloc(&w->dev.grpOf, 2 * (int64_t)np) ||
        DevAlloc(&w->dev.sapXY, ns) || DevAlloc(&w->dev.sapZ, ns) || DevAlloc(&w->dev.sapAttr, ns) ||
        DevAlloc(&w->dev.pairFlags, np) || DevAlloc(&w->dev.bodyFlags, nb) ||
        DevAlloc(&w->dev.blockCounts, Blocks(np > nb ? np : nb, M3_COMPACT_THREADS)) || DevAlloc(&w->dev.root, nb) ||
        DevAlloc(&w->dev.bodyMasks, nb) || DevAlloc(&w->dev.bodyOwner, nb) ||
        DevAlloc(&v.slotColors, np) ||
        DevAlloc(&v.consBodies, 2 * (int64_t)np) || DevAlloc(&v.slotOf, np) ||
        DevAlloc(&v.slotTable, (int64_t)M3_GRAPH_COLORS * nb) || DevAlloc(&v.slotMask, nb) ||
        DevAlloc(&w->dev.barrier, 4) || DevAlloc(&w->dev.narrowBins, 8 + 8 * (int64_t)np) ||
        DevAlloc(&w->dev.flowRec, 6 * (int64_t)nb) || DevAlloc(&w->dev.flowRank, np) ||
        DevAlloc(&w->dev.warmAdd, 2 * (int64_t)np) ||
        DevAlloc(&w->dev.deltaRec, M3_DELTA_WORDS * (int64_t)nb) ||
        DevAlloc(&w->dev.cellOf, nb) || DevAlloc(&w->dev.cellLocal, nb) ||
        DevAlloc(&w->dev.cellBodies, (int64_t)M3_CELL_MAX * M3_CELL_BODIES) ||
        DevAlloc(&w->dev.cellCount, M3_CELL_MAX) || DevAlloc(&w->dev.cellConsCount, M3_CELL_MAX) ||
        DevAlloc(&w->dev.slotOwner, np) || DevAlloc(&w->dev.slotBodies, np) ||
        DevAlloc(&w->dev.cellDone, M3_CELL_MAX) ||
        DevAlloc(&w->dev.cellCons, (int64_t)M3_CELL_MAX * M3_CELL_CONS) ||
        DevAlloc(&w->hitRecBuf, np) || DevAlloc(&v.shapeRank, ns) || DevAlloc(&v.ccdStatics, ns) ||
        DevAlloc(&w->dev.ccdXf, nb) || DevAlloc(&w->dev.ccdHit, nb))
    {
        return -1;
    }
    HIP_TRY(hipHostMalloc((void**)&w->rankStage, (size_t)(ns > 0 ? ns : 1) * sizeof(int32_t),
                          hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&w->leafStage, (size_t)(ns > 0 ? ns : 1) * 6 * sizeof(double),
                          hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&w->leavesCopied, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&w->ranksCopied, hipEventDisableTiming));
    int32_t nj = c.jointCapacity > 0 ? c.jointCapacity : 1;
    if (DevAlloc(&v.jointAlive, nj) || DevAlloc(&v.jointType, nj) || DevAlloc(&v.jointBodyA, nj) ||
        DevAlloc(&v.jointBodyB, nj) || DevAlloc(&v.jointLocalA, nj) ||
        DevAlloc(&v.jointLocalB, nj) || DevAlloc(&v.jointCollide, nj) ||
        DevAlloc(&v.jointFrameQA, nj) || DevAlloc(&v.jointFrameQB, nj) ||
        DevAlloc(&v.jointFlags, nj) || DevAlloc(&v.jointLimits, nj) ||
        DevAlloc(&v.jointImpulse, nj) || DevAlloc(&v.jointPerpImpulse, nj) ||
        DevAlloc(&v.jointLimitImpulse, nj) || DevAlloc(&v.jointAngularImpulse, nj) ||
        DevAlloc(&v.jointSpringImpulse, nj) || DevAlloc(&v.jointNextA, nj) ||
        DevAlloc(&v.jointNextB, nj) || DevAlloc(&v.bodyJointHead, nb) ||
        DevAlloc(&v.jointCompStart, nj + 1) || DevAlloc(&v.jointCompList, nj) ||
        DevAlloc(&v.jcons, nj) || DevAlloc(&v.jointActiveFlag, nj))
    {
        return -1;
    }
    // Temp storage of the sweep-key radix sort.
    {
        size_t bytes = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, w->dev.sapKeysIn, w->dev.sapKeysOut, ns,
                                                  0, 64, w->stream));
        w->cubBytes = bytes > 0 ? bytes : 1;
        HIP_TRY(hipMalloc(&w->cubTemp, w->cubBytes));
    }
    HIP_TRY(hipHostMalloc((void**)&w->hostScalars, sizeof(m3Scalars), hipHostMallocDefault));
    memset(w->hostScalars, 0, sizeof(m3Scalars));

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
        {
            int a = 0, b = 0, c = 0;
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_islands, M3_COMPACT_THREADS, 0));
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_constraints, M3_COMPACT_THREADS, 0));
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_movers, M3_COMPACT_THREADS, 0));
            const int m = a < b ? (a < c ? a : c) : (b < c ? b : c);
            const int most = (m > 0 ? m : 1) * prop.multiProcessorCount;
            w->coopCompact = most < M3_COOP_COMPACT_BLOCKS ? most : M3_COOP_COMPACT_BLOCKS;
        }
        // Up to 40 workgroups of 256: one body per thread in the
        // integrate-and-warm-start phase for worlds up to 10,240 bodies,
        // while the grid barrier stays near its few-microsecond floor.  A
        // world with fewer bodies launches fewer (EnqueueStep).
        const char* env = getenv("MAUL3D_AMD_SUBSTEP_BLOCKS");
        int want = env != nullptr ? atoi(env) : 40;
        w->coopBlocks = want < 1 ? 1 : (want > maxBlocks ? maxBlocks : want);
        w->coopFixed = env != nullptr;
        w->coopEnvFixed = w->coopFixed;
        w->coopDefault = w->coopBlocks;
        w->coopMax = maxBlocks;
        const char* flowEnv = getenv("MAUL3D_AMD_FLOW");
        w->flow = flowEnv != nullptr && atoi(flowEnv) == 0 ? 0 : 1;
        // About 63 movers per cell (160 cells for the 10k rain, the fastest
        // of 48-256 measured there; MAUL3D_AMD_CELLS overrides), as a grid
        // of x slabs by z cells, within the co-resident limit.
        const char* cellEnv = getenv("MAUL3D_AMD_CELLS");
        int cells = cellEnv != nullptr ? atoi(cellEnv) : (nb + 62) / 63;
        const int cellMax = maxBlocks < M3_CELL_MAX ? maxBlocks : M3_CELL_MAX;
        cells = cells < 1 ? 1 : (cells > cellMax ? cellMax : cells);
        int gx = 1;
        while ((gx * 2) * (gx * 2) <= cells && gx * 2 <= M3_CELL_GX_MAX)
        {
            gx *= 2;
        }
        w->cellsX = gx;
        w->cellsZ = cells / gx;
    }
    HIP_TRY(hipMalloc(&w->devDev, 2 * sizeof(m3Dev)));
    w->devDirty = 1;
    w->cur = 0;
    v.keys = w->keysBuf[0];
    v.manifolds = w->mansBuf[0];
    v.oldKeys = w->keysBuf[1];
    v.oldManifolds = w->mansBuf[1];
    // The clears of DevAlloc ran on the null stream, which does not order
    // against the world's non-blocking stream.
    HIP_TRY(hipStreamSynchronize(nullptr));
    HIP_TRY(hipStreamSynchronize(w->stream));
    return 0;
}

extern "C" m3gpuWorld* m3gpu_world_create(const m3gpuCaps* caps)
{
    if (caps == nullptr || caps->bodyCapacity <= 0 || caps->shapeCapacity <= 0 ||
        caps->pairCapacity <= 0 || caps->shapeCapacity > M3_DEVICE_MAX_SHAPES)
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
static void EventBufsFree(m3gpuWorld* w);

extern "C" void m3gpu_world_destroy(m3gpuWorld* w)
{
    if (w == nullptr)
    {
        return;
    }
    BatchForget(w);
    (void)hipStreamSynchronize(w->stream);
    m3View& v = w->view;
    void* ptrs[] = {v.xf, v.lv, v.av, v.invMass, v.invI, v.inertia, v.lc, v.gravityScale,
                    v.linDamp, v.angDamp, v.bodyAlive, v.type, v.awake, v.sleepTime, v.bullet,
                    v.minExt, v.maxExt, v.force, v.torque, v.enabled, v.locks, v.sleepThreshold,
                    v.canSleep, v.hasTarget, v.target, v.shapeHead, v.shapeAlive, v.shapeBody,
                    v.shapeType, v.geom, v.friction, v.restitution, v.rolling, v.sensor, v.hitEvents, w->hitRecBuf,
                    v.preSolve, w->preCands, w->preCount, w->vetoKeys, w->vetoFlags,
                    v.hasOffset, v.localPos, v.localRot, v.surfaceVel, v.category, v.mask,
                    v.group, v.shapeNext, v.hullIndex, v.leaf, v.fresh, v.planes,
                    (void*)v.hulls, w->keysBuf[0], w->keysBuf[1], w->mansBuf[0], w->mansBuf[1],
                    v.sleepKeys, v.emit, v.satQueries, v.satList, w->dev.emitUnique, v.cons, v.consPair,
                    v.colors, v.colorLists, v.movers, v.deltaPos, v.deltaRot, v.com0, v.rot0,
                    v.parent, v.forced, v.flag, v.pending, v.sc, w->dev.sapKeysIn, w->dev.sapKeysOut,
                    w->cubTemp, w->dev.grp, w->dev.grpKeys, w->dev.grpOf, w->dev.sapXY, w->dev.sapZ, w->dev.sapAttr, w->dev.pairFlags, w->dev.bodyFlags, w->dev.blockCounts, w->dev.root, w->dev.narrowBins,
                    w->dev.bodyMasks, w->dev.bodyOwner, v.slotColors, v.consBodies, v.slotOf,
                    v.slotTable, v.slotMask, w->dev.barrier, w->dev.flowRec, w->dev.flowRank,
                    w->dev.warmAdd, w->dev.deltaRec, w->dev.cellOf, w->dev.cellLocal, w->dev.cellBodies, w->dev.cellCount,
                    w->dev.cellConsCount, w->dev.slotOwner, w->dev.slotBodies, w->dev.cellDone, w->dev.cellCons, w->devDev,
                    v.jointAlive, v.jointType, v.jointBodyA, v.jointBodyB, v.jointLocalA,
                    v.jointLocalB, v.jointCollide, v.jointFrameQA, v.jointFrameQB, v.jointFlags,
                    v.jointLimits, v.jointImpulse, v.jointPerpImpulse, v.jointLimitImpulse,
                    v.jointAngularImpulse, v.jointSpringImpulse, v.jointNextA, v.jointNextB,
                    v.bodyJointHead, v.jointCompStart, v.jointCompList, v.jcons, v.jointActiveFlag,
                    v.shapeRank, v.ccdStatics, w->dev.ccdXf, w->dev.ccdHit, w->tableKeys, w->tableMans};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i)
    {
        if (ptrs[i] != nullptr)
        {
            (void)hipFree(ptrs[i]);
        }
    }
    (void)hipHostFree(w->hostScalars);
    if (w->rankStage != nullptr)
    {
        (void)hipHostFree(w->rankStage);
    }
    if (w->leafStage != nullptr)
    {
        (void)hipHostFree(w->leafStage);
    }
    if (w->leavesCopied != nullptr)
    {
        (void)hipEventDestroy(w->leavesCopied);
    }
    if (w->ranksCopied != nullptr)
    {
        (void)hipEventDestroy(w->ranksCopied);
    }
    if (w->preStage != nullptr)
    {
        (void)hipHostFree(w->preStage);
        (void)hipHostFree(w->vetoStage);
        (void)hipEventDestroy(w->preGathered);
        (void)hipEventDestroy(w->vetoCopied);
    }
    EventBufsFree(w);
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
    if (w == nullptr || h == nullptr || JoinBatch(w) != 0)
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
        UP(v.hitEvents, h->shapeHitEvents, ns, uint8_t);
        UP(v.preSolve, h->shapePreSolve, ns, uint8_t);
        {
            // Hit records are kept (cleared by the narrowphase, written by the
            // impulse store) only in worlds where some shape opted in.
            int anyHit = 0;
            for (int32_t s = 0; s < ns && !anyHit; ++s)
            {
                anyHit = h->shapeAlive[s] != 0 && h->shapeHitEvents[s] != 0;
            }
            v.hitRecs = anyHit ? w->hitRecBuf : nullptr;
        }
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
        w->wideHulls = 0;
        w->boxHulls = 0;
        w->smallHulls = 1;
        for (int32_t k = 0; k < h->maxHull; ++k)
        {
            const m3HullData* hull = &((const m3HullData*)h->hullData)[k];
            int32_t e = hull->edgeCount;
            w->wideHulls |= e > 24;
            w->boxHulls |= e >= 24;
            if (hull->vertexCount > M3_SAT_BOX_VERTS || e > M3_SAT_BOX_HALF_EDGES)
            {
                w->smallHulls = 0;
            }
        }
        if (h->maxJoint > w->caps.jointCapacity)
        {
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_upload: joints exceed capacity");
            return -1;
        }
        int32_t mj = h->maxJoint;
        v.maxJoint = mj;
        v.jointCompCount = h->jointCompCount;
        UP(v.bodyJointHead, h->bodyJointHead, nb, int32_t);
        if (mj > 0)
        {
            UP(v.jointAlive, h->jointAlive, mj, uint8_t);
            UP(v.jointType, h->jointType, mj, uint8_t);
            UP(v.jointBodyA, h->jointBodyA, mj, int32_t);
            UP(v.jointBodyB, h->jointBodyB, mj, int32_t);
            UP(v.jointLocalA, h->jointLocalA, mj, m3Vec3);
            UP(v.jointLocalB, h->jointLocalB, mj, m3Vec3);
            UP(v.jointCollide, h->jointCollide, mj, uint8_t);
            UP(v.jointFrameQA, h->jointFrameQA, mj, m3Quat);
            UP(v.jointFrameQB, h->jointFrameQB, mj, m3Quat);
            UP(v.jointFlags, h->jointFlags, mj, uint8_t);
            UP(v.jointLimits, h->jointLimits, mj, m3Vec3);
            UP(v.jointNextA, h->jointNextA, mj, int32_t);
            UP(v.jointNextB, h->jointNextB, mj, int32_t);
            UP(v.jointCompStart, h->jointCompStart, h->jointCompCount + 1, int32_t);
            UP(v.jointCompList, h->jointCompList, mj, int32_t);
        }
        // The plane list, the static tree shapes (continuous targets) and
        // the census of what the continuous pass meets are host facts.
        const uint8_t* alive = h->shapeAlive;
        const uint8_t* stype = h->shapeType;
        int32_t* planes = (int32_t*)malloc((size_t)(ns > 0 ? ns : 1) * sizeof(int32_t));
        int32_t* statics = (int32_t*)malloc((size_t)(ns > 0 ? ns : 1) * sizeof(int32_t));
        if (planes == nullptr || statics == nullptr)
        {
            free(planes);
            free(statics);
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_upload: out of host memory");
            return -1;
        }
        int32_t pc = 0;
        int32_t sc = 0;
        int unsupported = 0;
        int needsTree = 0;
        int bullets = 0;
        int restitution = 0;
        for (int32_t s = 0; s < ns; ++s)
        {
            if (alive[s] == 0)
            {
                continue;
            }
            restitution |= h->shapeRestitution[s] > 0.0f;
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
                statics[sc++] = s;
                // Sensors never stop a sweep (src/solver.c:3102-3105).
                needsTree |= h->shapeSensor[s] == 0;
            }
        }
        for (int32_t b = 0; b < nb; ++b)
        {
            if (h->bodyAlive[b] != 0 && h->bulletFlags[b] != 0)
            {
                needsTree = 1;
                bullets = 1;
            }
        }
        w->ccdNeedsTree = needsTree;
        w->anyRestitution = restitution;
        w->dev.ccdBullets = bullets;
        v.ccdStaticCount = sc;
        if (Copy(w, v.ccdStatics, statics, (int64_t)sc * 4, hipMemcpyHostToDevice) != 0)
        {
            free(planes);
            free(statics);
            return -1;
        }
        HIP_TRY(hipStreamSynchronize(w->stream));
        free(statics);
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
        if (h->maxJoint > 0)
        {
            int32_t mj = h->maxJoint;
            UP(v.jointImpulse, h->jointImpulse, mj, m3Vec3);
            UP(v.jointPerpImpulse, h->jointPerpImpulse, mj, m3Vec3);
            UP(v.jointLimitImpulse, h->jointLimitImpulse, mj, m3Vec3);
            UP(v.jointAngularImpulse, h->jointAngularImpulse, mj, m3Vec3);
            UP(v.jointSpringImpulse, h->jointSpringImpulse, mj, m3Vec3);
        }
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
        int32_t fullQuery = w->fullQuery;
        UP(&v.sc->fullQuery, &fullQuery, 1, int32_t);
        HIP_TRY(hipStreamSynchronize(w->stream));
    }
    if (fields & M3GPU_FIELDS_PAIR_TABLE)
    {
        if (w->tableKeys == nullptr)
        {
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_upload: no pair table");
            return -1;
        }
        UP(w->tableKeys, h->pairKeys, v.pairCapacity, uint64_t);
        UP(w->tableMans, h->manifolds, v.pairCapacity, m3Manifold);
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    return RefreshDevCopies(w);
}

// A world whose substep barrier timed out holds no valid state.
static int CheckFault(m3gpuWorld* w)
{
    if (JoinBatch(w) != 0)
    {
        return -1;
    }
    HIP_TRY(hipMemcpyAsync(&w->hostScalars->fault, &w->view.sc->fault, sizeof(int32_t),
                           hipMemcpyDeviceToHost, w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    if (w->hostScalars->fault != 0)
    {
        snprintf(s_lastError, sizeof(s_lastError), "device step gave up: a substep grid barrier timed out or the coloring stalled");
        return -1;
    }
    return 0;
}

extern "C" int m3gpu_download(m3gpuWorld* w, m3gpuHostState* h, uint32_t fields)
{
    if (w == nullptr || h == nullptr || JoinBatch(w) != 0)
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
        if (v.maxJoint > 0)
        {
            int32_t mj = v.maxJoint;
            DOWN(h->jointImpulse, v.jointImpulse, mj, m3Vec3);
            DOWN(h->jointPerpImpulse, v.jointPerpImpulse, mj, m3Vec3);
            DOWN(h->jointLimitImpulse, v.jointLimitImpulse, mj, m3Vec3);
            DOWN(h->jointAngularImpulse, v.jointAngularImpulse, mj, m3Vec3);
            DOWN(h->jointSpringImpulse, v.jointSpringImpulse, mj, m3Vec3);
        }
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
    if (fields & M3GPU_FIELDS_PAIR_TABLE)
    {
        if (w->tableKeys == nullptr)
        {
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_download: no pair table");
            return -1;
        }
        DOWN(h->pairKeys, w->tableKeys, v.pairCapacity, uint64_t);
        DOWN(h->manifolds, w->tableMans, v.pairCapacity, m3Manifold);
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    return CheckFault(w);
}

// Stage names for the profile (event-timed on the world's stream).
enum m3Stage
{
    ST_BEGIN, ST_FILL, ST_FRESH, ST_DISCOVER_ALL, ST_REFRESH, ST_SAP_SORT, ST_SAP_GATHER, ST_SAP, ST_PLANES,
    ST_FROZEN, ST_EMIT_SORT, ST_UNIQUE, ST_PAIR_COUNT, ST_EXPAND, ST_STALL_RESTORE,
    ST_SLEEP_SELECT, ST_SAT_LIST, ST_SAT, ST_NARROW, ST_CAPTURE, ST_UNION, ST_ROOT, ST_WAKE_MARK, ST_WAKE_APPLY,
    ST_CONS_FLAGS, ST_CONS_SELECT, ST_CONS_BODIES, ST_COLOR, ST_PREPARE,
    ST_MOVER_FLAGS, ST_MOVER_SELECT, ST_CELLS, ST_SUBSTEPS, ST_CLEAR_FORCES, ST_RESTITUTION, ST_STORE, ST_CCD,
    ST_SLEEP_TIMERS, ST_SLEEP_READY, ST_SLEEP_FREEZE, ST_END,
    ST_COUNT
};

static const char* const kStageNames[] = {
    "begin", "fill_emit", "fresh_bounds", "discover_all", "refresh_proxies", "sap_sort", "sap_gather", "sap_sweep",
    "plane_pairs", "frozen_merge", "pair_sort", "pair_unique", "pair_count", "pair_expand",
    "stall_restore", "frozen_select", "hull_sat_list", "hull_sat", "narrowphase", "capture", "island_union", "island_root",
    "wake_mark", "wake_apply", "constraint_flags", "constraint_select", "constraint_bodies",
    "coloring", "prepare", "mover_flags", "mover_select", "cells", "substeps", "clear_forces",
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

// Both device copies of a world's kernel view, one per pair-buffer parity
// (a step with cur = p writes keysBuf[p] and carries from keysBuf[p ^ 1]).
// Refreshed after creation and every upload, once the stream is idle.
static int RefreshDevCopies(m3gpuWorld* w)
{
    m3Dev copies[2];
    for (int parity = 0; parity < 2; ++parity)
    {
        copies[parity] = w->dev;
        copies[parity].v = w->view;
        copies[parity].v.keys = w->keysBuf[parity];
        copies[parity].v.manifolds = w->mansBuf[parity];
        copies[parity].v.oldKeys = w->keysBuf[parity ^ 1];
        copies[parity].v.oldManifolds = w->mansBuf[parity ^ 1];
        copies[parity].v.satBoxes = 1;
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    HIP_TRY(hipMemcpy(w->devDev, copies, sizeof(copies), hipMemcpyHostToDevice));
    w->devDirty = 0;
    w->layout += 1;
    w->streamBusy = 0;
    return 0;
}

// Workgroups of the wide-hull SAT pass: one per listed pair up to about
// eight per CU over all worlds of the launch (the kernel strides over its
// world's list).
static int SatBlocks(int64_t pairs, int32_t worlds)
{
    int64_t cap = 2048 / (worlds > 0 ? worlds : 1);
    int64_t b = pairs < cap ? pairs : cap;
    return b < 1 ? 1 : (int)b;
}

static m3SubstepParams SubstepParams(const m3gpuStepParams* p)
{
    float h = p->dt / (float)p->substeps;
    m3SubstepParams sp;
    sp.substeps = p->substeps;
    sp.tailLimit = M3_TAIL_CONSTRAINTS;
    sp.flow = 0;
    sp.h = h;
    sp.invH = h > 0.0f ? 1.0f / h : 0.0f;
    sp.push = p->contactPushMaxSpeed;
    sp.maxLinear = p->maximumLinearSpeed;
    sp.maxAngular = p->maximumAngularSpeed;
    sp.gravity = m3V3(p->gravity[0], p->gravity[1], p->gravity[2]);
    return sp;
}

static int EventBufsCreate(m3gpuWorld* w)
{
    if (w->events != nullptr)
    {
        return 0;
    }
    m3EventBufs* e = (m3EventBufs*)calloc(1, sizeof(m3EventBufs));
    if (e == nullptr)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_events: out of host memory");
        return -1;
    }
    const int32_t np = w->caps.pairCapacity;
    const int32_t nb = w->caps.bodyCapacity;
    if (DevAlloc(&e->codeNew, np) || DevAlloc(&e->codeOld, np) || DevAlloc(&e->counts, 8) ||
        DevAlloc(&e->lists, 4 * (int64_t)np) || DevAlloc(&e->hits, np) || DevAlloc(&e->moves, nb) ||
        hipHostMalloc((void**)&e->hostCounts, 8 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_events: device allocation failed");
        w->events = e;
        EventBufsFree(w); // nothing half-built stays behind
        return -1;
    }
    // DevAlloc clears with hipMemset on the null stream, which does not order
    // against the world's non-blocking stream: finish the clears before the
    // event kernels write these buffers.
    HIP_TRY(hipStreamSynchronize(nullptr));
    w->events = e;
    return 0;
}

static void EventBufsFree(m3gpuWorld* w)
{
    m3EventBufs* e = w->events;
    if (e == nullptr)
    {
        return;
    }
    void* ptrs[] = {e->codeNew, e->codeOld, e->counts, e->lists, e->hits, e->moves};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i)
    {
        if (ptrs[i] != nullptr)
        {
            (void)hipFree(ptrs[i]);
        }
    }
    if (e->hostCounts != nullptr)
    {
        (void)hipHostFree(e->hostCounts);
    }
    free(e);
    w->events = nullptr;
}

extern "C" int m3gpu_events(m3gpuWorld* w, float hitThreshold, m3gpuEvents* out)
{
    if (w == nullptr || out == nullptr || JoinBatch(w) != 0)
    {
        return -1;
    }
    if (EventBufsCreate(w) != 0)
    {
        return -1;
    }
    if (w->devDirty && RefreshDevCopies(w) != 0)
    {
        return -1;
    }
    m3EventBufs& e = *w->events;
    const m3Dev* D = w->devDev + w->cur;
    const int32_t np = w->caps.pairCapacity;
    const int32_t nb = w->caps.bodyCapacity;
    const int32_t n = np > nb ? np : nb;
    hipStream_t st = w->stream;
    KL(k_event_codes, dim3(Blocks(n, 256)), dim3(256), 0, st, D, e, hitThreshold);
    HIP_TRY(hipGetLastError());
    KL(k_event_compact, dim3(1), dim3(M3_EVENT_THREADS), 0, st, D, e);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(e.hostCounts, e.counts, 7 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int k = 0; k < 7; ++k)
    {
        out->counts[k] = e.hostCounts[k];
    }
    uint64_t* dst[4] = {out->beginKeys, out->endKeys, out->sensorBeginKeys, out->sensorEndKeys};
    for (int k = 0; k < 4; ++k)
    {
        DOWN(dst[k], e.lists + (int64_t)k * np, out->counts[k], uint64_t);
    }
    DOWN(out->hits, e.hits, out->counts[4], m3gpuHit);
    DOWN(out->moves, e.moves, out->counts[5], m3gpuMove);
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

// ---------------------------------------------------------------------------
// The event read without allocation (m3gpu_events_noalloc): the same kernels
// over a scratch carved from one file-scope device array per GPU instead of
// the per-world buffers m3gpu_events allocates on its first call.  Every
// byte the kernels read is written earlier in the same call (codes for every
// pair and body slot, counts by the compaction), so nothing needs clearing;
// calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------

#define M3_NOALLOC_PAIRS 32768
#define M3_NOALLOC_BODIES 8192
#define M3_NOALLOC_DEVICES 64

static constexpr size_t Align256(size_t n) { return (n + 255) & ~(size_t)255; }
static constexpr size_t kNoAllocCodes = Align256(M3_NOALLOC_PAIRS);
static constexpr size_t kNoAllocCounts = Align256(8 * sizeof(int32_t));
static constexpr size_t kNoAllocLists = Align256(4 * (size_t)M3_NOALLOC_PAIRS * sizeof(uint64_t));
static constexpr size_t kNoAllocHits = Align256((size_t)M3_NOALLOC_PAIRS * sizeof(m3gpuHit));
static constexpr size_t kNoAllocMoves = Align256((size_t)M3_NOALLOC_BODIES * sizeof(m3gpuMove));
static constexpr size_t kNoAllocBytes =
    2 * kNoAllocCodes + kNoAllocCounts + kNoAllocLists + kNoAllocHits + kNoAllocMoves;

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

// The array's device address on each GPU (hipGetSymbolAddress, once per
// device with that device current); the host-side name is no address.
static uint8_t* s_scratchPtr[M3_NOALLOC_DEVICES];
static pthread_mutex_t s_scratchLock = PTHREAD_MUTEX_INITIALIZER;

static uint8_t* ScratchOnCurrentDevice(void)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= M3_NOALLOC_DEVICES)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_events_noalloc: no usable current device");
        return nullptr;
    }
    pthread_mutex_lock(&s_scratchLock);
    if (s_scratchPtr[dev] == nullptr)
    {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr)
        {
            pthread_mutex_unlock(&s_scratchLock);
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_events_noalloc: hipGetSymbolAddress: %s",
                     hipGetErrorString(e));
            return nullptr;
        }
        s_scratchPtr[dev] = (uint8_t*)ptr;
    }
    uint8_t* p = s_scratchPtr[dev];
    pthread_mutex_unlock(&s_scratchLock);
    return p;
}

extern "C" int m3gpu_events_noalloc(m3gpuWorld* w, float hitThreshold, m3gpuEvents* out)
{
    if (w == nullptr || out == nullptr || JoinBatch(w) != 0)
    {
        return -1;
    }
    const int32_t np = w->caps.pairCapacity;
    const int32_t nb = w->caps.bodyCapacity;
    if (np > M3_NOALLOC_PAIRS || nb > M3_NOALLOC_BODIES)
    {
        snprintf(s_lastError, sizeof(s_lastError),
                 "m3gpu_events_noalloc: the world's capacities (%d pairs, %d bodies) exceed the "
                 "static scratch (%d pairs, %d bodies); use m3gpu_events",
                 np, nb, M3_NOALLOC_PAIRS, M3_NOALLOC_BODIES);
        return -1;
    }
    uint8_t* base = ScratchOnCurrentDevice();
    if (base == nullptr)
    {
        return -1;
    }
    if (w->devDirty && RefreshDevCopies(w) != 0)
    {
        return -1;
    }
    m3EventBufs e;
    memset(&e, 0, sizeof(e));
    uint8_t* at = base;
    e.codeNew = at;
    at += kNoAllocCodes;
    e.codeOld = at;
    at += kNoAllocCodes;
    e.counts = (int32_t*)at;
    at += kNoAllocCounts;
    e.lists = (uint64_t*)at;
    at += kNoAllocLists;
    e.hits = (m3gpuHit*)at;
    at += kNoAllocHits;
    e.moves = (m3gpuMove*)at;
    const m3Dev* D = w->devDev + w->cur;
    const int32_t n = np > nb ? np : nb;
    hipStream_t st = w->stream;
    KL(k_event_codes, dim3(Blocks(n, 256)), dim3(256), 0, st, D, e, hitThreshold);
    HIP_TRY(hipGetLastError());
    KL(k_event_compact, dim3(1), dim3(M3_EVENT_THREADS), 0, st, D, e);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out->counts, e.counts, 7 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    uint64_t* dst[4] = {out->beginKeys, out->endKeys, out->sensorBeginKeys, out->sensorEndKeys};
    for (int k = 0; k < 4; ++k)
    {
        DOWN(dst[k], e.lists + (int64_t)k * np, out->counts[k], uint64_t);
    }
    DOWN(out->hits, e.hits, out->counts[4], m3gpuHit);
    DOWN(out->moves, e.moves, out->counts[5], m3gpuMove);
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

// Launches a kernel whose workgroups meet at grid barriers (GridBarrier),
// on a grid within the kernel's occupancy limit.  The barrier needs every
// workgroup resident at once, which such a grid gets on an in-order stream
// (the previous launch has drained the device); its waits are bounded, so
// a grid that was somehow not all resident faults the world rather than
// hanging.  hipLaunchCooperativeKernel guarantees residency but idles the
// device around each launch (about 11 us before and after on MI355X), so it
// is taken only when MAUL3D_AMD_COOP_LAUNCH=1.
static int GridLaunch(const void* fn, dim3 grid, dim3 block, void** args, hipStream_t st)
{
    static const int coop = [] {
        const char* env = getenv("MAUL3D_AMD_COOP_LAUNCH");
        return env != nullptr && atoi(env) != 0 ? 1 : 0;
    }();
    if (coop)
    {
        HIP_TRY(hipLaunchCooperativeKernel(fn, grid, block, args, 0, st));
    }
    else
    {
        HIP_TRY(hipLaunchKernel(fn, grid, block, args, 0, st));
    }
    return 0;
}

// A grid-barrier launch of 256-thread workgroups over `items`, at most
// `limit` workgroups (the kernels stride over chunks of the items).
static int CoopLaunch(const void* fn, int64_t items, int limit, void** args, hipStream_t st)
{
    int blocks = (int)Blocks(items, M3_COMPACT_THREADS);
    blocks = blocks > limit ? limit : blocks;
    return GridLaunch(fn, dim3(blocks), dim3(M3_COMPACT_THREADS), args, st);
}

// Enqueue one step.  Launch geometry only depends on host-known capacities
// and high-water marks; every data-dependent count is read on the device.
// The step through the island wake pass (pair list, narrowphase, islands);
// EnqueueStepB and EnqueueStepEnd enqueue the rest.  With `leaves` set, a
// copy of this step's leaf boxes into the pinned stage follows the
// broadphase (m3gpu_step_leaves waits for it).
static int EnqueueStepA(m3gpuWorld* w, const m3gpuStepParams* p, bool leaves)
{
    hipStream_t st = w->stream;
    if (w->devDirty && RefreshDevCopies(w) != 0)
    {
        return -1;
    }
    // Swap the pair buffers: last step's list becomes the carry source.
    w->cur ^= 1;
    m3View& v = w->view;
    v.keys = w->keysBuf[w->cur];
    v.manifolds = w->mansBuf[w->cur];
    v.oldKeys = w->keysBuf[w->cur ^ 1];
    v.oldManifolds = w->mansBuf[w->cur ^ 1];
    const m3Dev* D = w->devDev + w->cur;
    const m3Dev& hd = w->dev;
    const int T = 256;
    int32_t nb = v.maxBody;
    int32_t ns = v.maxShape;
    int32_t np = v.pairCapacity;
    int32_t ne = v.emitCapacity;

    // Step scalars, the emission sentinel and the broadphase's fresh boxes.
    STAGE(ST_FILL, KL(k_open, dim3(Blocks(ne > ns ? ne : ns, T)), dim3(T), 0, st, D));
    if (w->fullQuery)
    {
        // After a restore the frozen buffer is rebuilt from every sleeping
        // body (src/broad_phase.c:432-448).
        KL(k_reset_sleep, dim3(1), dim3(1), 0, st, D);
        STAGE(ST_DISCOVER_ALL, KL(k_pending_all, dim3(Blocks(nb, T)), dim3(T), 0, st, D));
        w->fullQuery = 0;
    }
    if (ns <= M3_RSORT_MAX)
    {
        // Runs of at most M3_RSORT_RUN keys; one thread per key to merge.
        int32_t runs = (int32_t)Blocks(ns, M3_RSORT_RUN);
        void* args[] = {(void*)&D, (void*)&runs};
        STAGE(ST_SAP_SORT, if (GridLaunch((const void*)k_sap_sort, dim3(Blocks(ns, M3_RSORT_THREADS)),
                                          dim3(M3_RSORT_THREADS), args, st) != 0) { return -1; });
    }
    else
    {
        STAGE(ST_SAP_SORT, HIP_TRY(hipcub::DeviceRadixSort::SortKeys(w->cubTemp, w->cubBytes, hd.sapKeysIn,
                                                                     hd.sapKeysOut, ns, 0, 64, st)));
    }
    STAGE(ST_SAP_GATHER, KL(k_sap_gather, dim3(Blocks(ns, T)), dim3(T), 0, st, D));
    STAGE(ST_SAP, KL(k_sap, dim3(Blocks(ns, M3_SAP_SHAPES)), dim3(M3_SAP_THREADS), 0, st, D));
    {
        const int64_t work = (int64_t)ns * (v.planeCount > 0 ? v.planeCount : 1);
        STAGE(ST_PLANES, KL(k_planes, dim3(Blocks(work, T)), dim3(T), 0, st, D));
    }
    if (leaves)
    {
        HIP_TRY(hipMemcpyAsync(w->leafStage, v.leaf, (size_t)ns * 6 * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipEventRecord(w->leavesCopied, st));
    }
    STAGE(ST_FROZEN, KL(k_frozen, dim3(Blocks(np, T)), dim3(T), 0, st, D));
    {
        void* args[] = {(void*)&D};
        STAGE(ST_EMIT_SORT, if (GridLaunch((const void*)k_emit_pairs, dim3(M3_SORT_GRID), dim3(M3_SORT_THREADS),
                                            args, st) != 0) { return -1; });
    }
    const int CT = M3_COMPACT_THREADS;
    // The pair list, its SAT listing and narrowphase bins, and the bodies'
    // step-start state.
    STAGE(ST_EXPAND, KL(k_pairs_settle, dim3(Blocks(np > nb ? np : nb, CT)), dim3(CT), 0, st, D,
                        (w->wideHulls || w->boxHulls) ? 1 : 0));
    STAGE(ST_SLEEP_SELECT, KL(k_compact_frozen, dim3(Blocks(np, CT)), dim3(CT), 0, st, D));
    // Narrowphase.
    if (w->wideHulls || w->boxHulls)
    {
        if (w->smallHulls)
        {
            // The small layout leaves the registers, not LDS, to bound the
            // resident workgroups: one per listed pair, up to 8192.
            STAGE(ST_SAT, KL(k_sat<SatLdsBox>, dim3(SatBlocks(np, 1) * 4), dim3(64), 0, st, D));
        }
        else
        {
            STAGE(ST_SAT, KL(k_sat<SatLds>, dim3(SatBlocks(np, 1)), dim3(64), 0, st, D));
        }
    }
    // The narrowphase, which also unites the bodies of each touching pair.
    STAGE(ST_NARROW, KL(k_narrow_binned, dim3(Blocks(np, 64) + M3_NARROW_CLASSES), dim3(64), 0, st, D));
    // Islands and wake.
    if (v.maxJoint > 0)
    {
        STAGE(ST_UNION, KL(k_union_joints, dim3(Blocks(v.maxJoint, T)), dim3(T), 0, st, D));
    }
    {
        void* args[] = {(void*)&D};
        STAGE(ST_ROOT, if (CoopLaunch((const void*)k_islands, nb, w->coopCompact, args, st) != 0) { return -1; });
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Constraints, coloring, prepare, the substeps, restitution and the impulse
// store: the step from after the island wake pass up to the continuous
// pass.  With `vetoKeys` (ascending) those pairs make no constraint.
static int EnqueueStepB(m3gpuWorld* w, const m3gpuStepParams* p, const uint64_t* vetoKeys, int32_t vetoCount)
{
    hipStream_t st = w->stream;
    const m3Dev* D = w->devDev + w->cur;
    const m3Dev& hd = w->dev;
    const m3View& v = w->view;
    const int T = 256;
    const int32_t nb = v.maxBody;
    const int32_t np = v.pairCapacity;
    const bool vetoes = vetoCount > 0;
    if (vetoes != (w->view.veto != nullptr))
    {
        // The flags are read through the device copies of the world's entry.
        w->view.veto = vetoes ? w->vetoFlags : nullptr;
        if (RefreshDevCopies(w) != 0)
        {
            return -1;
        }
    }
    if (vetoes)
    {
        if (w->vetoCopyPending)
        {
            HIP_TRY(hipEventSynchronize(w->vetoCopied));
            w->vetoCopyPending = 0;
        }
        memcpy(w->vetoStage, vetoKeys, (size_t)vetoCount * sizeof(uint64_t));
        HIP_TRY(hipMemcpyAsync(w->vetoKeys, w->vetoStage, (size_t)vetoCount * sizeof(uint64_t),
                               hipMemcpyHostToDevice, st));
        HIP_TRY(hipEventRecord(w->vetoCopied, st));
        w->vetoCopyPending = 1;
        KL(k_veto_mark, dim3(Blocks(np, T)), dim3(T), 0, st, D, w->vetoKeys, vetoCount, w->vetoFlags);
    }
    // Prepare and color.
    const m3SubstepParams sp = SubstepParams(p);
    const int CT = M3_COMPACT_THREADS;
    {
        void* args[] = {(void*)&D};
        STAGE(ST_CONS_FLAGS, if (CoopLaunch((const void*)k_constraints, np, w->coopCompact, args, st) != 0) { return -1; });
    }
    if (nb <= M3_COLOR_LDS_BODIES)
    {
        STAGE(ST_COLOR, KL(k_color<true>, dim3(1), dim3(M3_COLOR_THREADS), 0, st, D, 1));
    }
    else
    {
        STAGE(ST_COLOR, KL(k_color<false>, dim3(1), dim3(M3_COLOR_THREADS), 0, st, D, 1));
    }
    STAGE(ST_PREPARE, KL(k_prepare, dim3(Blocks(np, 64)), dim3(64), 0, st, D, p->contactHertz,
                         p->contactDampingRatio, sp.h, w->flow));
    if (v.maxJoint > 0)
    {
        STAGE(ST_PREPARE,
              KL(k_prepare_joints, dim3(Blocks(v.maxJoint, 64)), dim3(64), 0, st, D, sp.h));
    }
    {
        float dt = p->dt;
        void* args[] = {(void*)&D, (void*)&dt};
        STAGE(ST_MOVER_FLAGS, if (CoopLaunch((const void*)k_movers, nb, w->coopCompact, args, st) != 0) { return -1; });
    }
    // Substeps: one cooperative launch.
    {
        m3SubstepParams spArg = sp;
        if (w->tailLimit > 0)
        {
            spArg.tailLimit = w->tailLimit;
        }
        spArg.flow = w->flow;
        void* args[] = {(void*)&D, (void*)&spArg};
        int blocks = w->coopBlocks;
        if (!w->coopFixed)
        {
            int perBody = (nb + w->coopThreads - 1) / w->coopThreads;
            blocks = perBody < 1 ? 1 : (perBody < blocks ? perBody : blocks);
        }
        if (w->flow)
        {
            // One workgroup per cell; the colored form, taken when a step
            // does not fit the cells, runs on the same grid.
            blocks = w->cellsX * w->cellsZ;
            STAGE(ST_CELLS, KL(k_cells, dim3(1), dim3(M3_CELL_BINS), 0, st, D, w->cellsX, w->cellsZ));
        }
        STAGE(ST_SUBSTEPS, if (GridLaunch((const void*)k_substeps, dim3(blocks), dim3(w->coopThreads), args,
                                          st) != 0) { return -1; });
    }
    if (w->anyRestitution)
    {
        STAGE(ST_RESTITUTION, KL(k_restitution, dim3(1), dim3(1), 0, st, D, p->restitutionThreshold));
    }
    if (v.maxJoint > 0)
    {
        STAGE(ST_STORE, KL(k_store_joints, dim3(Blocks(v.maxJoint, T)), dim3(T), 0, st, D));
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// The continuous pass and the rest of the step.  `ranks` (maxShape entries,
// or null) is the depth-first rank of every shape in the broadphase tree
// after this step's leaf refresh; with it the continuous pass also sweeps
// against tree shapes.
static int EnqueueStepEnd(m3gpuWorld* w, const m3gpuStepParams* p, const int32_t* ranks)
{
    hipStream_t st = w->stream;
    const m3Dev* D = w->devDev + w->cur;
    const m3View& v = w->view;
    const int T = 256;
    const int32_t nb = v.maxBody;
    const int32_t ns = v.maxShape;
    const int tree = ranks != nullptr ? 1 : 0;
    if (tree != w->dev.v.treeCcd)
    {
        // The flag lives in the device copies of the world's entry.
        w->dev.v.treeCcd = tree;
        w->view.treeCcd = tree;
        if (RefreshDevCopies(w) != 0)
        {
            return -1;
        }
    }
    if (tree)
    {
        if (w->rankCopyPending)
        {
            HIP_TRY(hipEventSynchronize(w->ranksCopied));
            w->rankCopyPending = 0;
        }
        memcpy(w->rankStage, ranks, (size_t)ns * sizeof(int32_t));
        HIP_TRY(hipMemcpyAsync(v.shapeRank, w->rankStage, (size_t)ns * sizeof(int32_t),
                               hipMemcpyHostToDevice, st));
        HIP_TRY(hipEventRecord(w->ranksCopied, st));
        w->rankCopyPending = 1;
    }
    // Impulse store, the continuous pass and the sleep flags' reset.
    const int32_t np = v.pairCapacity;
    const int bullets = p->continuousEnabled && w->dev.ccdBullets;
    const int32_t finishBase = Blocks(np > nb ? np : nb, 64);
    STAGE(ST_STORE, KL(k_finish, dim3(finishBase + Blocks(nb, 64)), dim3(64), 0, st, D,
                       p->continuousEnabled ? 1 : 0, bullets ? 0 : 1, p->dt, p->sleepEnabled, finishBase));
    if (p->continuousEnabled)
    {
        if (w->dev.ccdBullets)
        {
            STAGE(ST_CCD, KL(k_ccd_bullets, dim3(Blocks(nb, 64)), dim3(64), 0, st, D));
            STAGE(ST_CCD, KL(k_ccd_apply, dim3(Blocks(nb, T)), dim3(T), 0, st, D));
        }
    }
    if (bullets)
    {
        STAGE(ST_SLEEP_TIMERS,
              KL(k_sleep_timers, dim3(Blocks(nb, T)), dim3(T), 0, st, D, p->dt, p->sleepEnabled));
    }
    STAGE(ST_END, KL(k_close, dim3(Blocks(nb, T)), dim3(T), 0, st, D, p->sleepEnabled));
    if (w->pairTable)
    {
        KL(k_pair_table, dim3(Blocks(v.pairCapacity, T)), dim3(T), 0, st, D, w->tableKeys, w->tableMans);
    }
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

extern "C" int m3gpu_set_pair_table(m3gpuWorld* w, int32_t on)
{
    if (w == nullptr)
    {
        return -1;
    }
    if (on && w->tableKeys == nullptr)
    {
        if (DevAlloc(&w->tableKeys, w->view.pairCapacity) || DevAlloc(&w->tableMans, w->view.pairCapacity))
        {
            return -1;
        }
        // The clears run on the null stream; finish them before an upload
        // or a step writes the table.
        HIP_TRY(hipStreamSynchronize(nullptr));
    }
    w->pairTable = on ? 1 : 0;
    return 0;
}

extern "C" int m3gpu_step_needs_tree(const m3gpuWorld* w, const m3gpuStepParams* p)
{
    return w != nullptr && p != nullptr && p->continuousEnabled && w->ccdNeedsTree;
}

extern "C" int m3gpu_step_supported(const m3gpuWorld* w, const m3gpuStepParams* p)
{
    return w != nullptr && p != nullptr && !(p->continuousEnabled && w->ccdUnsupported) &&
           !m3gpu_step_needs_tree(w, p);
}

static int StepArgs(const m3gpuWorld* w, const m3gpuStepParams* p, const char* who)
{
    if (w == nullptr || p == nullptr || !(p->dt > 0.0f) || p->substeps < 1 ||
        p->substeps > M3_MAX_SUBSTEPS)
    {
        snprintf(s_lastError, sizeof(s_lastError), "%s: bad arguments", who);
        return -1;
    }
    if (p->continuousEnabled && w->ccdUnsupported)
    {
        snprintf(s_lastError, sizeof(s_lastError),
                 "%s: continuous collision against mesh, voxel or heightfield shapes is not "
                 "simulated on the device path", who);
        return M3GPU_REFUSED;
    }
    return 0;
}

// Pinned staging and buffers of the pre-solve path, allocated at first use.
static int PreSolveBuffers(m3gpuWorld* w)
{
    if (w->preStage != nullptr)
    {
        return 0;
    }
    const int32_t np = w->view.pairCapacity;
    if (DevAlloc(&w->preCands, np) || DevAlloc(&w->preCount, 1) || DevAlloc(&w->vetoKeys, np) ||
        DevAlloc(&w->vetoFlags, np))
    {
        return -1;
    }
    HIP_TRY(hipHostMalloc((void**)&w->vetoStage, (size_t)np * sizeof(uint64_t), hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&w->preGathered, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&w->vetoCopied, hipEventDisableTiming));
    HIP_TRY(hipHostMalloc((void**)&w->preStage, (size_t)np * sizeof(m3gpuPreSolve), hipHostMallocDefault));
    // DevAlloc clears on the null stream, which does not order against the
    // world's non-blocking stream: finish the clears before the step's
    // copies and kernels write these buffers.
    HIP_TRY(hipStreamSynchronize(nullptr));
    return 0;
}

extern "C" int m3gpu_step_begin(m3gpuWorld* w, const m3gpuStepParams* p, uint32_t flags,
                                const uint64_t* vetoKeys, int32_t vetoCount)
{
    const int rc = StepArgs(w, p, "m3gpu_step_begin");
    if (rc != 0)
    {
        return rc;
    }
    const bool presolve = (flags & M3GPU_STEP_PRESOLVE) != 0;
    if ((presolve || vetoCount > 0) && PreSolveBuffers(w) != 0)
    {
        return -1;
    }
    if (vetoCount < 0 || vetoCount > w->view.pairCapacity || (vetoCount > 0 && vetoKeys == nullptr))
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_step_begin: bad veto list");
        return -1;
    }
    if (JoinBatch(w) != 0 || EnqueueStepA(w, p, (flags & M3GPU_STEP_LEAVES) != 0) != 0)
    {
        return -1;
    }
    w->streamBusy = 1;
    if (presolve)
    {
        const int T = 256;
        const m3Dev* D = w->devDev + w->cur;
        HIP_TRY(hipMemsetAsync(w->preCount, 0, sizeof(int32_t), w->stream));
        KL(k_presolve_gather, dim3(Blocks(w->view.pairCapacity, T)), dim3(T), 0, w->stream, D, w->preCands,
           w->preCount);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(w->preGathered, w->stream));
        return 0;
    }
    return EnqueueStepB(w, p, vetoKeys, vetoCount);
}

static int PreOrder(const void* a, const void* b)
{
    const int32_t x = ((const m3gpuPreSolve*)a)->pair;
    const int32_t y = ((const m3gpuPreSolve*)b)->pair;
    return (x > y) - (x < y);
}

extern "C" int m3gpu_step_presolve(m3gpuWorld* w, const m3gpuPreSolve** candidates, int32_t* count)
{
    if (w == nullptr || candidates == nullptr || count == nullptr || w->preStage == nullptr)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_step_presolve: no pre-solve step begun");
        return -1;
    }
    int32_t n = 0;
    HIP_TRY(hipEventSynchronize(w->preGathered));
    HIP_TRY(hipMemcpy(&n, w->preCount, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (n < 0 || n > w->view.pairCapacity)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_step_presolve: bad candidate count");
        return -1;
    }
    HIP_TRY(hipMemcpy(w->preStage, w->preCands, (size_t)n * sizeof(m3gpuPreSolve), hipMemcpyDeviceToHost));
    qsort(w->preStage, (size_t)n, sizeof(m3gpuPreSolve), PreOrder);
    *candidates = w->preStage;
    *count = n;
    return 0;
}

extern "C" int m3gpu_step_vetoes(m3gpuWorld* w, const m3gpuStepParams* p, const uint64_t* vetoKeys,
                                 int32_t vetoCount)
{
    const int rc = StepArgs(w, p, "m3gpu_step_vetoes");
    if (rc != 0)
    {
        return rc;
    }
    if (vetoCount < 0 || vetoCount > w->view.pairCapacity || (vetoCount > 0 && vetoKeys == nullptr))
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_step_vetoes: bad veto list");
        return -1;
    }
    return EnqueueStepB(w, p, vetoKeys, vetoCount);
}

extern "C" const double* m3gpu_step_leaves(m3gpuWorld* w)
{
    if (w == nullptr || hipEventSynchronize(w->leavesCopied) != hipSuccess)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_step_leaves: the copy failed");
        return nullptr;
    }
    return w->leafStage;
}

extern "C" int m3gpu_step_end(m3gpuWorld* w, const m3gpuStepParams* p, const int32_t* ranks)
{
    const int rc = StepArgs(w, p, "m3gpu_step_end");
    if (rc != 0)
    {
        return rc;
    }
    if (EnqueueStepEnd(w, p, ranks) != 0)
    {
        return -1;
    }
    return ProfileCollect(w, p->substeps);
}

extern "C" int m3gpu_step(m3gpuWorld* w, const m3gpuStepParams* p, m3gpuStepStats* stats)
{
    const int rc = StepArgs(w, p, "m3gpu_step");
    if (rc != 0)
    {
        return rc;
    }
    if (m3gpu_step_needs_tree(w, p))
    {
        snprintf(s_lastError, sizeof(s_lastError),
                 "m3gpu_step: the continuous pass meets tree shapes; step with m3gpu_step_begin, "
                 "the tree ranks and m3gpu_step_end");
        return M3GPU_REFUSED;
    }
    if (JoinBatch(w) != 0 || EnqueueStepA(w, p, false) != 0 || EnqueueStepB(w, p, nullptr, 0) != 0 ||
        EnqueueStepEnd(w, p, nullptr) != 0)
    {
        return -1;
    }
    w->streamBusy = 1;
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

// ---------------------------------------------------------------------------
// Batched stepping: independent worlds of one m3gpu_step_many call share
// every launch of the step (grid.y = world), each world in its own device
// buffers.  The world list of the last batched call stays cached with both
// parities of the kernel views, so repeated calls launch straight away.
// ---------------------------------------------------------------------------

struct m3gpuBatch
{
    int32_t count;
    int32_t capacity;
    m3gpuWorld** worlds;
    int* baseCur;        // each world's cur when the views were built
    int64_t* layouts;    // each world's layout version then
    int64_t steps;       // batch steps since the views were built
    m3Dev* hostViews;    // [2][capacity] staging
    m3Dev* devViews;     // [2][capacity]: entry q*capacity+k serves world k on steps s with s%2 == q
    hipStream_t stream;
    hipEvent_t done;
    int32_t maxBody;
    int32_t maxShape;
    int32_t maxPairs;
    int32_t maxEmit;
    int32_t maxJoint;
    int wideHulls;
    int64_t maxPlaneWork;
    int threads; // workgroup size of the per-world substep loop
    // The step's launch sequence captured once per parity and replayed.
    int useGraphs;
    hipGraphExec_t graph[2];
    m3gpuStepParams graphParams[2];
    int graphReady[2];
};

static m3gpuBatch s_batch;
static pthread_mutex_t s_batchLock = PTHREAD_MUTEX_INITIALIZER;
// The adjoint's process-lifetime arena and replay stream (defined with the
// adjoint below).
static void* s_adjArena;
static hipStream_t s_adjStream;
static size_t s_adjBytes;
static int s_batchThreadsWanted = 0; // set by m3gpu_set_batch_threads (0: default)

// Threads per world workgroup of the batched substep kernel: the tuned
// value, else MAUL3D_AMD_BATCH_THREADS, else 256.
static int BatchThreadsDefault(void)
{
    const char* env = getenv("MAUL3D_AMD_BATCH_THREADS");
    int t = s_batchThreadsWanted > 0 ? s_batchThreadsWanted : (env != nullptr ? atoi(env) : 256);
    t = t < 64 ? 64 : (t > M3_BLOCK_MAX_THREADS ? M3_BLOCK_MAX_THREADS : t);
    return t & ~63;
}

// A world whose latest work ran in the batch stream waits for it before
// anything else touches the world.
static int JoinBatch(m3gpuWorld* w)
{
    if (!w->batched)
    {
        return 0;
    }
    pthread_mutex_lock(&s_batchLock);
    hipError_t e = hipEventSynchronize(s_batch.done);
    w->batched = 0;
    pthread_mutex_unlock(&s_batchLock);
    if (e != hipSuccess)
    {
        snprintf(s_lastError, sizeof(s_lastError), "batch wait: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

static void BatchDropGraphs(void)
{
    for (int q = 0; q < 2; ++q)
    {
        if (s_batch.graphReady[q])
        {
            (void)hipGraphExecDestroy(s_batch.graph[q]);
            s_batch.graphReady[q] = 0;
        }
    }
}

static void BatchDrop(void)
{
    BatchDropGraphs();
    if (s_batch.count > 0)
    {
        (void)hipStreamSynchronize(s_batch.stream);
        for (int32_t k = 0; k < s_batch.count; ++k)
        {
            s_batch.worlds[k]->batched = 0;
        }
    }
    s_batch.count = 0;
}

static void BatchForget(m3gpuWorld* w)
{
    pthread_mutex_lock(&s_batchLock);
    for (int32_t k = 0; k < s_batch.count; ++k)
    {
        if (s_batch.worlds[k] == w)
        {
            BatchDrop();
            break;
        }
    }
    pthread_mutex_unlock(&s_batchLock);
}

static bool BatchEligible(m3gpuWorld* w)
{
    const m3View& v = w->view;
    return v.emitCapacity <= M3_BATCH_EMIT_MAX && v.maxShape <= M3_BATCH_SHAPES_MAX &&
           v.maxBody <= M3_COLOR_LDS_BODIES && Prof(w) == nullptr;
}

static int BatchReserve(int32_t n)
{
    if (n <= s_batch.capacity)
    {
        return 0;
    }
    if (s_batch.capacity == 0)
    {
        HIP_TRY(hipStreamCreateWithFlags(&s_batch.stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&s_batch.done, hipEventDisableTiming));
        const char* genv = getenv("MAUL3D_AMD_BATCH_GRAPHS");
        s_batch.useGraphs = genv == nullptr || atoi(genv) != 0;
        s_batch.threads = BatchThreadsDefault();
    }
    int32_t cap = n < 64 ? 64 : n;
    free(s_batch.worlds);
    free(s_batch.baseCur);
    free(s_batch.layouts);
    free(s_batch.hostViews);
    if (s_batch.devViews != nullptr)
    {
        HIP_TRY(hipFree(s_batch.devViews));
    }
    s_batch.worlds = (m3gpuWorld**)calloc((size_t)cap, sizeof(m3gpuWorld*));
    s_batch.baseCur = (int*)calloc((size_t)cap, sizeof(int));
    s_batch.layouts = (int64_t*)calloc((size_t)cap, sizeof(int64_t));
    s_batch.hostViews = (m3Dev*)calloc(2 * (size_t)cap, sizeof(m3Dev));
    if (s_batch.worlds == nullptr || s_batch.baseCur == nullptr || s_batch.layouts == nullptr ||
        s_batch.hostViews == nullptr)
    {
        snprintf(s_lastError, sizeof(s_lastError), "batch: out of host memory");
        s_batch.capacity = 0;
        return -1;
    }
    HIP_TRY(hipMalloc(&s_batch.devViews, 2 * (size_t)cap * sizeof(m3Dev)));
    s_batch.capacity = cap;
    return 0;
}

static bool BatchCurrent(m3gpuWorld* const* ws, int32_t n)
{
    if (s_batch.count != n)
    {
        return false;
    }
    for (int32_t k = 0; k < n; ++k)
    {
        m3gpuWorld* w = ws[k];
        if (s_batch.worlds[k] != w || w->devDirty || w->layout != s_batch.layouts[k] ||
            w->cur != (s_batch.baseCur[k] ^ (int)(s_batch.steps & 1)))
        {
            return false;
        }
    }
    return true;
}

static int BatchBuild(m3gpuWorld* const* ws, int32_t n)
{
    BatchDrop();
    if (BatchReserve(n) != 0)
    {
        return -1;
    }
    s_batch.maxBody = s_batch.maxShape = s_batch.maxPairs = s_batch.maxEmit = 0;
    s_batch.maxJoint = 0;
    s_batch.wideHulls = 0;
    s_batch.maxPlaneWork = 0;
    for (int32_t k = 0; k < n; ++k)
    {
        m3gpuWorld* w = ws[k];
        if (w->devDirty && RefreshDevCopies(w) != 0)
        {
            return -1;
        }
        HIP_TRY(hipStreamSynchronize(w->stream));
        w->streamBusy = 0;
        for (int q = 0; q < 2; ++q)
        {
            // Step s of the batch runs world k with cur = baseCur ^ 1 ^ (s & 1).
            int parity = w->cur ^ 1 ^ q;
            m3Dev& e = s_batch.hostViews[(size_t)q * s_batch.capacity + k];
            e = w->dev;
            e.v = w->view;
            e.v.keys = w->keysBuf[parity];
            e.v.manifolds = w->mansBuf[parity];
            e.v.oldKeys = w->keysBuf[parity ^ 1];
            e.v.oldManifolds = w->mansBuf[parity ^ 1];
            e.v.satBoxes = 0;
        }
        s_batch.worlds[k] = w;
        s_batch.baseCur[k] = w->cur;
        s_batch.layouts[k] = w->layout;
        const m3View& v = w->view;
        s_batch.maxBody = v.maxBody > s_batch.maxBody ? v.maxBody : s_batch.maxBody;
        s_batch.maxShape = v.maxShape > s_batch.maxShape ? v.maxShape : s_batch.maxShape;
        s_batch.maxPairs = v.pairCapacity > s_batch.maxPairs ? v.pairCapacity : s_batch.maxPairs;
        s_batch.maxEmit = v.emitCapacity > s_batch.maxEmit ? v.emitCapacity : s_batch.maxEmit;
        s_batch.maxJoint = v.maxJoint > s_batch.maxJoint ? v.maxJoint : s_batch.maxJoint;
        s_batch.wideHulls |= w->wideHulls;
        int64_t pw = (int64_t)v.maxShape * v.planeCount;
        s_batch.maxPlaneWork = pw > s_batch.maxPlaneWork ? pw : s_batch.maxPlaneWork;
    }
    HIP_TRY(hipMemcpy(s_batch.devViews, s_batch.hostViews,
                      2 * (size_t)s_batch.capacity * sizeof(m3Dev), hipMemcpyHostToDevice));
    s_batch.count = n;
    s_batch.steps = 0;
    return 0;
}

// The launch sequence of one batch step: the same stage order as
// EnqueueStep, each stage one launch over all worlds (grid.y = world).
// ---- tape records of a batched step (adjoint replays) ------------------------

// Step-start body state (before anything of the step runs).
__global__ void k_tape_start(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= v.maxBody)
    {
        return;
    }
    m3Tape* t = d.tape;
    t->xf0[b] = v.xf[b];
    t->lv0[b] = v.lv[b];
    t->av0[b] = v.av[b];
    t->force0[b] = v.force[b];
    t->slept[b] = 0;
    t->ccd[b] = 0;
    t->mover[b] = 0;
}

// The pair list after narrowphase and the carry's source of each point
// (the same search as m3NarrowPair's).
__global__ void k_tape_pairs(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.sc->pairCount)
    {
        return;
    }
    m3Tape* t = d.tape;
    const m3Manifold* m = &v.manifolds[i];
    uint64_t key = v.keys[i];
    t->keys[i] = key;
    t->man[i] = *m;
    t->carryPair[i] = -1;
    for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        t->carryPoint[4 * i + k] = -1;
    }
    int32_t oldCount = v.sc->oldPairCount;
    if (m->pointCount == 0 || oldCount == 0)
    {
        return;
    }
    const int32_t mid = m3FindKey(v.oldKeys, oldCount, key);
    if (mid < 0)
    {
        return;
    }
    t->carryPair[i] = mid;
    const m3Manifold* prev = &v.oldManifolds[mid];
    for (int32_t k = 0; k < m->pointCount; ++k)
    {
        for (int32_t o = 0; o < prev->pointCount; ++o)
        {
            if (prev->points[o].id == m->points[k].id)
            {
                t->carryPoint[4 * i + k] = (int8_t)o;
                break;
            }
        }
    }
}

// Counts, the constraints' pairs in slot order, color starts and movers.
__global__ void k_tape_solver(const m3Dev* __restrict__ ds)
{
    M3_WORLD;
    m3Tape* t = d.tape;
    int32_t nc = v.sc->constraintCount;
    int32_t nm = v.sc->moverCount;
    for (int32_t k = threadIdx.x; k < nc; k += blockDim.x)
    {
        t->consPair[k] = v.cons[k].manifoldIndex;
    }
    for (int32_t m = threadIdx.x; m < nm; m += blockDim.x)
    {
        t->mover[v.movers[m]] = 1;
    }
    if (threadIdx.x == 0)
    {
        t->counts[0] = v.sc->pairCount;
        t->counts[1] = nc;
        t->counts[2] = nm;
        for (int32_t k = 0; k < M3_GRAPH_COLORS + 2; ++k)
        {
            t->colorStarts[k] = v.sc->colorStarts[k];
        }
    }
}

static int BatchLaunches(hipStream_t st, const m3Dev* D, const m3gpuStepParams* p, bool taped = false)
{
    const uint32_t W = (uint32_t)s_batch.count;
    const int T = 256;
    const int BT = M3_BATCH_THREADS;
    const int32_t nb = s_batch.maxBody;
    const int32_t ns = s_batch.maxShape;
    const int32_t np = s_batch.maxPairs;
    const int32_t ne = s_batch.maxEmit;
    const m3SubstepParams sp = SubstepParams(p);
    if (taped)
    {
        KL(k_tape_start, dim3(Blocks(nb, T), W), dim3(T), 0, st, D);
    }
    KL(k_open, dim3(Blocks(ne > ns ? ne : ns, T), W), dim3(T), 0, st, D);
    KL(k_reset_sleep, dim3(1, W), dim3(1), 0, st, D);
    KL(k_pending_all, dim3(Blocks(nb, T), W), dim3(T), 0, st, D);
    KL(k_sap_sort_block, dim3(1, W), dim3(BT), 0, st, D);
    KL(k_sap_gather, dim3(Blocks(ns, T), W), dim3(T), 0, st, D);
    KL(k_sap, dim3(Blocks(ns, M3_SAP_SHAPES), W), dim3(M3_SAP_THREADS), 0, st, D);
    KL(k_planes, dim3(Blocks(s_batch.maxPlaneWork > ns ? s_batch.maxPlaneWork : ns, T), W), dim3(T), 0, st, D);
    KL(k_frozen, dim3(Blocks(np, T), W), dim3(T), 0, st, D);
    KL(k_emit_sort_block, dim3(1, W), dim3(BT), 0, st, D);
    KL(k_pairs_settle, dim3(Blocks(np > nb ? np : nb, M3_COMPACT_THREADS), W), dim3(M3_COMPACT_THREADS), 0, st,
       D, s_batch.wideHulls ? 1 : 0);
    KL(k_select_block, dim3(1, W), dim3(BT), 0, st, D, 0);
    if (s_batch.wideHulls)
    {
        KL(k_sat<SatLds>, dim3(SatBlocks(np, W), W), dim3(64), 0, st, D);
    }
    // The narrowphase binned by collision routine per world, which also
    // unites the bodies of each touching pair.
    KL(k_narrow_binned, dim3(Blocks(np, 64) + M3_NARROW_CLASSES, W), dim3(64), 0, st, D);
    if (taped)
    {
        KL(k_tape_pairs, dim3(Blocks(np, T), W), dim3(T), 0, st, D);
    }
    if (s_batch.maxJoint > 0)
    {
        KL(k_union_joints, dim3(Blocks(s_batch.maxJoint, T), W), dim3(T), 0, st, D);
    }
    KL(k_root, dim3(Blocks(nb, T), W), dim3(T), 0, st, D);
    KL(k_wake_apply, dim3(Blocks(nb, T), W), dim3(T), 0, st, D);
    KL(k_constraint_flags, dim3(Blocks(np, T), W), dim3(T), 0, st, D);
    KL(k_select_block, dim3(1, W), dim3(BT), 0, st, D, 1);
    KL(k_cons_bodies, dim3(Blocks(np, T), W), dim3(T), 0, st, D);
    KL(k_color<true>, dim3(1, W), dim3(M3_COLOR_THREADS), 0, st, D, 0);
    KL(k_prepare, dim3(Blocks(np, 64), W), dim3(64), 0, st, D, p->contactHertz,
       p->contactDampingRatio, sp.h, 0);
    if (s_batch.maxJoint > 0)
    {
        KL(k_prepare_joints, dim3(Blocks(s_batch.maxJoint, 64), W), dim3(64), 0, st, D, sp.h);
    }
    KL(k_mover_flags, dim3(Blocks(nb, T), W), dim3(T), 0, st, D, p->dt);
    KL(k_select_block, dim3(1, W), dim3(BT), 0, st, D, 2);
    if (taped)
    {
        KL(k_tape_solver, dim3(1, W), dim3(256), 0, st, D);
    }
    KL(k_substeps_block, dim3(1, W), dim3(s_batch.threads), 0, st, D, sp);
    KL(k_restitution, dim3(1, W), dim3(1), 0, st, D, p->restitutionThreshold);
    if (s_batch.maxJoint > 0)
    {
        KL(k_store_joints, dim3(Blocks(s_batch.maxJoint, T), W), dim3(T), 0, st, D);
    }
    const int32_t finishBase = Blocks(np > nb ? np : nb, 64);
    KL(k_finish, dim3(finishBase + Blocks(nb, 64), W), dim3(64), 0, st, D, p->continuousEnabled ? 1 : 0, 1,
       p->dt, p->sleepEnabled, finishBase);
    KL(k_close, dim3(Blocks(nb, T), W), dim3(T), 0, st, D, p->sleepEnabled);
    HIP_TRY(hipGetLastError());
    return 0;
}

// One step of every world in the batch.
static int BatchEnqueue(const m3gpuStepParams* p)
{
    hipStream_t st = s_batch.stream;
    const int q = (int)(s_batch.steps & 1);
    const m3Dev* D = s_batch.devViews + (size_t)q * s_batch.capacity;
    for (int32_t k = 0; k < s_batch.count; ++k)
    {
        m3gpuWorld* w = s_batch.worlds[k];
        w->cur ^= 1;
        m3View& v = w->view;
        v.keys = w->keysBuf[w->cur];
        v.manifolds = w->mansBuf[w->cur];
        v.oldKeys = w->keysBuf[w->cur ^ 1];
        v.oldManifolds = w->mansBuf[w->cur ^ 1];
        w->fullQuery = 0;
        w->batched = 1;
    }
    if (s_batch.useGraphs)
    {
        if (s_batch.graphReady[q] && memcmp(&s_batch.graphParams[q], p, sizeof(*p)) != 0)
        {
            HIP_TRY(hipGraphExecDestroy(s_batch.graph[q]));
            s_batch.graphReady[q] = 0;
        }
        if (!s_batch.graphReady[q])
        {
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = BatchLaunches(st, D, p);
            hipError_t e = hipStreamEndCapture(st, &g);
            if (rc != 0 || e != hipSuccess)
            {
                snprintf(s_lastError, sizeof(s_lastError), "batch graph capture failed: %s",
                         hipGetErrorString(e));
                return -1;
            }
            HIP_TRY(hipGraphInstantiate(&s_batch.graph[q], g, nullptr, nullptr, 0));
            HIP_TRY(hipGraphDestroy(g));
            s_batch.graphParams[q] = *p;
            s_batch.graphReady[q] = 1;
        }
        HIP_TRY(hipGraphLaunch(s_batch.graph[q], st));
    }
    else if (BatchLaunches(st, D, p) != 0)
    {
        return -1;
    }
    HIP_TRY(hipEventRecord(s_batch.done, st));
    s_batch.steps += 1;
    return 0;
}

// ---------------------------------------------------------------------------
// Launch tuning.  None of these change what a step computes (every
// participant count and workgroup size gives the same bits); they only
// choose launch shapes, so an autotuner can time candidates on a world's
// real state and keep the fastest.
// ---------------------------------------------------------------------------

extern "C" int m3gpu_set_substep_tuning(m3gpuWorld* w, int32_t blocks, int32_t tailLimit)
{
    if (w == nullptr || blocks < 0 || tailLimit < 0)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_set_substep_tuning: bad arguments");
        return -1;
    }
    if (blocks > w->coopMax)
    {
        snprintf(s_lastError, sizeof(s_lastError),
                 "m3gpu_set_substep_tuning: %d workgroups exceed the %d that can be co-resident",
                 blocks, w->coopMax);
        return -1;
    }
    w->coopBlocks = blocks > 0 ? blocks : w->coopDefault;
    w->coopFixed = blocks > 0 || w->coopEnvFixed;
    w->tailLimit = tailLimit;
    return 0;
}

extern "C" int m3gpu_get_substep_tuning(const m3gpuWorld* w, int32_t* blocks, int32_t* tailLimit)
{
    if (w == nullptr)
    {
        return -1;
    }
    if (blocks != nullptr)
    {
        int b = w->coopBlocks;
        if (!w->coopFixed)
        {
            const int perBody = (w->view.maxBody + w->coopThreads - 1) / w->coopThreads;
            b = perBody < 1 ? 1 : (perBody < b ? perBody : b);
        }
        *blocks = b;
    }
    if (tailLimit != nullptr)
    {
        *tailLimit = w->tailLimit > 0 ? w->tailLimit : M3_TAIL_CONSTRAINTS;
    }
    return 0;
}

extern "C" int m3gpu_set_batch_threads(int32_t threads)
{
    if (threads != 0 && (threads < 64 || threads > M3_BLOCK_MAX_THREADS || threads % 64 != 0))
    {
        snprintf(s_lastError, sizeof(s_lastError),
                 "m3gpu_set_batch_threads: threads must be 0 or a multiple of 64 in [64, %d]",
                 M3_BLOCK_MAX_THREADS);
        return -1;
    }
    pthread_mutex_lock(&s_batchLock);
    s_batchThreadsWanted = threads;
    if (s_batch.capacity > 0)
    {
        if (s_batch.count > 0)
        {
            (void)hipStreamSynchronize(s_batch.stream);
        }
        // The captured launch sequences hold the old workgroup size.
        BatchDropGraphs();
        s_batch.threads = BatchThreadsDefault();
    }
    pthread_mutex_unlock(&s_batchLock);
    return 0;
}

extern "C" int32_t m3gpu_get_batch_threads(void)
{
    pthread_mutex_lock(&s_batchLock);
    const int32_t t = s_batch.capacity > 0 ? s_batch.threads : BatchThreadsDefault();
    pthread_mutex_unlock(&s_batchLock);
    return t;
}

extern "C" int m3gpu_step_many(m3gpuWorld* const* worlds, const m3gpuStepParams* params,
                               int32_t count)
{
    if (count <= 0)
    {
        return 0;
    }
    if (worlds == nullptr || params == nullptr || count > 65535)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_step_many: bad arguments");
        return -1;
    }
    // One launch sequence serves the call when every world fits the block
    // kernels, all step with the same parameters and none appears twice.
    bool batch = count >= 2;
    for (int32_t k = 0; k < count && batch; ++k)
    {
        m3gpuWorld* w = worlds[k];
        const m3gpuStepParams* p = &params[k];
        if (w == nullptr || !(p->dt > 0.0f) || p->substeps < 1 || p->substeps > M3_MAX_SUBSTEPS ||
            !m3gpu_step_supported(w, p) || !BatchEligible(w) ||
            memcmp(p, &params[0], sizeof(*p)) != 0)
        {
            batch = false;
        }
    }
    if (batch)
    {
        // A world listed twice steps twice, in order: no batch then.
        static int64_t s_stamp = 0;
        pthread_mutex_lock(&s_batchLock);
        int64_t stamp = ++s_stamp;
        for (int32_t k = 0; k < count; ++k)
        {
            if (worlds[k]->stamp == stamp)
            {
                batch = false;
            }
            worlds[k]->stamp = stamp;
        }
        pthread_mutex_unlock(&s_batchLock);
    }
    if (!batch)
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
    pthread_mutex_lock(&s_batchLock);
    int rc = 0;
    if (!BatchCurrent(worlds, count))
    {
        rc = BatchBuild(worlds, count);
    }
    else
    {
        // worlds stepped on their own streams since the last batch step
        for (int32_t k = 0; k < count && rc == 0; ++k)
        {
            if (worlds[k]->streamBusy)
            {
                if (hipStreamSynchronize(worlds[k]->stream) != hipSuccess)
                {
                    rc = -1;
                }
                worlds[k]->streamBusy = 0;
            }
        }
    }
    if (rc == 0)
    {
        rc = BatchEnqueue(&params[0]);
    }
    pthread_mutex_unlock(&s_batchLock);
    return rc;
}

// Releases what the library holds for the whole process: the batch's
// captured launch graphs, views, stream and event, and the adjoint arena and
// stream.  The host world calls it from an atexit handler registered after
// the HIP runtime initialised, so it runs before the runtime's own teardown.
extern "C" void m3gpu_shutdown(void)
{
    pthread_mutex_lock(&s_batchLock);
    BatchDrop();
    if (s_batch.capacity > 0)
    {
        (void)hipStreamSynchronize(s_batch.stream);
        (void)hipFree(s_batch.devViews);
        (void)hipEventDestroy(s_batch.done);
        (void)hipStreamDestroy(s_batch.stream);
        free(s_batch.worlds);
        free(s_batch.baseCur);
        free(s_batch.layouts);
        free(s_batch.hostViews);
    }
    memset(&s_batch, 0, sizeof(s_batch));
    pthread_mutex_unlock(&s_batchLock);
    if (s_adjStream != nullptr)
    {
        (void)hipStreamSynchronize(s_adjStream);
        (void)hipStreamDestroy(s_adjStream);
        s_adjStream = nullptr;
    }
    if (s_adjArena != nullptr)
    {
        (void)hipFree(s_adjArena);
        s_adjArena = nullptr;
        s_adjBytes = 0;
    }
}

extern "C" int m3gpu_sync(m3gpuWorld* w)
{
    if (w == nullptr || JoinBatch(w) != 0)
    {
        return -1;
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    w->streamBusy = 0;
    return CheckFault(w);
}

extern "C" int m3gpu_last_stats(m3gpuWorld* w, m3gpuStepStats* stats)
{
    if (w == nullptr || stats == nullptr || JoinBatch(w) != 0)
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
    if (w != nullptr && JoinBatch(w) != 0)
    {
        return -1;
    }
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

// ---------------------------------------------------------------------------
// Trajectory gradients (m3gpu_adjoint_many)
// ---------------------------------------------------------------------------
//
// The worlds step through the batched launch sequence (one workgroup per
// world in every block kernel): a forward pass that checkpoints each world's
// mutable state every `checkpointEvery` steps, then, segment by segment from
// the last, a replay from the segment's checkpoint with the tape kernels
// recording each step's intermediate values, and the reverse pass of
// m3_adjoint.h over the segment, one 64-lane workgroup per world (lanes over
// the constraints of a color, the bodies, the pairs).  Each world's state and
// parameters are restored at the end.

#define M3_ADJ_LANES 64

struct m3AdjCtx
{
    m3View v;
    uint64_t* keysBuf[2];
    m3Manifold* mansBuf[2];
    m3SerialParams sp;
    int32_t steps;
    int32_t every;
    int32_t segs;
    int32_t forwardOnly;
    int32_t cur0;
    double target[3];
    const m3Vec3* force;
    int32_t offKind;
    int32_t offBody;
    int32_t offComp;
    m3real offValue;
    m3real saveInvMass;
    m3Mat3 saveInvI;
    m3Mat3 saveInertia;
    m3real saveFriction;
    uint8_t* ckpt;
    int64_t ckptStride;
    m3Tape* tapes;
    m3AdjWorld aw;
    m3real* comBar; // [maxBody]
    double loss;
    float* grads; // [maxBody * 11]
};

__host__ __device__ static inline int64_t AdjAlign(int64_t x) { return (x + 15) & ~(int64_t)15; }

__host__ static int64_t AdjCheckpointBytes(int32_t nb, int32_t ns, int32_t np)
{
    int64_t b = 32;
    b += AdjAlign((int64_t)nb * sizeof(m3Transform));
    b += 4 * AdjAlign((int64_t)nb * sizeof(m3Vec3));
    b += AdjAlign((int64_t)nb * sizeof(m3real));
    b += 2 * AdjAlign((int64_t)nb);
    b += 2 * AdjAlign((int64_t)np * sizeof(uint64_t));
    b += AdjAlign((int64_t)np * sizeof(m3Manifold));
    b += AdjAlign((int64_t)ns * 6 * sizeof(double));
    return b;
}

template <typename T> __device__ static inline void AdjCopy(T* dst, const T* src, int64_t n, int lane)
{
    for (int64_t i = lane; i < n; i += M3_ADJ_LANES)
    {
        dst[i] = src[i];
    }
}

// Saves (save != 0) or loads world k's mutable state at checkpoint slot
// `slot` (slot < 0: the state before the call).  One workgroup per world.
__global__ void __launch_bounds__(M3_ADJ_LANES) k_adj_checkpoint(m3AdjCtx* ctxs, int32_t slot, int save)
{
    m3AdjCtx* c = &ctxs[blockIdx.x];
    const int lane = threadIdx.x;
    m3View& v = c->v;
    int32_t nb = v.maxBody;
    int32_t ns = v.maxShape;
    int32_t np = v.pairCapacity;
    int32_t k = slot < 0 ? c->segs : slot;
    int32_t cur = slot < 0 ? c->cur0 : (c->cur0 ^ ((slot * c->every) & 1));
    uint8_t* p = c->ckpt + (int64_t)k * c->ckptStride;
    int32_t* hdr = (int32_t*)p;
    uint64_t* step = (uint64_t*)(p + 16);
    uint8_t* q = p + 32;
#define ADJ_FIELD(ptr, T, n)                                                                       \
    T* ptr = (T*)q;                                                                                \
    q += AdjAlign((int64_t)(n) * (int64_t)sizeof(T))
    ADJ_FIELD(xf, m3Transform, nb);
    ADJ_FIELD(lv, m3Vec3, nb);
    ADJ_FIELD(av, m3Vec3, nb);
    ADJ_FIELD(fo, m3Vec3, nb);
    ADJ_FIELD(to, m3Vec3, nb);
    ADJ_FIELD(st, m3real, nb);
    ADJ_FIELD(aw, uint8_t, nb);
    ADJ_FIELD(pe, uint8_t, nb);
    ADJ_FIELD(keys, uint64_t, np);
    ADJ_FIELD(sk, uint64_t, np);
    ADJ_FIELD(mans, m3Manifold, np);
    ADJ_FIELD(leaf, double, 6 * (int64_t)ns);
#undef ADJ_FIELD
    if (save)
    {
        int32_t pc = v.sc->pairCount;
        int32_t zc = v.sc->sleepCount;
        if (lane == 0)
        {
            hdr[0] = cur;
            hdr[1] = pc;
            hdr[2] = zc;
            hdr[3] = v.sc->fullQuery;
            *step = v.sc->stepCount;
        }
        AdjCopy(xf, v.xf, nb, lane);
        AdjCopy(lv, v.lv, nb, lane);
        AdjCopy(av, v.av, nb, lane);
        AdjCopy(fo, v.force, nb, lane);
        AdjCopy(to, v.torque, nb, lane);
        AdjCopy(st, v.sleepTime, nb, lane);
        AdjCopy(aw, v.awake, nb, lane);
        AdjCopy(pe, v.pending, nb, lane);
        AdjCopy(keys, c->keysBuf[cur], pc, lane);
        AdjCopy(mans, c->mansBuf[cur], pc, lane);
        AdjCopy(sk, v.sleepKeys, zc, lane);
        AdjCopy(leaf, v.leaf, 6 * (int64_t)ns, lane);
    }
    else
    {
        int32_t pc = hdr[1];
        int32_t zc = hdr[2];
        __syncthreads();
        if (lane == 0)
        {
            v.sc->pairCount = pc;
            v.sc->sleepCount = zc;
            v.sc->fullQuery = hdr[3];
            v.sc->stepCount = *step;
            v.sc->stall = 0;
        }
        AdjCopy(v.xf, xf, nb, lane);
        AdjCopy(v.lv, lv, nb, lane);
        AdjCopy(v.av, av, nb, lane);
        AdjCopy(v.force, fo, nb, lane);
        AdjCopy(v.torque, to, nb, lane);
        AdjCopy(v.sleepTime, st, nb, lane);
        AdjCopy(v.awake, aw, nb, lane);
        AdjCopy(v.pending, pe, nb, lane);
        AdjCopy(c->keysBuf[cur], keys, pc, lane);
        AdjCopy(c->mansBuf[cur], mans, pc, lane);
        AdjCopy(v.sleepKeys, sk, zc, lane);
        AdjCopy(v.leaf, leaf, 6 * (int64_t)ns, lane);
    }
}

// The requested parameter offset (finite differences), after the original
// values are saved for the restore.
__global__ void k_adj_offset(m3AdjCtx* ctxs)
{
    m3AdjCtx* c = &ctxs[blockIdx.x];
    m3View& v = c->v;
    int32_t b = c->offBody;
    if (threadIdx.x != 0 || c->offKind == 0 || b < 0 || b >= v.maxBody)
    {
        return;
    }
    m3real d = c->offValue;
    c->saveInvMass = v.invMass[b];
    c->saveInvI = v.invI[b];
    c->saveInertia = v.inertia[b];
    int32_t shape = v.shapeHead[b];
    c->saveFriction = shape >= 0 ? v.friction[shape] : 0.0f;
    m3Vec3* vec = c->offKind == 1 ? &v.lv[b] : (c->offKind == 2 ? &v.av[b] : nullptr);
    if (vec != nullptr)
    {
        if (c->offComp == 0)
            vec->x += d;
        else if (c->offComp == 1)
            vec->y += d;
        else
            vec->z += d;
    }
    if (c->offKind == 3)
    {
        m3real m = 1.0f / v.invMass[b];
        m3real m2 = m + d;
        v.invMass[b] = 1.0f / m2;
        m3real si = m / m2;
        m3real sf = m2 / m;
        v.invI[b].cx = m3MulSV3(si, v.invI[b].cx);
        v.invI[b].cy = m3MulSV3(si, v.invI[b].cy);
        v.invI[b].cz = m3MulSV3(si, v.invI[b].cz);
        v.inertia[b].cx = m3MulSV3(sf, v.inertia[b].cx);
        v.inertia[b].cy = m3MulSV3(sf, v.inertia[b].cy);
        v.inertia[b].cz = m3MulSV3(sf, v.inertia[b].cz);
    }
    if (c->offKind == 4 && shape >= 0)
    {
        v.friction[shape] += d;
    }
}

// The applied force of every dynamic body (cleared by each step).
__global__ void k_adj_force(m3AdjCtx* ctxs)
{
    m3AdjCtx* c = &ctxs[blockIdx.x];
    m3View& v = c->v;
    for (int32_t b = threadIdx.x; b < v.maxBody; b += blockDim.x)
    {
        if (v.bodyAlive[b] && v.type[b] == M3_DYNAMIC)
        {
            v.force[b] = c->force[b];
        }
    }
}

// Loss of the final mass-weighted centre of mass and the reverse pass's
// starting bars.
__global__ void __launch_bounds__(M3_ADJ_LANES) k_adj_loss(m3AdjCtx* ctxs)
{
    m3AdjCtx* c = &ctxs[blockIdx.x];
    m3View& v = c->v;
    const int lane = threadIdx.x;
    __shared__ double C[3];
    __shared__ double M;
    if (lane == 0)
    {
        double m0 = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
        for (int32_t b = 0; b < v.maxBody; ++b)
        {
            if (!v.bodyAlive[b] || v.type[b] != M3_DYNAMIC || v.invMass[b] == 0.0f)
            {
                continue;
            }
            double m = 1.0 / (double)v.invMass[b];
            m3Vec3 r = m3RotateVec3(v.xf[b].q, v.lc[b]);
            cx += m * (v.xf[b].p.x + (double)r.x);
            cy += m * (v.xf[b].p.y + (double)r.y);
            cz += m * (v.xf[b].p.z + (double)r.z);
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
    __syncthreads();
    if (c->forwardOnly)
    {
        return;
    }
    m3AdjWorld* a = &c->aw;
    m3real Cb[3];
    for (int i = 0; i < 3; ++i)
    {
        Cb[i] = (m3real)(2.0 * (C[i] - c->target[i]));
    }
    for (int32_t b = lane; b < v.maxBody; b += M3_ADJ_LANES)
    {
        memset(&a->body[b], 0, sizeof(m3BodyBar));
        memset(&a->param[b], 0, sizeof(m3ParamBar));
        c->comBar[b] = 0.0f;
        if (!v.bodyAlive[b] || v.type[b] != M3_DYNAMIC)
        {
            continue;
        }
        m3real m = 1.0f / v.invMass[b];
        m3Vec3 cbar = m3V3(Cb[0] * m / (m3real)M, Cb[1] * m / (m3real)M, Cb[2] * m / (m3real)M);
        a->body[b].p = cbar;
        m3AdjRotate(v.xf[b].q, v.lc[b], cbar, &a->body[b].q);
        m3Vec3 r = m3RotateVec3(v.xf[b].q, v.lc[b]);
        double dx = v.xf[b].p.x + (double)r.x - C[0];
        double dy = v.xf[b].p.y + (double)r.y - C[1];
        double dz = v.xf[b].p.z + (double)r.z - C[2];
        c->comBar[b] = (m3real)((Cb[0] * dx + Cb[1] * dy + Cb[2] * dz) / M);
    }
    for (int32_t i = lane; i < v.pairCapacity; i += M3_ADJ_LANES)
    {
        memset(&a->store[i], 0, sizeof(m3ManBar));
    }
}

// LDS of the reverse pass: body and parameter bars for the segment, and
// the step's re-prepared constraints, their bars and contributions when
// they fit (M3_ADJ_LDS_CONS of them).
#define M3_ADJ_LDS_CONS 64

__host__ __device__ static inline int64_t AdjLdsBytes(int32_t nb)
{
    return AdjAlign((int64_t)nb * sizeof(m3BodyBar)) + AdjAlign((int64_t)nb * sizeof(m3ParamBar)) +
           AdjAlign((int64_t)nb * sizeof(m3Vec3)) + AdjAlign((int64_t)nb * sizeof(m3Quat)) +
           AdjAlign((int64_t)M3_ADJ_LDS_CONS * sizeof(m3Constraint)) +
           AdjAlign((int64_t)M3_ADJ_LDS_CONS * sizeof(m3ConBar)) +
           AdjAlign((int64_t)M3_ADJ_LDS_CONS * sizeof(m3ConContrib));
}

// Reverse pass over segment `seg` (its tape is recorded).
__global__ void __launch_bounds__(M3_ADJ_LANES) k_adj_reverse(m3AdjCtx* ctxs, int32_t seg, int32_t region)
{
    extern __shared__ __align__(16) uint8_t s_adj[];
    m3AdjCtx* c = &ctxs[blockIdx.x];
    const int lane = threadIdx.x;
    const int32_t nb = c->v.maxBody;
    uint8_t* q = s_adj;
    m3BodyBar* body = (m3BodyBar*)q;
    q += AdjAlign((int64_t)nb * sizeof(m3BodyBar));
    m3ParamBar* param = (m3ParamBar*)q;
    q += AdjAlign((int64_t)nb * sizeof(m3ParamBar));
    m3Vec3* ccdP = (m3Vec3*)q;
    q += AdjAlign((int64_t)nb * sizeof(m3Vec3));
    m3Quat* ccdQ = (m3Quat*)q;
    q += AdjAlign((int64_t)nb * sizeof(m3Quat));
    m3Constraint* prep = (m3Constraint*)q;
    q += AdjAlign((int64_t)M3_ADJ_LDS_CONS * sizeof(m3Constraint));
    m3ConBar* cons = (m3ConBar*)q;
    q += AdjAlign((int64_t)M3_ADJ_LDS_CONS * sizeof(m3ConBar));
    m3ConContrib* contrib = (m3ConContrib*)q;
    for (int32_t b = lane; b < nb; b += M3_ADJ_LANES)
    {
        body[b] = c->aw.body[b];
        param[b] = c->aw.param[b];
    }
    __shared__ m3View sv;
    if (lane == 0)
    {
        sv = c->v;
    }
    m3AdjWorld a = c->aw;
    a.body = body;
    a.param = param;
    a.ccdP = ccdP;
    a.ccdQ = ccdQ;
    a.unsupported = 0;
    a.ccdPulled = 0;
    __syncthreads();
    int32_t s0 = seg * c->every;
    int32_t s1 = s0 + c->every < c->steps ? s0 + c->every : c->steps;
    for (int32_t s = s1 - 1; s >= s0; --s)
    {
        const m3Tape* t = &c->tapes[region * c->every + (s - s0)];
        int inLds = t->counts[1] <= M3_ADJ_LDS_CONS;
        a.prep = inLds ? prep : c->aw.prep;
        a.cons = inLds ? cons : c->aw.cons;
        a.contrib = inLds ? contrib : c->aw.contrib;
        __syncthreads();
        if (lane == 0)
        {
            m3AdjTapeView(&sv, t);
        }
        __syncthreads();
        m3AdjStepPar(&sv, &c->sp, t, &a, lane, M3_ADJ_LANES, m3BlockSync{});
        m3ManBar* sw = a.store;
        a.store = a.storePrev;
        a.storePrev = sw;
    }
    for (int32_t b = lane; b < nb; b += M3_ADJ_LANES)
    {
        c->aw.body[b] = body[b];
        c->aw.param[b] = param[b];
    }
    if (a.unsupported != 0)
    {
        atomicAdd(&c->aw.unsupported, a.unsupported);
    }
    if (a.ccdPulled != 0)
    {
        atomicAdd(&c->aw.ccdPulled, a.ccdPulled);
    }
    if (lane == 0)
    {
        c->aw.store = a.store;
        c->aw.storePrev = a.storePrev;
    }
}

// Gradients, then the original parameters back (the state is restored by
// the checkpoint kernel).
__global__ void __launch_bounds__(M3_ADJ_LANES) k_adj_finish(m3AdjCtx* ctxs)
{
    m3AdjCtx* c = &ctxs[blockIdx.x];
    m3View& v = c->v;
    if (!c->forwardOnly)
    {
        for (int32_t b = threadIdx.x; b < v.maxBody; b += M3_ADJ_LANES)
        {
            float* g = &c->grads[11 * b];
            const m3BodyBar* bb = &c->aw.body[b];
            const m3ParamBar* pb = &c->aw.param[b];
            int dyn = v.bodyAlive[b] && v.type[b] == M3_DYNAMIC;
            m3real im = v.invMass[b];
            m3real m = dyn ? 1.0f / im : 0.0f;
            g[0] = bb->v.x;
            g[1] = bb->v.y;
            g[2] = bb->v.z;
            g[3] = bb->w.x;
            g[4] = bb->w.y;
            g[5] = bb->w.z;
            g[6] = dyn ? pb->invMass * (-im * im) + pb->inertia / m - pb->invI / m + c->comBar[b] : 0.0f;
            g[7] = pb->friction;
            g[8] = pb->force.x;
            g[9] = pb->force.y;
            g[10] = pb->force.z;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && c->offKind != 0 && c->offBody >= 0 && c->offBody < v.maxBody)
    {
        int32_t b = c->offBody;
        v.invMass[b] = c->saveInvMass;
        v.invI[b] = c->saveInvI;
        v.inertia[b] = c->saveInertia;
        int32_t shape = v.shapeHead[b];
        if (shape >= 0)
        {
            v.friction[shape] = c->saveFriction;
        }
    }
}

// Device arena of the adjoint, kept between calls and grown as needed, and
// the stream its reverse passes run on.

static int AdjReserve(size_t bytes)
{
    if (bytes <= s_adjBytes)
    {
        return 0;
    }
    if (s_adjArena != nullptr)
    {
        HIP_TRY(hipFree(s_adjArena));
        s_adjArena = nullptr;
        s_adjBytes = 0;
    }
    HIP_TRY(hipMalloc(&s_adjArena, bytes));
    s_adjBytes = bytes;
    return 0;
}

static int64_t TapeStepBytes(int32_t nb, int32_t np, int32_t S)
{
    int64_t b = 0;
    b += 2 * AdjAlign((int64_t)nb * sizeof(m3Transform)); // xf0, xfEnd
    b += 3 * AdjAlign((int64_t)nb * sizeof(m3Vec3));      // lv0, av0, force0
    b += 3 * AdjAlign((int64_t)nb);                       // slept, ccd, mover
    b += AdjAlign((int64_t)np * sizeof(uint64_t));
    b += AdjAlign((int64_t)np * sizeof(m3Manifold));
    b += 2 * AdjAlign((int64_t)np * sizeof(int32_t)); // carryPair, consPair
    b += AdjAlign((int64_t)np * 4);                   // carryPoint
    b += AdjAlign(4 * sizeof(int32_t)) + AdjAlign((M3_GRAPH_COLORS + 2) * sizeof(int32_t));
    int64_t sb = (int64_t)S * nb;
    b += AdjAlign(sb * sizeof(m3Transform));
    b += 8 * AdjAlign(sb * sizeof(m3Vec3)); // slv sav sdp ilv iav plv pav rdp
    b += 2 * AdjAlign(sb * sizeof(m3Quat)); // sdq rdq
    b += AdjAlign((int64_t)S * 2 * np * sizeof(m3TapePass));
    b += AdjAlign((int64_t)np * sizeof(m3TapePass));
    return b;
}

static int64_t BarsBytes(int32_t nb, int32_t np)
{
    return AdjAlign((int64_t)nb * sizeof(m3BodyBar)) + AdjAlign((int64_t)nb * sizeof(m3ParamBar)) +
           AdjAlign((int64_t)np * sizeof(m3ConBar)) + 2 * AdjAlign((int64_t)np * sizeof(m3ManBar)) +
           AdjAlign((int64_t)np * sizeof(m3Constraint)) + AdjAlign((int64_t)np * sizeof(m3ConContrib)) +
           AdjAlign((int64_t)nb * sizeof(m3Vec3)) + AdjAlign((int64_t)nb * sizeof(m3Quat)) +
           AdjAlign((int64_t)nb * sizeof(m3real)) + AdjAlign((int64_t)nb * 11 * sizeof(float)) +
           AdjAlign((int64_t)nb * sizeof(m3Vec3));
}

template <typename T> static T* Carve(uint8_t** p, int64_t n)
{
    T* r = (T*)*p;
    *p += AdjAlign(n * (int64_t)sizeof(T));
    return r;
}

static void CarveTape(m3Tape* t, uint8_t** p, int32_t nb, int32_t np, int32_t S)
{
    t->bodyCap = nb;
    t->pairCap = np;
    t->substeps = S;
    t->xf0 = Carve<m3Transform>(p, nb);
    t->xfEnd = Carve<m3Transform>(p, nb);
    t->lv0 = Carve<m3Vec3>(p, nb);
    t->av0 = Carve<m3Vec3>(p, nb);
    t->force0 = Carve<m3Vec3>(p, nb);
    t->slept = Carve<uint8_t>(p, nb);
    t->ccd = Carve<uint8_t>(p, nb);
    t->mover = Carve<uint8_t>(p, nb);
    t->keys = Carve<uint64_t>(p, np);
    t->man = Carve<m3Manifold>(p, np);
    t->carryPair = Carve<int32_t>(p, np);
    t->consPair = Carve<int32_t>(p, np);
    t->carryPoint = Carve<int8_t>(p, 4 * (int64_t)np);
    t->counts = Carve<int32_t>(p, 4);
    t->colorStarts = Carve<int32_t>(p, M3_GRAPH_COLORS + 2);
    int64_t sb = (int64_t)S * nb;
    t->sxf = Carve<m3Transform>(p, sb);
    t->slv = Carve<m3Vec3>(p, sb);
    t->sav = Carve<m3Vec3>(p, sb);
    t->sdp = Carve<m3Vec3>(p, sb);
    t->ilv = Carve<m3Vec3>(p, sb);
    t->iav = Carve<m3Vec3>(p, sb);
    t->plv = Carve<m3Vec3>(p, sb);
    t->pav = Carve<m3Vec3>(p, sb);
    t->rdp = Carve<m3Vec3>(p, sb);
    t->sdq = Carve<m3Quat>(p, sb);
    t->rdq = Carve<m3Quat>(p, sb);
    t->pass = Carve<m3TapePass>(p, (int64_t)S * 2 * np);
    t->fin = Carve<m3TapePass>(p, np);
}

static int AdjointRun(m3gpuWorld* const* ws, const m3gpuStepParams* ps, const m3gpuAdjointRequest* rqs,
                      int32_t count)
{
    const m3gpuAdjointRequest* r0 = &rqs[0];
    const int32_t steps = r0->steps;
    const int32_t every = r0->checkpointEvery > 0 ? r0->checkpointEvery : 10;
    const int32_t segs = (steps + every - 1) / every;
    const int forwardOnly = r0->forwardOnly;
    // The last `resident` segments keep the tapes the forward pass records;
    // earlier segments are replayed from their checkpoints into region 0.
    const char* renv = getenv("MAUL3D_AMD_ADJOINT_RESIDENT");
    int32_t resident = renv != nullptr ? atoi(renv) : 1;
    resident = resident < 1 ? 1 : (resident > segs ? segs : resident);
    // Replayed segments alternate between two regions so the replay of one
    // segment overlaps the reverse pass of the next.
    const int32_t spare = resident < segs ? (segs - resident > 1 ? 2 : 1) : 0;
    const int32_t regions = resident + spare;
    auto regionOf = [&](int32_t seg) -> int32_t {
        if (seg >= segs - resident)
        {
            return spare + seg - (segs - resident);
        }
        return seg % spare;
    };
    for (int32_t k = 0; k < count; ++k)
    {
        m3gpuWorld* w = ws[k];
        if (w == nullptr || rqs[k].steps != steps || rqs[k].checkpointEvery != r0->checkpointEvery ||
            rqs[k].forwardOnly != forwardOnly || memcmp(&ps[k], &ps[0], sizeof(ps[0])) != 0 ||
            !BatchEligible(w) || !m3gpu_step_supported(w, &ps[k]) || steps < 1)
        {
            snprintf(s_lastError, sizeof(s_lastError),
                     "m3gpu_adjoint_many: worlds must share steps, checkpoints and parameters and fit the batched kernels");
            return 1;
        }
        for (int32_t j = 0; j < k; ++j)
        {
            if (ws[j] == w)
            {
                snprintf(s_lastError, sizeof(s_lastError), "m3gpu_adjoint_many: a world appears twice");
                return 1;
            }
        }
    }
    if (BatchBuild(ws, count) != 0)
    {
        return -1;
    }
    hipStream_t st = s_batch.stream;
    // Arena: per world checkpoints, force, bars, tapes; then the contexts and
    // the per-step kernel views.
    std::vector<int64_t> offs(count + 1, 0);
    for (int32_t k = 0; k < count; ++k)
    {
        const m3View& v = ws[k]->view;
        int64_t bytes = (int64_t)(segs + 1) * AdjCheckpointBytes(v.maxBody, v.maxShape, v.pairCapacity);
        bytes += AdjAlign((int64_t)v.maxBody * sizeof(m3Vec3));
        bytes += BarsBytes(v.maxBody, v.pairCapacity);
        if (!forwardOnly)
        {
            bytes += (int64_t)regions * every * TapeStepBytes(v.maxBody, v.pairCapacity, ps[k].substeps);
            bytes += AdjAlign((int64_t)regions * every * sizeof(m3Tape));
        }
        offs[k + 1] = offs[k] + AdjAlign(bytes);
    }
    int64_t ctxBytes = AdjAlign((int64_t)count * sizeof(m3AdjCtx));
    int64_t viewBytes = AdjAlign((int64_t)steps * count * sizeof(m3Dev));
    if (AdjReserve((size_t)(offs[count] + ctxBytes + 2 * viewBytes)) != 0)
    {
        return -1;
    }
    uint8_t* arena = (uint8_t*)s_adjArena;
    m3AdjCtx* dctx = (m3AdjCtx*)(arena + offs[count]);
    m3Dev* dFwd = (m3Dev*)(arena + offs[count] + ctxBytes);
    m3Dev* dRep = (m3Dev*)(arena + offs[count] + ctxBytes + viewBytes);
    std::vector<m3AdjCtx> ctx(count);
    std::vector<m3Tape> tapes;
    std::vector<float> hforce;
    std::vector<m3Dev> hv((size_t)steps * count);
    std::vector<m3Dev> hr(forwardOnly ? 0 : (size_t)steps * count);
    for (int32_t k = 0; k < count; ++k)
    {
        m3gpuWorld* w = ws[k];
        const m3gpuAdjointRequest* r = &rqs[k];
        m3AdjCtx& c = ctx[k];
        memset(&c, 0, sizeof(c));
        const m3View& v = w->view;
        int32_t nb = v.maxBody, ns = v.maxShape, np = v.pairCapacity;
        c.v = v;
        c.keysBuf[0] = w->keysBuf[0];
        c.keysBuf[1] = w->keysBuf[1];
        c.mansBuf[0] = w->mansBuf[0];
        c.mansBuf[1] = w->mansBuf[1];
        c.cur0 = w->cur;
        const m3gpuStepParams* p = &ps[k];
        c.sp.dt = p->dt;
        c.sp.substeps = p->substeps;
        c.sp.gravity = m3V3(p->gravity[0], p->gravity[1], p->gravity[2]);
        c.sp.contactHertz = p->contactHertz;
        c.sp.contactDampingRatio = p->contactDampingRatio;
        c.sp.contactPushMaxSpeed = p->contactPushMaxSpeed;
        c.sp.restitutionThreshold = p->restitutionThreshold;
        c.sp.maximumLinearSpeed = p->maximumLinearSpeed;
        c.sp.maximumAngularSpeed = p->maximumAngularSpeed;
        c.sp.sleepEnabled = p->sleepEnabled;
        c.sp.continuousEnabled = p->continuousEnabled;
        c.steps = steps;
        c.every = every;
        c.segs = segs;
        c.forwardOnly = forwardOnly;
        for (int i = 0; i < 3; ++i)
        {
            c.target[i] = (double)r->target[i];
        }
        c.offKind = r->offsetKind;
        c.offBody = r->offsetBody;
        c.offComp = r->offsetComp;
        c.offValue = (m3real)r->offsetValue;
        uint8_t* q = arena + offs[k];
        c.ckptStride = AdjCheckpointBytes(nb, ns, np);
        c.ckpt = Carve<uint8_t>(&q, (int64_t)(segs + 1) * c.ckptStride);
        m3Vec3* force = Carve<m3Vec3>(&q, nb);
        c.force = force;
        hforce.assign((size_t)nb * 3, 0.0f);
        if (r->force != nullptr)
        {
            memcpy(hforce.data(), r->force, (size_t)nb * 3 * sizeof(float));
        }
        HIP_TRY(hipMemcpy(force, hforce.data(), (size_t)nb * sizeof(m3Vec3), hipMemcpyHostToDevice));
        c.aw.body = Carve<m3BodyBar>(&q, nb);
        c.aw.param = Carve<m3ParamBar>(&q, nb);
        c.aw.cons = Carve<m3ConBar>(&q, np);
        c.aw.store = Carve<m3ManBar>(&q, np);
        c.aw.storePrev = Carve<m3ManBar>(&q, np);
        c.aw.prep = Carve<m3Constraint>(&q, np);
        c.aw.contrib = Carve<m3ConContrib>(&q, np);
        c.aw.ccdP = Carve<m3Vec3>(&q, nb);
        c.aw.ccdQ = Carve<m3Quat>(&q, nb);
        c.comBar = Carve<m3real>(&q, nb);
        c.grads = Carve<float>(&q, (int64_t)nb * 11);
        if (!forwardOnly)
        {
            m3Tape* dt = Carve<m3Tape>(&q, (int64_t)regions * every);
            tapes.resize((size_t)regions * every);
            for (int32_t s = 0; s < regions * every; ++s)
            {
                CarveTape(&tapes[s], &q, nb, np, p->substeps);
            }
            HIP_TRY(hipMemcpy(dt, tapes.data(), tapes.size() * sizeof(m3Tape), hipMemcpyHostToDevice));
            c.tapes = dt;
        }
        // Kernel views of every step: step s writes buffer cur0 ^ ((s + 1) & 1).
        for (int32_t s = 0; s < steps; ++s)
        {
            int par = w->cur ^ ((s + 1) & 1);
            m3Dev e = w->dev;
            e.v = w->view;
            e.v.keys = w->keysBuf[par];
            e.v.manifolds = w->mansBuf[par];
            e.v.oldKeys = w->keysBuf[par ^ 1];
            e.v.oldManifolds = w->mansBuf[par ^ 1];
            e.tape = nullptr;
            hv[(size_t)s * count + k] = e;
            if (!forwardOnly)
            {
                e.tape = c.tapes + (size_t)regionOf(s / every) * every + (s % every);
                hr[(size_t)s * count + k] = e;
            }
        }
    }
    HIP_TRY(hipMemcpy(dctx, ctx.data(), count * sizeof(m3AdjCtx), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dFwd, hv.data(), hv.size() * sizeof(m3Dev), hipMemcpyHostToDevice));
    if (!forwardOnly)
    {
        HIP_TRY(hipMemcpy(dRep, hr.data(), hr.size() * sizeof(m3Dev), hipMemcpyHostToDevice));
    }
    const dim3 G(count);
    const dim3 L(M3_ADJ_LANES);
    int32_t nbMax = 0;
    for (int32_t k = 0; k < count; ++k)
    {
        nbMax = ws[k]->view.maxBody > nbMax ? ws[k]->view.maxBody : nbMax;
    }
    const size_t ldsBytes = (size_t)AdjLdsBytes(nbMax);
    if (ldsBytes > 160 * 1024)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_adjoint_many: worlds too large for the reverse pass");
        return 1;
    }
    if (ldsBytes > 64 * 1024)
    {
        HIP_TRY(hipFuncSetAttribute((const void*)k_adj_reverse, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)ldsBytes));
    }
    hipLaunchKernelGGL(k_adj_checkpoint, G, L, 0, st, dctx, -1, 1);
    hipLaunchKernelGGL(k_adj_offset, G, dim3(1), 0, st, dctx);
    for (int32_t s = 0; s < steps; ++s)
    {
        if (s % every == 0)
        {
            hipLaunchKernelGGL(k_adj_checkpoint, G, L, 0, st, dctx, s / every, 1);
        }
        hipLaunchKernelGGL(k_adj_force, G, L, 0, st, dctx);
        bool taped = !forwardOnly && s / every >= segs - resident;
        if (BatchLaunches(st, (taped ? dRep : dFwd) + (size_t)s * count, &ps[0], taped) != 0)
        {
            return -1;
        }
    }
    hipLaunchKernelGGL(k_adj_loss, G, L, 0, st, dctx);
    if (!forwardOnly)
    {
        // Replays run on the batch stream, reverse passes on a second
        // stream: reverse(seg) waits for replay(seg); a replay into a region
        // waits for the reverse pass that last read it.
        if (s_adjStream == nullptr)
        {
            HIP_TRY(hipStreamCreateWithFlags(&s_adjStream, hipStreamNonBlocking));
        }
        hipStream_t rs = s_adjStream;
        std::vector<hipEvent_t> replayed(segs), reversed(segs);
        for (int32_t k = 0; k < segs; ++k)
        {
            HIP_TRY(hipEventCreateWithFlags(&replayed[k], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&reversed[k], hipEventDisableTiming));
        }
        HIP_TRY(hipEventRecord(replayed[segs - 1], st)); // the forward pass (resident tapes)
        for (int32_t seg = segs - 1; seg >= 0; --seg)
        {
            if (seg < segs - resident)
            {
                int32_t user = seg + spare; // previous reader of this segment's region
                if (spare == 2 && user < segs - resident)
                {
                    HIP_TRY(hipStreamWaitEvent(st, reversed[user], 0));
                }
                hipLaunchKernelGGL(k_adj_checkpoint, G, L, 0, st, dctx, seg, 0);
                int32_t s0 = seg * every;
                int32_t s1 = s0 + every < steps ? s0 + every : steps;
                for (int32_t s = s0; s < s1; ++s)
                {
                    hipLaunchKernelGGL(k_adj_force, G, L, 0, st, dctx);
                    if (BatchLaunches(st, dRep + (size_t)s * count, &ps[0], true) != 0)
                    {
                        return -1;
                    }
                }
                HIP_TRY(hipEventRecord(replayed[seg], st));
            }
            else
            {
                HIP_TRY(hipEventRecord(replayed[seg], st));
            }
            HIP_TRY(hipStreamWaitEvent(rs, replayed[seg], 0));
            hipLaunchKernelGGL(k_adj_reverse, G, L, ldsBytes, rs, dctx, seg, regionOf(seg));
            HIP_TRY(hipEventRecord(reversed[seg], rs));
            if (spare == 1 && seg < segs - resident)
            {
                // One replay region: the next replay waits for this reverse pass.
                HIP_TRY(hipStreamWaitEvent(st, reversed[seg], 0));
            }
        }
        HIP_TRY(hipStreamWaitEvent(st, reversed[0], 0));
        HIP_TRY(hipStreamSynchronize(st));
        for (int32_t k = 0; k < segs; ++k)
        {
            (void)hipEventDestroy(replayed[k]);
            (void)hipEventDestroy(reversed[k]);
        }
    }
    hipLaunchKernelGGL(k_adj_finish, G, L, 0, st, dctx);
    hipLaunchKernelGGL(k_adj_checkpoint, G, L, 0, st, dctx, -1, 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctx.data(), dctx, count * sizeof(m3AdjCtx), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<float> g;
    for (int32_t k = 0; k < count; ++k)
    {
        const m3AdjCtx& c = ctx[k];
        const m3gpuAdjointRequest* r = &rqs[k];
        int32_t nb = ws[k]->view.maxBody;
        if (r->loss)
        {
            *r->loss = c.loss;
        }
        if (r->unsupported)
        {
            *r->unsupported = c.aw.unsupported;
        }
        if (r->ccdPulled)
        {
            *r->ccdPulled = c.aw.ccdPulled;
        }
        if (forwardOnly)
        {
            continue;
        }
        g.resize((size_t)nb * 11);
        HIP_TRY(hipMemcpy(g.data(), c.grads, g.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int32_t b = 0; b < nb; ++b)
        {
            const float* e = &g[11 * (size_t)b];
            if (r->gradLinearVelocity)
            {
                memcpy(&r->gradLinearVelocity[3 * b], e, 3 * sizeof(float));
            }
            if (r->gradAngularVelocity)
            {
                memcpy(&r->gradAngularVelocity[3 * b], e + 3, 3 * sizeof(float));
            }
            if (r->gradMass)
            {
                r->gradMass[b] = e[6];
            }
            if (r->gradFriction)
            {
                r->gradFriction[b] = e[7];
            }
            if (r->gradForce)
            {
                memcpy(&r->gradForce[3 * b], e + 8, 3 * sizeof(float));
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Trajectory gradients in double precision (m3_adj64.h): one 64-lane
// workgroup per world over its image in the adjoint arena.
//
// Switched off (M3_ADJ64_DEVICE 0): its first run on the MI355X ended in a
// GPU memory fault whose cause is not found.  The same image and driver run
// clean on the CPU (one lane and 64 threaded lanes, exact-size allocations
// under AddressSanitizer) and equal the FP64 twin bit for bit; the kernel
// needs 20.9 KB of scratch per lane, four times any kernel run here so far.
// Until the cause is shown, precision 64 is refused.
// ---------------------------------------------------------------------------

#ifndef M3_ADJ64_DEVICE
#define M3_ADJ64_DEVICE 0
#endif

#if M3_ADJ64_DEVICE
__global__ void __launch_bounds__(M3_ADJ_LANES) k_adj64(m3Adj64World* worlds)
{
    __shared__ m3Adj64Shared sh;
    Adj64Body(&worlds[blockIdx.x], (int)threadIdx.x, M3_ADJ_LANES, m3BlockSync{}, &sh);
}

static int Adjoint64Run(m3gpuWorld* const* ws, const m3gpuStepParams* ps, const m3gpuAdjointRequest* rqs,
                        int32_t count)
{
    for (int32_t k = 0; k < count; ++k)
    {
        const m3gpuHostState* h = rqs[k].host;
        if (h == nullptr || rqs[k].steps < 1)
        {
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_adjoint_many: precision 64 needs the host state");
            return -1;
        }
        for (int32_t b = 0; b < h->maxBody; ++b)
        {
            if (h->bodyAlive[b] != 0 && h->bodyHasTarget[b] != 0)
            {
                snprintf(s_lastError, sizeof(s_lastError),
                         "m3gpu_adjoint_many: kinematic targets are outside the double-precision pass");
                return -1;
            }
        }
    }
    // Sizing pass, then one image: world records first, then each world's
    // arrays.
    std::vector<int64_t> base(count + 1);
    const int64_t recBytes = Adj64Align((int64_t)count * sizeof(m3Adj64World));
    base[0] = recBytes;
    std::vector<m3Adj64World> recs(count);
    for (int32_t k = 0; k < count; ++k)
    {
        m3Adj64Carver B{nullptr, nullptr, 0};
        Adj64Carve(B, ws[k]->caps.pairCapacity, &ps[k], &rqs[k], &recs[k]);
        base[k + 1] = base[k] + B.off;
    }
    const int64_t total = base[count];
    if (AdjReserve((size_t)total) != 0)
    {
        return -1;
    }
    uint8_t* image = (uint8_t*)calloc((size_t)total, 1);
    if (image == nullptr)
    {
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_adjoint_many: out of host memory");
        return -1;
    }
    uint8_t* dev = (uint8_t*)s_adjArena;
    for (int32_t k = 0; k < count; ++k)
    {
        m3Adj64Carver B{image + base[k], dev + base[k], 0};
        Adj64Carve(B, ws[k]->caps.pairCapacity, &ps[k], &rqs[k], &recs[k]);
    }
    memcpy(image, recs.data(), (size_t)count * sizeof(m3Adj64World));
    if (s_adjStream == nullptr)
    {
        const hipError_t e = hipStreamCreateWithFlags(&s_adjStream, hipStreamNonBlocking);
        if (e != hipSuccess)
        {
            free(image);
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_adjoint_many: %s", hipGetErrorString(e));
            return -1;
        }
    }
    hipStream_t st = s_adjStream;
    hipError_t e = hipMemcpyAsync(dev, image, (size_t)total, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
    {
        hipLaunchKernelGGL(k_adj64, dim3(count), dim3(M3_ADJ_LANES), 0, st, (m3Adj64World*)dev);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
    {
        e = hipMemcpyAsync(image, dev, (size_t)total, hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess)
    {
        e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess)
    {
        free(image);
        snprintf(s_lastError, sizeof(s_lastError), "m3gpu_adjoint_many: %s", hipGetErrorString(e));
        return -1;
    }
    int rc = 0;
    for (int32_t k = 0; k < count; ++k)
    {
        const m3Adj64World* c = (const m3Adj64World*)image + k;
        const m3gpuAdjointRequest* r = &rqs[k];
        if (c->failed != 0)
        {
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_adjoint_many: a step stalled");
            rc = -1;
            continue;
        }
        if (r->loss != nullptr)
        {
            *r->loss = c->loss;
        }
        if (r->forwardOnly)
        {
            continue;
        }
        const float* g = (const float*)(image + ((uint8_t*)c->grads - dev));
        const int32_t nb = r->host->maxBody;
        for (int32_t b = 0; b < nb; ++b)
        {
            const float* gb = g + 11 * b;
            for (int i = 0; i < 3; ++i)
            {
                if (r->gradLinearVelocity != nullptr)
                {
                    r->gradLinearVelocity[3 * b + i] = gb[i];
                }
                if (r->gradAngularVelocity != nullptr)
                {
                    r->gradAngularVelocity[3 * b + i] = gb[3 + i];
                }
                if (r->gradForce != nullptr)
                {
                    r->gradForce[3 * b + i] = gb[8 + i];
                }
            }
            if (r->gradMass != nullptr)
            {
                r->gradMass[b] = gb[6];
            }
            if (r->gradFriction != nullptr)
            {
                r->gradFriction[b] = gb[7];
            }
        }
        if (r->unsupported != nullptr)
        {
            *r->unsupported = c->aw.unsupported;
        }
        if (r->ccdPulled != nullptr)
        {
            *r->ccdPulled = c->aw.ccdPulled;
        }
    }
    free(image);
    return rc;
}
#endif // M3_ADJ64_DEVICE

extern "C" int m3gpu_adjoint_many(m3gpuWorld* const* ws, const m3gpuStepParams* ps,
                                  const m3gpuAdjointRequest* rqs, int32_t count)
{
    if (ws == nullptr || ps == nullptr || rqs == nullptr || count <= 0)
    {
        return -1;
    }
    for (int32_t k = 0; k < count; ++k)
    {
        if (ws[k] == nullptr || JoinBatch(ws[k]) != 0)
        {
            return -1;
        }
        HIP_TRY(hipStreamSynchronize(ws[k]->stream));
        ws[k]->streamBusy = 0;
    }
    for (int32_t k = 0; k < count; ++k)
    {
        if ((rqs[k].precision == 64) != (rqs[0].precision == 64))
        {
            snprintf(s_lastError, sizeof(s_lastError), "m3gpu_adjoint_many: requests mix precisions");
            return -1;
        }
    }
#if !M3_ADJ64_DEVICE
    if (rqs[0].precision == 64)
    {
        snprintf(s_lastError, sizeof(s_lastError),
                 "m3gpu_adjoint_many: precision 64 is switched off on the device");
        return M3GPU_REFUSED;
    }
#endif
    pthread_mutex_lock(&s_batchLock);
#if M3_ADJ64_DEVICE
    int rc = rqs[0].precision == 64 ? Adjoint64Run(ws, ps, rqs, count) : AdjointRun(ws, ps, rqs, count);
#else
    int rc = AdjointRun(ws, ps, rqs, count);
#endif
    BatchDrop();
    pthread_mutex_unlock(&s_batchLock);
    return rc;
}
