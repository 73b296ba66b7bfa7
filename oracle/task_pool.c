/* Host thread pool behind the reference engine's task hooks
 * (include/maul3d/world.h:14-26, m3WorldDef.enqueueTask / finishTask /
 * userTaskContext at :60-62).  Test infrastructure: it lets the CPU baseline
 * time the reference build with its parallel narrowphase
 * (src/manifold.c:2932-2937, minRange 16) and colored solve
 * (src/solver.c:2867-2870, minRange 8, colors of at least 16 constraints),
 * and lets the tests check that a threaded reference world hashes like the
 * serial one.  Nothing in the product links it.
 *
 * The contract the hooks follow: enqueueTask splits [0, itemCount) into
 * ranges of at least minRange items, every range runs exactly once on some
 * thread, and finishTask blocks until all of them have.  The engine calls
 * finishTask right after each enqueueTask, so the pool holds one job at a
 * time.
 *
 * Scheduling: one 64-bit word packs the job generation (high half) and the
 * next unclaimed chunk (low half).  A thread claims a chunk by compare-and-swap
 * on that word, so a thread still holding an old generation can never claim a
 * chunk of a newer job; a claimed chunk keeps its job alive (finishTask waits
 * for every chunk's completion count), so the job's parameters stay valid
 * while it runs.  Idle workers spin, then yield, then sleep on a condition
 * variable; the caller of finishTask works on chunks too.
 */
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>

typedef void m3TaskFn(int32_t startIndex, int32_t endIndex, void* taskContext);

typedef struct m3PoolJob
{
    m3TaskFn* fn;
    void* ctx;
    int32_t count;
    int32_t chunk;
    int32_t chunks;
} m3PoolJob;

typedef struct m3Pool
{
    _Atomic uint64_t word;   /* generation << 32 | next chunk */
    _Atomic int32_t done;    /* chunks finished in the current job */
    _Atomic int32_t stop;
    _Atomic int32_t sleepers;
    m3PoolJob jobs[2];       /* indexed by generation & 1 */
    int32_t threads;         /* including the thread that calls finishTask */
    pthread_mutex_t lock;
    pthread_cond_t wake;
    pthread_t* workers;
} m3Pool;

/* Claims and runs chunks of generation gen until none is left; returns the
 * number of chunks this thread ran. */
static int32_t RunChunks(m3Pool* pool, uint32_t gen)
{
    int32_t ran = 0;
    for (;;)
    {
        uint64_t v = atomic_load_explicit(&pool->word, memory_order_acquire);
        if ((uint32_t)(v >> 32) != gen)
            return ran;
        const m3PoolJob* job = &pool->jobs[gen & 1u];
        uint32_t idx = (uint32_t)v;
        if ((int32_t)idx >= job->chunks)
            return ran;
        if (!atomic_compare_exchange_weak_explicit(&pool->word, &v, v + 1, memory_order_acq_rel,
                                                   memory_order_acquire))
            continue;
        int32_t begin = (int32_t)idx * job->chunk;
        int32_t end = begin + job->chunk;
        if (end > job->count)
            end = job->count;
        job->fn(begin, end, job->ctx);
        atomic_fetch_add_explicit(&pool->done, 1, memory_order_release);
        ++ran;
    }
}

static void* Worker(void* arg)
{
    m3Pool* pool = (m3Pool*)arg;
    uint32_t seen = 0;
    while (!atomic_load_explicit(&pool->stop, memory_order_acquire))
    {
        uint32_t gen = (uint32_t)(atomic_load_explicit(&pool->word, memory_order_acquire) >> 32);
        if (gen != seen)
        {
            seen = gen;
            RunChunks(pool, gen);
            continue;
        }
        /* Idle: spin, then yield, then sleep until the next job. */
        int32_t spins = 0;
        while (spins < 4000 &&
               (uint32_t)(atomic_load_explicit(&pool->word, memory_order_acquire) >> 32) == seen &&
               !atomic_load_explicit(&pool->stop, memory_order_acquire))
        {
            if (spins > 2000)
                sched_yield();
            ++spins;
        }
        if (spins < 4000)
            continue;
        pthread_mutex_lock(&pool->lock);
        atomic_fetch_add(&pool->sleepers, 1);
        while ((uint32_t)(atomic_load(&pool->word) >> 32) == seen && !atomic_load(&pool->stop))
            pthread_cond_wait(&pool->wake, &pool->lock);
        atomic_fetch_sub(&pool->sleepers, 1);
        pthread_mutex_unlock(&pool->lock);
    }
    return NULL;
}

/* m3EnqueueTaskFn: publishes one job and returns the pool as its handle. */
void* m3pool_enqueue(m3TaskFn* task, int32_t itemCount, int32_t minRange, void* taskContext,
                     void* userContext)
{
    m3Pool* pool = (m3Pool*)userContext;
    uint32_t gen = (uint32_t)(atomic_load(&pool->word) >> 32) + 1u;
    m3PoolJob* job = &pool->jobs[gen & 1u];
    int32_t grain = minRange > 0 ? minRange : 1;
    /* About four chunks per thread, never below the engine's grain. */
    int32_t target = pool->threads * 4;
    int32_t chunk = (itemCount + target - 1) / target;
    if (chunk < grain)
        chunk = grain;
    job->fn = task;
    job->ctx = taskContext;
    job->count = itemCount;
    job->chunk = chunk;
    job->chunks = itemCount > 0 ? (itemCount + chunk - 1) / chunk : 0;
    atomic_store_explicit(&pool->done, 0, memory_order_relaxed);
    /* Sequentially consistent, paired with the sleepers count: either this
     * thread sees a worker that is about to sleep, or that worker sees the new
     * generation before it waits. */
    atomic_store(&pool->word, (uint64_t)gen << 32);
    if (atomic_load(&pool->sleepers) > 0)
    {
        pthread_mutex_lock(&pool->lock);
        pthread_cond_broadcast(&pool->wake);
        pthread_mutex_unlock(&pool->lock);
    }
    return pool;
}

/* m3FinishTaskFn: the caller works on the job, then waits for every chunk. */
void m3pool_finish(void* userTask, void* userContext)
{
    (void)userTask;
    m3Pool* pool = (m3Pool*)userContext;
    uint32_t gen = (uint32_t)(atomic_load(&pool->word) >> 32);
    RunChunks(pool, gen);
    const int32_t chunks = pool->jobs[gen & 1u].chunks;
    while (atomic_load_explicit(&pool->done, memory_order_acquire) < chunks)
        ;
}

/* A pool of `threads` threads in all: threads - 1 workers plus the caller of
 * finishTask.  Returns NULL on failure. */
void* m3pool_create(int32_t threads)
{
    if (threads < 1)
        return NULL;
    m3Pool* pool = (m3Pool*)calloc(1, sizeof(m3Pool));
    if (pool == NULL)
        return NULL;
    pool->threads = threads;
    pthread_mutex_init(&pool->lock, NULL);
    pthread_cond_init(&pool->wake, NULL);
    pool->workers = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    if (pool->workers == NULL)
    {
        free(pool);
        return NULL;
    }
    for (int32_t i = 0; i + 1 < threads; ++i)
    {
        if (pthread_create(&pool->workers[i], NULL, Worker, pool) != 0)
        {
            pool->threads = i + 1;
            break;
        }
    }
    return pool;
}

void m3pool_destroy(void* handle)
{
    m3Pool* pool = (m3Pool*)handle;
    if (pool == NULL)
        return;
    atomic_store(&pool->stop, 1);
    pthread_mutex_lock(&pool->lock);
    pthread_cond_broadcast(&pool->wake);
    pthread_mutex_unlock(&pool->lock);
    for (int32_t i = 0; i + 1 < pool->threads; ++i)
        pthread_join(pool->workers[i], NULL);
    pthread_cond_destroy(&pool->wake);
    pthread_mutex_destroy(&pool->lock);
    free(pool->workers);
    free(pool);
}

/* Pointers to the two hooks, for callers (ctypes) that fill an m3WorldDef. */
void* m3pool_enqueue_fn(void) { return (void*)m3pool_enqueue; }
void* m3pool_finish_fn(void) { return (void*)m3pool_finish; }
