/*
 * par.c -- a parallel for over [0, n) on POSIX threads with a shared work
 * counter (dynamic schedule), for timing the oracle on every host core.
 *
 * TEST INFRASTRUCTURE ONLY.  Plain pthreads rather than OpenMP: the oracle
 * is loaded into processes that already hold torch's own OpenMP runtime.
 * Results never depend on the thread count: every index is one independent
 * task whose arithmetic is the same in any order.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>

typedef struct {
    opar_fn fn;
    void* ctx;
    uint32_t n, chunk;
    atomic_uint next;
} opar_job;

typedef struct {
    opar_job* job;
    int tid;
} opar_arg;

static void* opar_worker(void* p) {
    opar_arg* a = (opar_arg*)p;
    opar_job* j = a->job;
    for (;;) {
        uint32_t b = atomic_fetch_add(&j->next, j->chunk);
        if (b >= j->n) break;
        uint32_t e = j->n - b < j->chunk ? j->n : b + j->chunk;
        for (uint32_t i = b; i < e; ++i) j->fn(j->ctx, i, a->tid);
    }
    return 0;
}

int oracle_parallel_for(uint32_t n, uint32_t chunk, int threads, opar_fn fn, void* ctx) {
    if (threads > OPAR_MAX_THREADS) threads = OPAR_MAX_THREADS;
    if (chunk == 0) chunk = 1;
    if (threads <= 1 || n <= chunk) {
        for (uint32_t i = 0; i < n; ++i) fn(ctx, i, 0);
        return 1;
    }
    opar_job job;
    job.fn = fn;
    job.ctx = ctx;
    job.n = n;
    job.chunk = chunk;
    atomic_init(&job.next, 0u);
    pthread_t th[OPAR_MAX_THREADS];
    opar_arg args[OPAR_MAX_THREADS];
    int started = 0;
    for (int t = 1; t < threads; ++t) {
        args[t].job = &job;
        args[t].tid = t;
        if (pthread_create(&th[t], 0, opar_worker, &args[t]) != 0) break;
        started = t;
    }
    args[0].job = &job;
    args[0].tid = 0;
    opar_worker(&args[0]);
    for (int t = 1; t <= started; ++t) pthread_join(th[t], 0);
    return started + 1;
}
