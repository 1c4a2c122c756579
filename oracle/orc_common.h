/*
 * orc_common.h — shared helpers of the CPU oracle (TEST INFRASTRUCTURE).
 *
 * The oracle is a plain-C restatement of Hammerhead.jl's WIDIM correlation
 * path, used only by tests/, the smoke check and the CPU-baseline leg of
 * bench.py as the checker. It is never linked into, loaded by or called from
 * the product library. See orc_impl.h for the per-function citations.
 */
#ifndef ORC_COMMON_H
#define ORC_COMMON_H

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/hammerhead_mi355x.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)

/* UQ_OFFSETS (src/uncertainty.jl:23-26): the half-plane offsets (dr, dc),
 * dc = 0..4, dr = 1..4 for dc == 0 else -4..4, stably sorted by ring
 * max(|dr|, |dc|). Filled by orc_uq_offsets(). */
#define ORC_UQ_MAXOFF 4
#define ORC_UQ_NOFF 40

/* Simple static-partition parallel for over [0, n). */
typedef void (*orc_range_fn)(void *arg, int begin, int end);

typedef struct {
    orc_range_fn fn;
    void *arg;
    int begin, end;
} orc_task;

static void *orc_task_run(void *p)
{
    orc_task *t = (orc_task *)p;
    if (t->begin < t->end) t->fn(t->arg, t->begin, t->end);
    return NULL;
}

static void orc_parallel_for(int n, int nthreads, orc_range_fn fn, void *arg)
{
    if (nthreads <= 1 || n < 2) {
        fn(arg, 0, n);
        return;
    }
    if (nthreads > n) nthreads = n;
    pthread_t th[256];
    orc_task tk[256];
    if (nthreads > 256) nthreads = 256;
    int chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        tk[t].fn = fn;
        tk[t].arg = arg;
        tk[t].begin = t * chunk;
        tk[t].end = (t + 1) * chunk < n ? (t + 1) * chunk : n;
        pthread_create(&th[t], NULL, orc_task_run, &tk[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

static int orc_uq_offsets(int off[ORC_UQ_NOFF][2], int ringstart[ORC_UQ_MAXOFF + 1])
{
    int tmp[ORC_UQ_NOFF][2];
    int n = 0;
    for (int dc = 0; dc <= ORC_UQ_MAXOFF; dc++) {
        int lo = dc == 0 ? 1 : -ORC_UQ_MAXOFF;
        for (int dr = lo; dr <= ORC_UQ_MAXOFF; dr++) {
            tmp[n][0] = dr;
            tmp[n][1] = dc;
            n++;
        }
    }
    /* stable selection by ring */
    int k = 0;
    for (int ring = 1; ring <= ORC_UQ_MAXOFF; ring++) {
        ringstart[ring - 1] = k;
        for (int i = 0; i < n; i++) {
            int a = abs(tmp[i][0]), b = abs(tmp[i][1]);
            if ((a > b ? a : b) == ring) {
                off[k][0] = tmp[i][0];
                off[k][1] = tmp[i][1];
                k++;
            }
        }
    }
    ringstart[ORC_UQ_MAXOFF] = k;
    return k;
}

/* Grid starts of pass_grid (src/pipeline.jl:529-555): Julia range
 * (1+m):(w-ov):(n-w-m+1), 1-based. Returns the count. */
static int orc_range_count(int start, int step, int stop)
{
    if (step <= 0 || stop < start) return 0;
    return (stop - start) / step + 1;
}

static int orc_cmpd(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* A fixed pseudo-random value in [-1, 1) for index k (splitmix64 finalizer):
 * the deterministic perturbation of the test diagnostics. */
static double orc_hash_unit(size_t k)
{
    uint64_t z = (uint64_t)k * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

#endif
