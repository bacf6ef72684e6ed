/* par.c - the host parallel for of par.h (pthreads, one thread per part, joined at the end) */
#include <pthread.h>
#include <stdlib.h>
#include <unistd.h>
#include "par.h"

#define PAR_MAX 16

int par_parts(int64_t n, int64_t grain) {
    long nc = sysconf(_SC_NPROCESSORS_ONLN);
    if (nc < 1) nc = 1;
    if (nc > PAR_MAX) nc = PAR_MAX;
    if (grain < 1) grain = 1;
    int64_t p = n / grain;
    if (p > nc) p = nc;
    return p < 1 ? 1 : (int)p;
}

typedef struct { par_fn fn; void *ctx; int64_t lo, hi; int part; } ParJob;
static void *par_run(void *a) {
    ParJob *j = a;
    j->fn(j->ctx, j->lo, j->hi, j->part);
    return NULL;
}

void par_for(int64_t n, int64_t grain, par_fn fn, void *ctx) {
    if (n <= 0) return;
    const int P = par_parts(n, grain);
    if (P == 1) { fn(ctx, 0, n, 0); return; }
    ParJob job[PAR_MAX];
    pthread_t th[PAR_MAX];
    int started[PAR_MAX];
    for (int p = 0; p < P; p++) {
        job[p].fn = fn; job[p].ctx = ctx; job[p].part = p;
        job[p].lo = n * p / P; job[p].hi = n * (p + 1) / P;
        started[p] = p > 0 && pthread_create(&th[p], NULL, par_run, &job[p]) == 0;
    }
    par_run(&job[0]);
    for (int p = 1; p < P; p++) {
        if (started[p]) pthread_join(th[p], NULL);
        else par_run(&job[p]);              /* a thread that could not start: its range here */
    }
}
