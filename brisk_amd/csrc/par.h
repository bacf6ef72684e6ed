/* par.h - a parallel for over index ranges on the host cores (pthreads), for the setup passes
 * over a problem's entries and dense constraint matrices. fn(ctx, lo, hi) runs on [lo, hi) for
 * consecutive ranges of at least `grain` indices, on up to 16 threads; the calls of one
 * par_for share no ranges, so a function that writes only its own range needs no locks. */
#ifndef BRISK_PAR_H
#define BRISK_PAR_H
#include <stdint.h>
typedef void (*par_fn)(void *ctx, int64_t lo, int64_t hi, int part);
/* the number of parts par_for splits n indices into (1 when n < 2 grain) */
int par_parts(int64_t n, int64_t grain);
/* runs fn over [0, n) in par_parts(n, grain) parts; part p covers [p n / P, (p + 1) n / P) */
void par_for(int64_t n, int64_t grain, par_fn fn, void *ctx);
#endif
