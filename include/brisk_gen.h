/* brisk_gen.h - seeded generators of the SDP instances the GPU solver is measured and tested on.
 *
 * These have no counterpart in the reference's library: the reference generates its test
 * problems in scripts (tools/timelimit_test.sh:12-24 for the theta SDP of a random graph,
 * python/tests/test_brisk.py:26-40 for max-cut). They return the numbers of an SDPA sparse
 * file in exactly the layout brisk_solve_data takes (capi.c:175-192): 1-based, mat 0 = F0
 * (C = -F0), mat 1..m = A_i, c = b, upper-triangle entries, LP blocks negative in bs.
 * Every random number is a pure function of (seed, stream, counter), so an instance is the
 * same on every machine. */
#ifndef BRISK_GEN_H
#define BRISK_GEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int m, nblk;
    int *bs;            /* block sizes, LP blocks negative */
    double *c;          /* b (m) */
    int64_t nnz;
    int *mat, *blk, *i, *j;
    double *v;
} BriskInstance;

/* Lovasz theta of G(n, p): edges i<j in row-major order, each with probability p */
int brisk_gen_theta(int n, double p, uint64_t seed, BriskInstance *out);
/* theta of an explicit graph: ne edges as 0-based pairs */
int brisk_gen_theta_edges(int n, int64_t ne, const int *edges, BriskInstance *out);
/* max-cut SDP of G(n, p) with weights uniform in {-1, +1}: C = -L/4, diag X = 1 */
int brisk_gen_maxcut(int n, double p, uint64_t seed, BriskInstance *out);
/* max-cut SDP of the cycle C_n with unit weights */
int brisk_gen_maxcut_cycle(int n, BriskInstance *out);
/* dense random SDP: one n x n block, m dense A_i ~ N(0, 1/n), feasible from X0, Z0, y0 */
int brisk_gen_dense(int n, int m, uint64_t seed, BriskInstance *out);
/* multi-block sparse SDP: nb blocks of order bn plus an LP block of nlp (0: none), m rows */
int brisk_gen_multi(int nb, int bn, int nlp, int m, uint64_t seed, BriskInstance *out);
void brisk_gen_free(BriskInstance *inst);

/* the generators' random numbers, for tests and stream generators in other languages */
double brisk_gen_u01(uint64_t seed, uint64_t stream, uint64_t i);
double brisk_gen_gauss(uint64_t seed, uint64_t stream, uint64_t i);

#ifdef __cplusplus
}
#endif
#endif
