/* probe.c - kernel-level entry points for the parity tests (include/brisk_probe.h).
 *
 * Each call builds the problem exactly as a solve does (hprob_build: scaling, constraint lists,
 * groups; solve_setup: the device heap layout), places caller data in the solve's arrays, runs
 * one kernel family of the iteration (ipm_probe_*) and returns the result on the data as read:
 * the Schur complement M_ij = <A_i, X A_j Z^-1>, the operators A(D) and A*(y), the Lanczos
 * bound of the step-length test, the DIMACS errors, the end phase's metric corrections. The solver scales A_i by d_i = 1/||A_i||_F,
 * X by 1/bs, y_i by 1/(cs d_i) (problem.c); the probes apply and undo those factors.
 *
 * Block data is packed as the result accessors return it: per input block, n x n column-major
 * (SDP) or n numbers (LP); internally every LP variable is a 1 x 1 block. */
#include <stdlib.h>
#include <string.h>
#include "brisk_amd.h"
#include "brisk_probe.h"
#include "solver.h"

#define API __attribute__((visibility("default")))

int brisk_heap_ensure(size_t need);
int ipm_probe_schur(Solve *s, int prec);
int ipm_probe_aop(Solve *s);
int ipm_probe_aty(Solve *s);
int ipm_probe_lammin(Solve *s, int mode, double *out);
int ipm_probe_dimacs(Solve *s);
int ipm_probe_correct(Solve *s, int half, double *steps);

static int64_t packed_len(const HProb *P) {
    int64_t t = 0;
    for (int k = 0; k < P->nblk_in; k++) t += P->bs_in[k] > 0 ? (int64_t)P->bs_in[k] * P->bs_in[k] : -P->bs_in[k];
    return t;
}
/* packed input blocks -> block arena (times f), and back */
static void to_arena(const HProb *P, const double *in, double f, double *ar) {
    int64_t o = 0;
    memset(ar, 0, sizeof(double) * (size_t)P->T);
    for (int k = 0; k < P->nblk_in; k++) {
        const int kb = P->in_first[k];
        if (P->bs_in[k] > 0) {
            const int64_t len = (int64_t)P->bs_in[k] * P->bs_in[k];
            for (int64_t q = 0; q < len; q++) ar[P->boff[kb] + q] = f * in[o + q];
            o += len;
        } else {
            for (int i = 0; i < -P->bs_in[k]; i++) ar[P->boff[kb + i]] = f * in[o + i];
            o += -P->bs_in[k];
        }
    }
}
static void from_arena(const HProb *P, const double *ar, double *out) {
    int64_t o = 0;
    for (int k = 0; k < P->nblk_in; k++) {
        const int kb = P->in_first[k];
        if (P->bs_in[k] > 0) {
            const int64_t len = (int64_t)P->bs_in[k] * P->bs_in[k];
            memcpy(out + o, ar + P->boff[kb], sizeof(double) * len);
            o += len;
        } else {
            for (int i = 0; i < -P->bs_in[k]; i++) out[o + i] = ar[P->boff[kb + i]];
            o += -P->bs_in[k];
        }
    }
}

/* the solve of a probe: built and set up on the device (0), or an error code */
static int probe_open(Solve *s, const BriskProblem *p) {
    memset(s, 0, sizeof(*s));
    s->o.tol = 1e-8; s->o.maxit = 100; s->o.mixed_until = 1e-6; s->o.apply = -1;
    s->o.prec = PREC_F64;
    if (!p) return -1;
    const int e = hprob_build(&s->P, p->m, p->nblk, p->bs, p->c, p->nnz, p->mat, p->blk, p->i, p->j, p->v);
    if (e) { hprob_free(&s->P); return e; }
    if (bk_init() || brisk_heap_ensure((size_t)solve_heap_bytes(&s->P) + ((size_t)64 << 20)) || solve_setup(s)) {
        hprob_free(&s->P);
        return -2;
    }
    return 0;
}
static void probe_close(Solve *s) {
    solve_release(s);
    ipm_release_scratch();
    hprob_free(&s->P);
    free(s->res_X); free(s->res_Z); free(s->res_y);
}
static int upload_blocks(Solve *s, int64_t off, const double *packed, double f) {
    double *ar = malloc(sizeof(double) * (size_t)(s->P.T + 1));
    to_arena(&s->P, packed, f, ar);
    const int rc = bk_h2d((double *)bk_heap() + off, ar, sizeof(double) * (size_t)s->P.T);
    free(ar);
    return rc;
}

API int brisk_probe_schur(const BriskProblem *p, const double *X, const double *Z, int prec, double *M) {
    Solve s;
    int rc = probe_open(&s, p);
    if (rc) return rc;
    const int m = s.P.m;
    if (upload_blocks(&s, s.X, X, 1.0) || upload_blocks(&s, s.Z, Z, 1.0) || ipm_probe_schur(&s, prec) ||
        bk_d2h(M, (double *)bk_heap() + s.M, sizeof(double) * (size_t)m * m)) rc = -4;
    else
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) M[i + (size_t)j * m] /= s.P.d[i] * s.P.d[j];
    probe_close(&s);
    return rc;
}

API int brisk_probe_aop(const BriskProblem *p, const double *D, double *out) {
    Solve s;
    int rc = probe_open(&s, p);
    if (rc) return rc;
    if (upload_blocks(&s, s.W1, D, 1.0) || ipm_probe_aop(&s) || bk_d2h(out, (double *)bk_heap() + s.rp, sizeof(double) * s.P.m)) rc = -4;
    else
        for (int i = 0; i < s.P.m; i++) out[i] /= s.P.d[i];
    probe_close(&s);
    return rc;
}

API int brisk_probe_aty(const BriskProblem *p, const double *y, double *out) {
    Solve s;
    int rc = probe_open(&s, p);
    if (rc) return rc;
    double *ys = malloc(sizeof(double) * (s.P.m + 1)), *ar = malloc(sizeof(double) * (size_t)(s.P.T + 1));
    for (int i = 0; i < s.P.m; i++) ys[i] = y[i] / s.P.d[i];
    if (bk_h2d((double *)bk_heap() + s.y, ys, sizeof(double) * s.P.m) || ipm_probe_aty(&s) ||
        bk_d2h(ar, (double *)bk_heap() + s.W1, sizeof(double) * (size_t)s.P.T)) rc = -4;
    else
        from_arena(&s.P, ar, out);
    free(ys); free(ar);
    probe_close(&s);
    return rc;
}

API int brisk_probe_lammin(const BriskProblem *p, const double *X, const double *dX, int mode, double *out) {
    Solve s;
    int rc = probe_open(&s, p);
    if (rc) return rc;
    if (upload_blocks(&s, s.X, X, 1.0) || (mode && upload_blocks(&s, s.dX, dX, 1.0)) || ipm_probe_lammin(&s, mode, out)) rc = -4;
    probe_close(&s);
    return rc;
}

API int brisk_probe_dimacs(const BriskProblem *p, const double *X, const double *y, double *err, double *obj) {
    Solve s;
    int rc = probe_open(&s, p);
    if (rc) return rc;
    double *ys = malloc(sizeof(double) * (s.P.m + 1));
    for (int i = 0; i < s.P.m; i++) ys[i] = y[i] / (s.P.cs * s.P.d[i]);
    if (upload_blocks(&s, s.X, X, 1.0 / s.P.bs) || bk_h2d((double *)bk_heap() + s.y, ys, sizeof(double) * s.P.m) || ipm_probe_dimacs(&s)) rc = -4;
    else {
        for (int e = 0; e < 6; e++) err[e] = s.res_err[e];
        obj[0] = s.res_pobj; obj[1] = s.res_dobj;
    }
    free(ys);
    probe_close(&s);
    return rc;
}

API int brisk_probe_correct(const BriskProblem *p, const double *X, const double *y, int half, double *Xout, double *steps) {
    Solve s;
    int rc = probe_open(&s, p);
    if (rc) return rc;
    double *ys = malloc(sizeof(double) * (s.P.m + 1)), *ar = malloc(sizeof(double) * (size_t)(s.P.T + 1));
    for (int i = 0; i < s.P.m; i++) ys[i] = y[i] / (s.P.cs * s.P.d[i]);
    if (upload_blocks(&s, s.X, X, 1.0 / s.P.bs) || bk_h2d((double *)bk_heap() + s.y, ys, sizeof(double) * s.P.m) || ipm_probe_correct(&s, half, steps) ||
        bk_d2h(ar, (double *)bk_heap() + s.X, sizeof(double) * (size_t)s.P.T)) rc = -4;
    else {
        for (int64_t q = 0; q < s.P.T; q++) ar[q] *= s.P.bs;
        from_arena(&s.P, ar, Xout);
    }
    free(ys); free(ar);
    probe_close(&s);
    return rc;
}

API int64_t brisk_probe_packed_len(const BriskProblem *p) {
    int64_t t = 0;
    if (!p) return -1;
    for (int k = 0; k < p->nblk; k++) t += p->bs[k] > 0 ? (int64_t)p->bs[k] * p->bs[k] : -p->bs[k];
    return t;
}
