/* capi.c - the library's C ABI (include/brisk_amd.h): the reference's flat interface
 * (MilanKorda/brisk capi.c) on top of the GPU solver, plus the batch entry.
 *
 * Exit codes and status strings follow the reference (capi.c:171-174, main.c:26-28); a bad
 * option returns 1, unreadable data 2, a GPU failure < 0 with an empty result. */
#include <math.h>
#include <pthread.h>
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "brisk_amd.h"
#include "solver.h"

#define API __attribute__((visibility("default")))

struct BriskResult {
    int status;
    char status_str[40];
    int exit_code, iters, have_x, m, nblk;
    int *bs;
    double pobj, dobj, err[6], time;
    double *y, *Xp, *Zp;
    double **X, **Z;
    double phases[4];
    int counts[3];
    int ntrace;
    double tr[3][128];
    char cause[160];
    int n_resolves;             /* re-solves (none here: one method per solve) and their time */
    double t_resolves;
};

static const char *g_status_str[] = { "OPTIMAL", "PRIMAL INFEASIBLE", "DUAL INFEASIBLE", "ITERATION LIMIT",
                                      "NUMERICAL DIFFICULTIES", "SOLVED TO REDUCED ACCURACY", "TIME LIMIT" };
static const int g_exit_of[] = { 0, 11, 12, 13, 14, 10, 15 };

static int g_out_mode = 0;
static int (*g_out_cb)(const char *, int) = NULL;
static volatile sig_atomic_t g_stop = 0;
static int g_stop_owner = 0;

API void brisk_set_output(int mode, int (*cb)(const char *, int)) { g_out_mode = mode; g_out_cb = cb; }
API const char *brisk_version(void) { return "1.7-mi355x"; }
API void brisk_interrupt(void) { g_stop = 1; }
API int brisk_interrupted(void) { return g_stop != 0; }
API void brisk_clear_interrupt(void) { g_stop = 0; }
API void brisk_set_interrupt_owner(int caller) { g_stop_owner = caller != 0; }
int brisk_stop_requested(void) { return g_stop != 0; }
/* read at every call, so that a test can set and clear BRISK_FAULT within one process */
int brisk_fault(const char *name) {
    const char *f = getenv("BRISK_FAULT");
    return f && strstr(f, name) != NULL;
}

static void say(int is_err, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
#include <stdarg.h>
static void say(int is_err, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (g_out_mode == 2 && g_out_cb) { g_out_cb(buf, is_err); return; }
    if (g_out_mode == 1 && !is_err) return;
    fputs(buf, is_err ? stderr : stdout);
}

static void result_clear(BriskResult *r) {
    free(r->bs); free(r->y); free(r->Xp); free(r->Zp); free(r->X); free(r->Z);
    memset(r, 0, sizeof(*r));
    r->status = -1;
}
/* a result that was not solved: status -1, exit code -1 (the reference's code for an aborted
 * call, capi.c:171-174), and its cause */
void brisk_result_unsolved(BriskResult *r, const char *cause) {
    if (!r) return;
    result_clear(r);
    r->exit_code = -1;
    snprintf(r->cause, sizeof r->cause, "%s", cause ? cause : "");
}
void brisk_result_move(BriskResult *dst, BriskResult *src) {
    result_clear(dst);
    *dst = *src;
    memset(src, 0, sizeof(*src));
    src->status = -1;
}
API BriskResult *brisk_result_new(void) {
    BriskResult *r = calloc(1, sizeof(BriskResult));
    if (r) r->status = -1;
    return r;
}
API void brisk_result_delete(BriskResult *r) { if (r) { result_clear(r); free(r); } }
API int brisk_result_status(const BriskResult *r) { return r ? r->status : -1; }
API const char *brisk_result_status_str(const BriskResult *r) { return r && r->status >= 0 ? r->status_str : "NOT SOLVED"; }
API int brisk_result_exit_code(const BriskResult *r) { return r ? r->exit_code : -1; }
API int brisk_result_iterations(const BriskResult *r) { return r ? r->iters : 0; }
API int brisk_result_m(const BriskResult *r) { return r ? r->m : 0; }
API int brisk_result_nblk(const BriskResult *r) { return r ? r->nblk : 0; }
API int brisk_result_blocksize(const BriskResult *r, int k) { return r && r->bs && k >= 0 && k < r->nblk ? r->bs[k] : 0; }
API double brisk_result_pobj(const BriskResult *r) { return r ? r->pobj : 0; }
API double brisk_result_dobj(const BriskResult *r) { return r ? r->dobj : 0; }
API double brisk_result_time(const BriskResult *r) { return r ? r->time : 0; }
API int brisk_result_have_x(const BriskResult *r) { return r ? r->have_x : 0; }
API void brisk_result_dimacs(const BriskResult *r, double *err) { for (int e = 0; e < 6; e++) err[e] = r ? r->err[e] : 0; }
API const double *brisk_result_y(const BriskResult *r) { return r ? r->y : NULL; }
API const double *brisk_result_X(const BriskResult *r, int k) { return r && r->have_x && r->X && k >= 0 && k < r->nblk ? r->X[k] : NULL; }
API const double *brisk_result_Z(const BriskResult *r, int k) { return r && r->Z && k >= 0 && k < r->nblk ? r->Z[k] : NULL; }
API const char *brisk_result_cause(const BriskResult *r) { return r ? r->cause : ""; }
/* re-solves of a result (reference capi.c:121: the method retries and fallbacks, brisk.h:122):
 * the GPU path runs one method per solve and retries none, so the count and its time are 0 for
 * every result, solved or not */
API int brisk_result_resolves(const BriskResult *r, double *t) { if (t) *t = r ? r->t_resolves : 0; return r ? r->n_resolves : 0; }
API void brisk_result_phases(const BriskResult *r, double *out) { for (int k = 0; k < 4; k++) out[k] = r ? r->phases[k] : 0; }
API void brisk_result_counts(const BriskResult *r, int *out) { for (int k = 0; k < 3; k++) out[k] = r ? r->counts[k] : 0; }
API int brisk_result_trace(const BriskResult *r, double *po, double *dob, double *mu, int n) {
    if (!r) return 0;
    const int c = r->ntrace < n ? r->ntrace : n;
    for (int k = 0; k < c; k++) { if (po) po[k] = r->tr[0][k]; if (dob) dob[k] = r->tr[1][k]; if (mu) mu[k] = r->tr[2][k]; }
    return c;
}

/* ---- options ---- */
int brisk_parse_options(int nopt, const char *const *opts, SolveOpts *o, int m, int *slots, double *heapgb) {
    o->tol = 1e-8; o->maxit = 100; o->verbose = 0; o->mixed_until = 1e-6; o->time_limit = 0; o->hsd = -1; o->hsd_first = 0; o->hsd_first_c = 12.0; o->hsd_first_asm = 0.5; o->dir = 0; o->apply = -1; o->tput = 0; o->polish = 1; o->polish_r = 1; o->polish_x = 1;
    o->prec = m >= 1000 ? PREC_MIXED : PREC_F64;
    for (int k = 0; k < nopt; k++) {
        const char *a = opts[k];
        if (!a) return 1;
        const char *nx = k + 1 < nopt ? opts[k + 1] : NULL;
#define NEED() do { if (!nx) { say(1, "brisk: option %s needs a value\n", a); return 1; } k++; } while (0)
        if (!strcmp(a, "-prec")) {
            NEED();
            if (!strcmp(nx, "double") || !strcmp(nx, "fp64")) o->prec = PREC_F64;
            else if (!strcmp(nx, "mixed")) o->prec = PREC_MIXED;
            else { say(1, "brisk: -prec takes double or mixed\n"); return 1; }
        } else if (!strcmp(a, "-tol")) { NEED(); o->tol = atof(nx); if (!(o->tol > 0)) return 1; }
        else if (!strcmp(a, "-maxit")) { NEED(); o->maxit = atoi(nx); if (o->maxit < 0) return 1; }
        else if (!strcmp(a, "-timelimit")) { NEED(); o->time_limit = atof(nx); }
        else if (!strcmp(a, "-mixeduntil")) { NEED(); o->mixed_until = atof(nx); }
        else if (!strcmp(a, "-slots")) { NEED(); if (slots) *slots = atoi(nx); }
        else if (!strcmp(a, "-heapgb")) { NEED(); if (heapgb) *heapgb = atof(nx); }
        else if (!strcmp(a, "-hsd")) {
            /* the reference's flag (-hsd: the embedding), or -hsd 0 | 1 */
            if (nx && (!strcmp(nx, "0") || !strcmp(nx, "1"))) { o->hsd = atoi(nx); k++; } else o->hsd = 1;
        }
        else if (!strcmp(a, "-nohsd")) o->hsd = 0;
        else if (!strcmp(a, "-dir")) {
            NEED();
            if (!strcmp(nx, "nt")) o->dir = 1;
            else if (!strcmp(nx, "hkm")) o->dir = 0;
            else if (!strcmp(nx, "auto")) o->dir = -1;
            else { say(1, "brisk: -dir must be auto, hkm or nt\n"); return 1; }
        }
        else if (!strcmp(a, "-hsdfirst")) { NEED(); o->hsd_first = atoi(nx) > 0; }
        else if (!strcmp(a, "-hsdfirstc")) { NEED(); o->hsd_first_c = atof(nx); }
        else if (!strcmp(a, "-hsdfirstasm")) { NEED(); o->hsd_first_asm = atof(nx); }
        else if (!strcmp(a, "-nopolish")) o->polish = 0;
        else if (!strcmp(a, "-nopolishr")) o->polish_r = 0;
        else if (!strcmp(a, "-nopolishx")) o->polish_x = 0;
        else if (!strcmp(a, "-throughput")) { NEED(); o->tput = atoi(nx) > 0; }
        else if (!strcmp(a, "-schurapply")) {
            NEED();
            if (!strcmp(nx, "auto")) o->apply = -1;
            else if (!strcmp(nx, "inverse")) o->apply = 0;
            else if (!strcmp(nx, "solves")) o->apply = 1;
            else { say(1, "brisk: -schurapply takes auto, inverse or solves\n"); return 1; }
        }
        else if (!strcmp(a, "-threads") || !strcmp(a, "-bm") || !strcmp(a, "-gpu")) { NEED(); }
        else if (!strcmp(a, "-v")) o->verbose = 1;
        else if (!strcmp(a, "-q")) o->verbose = 0;
        else { say(1, "brisk: unknown option %s\n", a); return 1; }
#undef NEED
    }
    return 0;
}

/* device heap: sized for the work at hand (grown only when nothing lives in it) */
int brisk_heap_ensure(size_t need) {
    const size_t cur = bk_heap_size();
    if (cur >= need && heap_largest_free() >= need) return 0;
    ipm_release_scratch();
    size_t want = need + (need >> 3) + ((size_t)64 << 20);
    const size_t freeb = bk_mem_free() + cur;
    if (want > freeb) want = freeb - ((size_t)512 << 20);
    if (want < need) return -1;
    return heap_setup(want);
}

void brisk_fill_result(BriskResult *r, Solve *s, double t_setup, double t_iter, double t_dimacs, double t_down) {
    result_clear(r);
    const HProb *P = &s->P;
    r->status = s->status;
    snprintf(r->status_str, sizeof r->status_str, "%s", s->status >= 0 && s->status <= 6 ? g_status_str[s->status] : "NOT SOLVED");
    r->exit_code = s->status >= 0 && s->status <= 6 ? g_exit_of[s->status] : -1;
    r->iters = s->it;
    r->m = P->m; r->nblk = P->nblk_in;
    r->bs = malloc(sizeof(int) * P->nblk_in);
    memcpy(r->bs, P->bs_in, sizeof(int) * P->nblk_in);
    r->pobj = s->res_pobj; r->dobj = s->res_dobj;
    for (int e = 0; e < 6; e++) r->err[e] = s->res_err[e];
    r->phases[0] = t_setup; r->phases[1] = t_iter; r->phases[2] = t_dimacs; r->phases[3] = t_down;
    r->time = t_setup + t_iter + t_dimacs + t_down;
    r->counts[0] = s->nfactor_f32; r->counts[1] = s->nfactor_f64; r->counts[2] = s->pcg_total;
    r->ntrace = s->it + 1 < 128 ? s->it + 1 : 128;
    for (int k = 0; k < r->ntrace; k++) { r->tr[0][k] = s->tr_pobj[k]; r->tr[1][k] = s->tr_dobj[k]; r->tr[2][k] = s->tr_mu[k]; }
    r->y = s->res_y; s->res_y = NULL;
    r->Xp = s->res_X; s->res_X = NULL;
    r->Zp = s->res_Z; s->res_Z = NULL;
    r->have_x = r->Xp != NULL;
    r->X = malloc(sizeof(double *) * P->nblk_in);
    r->Z = malloc(sizeof(double *) * P->nblk_in);
    int64_t o = 0;
    for (int k = 0; k < P->nblk_in; k++) {
        r->X[k] = r->Xp ? r->Xp + o : NULL;
        r->Z[k] = r->Zp ? r->Zp + o : NULL;
        o += P->bs_in[k] > 0 ? (int64_t)P->bs_in[k] * P->bs_in[k] : -P->bs_in[k];
    }
    if (r->status != 0)
        snprintf(r->cause, sizeof r->cause, "DIMACS error %.1e after %d iterations", acc_err(r->err), r->iters);
}

static int solve_many(int n, const BriskProblem *pr, int nopt, const char *const *opts, BriskResult **res) {
    if (n <= 0) return 0;
    if (!g_stop_owner) g_stop = 0;
    const double t0 = bk_now();
    Solve *S = calloc(n, sizeof(Solve));
    Solve **SP = malloc(sizeof(Solve *) * n);
    int rc = 0, built = 0, setup = 0;
    size_t need = 0;
    for (int q = 0; q < n; q++) {
        for (int k = 0; k < 1; k++) if (res[q]) result_clear(res[q]);
        int mq = pr[q].m;
        if (brisk_parse_options(nopt, opts, &S[q].o, mq, NULL, NULL)) { rc = 1; goto out; }
    }
    for (int q = 0; q < n; q++) {
        const BriskProblem *p = &pr[q];
        const int e = hprob_build(&S[q].P, p->m, p->nblk, p->bs, p->c, p->nnz, p->mat, p->blk, p->i, p->j, p->v);
        built = q + 1;
        if (e) { rc = e; goto out; }
        if (S[q].o.hsd < 0) S[q].o.hsd = hsd_first_rule(&S[q].P, &S[q].o);
        S[q].P.want_hsd = S[q].o.hsd;
        if (S[q].o.dir < 0) S[q].o.dir = nt_rule(&S[q].P);
        S[q].P.want_nt = !S[q].o.hsd && S[q].o.dir == 1;
        need += (size_t)solve_heap_bytes(&S[q].P);
        SP[q] = &S[q];
        S[q].id = q;
    }
    const double tb = bk_now();
    if (bk_init()) { say(1, "brisk: no usable GPU: %s\n", bk_error()); rc = -3; goto out; }
    if (brisk_heap_ensure(need + ((size_t)8 << 20))) { say(1, "brisk: device memory: %zu bytes needed\n", need); rc = -2; goto out; }
    const double th = bk_now();
    for (int q = 0; q < n; q++) {
        if (solve_setup(&S[q])) { say(1, "brisk: device setup failed: %s\n", bk_error()); rc = -2; goto out; }
        setup = q + 1;
    }
    bk_sync();
    const double t1 = bk_now();
    if (S[0].o.verbose)
        say(0, "setup: problem build %.4f s, device and heap %.4f s, upload %.4f s\n", tb - t0, th - tb, t1 - th);
    if (ipm_start_batch(SP, n)) { rc = -4; goto out; }
    for (int guard = 0; guard < 10000; guard++) {
        int any = 0;
        for (int q = 0; q < n; q++) any |= !S[q].done;
        if (!any) break;
        if (g_stop) for (int q = 0; q < n; q++) if (!S[q].done) { S[q].done = 1; S[q].status = 6; }
        if (ipm_iterate_batch(SP, n)) { rc = -4; goto out; }
    }
    bk_sync();
    const double t2 = bk_now();
    {
        int *st0 = malloc(sizeof(int) * n);
        for (int q = 0; q < n; q++) st0[q] = S[q].status;
        const int e = ipm_dimacs(SP, n) || ipm_end_corrections(SP, n, st0) || ipm_polish(SP, n, st0);
        free(st0);
        if (e) { rc = -4; goto out; }
    }
    bk_sync();
    const double t3 = bk_now();
    if (ipm_finish_batch(SP, n)) { rc = -4; goto out; }
    const double t4 = bk_now();
    for (int q = 0; q < n; q++) if (res[q]) brisk_fill_result(res[q], &S[q], t1 - t0, t2 - t1, t3 - t2, (t4 - t3) / n);
    if (bk_failed()) rc = -4;
out:
    if (rc < 0 && bk_failed()) say(1, "brisk: GPU error: %s\n", bk_error());
    for (int q = 0; q < setup; q++) solve_release(&S[q]);
    for (int q = 0; q < built; q++) { hprob_free(&S[q].P); free(S[q].res_X); free(S[q].res_Z); free(S[q].res_y); }
    free(S); free(SP);
    return rc;
}

API int brisk_solve_data(int m, int nblk, const int *bs, const double *c, int64_t nnz, const int *mat, const int *blk,
                         const int *i, const int *j, const double *v, int nopt, const char *const *opts, BriskResult *res) {
    if (!res) return -1;
    if (nnz < 0) { result_clear(res); return 2; }
    BriskProblem p = { m, nblk, bs, c, nnz, mat, blk, i, j, v };
    BriskResult *rr[1] = { res };
    const int rc = solve_many(1, &p, nopt, opts, rr);
    if (rc) return rc;
    return res->exit_code;
}

API int brisk_solve_batch(int n, const BriskProblem *probs, int nopt, const char *const *opts, BriskResult **res) {
    if (n < 0 || (n > 0 && (!probs || !res))) return -1;
    return solve_many(n, probs, nopt, opts, res);
}


API int brisk_solve_file(const char *fname, int nopt, const char *const *opts, BriskResult *res) {
    if (!res) return -1;
    result_clear(res);
    BriskProblem p;
    void *own = NULL;
    if (!fname || brisk_read_sdpa(fname, &p, &own)) { free(own); say(1, "brisk: cannot read %s\n", fname ? fname : "(null)"); return 2; }
    const int rc = brisk_solve_data(p.m, p.nblk, p.bs, p.c, p.nnz, p.mat, p.blk, p.i, p.j, p.v, nopt, opts, res);
    free(own);
    return rc;
}
