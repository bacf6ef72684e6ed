/* serve.c - the persistent request loop (continuous batching of SDP solves on one GPU).
 *
 * The reference has no counterpart (its library is not reentrant, capi.c:19-20). Requests queue
 * in arrival order; every step admits waiting requests while solve slots and device memory
 * allow (each solve's projected bytes: its block arrays, its Schur complement and factor, its
 * vectors - solve_heap_bytes), starts the newcomers, advances every solve in flight by one
 * interior-point iteration in one set of batched launches, and finishes the solves that ended.
 * A solve's arithmetic does not depend on its batch-mates: every reduction is per solve in a
 * fixed order, so a result is the same alone and in the mix.
 *
 * Submission checks the data and copies it; the problem's host structures (scaling, constraint
 * lists, groups: hprob_build) are built by a pool of worker threads while the GPU runs the
 * solves in flight, and a request is admitted once its build is done. */
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <pthread.h>
#include "brisk_amd.h"
#include "solver.h"

#define API __attribute__((visibility("default")))

int brisk_parse_options(int nopt, const char *const *opts, SolveOpts *o, int m, int *slots, double *heapgb);
int brisk_heap_ensure(size_t need);
void brisk_fill_result(BriskResult *r, Solve *s, double t_setup, double t_iter, double t_dimacs, double t_down);
void brisk_result_unsolved(BriskResult *r, const char *cause);

/* a request's build job: its own copy of the data, built into s->P by a worker thread (state 0
 * waiting, 1 built, -1 failed: then the request ends unsolved); at a fixed address, since the
 * request table moves as it grows */
typedef struct {
    Solve *s;
    int m, nblk;
    int64_t nnz;
    int *bs, *mat, *blk, *ii, *jj;
    double *c, *v;
    int fault;                  /* BRISK_FAULT=build at submission: the build fails (tests of that path) */
    int state;
} Build;

typedef struct {
    int state;                  /* 0 waiting, 1 in flight, 2 finished, 3 taken */
    Solve *s;
    Build *b;
    double t_submit, t_admit, t_done;
    BriskResult *res;
} Req;

#define SERVE_WORKERS 8

struct BriskServe {
    int nopt;
    char **opts;
    int slots;
    Req *req;
    int nreq, capreq;
    int *queue, qhead, qtail, qcap;   /* waiting request ids, FIFO */
    Solve **fly;                      /* solves in flight */
    int *fly_id;
    int nfly;
    int *fin, nfin, capfin;           /* finished ids not yet reported by brisk_serve_done */
    /* the build pool: jobs in submission order, a lock and two conditions (a job queued, a job
     * done) */
    pthread_t th[SERVE_WORKERS];
    int nth, stop;
    Build **jobs;
    int jhead, jtail, jcap;
    pthread_mutex_t lk;
    pthread_cond_t cv_job, cv_done;
};

static void build_free_data(Build *b) {
    free(b->bs); free(b->mat); free(b->blk); free(b->ii); free(b->jj); free(b->c); free(b->v);
    b->bs = b->mat = b->blk = b->ii = b->jj = NULL; b->c = b->v = NULL;
}
/* the problem's host structures and the method and direction rules (as brisk_solve_data does) */
static int build_run(Build *b) {
    Solve *s = b->s;
    if (b->fault) return -1;
    if (hprob_build(&s->P, b->m, b->nblk, b->bs, b->c, b->nnz, b->mat, b->blk, b->ii, b->jj, b->v)) return -1;
    if (s->o.hsd < 0) s->o.hsd = hsd_first_rule(&s->P, &s->o);
    s->P.want_hsd = s->o.hsd;
    if (s->o.dir < 0) s->o.dir = nt_rule(&s->P);
    s->P.want_nt = !s->o.hsd && s->o.dir == 1;
    return 0;
}
static void *build_worker(void *arg) {
    BriskServe *sv = arg;
    pthread_mutex_lock(&sv->lk);
    for (;;) {
        while (!sv->stop && sv->jhead == sv->jtail) pthread_cond_wait(&sv->cv_job, &sv->lk);
        if (sv->stop) break;
        Build *b = sv->jobs[sv->jhead++];
        pthread_mutex_unlock(&sv->lk);
        const int st = build_run(b) ? -1 : 1;
        build_free_data(b);
        pthread_mutex_lock(&sv->lk);
        b->state = st;
        pthread_cond_broadcast(&sv->cv_done);
    }
    pthread_mutex_unlock(&sv->lk);
    return NULL;
}
/* the build state of a request; with wait, blocks until its build has ended */
static int build_state(BriskServe *sv, Build *b, int wait) {
    pthread_mutex_lock(&sv->lk);
    if (wait) while (b->state == 0) pthread_cond_wait(&sv->cv_done, &sv->lk);
    const int st = b->state;
    pthread_mutex_unlock(&sv->lk);
    return st;
}

API BriskServe *brisk_serve_new(int nopt, const char *const *opts) {
    SolveOpts o;
    int slots = 64;
    double heapgb = 0;
    if (brisk_parse_options(nopt, opts, &o, 0, &slots, &heapgb)) return NULL;
    if (slots < 1) slots = 1;
    if (bk_init()) return NULL;
    if (heapgb > 0 && brisk_heap_ensure((size_t)(heapgb * (double)(1ull << 30)))) return NULL;
    BriskServe *sv = calloc(1, sizeof(BriskServe));
    /* the requests' options: the plan choices for throughput (-throughput 1) unless the caller's
     * options, which follow and so take precedence, say otherwise */
    sv->nopt = nopt + 2;
    sv->opts = malloc(sizeof(char *) * (nopt + 3));
    sv->opts[0] = strdup("-throughput"); sv->opts[1] = strdup("1");
    for (int k = 0; k < nopt; k++) sv->opts[k + 2] = strdup(opts[k]);
    sv->slots = slots;
    sv->fly = malloc(sizeof(Solve *) * slots);
    sv->fly_id = malloc(sizeof(int) * slots);
    pthread_mutex_init(&sv->lk, NULL);
    pthread_cond_init(&sv->cv_job, NULL);
    pthread_cond_init(&sv->cv_done, NULL);
    long nc = sysconf(_SC_NPROCESSORS_ONLN);
    const int want = nc > 2 ? (nc - 1 < SERVE_WORKERS ? (int)nc - 1 : SERVE_WORKERS) : 1;
    for (int t = 0; t < want; t++)
        if (pthread_create(&sv->th[sv->nth], NULL, build_worker, sv) == 0) sv->nth++;
    if (heapgb <= 0) {
        /* default: most of the free device memory */
        const size_t f = bk_mem_free() + bk_heap_size();
        size_t want = f > ((size_t)4 << 30) ? f - ((size_t)4 << 30) : f / 2;
        if (want > ((size_t)96 << 30)) want = (size_t)96 << 30;
        if (brisk_heap_ensure(want)) { brisk_serve_delete(sv); return NULL; }
    }
    return sv;
}

static void *dup_mem(const void *p, size_t n) {
    void *q = malloc(n ? n : 1);
    if (q && n) memcpy(q, p, n);
    return q;
}
API int brisk_serve_submit(BriskServe *sv, const BriskProblem *p) {
    if (!sv || !p) return -1;
    if (hprob_check(p->m, p->nblk, p->bs, p->c, p->nnz, p->mat, p->blk, p->i, p->j, p->v)) return -2;
    Solve *s = calloc(1, sizeof(Solve));
    if (brisk_parse_options(sv->nopt, (const char *const *)sv->opts, &s->o, p->m, NULL, NULL)) { free(s); return -1; }
    Build *b = calloc(1, sizeof(Build));
    const size_t nz = (size_t)p->nnz;
    b->s = s; b->m = p->m; b->nblk = p->nblk; b->nnz = p->nnz;
    b->fault = brisk_fault("build");    /* read here: the workers do not read the environment */
    b->bs = dup_mem(p->bs, sizeof(int) * p->nblk);
    b->c = dup_mem(p->c, sizeof(double) * p->m);
    b->mat = dup_mem(p->mat, sizeof(int) * nz); b->blk = dup_mem(p->blk, sizeof(int) * nz);
    b->ii = dup_mem(p->i, sizeof(int) * nz); b->jj = dup_mem(p->j, sizeof(int) * nz);
    b->v = dup_mem(p->v, sizeof(double) * nz);
    if (sv->nreq == sv->capreq) { sv->capreq = sv->capreq ? 2 * sv->capreq : 256; sv->req = realloc(sv->req, sizeof(Req) * sv->capreq); }
    const int id = sv->nreq++;
    Req *r = &sv->req[id];
    memset(r, 0, sizeof(*r));
    r->s = s;
    r->b = b;
    s->id = id;
    r->t_submit = bk_now();
    if (sv->nth > 0) {
        pthread_mutex_lock(&sv->lk);
        if (sv->jtail == sv->jcap) {
            if (sv->jhead > 0) { memmove(sv->jobs, sv->jobs + sv->jhead, sizeof(Build *) * (sv->jtail - sv->jhead)); sv->jtail -= sv->jhead; sv->jhead = 0; }
            if (sv->jtail == sv->jcap) { sv->jcap = sv->jcap ? 2 * sv->jcap : 256; sv->jobs = realloc(sv->jobs, sizeof(Build *) * sv->jcap); }
        }
        sv->jobs[sv->jtail++] = b;
        pthread_cond_signal(&sv->cv_job);
        pthread_mutex_unlock(&sv->lk);
    } else {                            /* no worker thread: built here */
        b->state = build_run(b) ? -1 : 1;
        build_free_data(b);
    }
    if (sv->qtail == sv->qcap) {
        if (sv->qhead > 0) { memmove(sv->queue, sv->queue + sv->qhead, sizeof(int) * (sv->qtail - sv->qhead)); sv->qtail -= sv->qhead; sv->qhead = 0; }
        if (sv->qtail == sv->qcap) { sv->qcap = sv->qcap ? 2 * sv->qcap : 256; sv->queue = realloc(sv->queue, sizeof(int) * sv->qcap); }
    }
    sv->queue[sv->qtail++] = id;
    return id;
}

static void fin_push(BriskServe *sv, int id) {
    if (sv->nfin == sv->capfin) { sv->capfin = sv->capfin ? 2 * sv->capfin : 256; sv->fin = realloc(sv->fin, sizeof(int) * sv->capfin); }
    sv->fin[sv->nfin++] = id;
}

static void finish_done(BriskServe *sv) {
    int nd = 0;
    for (int q = 0; q < sv->nfly; q++) nd += sv->fly[q]->done;
    if (!nd) return;
    Solve **D = malloc(sizeof(Solve *) * nd);
    int k = 0;
    for (int q = 0; q < sv->nfly; q++) if (sv->fly[q]->done) D[k++] = sv->fly[q];
    const double t0 = bk_now();
    const int bad = ipm_end_batch(D, nd);
    const double t1 = bk_now();
    for (int q = 0; q < nd; q++) {
        Req *r = &sv->req[D[q]->id];
        r->res = brisk_result_new();
        if (bad) D[q]->status = 4;
        brisk_fill_result(r->res, D[q], 0.0, t0 - r->t_admit, (t1 - t0) / nd, 0.0);
        r->t_done = bk_now();
        r->state = 2;
        fin_push(sv, D[q]->id);
        solve_release(D[q]);
        hprob_free(&D[q]->P);
        free(D[q]->res_X); free(D[q]->res_Z); free(D[q]->res_y);
        free(D[q]);
        r->s = NULL;
        free(r->b);
        r->b = NULL;
    }
    int w = 0;
    for (int q = 0; q < sv->nfly; q++) {
        if (sv->req[sv->fly_id[q]].state == 2) continue;
        sv->fly[w] = sv->fly[q]; sv->fly_id[w] = sv->fly_id[q]; w++;
    }
    sv->nfly = w;
    free(D);
}

/* a request that cannot be solved ends unsolved (status -1, exit code -1), its cause given */
static void finish_unsolved(BriskServe *sv, int id, const char *cause) {
    Req *r = &sv->req[id];
    if (r->s) {
        hprob_free(&r->s->P);
        free(r->s->res_X); free(r->s->res_Z); free(r->s->res_y);
        free(r->s);
        r->s = NULL;
    }
    if (r->b) { build_free_data(r->b); free(r->b); r->b = NULL; }
    r->res = brisk_result_new();
    brisk_result_unsolved(r->res, cause);
    r->t_done = bk_now();
    r->state = 2;
    fin_push(sv, id);
}

API int brisk_serve_step(BriskServe *sv) {
    if (!sv) return -1;
    /* admission: in arrival order while a slot and the solve's projected memory are free */
    const double t_adm = bk_prof_on() ? bk_now() : 0;
    Solve **nw = malloc(sizeof(Solve *) * sv->slots);
    int nn = 0;
    while (sv->qhead < sv->qtail && sv->nfly + nn < sv->slots) {
        const int id = sv->queue[sv->qhead];
        Solve *s = sv->req[id].s;
        /* its build: waited for only when nothing would run meanwhile */
        const int bst = build_state(sv, sv->req[id].b, sv->nfly + nn == 0);
        if (bst == 0) break;
        if (bst < 0) {
            sv->qhead++;
            finish_unsolved(sv, id, "the problem's host structures could not be built");
            continue;
        }
        const size_t need = (size_t)solve_heap_bytes(&s->P) + ((size_t)1 << 20);
        if (heap_largest_free() < need) {
            if (sv->nfly + nn == 0 && brisk_heap_ensure(need)) {
                /* larger than the device memory */
                sv->qhead++;
                finish_unsolved(sv, id, "the solve needs more device memory than the GPU has");
                continue;
            }
            if (heap_largest_free() < need) break;
        }
        if (solve_setup(s)) {
            /* with solves in flight the request waits for them to free memory; with none, no
             * later step can do better: it ends unsolved */
            if (sv->nfly + nn > 0) break;
            sv->qhead++;
            finish_unsolved(sv, id, "device setup of the solve failed");
            continue;
        }
        sv->qhead++;
        sv->req[id].state = 1;
        sv->req[id].t_admit = bk_now();
        nw[nn++] = s;
    }
    if (t_adm > 0) bk_prof_host(48, bk_now() - t_adm);
    if (nn) {
        if (ipm_start_batch(nw, nn)) {
            /* a GPU failure: the newcomers end unsolved, their device memory released */
            for (int q = 0; q < nn; q++) { solve_release(nw[q]); finish_unsolved(sv, nw[q]->id, "the start point failed on the GPU"); }
            free(nw);
            return -1;
        }
        for (int q = 0; q < nn; q++) { sv->fly[sv->nfly] = nw[q]; sv->fly_id[sv->nfly] = nw[q]->id; sv->nfly++; }
    }
    free(nw);
    const double t_fin = bk_prof_on() ? bk_now() : 0;
    finish_done(sv);
    if (t_fin > 0) bk_prof_host(51, bk_now() - t_fin);
    if (sv->nfly) {
        if (ipm_iterate_batch(sv->fly, sv->nfly)) return -1;
        finish_done(sv);
    }
    return (sv->qtail - sv->qhead) + sv->nfly;
}

API int brisk_serve_take(BriskServe *sv, int id, BriskResult *res, double *latency) {
    if (!sv || id < 0 || id >= sv->nreq) return -1;
    Req *r = &sv->req[id];
    if (r->state < 2) return 0;
    if (r->state == 3) return -1;
    if (r->res && res) {
        /* move: swap the contents */
        BriskResult *tmp = r->res;
        extern void brisk_result_move(BriskResult *dst, BriskResult *src);
        brisk_result_move(res, tmp);
        brisk_result_delete(tmp);
    } else if (r->res) brisk_result_delete(r->res);
    r->res = NULL;
    if (latency) *latency = r->t_done - r->t_submit;
    r->state = 3;
    return 1;
}
API int brisk_serve_inflight(const BriskServe *sv) { return sv ? sv->nfly : 0; }

/* ids of the requests finished since the last call, in finishing order (at most cap; the rest
 * stay for the next call) */
API int brisk_serve_done(BriskServe *sv, int *ids, int cap) {
    if (!sv || !ids || cap <= 0) return 0;
    const int k = sv->nfin < cap ? sv->nfin : cap;
    memcpy(ids, sv->fin, sizeof(int) * k);
    memmove(sv->fin, sv->fin + k, sizeof(int) * (sv->nfin - k));
    sv->nfin -= k;
    return k;
}

API void brisk_serve_delete(BriskServe *sv) {
    if (!sv) return;
    pthread_mutex_lock(&sv->lk);
    sv->stop = 1;
    pthread_cond_broadcast(&sv->cv_job);
    pthread_mutex_unlock(&sv->lk);
    for (int t = 0; t < sv->nth; t++) pthread_join(sv->th[t], NULL);
    pthread_mutex_destroy(&sv->lk);
    pthread_cond_destroy(&sv->cv_job);
    pthread_cond_destroy(&sv->cv_done);
    for (int q = 0; q < sv->nfly; q++) { solve_release(sv->fly[q]); }
    for (int i = 0; i < sv->nreq; i++) {
        Req *r = &sv->req[i];
        if (r->s) { hprob_free(&r->s->P); free(r->s->res_X); free(r->s->res_Z); free(r->s->res_y); free(r->s); }
        if (r->b) { build_free_data(r->b); free(r->b); }
        if (r->res) brisk_result_delete(r->res);
    }
    for (int k = 0; k < sv->nopt; k++) free(sv->opts[k]);
    free(sv->opts); free(sv->req); free(sv->queue); free(sv->fly); free(sv->fly_id); free(sv->fin); free(sv->jobs);
    free(sv);
}
