/* problem.c - the problem from SDPA sparse triplets (host, C), and the device heap allocator.
 *
 * Conventions of the reference's in-memory entry (brisk.h:128-140, capi.c:175-192): 1-based
 * indices, mat 0 = F0 with C = -F0, mat 1..m = A_i, b = c, entries below the diagonal swapped
 * to the upper triangle, LP blocks (negative sizes) diagonal only, zero values skipped.
 * Scaling follows problem.c:1866-1909 of the reference: d_i = 1/||A_i||_F, bs = max(1,
 * ||d o b||_2), cs = max(1, ||C||_F); b~ = d o b / bs, C~ = C / cs, du_i = bs / d_i.
 *
 * Layout: every SDP block is one internal block; every variable of an LP block is an internal
 * 1 x 1 block, so the cones are handled by one set of batched kernels. Block k occupies n_k^2
 * elements at boff[k] of every block arena. */
#include <math.h>
#include <stdlib.h>
#include <pthread.h>
#include <string.h>
#include "par.h"
#include "solver.h"

typedef struct { int con, k, r, c; double v; } Trip;

/* ---- the setup passes over the entries, in parallel on the host cores ---- */
typedef struct {
    int m, nblk, n;
    const int *bs, *mat, *blk, *ii, *jj;
    const double *v;
    double *Ad, *C, *d;
    const int64_t *rb, *re;         /* dense route: the entries of matrix t (0: F0) at [rb[t], re[t]) */
    int64_t *chg[16];               /* per part: the positions q > 0 where the matrix number changes (at most m + 1) */
    int bad[16];
    int64_t ncon[16], nchg[16];
} SetupCtx;
/* the entries valid (indices within their blocks, LP entries diagonal, finite values); the
 * count of nonzero constraint entries; where the matrix number changes (the first m + 1 places,
 * nchg counting all of them) */
/* an entry that names no matrix or block, lies outside its block, is off the diagonal of an LP
 * block, or is not finite */
static inline int entry_bad(int m, int nblk, const int *bs, int t, int blk1, int i, int j, double v) {
    const int k = blk1 - 1;
    if (t < 0 || t > m || k < 0 || k >= nblk) return 1;
    const int n = abs(bs[k]);
    if (i < 1 || j < 1 || i > n || j > n) return 1;
    if (bs[k] < 0 && i != j) return 1;
    return !isfinite(v);
}
int hprob_check(int m, int nblk, const int *bs, const double *c, int64_t nnz, const int *mat, const int *blk, const int *ii,
                const int *jj, const double *v) {
    if (m < 0 || nblk < 1 || !bs || (m > 0 && !c) || nnz < 0) return 2;
    if (nnz > 0 && (!mat || !blk || !ii || !jj || !v)) return 2;
    for (int k = 0; k < nblk; k++) if (bs[k] == 0) return 2;
    for (int64_t q = 0; q < nnz; q++) if (entry_bad(m, nblk, bs, mat[q], blk[q], ii[q], jj[q], v[q])) return 2;
    return 0;
}
static void check_range(void *c_, int64_t lo, int64_t hi, int part) {
    SetupCtx *c = c_;
    int bad = 0;
    int64_t ncon = 0, nchg = 0;
    int64_t *chg = malloc(sizeof(int64_t) * ((size_t)c->m + 2));
    for (int64_t q = lo; q < hi; q++) {
        if (entry_bad(c->m, c->nblk, c->bs, c->mat[q], c->blk[q], c->ii[q], c->jj[q], c->v[q])) { bad = 1; break; }
        const int t = c->mat[q];
        if (c->v[q] != 0 && t > 0) ncon++;
        if (q > 0 && t != c->mat[q - 1]) {
            if (chg && nchg <= c->m) chg[nchg] = q;
            nchg++;
        }
    }
    c->bad[part] = bad; c->ncon[part] = ncon; c->chg[part] = chg;
    c->nchg[part] = chg ? nchg : (int64_t)c->m + 1;     /* (no list: taken as not grouped) */
}
/* the dense route's host columns: the largest buffer seen (at most 32 GiB) is kept when its
 * problem is freed and handed to the next dense problem that fits in it, so that a repeated
 * solve of that size writes memory that is already mapped (the first touch of a fresh 2.5 GB
 * buffer costs more than building the columns). Problems are built on the serving loop's worker
 * threads too: the kept buffer is taken and given under a lock. */
#define AD_KEEP_MAX ((size_t)4 << 30)           /* doubles */
static double *g_ad_keep;
static size_t g_ad_keep_n;
static pthread_mutex_t g_ad_lock = PTHREAD_MUTEX_INITIALIZER;
static double *ad_take(size_t n, size_t *cap) {
    pthread_mutex_lock(&g_ad_lock);
    if (g_ad_keep && g_ad_keep_n >= n) {
        double *p = g_ad_keep;
        *cap = g_ad_keep_n;
        g_ad_keep = NULL; g_ad_keep_n = 0;
        pthread_mutex_unlock(&g_ad_lock);
        return p;
    }
    pthread_mutex_unlock(&g_ad_lock);
    *cap = n;
    return malloc(sizeof(double) * n);
}
static void ad_give(double *p, size_t cap) {
    if (!p) return;
    double *drop = p;
    pthread_mutex_lock(&g_ad_lock);
    if (cap <= AD_KEEP_MAX && (!g_ad_keep || cap > g_ad_keep_n)) {
        drop = g_ad_keep;
        g_ad_keep = p; g_ad_keep_n = cap;
    }
    pthread_mutex_unlock(&g_ad_lock);
    free(drop);
}
/* 1 / ||A||_F of an n x n column of Ad, or 1 for a zero column */
static double column_dinv(const double *A, int64_t nn) {
    double s2 = 0;
    for (int64_t q = 0; q < nn; q++) s2 += A[q] * A[q];
    return s2 > 0 ? 1.0 / sqrt(s2) : 1.0;
}
/* dense route, matrices lo .. hi - 1 (entries grouped by matrix): each column of Ad built from
 * its entries (full symmetric storage) and d = 1 / its norm taken. The norm is summed during
 * the scatter when no position of the column is written twice (a second write finds a nonzero;
 * the entries' squares then are the column's), else taken from the column afterwards. The
 * columns stay unscaled here: the device scales them by d as it makes their f32 copy */
static void dense_columns(void *c_, int64_t lo, int64_t hi, int part) {
    SetupCtx *c = c_;
    (void)part;
    const int n = c->n;
    const int64_t nn = (int64_t)n * n;
    for (int64_t t = lo; t < hi; t++) {
        double *A = c->Ad + t * nn;
        memset(A, 0, sizeof(double) * nn);
        double s2 = 0;
        int twice = 0;
        for (int64_t q = c->rb[t + 1]; q < c->re[t + 1]; q++) {
            const double v = c->v[q];
            if (v == 0) continue;
            int i = c->ii[q] - 1, j = c->jj[q] - 1;
            if (i > j) { const int s = i; i = j; j = s; }
            double *u = A + i + (int64_t)j * n;
            twice |= *u != 0;
            *u += v;
            if (i != j) { A[j + (int64_t)i * n] += v; s2 += 2 * v * v; }
            else s2 += v * v;
        }
        c->d[t] = twice ? column_dinv(A, nn) : s2 > 0 ? 1.0 / sqrt(s2) : 1.0;
    }
}

static int cmp_con(const void *a, const void *b) {
    const Trip *x = a, *y = b;
    if (x->con != y->con) return x->con < y->con ? -1 : 1;
    if (x->k != y->k) return x->k < y->k ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    return 0;
}
static int cmp_pos(const void *a, const void *b) {
    const Trip *x = a, *y = b;
    if (x->k != y->k) return x->k < y->k ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    if (x->con != y->con) return x->con < y->con ? -1 : 1;
    return 0;
}

void hprob_free(HProb *P) {
    free(P->bs_in); free(P->in_first); free(P->bn); free(P->boff);
    free(P->b); free(P->b0); free(P->d); free(P->du); free(P->C); free(P->cpos); free(P->cval);
    free(P->cptr); free(P->ent); free(P->gptr); free(P->grp); free(P->pptr); free(P->ppos); free(P->pcon); free(P->pval);
    ad_give(P->Ad, P->ad_cap);
    free(P->bdiag); free(P->pairs);
    memset(P, 0, sizeof(*P));
}

/* the entries of each matrix t = 0 .. m as one range [rb[t], re[t]) (empty: rb = re = 0) when
 * every matrix's entries are contiguous, in any order of the matrices (an SDPA file lists them
 * so); the change positions come from the parts of the validity pass. Returns 0, or 1 when the
 * entries are not grouped that way */
static int matrix_ranges(const SetupCtx *c, int64_t nnz, int64_t *rb, int64_t *re) {
    const int m = c->m, P_ = par_parts(nnz, 1 << 20);
    int64_t tot = 0;
    for (int p = 0; p < P_; p++) tot += c->nchg[p];
    if (tot > m) return 1;
    memset(rb, 0, sizeof(int64_t) * (m + 1));
    memset(re, 0, sizeof(int64_t) * (m + 1));
    int64_t a = 0;                                   /* the current run starts at a */
    for (int p = 0; p <= P_ && nnz > 0; p++) {
        const int64_t nc = p < P_ ? c->nchg[p] : 1;
        for (int64_t r = 0; r < nc; r++) {
            const int64_t b = p < P_ ? c->chg[p][r] : nnz;
            const int t = c->mat[a];
            if (re[t] > 0) return 1;                /* a second run of matrix t */
            rb[t] = a; re[t] = b;
            a = b;
        }
    }
    return 0;
}

/* dense route: the scaled A_i in full symmetric storage, column i of Ad = vec(A_i); with each
 * matrix's entries together (rb, re: matrix_ranges) the columns are built, measured and scaled
 * in parallel, each by one thread while it is in cache */
static void finish_scaling(HProb *P, int have_d);
static int hprob_build_dense(HProb *P, int m, int n, int64_t nnz, const int *mat, const int *blk, const int *ii, const int *jj,
                             const double *v, const int64_t *rb, const int64_t *re) {
    const int64_t nn = (int64_t)n * n;
    P->dense = 1;
    P->Ad = ad_take((size_t)nn * m + 1, &P->ad_cap);
    if (!P->Ad) return -1;
    (void)blk;
    int have_d = 0;
    if (rb) {
        for (int64_t e = rb[0]; e < re[0]; e++) {          /* C = -F0 */
            if (v[e] == 0) continue;
            int i = ii[e] - 1, j = jj[e] - 1;
            if (i > j) { const int s = i; i = j; j = s; }
            P->C[i + (int64_t)j * n] += -v[e];
            if (i != j) P->C[j + (int64_t)i * n] += -v[e];
        }
        SetupCtx c;
        memset(&c, 0, sizeof(c));
        c.n = n; c.ii = ii; c.jj = jj; c.v = v; c.Ad = P->Ad; c.d = P->d; c.rb = rb; c.re = re;
        par_for(m, 16, dense_columns, &c);
        have_d = 1;
    } else {
        memset(P->Ad, 0, sizeof(double) * (size_t)nn * m);
        for (int64_t q = 0; q < nnz; q++) {
            if (v[q] == 0) continue;
            int i = ii[q] - 1, j = jj[q] - 1;
            if (i > j) { const int s = i; i = j; j = s; }
            if (mat[q] == 0) {
                P->C[i + (int64_t)j * n] += -v[q];
                if (i != j) P->C[j + (int64_t)i * n] += -v[q];
                continue;
            }
            double *A = P->Ad + (int64_t)(mat[q] - 1) * nn;
            A[i + (int64_t)j * n] += v[q];
            if (i != j) A[j + (int64_t)i * n] += v[q];
        }
    }
    finish_scaling(P, have_d);
    P->cptr = calloc(m + 1, sizeof(int));
    P->ent = malloc(sizeof(SEnt));
    P->gptr = calloc(m + 1, sizeof(int)); P->grp = malloc(sizeof(GrpRef)); P->ngrp = 0; P->gp_tot = 0;
    P->pptr = calloc(1, sizeof(int)); P->ppos = malloc(sizeof(int) * 2); P->pcon = malloc(sizeof(int)); P->pval = malloc(sizeof(double));
    return 0;
}

/* norms on the data as read, row equilibration and scaling of b and C (problem.c:1866-1909);
 * the dense route's d_i = 1 / ||A_i|| from its columns of Ad (have_d: taken already); the
 * columns themselves are scaled on the device */
static void finish_scaling(HProb *P, int have_d) {
    const int m = P->m;
    double c2 = 0, c1 = 0;
    for (int64_t q = 0; q < P->T; q++) { c2 += P->C[q] * P->C[q]; c1 += fabs(P->C[q]); }
    P->normC2 = sqrt(c2); P->normC1 = c1;
    P->normb1 = P->normb2 = 0;
    for (int i = 0; i < m; i++) { P->normb1 += fabs(P->b0[i]); P->normb2 += P->b0[i] * P->b0[i]; }
    P->normb2 = sqrt(P->normb2);
    if (!have_d)
        for (int i = 0; i < m; i++) P->d[i] = column_dinv(P->Ad + (int64_t)i * P->T, P->T);
    double nbv = 0;
    for (int i = 0; i < m; i++) nbv += P->d[i] * P->b[i] * P->d[i] * P->b[i];
    nbv = sqrt(nbv);
    P->bs = nbv > 1 ? nbv : 1.0;
    for (int i = 0; i < m; i++) { P->b[i] *= P->d[i]; P->du[i] = P->bs / P->d[i]; }
    P->cs = P->normC2 > 1 ? P->normC2 : 1.0;
    for (int i = 0; i < m; i++) P->b[i] /= P->bs;
    for (int64_t q = 0; q < P->T; q++) P->C[q] /= P->cs;
    double nb2 = 0;
    for (int i = 0; i < m; i++) nb2 += P->b[i] * P->b[i];
    P->nbs = sqrt(nb2);
}

/* P->pairs: the pairs (a, b), a >= b, of constraints with a block in common, by column b and rows
 * a ascending (2 ints each), when they are at most a quarter of the lower triangle; else none
 * (npair 0) and the assembly visits every pair. Built from the lists of the constraints of each
 * block (the groups), with a stamp per row against pairs that share several blocks. */
/* col[0 .. nc) holds runs in descending order, run r at [rb[r], rb[r + 1]), no value twice: into
 * ascending order through tmp, by merging each reversed run into the result so far */
static void merge_runs(int *col, int nc, const int *rb, int nrun, int *tmp) {
    int len = 0;                                 /* col[0 .. len) sorted ascending */
    for (int r = 0; r < nrun; r++) {
        const int b0 = rb[r], b1 = rb[r + 1];
        int i = 0, j = b1 - 1, o = 0;            /* the run read from its end: ascending */
        while (i < len && j >= b0) tmp[o++] = col[i] < col[j] ? col[i++] : col[j--];
        while (i < len) tmp[o++] = col[i++];
        while (j >= b0) tmp[o++] = col[j--];
        memcpy(col, tmp, sizeof(int) * o);
        len = o;
    }
    (void)nc;
}
static int build_pairs(HProb *P) {
    const int m = P->m, nb = P->nb;
    P->pairs = NULL; P->npair = 0;
    if (m < 64) return 0;
    int *bcnt = calloc(nb + 1, sizeof(int)), *bptr = malloc(sizeof(int) * (nb + 1)), *bcon = malloc(sizeof(int) * (P->ngrp + 1));
    int *stamp = malloc(sizeof(int) * m), *col = malloc(sizeof(int) * m), *tmp = malloc(sizeof(int) * m);
    int maxg = 1;
    for (int a = 0; a < m; a++) if (P->gptr[a + 1] - P->gptr[a] > maxg) maxg = P->gptr[a + 1] - P->gptr[a];
    int *rb = malloc(sizeof(int) * (maxg + 2));
    if (!bcnt || !bptr || !bcon || !stamp || !col || !tmp || !rb) {
        free(bcnt); free(bptr); free(bcon); free(stamp); free(col); free(tmp); free(rb);
        return -1;
    }
    for (int g = 0; g < P->ngrp; g++) bcnt[P->grp[g].k]++;
    bptr[0] = 0;
    for (int k = 0; k < nb; k++) bptr[k + 1] = bptr[k] + bcnt[k];
    for (int k = 0; k < nb; k++) bcnt[k] = bptr[k];
    for (int a = 0; a < m; a++)
        for (int g = P->gptr[a]; g < P->gptr[a + 1]; g++) bcon[bcnt[P->grp[g].k]++] = a;   /* ascending a per block */
    const double cap = 0.25 * (double)m * (m + 1) / 2.0;
    for (int pass = 0; pass < 2; pass++) {
        int64_t np = 0;
        for (int a = 0; a < m; a++) stamp[a] = -1;
        for (int b = 0; b < m; b++) {
            int nc = 0, nrun = 0;
            for (int g = P->gptr[b]; g < P->gptr[b + 1]; g++) {
                const int k = P->grp[g].k;
                rb[nrun++] = nc;
                for (int t = bptr[k + 1] - 1; t >= bptr[k] && bcon[t] >= b; t--) {
                    const int a = bcon[t];
                    if (stamp[a] != b) { stamp[a] = b; col[nc++] = a; }
                }
            }
            rb[nrun] = nc;
            if (pass == 1) {
                merge_runs(col, nc, rb, nrun, tmp);
                for (int t = 0; t < nc; t++) { P->pairs[2 * (np + t)] = col[t]; P->pairs[2 * (np + t) + 1] = b; }
            }
            np += nc;
            if (pass == 0 && np > cap) break;
        }
        if (pass == 0) {
            if (np > cap || np == 0) break;
            P->pairs = malloc(sizeof(int) * 2 * (size_t)np);
            if (!P->pairs) { free(bcnt); free(bptr); free(bcon); free(stamp); free(col); free(tmp); free(rb); return -1; }
        } else P->npair = np;
    }
    free(bcnt); free(bptr); free(bcon); free(stamp); free(col); free(tmp); free(rb);
    return 0;
}

/* returns 0, or 2 for data that is not a valid problem (the reference's exit code for bad data) */
/* groups of entries per (constraint, block) and the heavy ones: a group of 3 or more entries
 * of an SDP block keeps its product P = X A Zi (one MFMA GEMM per group), so that a pair of
 * constraints costs 2 |A_light| loads instead of 4 |A_a| |A_b|. The heaviest groups are chosen
 * first, within a budget of device memory for their P and T arrays (the larger of 256 MB and
 * 16 m^2 bytes). */
static __thread const GrpRef *g_sort_grp;   /* the groups being ranked (set around the qsort; per thread) */
static int cmp_grp_cnt(const void *a, const void *b) {
    const GrpRef *G = g_sort_grp;
    const int x = *(const int *)a, y = *(const int *)b;
    if (G[x].cnt != G[y].cnt) return G[y].cnt - G[x].cnt;
    return x - y;
}
static int build_groups(HProb *P) {
    const int m = P->m;
    int ng = 0;
    for (int a = 0; a < m; a++)
        for (int e = P->cptr[a]; e < P->cptr[a + 1]; e++)
            if (e == P->cptr[a] || P->ent[e].k != P->ent[e - 1].k) ng++;
    P->ngrp = ng;
    P->gptr = calloc(m + 1, sizeof(int));
    P->grp = malloc(sizeof(GrpRef) * (ng + 1));
    if (!P->gptr || !P->grp) return -1;
    int g = 0;
    for (int a = 0; a < m; a++) {
        P->gptr[a] = g;
        for (int e = P->cptr[a]; e < P->cptr[a + 1]; e++) {
            if (e == P->cptr[a] || P->ent[e].k != P->ent[e - 1].k) {
                GrpRef r = { P->ent[e].k, e, 0, 0, -1 };
                P->grp[g++] = r;
            }
            P->grp[g - 1].cnt++;
        }
    }
    P->gptr[m] = g;
    P->gp_tot = 0;
    int *cand = malloc(sizeof(int) * (ng + 1)), nc = 0;
    for (int q = 0; q < ng; q++) if (P->grp[q].cnt >= 3 && P->bn[P->grp[q].k] > 1) cand[nc++] = q;
    g_sort_grp = P->grp;
    qsort(cand, nc, sizeof(int), cmp_grp_cnt);
    g_sort_grp = NULL;
    const int64_t budget = (int64_t)fmax(256.0 * (1 << 20), 16.0 * (double)m * m);
    for (int t = 0; t < nc; t++) {
        GrpRef *G = &P->grp[cand[t]];
        const int64_t n = P->bn[G->k];
        if ((P->gp_tot + n * n) * 16 > budget) continue;
        G->poff = P->gp_tot;
        P->gp_tot += n * n;
    }
    free(cand);
    return 0;
}

/* C's nonzeros sorted by position (stable: the entries of one position keep their order, so
 * that their sum is the dense accumulation's) and merged; zero sums kept (as the dense arena
 * keeps them) */
typedef struct { int64_t pos, seq; double v; } CEnt;
static int cmp_cent(const void *a_, const void *b_) {
    const CEnt *a = a_, *b = b_;
    return a->pos < b->pos ? -1 : a->pos > b->pos ? 1 : (a->seq < b->seq ? -1 : a->seq > b->seq);
}
static void c_merge(HProb *P) {
    CEnt *e = malloc(sizeof(CEnt) * (P->nc + 1));
    for (int64_t q = 0; q < P->nc; q++) { e[q].pos = P->cpos[q]; e[q].seq = q; e[q].v = P->cval[q]; }
    qsort(e, P->nc, sizeof(CEnt), cmp_cent);
    int64_t w = 0;
    for (int64_t q = 0; q < P->nc; q++) {
        if (w > 0 && P->cpos[w - 1] == e[q].pos) P->cval[w - 1] += e[q].v;
        else { P->cpos[w] = e[q].pos; P->cval[w] = 0.0 + e[q].v; w++; }
    }
    P->nc = w;
    free(e);
}
int hprob_build(HProb *P, int m, int nblk, const int *bs, const double *c, int64_t nnz, const int *mat, const int *blk,
                const int *ii, const int *jj, const double *v) {
    memset(P, 0, sizeof(*P));
    if (m < 0 || nblk < 1 || !bs || (m > 0 && !c) || nnz < 0) return 2;
    if (nnz > 0 && (!mat || !blk || !ii || !jj || !v)) return 2;
    P->m = m; P->nblk_in = nblk;
    P->bs_in = malloc(sizeof(int) * nblk);
    P->in_first = malloc(sizeof(int) * nblk);
    int nb = 0;
    for (int k = 0; k < nblk; k++) {
        if (bs[k] == 0) return 2;
        P->bs_in[k] = bs[k];
        P->in_first[k] = nb;
        nb += bs[k] > 0 ? 1 : -bs[k];
        if (bs[k] > 0) P->any_sdp = 1;
    }
    P->nb = nb;
    P->bn = malloc(sizeof(int) * nb);
    P->boff = malloc(sizeof(int64_t) * nb);
    int64_t off = 0;
    for (int k = 0, q = 0; k < nblk; k++) {
        const int cnt = bs[k] > 0 ? 1 : -bs[k], n = bs[k] > 0 ? bs[k] : 1;
        for (int t = 0; t < cnt; t++, q++) {
            P->bn[q] = n; P->boff[q] = off; off += (int64_t)n * n;
            if (n > P->nmax) P->nmax = n;
            P->Ntot += n;
        }
    }
    P->T = off;
    P->b = malloc(sizeof(double) * (m + 1)); P->b0 = malloc(sizeof(double) * (m + 1));
    P->d = malloc(sizeof(double) * (m + 1)); P->du = malloc(sizeof(double) * (m + 1));
    for (int i = 0; i < m; i++) P->b[i] = P->b0[i] = c[i];
    /* triplets of the constraints (internal block, upper position): checked in parallel */
    int64_t ncon = 0;
    SetupCtx sc;
    memset(&sc, 0, sizeof(sc));
    sc.m = m; sc.nblk = nblk; sc.bs = bs; sc.mat = mat; sc.blk = blk; sc.ii = ii; sc.jj = jj; sc.v = v;
    par_for(nnz, 1 << 20, check_range, &sc);
    const int nparts = nnz > 0 ? par_parts(nnz, 1 << 20) : 0;
    int bad = 0;
    for (int p = 0; p < nparts; p++) { bad |= sc.bad[p]; ncon += sc.ncon[p]; }
    if (bad) { for (int p = 0; p < nparts; p++) free(sc.chg[p]); return 2; }
    /* C: dense in the block arena, or by its nonzeros where they are few (their count with the
     * mirrored off-diagonal entries at most an eighth of the arena; the dense route keeps it dense) */
    /* dense route: one SDP block whose constraints fill at least a quarter of their upper
     * triangles (the reference's route 2, solver.c:2968-3016, taken for every constraint) */
    const int dense_route = nblk == 1 && bs[0] > 1 && m > 0 && (double)ncon >= 0.25 * (double)m * bs[0] * (bs[0] + 1) / 2.0 &&
                            (double)m * bs[0] * bs[0] * 8.0 < 64e9;
    int64_t ncin = 0;
    if (!dense_route)
        for (int64_t q = 0; q < nnz; q++) if (mat[q] == 0 && v[q] != 0) ncin += ii[q] == jj[q] ? 1 : 2;
    const int c_sparse = !dense_route && 8 * ncin <= off;
    if (!c_sparse) P->C = calloc(off + 1, sizeof(double));
    else { P->cpos = malloc(sizeof(int64_t) * (ncin + 1)); P->cval = malloc(sizeof(double) * (ncin + 1)); }
    if (dense_route) {
        int64_t *rb = malloc(sizeof(int64_t) * (m + 1)), *re = malloc(sizeof(int64_t) * (m + 1));
        const int grouped = !matrix_ranges(&sc, nnz, rb, re);
        for (int p = 0; p < nparts; p++) free(sc.chg[p]);
        const int e = hprob_build_dense(P, m, bs[0], nnz, mat, blk, ii, jj, v, grouped ? rb : NULL, grouped ? re : NULL);
        free(rb); free(re);
        return e;
    }
    for (int p = 0; p < nparts; p++) free(sc.chg[p]);
    Trip *tr = malloc(sizeof(Trip) * (ncon + 1));
    int64_t nt = 0;
    for (int64_t q = 0; q < nnz; q++) {
        if (v[q] == 0) continue;
        const int t = mat[q], k0 = blk[q] - 1;
        int i = ii[q] - 1, j = jj[q] - 1;
        if (i > j) { const int s = i; i = j; j = s; }
        int k, r, cc;
        if (bs[k0] > 0) { k = P->in_first[k0]; r = i; cc = j; }
        else { k = P->in_first[k0] + i; r = 0; cc = 0; }
        if (t == 0) {
            const int n = P->bn[k];
            const int64_t p0 = P->boff[k] + r + (int64_t)cc * n, p1 = P->boff[k] + cc + (int64_t)r * n;
            if (!c_sparse) {
                P->C[p0] += -v[q];
                if (r != cc) P->C[p1] += -v[q];
            } else {
                P->cpos[P->nc] = p0; P->cval[P->nc++] = -v[q];
                if (r != cc) { P->cpos[P->nc] = p1; P->cval[P->nc++] = -v[q]; }
            }
            continue;
        }
        tr[nt].con = t - 1; tr[nt].k = k; tr[nt].r = r; tr[nt].c = cc; tr[nt].v = v[q]; nt++;
    }
    if (c_sparse) c_merge(P);
    qsort(tr, nt, sizeof(Trip), cmp_con);
    /* merge duplicate positions of one constraint */
    int64_t w = 0;
    for (int64_t q = 0; q < nt; q++) {
        if (w > 0 && tr[w - 1].con == tr[q].con && tr[w - 1].k == tr[q].k && tr[w - 1].r == tr[q].r && tr[w - 1].c == tr[q].c) tr[w - 1].v += tr[q].v;
        else tr[w++] = tr[q];
    }
    nt = w;
    if (nt > 0x7fffffff) { free(tr); return 2; }
    /* norms on the data as read (the nonzeros in the arena's order: the same sums) */
    double c2 = 0, c1 = 0;
    if (!c_sparse) for (int64_t q = 0; q < P->T; q++) { c2 += P->C[q] * P->C[q]; c1 += fabs(P->C[q]); }
    else for (int64_t q = 0; q < P->nc; q++) { c2 += P->cval[q] * P->cval[q]; c1 += fabs(P->cval[q]); }
    P->normC2 = sqrt(c2); P->normC1 = c1;
    for (int i = 0; i < m; i++) { P->normb1 += fabs(P->b0[i]); P->normb2 += P->b0[i] * P->b0[i]; }
    P->normb2 = sqrt(P->normb2);
    /* row equilibration (full symmetric expansion: off-diagonal entries count twice) */
    double *nrm2 = calloc(m + 1, sizeof(double));
    for (int64_t q = 0; q < nt; q++) nrm2[tr[q].con] += (tr[q].r == tr[q].c ? 1.0 : 2.0) * tr[q].v * tr[q].v;
    for (int i = 0; i < m; i++) P->d[i] = nrm2[i] > 0 ? 1.0 / sqrt(nrm2[i]) : 1.0;
    free(nrm2);
    double nbv = 0;
    for (int i = 0; i < m; i++) nbv += P->d[i] * P->b[i] * P->d[i] * P->b[i];
    nbv = sqrt(nbv);
    P->bs = nbv > 1 ? nbv : 1.0;
    for (int64_t q = 0; q < nt; q++) tr[q].v *= P->d[tr[q].con];
    for (int i = 0; i < m; i++) { P->b[i] *= P->d[i]; P->du[i] = P->bs / P->d[i]; }
    P->cs = P->normC2 > 1 ? P->normC2 : 1.0;
    for (int i = 0; i < m; i++) P->b[i] /= P->bs;
    if (!c_sparse) for (int64_t q = 0; q < P->T; q++) P->C[q] /= P->cs;
    else for (int64_t q = 0; q < P->nc; q++) P->cval[q] /= P->cs;
    double nb2 = 0;
    for (int i = 0; i < m; i++) nb2 += P->b[i] * P->b[i];
    P->nbs = sqrt(nb2);
    /* constraint CSR */
    P->nent = (int)nt;
    P->cptr = calloc(m + 1, sizeof(int));
    P->ent = malloc(sizeof(SEnt) * (nt + 1));
    for (int64_t q = 0; q < nt; q++) {
        P->cptr[tr[q].con + 1]++;
        P->ent[q].k = tr[q].k; P->ent[q].r = tr[q].r; P->ent[q].c = tr[q].c; P->ent[q].pad = 0; P->ent[q].v = tr[q].v;
    }
    for (int i = 0; i < m; i++) P->cptr[i + 1] += P->cptr[i];
    if (build_groups(P)) { free(tr); return 2; }
    P->diag = m > 0;
    for (int i = 0; i < m && P->diag; i++) {
        if (P->cptr[i + 1] - P->cptr[i] != 1) P->diag = 0;
        else if (P->ent[P->cptr[i]].r != P->ent[P->cptr[i]].c) P->diag = 0;
    }
    /* the pairs of constraints that share a block, when they are few (many blocks): the Schur
     * assembly then visits only those (the other entries of M are zero) */
    if (!P->diag && build_pairs(P)) { free(tr); return 2; }
    /* blocks every constraint entry of which is diagonal (every LP variable; every SDP block of
     * a max-cut problem): there the dual direction dZ = theta R0 - A*(dy) is diagonal too */
    P->bdiag = malloc(sizeof(int) * (P->nb + 1));
    for (int k = 0; k < P->nb; k++) P->bdiag[k] = 1;
    for (int q = 0; q < P->nent; q++) if (P->ent[q].r != P->ent[q].c) P->bdiag[P->ent[q].k] = 0;
    /* positions: unique (k, r, c) with their constraints in ascending order */
    qsort(tr, nt, sizeof(Trip), cmp_pos);
    int npos = 0;
    for (int64_t q = 0; q < nt; q++)
        if (q == 0 || tr[q].k != tr[q - 1].k || tr[q].r != tr[q - 1].r || tr[q].c != tr[q - 1].c) npos++;
    P->npos = npos;
    P->pptr = malloc(sizeof(int) * (npos + 1));
    P->ppos = malloc(sizeof(int) * 2 * (npos + 1));
    P->pcon = malloc(sizeof(int) * (nt + 1));
    P->pval = malloc(sizeof(double) * (nt + 1));
    int p = -1;
    for (int64_t q = 0; q < nt; q++) {
        if (q == 0 || tr[q].k != tr[q - 1].k || tr[q].r != tr[q - 1].r || tr[q].c != tr[q - 1].c) {
            p++;
            P->pptr[p] = (int)q;
            const int n = P->bn[tr[q].k];
            const int64_t o1 = P->boff[tr[q].k] + tr[q].r + (int64_t)tr[q].c * n;
            const int64_t o2 = tr[q].r == tr[q].c ? -1 : P->boff[tr[q].k] + tr[q].c + (int64_t)tr[q].r * n;
            if (o1 > 0x7fffffff || o2 > 0x7fffffff) { free(tr); return 2; }
            P->ppos[2 * p] = (int)o1; P->ppos[2 * p + 1] = (int)o2;
        }
        P->pcon[q] = tr[q].con; P->pval[q] = tr[q].v;
    }
    P->pptr[npos] = (int)nt;
    free(tr);
    return 0;
}

/* ---- heap allocator: first fit over a sorted free list (host-side bookkeeping only) ---- */
typedef struct { int64_t off, len; } Span;
static Span *g_free = NULL, *g_used = NULL;
static int g_nfree = 0, g_nused = 0, g_capf = 0, g_capu = 0;
static size_t g_heap_bytes = 0;

int heap_setup(size_t bytes) {
    if (g_heap_bytes >= bytes && g_heap_bytes) return 0;
    if (g_nused > 0) return -1;        /* cannot move live arrays */
    if (bk_heap_init(bytes)) return -1;
    g_heap_bytes = bytes;
    free(g_free); free(g_used);
    g_capf = g_capu = 64;
    g_free = malloc(sizeof(Span) * g_capf); g_used = malloc(sizeof(Span) * g_capu);
    g_nfree = 1; g_nused = 0;
    g_free[0].off = 0; g_free[0].len = (int64_t)bytes;
    return 0;
}
int64_t heap_alloc(size_t bytes) {
    const int64_t need = (int64_t)((bytes + 255) & ~(size_t)255) + (bytes == 0 ? 256 : 0);
    for (int i = 0; i < g_nfree; i++) {
        if (g_free[i].len < need) continue;
        const int64_t off = g_free[i].off;
        g_free[i].off += need; g_free[i].len -= need;
        if (g_free[i].len == 0) { memmove(g_free + i, g_free + i + 1, sizeof(Span) * (g_nfree - i - 1)); g_nfree--; }
        if (g_nused == g_capu) { g_capu *= 2; g_used = realloc(g_used, sizeof(Span) * g_capu); }
        g_used[g_nused].off = off; g_used[g_nused].len = need; g_nused++;
        return off;
    }
    return -1;
}
void heap_free(int64_t off) {
    int u = -1;
    for (int i = 0; i < g_nused; i++) if (g_used[i].off == off) { u = i; break; }
    if (u < 0) return;
    Span s = g_used[u];
    g_used[u] = g_used[--g_nused];
    /* insert into the sorted free list and merge neighbours */
    int pos = 0;
    while (pos < g_nfree && g_free[pos].off < s.off) pos++;
    if (g_nfree == g_capf) { g_capf *= 2; g_free = realloc(g_free, sizeof(Span) * g_capf); }
    memmove(g_free + pos + 1, g_free + pos, sizeof(Span) * (g_nfree - pos));
    g_free[pos] = s; g_nfree++;
    if (pos + 1 < g_nfree && g_free[pos].off + g_free[pos].len == g_free[pos + 1].off) {
        g_free[pos].len += g_free[pos + 1].len;
        memmove(g_free + pos + 1, g_free + pos + 2, sizeof(Span) * (g_nfree - pos - 2)); g_nfree--;
    }
    if (pos > 0 && g_free[pos - 1].off + g_free[pos - 1].len == g_free[pos].off) {
        g_free[pos - 1].len += g_free[pos].len;
        memmove(g_free + pos, g_free + pos + 1, sizeof(Span) * (g_nfree - pos - 1)); g_nfree--;
    }
}
size_t heap_largest_free(void) {
    int64_t mx = 0;
    for (int i = 0; i < g_nfree; i++) if (g_free[i].len > mx) mx = g_free[i].len;
    return (size_t)mx;
}
