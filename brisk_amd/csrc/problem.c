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
#include <string.h>
#include "solver.h"

typedef struct { int con, k, r, c; double v; } Trip;

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
    free(P->b); free(P->b0); free(P->d); free(P->du); free(P->C);
    free(P->cptr); free(P->ent); free(P->gptr); free(P->grp); free(P->pptr); free(P->ppos); free(P->pcon); free(P->pval); free(P->Ad);
    free(P->bdiag);
    memset(P, 0, sizeof(*P));
}

static void finish_scaling(HProb *P);
/* dense route: the scaled A_i in full symmetric storage, column i of Ad = vec(A_i) */
static int hprob_build_dense(HProb *P, int m, int n, int64_t nnz, const int *mat, const int *blk, const int *ii, const int *jj, const double *v) {
    const int64_t nn = (int64_t)n * n;
    P->dense = 1;
    P->Ad = calloc((size_t)nn * m + 1, sizeof(double));
    if (!P->Ad) return -1;
    for (int64_t q = 0; q < nnz; q++) {
        if (v[q] == 0) continue;
        int i = ii[q] - 1, j = jj[q] - 1;
        if (i > j) { const int s = i; i = j; j = s; }
        (void)blk;
        if (mat[q] == 0) {
            P->C[i + (int64_t)j * n] += -v[q];
            if (i != j) P->C[j + (int64_t)i * n] += -v[q];
            continue;
        }
        double *A = P->Ad + (int64_t)(mat[q] - 1) * nn;
        A[i + (int64_t)j * n] += v[q];
        if (i != j) A[j + (int64_t)i * n] += v[q];
    }
    finish_scaling(P);
    /* scale the columns by d_i */
    for (int c = 0; c < m; c++) {
        double *A = P->Ad + (int64_t)c * nn;
        const double dc = P->d[c];
        for (int64_t q = 0; q < nn; q++) A[q] *= dc;
    }
    P->cptr = calloc(m + 1, sizeof(int));
    P->ent = malloc(sizeof(SEnt));
    P->gptr = calloc(m + 1, sizeof(int)); P->grp = malloc(sizeof(GrpRef)); P->ngrp = 0; P->gp_tot = 0;
    P->pptr = calloc(1, sizeof(int)); P->ppos = malloc(sizeof(int) * 2); P->pcon = malloc(sizeof(int)); P->pval = malloc(sizeof(double));
    return 0;
}

/* norms on the data as read, row equilibration and scaling of b and C (problem.c:1866-1909);
 * the dense route's norms of A_i come from Ad */
static void finish_scaling(HProb *P) {
    const int m = P->m;
    double c2 = 0, c1 = 0;
    for (int64_t q = 0; q < P->T; q++) { c2 += P->C[q] * P->C[q]; c1 += fabs(P->C[q]); }
    P->normC2 = sqrt(c2); P->normC1 = c1;
    P->normb1 = P->normb2 = 0;
    for (int i = 0; i < m; i++) { P->normb1 += fabs(P->b0[i]); P->normb2 += P->b0[i] * P->b0[i]; }
    P->normb2 = sqrt(P->normb2);
    const int64_t nn = P->T;
    for (int i = 0; i < m; i++) {
        double s2 = 0;
        const double *A = P->Ad + (int64_t)i * nn;
        for (int64_t q = 0; q < nn; q++) s2 += A[q] * A[q];
        P->d[i] = s2 > 0 ? 1.0 / sqrt(s2) : 1.0;
    }
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

/* returns 0, or 2 for data that is not a valid problem (the reference's exit code for bad data) */
/* groups of entries per (constraint, block) and the heavy ones: a group of 3 or more entries
 * of an SDP block keeps its product P = X A Zi (one MFMA GEMM per group), so that a pair of
 * constraints costs 2 |A_light| loads instead of 4 |A_a| |A_b|. The heaviest groups are chosen
 * first, within a budget of device memory for their P and T arrays (the larger of 256 MB and
 * 16 m^2 bytes). */
static const GrpRef *g_sort_grp;      /* the groups being ranked (set around the qsort) */
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
    P->C = calloc(off + 1, sizeof(double));
    P->b = malloc(sizeof(double) * (m + 1)); P->b0 = malloc(sizeof(double) * (m + 1));
    P->d = malloc(sizeof(double) * (m + 1)); P->du = malloc(sizeof(double) * (m + 1));
    for (int i = 0; i < m; i++) P->b[i] = P->b0[i] = c[i];
    /* triplets of the constraints (internal block, upper position) */
    int64_t ncon = 0;
    for (int64_t q = 0; q < nnz; q++) {
        const int t = mat[q], k = blk[q] - 1;
        if (t < 0 || t > m || k < 0 || k >= nblk) return 2;
        const int n = abs(bs[k]);
        if (ii[q] < 1 || jj[q] < 1 || ii[q] > n || jj[q] > n) return 2;
        if (bs[k] < 0 && ii[q] != jj[q]) return 2;
        if (!isfinite(v[q])) return 2;
        if (v[q] != 0 && t > 0) ncon++;
    }
    /* dense route: one SDP block whose constraints fill at least a quarter of their upper
     * triangles (the reference's route 2, solver.c:2968-3016, taken for every constraint) */
    if (nblk == 1 && bs[0] > 1 && m > 0 && (double)ncon >= 0.25 * (double)m * bs[0] * (bs[0] + 1) / 2.0 &&
        (double)m * bs[0] * bs[0] * 8.0 < 64e9)
        return hprob_build_dense(P, m, bs[0], nnz, mat, blk, ii, jj, v);
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
            P->C[P->boff[k] + r + (int64_t)cc * n] += -v[q];
            if (r != cc) P->C[P->boff[k] + cc + (int64_t)r * n] += -v[q];
            continue;
        }
        tr[nt].con = t - 1; tr[nt].k = k; tr[nt].r = r; tr[nt].c = cc; tr[nt].v = v[q]; nt++;
    }
    qsort(tr, nt, sizeof(Trip), cmp_con);
    /* merge duplicate positions of one constraint */
    int64_t w = 0;
    for (int64_t q = 0; q < nt; q++) {
        if (w > 0 && tr[w - 1].con == tr[q].con && tr[w - 1].k == tr[q].k && tr[w - 1].r == tr[q].r && tr[w - 1].c == tr[q].c) tr[w - 1].v += tr[q].v;
        else tr[w++] = tr[q];
    }
    nt = w;
    if (nt > 0x7fffffff) { free(tr); return 2; }
    /* norms on the data as read */
    double c2 = 0, c1 = 0;
    for (int64_t q = 0; q < P->T; q++) { c2 += P->C[q] * P->C[q]; c1 += fabs(P->C[q]); }
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
    for (int64_t q = 0; q < P->T; q++) P->C[q] /= P->cs;
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
