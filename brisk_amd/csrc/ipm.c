/* ipm.c - the primal-dual interior-point iteration on the GPU, batched over solves (host C).
 *
 * The method is the reference's infeasible-start Mehrotra predictor-corrector with the HKM
 * direction (solver.c:13507-15189); each phase below cites the lines it follows. All matrix and
 * vector work runs as batched HIP launches over every active solve; each solve keeps its own
 * scalars on the host, read back from a single batch scalar buffer once per phase.
 *
 * Precision: in FP64 mode everything is FP64. In mixed mode the Schur complement is factored in
 * FP32 (equilibrated, on MFMA) while max(relgap, pinf, dinf) exceeds mixed_until, and every
 * solve with it is refined by PCG against the FP64 product M v (reference solver.c:5266-5354,
 * 5483-5622); a factor whose PCG needs more than 4 steps or misses its tolerance switches the
 * solve to FP64 for good (reference solver.c:5620). */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "solver.h"

#define NSC 64                  /* scalars per solve in the batch scalar buffer */
#define FLD(s, f) (((Solve *)(s))->f)

void solve_log(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
}

/* ---- growable item lists --------------------------------------------------------------------- */
#define LIST(T, name) typedef struct { T *v; int n, cap; } name; \
    static void name##_add(name *L, T x) { if (L->n == L->cap) { L->cap = L->cap ? 2 * L->cap : 64; L->v = realloc(L->v, sizeof(T) * L->cap); } L->v[L->n++] = x; } \
    static void name##_clear(name *L) { L->n = 0; }
LIST(DotItem, DotL)
LIST(LinItem, LinL)
LIST(GemmItem, GemmL)
LIST(MItem, MItL)
LIST(GemvItem, GemvL)
LIST(ConItem, ConL)
LIST(LzItem, LzL)

static DotL g_dl; static LinL g_ll; static GemmL g_gl; static MItL g_ml; static GemvL g_vl; static ConL g_cl; static LzL g_zl;

/* ---- batch scratch (device) ----------------------------------------------------------------- */
static int64_t g_bsc = -1, g_bint = -1, g_bdbl = -1, g_bab = -1;   /* byte offsets */
static int g_bsc_cap = 0, g_bint_cap = 0, g_bab_cap = 0;
static double *g_hab = NULL;
static double *g_hsc = NULL;                            /* host mirror of the batch scalars */
static int *g_hint = NULL;
static double *g_hdbl = NULL;

static int batch_scratch(int ns, int nslots) {
    if (ns > g_bsc_cap) {
        if (g_bsc >= 0) heap_free(g_bsc);
        g_bsc_cap = ns + 16;
        g_bsc = heap_alloc(sizeof(double) * NSC * g_bsc_cap);
        bk_sync();
        bk_host_free(g_hsc);
        g_hsc = bk_host_alloc(sizeof(double) * NSC * g_bsc_cap);
        bk_dot_mirror(g_bsc / 8, (int64_t)NSC * g_bsc_cap, g_hsc);
        if (g_bsc < 0 || !g_hsc) return -1;
    }
    if (nslots > g_bint_cap) {
        if (g_bint >= 0) heap_free(g_bint);
        if (g_bdbl >= 0) heap_free(g_bdbl);
        g_bint_cap = nslots + 64;
        g_bint = heap_alloc(sizeof(int) * g_bint_cap);
        g_bdbl = heap_alloc(sizeof(double) * g_bint_cap);
        bk_host_free(g_hint); bk_host_free(g_hdbl);
        g_hint = bk_host_alloc(sizeof(int) * g_bint_cap);
        g_hdbl = bk_host_alloc(sizeof(double) * g_bint_cap);
        if (g_bint < 0 || g_bdbl < 0 || !g_hint || !g_hdbl) return -1;
    }
    return 0;
}
void ipm_release_scratch(void) {
    bk_sync();
    if (g_bsc >= 0) heap_free(g_bsc);
    if (g_bint >= 0) heap_free(g_bint);
    if (g_bdbl >= 0) heap_free(g_bdbl);
    if (g_bab >= 0) heap_free(g_bab);
    g_bsc = g_bint = g_bdbl = g_bab = -1;
    bk_dot_mirror(0, 0, NULL);
    bk_host_free(g_hsc); g_hsc = NULL;
    g_bsc_cap = g_bint_cap = g_bab_cap = 0;
}
/* Lanczos coefficients of every block of a batch: 64 doubles per block */
static int batch_ab(int nblocks) {
    if (nblocks <= g_bab_cap) return 0;
    if (g_bab >= 0) heap_free(g_bab);
    g_bab_cap = nblocks + 256;
    g_bab = heap_alloc(sizeof(double) * 64 * g_bab_cap);
    bk_host_free(g_hab);
    g_hab = bk_host_alloc(sizeof(double) * 64 * g_bab_cap);
    return g_bab < 0 ? -1 : 0;
}
static inline int64_t SC(int q, int k) { return g_bsc / 8 + (int64_t)q * NSC + k; }  /* double offset of scalar k of batch slot q */
static inline double HS(int q, int k) { return g_hsc[(size_t)q * NSC + k]; }
static void *heapp(void) { return bk_heap(); }
static int *dint(void) { return (int *)((char *)bk_heap() + g_bint); }
static double *ddbl(void) { return (double *)((char *)bk_heap() + g_bdbl); }
/* the scalars of the batch: the dot products write them to the pinned host mirror directly */
static int read_scalars(int ns) { (void)ns; return bk_sync(); }

/* ---- solve setup ----------------------------------------------------------------------------- */
typedef struct { int64_t dinv_tot, lz_tot; int64_t *dboff, *lzoff; } BlkAux;

static int64_t arr(Solve *S, size_t bytes, int *fail) {
    const int64_t o = heap_alloc(bytes);
    if (o < 0) { *fail = 1; return 0; }
    return o;
}
/* per-solve block auxiliaries are kept in a side table keyed by the solve */
typedef struct { Solve *S; int64_t *dboff, *lzoff; int64_t allocs[96]; int nalloc; } SAux;
static SAux *g_aux = NULL;
static int g_naux = 0, g_capaux = 0;
static SAux *aux_of(Solve *S) {
    for (int i = 0; i < g_naux; i++) if (g_aux[i].S == S) return &g_aux[i];
    return NULL;
}

/* Mwork: the factorization's n x n scratch, and between factorizations the partial sums of the
 * symmetric products with M (T x T slots of 64, T = ceil(m / 64)) */
static int64_t mwork_elems(int64_t m) {
    const int64_t t = (m + 63) / 64, a = m * m, b = 3 * t * t * 64;   /* (three sets of partial-sum slots: schur_solve_k) */
    return a > b ? a : b;
}
/* The automatic choice of the method (reference brisk_solve_body, solver.c:13519-13568): the
 * self-dual embedding first unless its extra dense work per iteration, hsd_first_c (12) sum n^3
 * over the SDP blocks (the C and residual products, the second endpoint), is large next to the
 * factorization of the Schur complement (m^3/3): it is first when that work is at most m^3/3, or
 * at most hsd_first_asm (0.5) of m^3/3 plus the assembly, or when the two together stay below 1e7
 * flops. The assembly's flops are this build's own count of its routes: on the dense route
 * 4 m n^3 + m^2 n^2 (the products P_i = L' A_i L and M = P'P); elsewhere 2 n^3 + n^2 cnt per
 * heavy group (its product X A Zi) and two per pair of entries that meet in a block. */
int hsd_first_rule(const HProb *P, const SolveOpts *o) {
    if (!o->hsd_first) return 0;
    double n3 = 0, asm_fl = 0;
    const double mm = P->m;
    for (int k = 0; k < P->nb; k++) if (P->bn[k] > 1) { const double n = P->bn[k]; n3 += n * n * n; }
    if (P->dense) {
        const double n = P->bn[0];
        asm_fl = 4.0 * mm * n * n * n + mm * mm * n * n;
    } else {
        double *per = calloc(P->nb + 1, sizeof(double));
        for (int g = 0; g < P->ngrp; g++) {
            const GrpRef *r = &P->grp[g];
            const double n = P->bn[r->k];
            per[r->k] += r->cnt;
            if (r->poff >= 0) asm_fl += 2.0 * n * n * n + n * n * r->cnt;
        }
        for (int k = 0; k < P->nb; k++) asm_fl += 2.0 * per[k] * per[k];
        free(per);
    }
    const double fac = mm * mm * mm / 3.0, dense = o->hsd_first_c * n3;
    return dense <= fac || dense <= o->hsd_first_asm * (fac + asm_fl) || fac + dense < 1e7;
}

/* The automatic choice of the direction of the infeasible-start method (reference
 * brisk_solve_body, solver.c:13582-13592): Nesterov-Todd, whose scaling costs about 15 n^3 more
 * dense work per SDP block and iteration than HKM, where that is small next to the Schur
 * factorization and HKM's own dense work (15 sum n^3 < 0.25 (m^3/3 + 5 sum n^3)) or at most 1e8
 * flops in all. */
int nt_rule(const HProb *P) {
    double n3 = 0;
    const double mm = P->m;
    for (int k = 0; k < P->nb; k++) if (P->bn[k] > 1) { const double n = P->bn[k]; n3 += n * n * n; }
    return 15.0 * n3 < 0.25 * (mm * mm * mm / 3.0 + 5.0 * n3) || 15.0 * n3 <= 1e8;
}

int64_t solve_heap_bytes(const HProb *P) {
    const int64_t T = P->T, m = P->m;
    int64_t dinv = 0, lz = 0;
    for (int k = 0; k < P->nb; k++) { dinv += (int64_t)((P->bn[k] + 63) / 64) * 4096; lz += P->bn[k]; }
    int64_t b = 0;
    b += 19 * (T * 8 + 256);                              /* block arena arrays */
    b += 2 * (dinv * 8 + 256);
    b += 2 * (BK_LZ_K + 3) * lz * 8 + 1024;
    b += 4 * (m * m * 8 + 256) + (mwork_elems(m) * 8 + 256) + (((m + 63) / 64) * 4096 * 8 + 256);
    b += 22 * (m * 8 + 256) + NSC * 8 + 256;
    b += (m + 1) * 4 + (int64_t)P->nent * (sizeof(SEnt) + 12) + (int64_t)P->nb * sizeof(BlkRef) + (P->npos + 1) * 12 + 2048;
    b += (m + 1) * 4 + (int64_t)(P->ngrp + 1) * sizeof(GrpRef) + 2 * (P->gp_tot * 8 + 256) + 512;
    b += P->npair * 8 + 256;
    if (P->want_hsd) b += 9 * (T * 8 + 256) + 8 * (m * 8 + 256) + 16 * m * 8 + 256 + 4 * (BK_LZ_K + 3) * lz * 8 + (P->dense ? 4 * T * 8 + 256 : 0);   /* (and four more Lanczos sides) */
    if (P->want_nt) b += 6 * (T * 8 + 256) + 4 * lz * 8 + 256;
    if (P->dense) {
        const int64_t nn = T;
        b += nn * m * (8 + 4 + 8 + 8) + 4 * 256;      /* Ad, Af, P, B */
        b += 32 * m * m * 8 + 256;                     /* split-K partials (at most 32) */
        b += 2 * nn * 4 + 2 * nn * 8 + 4 * 256;        /* f32 factors, M v scratch */
    }
    return b;
}

int solve_setup(Solve *S) {
    HProb *P = &S->P;
    if (brisk_fault("setup")) return -1;
    if (g_naux == g_capaux) { g_capaux = g_capaux ? 2 * g_capaux : 16; g_aux = realloc(g_aux, sizeof(SAux) * g_capaux); }
    SAux *A = &g_aux[g_naux++];
    memset(A, 0, sizeof(*A));
    A->S = S;
    A->dboff = malloc(sizeof(int64_t) * (P->nb + 1));
    A->lzoff = malloc(sizeof(int64_t) * (P->nb + 1));
    int64_t dinv = 0, lz = 0;
    for (int k = 0; k < P->nb; k++) { A->dboff[k] = dinv; A->lzoff[k] = lz; dinv += (int64_t)((P->bn[k] + 63) / 64) * 4096; lz += P->bn[k]; }
    int fail = 0;
    const int64_t T = P->T, m = P->m;
#define ALLOC(field, bytes) do { const int64_t o_ = heap_alloc(bytes); if (o_ < 0) fail = 1; else { A->allocs[A->nalloc++] = o_; S->field = o_ / 8; } } while (0)
#define ALLOCB(field, bytes) do { const int64_t o_ = heap_alloc(bytes); if (o_ < 0) fail = 1; else { A->allocs[A->nalloc++] = o_; S->field = o_; } } while (0)
    ALLOC(X, T * 8); ALLOC(Z, T * 8); ALLOC(Zi, T * 8); ALLOC(LX, T * 8); ALLOC(LZ, T * 8); ALLOC(LXi, T * 8); ALLOC(LZi, T * 8); ALLOC(LXiT, T * 8); ALLOC(LZiT, T * 8);
    ALLOC(dX, T * 8); ALLOC(dZ, T * 8); ALLOC(G, T * 8); ALLOC(Q, T * 8); ALLOC(R0, T * 8); ALLOC(F, T * 8); ALLOC(W1, T * 8);
    ALLOC(W2, T * 8); ALLOC(C, T * 8); ALLOC(Xb, T * 8);
    ALLOC(DinvX, dinv * 8); ALLOC(DinvZ, dinv * 8);
    /* Lanczos bases for two sides (primal, dual), six for the embedding's sigma search */
    const int64_t nsd = S->o.hsd ? 6 : 2;
    ALLOC(lzQ, nsd * (int64_t)BK_LZ_K * lz * 8); ALLOC(lzw, nsd * lz * 8); ALLOC(lzt1, nsd * lz * 8); ALLOC(lzt2, nsd * lz * 8);
    ALLOC(M, m * m * 8); ALLOC(MF, m * m * 8); ALLOC(MLi, m * m * 8); ALLOC(MLiT, m * m * 8); ALLOC(Mwork, mwork_elems(m) * 8);
    ALLOC(MDinv, ((m + 63) / 64) * 4096 * 8); ALLOC(Dm, m * 8);
    ALLOC(y, m * 8); ALLOC(dy, m * 8); ALLOC(rp, m * 8); ALLOC(rhs, m * 8); ALLOC(a0, m * 8); ALLOC(h, m * 8); ALLOC(qv, m * 8);
    ALLOC(ev, m * 8); ALLOC(wv, m * 8); ALLOC(bvec, m * 8); ALLOC(duv, m * 8); ALLOC(pr, m * 8); ALLOC(pz, m * 8); ALLOC(pp, m * 8); ALLOC(pq, m * 8);
    ALLOC(pxb, m * 8); ALLOC(ps, m * 8); ALLOC(t1, m * 8); ALLOC(t2, m * 8); ALLOC(yb, m * 8);
    ALLOCB(cptr, (m + 1) * 4); ALLOCB(ent, (int64_t)(P->nent + 1) * sizeof(SEnt)); ALLOCB(blk, (int64_t)P->nb * sizeof(BlkRef));
    ALLOCB(pptr, (int64_t)(P->npos + 1) * 4); ALLOCB(ppos, (int64_t)(P->npos + 1) * 8); ALLOCB(pcon, (int64_t)(P->nent + 1) * 4);
    ALLOC(pval, (int64_t)(P->nent + 1) * 8);
    ALLOCB(gptr, (m + 1) * 4); ALLOCB(grp, (int64_t)(P->ngrp + 1) * sizeof(GrpRef));
    if (P->gp_tot > 0) { ALLOC(Pg, P->gp_tot * 8); ALLOC(Tg, P->gp_tot * 8); }
    if (P->npair > 0) ALLOCB(pairs, P->npair * 8);
    S->hsd = S->o.hsd;
    if (S->hsd) {
        ALLOC(Rd, T * 8); ALLOC(CX, T * 8); ALLOC(dX0, T * 8); ALLOC(dZ0, T * 8); ALLOC(dX1, T * 8); ALLOC(dZ1, T * 8);
        ALLOC(hv, m * 8); ALLOC(hg, m * 8); ALLOC(ht, m * 8); ALLOC(hy0, m * 8); ALLOC(hy1, m * 8);
        ALLOC(Qp, T * 8); ALLOC(dX0s, T * 8); ALLOC(dZ0s, T * 8); ALLOC(hy0s, m * 8);
        ALLOC(rhs2, m * 8); ALLOC(rhs3, m * 8); ALLOC(xvec, 2 * 8 * m * 8);
        if (P->dense) ALLOC(xom, 4 * T * 8);
    }
    S->nt = !S->hsd && S->o.dir == 1;
    if (S->nt) {
        ALLOC(Wn, T * 8); ALLOC(Gn, T * 8); ALLOC(GnT, T * 8); ALLOC(Gi, T * 8); ALLOC(ntV, T * 8); ALLOC(ntA, T * 8);
        ALLOC(lam, 4 * lz * 8);
    }
    S->SX = S->X; S->SZi = S->Zi; S->SLX = S->LX; S->SLZi = S->LZi;
    S->tau = 1.0; S->kap = 0.0;
    if (P->dense) {
        const int64_t nn = T;
        const int tl = (int)((m + 127) / 128), lower = tl * (tl + 1) / 2;
        int ns_ = (2048 + lower - 1) / lower;
        if (ns_ < 1) ns_ = 1;
        if (ns_ > 32) ns_ = 32;
        if (ns_ > nn / 256) ns_ = nn / 256 > 0 ? (int)(nn / 256) : 1;
        S->nsplit = ns_;
        S->kchunk = (int)((nn + ns_ - 1) / ns_);
        ALLOC(Ad, nn * m * 8); ALLOC(Af, nn * m * 4); ALLOC(Pm, nn * m * 8); ALLOC(Bm, nn * m * 8);
        ALLOC(Mpart, (int64_t)S->nsplit * m * m * 8); ALLOC(LXf, nn * 4); ALLOC(LZif, nn * 4); ALLOC(Om1, nn * 8); ALLOC(Om2, nn * 8);
    }
#undef ALLOC
#undef ALLOCB
    if (fail) { solve_release(S); return -1; }
    char *H = heapp();
    int rc = 0;
    if (P->C) rc |= bk_h2d_async(H + S->C * 8, P->C, T * 8);
    else {
        /* C by its nonzeros: the arena zeroed, positions and values staged in W2 and W1 (free until
         * the first iteration; nc <= T / 8), then scattered */
        rc |= bk_zero(H + S->C * 8, (size_t)T * 8);
        if (P->nc > 0) {
            rc |= bk_h2d_async(H + S->W2 * 8, P->cpos, P->nc * 8);
            rc |= bk_h2d_async(H + S->W1 * 8, P->cval, P->nc * 8);
            rc |= bk_scatter_f64((const int64_t *)(H + S->W2 * 8), (const double *)(H + S->W1 * 8), P->nc, (double *)(H + S->C * 8));
        }
    }
    rc |= bk_h2d_async(H + S->bvec * 8, P->b, m * 8);
    rc |= bk_h2d_async(H + S->duv * 8, P->du, m * 8);
    rc |= bk_h2d_async(H + S->cptr, P->cptr, (m + 1) * 4);
    if (P->nent) rc |= bk_h2d_async(H + S->ent, P->ent, (int64_t)P->nent * sizeof(SEnt));
    BlkRef *br = malloc(sizeof(BlkRef) * P->nb);
    for (int k = 0; k < P->nb; k++) { br[k].n = P->bn[k]; br[k].pad = 0; br[k].off = P->boff[k]; }
    rc |= bk_h2d_async(H + S->blk, br, (int64_t)P->nb * sizeof(BlkRef));
    free(br);
    rc |= bk_h2d_async(H + S->pptr, P->pptr, (int64_t)(P->npos + 1) * 4);
    if (P->npos) rc |= bk_h2d_async(H + S->ppos, P->ppos, (int64_t)P->npos * 8);
    if (P->nent) { rc |= bk_h2d_async(H + S->pcon, P->pcon, (int64_t)P->nent * 4); rc |= bk_h2d_async(H + S->pval * 8, P->pval, (int64_t)P->nent * 8); }
    rc |= bk_h2d_async(H + S->gptr, P->gptr, (m + 1) * 4);
    if (P->ngrp) {
        /* heavy groups: P offsets made absolute (doubles from the heap base) */
        GrpRef *gr = malloc(sizeof(GrpRef) * P->ngrp);
        for (int g = 0; g < P->ngrp; g++) { gr[g] = P->grp[g]; if (gr[g].poff >= 0) gr[g].poff += S->Pg; }
        rc |= bk_h2d_async(H + S->grp, gr, (int64_t)P->ngrp * sizeof(GrpRef));
        free(gr);
    }
    if (P->npair > 0) {
        /* the pairs, and M's zeros outside them (the assembly writes the pairs only) */
        rc |= bk_h2d_async(H + S->pairs, P->pairs, P->npair * 8);
        rc |= bk_zero(H + S->M * 8, (size_t)m * m * 8);
    }
    if (P->dense) {
        /* the columns as read, then scaled by d (staged in t1) with their f32 copy made */
        rc |= bk_h2d_async(H + S->Ad * 8, P->Ad, T * m * 8);
        rc |= bk_h2d_async(H + S->t1 * 8, P->d, m * 8);
        rc |= bk_colscale_d2f(T, m, S->Ad, 2 * S->Af, S->t1);
    }
    if (rc) { solve_release(S); return -1; }
    S->status = -1;
    return 0;
}

void solve_release(Solve *S) {
    SAux *A = aux_of(S);
    if (!A) return;
    bk_sync();
    for (int i = 0; i < A->nalloc; i++) heap_free(A->allocs[i]);
    free(A->dboff); free(A->lzoff);
    *A = g_aux[--g_naux];
}

/* ---- batched building blocks ------------------------------------------------------------- */
#define OFF(f) offsetof(Solve, f)
static inline int64_t F_(Solve *S, size_t f) { return *(int64_t *)((char *)S + f); }
static ConItem con_item(Solve *S) {
    ConItem c;
    memset(&c, 0, sizeof(c));
    c.m = S->P.m; c.nblk = S->P.nb;
    c.cptr = S->cptr; c.ent = S->ent; c.blk = S->blk;
    c.pptr = S->pptr; c.ppos = S->ppos; c.pcon = S->pcon; c.pval = S->pval;
    c.gptr = S->gptr; c.grp = S->grp;
    c.npos = S->P.npos; c.diag = S->P.diag;
    c.pairs = S->pairs; c.npair = S->P.npair;
    return c;
}
/* out = alpha A(D) + beta out, D and out given as double offsets per solve */
static int aop(Solve **S, int ns, const int *act, size_t fD, size_t fout, double alpha, double beta) {
    /* dense route: out = alpha Ad' vec(D) + beta out (one wave per constraint) */
    GemvL_clear(&g_vl);
    for (int q = 0; q < ns; q++) {
        if ((act && !act[q]) || !S[q]->P.dense) continue;
        const HProb *P = &S[q]->P;
        GemvItem g = { (int)P->T, P->m, (int)P->T, S[q]->Ad, F_(S[q], fD), F_(S[q], fout), 0 };
        GemvL_add(&g_vl, g);
    }
    if (g_vl.n) { double *H = heapp(); if (bk_gemv_f64(1, alpha, beta, H, H, H, g_vl.v, g_vl.n)) return -1; }
    ConL_clear(&g_cl);
    for (int q = 0; q < ns; q++) {
        if ((act && !act[q]) || S[q]->P.dense) continue;
        ConItem c = con_item(S[q]);
        c.D = *(int64_t *)((char *)S[q] + fD); c.out = *(int64_t *)((char *)S[q] + fout); c.alpha = alpha; c.beta = beta;
        ConL_add(&g_cl, c);
    }
    return bk_aop(g_cl.v, g_cl.n);
}
/* D += alpha A*(y) at the touched positions */
static int aty(Solve **S, int ns, const int *act, size_t fy, size_t fD, double alpha) {
    /* dense route: vec(D) += alpha Ad y (one thread per entry of D, the constraints in order) */
    GemvL_clear(&g_vl);
    for (int q = 0; q < ns; q++) {
        if ((act && !act[q]) || !S[q]->P.dense) continue;
        const HProb *P = &S[q]->P;
        GemvItem g = { (int)P->T, P->m, (int)P->T, S[q]->Ad, F_(S[q], fy), F_(S[q], fD), 0 };
        GemvL_add(&g_vl, g);
    }
    if (g_vl.n) { double *H = heapp(); if (bk_gemv_f64(0, alpha, 1.0, H, H, H, g_vl.v, g_vl.n)) return -1; }
    ConL_clear(&g_cl);
    for (int q = 0; q < ns; q++) {
        if ((act && !act[q]) || S[q]->P.dense) continue;
        ConItem c = con_item(S[q]);
        c.y = *(int64_t *)((char *)S[q] + fy); c.D = *(int64_t *)((char *)S[q] + fD); c.alpha = alpha;
        ConL_add(&g_cl, c);
    }
    return bk_aty(g_cl.v, g_cl.n);
}

/* lincomb over the block arena (len T) or vectors (len m): out = sum c_t in_t */
static void lin_add(int64_t len, int64_t out, int64_t i0, double c0, int64_t i1, double c1, int64_t i2, double c2, int64_t i3, double c3) {
    LinItem L;
    L.len = len; L.out = out; L.in[0] = i0; L.in[1] = i1; L.in[2] = i2; L.in[3] = i3;
    L.c[0] = c0; L.c[1] = c1; L.c[2] = c2; L.c[3] = c3;
    LinL_add(&g_ll, L);
}
static void dot_add(int64_t len, int64_t x, int64_t y, int64_t out) {
    DotItem d = { len, x, y, out };
    DotL_add(&g_dl, d);
}
/* per-block products: out_k = op(A_k) op(B_k) over every block of the active solves */
static int blk_gemm(Solve **S, int ns, const int *act, int ta, int tb, int tri, double alpha, double beta, size_t fA, size_t fB, size_t fC) {
    GemmL_clear(&g_gl);
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        const HProb *P = &S[q]->P;
        const int64_t a = F_(S[q], fA), b = F_(S[q], fB), c = F_(S[q], fC);
        for (int k = 0; k < P->nb; k++) {
            const int n = P->bn[k];
            GemmItem g = { n, n, n, n, n, n, a + P->boff[k], b + P->boff[k], c + P->boff[k], -1, tri };
            GemmL_add(&g_gl, g);
        }
    }
    double *H = heapp();
    return bk_gemm_f64(ta, tb, alpha, beta, H, H, H, g_gl.v, g_gl.n, NULL);
}
/* Out = Zi V X per block of the active solves (Zi, X the scaling pair: W V W for the Nesterov-Todd direction), V a dual-side direction: dZ = theta R0 - A*(dy)
 * (R0 = -shift I at the start, C - Z - A*(y) after a re-base), or the refinement's A*(w). Each
 * lies in the span of I and the A_i, so it is diagonal in a block whose constraint entries all
 * are (max-cut, LP variables): there Out = Zi (Diag(V) X), a row scaling into W1 and one GEMM;
 * elsewhere W1 = Zi V and Out = W1 X, and W1 keeps Zi V (the corrector's Zi dZ dX reuses it). */
/* dgo: only the diagonal of Out is wanted in the blocks where V is diagonal (their constraint
 * entries are all diagonal, so A(Out) reads nothing else there): Out(r, r) = sum_t Zi(t, r) V(t, t)
 * X(t, r) by one reduction per row instead of the product, W1 not written for them */
static int zi_v_x_(Solve **S, int ns, const int *act, size_t fV, size_t fX, size_t fOut, int dgo) {
    double *H = heapp();
    static MItL dl;
    GemmL_clear(&g_gl);
    MItL_clear(&g_ml);
    MItL_clear(&dl);
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        const HProb *P = &S[q]->P;
        const int64_t v = F_(S[q], fV), x = F_(S[q], fX), w = S[q]->W1, zi = S[q]->SZi, out = F_(S[q], fOut);
        for (int k = 0; k < P->nb; k++) {
            const int n = P->bn[k];
            const int64_t o = P->boff[k];
            if (P->bdiag && P->bdiag[k]) {
                if (dgo) { MItem it = { n, 0, zi + o, v + o, x + o, out + o }; MItL_add(&dl, it); }
                else { MItem it = { n, 0, v + o, x + o, w + o, 0 }; MItL_add(&g_ml, it); }
            } else { GemmItem g = { n, n, n, n, n, n, zi + o, v + o, w + o, -1, 0 }; GemmL_add(&g_gl, g); }
        }
    }
    if (g_gl.n && bk_gemm_f64(0, 0, 1.0, 0.0, H, H, H, g_gl.v, g_gl.n, NULL)) return -1;
    if (g_ml.n && bk_rowscale(g_ml.v, g_ml.n)) return -1;
    if (dl.n && bk_zvx_diag(dl.v, dl.n)) return -1;
    GemmL_clear(&g_gl);
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        const HProb *P = &S[q]->P;
        const int64_t x = F_(S[q], fX), w = S[q]->W1, zi = S[q]->SZi, out = F_(S[q], fOut);
        for (int k = 0; k < P->nb; k++) {
            const int n = P->bn[k];
            const int64_t o = P->boff[k];
            const int dg = P->bdiag && P->bdiag[k];
            if (dg && dgo) continue;
            GemmItem g = { n, n, n, n, n, n, dg ? zi + o : w + o, dg ? w + o : x + o, out + o, -1, 0 };
            GemmL_add(&g_gl, g);
        }
    }
    return g_gl.n ? bk_gemm_f64(0, 0, 1.0, 0.0, H, H, H, g_gl.v, g_gl.n, NULL) : 0;
}
static int zi_v_x(Solve **S, int ns, const int *act, size_t fV, size_t fX, size_t fOut) {
    return zi_v_x_(S, ns, act, fV, fX, fOut, 0);
}
/* the corrector's second-order term Q = Zi dZ dX of the affine directions: W1 dX with W1 = Zi dZ
 * left by the predictor's zi_v_x, and Zi (Diag(dZ) dX) in the blocks where dZ is diagonal */
static int zi_dz_dx(Solve **S, int ns, const int *act) {
    double *H = heapp();
    MItL_clear(&g_ml);
    GemmL_clear(&g_gl);
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        const Solve *s = S[q];
        const HProb *P = &s->P;
        for (int k = 0; k < P->nb; k++) {
            const int n = P->bn[k];
            const int64_t o = P->boff[k];
            const int dg = P->bdiag && P->bdiag[k];
            if (dg) { MItem it = { n, 0, s->dZ + o, s->dX + o, s->W1 + o, 0 }; MItL_add(&g_ml, it); }
            GemmItem g = { n, n, n, n, n, n, dg ? s->Zi + o : s->W1 + o, dg ? s->W1 + o : s->dX + o, s->Q + o, -1, 0 };
            GemmL_add(&g_gl, g);
        }
    }
    if (g_ml.n && bk_rowscale(g_ml.v, g_ml.n)) return -1;
    return bk_gemm_f64(0, 0, 1.0, 0.0, H, H, H, g_gl.v, g_gl.n, NULL);
}
static int blk_sym_(Solve **S, int ns, const int *act, size_t f, int mirror) {
    MItL_clear(&g_ml);
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        const HProb *P = &S[q]->P;
        for (int k = 0; k < P->nb; k++) {
            if (P->bn[k] < 2) continue;
            MItem it = { P->bn[k], 0, F_(S[q], f) + P->boff[k], 0, 0, 0, mirror, 0 };
            MItL_add(&g_ml, it);
        }
    }
    return bk_symmetrize(g_ml.v, g_ml.n);
}
/* G = (G + G')/2 per block */
static int blk_sym(Solve **S, int ns, const int *act, size_t f) { return blk_sym_(S, ns, act, f, 0); }
/* Zi = LZi' LZi per block: the product's lower triangle (K range k >= max(i, j); the tiles above
 * the diagonal skipped), mirrored into the upper one */
static int form_zi(Solve **S, int ns, const int *act) {
    if (blk_gemm(S, ns, act, 1, 0, 1 | 2 | 8, 1.0, 0.0, OFF(LZi), OFF(LZi), OFF(Zi))) return -1;
    return blk_sym_(S, ns, act, OFF(Zi), 1);
}
/* Cholesky of every block of the active (solve, side) pairs in one batched factorisation:
 * L = V (copied), factored in place; side s uses the fields fV[s], fL[s], fD[s] and the scratch
 * fW[s]; act and ok are indexed side * ns + q (ok = 1 when every block is positive definite) */
static int chol_sides_(Solve **S, int ns, int nside, const int *act, const size_t *fV, const size_t *fL, const size_t *fD,
                      const size_t *fW, int *ok) {
    LinL_clear(&g_ll);
    for (int sd = 0; sd < nside; sd++)
        for (int q = 0; q < ns; q++) if (!act || act[sd * ns + q]) lin_add(S[q]->P.T, F_(S[q], fL[sd]), F_(S[q], fV[sd]), 1.0, -1, 0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) return -1;
    MItL_clear(&g_ml);
    int slot = 0;
    int *first = malloc(sizeof(int) * (nside * ns + 1));
    for (int sd = 0; sd < nside; sd++)
        for (int q = 0; q < ns; q++) {
            first[sd * ns + q] = slot;
            if (act && !act[sd * ns + q]) continue;
            const HProb *P = &S[q]->P;
            SAux *A = aux_of(S[q]);
            for (int k = 0; k < P->nb; k++) {
                MItem it = { P->bn[k], slot++, F_(S[q], fL[sd]) + P->boff[k], F_(S[q], fD[sd]) + A->dboff[k], F_(S[q], fW[sd]) + P->boff[k], 0,
                             S[q]->o.tput ? BK_POTRF_SPLIT : 0, 0 };
                MItL_add(&g_ml, it);
            }
        }
    first[nside * ns] = slot;
    if (batch_scratch(ns, slot + 1)) { free(first); return -1; }
    bk_zero(dint(), sizeof(int) * (slot + 1));
    double *H = heapp();
    if (bk_potrf_f64(g_ml.v, g_ml.n, H, H, H, dint(), ddbl())) { free(first); return -1; }
    if (bk_d2h(g_hint, dint(), sizeof(int) * (slot + 1))) { free(first); return -1; }
    for (int u = 0; u < nside * ns; u++) {
        ok[u] = 1;
        if (act && !act[u]) continue;
        for (int t = first[u]; t < first[u + 1]; t++) if (g_hint[t]) ok[u] = 0;
    }
    free(first);
    return 0;
}
static int chol_sides(Solve **S, int ns, int nside, const int *act, const size_t *fV, const size_t *fL, const size_t *fD,
                      const size_t *fW, int *ok) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    const int r_ = chol_sides_(S, ns, nside, act, fV, fL, fD, fW, ok);
    if (t_ > 0) bk_prof_host(49, bk_now() - t_);
    return r_;
}
static int blk_chol(Solve **S, int ns, const int *act, size_t fV, size_t fL, size_t fDinv, int *ok) {
    const size_t a[1] = { fV }, b[1] = { fL }, c[1] = { fDinv }, w[1] = { OFF(W2) };
    return chol_sides(S, ns, 1, act, a, b, c, w, ok);
}
/* Cholesky of V + h_q I (h_q in g_hdbl[q]) for the active solves, built in dX: V is left as it is */
static int shifted_chol(Solve **S, int ns, const int *act, size_t fV, size_t fL, size_t fDinv, int *ok) {
    LinL_clear(&g_ll);
    MItL_clear(&g_ml);
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        lin_add(S[q]->P.T, S[q]->dX, F_(S[q], fV), 1.0, -1, 0, -1, 0, -1, 0);
        for (int k = 0; k < S[q]->P.nb; k++) { MItem it = { S[q]->P.bn[k], q, S[q]->dX + S[q]->P.boff[k], 0, 0, 0 }; MItL_add(&g_ml, it); }
    }
    if (!g_ml.n) { for (int q = 0; q < ns; q++) ok[q] = 1; return 0; }
    if (bk_lincomb(g_ll.v, g_ll.n) || bk_h2d(ddbl(), g_hdbl, sizeof(double) * ns) || bk_add_diag(g_ml.v, g_ml.n, g_bdbl / 8)) return -1;
    return blk_chol(S, ns, act, OFF(dX), fL, fDinv, ok);
}
/* Li = L^{-1} and LiT = Li' per block for the active (solve, side) pairs (after a successful
 * factorisation), all in one batched triangular inversion and one transpose */
static int inv_sides(Solve **S, int ns, int nside, const int *act, const size_t *fL, const size_t *fD, const size_t *fLi,
                     const size_t *fLiT, const size_t *fW) {
    MItL_clear(&g_ml);
    for (int sd = 0; sd < nside; sd++)
        for (int q = 0; q < ns; q++) {
            if (act && !act[sd * ns + q]) continue;
            const HProb *P = &S[q]->P;
            SAux *A = aux_of(S[q]);
            for (int k = 0; k < P->nb; k++) {
                MItem it = { P->bn[k], 0, F_(S[q], fL[sd]) + P->boff[k], F_(S[q], fD[sd]) + A->dboff[k], F_(S[q], fW[sd]) + P->boff[k],
                             F_(S[q], fLi[sd]) + P->boff[k] };
                MItL_add(&g_ml, it);
            }
        }
    if (!g_ml.n) return 0;
    if (batch_scratch(ns, 1)) return -1;
    bk_zero(dint(), sizeof(int));
    double *H = heapp();
    if (bk_trtri_f64(g_ml.v, g_ml.n, H, H, H, H, dint())) return -1;
    /* the transposes: every product with Li then reads columns (one wave per output) */
    for (int t = 0, c = 0; t < nside * ns; t++) {
        const int sd = t / ns, q = t % ns;
        if (act && !act[t]) continue;
        for (int k = 0; k < S[q]->P.nb; k++, c++) { g_ml.v[c].off = F_(S[q], fLi[sd]) + S[q]->P.boff[k]; g_ml.v[c].off4 = F_(S[q], fLiT[sd]) + S[q]->P.boff[k]; }
    }
    return bk_transpose_f64(g_ml.v, g_ml.n);
}
static const size_t F_V[2] = { offsetof(Solve, X), offsetof(Solve, Z) }, F_L[2] = { offsetof(Solve, LX), offsetof(Solve, LZ) },
                    F_D[2] = { offsetof(Solve, DinvX), offsetof(Solve, DinvZ) }, F_Li[2] = { offsetof(Solve, LXi), offsetof(Solve, LZi) },
                    F_LiT[2] = { offsetof(Solve, LXiT), offsetof(Solve, LZiT) }, F_W[2] = { offsetof(Solve, W2), offsetof(Solve, G) },
                    F_dV[2] = { offsetof(Solve, dX), offsetof(Solve, dZ) };

/* ---- Lanczos step-length bound (reference solver.c:6083-6225) ----------------------------- */
/* smallest eigenvalue of the k x k tridiagonal (a, b), b >= 0, by Sturm bisection to full
 * precision, and the last entry of its unit eigenvector by two inverse iterations just below it
 * (the LDL' solve of the positive definite T - sigma I), as the device's Ritz step for blocks of
 * order <= 256 takes them (k <= 30) */
static void tri_min(int k, const double *a, const double *b, double *theta, double *zlast) {
    double lo = INFINITY, hi = -INFINITY, scale = 0;
    for (int u = 0; u < k; u++) {
        const double r = (u > 0 ? fabs(b[u - 1]) : 0.0) + (u + 1 < k ? fabs(b[u]) : 0.0);
        lo = fmin(lo, a[u] - r); hi = fmax(hi, a[u] + r);
        scale = fmax(scale, fabs(a[u]));
    }
    for (int it = 0; it < 200; it++) {
        const double w = hi - lo;
        if (!(w > 2e-16 * fmax(fabs(lo), fabs(hi))) || w <= 1e-300) break;
        const double x = lo + 0.5 * w;
        int cnt = 0;
        double q = a[0] - x;
        if (q < 0) cnt++;
        for (int u = 1; u < k; u++) {
            if (fabs(q) < 1e-300) q = q < 0 ? -1e-300 : 1e-300;
            q = a[u] - x - b[u - 1] * b[u - 1] / q;
            if (q < 0) cnt++;
        }
        if (cnt >= 1) hi = x; else lo = x;
    }
    *theta = 0.5 * (lo + hi);
    const double sig = lo - (hi - lo) - 1e-15 * (scale + 1e-300);
    double x[BK_LZ_K], dd[BK_LZ_K];
    for (int u = 0; u < k; u++) x[u] = 1.0;
    for (int iter = 0; iter < 2; iter++) {
        dd[0] = a[0] - sig;
        for (int u = 1; u < k; u++) {
            if (fabs(dd[u - 1]) < 1e-300) dd[u - 1] = 1e-300;
            const double l = b[u - 1] / dd[u - 1];
            dd[u] = a[u] - sig - l * b[u - 1];
            x[u] -= l * x[u - 1];
        }
        if (fabs(dd[k - 1]) < 1e-300) dd[k - 1] = 1e-300;
        x[k - 1] /= dd[k - 1];
        for (int u = k - 2; u >= 0; u--) x[u] = (x[u] - b[u] * x[u + 1]) / dd[u];
        double nx = 0;
        for (int u = 0; u < k; u++) nx += x[u] * x[u];
        nx = 1.0 / sqrt(nx);
        for (int u = 0; u < k; u++) x[u] *= nx;
    }
    *zlast = x[k - 1];
}
/* Lanczos on T = Li V Li' (operator form: three products per step, Li and its transpose
 * stored) or on T = V (fLi == 0), per block of every active solve, for nside independent sides
 * (the primal and dual steps run in the same launches); every 3 steps from the 6th the
 * coefficients come back and a block whose extreme Ritz value is resolved to lz_tol stops
 * (reference solver.c:6083-6225). out[side * ns + q] = min over the blocks of solve q of the
 * bound theta - res of lambda_min(T) (want_step == 0) or the step 1/max(0, -lambda) (1). */
static int lanczos_min_(Solve **S, int ns, const int *act, int nside, const size_t *fLi, const size_t *fLiT, const size_t *fV,
                       int want_step, double lz_tol, double *out) {
    LzL_clear(&g_zl);
    int kmax = 0, nbt = 0;
    for (int q = 0; q < ns; q++) if (!act || act[q]) nbt += nside * S[q]->P.nb;
    if (batch_ab(nbt + 1)) return -1;
    int *kq = malloc(sizeof(int) * (nbt + 1)), *kb = malloc(sizeof(int) * (nbt + 1)), *ks = malloc(sizeof(int) * (nbt + 1));
    int *live = malloc(sizeof(int) * (nbt + 1));
    double *lam = malloc(sizeof(double) * (nbt + 1));
    for (int side = 0; side < nside; side++)
        for (int q = 0; q < ns; q++) {
            if (act && !act[q]) continue;
            const HProb *P = &S[q]->P;
            SAux *A = aux_of(S[q]);
            int64_t lzn = 0;
            for (int k = 0; k < P->nb; k++) lzn += P->bn[k];
            for (int k = 0; k < P->nb; k++) {
                const int n = P->bn[k], K = n < BK_LZ_K ? n : BK_LZ_K;
                LzItem z = { n, K, S[q]->lzQ + (int64_t)BK_LZ_K * (side * lzn + A->lzoff[k]), S[q]->lzw + side * lzn + A->lzoff[k],
                             g_bab / 8 + 64 * (int64_t)g_zl.n };
                kq[g_zl.n] = q; kb[g_zl.n] = k; ks[g_zl.n] = side; live[g_zl.n] = 1; lam[g_zl.n] = 0;
                LzL_add(&g_zl, z);
                if (K > kmax) kmax = K;
            }
        }
    const int nz = g_zl.n;
    int rc = -1;
    double *H = heapp();
    LzItem *zz = malloc(sizeof(LzItem) * (nz + 1));
    LzfItem *fz = malloc(sizeof(LzfItem) * (nz + 1));
    int *fused = calloc(nz + 1, sizeof(int));
    int nf = 0, nrest = 0;
    /* blocks of order <= 256: the whole bound in one fused launch; the rest step below */
    kmax = 0;
    for (int c = 0; c < nz; c++) {
        const LzItem *z = &g_zl.v[c];
        if (z->n <= 256) {
            Solve *s = S[kq[c]];
            const int sd = ks[c];
            const int64_t bo = s->P.boff[kb[c]];
            LzfItem f;
            memset(&f, 0, sizeof(f));
            f.n = z->n; f.K = z->K; f.mode = fLi[sd] ? 1 : 0;
            f.Li = fLi[sd] ? F_(s, fLi[sd]) + bo : 0; f.LiT = fLiT[sd] ? F_(s, fLiT[sd]) + bo : 0; f.V = F_(s, fV[sd]) + bo;
            /* blocks of order <= 32 run every step (K = n up to 30): the reference takes the
             * exact eigenvalue there (small_lammin, solver.c:6085-6089), and n steps give it */
            f.out = z->ab; f.tol = z->n <= 32 ? 0.0 : lz_tol;
            fz[nf++] = f;
            fused[c] = 1; live[c] = 0;
        } else {
            zz[nrest++] = *z;
            if (z->K > kmax) kmax = z->K;
        }
    }
    if (nf && bk_lanczos_fused(fz, nf)) goto out;
    if (nrest && bk_lanczos_start(zz, nrest)) goto out;
    for (int j = 0; j < kmax; j++) {
        /* w = T q_j for the live blocks: the stages' items in one list (stage-major), one staging */
        const int nstage = fLi[0] ? 3 : 1;
        GemvL_clear(&g_vl);
        for (int stage = 0; stage < nstage; stage++) {
            for (int c = 0; c < nz; c++) {
                const LzItem *z = &g_zl.v[c];
                if (!live[c] || j >= z->K) continue;
                Solve *s = S[kq[c]];
                SAux *A = aux_of(s);
                const int k = kb[c], sd = ks[c];
                int64_t lzn = 0;
                for (int b = 0; b < s->P.nb; b++) lzn += s->P.bn[b];
                const int64_t bo = s->P.boff[k], qj = z->Q + (int64_t)j * z->n;
                const int64_t t1 = s->lzt1 + sd * lzn + A->lzoff[k], t2 = s->lzt2 + sd * lzn + A->lzoff[k];
                GemvItem g;
                g.M = g.N = g.lda = z->n;
                if (!fLi[sd]) { g.a = F_(s, fV[sd]) + bo; g.x = qj; g.y = z->w; g.tri = 0; }
                else if (stage == 0) { g.a = F_(s, fLi[sd]) + bo; g.x = qj; g.y = t1; g.tri = 1; }     /* t1 = Li' q */
                else if (stage == 1) { g.a = F_(s, fV[sd]) + bo; g.x = t1; g.y = t2; g.tri = 0; }      /* t2 = V t1 */
                else { g.a = F_(s, fLiT[sd]) + bo; g.x = t2; g.y = z->w; g.tri = 2; }                  /* w = Li t2 */
                GemvL_add(&g_vl, g);
            }
        }
        if (bk_gemv_chain_f64(H, g_vl.v, g_vl.n / nstage, nstage)) goto out;
        int nl = 0;
        for (int c = 0; c < nz; c++) if (live[c] && j < g_zl.v[c].K) zz[nl++] = g_zl.v[c];
        if (bk_lanczos_step(zz, nl, j)) goto out;
        const int kk = j + 1;
        const int check = kk == kmax || (kk >= 6 && kk % 3 == 0);
        if (!check) continue;
        if (bk_d2h(g_hab, H + g_bab / 8, sizeof(double) * 64 * nz)) goto out;
        int any = 0;
        for (int c = 0; c < nz; c++) {
            if (!live[c]) continue;
            const LzItem *z = &g_zl.v[c];
            const int K = z->K, kc = kk < K ? kk : K;
            const double *al = g_hab + 64 * c, *be = al + 32;
            int m = kc, brk = 0;
            double scale = 0;
            for (int t = 0; t < kc; t++) {
                if (fabs(al[t]) > scale) scale = fabs(al[t]);
                if (!(be[t] > 1e-12 * (scale + 1e-300))) { m = t + 1; brk = 1; break; }
            }
            int bad = 0;
            for (int t = 0; t < m; t++) if (!isfinite(al[t]) || !isfinite(be[t])) bad = 1;
            if (bad) { lam[c] = -1e300; live[c] = 0; continue; }
            double th, zl;
            tri_min(m, al, be, &th, &zl);
            const double res = (brk || m >= z->n) ? 0.0 : fabs(be[m - 1] * zl);
            lam[c] = th - res;
            if (brk || kc >= K || res <= lz_tol * fabs(th) + 1e-14 * scale) live[c] = 0;
            else any = 1;
        }
        if (!any) break;
    }
    if (nf) {
        if (bk_d2h(g_hab, H + g_bab / 8, sizeof(double) * 64 * nz)) goto out;
        for (int c = 0; c < nz; c++) if (fused[c]) lam[c] = g_hab[64 * c];
    }
    for (int side = 0; side < nside; side++)
        for (int q = 0; q < ns; q++) if (!act || act[q]) out[side * ns + q] = want_step ? 1e30 : INFINITY;
    for (int c = 0; c < nz; c++) {
        double *o = &out[ks[c] * ns + kq[c]];
        if (want_step) { const double a = lam[c] >= 0 ? 1e30 : (lam[c] <= -1e299 ? 0.0 : -1.0 / lam[c]); if (a < *o) *o = a; }
        else if (lam[c] < *o) *o = lam[c];
    }
    rc = 0;
out:
    free(kq); free(kb); free(ks); free(live); free(lam); free(zz); free(fz); free(fused);
    return rc;
}
static int lanczos_min(Solve **S, int ns, const int *act, int nside, const size_t *fLi, const size_t *fLiT, const size_t *fV,
                       int want_step, double lz_tol, double *out) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    const int r_ = lanczos_min_(S, ns, act, nside, fLi, fLiT, fV, want_step, lz_tol, out);
    if (t_ > 0) bk_prof_host(50, bk_now() - t_);
    return r_;
}
/* alpha_max per solve for X + ap dX >= 0 and Z + ad dZ >= 0 given the inverse Cholesky factors
 * (reference maxstep, solver.c:6300-6388): both sides in the same launches */
static int maxstep2_(Solve **S, int ns, const int *act, double *ap, double *ad) {
    const size_t fLi[2] = { OFF(LXi), OFF(LZi) }, fLiT[2] = { OFF(LXiT), OFF(LZiT) }, fV[2] = { OFF(dX), OFF(dZ) };
    double *o = malloc(sizeof(double) * 2 * ns);
    const int rc = lanczos_min(S, ns, act, 2, fLi, fLiT, fV, 1, 0.02, o);
    for (int q = 0; q < ns; q++) { ap[q] = o[q]; ad[q] = o[ns + q]; }
    free(o);
    return rc;
}
static int maxstep2(Solve **S, int ns, const int *act, double *ap, double *ad) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    const int r_ = maxstep2_(S, ns, act, ap, ad);
    if (t_ > 0) bk_prof_host(43, bk_now() - t_);
    return r_;
}

/* ---- the Schur complement: assemble, factor, solve ------------------------------------------ */
/* dense route (every constraint dense in one block; the reference's route 2, solver.c:2968-3016,
 * restated as a Gram product): with X = R R' (R = LX) and Zi = S S' (S = LZi'),
 * M_ij = tr(A_i X A_j Zi) = <P_i, P_j> with P_i = R' A_i S. Two batched GEMMs form every P_i,
 * then M = P'P is a split-K product (partials summed in a fixed order). prec 0: f32 MFMA on f32
 * copies, 1: f64. gram: the start point's A A' (X = Zi = I, P = A). */
static int schur_dense(Solve *s, int prec, int gram) {
    const HProb *P = &s->P;
    const int n = P->bn[0], m = P->m;
    const int64_t nn = (int64_t)n * n, mul = prec ? 1 : 2;
    double *H = heapp();
    int64_t Pb;
    GemmL_clear(&g_gl);
    if (gram) Pb = prec ? s->Ad : 2 * s->Af;
    else {
        /* HKM: X = LX LX', Zi = LZi' LZi (triangular: the products' K ranges trimmed); NT: W = G G'
         * with SLX = G and SLZi = G' (full) */
        int64_t lx = s->SLX, lz = s->SLZi;
        const int ta = s->SLX == s->LX ? 2 : 0, tb = s->SLZi == s->LZi ? 16 : 0;
        if (!prec) {
            if (bk_d2f(nn, s->SLX, 2 * s->LXf) || bk_d2f(nn, s->SLZi, 2 * s->LZif)) return -1;
            lx = 2 * s->LXf; lz = 2 * s->LZif;
        }
        const int64_t A = prec ? s->Ad : 2 * s->Af, B = mul * s->Bm;
        Pb = mul * s->Pm;
        for (int i = 0; i < m; i++) { GemmItem g = { n, n, n, n, n, n, lx, A + i * nn, B + i * nn, -1, ta }; GemmL_add(&g_gl, g); }
        if (prec ? bk_gemm_f64(1, 0, 1.0, 0.0, H, H, H, g_gl.v, g_gl.n, NULL)
                 : bk_gemm_f32(1, 0, 1.0f, 0.0f, (float *)H, (float *)H, (float *)H, g_gl.v, g_gl.n, NULL)) return -1;
        GemmL_clear(&g_gl);
        for (int i = 0; i < m; i++) { GemmItem g = { n, n, n, n, n, n, B + i * nn, lz, Pb + i * nn, -1, tb }; GemmL_add(&g_gl, g); }
        if (prec ? bk_gemm_f64(0, 1, 1.0, 0.0, H, H, H, g_gl.v, g_gl.n, NULL)
                 : bk_gemm_f32(0, 1, 1.0f, 0.0f, (float *)H, (float *)H, (float *)H, g_gl.v, g_gl.n, NULL)) return -1;
    }
    GemmL_clear(&g_gl);
    const int64_t part = mul * s->Mpart;
    for (int c = 0; c < s->nsplit; c++) {
        const int64_t k0 = (int64_t)c * s->kchunk;
        const int kc = (int)((nn - k0) < s->kchunk ? (nn - k0) : s->kchunk);
        if (kc <= 0) break;
        GemmItem g = { m, m, kc, (int)nn, (int)nn, m, Pb + k0, Pb + k0, part + (int64_t)c * m * m, -1, 1 };
        GemmL_add(&g_gl, g);
    }
    if (prec ? bk_gemm_f64(1, 0, 1.0, 0.0, H, H, H, g_gl.v, g_gl.n, NULL)
             : bk_gemm_f32(1, 0, 1.0f, 0.0f, (float *)H, (float *)H, (float *)H, g_gl.v, g_gl.n, NULL)) return -1;
    if (bk_splitk_reduce(prec, part, g_gl.n, (int64_t)m * m, m, s->M)) return -1;
    s->m_exact = prec;
    return 0;
}
/* the products of the heavy groups at the current scaling pair (X, Zi, or W, W): T = A Zi_k, then P = X_k T (MFMA) */
static int heavy_products(Solve **S, int ns, const int *act) {
    int ngi = 0;
    for (int q = 0; q < ns; q++) if ((!act || act[q]) && !S[q]->P.dense && S[q]->P.gp_tot > 0) ngi = 1;
    if (!ngi) return 0;
    GrpItem *gi = NULL;
    GemmItem *gm = NULL;
    int n = 0, cap = 0;
    for (int q = 0; q < ns; q++) {
        if ((act && !act[q]) || S[q]->P.dense || S[q]->P.gp_tot <= 0) continue;
        Solve *s = S[q];
        const HProb *P = &s->P;
        for (int g = 0; g < P->ngrp; g++) {
            const GrpRef *G = &P->grp[g];
            if (G->poff < 0) continue;
            if (n == cap) { cap = cap ? 2 * cap : 256; gi = realloc(gi, sizeof(GrpItem) * cap); gm = realloc(gm, sizeof(GemmItem) * cap); }
            const int nk = P->bn[G->k];
            const int64_t boff = P->boff[G->k];
            GrpItem it = { nk, G->cnt, s->ent + (int64_t)G->beg * (int64_t)sizeof(SEnt), s->SZi + boff, s->Tg + G->poff };
            gi[n] = it;
            GemmItem m;
            memset(&m, 0, sizeof(m));
            m.M = nk; m.N = nk; m.K = nk; m.lda = nk; m.ldb = nk; m.ldc = nk;
            m.a = s->SX + boff; m.b = s->Tg + G->poff; m.c = s->Pg + G->poff;
            m.guard = -1; m.tri = 0;
            gm[n] = m;
            n++;
        }
    }
    double *H = heapp();
    int rc = bk_grp_t(gi, n);
    if (!rc) rc = bk_gemm_f64(0, 0, 1.0, 0.0, H, H, H, gm, n, NULL);
    free(gi); free(gm);
    return rc;
}

static int schur_assemble_(Solve **S, int ns, const int *act, int gram) {
    if (heavy_products(S, ns, act)) return -1;
    ConL_clear(&g_cl);
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        if (S[q]->P.dense) {
            const int prec = gram || S[q]->o.prec == PREC_F64 || S[q]->want_double;
            if (schur_dense(S[q], prec, gram)) return -1;
            continue;
        }
        S[q]->m_exact = 1;
        ConItem c = con_item(S[q]);
        c.X = S[q]->SX; c.Zi = S[q]->SZi; c.M = S[q]->M;
        ConL_add(&g_cl, c);
    }
    return bk_schur(g_cl.v, g_cl.n);
}
static int schur_assemble(Solve **S, int ns, const int *act, int gram) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    const int r_ = schur_assemble_(S, ns, act, gram);
    if (t_ > 0) bk_prof_host(40, bk_now() - t_);
    return r_;
}

/* factor M of the active solves (reference solver.c:5266-5420): f32 when allowed (accepted if
 * the smallest squared pivot is >= 1e-6), else f64 with the shift ladder and pivot test.
 * fail[q] = 1 when no factorisation was found. */
static TrsvItem trsv_item(const Solve *s, int prec) {
    const int64_t mul = prec ? 1 : 2;              /* element offsets in the factor's type */
    TrsvItem t = { s->P.m, 0, 0, mul * s->MF, mul * s->MLi, mul * s->MLiT, s->t1, s->t2 };
    return t;
}
static int schur_factor_(Solve **S, int ns, const int *act, int force64, int *fail) {
    int *todo = calloc(ns, sizeof(int)), *f64 = calloc(ns, sizeof(int)), *ran = calloc(ns, sizeof(int));
    double *reg = calloc(ns, sizeof(double));
    if (batch_scratch(ns, 3 * ns + 1)) { free(todo); free(f64); free(ran); free(reg); return -1; }
    double *H = heapp();
    for (int q = 0; q < ns; q++) {
        fail[q] = 0;
        if (act && !act[q]) continue;
        todo[q] = 1;
        Solve *s = S[q];
        f64[q] = force64 || s->o.prec == PREC_F64 || s->want_double;
        const double floor_ = fmax(1e-15, s->P.m * 2.220446049250313e-16);
        reg[q] = f64[q] && s->reg_hint ? fmax(floor_, 0.01 * s->reg_ok) : 0;
    }
    /* each attempt factors the f32 solves on the library stream and the f64 ones on the side
     * stream at the same time (two independent latency-bound chains). Shifts at ddbl[q], smallest
     * squared pivots of the f64 factors at ddbl[ns + q], of the f32 factors at the floats from
     * ddbl[2 ns], factor info at dint[q]. */
    double *pm64 = ddbl() + ns;
    float *pm32 = (float *)(ddbl() + 2 * ns);
    double tA = bk_prof_on() ? bk_now() : 0;
    for (int attempt = 0; attempt < 12; attempt++) {
        int cnt[2] = { 0, 0 };
        for (int q = 0; q < ns; q++) {
            ran[q] = todo[q] ? f64[q] : -1;
            g_hdbl[q] = todo[q] ? reg[q] : 0;
            if (todo[q]) cnt[f64[q]]++;
        }
        if (!cnt[0] && !cnt[1]) break;
        if (bk_h2d(ddbl(), g_hdbl, sizeof(double) * ns)) goto err;
        bk_zero(dint(), sizeof(int) * ns);
        if (cnt[0] && cnt[1] && bk_fork()) goto err;
        for (int prec = 0; prec < 2; prec++) {
            if (!cnt[prec]) continue;
            bk_side(prec == 1);
            /* equilibrate into MF (f32 view of the same bytes for the f32 factor) */
            MItL_clear(&g_ml);
            for (int q = 0; q < ns; q++) {
                if (ran[q] != prec) continue;
                Solve *s = S[q];
                const int64_t f = prec ? s->MF : 2 * s->MF;      /* element offset in the factor's type */
                MItem it = { s->P.m, q, s->M, s->Dm, f };
                MItL_add(&g_ml, it);
            }
            if (bk_equil(g_ml.v, g_ml.n, prec, g_bdbl / 8)) { bk_join(); goto err; }
            /* factor in place: items now point to MF in the factor's element type */
            MItem *fi = malloc(sizeof(MItem) * g_ml.n);
            for (int t = 0; t < g_ml.n; t++) {
                const int q = g_ml.v[t].slot;
                Solve *s = S[q];
                const int64_t mul = prec ? 1 : 2;
                fi[t].n = s->P.m; fi[t].slot = q; fi[t].off = mul * s->MF; fi[t].off2 = mul * s->MDinv; fi[t].off3 = mul * s->Mwork;
                fi[t].off4 = 0; fi[t].pad = 0;
                fi[t].opt = s->o.tput ? (s->P.m >= 1024 ? 256 : 128) | BK_POTRF_SPLIT : 0;
            }
            int rc = prec ? bk_potrf_f64(fi, g_ml.n, H, H, H, dint(), pm64)
                          : bk_potrf_f32(fi, g_ml.n, (float *)H, (float *)H, (float *)H, dint(), pm32);
            free(fi);
            if (rc) { bk_join(); goto err; }
        }
        if (bk_join()) goto err;
        {   /* the factors' failure columns and smallest pivots, one wait for both */
            void *dst[2] = { g_hint, g_hdbl + ns };
            const void *src[2] = { dint(), pm64 };
            const size_t by[2] = { sizeof(int) * ns, sizeof(double) * 2 * ns };
            if (bk_d2h_many(2, dst, src, by)) goto err;
        }
        const float *hp32 = (const float *)(g_hdbl + 2 * ns);
        for (int q = 0; q < ns; q++) {
            if (ran[q] < 0) continue;
            const int prec = ran[q];
            Solve *s = S[q];
            const double pmin = prec ? g_hdbl[ns + q] : (double)hp32[q];
            const double floor_ = fmax(1e-15, s->P.m * 2.220446049250313e-16);
            if (!prec) {
                /* an f32 factor that fails (or has a pivot below 1e-6) is redone in f64 */
                if (g_hint[q] == 0 && pmin >= 1e-6) { todo[q] = 0; s->f32 = 1; s->reg = 0; s->nfactor_f32++; }
                else { f64[q] = 1; s->want_double = 1; reg[q] = s->reg_hint ? fmax(floor_, 0.01 * s->reg_ok) : 0; }
                continue;
            }
            if (g_hint[q] == 0) {
                if (pmin >= 1e-14 || reg[q] >= floor_) {
                    todo[q] = 0; s->f32 = 0; s->reg = reg[q]; s->reg_hint = reg[q] > 0; s->reg_ok = reg[q]; s->nfactor_f64++;
                    continue;
                }
                reg[q] = fmax(floor_, reg[q] * 100);
            } else reg[q] = reg[q] == 0 ? floor_ : reg[q] * 100;
            if (attempt >= 10) { todo[q] = 0; fail[q] = 1; }
        }
    }
    for (int q = 0; q < ns; q++) if (todo[q]) fail[q] = 1;
    if (tA > 0) { bk_sync(); bk_prof_host(54, bk_now() - tA); tA = bk_now(); }
    /* every later solve with the factor is either two products with the explicit inverse and its
     * transpose, or two blocked triangular solves with the factor itself (the inverses of its
     * 256 x 256 diagonal blocks only): whichever the cost model below finds cheaper for the
     * precision group. Again the f32 factors on the library stream and the f64 ones on the side
     * stream. */
    {
        int cnt[2] = { 0, 0 };
        for (int q = 0; q < ns; q++) {
            if ((act && !act[q]) || fail[q]) continue;
            Solve *s = S[q];
            const int prec = s->f32 ? 0 : 1;
            cnt[prec]++;
            /* a solve's -schurapply option (inverse, solves, or auto: the cost model below, on the
             * solve's own order and precision, so that its arithmetic does not depend on the
             * batch). The inverse: m^3 / 3 flops at the rates measured for it (f32 3e13, f64
             * 1.5e13 flop/s) plus 0.3 ms of launches; the solves: two launches of about 15 us per
             * 256-wide block column (measured over serving batches) for each of the 6 (f32) or 3
             * (f64) preconditioner applications of a PCG. The solves lengthen every PCG round by
             * their launch chain, so they are taken for an f64 factor whose inverse costs over
             * 3 ms and twice the solves (m above about 5000), for an f32 factor only past 15 ms
             * and three times the solves. */
            const double m = s->P.m;
            const double t_inv = m * m * m / 3.0 / (prec ? 1.5e13 : 3.0e13) + 3e-4;
            const double t_sv = (prec ? 3 : 6) * 2.0 * ((s->P.m + 255) / 256) * 15e-6;
            /* with -throughput 1 (the serving loop's default) every factor takes the solves: in a
             * large batch their launches are shared by every solve in flight while the inverses'
             * flops add up (the serving stream, 1024 requests: 411 solves/s with the f64 factors
             * of order 1500 or more on the solves, 439 with all of them, 371 with none) */
            const int model = s->o.tput ? 1 : prec ? t_inv > 3e-3 && t_sv < t_inv / 2 : t_inv > 15e-3 && t_sv < t_inv / 3;
            s->trsv = s->o.apply < 0 ? model : s->o.apply;
        }
        bk_zero(dint(), sizeof(int) * ns);
        if (cnt[0] && cnt[1] && bk_fork()) goto err;
        for (int prec = 0; prec < 2; prec++) {
            if (!cnt[prec]) continue;
            bk_side(prec == 1);
            for (int mode = 0; mode < 2; mode++) {
                MItL_clear(&g_ml);
                for (int q = 0; q < ns; q++) {
                    if ((act && !act[q]) || fail[q] || S[q]->f32 == prec || S[q]->trsv != mode) continue;
                    Solve *s = S[q];
                    const int64_t mul = prec ? 1 : 2;
                    MItem it = { s->P.m, q, mul * s->MF, mul * s->MDinv, mul * s->Mwork, mul * s->MLi };
                    MItL_add(&g_ml, it);
                }
                if (!g_ml.n) continue;
                int rc;
                if (mode) {
                    rc = prec ? bk_trtri_diag_f64(g_ml.v, g_ml.n, H, H, H, H, dint())
                              : bk_trtri_diag_f32(g_ml.v, g_ml.n, (float *)H, (float *)H, (float *)H, (float *)H, dint());
                    if (rc) { bk_join(); goto err; }
                    TrsvItem *tv = malloc(sizeof(TrsvItem) * (g_ml.n + 1));
                    for (int t = 0; t < g_ml.n; t++) tv[t] = trsv_item(S[g_ml.v[t].slot], prec);
                    rc = prec ? bk_trsv_prep_f64(tv, g_ml.n, H) : bk_trsv_prep_f32(tv, g_ml.n, (float *)H);
                    free(tv);
                    if (rc) { bk_join(); goto err; }
                    continue;
                }
                rc = prec ? bk_trtri_f64(g_ml.v, g_ml.n, H, H, H, H, dint())
                          : bk_trtri_f32(g_ml.v, g_ml.n, (float *)H, (float *)H, (float *)H, (float *)H, dint());
                if (rc) { bk_join(); goto err; }
                for (int t = 0; t < g_ml.n; t++) {
                    const Solve *s = S[g_ml.v[t].slot];
                    const int64_t mul = prec ? 1 : 2;
                    g_ml.v[t].off = mul * s->MLi; g_ml.v[t].off4 = mul * s->MLiT;
                }
                rc = prec ? bk_transpose_f64(g_ml.v, g_ml.n) : bk_transpose_f32(g_ml.v, g_ml.n);
                if (rc) { bk_join(); goto err; }
            }
        }
        if (bk_join()) goto err;
    }
    if (tA > 0) { bk_sync(); bk_prof_host(55, bk_now() - tA); }
    free(todo); free(f64); free(ran); free(reg);
    return 0;
err:
    free(todo); free(f64); free(ran); free(reg);
    return -1;
}
static int schur_factor(Solve **S, int ns, const int *act, int force64, int *fail) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    const int r_ = schur_factor_(S, ns, act, force64, fail);
    if (t_ > 0) bk_prof_host(41, bk_now() - t_);
    return r_;
}

/* z = M^{-1} r through the factor: z = D Li' Li D r (two products with the inverse, or the two
 * blocked triangular solves with L) */
static int schur_prec(Solve **S, int ns, const int *act, size_t fr, size_t fz) {
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (!act || act[q]) lin_add(S[q]->P.m, S[q]->t1, S[q]->Dm, 1.0, F_(S[q], fr), 0, -1, 0, -1, 0);
    if (bk_vmul(g_ll.v, g_ll.n)) return -1;
    double *H = heapp();
    for (int prec = 0; prec < 2; prec++) {
        int nt = 0;
        for (int q = 0; q < ns; q++) if ((!act || act[q]) && S[q]->f32 != prec && S[q]->trsv) nt++;
        if (!nt) continue;
        TrsvItem *tv = malloc(sizeof(TrsvItem) * nt);
        nt = 0;
        for (int q = 0; q < ns; q++) if ((!act || act[q]) && S[q]->f32 != prec && S[q]->trsv) tv[nt++] = trsv_item(S[q], prec);
        const int rc = prec ? bk_trsv_f64(tv, nt, H, H) : bk_trsv_f32(tv, nt, (const float *)H, H);
        free(tv);
        if (rc) return -1;
    }
    for (int pass = 0; pass < 2; pass++) {
        for (int prec = 0; prec < 2; prec++) {
            GemvL_clear(&g_vl);
            for (int q = 0; q < ns; q++) {
                if ((act && !act[q]) || S[q]->f32 == prec || S[q]->trsv) continue;
                Solve *s = S[q];
                const int64_t mul = prec ? 1 : 2;
                /* pass 0: t2 = Li t1 (columns of Li' = rows of Li); pass 1: t1 = Li' t2 */
                GemvItem g = { s->P.m, s->P.m, s->P.m, mul * (pass == 0 ? s->MLiT : s->MLi), pass == 0 ? s->t1 : s->t2, pass == 0 ? s->t2 : s->t1, pass == 0 ? 2 : 1 };
                GemvL_add(&g_vl, g);
            }
            if (!g_vl.n) continue;
            if (prec ? bk_gemv_f64(1, 1.0, 0.0, H, H, H, g_vl.v, g_vl.n) : bk_gemv_mixed(1, 1.0, 0.0, (float *)H, H, H, g_vl.v, g_vl.n)) return -1;
        }
    }
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (!act || act[q]) lin_add(S[q]->P.m, F_(S[q], fz), S[q]->Dm, 1.0, S[q]->t1, 0, -1, 0, -1, 0);
    return bk_vmul(g_ll.v, g_ll.n);
}
/* y = alpha M x + beta y in FP64: a product with the stored M where it is exact, else the
 * operator form M x = A(X A*(x) Zi) on the f64 data (reference mvfree_apply, solver.c:4755-4773) */
static int schur_mv(Solve **S, int ns, const int *act, size_t fx, size_t fy, double alpha, double beta) {
    /* stored M: the symmetric product from its lower triangle, with the partial sums in Mwork
     * (free between factorizations) */
    SymvItem *sv = malloc(sizeof(SymvItem) * (ns + 1));
    int nsv = 0, nop = 0;
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        Solve *s = S[q];
        if (!s->m_exact) { nop++; continue; }
        SymvItem g = { s->P.m, s->P.m, s->M, F_(s, fx), F_(s, fy), s->Mwork };
        sv[nsv++] = g;
    }
    double *H = heapp();
    const int rs = nsv ? bk_symv_f64(alpha, beta, H, H, H, H, sv, nsv) : 0;
    free(sv);
    if (rs) return -1;
    if (!nop) return 0;
    int *op = calloc(ns, sizeof(int));
    for (int q = 0; q < ns; q++) op[q] = (!act || act[q]) && !S[q]->m_exact;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (op[q]) lin_add(S[q]->P.T, S[q]->Om1, -1, 0, -1, 0, -1, 0, -1, 0);
    int rc = bk_lincomb(g_ll.v, g_ll.n);
    rc = rc || aty(S, ns, op, fx, OFF(Om1), 1.0);
    rc = rc || blk_gemm(S, ns, op, 0, 0, 0, 1.0, 0.0, OFF(SX), OFF(Om1), OFF(Om2));
    rc = rc || blk_gemm(S, ns, op, 0, 0, 0, 1.0, 0.0, OFF(Om2), OFF(SZi), OFF(Om1));
    rc = rc || aop(S, ns, op, OFF(Om1), fy, alpha, beta);
    free(op);
    return rc ? -1 : 0;
}
/* batched PCG on M x = rhs (reference schur_solve_, solver.c:5483-5622); relres[q] out */
static int schur_solve_(Solve **S, int ns, const int *act0, size_t frhs, size_t fx, double tol_rel, double *relres) {
    int *act = malloc(sizeof(int) * ns), *used = calloc(ns, sizeof(int)), *stall = calloc(ns, sizeof(int)), *maxit = malloc(sizeof(int) * ns);
    double *nb = malloc(sizeof(double) * ns), *tol = malloc(sizeof(double) * ns), *best = malloc(sizeof(double) * ns), *rz = malloc(sizeof(double) * ns);
    double *sgain = malloc(sizeof(double) * ns);
    int *smax = malloc(sizeof(int) * ns);
    int rc = -1;
    if (batch_scratch(ns, 1)) goto out;
    for (int q = 0; q < ns; q++) act[q] = !act0 || act0[q];
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (act[q]) dot_add(S[q]->P.m, F_(S[q], frhs), F_(S[q], frhs), SC(q, 0));
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        nb[q] = sqrt(HS(q, 0));
        relres[q] = 0;
        if (nb[q] == 0) { act[q] = 0; lin_add(S[q]->P.m, F_(S[q], fx), -1, 0, -1, 0, -1, 0, -1, 0); }
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (schur_prec(S, ns, act, frhs, fx)) goto out;
    /* r = rhs - M x */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.m, S[q]->pr, F_(S[q], frhs), 1.0, -1, 0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (schur_mv(S, ns, act, fx, OFF(pr), -1.0, 1.0)) goto out;
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (act[q]) dot_add(S[q]->P.m, S[q]->pr, S[q]->pr, SC(q, 0));
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        Solve *s = S[q];
        const double nr = sqrt(HS(q, 0));
        tol[q] = fmax(tol_rel, s->f32 ? 1e-10 : 1e-13) * nb[q];
        if (nr <= tol[q]) { act[q] = 0; relres[q] = nr / nb[q]; continue; }
        best[q] = nr;
        const int stag = s->reg == 0 && !s->f32;
        maxit[q] = (s->reg > 0 || s->f32) ? 50 : 10;
        sgain[q] = stag ? 0.5 : 0.9; smax[q] = stag ? 1 : 4;
    }
    {
        int any = 0;
        for (int q = 0; q < ns; q++) any |= act[q];
        int *was = malloc(sizeof(int) * ns);
        memcpy(was, act, sizeof(int) * ns);
        /* Chronopoulos-Gear form of preconditioned CG: one reduction per step (r.u, w.u, r.r
         * with u = prec(r), w = M u) instead of three; p and s = M p follow by recurrence */
        double *al = calloc(ns, sizeof(double)), *be = calloc(ns, sizeof(double)), *ga = calloc(ns, sizeof(double));
        if (any) {
            /* xb = x; u = prec(r); w = M u; gamma = r.u, delta = w.u */
            LinL_clear(&g_ll);
            for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.m, S[q]->pxb, F_(S[q], fx), 1.0, -1, 0, -1, 0, -1, 0);
            if (bk_lincomb(g_ll.v, g_ll.n)) { free(was); free(al); free(be); free(ga); goto out; }
            if (schur_prec(S, ns, act, OFF(pr), OFF(pz)) || schur_mv(S, ns, act, OFF(pz), OFF(pq), 1.0, 0.0)) { free(was); free(al); free(be); free(ga); goto out; }
            DotL_clear(&g_dl);
            for (int q = 0; q < ns; q++) if (act[q]) { dot_add(S[q]->P.m, S[q]->pr, S[q]->pz, SC(q, 0)); dot_add(S[q]->P.m, S[q]->pq, S[q]->pz, SC(q, 1)); }
            if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) { free(was); free(al); free(be); free(ga); goto out; }
            any = 0;
            for (int q = 0; q < ns; q++) {
                if (!act[q]) continue;
                const double g = HS(q, 0), d = HS(q, 1);
                if (!(d > 0)) { act[q] = 0; continue; }
                ga[q] = g; al[q] = g / d; be[q] = 0;
                any = 1;
            }
        }
        for (int k = 0; any; k++) {
            if (bk_prof_on()) {        /* profile counters: PCG rounds (class 52), active solves summed over rounds (53) */
                int na = 0;
                for (int q = 0; q < ns; q++) na += act[q];
                bk_prof_host(52, 0.0); bk_prof_host(53, 1e-3 * na);
            }
            /* x += a (u + b p), r -= a (w + b s) (old p, s); then p = u + b p, s = w + b s */
            LinL_clear(&g_ll);
            for (int q = 0; q < ns; q++) {
                if (!act[q]) continue;
                Solve *s = S[q];
                const double a = al[q], ab = al[q] * be[q];
                if (k == 0) {
                    lin_add(s->P.m, F_(s, fx), F_(s, fx), 1.0, s->pz, a, -1, 0, -1, 0);
                    lin_add(s->P.m, s->pr, s->pr, 1.0, s->pq, -a, -1, 0, -1, 0);
                } else {
                    lin_add(s->P.m, F_(s, fx), F_(s, fx), 1.0, s->pz, a, s->pp, ab, -1, 0);
                    lin_add(s->P.m, s->pr, s->pr, 1.0, s->pq, -a, s->ps, -ab, -1, 0);
                }
                used[q]++;
                s->pcg_total++;
            }
            if (bk_lincomb(g_ll.v, g_ll.n)) { free(was); free(al); free(be); free(ga); goto out; }
            LinL_clear(&g_ll);
            for (int q = 0; q < ns; q++) {
                if (!act[q]) continue;
                Solve *s = S[q];
                if (k == 0) {
                    lin_add(s->P.m, s->pp, s->pz, 1.0, -1, 0, -1, 0, -1, 0);
                    lin_add(s->P.m, s->ps, s->pq, 1.0, -1, 0, -1, 0, -1, 0);
                } else {
                    lin_add(s->P.m, s->pp, s->pz, 1.0, s->pp, be[q], -1, 0, -1, 0);
                    lin_add(s->P.m, s->ps, s->pq, 1.0, s->ps, be[q], -1, 0, -1, 0);
                }
            }
            if (bk_lincomb(g_ll.v, g_ll.n)) { free(was); free(al); free(be); free(ga); goto out; }
            /* u = prec(r), w = M u, and the one reduction: r.r, r.u, w.u */
            if (schur_prec(S, ns, act, OFF(pr), OFF(pz)) || schur_mv(S, ns, act, OFF(pz), OFF(pq), 1.0, 0.0)) { free(was); free(al); free(be); free(ga); goto out; }
            DotL_clear(&g_dl);
            for (int q = 0; q < ns; q++) if (act[q]) {
                dot_add(S[q]->P.m, S[q]->pr, S[q]->pr, SC(q, 0));
                dot_add(S[q]->P.m, S[q]->pr, S[q]->pz, SC(q, 1));
                dot_add(S[q]->P.m, S[q]->pq, S[q]->pz, SC(q, 2));
            }
            if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) { free(was); free(al); free(be); free(ga); goto out; }
            LinL_clear(&g_ll);
            any = 0;
            for (int q = 0; q < ns; q++) {
                if (!act[q]) continue;
                const double nr = sqrt(HS(q, 0));
                if (nr < best[q]) {
                    stall[q] = nr > sgain[q] * best[q] ? stall[q] + 1 : 0;
                    best[q] = nr;
                    lin_add(S[q]->P.m, S[q]->pxb, F_(S[q], fx), 1.0, -1, 0, -1, 0, -1, 0);
                } else stall[q]++;
                if (nr <= tol[q] || stall[q] >= smax[q] || used[q] >= maxit[q]) { act[q] = 0; continue; }
                const double gn = HS(q, 1), d = HS(q, 2);
                const double b2 = gn / ga[q], den = d - b2 * gn / al[q];
                if (!(den > 0) || !isfinite(b2)) { act[q] = 0; continue; }
                be[q] = b2; al[q] = gn / den; ga[q] = gn;
                any = 1;
            }
            if (bk_lincomb(g_ll.v, g_ll.n)) { free(was); free(al); free(be); free(ga); goto out; }
        }
        free(al); free(be); free(ga);
        /* x = xb; true residual */
        any = 0;
        for (int q = 0; q < ns; q++) any |= was[q];
        if (any) {
            LinL_clear(&g_ll);
            for (int q = 0; q < ns; q++) if (was[q]) {
                lin_add(S[q]->P.m, F_(S[q], fx), S[q]->pxb, 1.0, -1, 0, -1, 0, -1, 0);
                lin_add(S[q]->P.m, S[q]->pr, F_(S[q], frhs), 1.0, -1, 0, -1, 0, -1, 0);
            }
            if (bk_lincomb(g_ll.v, g_ll.n)) { free(was); goto out; }
            if (schur_mv(S, ns, was, fx, OFF(pr), -1.0, 1.0)) { free(was); goto out; }
            DotL_clear(&g_dl);
            for (int q = 0; q < ns; q++) if (was[q]) dot_add(S[q]->P.m, S[q]->pr, S[q]->pr, SC(q, 0));
            if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) { free(was); goto out; }
            for (int q = 0; q < ns; q++) {
                if (!was[q]) continue;
                const double rel = sqrt(HS(q, 0)) / nb[q];
                relres[q] = rel;
                Solve *s = S[q];
                if (s->f32 && (used[q] > 4 || best[q] > tol[q] || rel > fmax(1e-9, 10 * tol_rel))) s->want_double = 1;
            }
        }
        free(was);
    }
    rc = 0;
out:
    free(act); free(used); free(stall); free(maxit); free(nb); free(tol); free(best); free(rz); free(sgain); free(smax);
    return rc;
}
static int schur_solve(Solve **S, int ns, const int *act0, size_t frhs, size_t fx, double tol_rel, double *relres) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    const int r_ = schur_solve_(S, ns, act0, frhs, fx, tol_rel, relres);
    if (t_ > 0) bk_prof_host(42, bk_now() - t_);
    return r_;
}

/* k right-hand sides per solve (k <= 3) in one batched PCG: the solves are entered k times, as
 * copies whose PCG vectors, preconditioner scratch, partial-sum slots of the symmetric product
 * (the j-th set of ceil(m/64)^2 slots in Mwork) and operator-form scratch are their own (the
 * solve's extra vectors), so that every launch covers all k systems; right-hand side j at field
 * fr[j], solution at fx[j]. The PCG counts and the switch to f64 go back to the solves. */
static int schur_solve_k(Solve **S, int ns, const int *act, int k, const size_t *fr, const size_t *fx, double tol_rel) {
    Solve *C = malloc(sizeof(Solve) * (size_t)ns * k);
    Solve **SS = malloc(sizeof(Solve *) * (size_t)ns * k);
    int *a = calloc((size_t)ns * k, sizeof(int));
    double *rel = calloc((size_t)ns * k, sizeof(double));
    for (int j = 0; j < k; j++)
        for (int q = 0; q < ns; q++) {
            const int u = j * ns + q;
            Solve *c = &C[u];
            *c = *S[q];
            SS[u] = c;
            a[u] = !act || act[q];
            if (!a[u]) continue;
            c->rhs = F_(S[q], fr[j]); c->wv = F_(S[q], fx[j]);
            if (j > 0) {
                const int64_t m = S[q]->P.m, base = S[q]->xvec + (int64_t)(j - 1) * 8 * m, tt = (m + 63) / 64;
                c->pr = base; c->pz = base + m; c->pp = base + 2 * m; c->pq = base + 3 * m; c->ps = base + 4 * m; c->pxb = base + 5 * m;
                c->t1 = base + 6 * m; c->t2 = base + 7 * m;
                c->Mwork = S[q]->Mwork + (int64_t)j * tt * tt * 64;
                if (S[q]->P.dense) { c->Om1 = S[q]->xom + (int64_t)(j - 1) * 2 * S[q]->P.T; c->Om2 = c->Om1 + S[q]->P.T; }
            }
        }
    const int rc = schur_solve(SS, ns * k, a, OFF(rhs), OFF(wv), tol_rel, rel);
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        const int p0 = S[q]->pcg_total;          /* (every copy started from it) */
        for (int j = 0; j < k; j++) {
            S[q]->pcg_total += C[j * ns + q].pcg_total - p0;
            if (C[j * ns + q].want_double) S[q]->want_double = 1;
        }
    }
    free(C); free(SS); free(a); free(rel);
    return rc;
}

/* ---- start point (reference solver.c:13940-14210) ----------------------------------------- */
static int set_identity(Solve **S, int ns, const int *act, size_t f) {
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (!act || act[q]) lin_add(S[q]->P.T, F_(S[q], f), -1, 0, -1, 0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) return -1;
    MItL_clear(&g_ml);
    for (int q = 0; q < ns; q++) {
        if (act && !act[q]) continue;
        for (int k = 0; k < S[q]->P.nb; k++) { MItem it = { S[q]->P.bn[k], q, F_(S[q], f) + S[q]->P.boff[k], 0, 0 }; MItL_add(&g_ml, it); }
    }
    for (int q = 0; q < ns; q++) g_hdbl[q] = 1.0;
    if (bk_h2d(ddbl(), g_hdbl, sizeof(double) * ns)) return -1;
    return bk_add_diag(g_ml.v, g_ml.n, g_bdbl / 8);
}
/* lambda_min of V per solve (Lanczos on V itself, conservative: theta - residual bound) */
static int lammin_all(Solve **S, int ns, const int *act, size_t fV, double *lam) {
    const size_t z[1] = { 0 }, v[1] = { fV };
    return lanczos_min(S, ns, act, 1, z, z, v, 0, 0.02, lam);
}

static int ipm_start_batch_(Solve **S, int ns);
int ipm_start_batch(Solve **S, int ns) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    const int r_ = ipm_start_batch_(S, ns);
    if (t_ > 0) bk_prof_host(46, bk_now() - t_);
    return r_;
}
static int ipm_start_batch_(Solve **S, int ns) {
    int rc = -1;
    int *fail = calloc(ns, sizeof(int)), *ok = calloc(ns, sizeof(int)), *act = calloc(ns, sizeof(int));
    double *rel = calloc(ns, sizeof(double)), *lx = calloc(ns, sizeof(double)), *lz = calloc(ns, sizeof(double));
    double *extra = calloc(ns, sizeof(double));
    int nsl = 1;
    for (int q = 0; q < ns; q++) nsl += S[q]->P.nb;
    if (batch_scratch(ns, nsl + ns)) goto out;
    for (int q = 0; q < ns; q++) {
        Solve *s = S[q];
        s->t_start = bk_now();
        s->status = -1; s->done = 0; s->it = 0; s->stall = 0; s->it_best = 0;
        s->best_score = 1e300; s->theta = 1.0; s->alp = s->ald = 0;
        s->want_double = 0; s->reg_hint = 0; s->reg_ok = 0; s->reg = 0; s->f32 = 0;
        s->pcg_total = 0; s->nfactor_f32 = s->nfactor_f64 = 0;
    }
    /* X = I, Zi = I; rhs = A(C); M = A A' (route at X = Zi = I) */
    if (set_identity(S, ns, NULL, OFF(X)) || set_identity(S, ns, NULL, OFF(Zi))) goto out;
    if (aop(S, ns, NULL, OFF(C), OFF(rhs), 1.0, 0.0)) goto out;
    if (schur_assemble(S, ns, NULL, 1)) goto out;
    if (schur_factor(S, ns, NULL, 1, fail)) goto out;
    for (int q = 0; q < ns; q++) if (fail[q]) { S[q]->status = 4; S[q]->done = 1; }
    for (int q = 0; q < ns; q++) act[q] = !S[q]->done;
    if (schur_solve(S, ns, act, OFF(bvec), OFF(dy), 1e-8, rel)) goto out;
    if (schur_solve(S, ns, act, OFF(rhs), OFF(y), 1e-8, rel)) goto out;
    /* X = A*(dy); Z = C - A*(y) */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        lin_add(S[q]->P.T, S[q]->X, -1, 0, -1, 0, -1, 0, -1, 0);
        lin_add(S[q]->P.T, S[q]->Z, S[q]->C, 1.0, -1, 0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (aty(S, ns, act, OFF(dy), OFF(X), 1.0) || aty(S, ns, act, OFF(y), OFF(Z), -1.0)) goto out;
    if (lammin_all(S, ns, act, OFF(X), lx) || lammin_all(S, ns, act, OFF(Z), lz)) goto out;
    MItL_clear(&g_ml);
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        const double ax = -lx[q], az = -lz[q];
        const double shx = ax >= -1e-8 ? 1.0 + ax : 0.0;
        const double shz = fmax(0.0, 1.0 + az);
        g_hdbl[2 * q] = shx; g_hdbl[2 * q + 1] = shz;
        for (int k = 0; k < S[q]->P.nb; k++) {
            MItem a = { S[q]->P.bn[k], 2 * q, S[q]->X + S[q]->P.boff[k], 0, 0 }; MItL_add(&g_ml, a);
            MItem b = { S[q]->P.bn[k], 2 * q + 1, S[q]->Z + S[q]->P.boff[k], 0, 0 }; MItL_add(&g_ml, b);
        }
    }
    if (bk_h2d(ddbl(), g_hdbl, sizeof(double) * 2 * ns)) goto out;
    if (bk_add_diag(g_ml.v, g_ml.n, g_bdbl / 8)) goto out;
    /* Cholesky of X and Z; a solve whose factorisation fails gets 1e-8, 1e-7, ... added to the
     * diagonal of that side */
    {
        int *need = calloc(2 * ns, sizeof(int)), *ok2 = calloc(2 * ns, sizeof(int));
        double *ex = calloc(2 * ns, sizeof(double));
        for (int t = 0; t < 2 * ns; t++) { need[t] = act[t % ns]; ex[t] = 1e-8; }
        for (int tries = 0; tries < 30; tries++) {
            int any = 0;
            for (int t = 0; t < 2 * ns; t++) any |= need[t];
            if (!any) break;
            if (chol_sides(S, ns, 2, need, F_V, F_L, F_D, F_W, ok2)) { free(need); free(ok2); free(ex); goto out; }
            MItL_clear(&g_ml);
            for (int t = 0; t < 2 * ns; t++) {
                if (!need[t]) continue;
                if (ok2[t]) { need[t] = 0; continue; }
                const int sd = t / ns, q = t % ns;
                g_hdbl[t] = ex[t]; ex[t] *= 10;
                for (int k = 0; k < S[q]->P.nb; k++) { MItem a = { S[q]->P.bn[k], t, F_(S[q], F_V[sd]) + S[q]->P.boff[k], 0, 0, 0 }; MItL_add(&g_ml, a); }
            }
            if (g_ml.n && (bk_h2d(ddbl(), g_hdbl, sizeof(double) * 2 * ns) || bk_add_diag(g_ml.v, g_ml.n, g_bdbl / 8))) { free(need); free(ok2); free(ex); goto out; }
        }
        for (int t = 0; t < 2 * ns; t++) if (need[t]) { S[t % ns]->status = 4; S[t % ns]->done = 1; act[t % ns] = 0; }
        free(need); free(ok2); free(ex);
    }
    {
        int *a2 = malloc(sizeof(int) * 2 * ns);
        for (int t = 0; t < 2 * ns; t++) a2[t] = act[t % ns];
        const int e = inv_sides(S, ns, 2, a2, F_L, F_D, F_Li, F_LiT, F_W);
        free(a2);
        if (e) goto out;
    }
    /* R0 = C - Z - A*(y); theta = 1 */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.T, S[q]->R0, S[q]->C, 1.0, S[q]->Z, -1.0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (aty(S, ns, act, OFF(y), OFF(R0), -1.0)) goto out;
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (act[q]) dot_add(S[q]->P.T, S[q]->R0, S[q]->R0, SC(q, 0));
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    for (int q = 0; q < ns; q++) if (act[q]) { S[q]->r0norm = sqrt(HS(q, 0)); S[q]->theta = 1.0; }
    /* best iterate starts at the start point */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        lin_add(S[q]->P.T, S[q]->Xb, S[q]->X, 1.0, -1, 0, -1, 0, -1, 0);
        lin_add(S[q]->P.m, S[q]->yb, S[q]->y, 1.0, -1, 0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    rc = 0;
out:
    free(fail); free(ok); free(act); free(rel); free(lx); free(lz); free(extra);
    return rc;
}

/* ---- one iteration (reference solver.c:14394-15189) --------------------------------------- */
static int ipm_iterate_batch_(Solve **S, int ns);
static int ipm_iterate_hsd_(Solve **S, int ns);
int ipm_iterate_batch(Solve **S, int ns) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    /* the infeasible-start solves and the self-dual ones advance in two batches */
    Solve **a = malloc(sizeof(Solve *) * (ns + 1)), **b = malloc(sizeof(Solve *) * (ns + 1));
    int na = 0, nb = 0;
    for (int q = 0; q < ns; q++) { if (S[q]->hsd) b[nb++] = S[q]; else a[na++] = S[q]; }
    int r_ = 0;
    if (na) r_ = ipm_iterate_batch_(a, na);
    if (!r_ && nb) r_ = ipm_iterate_hsd_(b, nb);
    free(a); free(b);
    if (t_ > 0) bk_prof_host(45, bk_now() - t_);
    return r_;
}
int ipm_iterate_hsd(Solve **S, int ns) { return ipm_iterate_hsd_(S, ns); }
/* ---- the Nesterov-Todd direction (reference nt_scaling, nt_toX / nt_toZ / nt_fromS,
 * solver.c:7813-8068) ---------------------------------------------------------------------- */
/* per block with X = L_X L_X' and Z = L_Z L_Z': L_X' Z L_X = (L_Z' L_X)' (L_Z' L_X) = V diag(lambda^2)
 * V' (the eigenvalues are those of X Z), G = L_X V diag(lambda^-1/2), W = G G' (W Z W = X),
 * G^-1 = diag(lambda^1/2) V' L_X^-1; LP variables (blocks of order 1) come out as W = sqrt(x / z).
 * The Schur complement and the direction's products then take (W, W) in the place of (X, Zi)
 * (the scaling pair SX, SZi; SLX = G, SLZi = G' for the dense route). A solve whose eigenvalues
 * are not all positive keeps HKM for the iteration. */
static int nt_scaling(Solve **S, int ns, const int *act) {
    int *a = calloc(ns, sizeof(int)), any = 0, rc = -1;
    for (int q = 0; q < ns; q++) { a[q] = act[q] && S[q]->nt; any |= a[q]; }
    if (!any) { free(a); return 0; }
    double *H = heapp();
    EigItem *ei = NULL;
    ScaleItem *si = NULL;
    int ne = 0, cap = 0;
    if (blk_gemm(S, ns, a, 1, 0, 0, 1.0, 0.0, OFF(LZ), OFF(LX), OFF(Gn)) || blk_gemm(S, ns, a, 1, 0, 0, 1.0, 0.0, OFF(Gn), OFF(Gn), OFF(ntA)) ||
        blk_sym(S, ns, a, OFF(ntA))) goto out;
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        SAux *A = aux_of(s);
        for (int k = 0; k < s->P.nb; k++) {
            if (ne == cap) { cap = cap ? 2 * cap : 256; ei = realloc(ei, sizeof(EigItem) * cap); si = realloc(si, sizeof(ScaleItem) * cap); }
            const int64_t o = s->P.boff[k];
            EigItem e = { s->P.bn[k], 0, s->ntA + o, s->ntV + o, s->lam + A->lzoff[k] };
            ScaleItem c = { s->P.bn[k], 3, s->lam + A->lzoff[k], s->lam + (int64_t)s->P.Ntot + 3 * A->lzoff[k] };
            ei[ne] = e; si[ne] = c; ne++;
        }
    }
    if (bk_syev_f64(ei, ne, H) || bk_nt_scale(si, ne, H)) goto out;
    /* an eigenvalue <= 0 leaves a NaN among the solve's lambda^-1/2 */
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (a[q]) { const int64_t nt = (int64_t)S[q]->P.Ntot; dot_add(3 * nt, S[q]->lam + nt, S[q]->lam + nt, SC(q, 0)); }
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    any = 0;
    for (int q = 0; q < ns; q++) if (a[q]) { if (!isfinite(HS(q, 0))) a[q] = 0; else any = 1; }
    if (!any) { rc = 0; goto out; }
    /* G = L_X V diag(lambda^-1/2); W = G G'; G'; G^-1 = diag(lambda^1/2) V' L_X^-1 */
    if (blk_gemm(S, ns, a, 0, 0, 4, 1.0, 0.0, OFF(LX), OFF(ntV), OFF(Gn)) || blk_gemm(S, ns, a, 1, 0, 8, 1.0, 0.0, OFF(ntV), OFF(LXi), OFF(Gi))) goto out;
    ne = 0;
    MItL_clear(&g_ml);
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        SAux *A = aux_of(s);
        for (int k = 0; k < s->P.nb; k++) {
            if (ne + 2 > cap) { cap = 2 * cap + 2; ei = realloc(ei, sizeof(EigItem) * cap); si = realloc(si, sizeof(ScaleItem) * cap); }
            const int n = s->P.bn[k];
            const int64_t o = s->P.boff[k], u = s->lam + (int64_t)s->P.Ntot + 3 * A->lzoff[k];
            ScaleItem c0 = { n, 0, s->Gn + o, u + n }, c1 = { n, 1, s->Gi + o, u + 2 * n };
            si[ne++] = c0; si[ne++] = c1;
            MItem it = { n, 0, s->Gn + o, 0, 0, s->GnT + o };
            MItL_add(&g_ml, it);
        }
    }
    if (bk_nt_scale(si, ne, H) || bk_transpose_f64(g_ml.v, g_ml.n)) goto out;
    if (blk_gemm(S, ns, a, 0, 1, 0, 1.0, 0.0, OFF(Gn), OFF(Gn), OFF(Wn)) || blk_sym(S, ns, a, OFF(Wn))) goto out;
    for (int q = 0; q < ns; q++) if (a[q]) { Solve *s = S[q]; s->SX = s->Wn; s->SZi = s->Wn; s->SLX = s->Gn; s->SLZi = s->GnT; }
    rc = 0;
out:
    free(a); free(ei); free(si);
    return rc;
}
/* the corrector's second-order term in the Nesterov-Todd frame (reference solver.c:9518-9524,
 * 7990-8068): dXt = G^-1 dX G^-T, dZt = G' dZ G, R = sym(dXt dZt) with R(i, j) scaled by
 * 2 / (lambda_i + lambda_j) (the inverse of the Lyapunov operator of diag(lambda)), Q = G R G' */
static int nt_corr(Solve **S, int ns, const int *a) {
    double *H = heapp();
    if (blk_gemm(S, ns, a, 0, 0, 0, 1.0, 0.0, OFF(Gi), OFF(dX), OFF(W1)) || blk_gemm(S, ns, a, 0, 1, 0, 1.0, 0.0, OFF(W1), OFF(Gi), OFF(W2)) ||
        blk_gemm(S, ns, a, 0, 0, 0, 1.0, 0.0, OFF(GnT), OFF(dZ), OFF(W1)) || blk_gemm(S, ns, a, 0, 0, 0, 1.0, 0.0, OFF(W1), OFF(Gn), OFF(F)) ||
        blk_gemm(S, ns, a, 0, 0, 0, 1.0, 0.0, OFF(W2), OFF(F), OFF(Q)) || blk_sym(S, ns, a, OFF(Q))) return -1;
    ScaleItem *si = NULL;
    int ne = 0, cap = 0;
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        SAux *A = aux_of(s);
        for (int k = 0; k < s->P.nb; k++) {
            if (ne == cap) { cap = cap ? 2 * cap : 256; si = realloc(si, sizeof(ScaleItem) * cap); }
            ScaleItem c = { s->P.bn[k], 2, s->Q + s->P.boff[k], s->lam + (int64_t)s->P.Ntot + 3 * A->lzoff[k] };
            si[ne++] = c;
        }
    }
    const int rc = bk_nt_scale(si, ne, H);
    free(si);
    if (rc) return -1;
    if (blk_gemm(S, ns, a, 0, 0, 0, 1.0, 0.0, OFF(Gn), OFF(Q), OFF(W1)) || blk_gemm(S, ns, a, 0, 1, 0, 1.0, 0.0, OFF(W1), OFF(Gn), OFF(Q))) return -1;
    return 0;
}

static int ipm_iterate_batch_(Solve **S, int ns) {
    int rc = -1;
    int *act = calloc(ns, sizeof(int)), *fail = calloc(ns, sizeof(int)), *okp = calloc(ns, sizeof(int)), *okd = calloc(ns, sizeof(int));
    int *hasH = calloc(ns, sizeof(int)), *ract = calloc(ns, sizeof(int)), *tact = calloc(ns, sizeof(int));
    double *rel = calloc(ns, sizeof(double)), *ap = calloc(ns, sizeof(double)), *ad = calloc(ns, sizeof(double));
    double *smu = calloc(ns, sizeof(double)), *gam = calloc(ns, sizeof(double)), *apr = calloc(ns, sizeof(double)), *adr = calloc(ns, sizeof(double));
    double *ne = calloc(ns, sizeof(double)), *nrp = calloc(ns, sizeof(double)), *stepped = calloc(ns, sizeof(double));
    for (int q = 0; q < ns; q++) act[q] = !S[q]->done;
    int nsl = 1;
    for (int q = 0; q < ns; q++) nsl += S[q]->P.nb;
    if (batch_scratch(ns, nsl + ns)) goto out;

    /* Zi = LZi' LZi; the Nesterov-Todd scaling where that direction is taken */
    if (form_zi(S, ns, act)) goto out;
    if (nt_scaling(S, ns, act)) goto out;
    /* residuals and objectives (solver.c:14426-14482) */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        lin_add(S[q]->P.m, S[q]->rp, S[q]->bvec, 1.0, -1, 0, -1, 0, -1, 0);
        lin_add(S[q]->P.T, S[q]->F, S[q]->Z, 1.0, S[q]->C, -1.0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (aop(S, ns, act, OFF(X), OFF(rp), -1.0, 1.0)) goto out;
    if (aty(S, ns, act, OFF(y), OFF(F), 1.0)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        lin_add(S[q]->P.T, S[q]->W1, S[q]->F, 1.0, S[q]->C, 1.0, -1, 0, -1, 0);                 /* A'y + Z */
        lin_add(S[q]->P.T, S[q]->W2, S[q]->F, 1.0, S[q]->R0, S[q]->theta, -1, 0, -1, 0);        /* F + theta R0 */
    }
    /* t1 = du o rp (the residual on the data as read), t2 = b - rp = A(X) */
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.m, S[q]->t2, S[q]->bvec, 1.0, S[q]->rp, -1.0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.m, S[q]->t1, S[q]->rp, 1.0, S[q]->duv, 0, -1, 0, -1, 0);
    if (bk_vmul(g_ll.v, g_ll.n)) goto out;
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (act[q]) {
        Solve *s = S[q];
        const int64_t T = s->P.T;
        dot_add(T, s->W1, s->W1, SC(q, 0));
        dot_add(T, s->C, s->X, SC(q, 1));
        dot_add(T, s->X, s->Z, SC(q, 2));
        dot_add(T, s->F, s->F, SC(q, 3));
        dot_add(T, s->W2, s->W2, SC(q, 4));
        dot_add(s->P.m, s->bvec, s->y, SC(q, 5));
        dot_add(s->P.m, s->t1, s->t1, SC(q, 6));
        dot_add(s->P.m, s->t2, s->t2, SC(q, 7));
    }
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        Solve *s = S[q];
        const HProb *P = &s->P;
        const double ncr2 = HS(q, 0), pobj = HS(q, 1), xz = HS(q, 2), rdn2 = HS(q, 3), dev2 = HS(q, 4);
        const double dobj = HS(q, 5), rpo = sqrt(HS(q, 6)), nAX = sqrt(HS(q, 7));
        const double mu = xz / P->Ntot;
        const double rdo = P->cs * sqrt(rdn2), sc = P->bs * P->cs;
        const double po = sc * pobj, dob = sc * dobj;
        s->pinf = rpo / (1 + P->normb2);
        s->dinf = rdo / (1 + P->normC2);
        const double oden = 1 + fabs(po) + fabs(dob);
        s->relgap = fabs(po - dob) / oden;
        s->relcomp = sc * xz / oden;
        s->pobj = po; s->dobj = dob; s->mu = mu;
        double score = iterate_score(s->relgap, s->relcomp, s->pinf, s->dinf);
        if (!isfinite(score)) score = INFINITY;
        s->score = score;
        const int it = s->it;
        if (it < 128) { s->tr_pobj[it] = po; s->tr_dobj[it] = dob; s->tr_mu[it] = sc * mu; }
        if (s->o.verbose)
            solve_log("%4d %+.10e %+.10e %.1e %.1e %.1e %.1e %.3f %.3f %s %.3fs\n", it, po, dob, s->pinf, s->dinf, s->relgap, sc * mu,
                      s->alp, s->ald, s->f32 ? "f32" : "f64", bk_now() - s->t_start);
        int st = -1;
        if (!isfinite(po) || !isfinite(dob)) st = 4;
        else {
            if (score < s->best_score) {
                s->best_score = score; s->it_best = it;
                LinL_clear(&g_ll);
                lin_add(P->T, s->Xb, s->X, 1.0, -1, 0, -1, 0, -1, 0);
                lin_add(P->m, s->yb, s->y, 1.0, -1, 0, -1, 0, -1, 0);
                if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
            }
            if (score < s->o.tol) st = 0;
            else if (dobj > 0 && sqrt(ncr2) / dobj < 1e-8 && it > 5) st = 1;
            else if (pobj < 0 && nAX / (-pobj) < 1e-8 && it > 5) st = 2;
            else if (it >= s->o.maxit) st = 3;
            else if (s->stall >= 3) st = 4;
            else if (it - s->it_best >= 5 && s->best_score < 1e-6) st = 4;
            else if (it - s->it_best >= 8 && s->best_score < 1e-4) st = 4;
            else if (it - s->it_best >= 12) st = 4;
            else if (s->o.time_limit > 0 && bk_now() - s->t_start > s->o.time_limit) st = 6;
        }
        if (st >= 0) { s->status = st; s->done = 1; act[q] = 0; continue; }
        /* re-base the dual residual (solver.c:14614-14636) */
        const double dev = sqrt(dev2);
        if (dev > 1e-10 * (1.0 + s->theta * s->r0norm) && dev > 1e-12) {
            s->theta = 1.0; s->r0norm = sqrt(rdn2);
            LinL_clear(&g_ll);
            lin_add(P->T, s->R0, s->F, -1.0, -1, 0, -1, 0, -1, 0);
            if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
        }
        if (s->theta * s->r0norm <= 1e-13) s->theta = 0.0;
        hasH[q] = s->theta > 0;
        /* mixed precision: the f32 factor only far from the optimum */
        if (s->o.prec == PREC_MIXED && fmax(s->relgap, fmax(s->pinf, s->dinf)) <= s->o.mixed_until) s->want_double = 1;
    }
    {
        int any = 0;
        for (int q = 0; q < ns; q++) any |= act[q];
        if (!any) { rc = 0; goto out; }
    }
    /* Schur complement and its factor */
    if (schur_assemble(S, ns, act, 0)) goto out;
    if (schur_factor(S, ns, act, 0, fail)) goto out;
    for (int q = 0; q < ns; q++) if (act[q] && fail[q]) { S[q]->status = 4; S[q]->done = 1; act[q] = 0; }
    /* a0 = A(Zi); h = A(Zi theta R0 X) */
    if (aop(S, ns, act, OFF(Zi), OFF(a0), 1.0, 0.0)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        if (hasH[q]) lin_add(S[q]->P.T, S[q]->dZ, S[q]->R0, S[q]->theta, -1, 0, -1, 0, -1, 0);
        else lin_add(S[q]->P.m, S[q]->h, -1, 0, -1, 0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (zi_v_x_(S, ns, hasH, OFF(dZ), OFF(SX), OFF(G), 1)) goto out;
    if (aop(S, ns, hasH, OFF(G), OFF(h), 1.0, 0.0)) goto out;
    /* predictor: M dy = b + h (tol 1e-7); dZ = theta R0 - A*(dy); dX = -X - sym(Zi dZ X) */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        lin_add(S[q]->P.m, S[q]->rhs, S[q]->bvec, 1.0, S[q]->h, 1.0, -1, 0, -1, 0);
        lin_add(S[q]->P.T, S[q]->dZ, S[q]->R0, S[q]->theta, -1, 0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (schur_solve(S, ns, act, OFF(rhs), OFF(dy), 1e-7, rel)) goto out;
    if (aty(S, ns, act, OFF(dy), OFF(dZ), -1.0)) goto out;
    if (zi_v_x(S, ns, act, OFF(dZ), OFF(SX), OFF(G))) goto out;
    if (blk_sym(S, ns, act, OFF(G))) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.T, S[q]->dX, S[q]->G, -1.0, S[q]->X, -1.0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (maxstep2(S, ns, act, ap, ad)) goto out;
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (act[q]) {
        Solve *s = S[q];
        dot_add(s->P.T, s->X, s->dZ, SC(q, 0));
        dot_add(s->P.T, s->dX, s->Z, SC(q, 1));
        dot_add(s->P.T, s->dX, s->dZ, SC(q, 2));
        dot_add(s->P.T, s->X, s->Z, SC(q, 3));
    }
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        Solve *s = S[q];
        const double a_p = fmin(1.0, ap[q]), a_d = fmin(1.0, ad[q]);
        const double xdz = HS(q, 0), dxz = HS(q, 1), dxdz = HS(q, 2), xz = HS(q, 3);
        const double mu = xz / s->P.Ntot;
        const double mu_a = (xz + a_p * dxz + a_d * xdz + a_p * a_d * dxdz) / s->P.Ntot;
        const double frac = fmax(0.0, mu_a / mu), mina = fmin(a_p, a_d);
        const double expon = !s->P.any_sdp ? 3.0 : (s->relgap > 1e-6 ? fmax(1.0, 3 * mina * mina) : 3.0);
        const double sigma = fmin(1.0, pow(frac, expon));
        gam[q] = 0.9 + (0.99 - 0.9) * mina;
        smu[q] = sigma * mu;
    }
    /* corrector (solver.c:14997-15055): Q = Zi dZ dX (W1 = Zi dZ still holds from the predictor
     * in the blocks where dZ is not diagonal); rhs = b - smu a0 + A(Q) + h */
    {
        int *hk = calloc(ns, sizeof(int)), *nt = calloc(ns, sizeof(int)), anyh = 0, anyn = 0;
        for (int q = 0; q < ns; q++) if (act[q]) { if (S[q]->SX == S[q]->Wn && S[q]->nt) { nt[q] = 1; anyn = 1; } else { hk[q] = 1; anyh = 1; } }
        const int e = (anyh && zi_dz_dx(S, ns, hk)) || (anyn && nt_corr(S, ns, nt));
        free(hk); free(nt);
        if (e) goto out;
    }
    if (aop(S, ns, act, OFF(Q), OFF(qv), 1.0, 0.0)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        lin_add(S[q]->P.m, S[q]->rhs, S[q]->bvec, 1.0, S[q]->a0, -smu[q], S[q]->qv, 1.0, S[q]->h, 1.0);
        lin_add(S[q]->P.T, S[q]->dZ, S[q]->R0, S[q]->theta, -1, 0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (schur_solve(S, ns, act, OFF(rhs), OFF(dy), 0.0, rel)) goto out;
    if (aty(S, ns, act, OFF(dy), OFF(dZ), -1.0)) goto out;
    if (zi_v_x(S, ns, act, OFF(dZ), OFF(SX), OFF(G))) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.T, S[q]->dX, S[q]->G, -1.0, S[q]->Q, -1.0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (blk_sym(S, ns, act, OFF(dX))) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.T, S[q]->dX, S[q]->dX, 1.0, S[q]->Zi, smu[q], S[q]->X, -1.0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    /* refinement enforcing A(dX) = rp (solver.c:15058-15105) */
    for (int q = 0; q < ns; q++) ract[q] = act[q];
    for (int pass = 0; pass < 3; pass++) {
        int any = 0;
        for (int q = 0; q < ns; q++) any |= ract[q];
        if (!any) break;
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (ract[q]) lin_add(S[q]->P.m, S[q]->ev, S[q]->rp, 1.0, -1, 0, -1, 0, -1, 0);
        if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
        if (aop(S, ns, ract, OFF(dX), OFF(ev), -1.0, 1.0)) goto out;
        DotL_clear(&g_dl);
        for (int q = 0; q < ns; q++) if (ract[q]) { dot_add(S[q]->P.m, S[q]->ev, S[q]->ev, SC(q, 0)); dot_add(S[q]->P.m, S[q]->rp, S[q]->rp, SC(q, 1)); }
        if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
        for (int q = 0; q < ns; q++) {
            if (!ract[q]) continue;
            ne[q] = sqrt(HS(q, 0)); nrp[q] = sqrt(HS(q, 1));
            if (!(ne[q] > 1e-2 * nrp[q] && ne[q] > 1e-3 * S[q]->o.tol * (1 + S[q]->P.nbs))) ract[q] = 0;
        }
        any = 0;
        for (int q = 0; q < ns; q++) any |= ract[q];
        if (!any) break;
        if (schur_solve(S, ns, ract, OFF(ev), OFF(wv), 0.0, rel)) goto out;
        /* Q = A*(wv); G = sym(Zi Q X); qv = ev - A(G) */
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (ract[q]) {
            lin_add(S[q]->P.T, S[q]->Q, -1, 0, -1, 0, -1, 0, -1, 0);
            lin_add(S[q]->P.m, S[q]->qv, S[q]->ev, 1.0, -1, 0, -1, 0, -1, 0);
        }
        if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
        if (aty(S, ns, ract, OFF(wv), OFF(Q), 1.0)) goto out;
        if (zi_v_x(S, ns, ract, OFF(Q), OFF(SX), OFF(G))) goto out;
        if (blk_sym(S, ns, ract, OFF(G))) goto out;
        if (aop(S, ns, ract, OFF(G), OFF(qv), -1.0, 1.0)) goto out;
        DotL_clear(&g_dl);
        for (int q = 0; q < ns; q++) if (ract[q]) dot_add(S[q]->P.m, S[q]->qv, S[q]->qv, SC(q, 0));
        if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) {
            if (!ract[q]) continue;
            const double ne2 = sqrt(HS(q, 0));
            if (!(ne2 < 0.5 * ne[q])) { ract[q] = 0; continue; }
            Solve *s = S[q];
            lin_add(s->P.m, s->dy, s->dy, 1.0, s->wv, 1.0, -1, 0, -1, 0);
            lin_add(s->P.T, s->dZ, s->dZ, 1.0, s->Q, -1.0, -1, 0, -1, 0);
            lin_add(s->P.T, s->dX, s->dX, 1.0, s->G, 1.0, -1, 0, -1, 0);
        }
        if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    }
    /* step lengths (solver.c:15108-15130) */
    if (maxstep2(S, ns, act, apr, adr)) goto out;
    for (int q = 0; q < ns; q++) if (act[q]) { ap[q] = fmin(1.0, gam[q] * fmin(1.0, apr[q])); ad[q] = fmin(1.0, gam[q] * fmin(1.0, adr[q])); }
    /* catastrophe guard (solver.c:15143-15160) */
    {
        int *g = calloc(ns, sizeof(int)), any = 0;
        for (int q = 0; q < ns; q++) if (act[q] && S[q]->pinf < 1e-6) { g[q] = 1; any = 1; }
        if (any) {
            LinL_clear(&g_ll);
            for (int q = 0; q < ns; q++) if (g[q]) lin_add(S[q]->P.m, S[q]->ev, S[q]->rp, 1.0, -1, 0, -1, 0, -1, 0);
            if (bk_lincomb(g_ll.v, g_ll.n) || aop(S, ns, g, OFF(dX), OFF(ev), -1.0, 1.0)) { free(g); goto out; }
            DotL_clear(&g_dl);
            for (int q = 0; q < ns; q++) if (g[q]) { dot_add(S[q]->P.m, S[q]->ev, S[q]->ev, SC(q, 0)); dot_add(S[q]->P.m, S[q]->rp, S[q]->rp, SC(q, 1)); }
            if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) { free(g); goto out; }
            for (int q = 0; q < ns; q++) {
                if (!g[q]) continue;
                const double nev = sqrt(HS(q, 0)), nr = sqrt(HS(q, 1));
                const double gg = fmax(1e3, 1e-6 / fmax(S[q]->pinf, 1e-300)), lim = gg * nr;
                if (nev > lim && nr > 0) { const double cap = lim / nev; if (cap < ap[q]) ap[q] = cap; if (S[q]->o.prec == PREC_MIXED) S[q]->want_double = 1; }
            }
        }
        free(g);
    }
    /* try_step: V += (a - stepped) dV, L = chol(V); x0.8 on failure (solver.c:6393-6446); the
     * primal and dual sides factor in the same launches */
    {
        int *ta = calloc(2 * ns, sizeof(int)), *okv = calloc(2 * ns, sizeof(int)), *ok = calloc(2 * ns, sizeof(int));
        double *st = calloc(2 * ns, sizeof(double));
        for (int t = 0; t < 2 * ns; t++) ta[t] = act[t % ns];
        for (int tries = 0; tries < 40; tries++) {
            int any = 0;
            for (int t = 0; t < 2 * ns; t++) any |= ta[t];
            if (!any) break;
            LinL_clear(&g_ll);
            for (int t = 0; t < 2 * ns; t++) {
                if (!ta[t]) continue;
                const int sd = t / ns, q = t % ns;
                double *a = sd ? &ad[q] : &ap[q];
                const double inc = *a - st[t];
                st[t] = *a;
                lin_add(S[q]->P.T, F_(S[q], F_V[sd]), F_(S[q], F_V[sd]), 1.0, F_(S[q], F_dV[sd]), inc, -1, 0, -1, 0);
            }
            if (bk_lincomb(g_ll.v, g_ll.n)) { free(ta); free(okv); free(ok); free(st); goto out; }
            if (chol_sides(S, ns, 2, ta, F_V, F_L, F_D, F_W, ok)) { free(ta); free(okv); free(ok); free(st); goto out; }
            for (int t = 0; t < 2 * ns; t++) {
                if (!ta[t]) continue;
                if (ok[t]) { ta[t] = 0; okv[t] = 1; }
                else { if (t / ns) ad[t % ns] *= 0.8; else ap[t % ns] *= 0.8; }
            }
        }
        for (int q = 0; q < ns; q++) { okp[q] = okv[q]; okd[q] = okv[ns + q]; }
        const int e = inv_sides(S, ns, 2, okv, F_L, F_D, F_Li, F_LiT, F_W);
        free(ta); free(okv); free(ok); free(st);
        if (e) goto out;
    }
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        Solve *s = S[q];
        if (!okp[q] || !okd[q]) { s->status = 4; s->done = 1; continue; }
        lin_add(s->P.m, s->y, s->y, 1.0, s->dy, ad[q], -1, 0, -1, 0);
        s->theta *= (1.0 - ad[q]);
        s->alp = ap[q]; s->ald = ad[q];
        s->stall = (ap[q] < 1e-6 && ad[q] < 1e-6) ? s->stall + 1 : 0;
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    for (int q = 0; q < ns; q++) if (act[q] && !S[q]->done) S[q]->it++;
    rc = 0;
out:
    for (int q = 0; q < ns; q++) { S[q]->SX = S[q]->X; S[q]->SZi = S[q]->Zi; S[q]->SLX = S[q]->LX; S[q]->SLZi = S[q]->LZi; }
    free(act); free(fail); free(okp); free(okd); free(hasH); free(ract); free(tact); free(rel); free(ap); free(ad); free(smu); free(gam);
    free(apr); free(adr); free(ne); free(nrp); free(stepped);
    return rc;
}

/* ---- DIMACS errors of the returned point on the data as read (the measure of postsolve,
 * postsolve.c:2744-2897), FP64 on the device. With X = bs X~, y_i = cs d_i y~_i and Z = C - A'y = cs (C~ - A~'y~):
 *   A(X) - b = du o (A~(X~) - b~), <C,X> = bs cs <C~,X~>, b'y = bs cs b~'y~, <X,Z> = bs cs <X~, Z~>.
 * lambda_min(X) and lambda_min(Z) are tested by Cholesky: a factor that exists proves the block
 * positive definite (error 0); otherwise the Lanczos bound gives the violation. */
static int dimacs_point(Solve **S, int ns);
int ipm_dimacs(Solve **S, int ns) {
    /* the returned point into X, y (the best iterate unless the run ended optimal) */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (S[q]->status != 0) {
        lin_add(S[q]->P.T, S[q]->X, S[q]->Xb, 1.0, -1, 0, -1, 0, -1, 0);
        lin_add(S[q]->P.m, S[q]->y, S[q]->yb, 1.0, -1, 0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) return -1;
    return dimacs_point(S, ns);
}
/* the DIMACS errors, objectives and status of the point (X, y) as it stands */
static int dimacs_point(Solve **S, int ns) {
    int rc = -1;
    int *act = calloc(ns, sizeof(int)), *okx = calloc(ns, sizeof(int)), *okz = calloc(ns, sizeof(int));
    double *lam = calloc(ns, sizeof(double));
    int nsl = 1;
    for (int q = 0; q < ns; q++) { act[q] = 1; nsl += S[q]->P.nb; }
    if (batch_scratch(ns, nsl + ns)) goto out;
    /* rp = A(X) - b (scaled); W1 = C - A'y (scaled Z on the data) */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) {
        lin_add(S[q]->P.m, S[q]->rp, S[q]->bvec, -1.0, -1, 0, -1, 0, -1, 0);
        lin_add(S[q]->P.T, S[q]->W1, S[q]->C, 1.0, -1, 0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (aop(S, ns, act, OFF(X), OFF(rp), 1.0, 1.0) || aty(S, ns, act, OFF(y), OFF(W1), -1.0)) goto out;
    /* t1 = du o rp: the residual on the data as read, its squared norm among the dots */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) lin_add(S[q]->P.m, S[q]->t1, S[q]->rp, 1.0, S[q]->duv, 0, -1, 0, -1, 0);
    if (bk_vmul(g_ll.v, g_ll.n)) goto out;
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) {
        Solve *s = S[q];
        dot_add(s->P.T, s->C, s->X, SC(q, 0));
        dot_add(s->P.T, s->X, s->W1, SC(q, 1));
        dot_add(s->P.m, s->bvec, s->y, SC(q, 2));
        dot_add(s->P.m, s->t1, s->t1, SC(q, 3));
    }
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    for (int q = 0; q < ns; q++) {
        Solve *s = S[q];
        const HProb *P = &s->P;
        const double sc = P->bs * P->cs;
        s->res_pobj = sc * HS(q, 0);
        s->res_dobj = sc * HS(q, 2);
        const double xz = sc * HS(q, 1);
        const double r2 = HS(q, 3);
        const double den = 1 + fabs(s->res_pobj) + fabs(s->res_dobj);
        s->res_err[0] = sqrt(r2) / (1 + P->normb1);
        s->res_err[1] = 0; s->res_err[2] = 0;
        s->res_err[3] = 0;
        s->res_err[4] = (s->res_pobj - s->res_dobj) / den;
        s->res_err[5] = xz / den;
    }
    /* cone membership: X (its factor is recomputed; the iterate's may be stale for the best point)
     * and Z on the data as read (W1) */
    if (blk_chol(S, ns, act, OFF(X), OFF(LX), OFF(DinvX), okx)) goto out;
    if (blk_chol(S, ns, act, OFF(W1), OFF(LZ), OFF(DinvZ), okz)) goto out;
    /* a side without a factor: lambda_min >= -h when V + h I factors. h = delta = 1e-12 (1 + |b|_1)
     * for X, 1e-12 (1 + |C|_1) for Z, then decades of delta up to 1e9 delta, three bisections in
     * the exponent (a factor 1.33, as ps_lammin_tol's ladder, postsolve.c:668-726) and bisections
     * of the interval until it is within 1e-4 of h: the reported -lambda_min is the smallest
     * shift that factored, within 1e-4 relative of the eigenvalue the reference computes for
     * blocks of order <= 500 (postsolve.c:685); beyond 1e9 delta (an error past 1e-3, a verdict
     * only) the Lanczos estimate */
    for (int side = 0; side < 2; side++) {
        int *ok = side ? okz : okx, *need = calloc(ns, sizeof(int)), *ok2 = calloc(ns, sizeof(int)), any = 0;
        double *dl = calloc(ns, sizeof(double)), *lo = calloc(ns, sizeof(double)), *hi = calloc(ns, sizeof(double));
        const size_t fV = side ? OFF(W1) : OFF(X), fL = side ? OFF(LZ) : OFF(LX), fD = side ? OFF(DinvZ) : OFF(DinvX);
        for (int q = 0; q < ns; q++) if (!ok[q]) {
            const HProb *P = &S[q]->P;
            dl[q] = side ? 1e-12 * (1 + P->normC1) / P->cs : 1e-12 * (1 + P->normb1) / P->bs;   /* scaled data */
            lo[q] = 0; hi[q] = -1; need[q] = 1; any = 1;
        }
        /* the ladder: h = 10^k delta, k = 0..9, for the solves not yet bounded */
        for (int k = 0; any && k <= 9; k++) {
            for (int q = 0; q < ns; q++) g_hdbl[q] = need[q] ? dl[q] * pow(10.0, k) : 0;
            if (shifted_chol(S, ns, need, fV, fL, fD, ok2)) { free(need); free(ok2); free(dl); free(lo); free(hi); goto out; }
            any = 0;
            for (int q = 0; q < ns; q++) if (need[q]) {
                if (ok2[q]) { need[q] = 0; hi[q] = g_hdbl[q]; lo[q] = k ? hi[q] / 10 : 0; }
                else any = 1;
            }
        }
        /* three bisections in the exponent, then of the interval, for the solves bounded past
         * delta (at most 16 more factorizations) */
        int *bis = calloc(ns, sizeof(int)), nb = 0;
        for (int q = 0; q < ns; q++) if (hi[q] > 0 && lo[q] > 0) { bis[q] = 1; nb++; }
        for (int b = 0; nb && b < 19; b++) {
            for (int q = 0; q < ns; q++) g_hdbl[q] = bis[q] ? (b < 3 ? sqrt(lo[q] * hi[q]) : 0.5 * (lo[q] + hi[q])) : 0;
            if (shifted_chol(S, ns, bis, fV, fL, fD, ok2)) { free(bis); free(need); free(ok2); free(dl); free(lo); free(hi); goto out; }
            nb = 0;
            for (int q = 0; q < ns; q++) {
                if (!bis[q]) continue;
                if (ok2[q]) hi[q] = g_hdbl[q]; else lo[q] = g_hdbl[q];
                if (b >= 2 && hi[q] - lo[q] <= 1e-4 * hi[q]) bis[q] = 0;
                else nb++;
            }
        }
        free(bis);
        for (int q = 0; q < ns; q++) if (!ok[q] && hi[q] > 0) {
            const HProb *P = &S[q]->P;
            if (side) S[q]->res_err[3] = P->cs * hi[q] / (1 + P->normC1);
            else S[q]->res_err[1] = P->bs * hi[q] / (1 + P->normb1);
        }
        if (any) {
            if (lammin_all(S, ns, need, fV, lam)) { free(need); free(ok2); free(dl); free(lo); free(hi); goto out; }
            for (int q = 0; q < ns; q++) if (need[q]) {
                const HProb *P = &S[q]->P;
                const double e = nmax(0.0, -(side ? P->cs : P->bs) * lam[q]) / (1 + (side ? P->normC1 : P->normb1));
                S[q]->res_err[side ? 3 : 1] = nmax(e, (side ? P->cs : P->bs) * dl[q] * 1e9 / (1 + (side ? P->normC1 : P->normb1)));
            }
        }
        free(need); free(ok2); free(dl); free(lo); free(hi);
    }
    /* status from the measured errors (main.c:2349-2359, brisk.h:187-189) */
    for (int q = 0; q < ns; q++) S[q]->status = status_from_errors(S[q]->status, S[q]->res_err, S[q]->o.tol);
    rc = 0;
out:
    free(act); free(okx); free(okz); free(lam);
    return rc;
}

/* ---- end: the returned point on the data as read ------------------------------------------ */
/* X = bs X~ and Z = cs (C~ - A~'y~) (left in W1 by the DIMACS evaluation) are scaled on the device
 * into G and Q (free after the evaluation) and downloaded through the pinned staging buffers:
 * straight into the result arrays when the problem is one SDP block (the arena's layout is the
 * result's), else into a host copy that is packed per input block. y_i = cs d_i y~_i on the host. */
/* the results of the solves to the host: X = bs X~, Z = cs W1 (the measure's Z on the data as
 * read), y_i = cs d_i y~_i; one linear-combination launch and one batch of reads for all */
int ipm_finish_batch(Solve **S, int ns) {
    if (ns <= 0) return 0;
    double *H = heapp();
    int rc = -1;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) {
        Solve *s = S[q];
        lin_add(s->P.T, s->G, s->X, s->P.bs, -1, 0, -1, 0, -1, 0);
        lin_add(s->P.T, s->Q, s->W1, s->P.cs, -1, 0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) return -1;
    void **dst = malloc(sizeof(void *) * 3 * ns);
    const void **src = malloc(sizeof(void *) * 3 * ns);
    size_t *by = malloc(sizeof(size_t) * 3 * ns);
    double **tx = calloc(ns, sizeof(double *)), **tz = calloc(ns, sizeof(double *));
    for (int q = 0; q < ns; q++) {
        Solve *s = S[q];
        const HProb *P = &s->P;
        int64_t tot = 0;
        for (int k = 0; k < P->nblk_in; k++) tot += P->bs_in[k] > 0 ? (int64_t)P->bs_in[k] * P->bs_in[k] : -P->bs_in[k];
        free(s->res_X); free(s->res_Z); free(s->res_y);
        s->res_X = malloc(sizeof(double) * (tot + 1));
        s->res_Z = malloc(sizeof(double) * (tot + 1));
        s->res_y = malloc(sizeof(double) * (P->m + 1));
        /* one SDP block: read straight into the results; else through the block arena */
        const int direct = P->nblk_in == 1 && P->bs_in[0] > 0 && tot == P->T;
        if (!direct) { tx[q] = malloc(sizeof(double) * (P->T + 1)); tz[q] = malloc(sizeof(double) * (P->T + 1)); }
        dst[3 * q] = s->res_y; src[3 * q] = H + s->y; by[3 * q] = sizeof(double) * P->m;
        dst[3 * q + 1] = direct ? s->res_X : tx[q]; src[3 * q + 1] = H + s->G; by[3 * q + 1] = sizeof(double) * P->T;
        dst[3 * q + 2] = direct ? s->res_Z : tz[q]; src[3 * q + 2] = H + s->Q; by[3 * q + 2] = sizeof(double) * P->T;
    }
    if (bk_d2h_many(3 * ns, dst, src, by)) goto out;
    for (int q = 0; q < ns; q++) {
        Solve *s = S[q];
        const HProb *P = &s->P;
        for (int i = 0; i < P->m; i++) s->res_y[i] *= P->cs * P->d[i];
        if (tx[q]) {
            int64_t o = 0;
            for (int k = 0; k < P->nblk_in; k++) {
                const int kb = P->in_first[k];
                if (P->bs_in[k] > 0) {
                    const int64_t len = (int64_t)P->bs_in[k] * P->bs_in[k];
                    memcpy(s->res_X + o, tx[q] + P->boff[kb], sizeof(double) * len);
                    memcpy(s->res_Z + o, tz[q] + P->boff[kb], sizeof(double) * len);
                    o += len;
                } else {
                    for (int i = 0; i < -P->bs_in[k]; i++) { s->res_X[o + i] = tx[q][P->boff[kb + i]]; s->res_Z[o + i] = tz[q][P->boff[kb + i]]; }
                    o += -P->bs_in[k];
                }
            }
        }
        s->t_solve = bk_now() - s->t_start;
    }
    rc = 0;
out:
    for (int q = 0; q < ns; q++) { free(tx[q]); free(tz[q]); }
    free(tx); free(tz); free(dst); free(src); free(by);
    return rc;
}
int ipm_finish(Solve *s) { return ipm_finish_batch(&s, 1); }

/* ---- the homogeneous self-dual embedding (reference hsd_run, solver.c:9803-12211) ------------
 * Variables X, y, Z, tau, kappa with rp = tau b - A(X), Rd = tau C - A*(y) - Z, rg = kappa - b'y +
 * <C, X>, mu = (<X, Z> + tau kappa) / (nu + 1) (solver.c:10357). Each iteration: the Schur factor at
 * (X, Zi); v = M^-1 (b + A(X C Zi)) for the tau column; the affine direction (sigma = 0) and its
 * step; the corrector family at sigma = 0 and 1 with the second-order term of the affine direction
 * (affine in sigma, solver.c:11356-11375); sigma chosen among sigma_m, 0.2 sigma_m and 3 sigma_m by
 * the predicted mu reduction alpha (1 - sigma) (solver.c:11195-11228); one step length for all
 * variables; dkappa = (sigma mu - tau kappa - kappa dtau) / tau (solver.c:11127); tau normalised
 * to 1 when it passes 10 (solver.c:12054-12069). The solution is (X, y) / tau. */
/* ---- the self-dual embedding's direction family ------------------------------------------- */
/* per-solve scalars of an embedding iteration (HSD_NV doubles each) */
#define HSD_NV 64
#define HV(q, k) V[HSD_NV * (size_t)(q) + (k)]
enum { MU, XZ, CXv, BY, RG, C1, C2, CZ, BV, GV, DEN, SIGM, DTA, DKA, DT0, DK0, CK, CTK, BESTS, BESTA, BESTM, BESTAM,
       DDT, DDK, DTC, DKC, TK2P, QDP, M0, HSIG, DT0S, DK0S, S0S, A0S, AM0S, DTF, DKF, T0, T1, T2, T3, T4 };
#define HSD_PASSES 4        /* second-order passes an iteration at most (reference Params.hsd_hoc, solver.c:179) */
/* b'u and A(X C Zi)'u into scalars 0 and 1 of the active solves */
static int hsd_tau_dots(Solve **S, int ns, const int *a, size_t fu, int unused) {
    (void)unused;
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        dot_add(s->P.m, s->bvec, F_(s, fu), SC(q, 0));
        dot_add(s->P.m, s->hg, F_(s, fu), SC(q, 1));
    }
    return (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) ? -1 : 0;
}
/* a member of the family from its u (at fy) with the scalars T0..T4 of each solve:
 *   dy = u + T0 v,  dZ = T1 Rd + T0 C - A*(dy),  dX = -sym(Zi dZ X) + T2 X + T3 Zi + T4 Q
 * (T0 the member's dtau; Q only with withq) */
static int hsd_form(Solve **S, int ns, const int *a, const double *V, size_t fy, size_t fdZ, size_t fdX, int withq) {
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        lin_add(s->P.m, F_(s, fy), F_(s, fy), 1.0, s->hv, HV(q, T0), -1, 0, -1, 0);
        lin_add(s->P.T, F_(s, fdZ), s->Rd, HV(q, T1), s->C, HV(q, T0), -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n) || aty(S, ns, a, fy, fdZ, -1.0)) return -1;
    /* (full products: dZ holds tau C and Rd, which are not diagonal where the constraints are) */
    if (blk_gemm(S, ns, a, 0, 0, 0, 1.0, 0.0, OFF(Zi), fdZ, OFF(W1)) || blk_gemm(S, ns, a, 0, 0, 0, 1.0, 0.0, OFF(W1), OFF(X), OFF(G)) ||
        blk_sym(S, ns, a, OFF(G))) return -1;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        lin_add(s->P.T, F_(s, fdX), s->G, -1.0, s->X, HV(q, T2), HV(q, T3) != 0 ? s->Zi : -1, HV(q, T3), withq ? s->Q : -1, HV(q, T4));
    }
    return bk_lincomb(g_ll.v, g_ll.n);
}
/* the sigma search on the family (reference HSD_SIGSEARCH, solver.c:11199-11231): the members at
 * sigma_m, 0.2 sigma_m and 3 sigma_m formed in (dX, dZ), (G, F), (W1, W2), the six step bounds in
 * one Lanczos launch; the candidate with the largest alpha (1 - sigma) goes to BESTS, BESTA
 * (the damped step), BESTAM (the step to the boundary), BESTM */
static int hsd_search(Solve **S, int ns, const int *a, double *V, double *apx, double *apz, double *sgc, double *lo6) {
    static const size_t fX[3] = { OFF(dX), OFF(G), OFF(W1) }, fZ[3] = { OFF(dZ), OFF(F), OFF(W2) };
    (void)apx; (void)apz;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        const double sm = HV(q, SIGM);
        sgc[3 * q] = sm; sgc[3 * q + 1] = fmin(0.95, fmax(1e-4, 0.2 * sm)); sgc[3 * q + 2] = fmin(0.95, fmax(1e-3, 3.0 * sm));
        for (int c = 0; c < 3; c++) {
            lin_add(s->P.T, F_(s, fX[c]), s->dX0, 1.0, s->dX1, sgc[3 * q + c], -1, 0, -1, 0);
            lin_add(s->P.T, F_(s, fZ[c]), s->dZ0, 1.0, s->dZ1, sgc[3 * q + c], -1, 0, -1, 0);
        }
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) return -1;
    const size_t fLi[6] = { OFF(LXi), OFF(LZi), OFF(LXi), OFF(LZi), OFF(LXi), OFF(LZi) };
    const size_t fLiT[6] = { OFF(LXiT), OFF(LZiT), OFF(LXiT), OFF(LZiT), OFF(LXiT), OFF(LZiT) };
    const size_t fV[6] = { fX[0], fZ[0], fX[1], fZ[1], fX[2], fZ[2] };
    if (lanczos_min(S, ns, a, 6, fLi, fLiT, fV, 1, 0.02, lo6)) return -1;
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        HV(q, BESTM) = -1; HV(q, BESTS) = HV(q, SIGM); HV(q, BESTA) = 0; HV(q, BESTAM) = 0;
        for (int c = 0; c < 3; c++) {
            const double sg = sgc[3 * q + c];
            const double dtau = HV(q, DT0) + sg * HV(q, DDT), dkap = HV(q, DK0) + sg * HV(q, DDK);
            double am = fmin(lo6[(2 * c) * ns + q], lo6[(2 * c + 1) * ns + q]);
            if (dtau < 0) am = fmin(am, -s->tau / dtau);
            if (dkap < 0) am = fmin(am, -s->kap / dkap);
            const double gam = 0.9 + (0.99 - 0.9) * fmin(1.0, am);
            const double al = fmin(1.0, gam * am), merit = al * (1.0 - sg);
            if (merit > HV(q, BESTM)) { HV(q, BESTM) = merit; HV(q, BESTS) = sg; HV(q, BESTA) = al; HV(q, BESTAM) = am; }
        }
    }
    return 0;
}
/* lambda_min((X + al dX)(Z + al dZ)) >= thr per solve (al, thr in ald[q], thr[q]): L = chol(X +
 * al dX) in W1, then L'(Z + al dZ) L - thr I must factor (in G); the reference's sdp_nb_ok
 * (solver.c:6470-6504). LP variables are blocks of order 1, where this is the product test.
 * DinvX, DinvZ, F and G serve as scratch (the step's factorizations rebuild the first two). */
static int hsd_nb_ok(Solve **S, int ns, const int *a, const double *ald, const double *thr, int *ok) {
    int *ok1 = calloc(ns, sizeof(int)), *a2 = calloc(ns, sizeof(int)), rc = -1, any = 0;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (a[q]) {
        Solve *s = S[q];
        lin_add(s->P.T, s->W1, s->X, 1.0, s->dX, ald[q], -1, 0, -1, 0);
        lin_add(s->P.T, s->W2, s->Z, 1.0, s->dZ, ald[q], -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    {
        const size_t fa[1] = { OFF(W1) }, fd[1] = { OFF(DinvX) }, fw[1] = { OFF(G) };
        if (chol_sides(S, ns, 1, a, fa, fa, fd, fw, ok1)) goto out;
    }
    for (int q = 0; q < ns; q++) { ok[q] = 0; if (a[q] && ok1[q]) { a2[q] = 1; any = 1; } }
    if (!any) { rc = 0; goto out; }
    if (blk_gemm(S, ns, a2, 0, 0, 0, 1.0, 0.0, OFF(W2), OFF(W1), OFF(F)) || blk_gemm(S, ns, a2, 1, 0, 0, 1.0, 0.0, OFF(W1), OFF(F), OFF(G)) ||
        blk_sym(S, ns, a2, OFF(G))) goto out;
    MItL_clear(&g_ml);
    for (int q = 0; q < ns; q++) {
        g_hdbl[q] = a2[q] ? -thr[q] : 0.0;
        if (!a2[q]) continue;
        for (int k = 0; k < S[q]->P.nb; k++) { MItem it = { S[q]->P.bn[k], q, S[q]->G + S[q]->P.boff[k], 0, 0, 0 }; MItL_add(&g_ml, it); }
    }
    if (bk_h2d(ddbl(), g_hdbl, sizeof(double) * ns) || bk_add_diag(g_ml.v, g_ml.n, g_bdbl / 8)) goto out;
    {
        const size_t fa[1] = { OFF(G) }, fl[1] = { OFF(G) }, fd[1] = { OFF(DinvZ) }, fw[1] = { OFF(F) };
        if (chol_sides(S, ns, 1, a2, fa, fl, fd, fw, ok1)) goto out;
    }
    for (int q = 0; q < ns; q++) if (a2[q]) ok[q] = ok1[q];
    rc = 0;
out:
    free(ok1); free(a2);
    return rc;
}
/* refinement of the chosen direction (reference solver.c:11599-11695): two passes enforcing the
 * primal rows A(dX) - b dtau = eta rp (eta = 1 - sigma). The miss e is corrected by the
 * embedding's Newton step for that residual alone: M w = e, dtau_e = -(b - hg)'w / denom, dy_e =
 * w + dtau_e v, dZ_e = dtau_e C - A*(dy_e), dX_e = -sym(Zi dZ_e X), dkappa_e = -kappa dtau_e / tau;
 * a pass is taken when it at least halves the miss and stops when the miss is below 1e-2 of
 * eta |rp| */
static int hsd_refine(Solve **S, int ns, const int *act, double *V) {
    int *ra = calloc(ns, sizeof(int)), rc = -1;
    double *ne = calloc(ns, sizeof(double)), *rel = calloc(ns, sizeof(double));
    for (int q = 0; q < ns; q++) ra[q] = act[q];
    for (int pass = 0; pass < 2; pass++) {
        int any = 0;
        for (int q = 0; q < ns; q++) any |= ra[q];
        if (!any) break;
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (ra[q]) lin_add(S[q]->P.m, S[q]->ev, S[q]->rp, 1.0 - HV(q, BESTS), S[q]->bvec, HV(q, DTF), -1, 0, -1, 0);
        if (bk_lincomb(g_ll.v, g_ll.n) || aop(S, ns, ra, OFF(dX), OFF(ev), -1.0, 1.0)) goto out;
        DotL_clear(&g_dl);
        for (int q = 0; q < ns; q++) if (ra[q]) { dot_add(S[q]->P.m, S[q]->ev, S[q]->ev, SC(q, 0)); dot_add(S[q]->P.m, S[q]->rp, S[q]->rp, SC(q, 1)); }
        if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
        any = 0;
        for (int q = 0; q < ns; q++) if (ra[q]) {
            ne[q] = sqrt(HS(q, 0));
            if (!(ne[q] > 1e-2 * (1.0 - HV(q, BESTS)) * sqrt(HS(q, 1)) && ne[q] > 0)) ra[q] = 0; else any = 1;
        }
        if (!any) break;
        if (schur_solve(S, ns, ra, OFF(ev), OFF(wv), 0.0, rel) || hsd_tau_dots(S, ns, ra, OFF(wv), -1)) goto out;
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (ra[q]) {
            Solve *s = S[q];
            const double dte = -(HS(q, 0) - HS(q, 1)) / HV(q, DEN);
            HV(q, T0) = dte;
            lin_add(s->P.m, s->wv, s->wv, 1.0, s->hv, dte, -1, 0, -1, 0);
            lin_add(s->P.T, s->F, s->C, dte, -1, 0, -1, 0, -1, 0);
            lin_add(s->P.m, s->qv, s->rp, 1.0 - HV(q, BESTS), s->bvec, HV(q, DTF) + dte, -1, 0, -1, 0);
        }
        if (bk_lincomb(g_ll.v, g_ll.n) || aty(S, ns, ra, OFF(wv), OFF(F), -1.0)) goto out;
        if (blk_gemm(S, ns, ra, 0, 0, 0, 1.0, 0.0, OFF(Zi), OFF(F), OFF(W1)) || blk_gemm(S, ns, ra, 0, 0, 0, 1.0, 0.0, OFF(W1), OFF(X), OFF(G)) ||
            blk_sym(S, ns, ra, OFF(G))) goto out;
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (ra[q]) lin_add(S[q]->P.T, S[q]->Q, S[q]->dX, 1.0, S[q]->G, -1.0, -1, 0, -1, 0);
        if (bk_lincomb(g_ll.v, g_ll.n) || aop(S, ns, ra, OFF(Q), OFF(qv), -1.0, 1.0)) goto out;
        DotL_clear(&g_dl);
        for (int q = 0; q < ns; q++) if (ra[q]) dot_add(S[q]->P.m, S[q]->qv, S[q]->qv, SC(q, 0));
        if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (ra[q]) {
            Solve *s = S[q];
            if (!(sqrt(HS(q, 0)) < 0.5 * ne[q])) { ra[q] = 0; continue; }
            const double dte = HV(q, T0);
            lin_add(s->P.T, s->dX, s->Q, 1.0, -1, 0, -1, 0, -1, 0);
            lin_add(s->P.T, s->dZ, s->dZ, 1.0, s->F, 1.0, -1, 0, -1, 0);
            lin_add(s->P.m, s->dy, s->dy, 1.0, s->wv, 1.0, -1, 0, -1, 0);
            HV(q, DTF) += dte; HV(q, DKF) -= s->kap * dte / s->tau;
        }
        if (g_ll.n && bk_lincomb(g_ll.v, g_ll.n)) goto out;
    }
    rc = 0;
out:
    free(ra); free(ne); free(rel);
    return rc;
}
/* the long step after passes that stood (reference solver.c:11818-11855): where the step to the
 * boundary is at least 0.9, 0.992 of it is tested first; inside the neighbourhood lambda_min(XZ)
 * >= 0.3 mu_new (and tau kappa >= 0.3 mu_new), 0.9995 and then 0.998 are tried, the first one
 * inside taken (0.992 if neither is) */
static int hsd_long_step(Solve **S, int ns, const int *act, double *V, const int *kept, double *alpha) {
    static const double gs[3] = { 0.9995, 0.998, 0.992 };
    int *stage = calloc(ns, sizeof(int)), *t = calloc(ns, sizeof(int)), *ok = calloc(ns, sizeof(int)), rc = -1;
    double *al = calloc(ns, sizeof(double)), *thr = calloc(ns, sizeof(double)), *sf = calloc(ns, sizeof(double));
    /* stage: 0 not taking part, 1 the short test, 2 and 3 the tests at 0.9995 and 0.998 */
    for (int q = 0; q < ns; q++) stage[q] = act[q] && kept[q] > 0 && HV(q, BESTAM) >= 0.9 ? 1 : 0;
    for (int round = 0; round < 3; round++) {
        int any = 0;
        for (int q = 0; q < ns; q++) {
            t[q] = 0;
            if (!stage[q]) continue;
            Solve *s = S[q];
            const double am = HV(q, BESTAM), sg = HV(q, BESTS), dt = HV(q, DTF), dk = HV(q, DKF), mu = HV(q, MU);
            for (;;) {
                const double a = fmin(1.0, (stage[q] == 1 ? gs[2] : gs[stage[q] - 2]) * am);
                if (stage[q] == 1 && !(a > alpha[q] && s->tau + a * dt > 0 && s->kap + a * dk > 0)) { stage[q] = 2; continue; }
                if (stage[q] == 4) { stage[q] = 0; break; }
                if (stage[q] >= 2 && !(a > alpha[q])) { stage[q] = 0; break; }
                if (stage[q] >= 2 && (s->tau + a * dt <= 0 || s->kap + a * dk <= 0)) { stage[q]++; continue; }
                const double mun = mu * (1.0 - a * (1.0 - sg));
                if (!((s->tau + a * dt) * (s->kap + a * dk) >= 0.3 * mun)) {
                    if (stage[q] == 1) { stage[q] = 0; break; }
                    stage[q]++; continue;
                }
                al[q] = a; thr[q] = 0.3 * mun; t[q] = 1; any = 1;
                break;
            }
        }
        if (!any) break;
        if (hsd_nb_ok(S, ns, t, al, thr, ok)) goto out;
        for (int q = 0; q < ns; q++) {
            if (!t[q]) continue;
            if (stage[q] == 1) {
                /* the short step inside: the longer ones are tried (0.992 stands if neither is inside) */
                if (ok[q]) { stage[q] = 2; sf[q] = al[q]; } else stage[q] = 0;
            } else if (ok[q]) { alpha[q] = al[q]; stage[q] = 0; }
            else if (stage[q] == 3) { if (sf[q] > alpha[q]) alpha[q] = sf[q]; stage[q] = 0; }
            else stage[q]++;
        }
    }
    /* a short test that stood with no longer step inside (or none left to try) */
    for (int q = 0; q < ns; q++) if (stage[q] >= 2 && sf[q] > alpha[q]) alpha[q] = sf[q];
    rc = 0;
out:
    free(stage); free(t); free(ok); free(al); free(thr); free(sf);
    return rc;
}
/* the neighbourhood safeguard (reference solver.c:11856-11882): tau kappa and the products of the
 * LP variables at least 1e-3 of the new mu, else the step x0.7 (at most ten times) */
static int hsd_safeguard(Solve **S, int ns, const int *act, double *V, double *alpha) {
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (act[q]) {
        Solve *s = S[q];
        dot_add(s->P.T, s->X, s->dZ, SC(q, 0));
        dot_add(s->P.T, s->dX, s->Z, SC(q, 1));
        dot_add(s->P.T, s->dX, s->dZ, SC(q, 2));
    }
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) return -1;
    double *H = heapp();
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        Solve *s = S[q];
        const HProb *P = &s->P;
        /* the LP variables (blocks of order 1) in runs of consecutive blocks */
        int nlp = 0;
        for (int k = 0; k < P->nb; k++) nlp += P->bn[k] == 1;
        double *lp = NULL;
        if (nlp) {
            lp = malloc(sizeof(double) * 4 * nlp);
            int c = 0;
            for (int k = 0; k < P->nb;) {
                if (P->bn[k] != 1) { k++; continue; }
                int e = k;
                while (e < P->nb && P->bn[e] == 1) e++;
                const int64_t o = P->boff[k], len = e - k;
                if (bk_d2h(lp + c, H + s->X + o, 8 * len) || bk_d2h(lp + nlp + c, H + s->dX + o, 8 * len) ||
                    bk_d2h(lp + 2 * nlp + c, H + s->Z + o, 8 * len) || bk_d2h(lp + 3 * nlp + c, H + s->dZ + o, 8 * len)) { free(lp); return -1; }
                c += (int)len; k = e;
            }
        }
        const double dt = HV(q, DTF), dk = HV(q, DKF);
        double a = alpha[q];
        for (int t2 = 0; t2 < 10; t2++) {
            const double tk = (s->tau + a * dt) * (s->kap + a * dk);
            const double munew = (HV(q, XZ) + a * (HS(q, 0) + HS(q, 1)) + a * a * HS(q, 2) + tk) / (P->Ntot + 1.0);
            double lo = tk;
            for (int i = 0; i < nlp; i++) {
                const double w = (lp[i] + a * lp[nlp + i]) * (lp[2 * nlp + i] + a * lp[3 * nlp + i]);
                if (w < lo) lo = w;
            }
            if (!(munew > 0) || lo >= 1e-3 * munew || a < 1e-3) break;
            a *= 0.7;
        }
        alpha[q] = a;
        free(lp);
    }
    return 0;
}

static int ipm_iterate_hsd_(Solve **S, int ns) {
    int rc = -1;
    int *act = calloc(ns, sizeof(int)), *fail = calloc(ns, sizeof(int)), *okp = calloc(ns, sizeof(int)), *okd = calloc(ns, sizeof(int));
    int *nrm = calloc(ns, sizeof(int));
    double *V = calloc(HSD_NV * (size_t)ns, sizeof(double));  /* per-solve scalars */
    double *rel = calloc(ns, sizeof(double)), *apx = calloc(ns, sizeof(double)), *apz = calloc(ns, sizeof(double));
    for (int q = 0; q < ns; q++) act[q] = !S[q]->done;
    int nsl = 1;
    for (int q = 0; q < ns; q++) nsl += S[q]->P.nb;
    if (batch_scratch(ns, nsl + ns)) goto out;
    /* Zi = LZi' LZi */
    if (form_zi(S, ns, act)) goto out;
    /* residuals: rp = tau b - A(X), Rd = tau C - Z - A*(y) */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        Solve *s = S[q];
        lin_add(s->P.m, s->rp, s->bvec, s->tau, -1, 0, -1, 0, -1, 0);
        lin_add(s->P.T, s->Rd, s->C, s->tau, s->Z, -1.0, -1, 0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (aop(S, ns, act, OFF(X), OFF(rp), -1.0, 1.0)) goto out;
    if (aty(S, ns, act, OFF(y), OFF(Rd), -1.0)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.m, S[q]->t2, S[q]->bvec, S[q]->tau, S[q]->rp, -1.0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.m, S[q]->t1, S[q]->rp, 1.0, S[q]->duv, 0, -1, 0, -1, 0);
    if (bk_vmul(g_ll.v, g_ll.n)) goto out;
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (act[q]) {
        Solve *s = S[q];
        dot_add(s->P.T, s->X, s->Z, SC(q, 0));
        dot_add(s->P.T, s->C, s->X, SC(q, 1));
        dot_add(s->P.T, s->Rd, s->Rd, SC(q, 2));
        dot_add(s->P.m, s->bvec, s->y, SC(q, 3));
        dot_add(s->P.m, s->t1, s->t1, SC(q, 4));
        dot_add(s->P.m, s->t2, s->t2, SC(q, 5));
    }
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        Solve *s = S[q];
        const HProb *P = &s->P;
        const double xz = HS(q, 0), cx = HS(q, 1), rdn2 = HS(q, 2), by = HS(q, 3), rpo = sqrt(HS(q, 4)), nAX = sqrt(HS(q, 5));
        if (s->kap <= 0) s->kap = fmax(xz / P->Ntot, 1e-12);           /* start: tau kappa = mu */
        const double tau = s->tau, kap = s->kap;
        HV(q, XZ) = xz; HV(q, CXv) = cx; HV(q, BY) = by;
        HV(q, MU) = (xz + tau * kap) / (P->Ntot + 1.0);
        HV(q, RG) = kap - by + cx;
        const double sc = P->bs * P->cs, po = sc * cx / tau, dob = sc * by / tau;
        s->pinf = rpo / tau / (1 + P->normb2);
        s->dinf = P->cs * sqrt(rdn2) / tau / (1 + P->normC2);
        const double oden = 1 + fabs(po) + fabs(dob);
        s->relgap = fabs(po - dob) / oden;
        s->relcomp = sc * xz / (tau * tau) / oden;
        s->pobj = po; s->dobj = dob; s->mu = xz / P->Ntot / (tau * tau);
        double score = iterate_score(s->relgap, s->relcomp, s->pinf, s->dinf);
        if (!isfinite(score)) score = INFINITY;
        s->score = score;
        const int it = s->it;
        if (it < 128) { s->tr_pobj[it] = po; s->tr_dobj[it] = dob; s->tr_mu[it] = sc * s->mu; }
        if (s->o.verbose)
            solve_log("%4d %+.10e %+.10e %.1e %.1e %.1e %.1e %.3f tau %.2e kap %.2e %s retries %d %.3fs\n", it, po, dob, s->pinf, s->dinf, s->relgap,
                      sc * s->mu, s->alp, tau, kap, s->f32 ? "f32" : "f64", s->nretry, bk_now() - s->t_start);
        int st = -1;
        if (!isfinite(po) || !isfinite(dob)) st = 4;
        else {
            if (score < s->best_score) {
                s->best_score = score; s->it_best = it;
                lin_add(P->T, s->Xb, s->X, 1.0 / tau, -1, 0, -1, 0, -1, 0);
                lin_add(P->m, s->yb, s->y, 1.0 / tau, -1, 0, -1, 0, -1, 0);
            }
            if (score < s->o.tol) st = 0;
            else if (it >= s->o.maxit) st = 3;
            else if (s->stall >= 3) st = 4;
            else if (it - s->it_best >= 5 && s->best_score < 1e-6) st = 4;
            else if (it - s->it_best >= 12) st = 4;
            else if (s->o.time_limit > 0 && bk_now() - s->t_start > s->o.time_limit) st = 6;
            (void)nAX;
        }
        if (st == 0) nrm[q] = 1;  /* the normalised iterate is the solution (scaled below) */
        if (st >= 0) { s->status = st; s->done = 1; act[q] = 0; continue; }
        if (s->o.prec == PREC_MIXED && fmax(s->relgap, fmax(s->pinf, s->dinf)) <= s->o.mixed_until) s->want_double = 1;
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (nrm[q]) {
        Solve *s = S[q];
        lin_add(s->P.T, s->X, s->X, 1.0 / s->tau, -1, 0, -1, 0, -1, 0);
        lin_add(s->P.m, s->y, s->y, 1.0 / s->tau, -1, 0, -1, 0, -1, 0);
        nrm[q] = 0;
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    {
        int any = 0;
        for (int q = 0; q < ns; q++) any |= act[q];
        if (!any) { rc = 0; goto out; }
    }
    if (schur_assemble(S, ns, act, 0)) goto out;
    if (schur_factor(S, ns, act, 0, fail)) goto out;
    for (int q = 0; q < ns; q++) if (act[q] && fail[q]) { S[q]->status = 4; S[q]->done = 1; act[q] = 0; }
    /* CX = Zi C X, G = Zi Rd X; hg = A(CX), ht = A(G), a0 = A(Zi); v = M^-1 (b + hg) */
    if (blk_gemm(S, ns, act, 0, 0, 0, 1.0, 0.0, OFF(Zi), OFF(C), OFF(W1))) goto out;
    if (blk_gemm(S, ns, act, 0, 0, 0, 1.0, 0.0, OFF(W1), OFF(X), OFF(CX))) goto out;
    if (blk_gemm(S, ns, act, 0, 0, 0, 1.0, 0.0, OFF(Zi), OFF(Rd), OFF(W1))) goto out;
    if (blk_gemm(S, ns, act, 0, 0, 0, 1.0, 0.0, OFF(W1), OFF(X), OFF(G))) goto out;
    if (aop(S, ns, act, OFF(CX), OFF(hg), 1.0, 0.0) || aop(S, ns, act, OFF(G), OFF(ht), 1.0, 0.0) || aop(S, ns, act, OFF(Zi), OFF(a0), 1.0, 0.0)) goto out;
    /* the three solves of the iteration whose right-hand sides do not depend on the second-order
     * term, in one PCG: v = M^-1 (b + hg) (the tau column), the affine direction's u =
     * M^-1 (tau b + ht) into dy, and the family's difference u1 = M^-1 (-rp - mu a0 - ht) into hy1 */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        Solve *s = S[q];
        lin_add(s->P.m, s->rhs, s->bvec, 1.0, s->hg, 1.0, -1, 0, -1, 0);
        lin_add(s->P.m, s->rhs2, s->bvec, s->tau, s->ht, 1.0, -1, 0, -1, 0);
        lin_add(s->P.m, s->rhs3, s->rp, -1.0, s->a0, -HV(q, MU), s->ht, -1.0, -1, 0);
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    {
        const size_t fr[3] = { OFF(rhs), OFF(rhs2), OFF(rhs3) }, fx[3] = { OFF(hv), OFF(dy), OFF(hy1) };
        if (schur_solve_k(S, ns, act, 3, fr, fx, 0.0)) goto out;
    }
    DotL_clear(&g_dl);
    for (int q = 0; q < ns; q++) if (act[q]) {
        Solve *s = S[q];
        dot_add(s->P.T, s->C, s->CX, SC(q, 0));
        dot_add(s->P.T, s->C, s->G, SC(q, 1));
        dot_add(s->P.T, s->C, s->Zi, SC(q, 2));
        dot_add(s->P.m, s->bvec, s->hv, SC(q, 3));
        dot_add(s->P.m, s->hg, s->hv, SC(q, 4));
    }
    if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
    for (int q = 0; q < ns; q++) if (act[q]) {
        HV(q, C2) = HS(q, 0); HV(q, C1) = HS(q, 1); HV(q, CZ) = HS(q, 2); HV(q, BV) = HS(q, 3); HV(q, GV) = HS(q, 4);
        HV(q, DEN) = HV(q, BV) + HV(q, C2) - HV(q, GV) + S[q]->kap / S[q]->tau;
    }
    /* The direction family of the embedding (reference solver.c:11240-11480). With w1(sigma) =
     * [tau b + A(X Rd Zi) + A(Q)] + sigma [-rp - mu A(Zi) - A(X Rd Zi)], M u = w1 has the solution
     * u0 + sigma u1 with u1 independent of the second-order term Q, and every part of the
     * direction is affine in sigma: the family is kept as its base (sigma 0: hy0, dX0, dZ0) and
     * its difference (hy1, dX1, dZ1), which is formed once an iteration.
     * 1. The affine direction (sigma 0, no Q) into dy, dX, dZ: its step and the Mehrotra sigma_m.
     * 2. The difference.
     * 3. Second-order passes: Q = sym(Zi dZ dX) from the current direction (the affine one first,
     *    then the family's member at the sigma the first search chose), u0 solved again, the base
     *    formed and sigma searched among sigma_m, 0.2 sigma_m and 3 sigma_m by the largest
     *    predicted reduction alpha (1 - sigma) of mu (the three candidates' step bounds in one
     *    Lanczos launch). A pass stands while Q contracts (|Q_k - Q_k-1| <= 0.9 |Q_k-1 - Q_k-2|)
     *    and its predicted reduction is at least the first family's, else the first family
     *    stands; at most four passes, none within a factor 100 of the tolerance and none for the
     *    rest of the solve after three iterations whose passes fell behind the first family
     *    (solver.c:11236-11247). */
    {
        double *sgc = calloc(3 * (size_t)ns, sizeof(double)), *lo6 = calloc(6 * (size_t)ns, sizeof(double));
        int *hl = calloc(ns, sizeof(int)), *nhoc = calloc(ns, sizeof(int)), *kept = calloc(ns, sizeof(int));
        int e = 0;
        /* 1. affine: u = M^-1 (tau b + A(X Rd Zi)) (in dy) */
        e = hsd_tau_dots(S, ns, act, OFF(dy), -1);
        for (int q = 0; q < ns && !e; q++) if (act[q]) {
            Solve *s = S[q];
            const double tau = s->tau, kap = s->kap;
            const double num = HV(q, RG) - HS(q, 0) - HV(q, CXv) - HV(q, C1) + HS(q, 1) - kap;
            HV(q, DTA) = num / HV(q, DEN);
            HV(q, DKA) = (-tau * kap - kap * HV(q, DTA)) / tau;
            HV(q, T0) = HV(q, DTA); HV(q, T1) = 1.0; HV(q, T2) = -1.0; HV(q, T3) = 0.0; HV(q, T4) = 0.0;
        }
        if (!e) e = hsd_form(S, ns, act, V, OFF(dy), OFF(dZ), OFF(dX), 0);
        if (!e) e = maxstep2(S, ns, act, apx, apz);
        if (!e) {
            DotL_clear(&g_dl);
            for (int q = 0; q < ns; q++) if (act[q]) {
                Solve *s = S[q];
                dot_add(s->P.T, s->X, s->dZ, SC(q, 0));
                dot_add(s->P.T, s->dX, s->Z, SC(q, 1));
                dot_add(s->P.T, s->dX, s->dZ, SC(q, 2));
            }
            if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) e = 1;
        }
        for (int q = 0; q < ns && !e; q++) if (act[q]) {
            Solve *s = S[q];
            const double tau = s->tau, kap = s->kap, dtau = HV(q, DTA), dkap = HV(q, DKA);
            double am = fmin(apx[q], apz[q]);
            if (dtau < 0) am = fmin(am, -tau / dtau);
            if (dkap < 0) am = fmin(am, -kap / dkap);
            const double aa = fmin(1.0, am);
            const double mua = (HV(q, XZ) + aa * (HS(q, 0) + HS(q, 1)) + aa * aa * HS(q, 2) + (tau + aa * dtau) * (kap + aa * dkap)) / (s->P.Ntot + 1.0);
            const double expon = s->relgap > 1e-6 ? fmax(1.0, 3.0 * aa * aa) : 3.0;
            HV(q, SIGM) = fmin(1.0, pow(fmax(mua / HV(q, MU), 0.0), expon));
            const int late = fmax(s->pinf, fmax(s->relgap, s->dinf)) <= 100.0 * s->o.tol;
            nhoc[q] = (s->hoc_nfail >= 3 || late) ? 0 : HSD_PASSES;
            HV(q, DTC) = dtau; HV(q, DKC) = dkap;
            hl[q] = 1;
        }
        /* 2. the difference: u1 = M^-1 (-rp - mu A(Zi) - A(X Rd Zi)) (in hy1) */
        if (!e) e = hsd_tau_dots(S, ns, act, OFF(hy1), -1);
        for (int q = 0; q < ns && !e; q++) if (act[q]) {
            Solve *s = S[q];
            const double tau = s->tau, kap = s->kap, mu = HV(q, MU);
            const double num = -HV(q, RG) - HS(q, 0) + mu * HV(q, CZ) + HV(q, C1) + HS(q, 1) + mu / tau;
            HV(q, DDT) = num / HV(q, DEN);
            HV(q, DDK) = (mu - kap * HV(q, DDT)) / tau;
            HV(q, T0) = HV(q, DDT); HV(q, T1) = -1.0; HV(q, T2) = 0.0; HV(q, T3) = mu; HV(q, T4) = 0.0;
        }
        if (!e) e = hsd_form(S, ns, act, V, OFF(hy1), OFF(dZ1), OFF(dX1), 0);
        /* 3. the passes */
        for (int hoc = 0; !e; hoc++) {
            int any = 0;
            for (int q = 0; q < ns; q++) any |= hl[q];
            if (!any) break;
            /* Q = sym(Zi dZ dX) at the current direction, A(Q), <C, Q>; the change against the last pass */
            if (blk_gemm(S, ns, hl, 0, 0, 0, 1.0, 0.0, OFF(Zi), OFF(dZ), OFF(W1)) || blk_gemm(S, ns, hl, 0, 0, 0, 1.0, 0.0, OFF(W1), OFF(dX), OFF(Q)) ||
                blk_sym(S, ns, hl, OFF(Q))) { e = 1; break; }
            LinL_clear(&g_ll);
            for (int q = 0; q < ns; q++) if (hl[q] && nhoc[q] > 0 && hoc > 0) lin_add(S[q]->P.T, S[q]->G, S[q]->Q, 1.0, S[q]->Qp, -1.0, -1, 0, -1, 0);
            if (g_ll.n && bk_lincomb(g_ll.v, g_ll.n)) { e = 1; break; }
            if (aop(S, ns, hl, OFF(Q), OFF(qv), 1.0, 0.0)) { e = 1; break; }
            DotL_clear(&g_dl);
            for (int q = 0; q < ns; q++) if (hl[q]) {
                Solve *s = S[q];
                dot_add(s->P.T, s->C, s->Q, SC(q, 2));
                if (nhoc[q] > 0) { if (hoc == 0) dot_add(s->P.T, s->Q, s->Q, SC(q, 3)); else dot_add(s->P.T, s->G, s->G, SC(q, 3)); }
            }
            if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) { e = 1; break; }
            LinL_clear(&g_ll);
            for (int q = 0; q < ns; q++) if (hl[q]) {
                Solve *s = S[q];
                const double tk2 = HV(q, DTC) * HV(q, DKC);
                if (nhoc[q] > 0) {
                    const double d = sqrt(HS(q, 3) + (hoc == 0 ? tk2 * tk2 : (tk2 - HV(q, TK2P)) * (tk2 - HV(q, TK2P))));
                    if (hoc > 0 && !(d <= 0.9 * HV(q, QDP))) { hl[q] = 0; continue; }   /* the last pass that stood keeps its family */
                    HV(q, QDP) = d; HV(q, TK2P) = tk2;
                    lin_add(s->P.T, s->Qp, s->Q, 1.0, -1, 0, -1, 0, -1, 0);
                }
                HV(q, CK) = -HS(q, 2); HV(q, CTK) = -tk2;
                lin_add(s->P.m, s->rhs, s->bvec, s->tau, s->ht, 1.0, s->qv, 1.0, -1, 0);
            }
            if (bk_lincomb(g_ll.v, g_ll.n)) { e = 1; break; }
            any = 0;
            for (int q = 0; q < ns; q++) any |= hl[q];
            if (!any) break;
            if (schur_solve(S, ns, hl, OFF(rhs), OFF(hy0), 0.0, rel) || hsd_tau_dots(S, ns, hl, OFF(hy0), -1)) { e = 1; break; }
            for (int q = 0; q < ns; q++) if (hl[q]) {
                Solve *s = S[q];
                const double tau = s->tau, kap = s->kap;
                const double num = HV(q, RG) - HS(q, 0) - HV(q, CXv) - HV(q, C1) + HS(q, 1) + HV(q, CK) + (-tau * kap + HV(q, CTK)) / tau;
                HV(q, DT0) = num / HV(q, DEN);
                HV(q, DK0) = (-tau * kap + HV(q, CTK) - kap * HV(q, DT0)) / tau;
                HV(q, T0) = HV(q, DT0); HV(q, T1) = 1.0; HV(q, T2) = -1.0; HV(q, T3) = 0.0; HV(q, T4) = -1.0;
            }
            if (hsd_form(S, ns, hl, V, OFF(hy0), OFF(dZ0), OFF(dX0), 1)) { e = 1; break; }
            /* the sigma search: the three candidates' directions in (dX, dZ), (G, F), (W1, W2) */
            if (hsd_search(S, ns, hl, V, apx, apz, sgc, lo6)) { e = 1; break; }
            LinL_clear(&g_ll);
            int restore = 0;
            for (int q = 0; q < ns; q++) if (hl[q]) {
                Solve *s = S[q];
                if (nhoc[q] == 0) { kept[q] = 0; hl[q] = 0; continue; }
                if (hoc == 0) {
                    HV(q, M0) = HV(q, BESTM); HV(q, HSIG) = HV(q, BESTS);
                    HV(q, DT0S) = HV(q, DT0); HV(q, DK0S) = HV(q, DK0);
                    HV(q, S0S) = HV(q, BESTS); HV(q, A0S) = HV(q, BESTA); HV(q, AM0S) = HV(q, BESTAM);
                    lin_add(s->P.T, s->dX0s, s->dX0, 1.0, -1, 0, -1, 0, -1, 0);
                    lin_add(s->P.T, s->dZ0s, s->dZ0, 1.0, -1, 0, -1, 0, -1, 0);
                    lin_add(s->P.m, s->hy0s, s->hy0, 1.0, -1, 0, -1, 0, -1, 0);
                }
                if (hoc == 0 || HV(q, BESTM) >= HV(q, M0)) kept[q] = hoc;
                else {
                    /* behind the first family: back to it */
                    lin_add(s->P.T, s->dX0, s->dX0s, 1.0, -1, 0, -1, 0, -1, 0);
                    lin_add(s->P.T, s->dZ0, s->dZ0s, 1.0, -1, 0, -1, 0, -1, 0);
                    lin_add(s->P.m, s->hy0, s->hy0s, 1.0, -1, 0, -1, 0, -1, 0);
                    HV(q, DT0) = HV(q, DT0S); HV(q, DK0) = HV(q, DK0S);
                    HV(q, BESTS) = HV(q, S0S); HV(q, BESTA) = HV(q, A0S); HV(q, BESTAM) = HV(q, AM0S);
                    kept[q] = 0; s->hoc_nfail++; hl[q] = 0; restore = 1;
                    continue;
                }
                if (hoc >= nhoc[q]) { hl[q] = 0; continue; }
                /* the next pass's direction: the family's member at the first search's sigma */
                const double sg = HV(q, HSIG);
                lin_add(s->P.T, s->dX, s->dX0, 1.0, s->dX1, sg, -1, 0, -1, 0);
                lin_add(s->P.T, s->dZ, s->dZ0, 1.0, s->dZ1, sg, -1, 0, -1, 0);
                HV(q, DTC) = HV(q, DT0) + sg * HV(q, DDT); HV(q, DKC) = HV(q, DK0) + sg * HV(q, DDK);
            }
            (void)restore;
            if (g_ll.n && bk_lincomb(g_ll.v, g_ll.n)) { e = 1; break; }
        }
        /* the chosen member: dy, dX, dZ; dtau, dkappa; the step */
        if (!e) {
            LinL_clear(&g_ll);
            for (int q = 0; q < ns; q++) if (act[q]) {
                Solve *s = S[q];
                const double sg = HV(q, BESTS);
                lin_add(s->P.T, s->dX, s->dX0, 1.0, s->dX1, sg, -1, 0, -1, 0);
                lin_add(s->P.T, s->dZ, s->dZ0, 1.0, s->dZ1, sg, -1, 0, -1, 0);
                lin_add(s->P.m, s->dy, s->hy0, 1.0, s->hy1, sg, -1, 0, -1, 0);
                HV(q, DTF) = HV(q, DT0) + sg * HV(q, DDT); HV(q, DKF) = HV(q, DK0) + sg * HV(q, DDK);
                apx[q] = apz[q] = HV(q, BESTA);
            }
            if (bk_lincomb(g_ll.v, g_ll.n)) e = 1;
        }
        if (!e) e = hsd_refine(S, ns, act, V);
        if (!e) e = hsd_long_step(S, ns, act, V, kept, apx);
        if (!e) e = hsd_safeguard(S, ns, act, V, apx);
        for (int q = 0; q < ns; q++) if (act[q]) apz[q] = apx[q];
        free(sgc); free(lo6); free(hl); free(nhoc); free(kept);
        if (e) goto out;
    }
    {
        int *ta = calloc(2 * ns, sizeof(int)), *okv = calloc(2 * ns, sizeof(int)), *ok = calloc(2 * ns, sizeof(int));
        double *st = calloc(2 * ns, sizeof(double));
        int e = 0;
        for (int t = 0; t < 2 * ns; t++) ta[t] = act[t % ns];
        for (int pass = 0; pass < 2 && !e; pass++) {
            /* pass 0: each side its own retries; pass 1: the longer side back to the common step */
            if (pass == 1) {
                int any = 0;
                for (int t = 0; t < 2 * ns; t++) {
                    const int q = t % ns;
                    ta[t] = 0;
                    if (!act[q] || !okv[q] || !okv[ns + q]) continue;
                    const double a = fmin(apx[q], apz[q]);
                    double *my = t < ns ? &apx[q] : &apz[q];
                    if (*my > a) { *my = a; ta[t] = 1; any = 1; }
                }
                if (!any) break;
            }
            for (int tries = 0; tries < 40; tries++) {
                int any = 0;
                for (int t = 0; t < 2 * ns; t++) any |= ta[t];
                if (!any) break;
                LinL_clear(&g_ll);
                for (int t = 0; t < 2 * ns; t++) {
                    if (!ta[t]) continue;
                    const int sd = t / ns, q = t % ns;
                    double *a = sd ? &apz[q] : &apx[q];
                    const double inc = *a - st[t];
                    st[t] = *a;
                    lin_add(S[q]->P.T, F_(S[q], F_V[sd]), F_(S[q], F_V[sd]), 1.0, F_(S[q], F_dV[sd]), inc, -1, 0, -1, 0);
                }
                if (bk_lincomb(g_ll.v, g_ll.n)) { e = 1; break; }
                if (chol_sides(S, ns, 2, ta, F_V, F_L, F_D, F_W, ok)) { e = 1; break; }
                for (int t = 0; t < 2 * ns; t++) {
                    if (!ta[t]) continue;
                    if (ok[t]) { ta[t] = 0; okv[t] = 1; }
                    else { okv[t] = 0; S[t % ns]->nretry++; if (t / ns) apz[t % ns] *= 0.8; else apx[t % ns] *= 0.8; }
                }
            }
        }
        for (int q = 0; q < ns; q++) { okp[q] = okv[q]; okd[q] = okv[ns + q]; }
        if (!e) {
            /* update y, tau, kappa; normalise tau to 1 past 10 (X and Z are refactored) */
            LinL_clear(&g_ll);
            int anyn = 0;
            for (int q = 0; q < ns; q++) {
                if (!act[q]) continue;
                Solve *s = S[q];
                if (!okp[q] || !okd[q]) { s->status = 4; s->done = 1; continue; }
                const double a = fmin(apx[q], apz[q]);
                const double dtau = HV(q, DTF), dkap = HV(q, DKF);
                lin_add(s->P.m, s->y, s->y, 1.0, s->dy, a, -1, 0, -1, 0);
                s->tau += a * dtau; s->kap += a * dkap;
                s->alp = s->ald = a;
                s->stall = a < 1e-6 ? s->stall + 1 : 0;
                if (!(s->tau > 0) || !(s->kap > 0)) { s->status = 4; s->done = 1; okp[q] = 0; continue; }
                if (s->tau > 10.0) {
                    const double f = 1.0 / s->tau;
                    lin_add(s->P.T, s->X, s->X, f, -1, 0, -1, 0, -1, 0);
                    lin_add(s->P.T, s->Z, s->Z, f, -1, 0, -1, 0, -1, 0);
                    nrm[q] = 1; anyn = 1;
                }
            }
            if (bk_lincomb(g_ll.v, g_ll.n)) e = 1;
            if (!e && anyn) {
                LinL_clear(&g_ll);
                for (int q = 0; q < ns; q++) if (nrm[q]) {
                    Solve *s = S[q];
                    lin_add(s->P.m, s->y, s->y, 1.0 / s->tau, -1, 0, -1, 0, -1, 0);
                    s->kap /= s->tau; s->tau = 1.0;
                }
                if (bk_lincomb(g_ll.v, g_ll.n)) e = 1;
                for (int t = 0; t < 2 * ns; t++) ta[t] = nrm[t % ns];
                if (!e && chol_sides(S, ns, 2, ta, F_V, F_L, F_D, F_W, ok)) e = 1;
            }
            for (int t = 0; t < 2 * ns; t++) okv[t] = act[t % ns] && okp[t % ns] && okd[t % ns] && !S[t % ns]->done;
            if (!e) e = inv_sides(S, ns, 2, okv, F_L, F_D, F_Li, F_LiT, F_W);
        }
        free(ta); free(okv); free(ok); free(st);
        if (e) goto out;
    }
    for (int q = 0; q < ns; q++) if (act[q] && !S[q]->done) S[q]->it++;
    rc = 0;
out:
    free(act); free(fail); free(okp); free(okd); free(nrm); free(V); free(rel); free(apx); free(apz);
    return rc;
}

int ipm_run_batch(Solve **S, int ns) {
    if (ipm_start_batch(S, ns)) return -1;
    for (int guard = 0; guard < 1000; guard++) {
        int any = 0;
        for (int q = 0; q < ns; q++) any |= !S[q]->done;
        if (!any) break;
        if (ipm_iterate_batch(S, ns)) return -1;
    }
    return ipm_end_batch(S, ns);
}
static int ipm_end_batch_(Solve **S, int ns);
int ipm_end_batch(Solve **S, int ns) {
    const double t_ = bk_prof_on() ? bk_now() : 0;
    const int r_ = ipm_end_batch_(S, ns);
    if (t_ > 0) bk_prof_host(47, bk_now() - t_);
    return r_;
}
/* The best-iterate polish of the reference (polish_solution, solver.c:7012-7040, applied to
 * the returned point at solver.c:15534-15545): X_p = X + A*(w) with (A A') w = b - A(X), the
 * projection of X onto A(X) = b in the Euclidean metric, y kept and Z = C - A*(y) as before. The
 * Gram matrix A A' is assembled at X = Zi = I and factored in FP64 as for the start point, and
 * the system solved by PCG to 1e-14. The polished point is kept when its acceptance error (the
 * measure the status is set by) falls below half the point's; else the point is put back. The
 * projection moves the primal residual (error 1) and leaves the other errors nearly as they were,
 * so it runs only where error 1 is at least half the acceptance error: elsewhere it cannot halve
 * it, and a keep-if-lower rule would decide near-ties by rounding (the reference keeps any gain of
 * its own measure, solver.c:7012-7040; the margin makes the choice stable). For the solves whose
 * run ended optimal, in difficulties or at reduced accuracy (not at the iteration or time limit,
 * as in the reference), unless -nopolish; in either precision, the projection being FP64 here
 * (the reference skips its mixed-precision solves, whose end phase is its own); a self-dual
 * embedding's point is (X, y) / tau, as returned. st0: the runs' statuses before their returned points were
 * measured. */
/* ---- the end phase's metric corrections (reference sqrt_restore, solver.c:7551-7655, and
 * polish_xmetric, 7657-7760) ------------------------------------------------------------------ */
#define RST_MARGIN 0.01
#define RST_MAXN 512        /* the square-root metric's eigendecomposition on blocks up to this order */
/* X <- X + a dX with dX = D W D or sym(X W Zi), W = A*(w) and M w = b - A(X) for the Schur
 * complement M = A (D (x) D) A' or A (X (x) Zi) A', assembled (the routes of the iteration),
 * factored in FP64 and solved to 1e-14:
 *  - half = 1, the restoration of primal feasibility in the square-root metric: D = X^{1/2} by
 *    the eigendecomposition (eigenvalues below 0 taken as 0), D = G G' with G = V diag(w^{1/4}),
 *    so the dense route's P_i = G' A_i G. The step a (from 1, halved up to eight times) keeps
 *    (1 - 0.01) I + a W positive definite: X^{1/2} (I + a W) X^{1/2} stays in the cone and no
 *    eigenvalue shrinks by more than 0.01. Three passes, until the residual falls below 1e-6
 *    of its first value.
 *  - half = 0, the polish in the metric of X: Zi the inverse of the point's Z = C - A*(y) (the
 *    primal half of an HKM step); a (halved up to six times) keeps X + a dX positive definite.
 *    Two passes.
 * One step per solve, the smallest over its blocks (the reference steps each block by its own).
 * Solves whose Z has no factor (half = 0), whose Schur factorization fails or whose step never
 * passes drop out with X as the last pass left it. X is corrected in place (the caller keeps a
 * copy); W1, dX, dZ, F, G, R0, W2, LX, LZ, LXi, Zi, LZi and the Schur arrays are scratch. steps
 * (NULL, or 3 per solve): the step of each pass taken, 0 for none. */
static int metric_correct(Solve **S, int ns, const int *act0, int half, double *steps) {
    int rc = -1;
    int *act = calloc(ns, sizeof(int)), *ok = calloc(ns, sizeof(int)), *tst = calloc(ns, sizeof(int)), *pd = calloc(2 * ns, sizeof(int));
    int *wd = calloc(ns, sizeof(int));
    double *nr0 = calloc(ns, sizeof(double)), *al = calloc(ns, sizeof(double)), *rel = calloc(ns, sizeof(double));
    EigItem *ei = NULL;
    ScaleItem *s4 = NULL, *s0 = NULL;
    double *H = heapp();
    for (int q = 0; q < ns; q++) { act[q] = act0[q]; wd[q] = S[q]->want_double; S[q]->want_double = 1; }
    if (steps) for (int q = 0; q < 3 * ns; q++) steps[q] = 0;
    if (!half) {
        /* Zi = (C - A*(y))^-1 of the point (in W1 after the measure): LZ LZ' = Z, LZi = LZ^-1 */
        const size_t fV[1] = { OFF(W1) }, fL[1] = { OFF(LZ) }, fD[1] = { OFF(DinvZ) }, fW[1] = { OFF(W2) };
        if (chol_sides(S, ns, 1, act, fV, fL, fD, fW, pd)) goto out;
        for (int q = 0; q < ns; q++) if (act[q] && !pd[q]) act[q] = 0;
        const size_t fLi[1] = { OFF(LZi) }, fLiT[1] = { OFF(LZiT) };
        if (inv_sides(S, ns, 1, act, fL, fD, fLi, fLiT, fW) || form_zi(S, ns, act)) goto out;
    }
    for (int pass = 0; pass < (half ? 3 : 2); pass++) {
        int any = 0;
        for (int q = 0; q < ns; q++) any |= act[q];
        if (!any) break;
        /* rhs = b - A(X) */
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.m, S[q]->rhs, S[q]->bvec, 1.0, -1, 0, -1, 0, -1, 0);
        if (bk_lincomb(g_ll.v, g_ll.n) || aop(S, ns, act, OFF(X), OFF(rhs), -1.0, 1.0)) goto out;
        DotL_clear(&g_dl);
        for (int q = 0; q < ns; q++) if (act[q]) dot_add(S[q]->P.m, S[q]->rhs, S[q]->rhs, SC(q, 0));
        if (bk_dots(g_dl.v, g_dl.n) || read_scalars(ns)) goto out;
        any = 0;
        for (int q = 0; q < ns; q++) if (act[q]) {
            const double nr = sqrt(HS(q, 0));
            if (pass == 0) nr0[q] = nr; else if (nr <= 1e-6 * nr0[q]) { act[q] = 0; continue; }
            any = 1;
        }
        if (!any) break;
        if (half) {
            /* X = V diag(w) V' (a copy of X in dZ; V into G, w into lzw), G <- V diag(max(w, 0)^{1/4}),
             * D = G G' into dX, G' into R0; the scaling pair (D, D) with the factors (G, G') */
            int ne = 0, cap = 0;
            LinL_clear(&g_ll);
            MItL_clear(&g_ml);
            for (int q = 0; q < ns; q++) if (act[q]) {
                Solve *s = S[q];
                SAux *A = aux_of(s);
                lin_add(s->P.T, s->dZ, s->X, 1.0, -1, 0, -1, 0, -1, 0);
                for (int k = 0; k < s->P.nb; k++) {
                    if (ne == cap) {
                        cap = 2 * cap + 64;
                        ei = realloc(ei, sizeof(EigItem) * cap); s4 = realloc(s4, sizeof(ScaleItem) * cap); s0 = realloc(s0, sizeof(ScaleItem) * cap);
                    }
                    const int n = s->P.bn[k];
                    const int64_t o = s->P.boff[k], w = s->lzw + A->lzoff[k], u = s->lzw + (int64_t)s->P.Ntot + A->lzoff[k];
                    EigItem e = { n, 0, s->dZ + o, s->G + o, w };
                    ScaleItem c4 = { n, 4, w, u }, c0 = { n, 0, s->G + o, u };
                    ei[ne] = e; s4[ne] = c4; s0[ne] = c0; ne++;
                    MItem it = { n, 0, s->G + o, 0, 0, s->R0 + o };
                    MItL_add(&g_ml, it);
                }
            }
            if (bk_lincomb(g_ll.v, g_ll.n) || bk_syev_f64(ei, ne, H) || bk_nt_scale(s4, ne, H) || bk_nt_scale(s0, ne, H) ||
                bk_transpose_f64(g_ml.v, g_ml.n) || blk_gemm(S, ns, act, 0, 1, 0, 1.0, 0.0, OFF(G), OFF(G), OFF(dX)) || blk_sym(S, ns, act, OFF(dX)))
                goto out;
            for (int q = 0; q < ns; q++) if (act[q]) { Solve *s = S[q]; s->SX = s->dX; s->SZi = s->dX; s->SLX = s->G; s->SLZi = s->R0; }
        } else {
            /* the factor of X as it stands (the dense route's P_i = LX' A_i LZi') */
            const size_t fV[1] = { OFF(X) }, fL[1] = { OFF(LX) }, fD[1] = { OFF(DinvX) }, fW[1] = { OFF(W2) };
            if (chol_sides(S, ns, 1, act, fV, fL, fD, fW, pd)) goto out;
            for (int q = 0; q < ns; q++) if (act[q] && !pd[q]) act[q] = 0;
            for (int q = 0; q < ns; q++) if (act[q]) { Solve *s = S[q]; s->SX = s->X; s->SZi = s->Zi; s->SLX = s->LX; s->SLZi = s->LZi; }
        }
        /* w = M^-1 (b - A(X)) into dy, W = A*(w) into F */
        if (schur_assemble(S, ns, act, 0) || schur_factor(S, ns, act, 1, pd)) goto out;
        for (int q = 0; q < ns; q++) if (act[q] && pd[q]) act[q] = 0;
        if (schur_solve(S, ns, act, OFF(rhs), OFF(dy), 1e-14, rel)) goto out;
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (act[q]) lin_add(S[q]->P.T, S[q]->F, -1, 0, -1, 0, -1, 0, -1, 0);
        if (bk_lincomb(g_ll.v, g_ll.n) || aty(S, ns, act, OFF(dy), OFF(F), 1.0)) goto out;
        /* dX: D W D into dZ, or sym(Zi W X) (= sym(X W Zi)) into G */
        const size_t fdx = half ? OFF(dZ) : OFF(G);
        if (half) {
            if (blk_gemm(S, ns, act, 0, 0, 0, 1.0, 0.0, OFF(F), OFF(dX), OFF(LXi)) || blk_gemm(S, ns, act, 0, 0, 0, 1.0, 0.0, OFF(dX), OFF(LXi), OFF(dZ)) ||
                blk_sym(S, ns, act, OFF(dZ))) goto out;
        } else if (zi_v_x(S, ns, act, OFF(F), OFF(X), OFF(G)) || blk_sym(S, ns, act, OFF(G))) goto out;
        /* the step: (1 - margin) I + a W, or X + a dX, positive definite (Cholesky in LZ) */
        for (int q = 0; q < ns; q++) { al[q] = 1.0; tst[q] = act[q]; ok[q] = 0; }
        for (int h = 0; h < (half ? 8 : 6); h++) {
            int nt = 0;
            LinL_clear(&g_ll);
            MItL_clear(&g_ml);
            for (int q = 0; q < ns; q++) {
                g_hdbl[q] = 0;
                if (!tst[q]) continue;
                Solve *s = S[q];
                nt++;
                if (half) {
                    lin_add(s->P.T, s->LZ, s->F, al[q], -1, 0, -1, 0, -1, 0);
                    g_hdbl[q] = 1.0 - RST_MARGIN;
                    for (int k = 0; k < s->P.nb; k++) { MItem it = { s->P.bn[k], q, s->LZ + s->P.boff[k], 0, 0, 0 }; MItL_add(&g_ml, it); }
                } else lin_add(s->P.T, s->LZ, s->X, 1.0, F_(s, fdx), al[q], -1, 0, -1, 0);
            }
            if (!nt) break;
            if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
            if (half && (bk_h2d(ddbl(), g_hdbl, sizeof(double) * ns) || bk_add_diag(g_ml.v, g_ml.n, g_bdbl / 8))) goto out;
            const size_t fa[1] = { OFF(LZ) }, fd[1] = { OFF(DinvZ) }, fw[1] = { OFF(W2) };
            if (chol_sides(S, ns, 1, tst, fa, fa, fd, fw, pd)) goto out;
            for (int q = 0; q < ns; q++) if (tst[q]) { if (pd[q]) { ok[q] = 1; tst[q] = 0; } else al[q] *= 0.5; }
        }
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (act[q]) {
            if (S[q]->o.verbose)
                solve_log("%s, pass %d: residual %.3e, solve residual %.1e, step %.3g%s\n", half ? "restoration (square-root metric)" : "polish (X metric)",
                          pass, sqrt(HS(q, 0)), rel[q], ok[q] ? al[q] : 0.0, ok[q] ? "" : " (no step keeps X in the cone)");
            if (!ok[q]) { act[q] = 0; continue; }
            lin_add(S[q]->P.T, S[q]->X, S[q]->X, 1.0, F_(S[q], fdx), al[q], -1, 0, -1, 0);
            if (steps) steps[3 * q + pass] = al[q];
        }
        if (g_ll.n && bk_lincomb(g_ll.v, g_ll.n)) goto out;
    }
    rc = 0;
out:
    for (int q = 0; q < ns; q++) { Solve *s = S[q]; s->SX = s->X; s->SZi = s->Zi; s->SLX = s->LX; s->SLZi = s->LZi; s->want_double = wd[q]; }
    free(act); free(ok); free(tst); free(pd); free(wd); free(nr0); free(al); free(rel); free(ei); free(s4); free(s0);
    return rc;
}
/* The end phase's metric corrections of points that missed the tolerance (reference
 * solver.c:15574-15631): the restoration of primal feasibility in the square-root metric when
 * the primal residual is at least a quarter of the acceptance error (a restoration moves X in
 * proportion to that residual, so it cannot help an error that is the complementarity of a
 * feasible point) and every block has order <= 512; the polish in the metric of X when no
 * restoration was taken. Each is kept when it lowers the acceptance error; otherwise the point
 * and its measure go back. */
int ipm_end_corrections(Solve **S, int ns, const int *st0) {
    int rc = -1;
    int *cur = calloc(ns, sizeof(int)), *taken = calloc(ns, sizeof(int));
    double *old = calloc((size_t)ns * 9, sizeof(double));
    Solve **A = malloc(sizeof(Solve *) * (ns + 1));
    for (int half = 1; half >= 0; half--) {
        int na = 0;
        for (int q = 0; q < ns; q++) {
            Solve *s = S[q];
            const double acc = acc_err(s->res_err);
            int small = 1;
            for (int k = 0; k < s->P.nb; k++) if (s->P.bn[k] > RST_MAXN) small = 0;
            cur[q] = s->o.polish && (half ? s->o.polish_r : s->o.polish_x) && (st0[q] == 0 || st0[q] == 4 || st0[q] == 5) && !taken[q] && isfinite(acc) && acc > s->o.tol &&
                     (half ? small && s->res_err[0] >= 0.25 * acc : 1);
            na += cur[q];
        }
        if (!na) continue;
        /* the point and its measure kept: X in Xb, its Z (W1) in Q */
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) if (cur[q]) {
            Solve *s = S[q];
            lin_add(s->P.T, s->Xb, s->X, 1.0, -1, 0, -1, 0, -1, 0);
            lin_add(s->P.T, s->Q, s->W1, 1.0, -1, 0, -1, 0, -1, 0);
            for (int e = 0; e < 6; e++) old[9 * q + e] = s->res_err[e];
            old[9 * q + 6] = s->res_pobj; old[9 * q + 7] = s->res_dobj; old[9 * q + 8] = s->status;
        }
        if (bk_lincomb(g_ll.v, g_ll.n) || metric_correct(S, ns, cur, half, NULL)) goto out;
        na = 0;
        for (int q = 0; q < ns; q++) if (cur[q]) { S[q]->status = st0[q]; A[na++] = S[q]; }
        if (dimacs_point(A, na)) goto out;
        LinL_clear(&g_ll);
        for (int q = 0; q < ns; q++) {
            if (!cur[q]) continue;
            Solve *s = S[q];
            const int keep = acc_err(s->res_err) < acc_err(old + 9 * q);
            if (s->o.verbose)
                solve_log("%s: acceptance error %.3e -> %.3e (errors %.2e %.2e %.2e %.2e %.2e %.2e), %s\n",
                          half ? "restoration (square-root metric)" : "polish (X metric)", acc_err(old + 9 * q), acc_err(s->res_err),
                          s->res_err[0], s->res_err[1], s->res_err[2], s->res_err[3], s->res_err[4], s->res_err[5], keep ? "kept" : "not kept");
            if (keep) { taken[q] = 1; continue; }
            lin_add(s->P.T, s->X, s->Xb, 1.0, -1, 0, -1, 0, -1, 0);
            lin_add(s->P.T, s->W1, s->Q, 1.0, -1, 0, -1, 0, -1, 0);
            for (int e = 0; e < 6; e++) s->res_err[e] = old[9 * q + e];
            s->res_pobj = old[9 * q + 6]; s->res_dobj = old[9 * q + 7]; s->status = (int)old[9 * q + 8];
        }
        if (g_ll.n && bk_lincomb(g_ll.v, g_ll.n)) goto out;
    }
    rc = 0;
out:
    free(cur); free(taken); free(old); free(A);
    return rc;
}
int ipm_polish(Solve **S, int ns, const int *st0) {
    int rc = -1, na = 0;
    int *act = calloc(ns, sizeof(int)), *pol = calloc(ns, sizeof(int)), *fail = calloc(ns, sizeof(int));
    double *rel = calloc(ns, sizeof(double)), *old = calloc((size_t)ns * 9, sizeof(double));
    Solve **A = malloc(sizeof(Solve *) * (ns + 1));
    for (int q = 0; q < ns; q++) {
        Solve *s = S[q];
        pol[q] = act[q] = s->o.polish && (st0[q] == 0 || st0[q] == 4 || st0[q] == 5) &&
                          s->res_err[0] >= 0.5 * acc_err(s->res_err);
        na += act[q];
    }
    if (!na) { rc = 0; goto out; }
    /* the point and its measure kept: X in Xb, its Z (W1) in Q; rhs = b - A(X) */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (act[q]) {
        Solve *s = S[q];
        lin_add(s->P.T, s->Xb, s->X, 1.0, -1, 0, -1, 0, -1, 0);
        lin_add(s->P.T, s->Q, s->W1, 1.0, -1, 0, -1, 0, -1, 0);
        lin_add(s->P.m, s->rhs, s->bvec, 1.0, -1, 0, -1, 0, -1, 0);
        for (int e = 0; e < 6; e++) old[9 * q + e] = s->res_err[e];
        old[9 * q + 6] = s->res_pobj; old[9 * q + 7] = s->res_dobj; old[9 * q + 8] = s->status;
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (aop(S, ns, act, OFF(X), OFF(rhs), -1.0, 1.0)) goto out;
    /* the Gram factor at X = Zi = I, w = (A A')^{-1} (b - A(X)) */
    if (set_identity(S, ns, act, OFF(X)) || set_identity(S, ns, act, OFF(Zi))) goto out;
    if (schur_assemble(S, ns, act, 1) || schur_factor(S, ns, act, 1, fail)) goto out;
    for (int q = 0; q < ns; q++) if (act[q] && fail[q]) act[q] = 0;
    if (schur_solve(S, ns, act, OFF(rhs), OFF(dy), 1e-14, rel)) goto out;
    /* X_p = X + A*(w) (the solves whose Gram factor failed get their point back) */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) if (pol[q]) lin_add(S[q]->P.T, S[q]->X, S[q]->Xb, 1.0, -1, 0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    if (aty(S, ns, act, OFF(dy), OFF(X), 1.0)) goto out;
    na = 0;
    for (int q = 0; q < ns; q++) if (act[q]) { S[q]->status = st0[q]; A[na++] = S[q]; }
    if (na && dimacs_point(A, na)) goto out;
    /* kept when better; else the point and its measure back */
    LinL_clear(&g_ll);
    for (int q = 0; q < ns; q++) {
        if (!act[q]) continue;
        Solve *s = S[q];
        if (s->o.verbose)
            solve_log("polish: acceptance error %.3e -> %.3e (errors %.2e %.2e %.2e %.2e %.2e %.2e), %s\n", acc_err(old + 9 * q),
                      acc_err(s->res_err), s->res_err[0], s->res_err[1], s->res_err[2], s->res_err[3], s->res_err[4], s->res_err[5],
                      acc_err(s->res_err) < 0.5 * acc_err(old + 9 * q) ? "kept" : "not kept");
        if (acc_err(s->res_err) < 0.5 * acc_err(old + 9 * q)) continue;
        lin_add(s->P.T, s->X, s->Xb, 1.0, -1, 0, -1, 0, -1, 0);
        lin_add(s->P.T, s->W1, s->Q, 1.0, -1, 0, -1, 0, -1, 0);
        for (int e = 0; e < 6; e++) s->res_err[e] = old[9 * q + e];
        s->res_pobj = old[9 * q + 6]; s->res_dobj = old[9 * q + 7]; s->status = (int)old[9 * q + 8];
    }
    if (bk_lincomb(g_ll.v, g_ll.n)) goto out;
    rc = 0;
out:
    free(act); free(pol); free(fail); free(rel); free(old); free(A);
    return rc;
}
static int ipm_end_batch_(Solve **S, int ns) {
    int *st0 = malloc(sizeof(int) * (ns + 1));
    for (int q = 0; q < ns; q++) st0[q] = S[q]->status;
    const int e = ipm_dimacs(S, ns) || ipm_end_corrections(S, ns, st0) || ipm_polish(S, ns, st0);
    free(st0);
    if (e) return -1;
    return ipm_finish_batch(S, ns);
}

/* ---- kernel-level probes: one kernel family of the iteration on caller data, through the same
 * setup, layout and launches as the iteration (entry points of the parity tests, probe.c) ---- */
/* Zi = LZi' LZi from Z, and LX from X, as at the start of an iteration; then the Schur
 * complement at (X, Zi) in the precision asked for (prec 1 FP64; 0 the mixed path's FP32 dense
 * route), into s->M */
int ipm_probe_schur(Solve *s, int prec) {
    Solve *S[1] = { s };
    int ok[2] = { 0, 0 }, a2[2] = { 1, 1 };
    s->o.prec = prec ? PREC_F64 : PREC_MIXED;
    s->want_double = prec != 0;
    if (batch_scratch(1, s->P.nb + 2)) return -1;
    if (chol_sides(S, 1, 2, NULL, F_V, F_L, F_D, F_W, ok) || !ok[0] || !ok[1]) return -1;
    if (inv_sides(S, 1, 2, a2, F_L, F_D, F_Li, F_LiT, F_W)) return -1;
    if (form_zi(S, 1, NULL)) return -1;
    return schur_assemble(S, 1, NULL, 0) || bk_sync() ? -1 : 0;
}
/* s->rp = A(D) with D in s->W1 */
int ipm_probe_aop(Solve *s) {
    Solve *S[1] = { s };
    return aop(S, 1, NULL, OFF(W1), OFF(rp), 1.0, 0.0) || bk_sync() ? -1 : 0;
}
/* s->W1 = A*(y) with y in s->y */
int ipm_probe_aty(Solve *s) {
    Solve *S[1] = { s };
    LinL_clear(&g_ll);
    lin_add(s->P.T, s->W1, -1, 0, -1, 0, -1, 0, -1, 0);
    if (bk_lincomb(g_ll.v, g_ll.n)) return -1;
    return aty(S, 1, NULL, OFF(y), OFF(W1), 1.0) || bk_sync() ? -1 : 0;
}
/* the Lanczos bound of the step-length test, smallest over the solve's blocks: mode 0
 * lambda_min(X) (the start point's test), mode 1 lambda_min(LX^-1 dX LX^-T) with LX the
 * Cholesky factor of X (the test of maxstep); *out = theta - res */
int ipm_probe_lammin(Solve *s, int mode, double *out) {
    Solve *S[1] = { s };
    if (batch_scratch(1, s->P.nb + 2)) return -1;
    if (mode == 0) return lammin_all(S, 1, NULL, OFF(X), out);
    int ok = 0, a1 = 1;
    if (blk_chol(S, 1, NULL, OFF(X), OFF(LX), OFF(DinvX), &ok) || !ok) return -1;
    if (inv_sides(S, 1, 1, &a1, F_L, F_D, F_Li, F_LiT, F_W)) return -1;
    const size_t fLi[1] = { OFF(LXi) }, fLiT[1] = { OFF(LXiT) }, fV[1] = { OFF(dX) };
    return lanczos_min(S, 1, NULL, 1, fLi, fLiT, fV, 0, 0.02, out);
}
/* the DIMACS errors of (X, y) in s->X, s->y (scaled), as for a returned point */
/* the square-root-metric restoration (half 1, three passes) or the X-metric polish (half 0, two
 * passes) of the point (X, y) as the end phase runs it; steps[3]: the step of each pass */
int ipm_probe_correct(Solve *s, int half, double *steps) {
    Solve *S[1] = { s };
    const int one = 1;
    s->status = 0;
    return ipm_dimacs(S, 1) || metric_correct(S, 1, &one, half, steps) ? -1 : 0;
}
int ipm_probe_dimacs(Solve *s) {
    Solve *S[1] = { s };
    s->status = 0;
    return ipm_dimacs(S, 1);
}
