/* gen.c - seeded generators of the SDP instances the solver is measured on, as SDPA sparse
 * triplets (the numbers brisk_solve_data takes: mat 0 = F0 with C = -F0, mat 1..m = A_i,
 * 1-based block/row/column indices, upper triangle only, zeros skipped).
 *
 * Every random number is a pure function of (seed, stream, counter) - a counter-based hash
 * (SplitMix64 finaliser applied twice) - so any machine regenerates the same instance bit for
 * bit, with no state carried between calls.
 *
 * Families:
 *   theta   - Lovasz theta SDP of G(n, p): C = -J, tr X = 1, X_ij = 0 on edges (the form of the
 *             reference's tools/timelimit_test.sh:12-24: F0 = J, A_1 = I, A_e = E_ij).
 *   maxcut  - max-cut relaxation of a +-1 weighted G(n, p): C = -L/4, diag X = 1.
 *   dense   - one n x n block, m dense symmetric A_i with N(0, 1/n) entries, primal-dual
 *             feasible by construction from X0, Z0 = Q diag(U[0.1,1]) Q', y0 ~ N(0,1).
 *   multi   - nb SDP blocks of order bn plus an LP block of nlp variables; each A_i touches
 *             1-3 blocks with 4-12 random symmetric N(0,1) nonzeros each; feasible likewise.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "brisk_gen.h"

/* ---- counter-based random numbers ---------------------------------------------------- */
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t ctr64(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed * 0xD1B54A32D192ED03ull + stream * 0x8CB92BA72F3D8DD7ull) ^ (i * 0x9FB21C651E98DF25ull));
}
/* uniform in [0, 1) with 53 random bits */
static inline double u01(uint64_t seed, uint64_t stream, uint64_t i) {
    return (double)(ctr64(seed, stream, i) >> 11) * 0x1.0p-53;
}
/* standard normal by Box-Muller on the counters 2i, 2i+1 (cosine branch only) */
static inline double gauss(uint64_t seed, uint64_t stream, uint64_t i) {
    double u1 = u01(seed, stream, 2 * i), u2 = u01(seed, stream, 2 * i + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
double brisk_gen_u01(uint64_t seed, uint64_t stream, uint64_t i) { return u01(seed, stream, i); }
double brisk_gen_gauss(uint64_t seed, uint64_t stream, uint64_t i) { return gauss(seed, stream, i); }

enum { S_EDGE = 1, S_WEIGHT, S_A, S_QX, S_QZ, S_EX, S_EZ, S_Y0, S_BLKS, S_POS, S_VAL, S_LPX, S_LPZ, S_CNT };

/* ---- growable triplet list ------------------------------------------------------------ */
typedef struct { BriskInstance *I; int64_t cap; } TB;
static int tb_push(TB *t, int mat, int blk, int i, int j, double v) {
    if (v == 0.0) return 0;
    BriskInstance *I = t->I;
    if (I->nnz == t->cap) {
        int64_t nc = t->cap ? 2 * t->cap : 1024;
        int *a = realloc(I->mat, sizeof(int) * nc), *b, *c, *d;
        if (!a) return -1;
        I->mat = a;
        b = realloc(I->blk, sizeof(int) * nc); if (!b) return -1; I->blk = b;
        c = realloc(I->i, sizeof(int) * nc); if (!c) return -1; I->i = c;
        d = realloc(I->j, sizeof(int) * nc); if (!d) return -1; I->j = d;
        double *e = realloc(I->v, sizeof(double) * nc); if (!e) return -1; I->v = e;
        t->cap = nc;
    }
    const int64_t q = I->nnz++;
    I->mat[q] = mat; I->blk[q] = blk; I->i[q] = i; I->j[q] = j; I->v[q] = v;
    return 0;
}
static int tb_reserve(TB *t, int64_t n) {
    BriskInstance *I = t->I;
    if (n <= t->cap) return 0;
    I->mat = realloc(I->mat, sizeof(int) * n); I->blk = realloc(I->blk, sizeof(int) * n);
    I->i = realloc(I->i, sizeof(int) * n); I->j = realloc(I->j, sizeof(int) * n);
    I->v = realloc(I->v, sizeof(double) * n);
    if (!I->mat || !I->blk || !I->i || !I->j || !I->v) return -1;
    t->cap = n;
    return 0;
}

void brisk_gen_free(BriskInstance *I) {
    if (!I) return;
    free(I->bs); free(I->c); free(I->mat); free(I->blk); free(I->i); free(I->j); free(I->v);
    memset(I, 0, sizeof(*I));
}

/* ---- graph families ------------------------------------------------------------------- */
/* edges i < j (0-based) in row-major order; returns the count, fills E (2 per edge) */
static int64_t gnp_edges(int n, double p, uint64_t seed, int **E) {
    int64_t cnt = 0, cap = 1024;
    int *e = malloc(sizeof(int) * 2 * cap);
    if (!e) return -1;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (u01(seed, S_EDGE, (uint64_t)i * n + j) < p) {
                if (cnt == cap) { cap *= 2; int *f = realloc(e, sizeof(int) * 2 * cap); if (!f) { free(e); return -1; } e = f; }
                e[2 * cnt] = i; e[2 * cnt + 1] = j; cnt++;
            }
    *E = e;
    return cnt;
}

int brisk_gen_theta(int n, double p, uint64_t seed, BriskInstance *I) {
    memset(I, 0, sizeof(*I));
    if (n < 1) return -1;
    int *E = NULL;
    int64_t ne = gnp_edges(n, p, seed, &E);
    if (ne < 0) return -1;
    I->m = (int)(1 + ne); I->nblk = 1;
    I->bs = malloc(sizeof(int)); I->bs[0] = n;
    I->c = calloc(I->m, sizeof(double)); I->c[0] = 1.0;
    TB t = { I, 0 };
    if (tb_reserve(&t, (int64_t)n * (n + 1) / 2 + n + ne)) { free(E); return -1; }
    for (int i = 1; i <= n; i++) for (int j = i; j <= n; j++) tb_push(&t, 0, 1, i, j, 1.0);   /* F0 = J: C = -J */
    for (int i = 1; i <= n; i++) tb_push(&t, 1, 1, i, i, 1.0);                              /* A_1 = I */
    for (int64_t e = 0; e < ne; e++) tb_push(&t, (int)(2 + e), 1, E[2 * e] + 1, E[2 * e + 1] + 1, 1.0);
    free(E);
    return 0;
}

int brisk_gen_maxcut(int n, double p, uint64_t seed, BriskInstance *I) {
    memset(I, 0, sizeof(*I));
    if (n < 1) return -1;
    int *E = NULL;
    int64_t ne = gnp_edges(n, p, seed, &E);
    if (ne < 0) return -1;
    I->m = n; I->nblk = 1;
    I->bs = malloc(sizeof(int)); I->bs[0] = n;
    I->c = malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) I->c[i] = 1.0;
    double *deg = calloc(n, sizeof(double));
    TB t = { I, 0 };
    if (!deg || tb_reserve(&t, 2 * n + ne)) { free(E); free(deg); return -1; }
    /* F0 = L/4 (C = -L/4): off-diagonal -w/4, diagonal (weighted degree)/4 */
    for (int64_t e = 0; e < ne; e++) {
        const int a = E[2 * e], b = E[2 * e + 1];
        const double w = u01(seed, S_WEIGHT, (uint64_t)a * n + b) < 0.5 ? -1.0 : 1.0;
        deg[a] += w; deg[b] += w;
        tb_push(&t, 0, 1, a + 1, b + 1, -0.25 * w);
    }
    for (int i = 0; i < n; i++) tb_push(&t, 0, 1, i + 1, i + 1, 0.25 * deg[i]);
    for (int i = 0; i < n; i++) tb_push(&t, i + 1, 1, i + 1, i + 1, 1.0);
    free(E); free(deg);
    return 0;
}

/* cycle C_n with unit weights (known answers: max-cut SDP n for even n, (n/2)(1 + cos(pi/n)) odd) */
int brisk_gen_maxcut_cycle(int n, BriskInstance *I) {
    memset(I, 0, sizeof(*I));
    if (n < 3) return -1;
    I->m = n; I->nblk = 1;
    I->bs = malloc(sizeof(int)); I->bs[0] = n;
    I->c = malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) I->c[i] = 1.0;
    TB t = { I, 0 };
    for (int i = 0; i < n; i++) {
        int a = i, b = (i + 1) % n; if (a > b) { int s = a; a = b; b = s; }
        tb_push(&t, 0, 1, a + 1, b + 1, -0.25);
    }
    for (int i = 0; i < n; i++) tb_push(&t, 0, 1, i + 1, i + 1, 0.5);
    for (int i = 0; i < n; i++) tb_push(&t, i + 1, 1, i + 1, i + 1, 1.0);
    return 0;
}

/* theta SDP of an explicit edge list (0-based pairs) */
int brisk_gen_theta_edges(int n, int64_t ne, const int *edges, BriskInstance *I) {
    memset(I, 0, sizeof(*I));
    if (n < 1 || ne < 0) return -1;
    I->m = (int)(1 + ne); I->nblk = 1;
    I->bs = malloc(sizeof(int)); I->bs[0] = n;
    I->c = calloc(I->m, sizeof(double)); I->c[0] = 1.0;
    TB t = { I, 0 };
    for (int i = 1; i <= n; i++) for (int j = i; j <= n; j++) tb_push(&t, 0, 1, i, j, 1.0);
    for (int i = 1; i <= n; i++) tb_push(&t, 1, 1, i, i, 1.0);
    for (int64_t e = 0; e < ne; e++) {
        int a = edges[2 * e], b = edges[2 * e + 1]; if (a > b) { int s = a; a = b; b = s; }
        if (a < 0 || b >= n || a == b) return -1;
        tb_push(&t, (int)(2 + e), 1, a + 1, b + 1, 1.0);
    }
    return 0;
}

/* ---- dense random orthogonal factor ---------------------------------------------------- */
/* V = Q diag(e) Q' with Q the orthogonal factor of a Householder QR of a Gaussian matrix G
 * (G from stream sq), e_i ~ U[0.1, 1] from stream se; V is n x n column-major, full */
static int spd_from_qr(int n, uint64_t seed, int sq, int se, uint64_t salt, double *V) {
    double *A = malloc(sizeof(double) * (size_t)n * n), *Q = malloc(sizeof(double) * (size_t)n * n);
    double *vv = malloc(sizeof(double) * n), *e = malloc(sizeof(double) * n);
    if (!A || !Q || !vv || !e) { free(A); free(Q); free(vv); free(e); return -1; }
    for (size_t q = 0; q < (size_t)n * n; q++) A[q] = gauss(seed, sq, salt * 1000003ull * 1000003ull + q);
    for (int i = 0; i < n; i++) e[i] = 0.1 + 0.9 * u01(seed, se, salt * 1000003ull + i);
    /* Q = I, then apply the reflections H_0 .. H_{n-2} from the right: Q = H_0 H_1 ... */
    for (size_t q = 0; q < (size_t)n * n; q++) Q[q] = 0;
    for (int i = 0; i < n; i++) Q[i + (size_t)i * n] = 1;
    for (int k = 0; k < n - 1; k++) {
        double *a = A + (size_t)k * n;
        double nrm = 0;
        for (int i = k; i < n; i++) nrm += a[i] * a[i];
        nrm = sqrt(nrm);
        if (nrm == 0) continue;
        const double alpha = a[k] > 0 ? -nrm : nrm;
        for (int i = 0; i < n; i++) vv[i] = i < k ? 0 : a[i];
        vv[k] -= alpha;
        double vn = 0;
        for (int i = k; i < n; i++) vn += vv[i] * vv[i];
        if (vn == 0) continue;
        const double tau = 2.0 / vn;
        /* A <- H A for the remaining columns */
        for (int j = k; j < n; j++) {
            double *c = A + (size_t)j * n, s = 0;
            for (int i = k; i < n; i++) s += vv[i] * c[i];
            s *= tau;
            for (int i = k; i < n; i++) c[i] -= s * vv[i];
        }
        /* Q <- Q H */
        for (int r = 0; r < n; r++) {
            double s = 0;
            for (int i = k; i < n; i++) s += Q[r + (size_t)i * n] * vv[i];
            s *= tau;
            for (int i = k; i < n; i++) Q[r + (size_t)i * n] -= s * vv[i];
        }
    }
    /* V = Q diag(e) Q' */
    for (int j = 0; j < n; j++)
        for (int i = j; i < n; i++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += Q[i + (size_t)k * n] * e[k] * Q[j + (size_t)k * n];
            V[i + (size_t)j * n] = V[j + (size_t)i * n] = s;
        }
    free(A); free(Q); free(vv); free(e);
    return 0;
}

int brisk_gen_dense(int n, int m, uint64_t seed, BriskInstance *I) {
    memset(I, 0, sizeof(*I));
    if (n < 1 || m < 1) return -1;
    const size_t nn = (size_t)n * n, nu = (size_t)n * (n + 1) / 2;
    double *X0 = malloc(sizeof(double) * nn), *Z0 = malloc(sizeof(double) * nn), *C = calloc(nn, sizeof(double));
    double *Ai = malloc(sizeof(double) * nu);
    if (!X0 || !Z0 || !C || !Ai) { free(X0); free(Z0); free(C); free(Ai); return -1; }
    if (spd_from_qr(n, seed, S_QX, S_EX, 0, X0) || spd_from_qr(n, seed, S_QZ, S_EZ, 0, Z0)) { free(X0); free(Z0); free(C); free(Ai); return -1; }
    I->m = m; I->nblk = 1;
    I->bs = malloc(sizeof(int)); I->bs[0] = n;
    I->c = malloc(sizeof(double) * m);
    TB t = { I, 0 };
    if (tb_reserve(&t, (int64_t)(m + 1) * (int64_t)nu)) { free(X0); free(Z0); free(C); free(Ai); return -1; }
    const double sd = 1.0 / sqrt((double)n);
    for (int a = 0; a < m; a++) {
        /* upper entries (i <= j) column by column */
        double b = 0;
        const double y0 = gauss(seed, S_Y0, a);
        size_t q = 0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i <= j; i++, q++) {
                const double v = sd * gauss(seed, S_A, (uint64_t)a * nu + q);
                Ai[q] = v;
                b += (i == j ? 1.0 : 2.0) * v * X0[i + (size_t)j * n];
                C[i + (size_t)j * n] += y0 * v;
            }
        I->c[a] = b;
        q = 0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i <= j; i++, q++) tb_push(&t, a + 1, 1, i + 1, j + 1, Ai[q]);
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i <= j; i++) tb_push(&t, 0, 1, i + 1, j + 1, -(C[i + (size_t)j * n] + Z0[i + (size_t)j * n]));
    free(X0); free(Z0); free(C); free(Ai);
    return 0;
}

int brisk_gen_multi(int nb, int bn, int nlp, int m, uint64_t seed, BriskInstance *I) {
    memset(I, 0, sizeof(*I));
    if (nb < 0 || bn < 1 || nlp < 0 || m < 1 || nb + (nlp > 0) < 1) return -1;
    const int nblk = nb + (nlp > 0 ? 1 : 0);
    const size_t bnn = (size_t)bn * bn;
    double *X0 = malloc(sizeof(double) * bnn * (nb ? nb : 1)), *Z0 = malloc(sizeof(double) * bnn * (nb ? nb : 1));
    double *C = calloc(bnn * (nb ? nb : 1), sizeof(double));
    double *x0 = malloc(sizeof(double) * (nlp ? nlp : 1)), *z0 = malloc(sizeof(double) * (nlp ? nlp : 1)), *clp = calloc(nlp ? nlp : 1, sizeof(double));
    if (!X0 || !Z0 || !C || !x0 || !z0 || !clp) goto fail;
    for (int k = 0; k < nb; k++)
        if (spd_from_qr(bn, seed, S_QX, S_EX, 1 + k, X0 + k * bnn) || spd_from_qr(bn, seed, S_QZ, S_EZ, 1 + k, Z0 + k * bnn)) goto fail;
    for (int i = 0; i < nlp; i++) { x0[i] = 0.1 + 0.9 * u01(seed, S_LPX, i); z0[i] = 0.1 + 0.9 * u01(seed, S_LPZ, i); }
    I->m = m; I->nblk = nblk;
    I->bs = malloc(sizeof(int) * nblk);
    for (int k = 0; k < nb; k++) I->bs[k] = bn;
    if (nlp > 0) I->bs[nb] = -nlp;
    I->c = calloc(m, sizeof(double));
    TB t = { I, 0 };
    for (int a = 0; a < m; a++) {
        const double y0 = gauss(seed, S_Y0, a);
        uint64_t cnt = 0;
        const int nt = 1 + (int)(3.0 * u01(seed, S_BLKS, (uint64_t)a * 64 + cnt++));
        int chosen[3] = { -1, -1, -1 };
        for (int s = 0; s < nt; s++) {
            int k;
            do { k = (int)(nblk * u01(seed, S_BLKS, (uint64_t)a * 64 + cnt++)); }
            while (k == chosen[0] || k == chosen[1]);
            chosen[s] = k;
        }
        /* blocks in ascending order so the triplets of one constraint are grouped by block */
        for (int s = 0; s < nt; s++) for (int r = s + 1; r < nt; r++) if (chosen[r] < chosen[s]) { int w = chosen[r]; chosen[r] = chosen[s]; chosen[s] = w; }
        double b = 0;
        for (int s = 0; s < nt; s++) {
            const int k = chosen[s];
            const int isl = (nlp > 0 && k == nb);
            const int dim = isl ? nlp : bn;
            const int nz = 4 + (int)(9.0 * u01(seed, S_CNT, (uint64_t)a * 4 + s));
            int pi[12], pj[12], np = 0;
            uint64_t pc = 0;
            while (np < nz) {
                int i = (int)(dim * u01(seed, S_POS, ((uint64_t)a * 4 + s) * 4096 + pc++));
                int j = isl ? i : (int)(dim * u01(seed, S_POS, ((uint64_t)a * 4 + s) * 4096 + pc++));
                if (i > j) { int w = i; i = j; j = w; }
                int dup = 0;
                for (int q = 0; q < np; q++) if (pi[q] == i && pj[q] == j) dup = 1;
                if (dup) { if (pc > 4000) break; continue; }
                pi[np] = i; pj[np] = j; np++;
            }
            for (int q = 0; q < np; q++) {
                const double v = gauss(seed, S_VAL, ((uint64_t)a * 4 + s) * 16 + q);
                const int i = pi[q], j = pj[q];
                if (isl) { b += v * x0[i]; clp[i] += y0 * v; }
                else {
                    const double *X = X0 + k * bnn;
                    b += (i == j ? 1.0 : 2.0) * v * X[i + (size_t)j * bn];
                    C[k * bnn + i + (size_t)j * bn] += y0 * v;
                }
                tb_push(&t, a + 1, k + 1, i + 1, j + 1, v);
            }
        }
        I->c[a] = b;
    }
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < bn; j++)
            for (int i = 0; i <= j; i++)
                tb_push(&t, 0, k + 1, i + 1, j + 1, -(C[k * bnn + i + (size_t)j * bn] + Z0[k * bnn + i + (size_t)j * bn]));
    for (int i = 0; i < nlp; i++) tb_push(&t, 0, nb + 1, i + 1, i + 1, -(clp[i] + z0[i]));
    free(X0); free(Z0); free(C); free(x0); free(z0); free(clp);
    return 0;
fail:
    free(X0); free(Z0); free(C); free(x0); free(z0); free(clp);
    brisk_gen_free(I);
    return -1;
}
