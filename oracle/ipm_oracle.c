/* ipm_oracle.c - TEST INFRASTRUCTURE ONLY. A plain-C, single-threaded, FP64 restatement of the
 * reference's primal-dual interior-point iteration for block-diagonal SDP/LP (the infeasible-
 * start method with the HKM direction and Mehrotra's predictor-corrector), used as the checker
 * of the GPU path. Only tests/, the smoke entry and bench.py's CPU-baseline leg load it; the
 * product library never does.
 *
 * What it follows (reference file:line):
 *   scaling            problem.c:1866-1909 (d_i = 1/||A_i||_F, bs, cs, du)
 *   start point        solver.c:13940-14210 (least-norm X0, least-squares y0 on A A', shifts)
 *   R0 / theta         solver.c:14316-14329, re-base solver.c:14614-14636
 *   residuals/scores   solver.c:14426-14482, res_score solver.c:8089-8097
 *   stopping rules     solver.c:14567-14600 (tolerance, infeasibility, maxit, stall windows)
 *   Schur route 0      solver.c:2626-2660 (M_ab = sum v_p v_q X[c_p,r_q] Zi[c_q,r_p])
 *   factor             solver.c:5266-5420 (equilibration, FP64 shift ladder, pivot test)
 *   PCG solve          solver.c:5483-5622
 *   predictor          solver.c:14787-14818, sigma/gamma solver.c:14829-14872
 *   corrector          solver.c:14997-15055, refinement solver.c:15058-15105
 *   step lengths       solver.c:5947-6027 (tridiagonalisation + Sturm bisection, used here for
 *                      every block size: the exact lambda_min the Lanczos bound approximates);
 *                      the Lanczos bound itself, solver.c:6029-6225 (oracle_lanczos,
 *                      oracle_step_lambda), for the checks of the GPU's bound
 *   catastrophe guard  solver.c:15143-15160, try_step solver.c:6393-6446, update 15185-15188
 *   DIMACS errors      postsolve.c:3037-3112, status rule main.c:2349-2359, brisk.h:187-189
 *
 *   metric corrections solver.c:7551-7760 (sqrt_restore, polish_xmetric: oracle_correct)
 *
 * Deliberately absent (documented as out of scope): presolve, the NT direction and the
 * double-double endgame.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct { int n, lp; size_t off; } OBlk;          /* off into the packed block arrays */
typedef struct { int nnz; int *r, *c; double *v; } OEnt;  /* full symmetric expansion, 0-based */
typedef struct {
    int m, nb;
    OBlk *B;
    size_t tot;               /* packed length of all blocks (n*n SDP, n LP) */
    double Ntot;
    OEnt *A;                  /* A[k * m + i]: entries of constraint i in block k */
    double *C;                /* packed dense C (scaled) */
    double *b, *b0, *d, *du;
    double normb1, normb2, normC1, normC2, bs, cs;
} OProb;

/* ------------------------------------------------------------------------------------- */
static double dot(size_t n, const double *a, const double *b) { double s = 0; for (size_t i = 0; i < n; i++) s += a[i] * b[i]; return s; }

/* lower Cholesky in place (column-major); returns 0 or the failing column + 1 */
static int chol(int n, double *A) {
    for (int j = 0; j < n; j++) {
        double d = A[j + (size_t)j * n];
        for (int k = 0; k < j; k++) d -= A[j + (size_t)k * n] * A[j + (size_t)k * n];
        if (!(d > 0)) return j + 1;
        d = sqrt(d);
        A[j + (size_t)j * n] = d;
        for (int i = j + 1; i < n; i++) {
            double s = A[i + (size_t)j * n];
            for (int k = 0; k < j; k++) s -= A[i + (size_t)k * n] * A[j + (size_t)k * n];
            A[i + (size_t)j * n] = s / d;
        }
    }
    for (int j = 0; j < n; j++) for (int i = 0; i < j; i++) A[i + (size_t)j * n] = 0;
    return 0;
}
/* x <- L^{-1} x, x <- L^{-T} x */
static void lsolve(int n, const double *L, double *x) {
    for (int i = 0; i < n; i++) { double s = x[i]; for (int k = 0; k < i; k++) s -= L[i + (size_t)k * n] * x[k]; x[i] = s / L[i + (size_t)i * n]; }
}
static void ltsolve(int n, const double *L, double *x) {
    for (int i = n - 1; i >= 0; i--) { double s = x[i]; for (int k = i + 1; k < n; k++) s -= L[k + (size_t)i * n] * x[k]; x[i] = s / L[i + (size_t)i * n]; }
}
/* Vi = (L L')^{-1} (full) */
static void inv_from_chol(int n, const double *L, double *Vi) {
    memset(Vi, 0, sizeof(double) * (size_t)n * n);
    for (int j = 0; j < n; j++) { double *x = Vi + (size_t)j * n; x[j] = 1; lsolve(n, L, x); ltsolve(n, L, x); }
    for (int j = 0; j < n; j++) for (int i = j + 1; i < n; i++) { double v = 0.5 * (Vi[i + (size_t)j * n] + Vi[j + (size_t)i * n]); Vi[i + (size_t)j * n] = Vi[j + (size_t)i * n] = v; }
}
static void gemm_nn(int n, const double *A, const double *B, double *C) {     /* C = A B */
    for (int j = 0; j < n; j++) {
        double *c = C + (size_t)j * n;
        for (int i = 0; i < n; i++) c[i] = 0;
        for (int k = 0; k < n; k++) { const double bk = B[k + (size_t)j * n]; const double *a = A + (size_t)k * n; for (int i = 0; i < n; i++) c[i] += a[i] * bk; }
    }
}
static void symmetrize(int n, double *G) {
    for (int j = 0; j < n; j++) for (int i = j + 1; i < n; i++) { double v = 0.5 * (G[i + (size_t)j * n] + G[j + (size_t)i * n]); G[i + (size_t)j * n] = G[j + (size_t)i * n] = v; }
}
/* smallest eigenvalue of a symmetric matrix: Householder tridiagonalisation, then Sturm-count
 * bisection (the method of the reference's small_lammin, solver.c:5947-6027); B is destroyed.
 * The lower end of the final bracket is returned, less a rounding margin. */
double oracle_lammin(int n, double *B) {
    if (n == 1) return B[0];
    double *d = malloc(sizeof(double) * n), *e = calloc(n, sizeof(double)), *v = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n);
    for (int k = 0; k + 2 < n; k++) {
        const int r = n - k - 1;
        double *x = &B[(k + 1) + (size_t)k * n];
        double sig = 0;
        for (int i = 1; i < r; i++) sig += x[i] * x[i];
        if (sig == 0) { e[k] = x[0]; continue; }
        const double a = sqrt(x[0] * x[0] + sig), alpha = x[0] > 0 ? -a : a;
        v[0] = x[0] - alpha;
        for (int i = 1; i < r; i++) v[i] = x[i];
        const double beta = 2.0 / (v[0] * v[0] + sig);
        e[k] = alpha;
        double pv = 0;
        for (int i = 0; i < r; i++) {
            double t = 0;
            for (int j = 0; j < r; j++) t += B[(k + 1 + i) + (size_t)(k + 1 + j) * n] * v[j];
            p[i] = beta * t; pv += p[i] * v[i];
        }
        const double hb = 0.5 * beta * pv;
        for (int i = 0; i < r; i++) p[i] -= hb * v[i];
        for (int j = 0; j < r; j++)
            for (int i = j; i < r; i++) {
                double t = B[(k + 1 + i) + (size_t)(k + 1 + j) * n] - v[i] * p[j] - p[i] * v[j];
                B[(k + 1 + i) + (size_t)(k + 1 + j) * n] = t; B[(k + 1 + j) + (size_t)(k + 1 + i) * n] = t;
            }
    }
    for (int k = 0; k < n; k++) d[k] = B[k + (size_t)k * n];
    e[n - 2] = B[(n - 1) + (size_t)(n - 2) * n];
    double lo = INFINITY, hi = INFINITY, sc = 0;
    for (int k = 0; k < n; k++) {
        const double rad = (k > 0 ? fabs(e[k - 1]) : 0) + (k + 1 < n ? fabs(e[k]) : 0);
        lo = fmin(lo, d[k] - rad); hi = fmin(hi, d[k]); sc = fmax(sc, fabs(d[k]) + rad);
    }
    for (int it = 0; it < 200; it++) {
        const double w = hi - lo;
        if (w <= 1e-13 * fmax(fabs(lo), fabs(hi)) || w <= 1e-15 * sc) break;
        const double x = 0.5 * (lo + hi);
        int cnt = 0;
        double q = d[0] - x;
        if (q < 0) cnt++;
        for (int k = 1; k < n; k++) {
            if (fabs(q) < 1e-300) q = q < 0 ? -1e-300 : 1e-300;
            q = d[k] - x - e[k - 1] * e[k - 1] / q;
            if (q < 0) cnt++;
        }
        if (cnt >= 1) hi = x; else lo = x;
    }
    free(d); free(e); free(v); free(p);
    return lo - 1e-14 * sc;
}

/* ------------------------------------------------------------------------------------- */
static int prob_build(const OracleData *D, OProb *P) {
    memset(P, 0, sizeof(*P));
    const int m = D->m, nb = D->nblk;
    P->m = m; P->nb = nb;
    P->B = calloc(nb, sizeof(OBlk));
    size_t off = 0;
    for (int k = 0; k < nb; k++) {
        const int n = abs(D->bs[k]);
        P->B[k].n = n; P->B[k].lp = D->bs[k] < 0; P->B[k].off = off;
        off += P->B[k].lp ? (size_t)n : (size_t)n * n;
        P->Ntot += n;
    }
    P->tot = off;
    P->C = calloc(off + 1, sizeof(double));
    P->A = calloc((size_t)nb * m, sizeof(OEnt));
    P->b = malloc(sizeof(double) * (m + 1)); P->b0 = malloc(sizeof(double) * (m + 1));
    for (int i = 0; i < m; i++) P->b[i] = P->b0[i] = D->c[i];
    int *cnt = calloc((size_t)nb * m, sizeof(int));
    for (int64_t q = 0; q < D->nnz; q++) {
        const int t = D->mat[q], k = D->blk[q] - 1;
        int i = D->i[q] - 1, j = D->j[q] - 1;
        if (t < 0 || t > m || k < 0 || k >= nb || i < 0 || j < 0 || i >= P->B[k].n || j >= P->B[k].n) return 2;
        if (D->v[q] == 0) continue;
        if (P->B[k].lp && i != j) return 2;
        if (t == 0) continue;
        cnt[(size_t)k * m + t - 1] += (i == j || P->B[k].lp) ? 1 : 2;
    }
    for (size_t q = 0; q < (size_t)nb * m; q++)
        if (cnt[q]) { P->A[q].r = malloc(sizeof(int) * cnt[q]); P->A[q].c = malloc(sizeof(int) * cnt[q]); P->A[q].v = malloc(sizeof(double) * cnt[q]); }
    for (int64_t q = 0; q < D->nnz; q++) {
        const int t = D->mat[q], k = D->blk[q] - 1;
        int i = D->i[q] - 1, j = D->j[q] - 1;
        const double v = D->v[q];
        if (v == 0) continue;
        if (i > j) { int s = i; i = j; j = s; }
        const OBlk *B = &P->B[k];
        if (t == 0) {   /* C = -F0 */
            if (B->lp) P->C[B->off + i] += -v;
            else { P->C[B->off + i + (size_t)j * B->n] += -v; if (i != j) P->C[B->off + j + (size_t)i * B->n] += -v; }
            continue;
        }
        OEnt *E = &P->A[(size_t)k * m + t - 1];
        E->r[E->nnz] = i; E->c[E->nnz] = j; E->v[E->nnz] = v; E->nnz++;
        if (i != j && !B->lp) { E->r[E->nnz] = j; E->c[E->nnz] = i; E->v[E->nnz] = v; E->nnz++; }
    }
    free(cnt);
    /* norms on the data as read (problem.c:1866-1881) */
    double c2 = 0, c1 = 0;
    for (size_t q = 0; q < off; q++) { c2 += P->C[q] * P->C[q]; c1 += fabs(P->C[q]); }
    P->normC2 = sqrt(c2); P->normC1 = c1;
    for (int i = 0; i < m; i++) { P->normb1 += fabs(P->b0[i]); P->normb2 += P->b0[i] * P->b0[i]; }
    P->normb2 = sqrt(P->normb2);
    /* row equilibration and scaling (problem.c:1884-1909) */
    P->d = malloc(sizeof(double) * (m + 1)); P->du = malloc(sizeof(double) * (m + 1));
    for (int i = 0; i < m; i++) {
        double s = 0;
        for (int k = 0; k < nb; k++) { const OEnt *E = &P->A[(size_t)k * m + i]; for (int q = 0; q < E->nnz; q++) s += E->v[q] * E->v[q]; }
        P->d[i] = s > 0 ? 1.0 / sqrt(s) : 1.0;
    }
    double nbv = 0;
    for (int i = 0; i < m; i++) nbv += P->d[i] * P->b[i] * P->d[i] * P->b[i];
    nbv = sqrt(nbv);
    P->bs = nbv > 1 ? nbv : 1.0;
    for (int k = 0; k < nb; k++) for (int i = 0; i < m; i++) { OEnt *E = &P->A[(size_t)k * m + i]; for (int q = 0; q < E->nnz; q++) E->v[q] *= P->d[i]; }
    for (int i = 0; i < m; i++) { P->b[i] *= P->d[i]; P->du[i] = P->bs / P->d[i]; }
    P->cs = P->normC2 > 1 ? P->normC2 : 1.0;
    for (int i = 0; i < m; i++) P->b[i] /= P->bs;
    for (size_t q = 0; q < off; q++) P->C[q] /= P->cs;
    return 0;
}
static void prob_free(OProb *P) {
    for (size_t q = 0; q < (size_t)P->nb * P->m; q++) { free(P->A[q].r); free(P->A[q].c); free(P->A[q].v); }
    free(P->A); free(P->B); free(P->C); free(P->b); free(P->b0); free(P->d); free(P->du);
}

/* out[i] += alpha * <A_i, D> (D packed over blocks; nonsymmetric D is read as sym(D)) */
static void Aop(const OProb *P, const double *D, double *out, double alpha) {
    for (int i = 0; i < P->m; i++) {
        double s = 0;
        for (int k = 0; k < P->nb; k++) {
            const OEnt *E = &P->A[(size_t)k * P->m + i];
            const OBlk *B = &P->B[k];
            const double *Dk = D + B->off;
            for (int q = 0; q < E->nnz; q++) s += E->v[q] * (B->lp ? Dk[E->r[q]] : Dk[E->r[q] + (size_t)E->c[q] * B->n]);
        }
        out[i] += alpha * s;
    }
}
/* D = sum_i y_i A_i (packed) */
static void ATy(const OProb *P, const double *y, double *D) {
    memset(D, 0, sizeof(double) * P->tot);
    for (int k = 0; k < P->nb; k++) {
        const OBlk *B = &P->B[k];
        double *Dk = D + B->off;
        for (int i = 0; i < P->m; i++) {
            const OEnt *E = &P->A[(size_t)k * P->m + i];
            for (int q = 0; q < E->nnz; q++) {
                if (B->lp) Dk[E->r[q]] += y[i] * E->v[q];
                else Dk[E->r[q] + (size_t)E->c[q] * B->n] += y[i] * E->v[q];
            }
        }
    }
}
/* M (m x m, full) = sum_blocks route-0 products; LP: sum_p a_ip a_jp x_p zi_p */
void oracle_schur_packed(const void *Pv, const double *X, const double *Zi, double *M);
static void schur(const OProb *P, const double *X, const double *Zi, double *M) {
    const int m = P->m;
    memset(M, 0, sizeof(double) * (size_t)m * m);
    for (int k = 0; k < P->nb; k++) {
        const OBlk *B = &P->B[k];
        const double *Xk = X + B->off, *Zk = Zi + B->off;
        const int n = B->n;
        /* blocks whose constraints are dense: M_ab += <A_a, X A_b Zi> with the products formed
         * densely (the reference's route 2, solver.c:2968-3016) - the same sum as below */
        double ent = 0; int ncon = 0;
        for (int a = 0; a < m; a++) { const int e = P->A[(size_t)k * m + a].nnz; if (e) { ent += e; ncon++; } }
        if (!B->lp && ncon > 0 && ent / ncon > 2.0 * n) {
            const size_t nn = (size_t)n * n;
            double *Ad = calloc(nn, sizeof(double)), *G = malloc(sizeof(double) * nn), *W = malloc(sizeof(double) * nn);
            double *Gs = malloc(sizeof(double) * nn * m);
            for (int b = 0; b < m; b++) {
                const OEnt *Eb = &P->A[(size_t)k * m + b];
                if (!Eb->nnz) continue;
                memset(Ad, 0, sizeof(double) * nn);
                for (int q = 0; q < Eb->nnz; q++) Ad[Eb->r[q] + (size_t)Eb->c[q] * n] += Eb->v[q];
                gemm_nn(n, Xk, Ad, W);
                gemm_nn(n, W, Zk, Gs + (size_t)b * nn);
            }
            for (int a = 0; a < m; a++) {
                const OEnt *Ea = &P->A[(size_t)k * m + a];
                if (!Ea->nnz) continue;
                for (int b = 0; b <= a; b++) {
                    if (!P->A[(size_t)k * m + b].nnz) continue;
                    const double *Gb = Gs + (size_t)b * nn;
                    double s = 0;
                    for (int q = 0; q < Ea->nnz; q++) s += Ea->v[q] * Gb[Ea->c[q] + (size_t)Ea->r[q] * n];
                    M[a + (size_t)b * m] += s;
                }
            }
            free(Ad); free(G); free(W); free(Gs);
            continue;
        }
        for (int a = 0; a < m; a++) {
            const OEnt *Ea = &P->A[(size_t)k * m + a];
            if (!Ea->nnz) continue;
            for (int b = 0; b <= a; b++) {
                const OEnt *Eb = &P->A[(size_t)k * m + b];
                if (!Eb->nnz) continue;
                double s = 0;
                for (int p = 0; p < Ea->nnz; p++)
                    for (int q = 0; q < Eb->nnz; q++) {
                        if (B->lp) { if (Ea->r[p] == Eb->r[q]) s += Ea->v[p] * Eb->v[q] * Xk[Ea->r[p]] * Zk[Ea->r[p]]; }
                        else s += Ea->v[p] * Eb->v[q] * Xk[Ea->c[p] + (size_t)Eb->r[q] * n] * Zk[Eb->c[q] + (size_t)Ea->r[p] * n];
                    }
                M[a + (size_t)b * m] += s;
            }
        }
    }
    for (int a = 0; a < m; a++) for (int b = 0; b < a; b++) M[b + (size_t)a * m] = M[a + (size_t)b * m];
}

/* ---- the Schur factor and solve (solver.c:5266-5420, 5483-5622) ---- */
typedef struct { int m; double *M, *L, *D; double reg; int reg_hint; double reg_ok; } OSchur;
static int sc_factor(OSchur *S, int *nfail) {
    const int m = S->m;
    for (int i = 0; i < m; i++) { const double d = S->M[i + (size_t)i * m]; S->D[i] = d > 0 ? 1.0 / sqrt(d) : 1.0; }
    const double floor_ = fmax(1e-15, m * 2.220446049250313e-16);
    double reg = S->reg_hint ? fmax(floor_, 0.01 * S->reg_ok) : 0;
    for (int attempt = 0; attempt < 10; attempt++) {
        for (int j = 0; j < m; j++) for (int i = 0; i < m; i++) S->L[i + (size_t)j * m] = S->M[i + (size_t)j * m] * S->D[i] * S->D[j];
        for (int i = 0; i < m; i++) S->L[i + (size_t)i * m] += reg;
        if (chol(m, S->L) == 0) {
            double pmin = 1.0;
            for (int i = 0; i < m; i++) { double p = S->L[i + (size_t)i * m]; if (p * p < pmin) pmin = p * p; }
            if (pmin >= 1e-14 || reg >= floor_) { S->reg = reg; S->reg_hint = reg > 0; S->reg_ok = reg; return 0; }
            reg = fmax(floor_, reg * 100);
            continue;
        }
        if (nfail) (*nfail)++;
        reg = reg == 0 ? floor_ : reg * 100;
    }
    return 1;
}
static void sc_prec(const OSchur *S, const double *r, double *z) {
    for (int i = 0; i < S->m; i++) z[i] = r[i] * S->D[i];
    lsolve(S->m, S->L, z); ltsolve(S->m, S->L, z);
    for (int i = 0; i < S->m; i++) z[i] *= S->D[i];
}
static void sc_mv(const OSchur *S, const double *x, double *y) {
    const int m = S->m;
    for (int i = 0; i < m; i++) y[i] = 0;
    for (int j = 0; j < m; j++) { const double xj = x[j]; const double *c = S->M + (size_t)j * m; for (int i = 0; i < m; i++) y[i] += c[i] * xj; }
}
static double sc_solve(const OSchur *S, const double *rhs, double *x, double tol_rel) {
    const int m = S->m;
    double nb = sqrt(dot(m, rhs, rhs));
    if (nb == 0) { memset(x, 0, sizeof(double) * m); return 0; }
    double *r = malloc(sizeof(double) * m * 5), *z = r + m, *p = z + m, *q = p + m, *xb = q + m;
    sc_prec(S, rhs, x);
    sc_mv(S, x, q);
    for (int i = 0; i < m; i++) r[i] = rhs[i] - q[i];
    double nr = sqrt(dot(m, r, r));
    const double tol = fmax(tol_rel, 1e-13) * nb;
    if (nr <= tol) { free(r); return nr / nb; }
    double best = nr;
    memcpy(xb, x, sizeof(double) * m);
    sc_prec(S, r, z);
    memcpy(p, z, sizeof(double) * m);
    double rz = dot(m, r, z);
    const int stag = S->reg == 0;
    const int maxit = S->reg > 0 ? 50 : 10, smax = stag ? 1 : 4;
    const double sgain = stag ? 0.5 : 0.9;
    int stall = 0;
    for (int k = 0; k < maxit; k++) {
        sc_mv(S, p, q);
        const double pq = dot(m, p, q);
        if (!(pq > 0)) break;
        const double a = rz / pq;
        for (int i = 0; i < m; i++) { x[i] += a * p[i]; r[i] -= a * q[i]; }
        nr = sqrt(dot(m, r, r));
        if (nr < best) { stall = nr > sgain * best ? stall + 1 : 0; best = nr; memcpy(xb, x, sizeof(double) * m); }
        else stall++;
        if (nr <= tol || stall >= smax) break;
        sc_prec(S, r, z);
        const double rz2 = dot(m, r, z), beta = rz2 / rz;
        rz = rz2;
        for (int i = 0; i < m; i++) p[i] = z[i] + beta * p[i];
    }
    memcpy(x, xb, sizeof(double) * m);
    sc_mv(S, x, q);
    for (int i = 0; i < m; i++) r[i] = rhs[i] - q[i];
    const double rel = sqrt(dot(m, r, r)) / nb;
    free(r);
    return rel;
}

/* ---- the Lanczos bound of the step-length test (reference sdp_maxstep_lam, solver.c:6083-6225,
 * and ritz_bound, solver.c:6029-6038), restated for the kernel-level checks of the GPU's bound;
 * the oracle's own iteration uses the exact lambda_min (above) ---- */
/* the smallest eigenvalue theta of the k x k tridiagonal (a, b) by Sturm bisection, and the last
 * entry of its unit eigenvector by inverse iteration (dense LU with partial pivoting, k <= 64) */
static void tri_smallest(int k, const double *a, const double *b, double *theta, double *zlast) {
    double lo = INFINITY, hi = -INFINITY, sc = 0;
    for (int i = 0; i < k; i++) {
        const double rad = (i > 0 ? fabs(b[i - 1]) : 0) + (i + 1 < k ? fabs(b[i]) : 0);
        lo = fmin(lo, a[i] - rad); hi = fmax(hi, a[i] + rad); sc = fmax(sc, fabs(a[i]) + rad);
    }
    for (int it = 0; it < 300 && hi - lo > 1e-15 * sc; it++) {
        const double x = 0.5 * (lo + hi);
        int cnt = 0;
        double q = a[0] - x;
        if (q < 0) cnt++;
        for (int i = 1; i < k; i++) {
            if (q == 0) q = 1e-300;
            q = a[i] - x - b[i - 1] * b[i - 1] / q;
            if (q < 0) cnt++;
        }
        if (cnt >= 1) hi = x; else lo = x;
    }
    *theta = 0.5 * (lo + hi);
    double M[64 * 64], z[64];
    for (int i = 0; i < k; i++) z[i] = 1.0;
    for (int rep = 0; rep < 3; rep++) {
        const double sh = *theta - 1e-12 * (sc + 1e-300);
        memset(M, 0, sizeof(double) * 64 * 64);
        for (int i = 0; i < k; i++) {
            M[i + 64 * i] = a[i] - sh;
            if (i + 1 < k) M[i + 64 * (i + 1)] = M[(i + 1) + 64 * i] = b[i];
        }
        for (int c = 0; c < k; c++) {            /* LU with partial pivoting, applied to z */
            int pr = c;
            for (int r = c + 1; r < k; r++) if (fabs(M[r + 64 * c]) > fabs(M[pr + 64 * c])) pr = r;
            if (pr != c) {
                for (int j = 0; j < k; j++) { const double t = M[c + 64 * j]; M[c + 64 * j] = M[pr + 64 * j]; M[pr + 64 * j] = t; }
                const double t = z[c]; z[c] = z[pr]; z[pr] = t;
            }
            if (M[c + 64 * c] == 0) M[c + 64 * c] = 1e-300;
            for (int r = c + 1; r < k; r++) {
                const double f = M[r + 64 * c] / M[c + 64 * c];
                for (int j = c; j < k; j++) M[r + 64 * j] -= f * M[c + 64 * j];
                z[r] -= f * z[c];
            }
        }
        for (int r = k - 1; r >= 0; r--) {
            double t = z[r];
            for (int j = r + 1; j < k; j++) t -= M[r + 64 * j] * z[j];
            z[r] = t / M[r + 64 * r];
        }
        double nz = 0;
        for (int i = 0; i < k; i++) nz += z[i] * z[i];
        nz = 1.0 / sqrt(nz);
        for (int i = 0; i < k; i++) z[i] *= nz;
    }
    *zlast = z[k - 1];
}
/* Lanczos on T = L^{-1} D L^{-T} (L lower, n x n; T = D when L is NULL) from the reference's
 * fixed start 1 + 0.5 sin(1.7 i + 0.3): at most K steps (K <= 64); each w = T q_j, alpha_j = q_j'w,
 * two Gram-Schmidt passes over q_0..q_j, beta_j = |w|; a breakdown when beta_j <= 1e-12 max|alpha|
 * (beta_j = 0); from the 6th step, every 3rd, the Ritz test res <= tol |theta| + 1e-14 max|alpha|
 * stops it. Returns theta - res for the final tridiagonal (res = |beta_last z_last|), or -1e300
 * for coefficients that are not finite; info3 = { steps, theta, res } when given. */
double oracle_lanczos(int n, const double *L, const double *D, int K, double tol, double *info3) {
    if (K > n) K = n;
    if (K > 64) K = 64;
    double *Q = malloc(sizeof(double) * (size_t)n * (K + 1)), *w = malloc(sizeof(double) * n), *v = malloc(sizeof(double) * n);
    double al[64], be[64], theta = 0, zl = 0, res = 0, scale = 0;
    double nrm = 0;
    for (int i = 0; i < n; i++) { Q[i] = 1.0 + 0.5 * sin(1.7 * i + 0.3); nrm += Q[i] * Q[i]; }
    nrm = 1.0 / sqrt(nrm);
    for (int i = 0; i < n; i++) Q[i] *= nrm;
    int kk = 0;
    for (int j = 0; j < K; j++) {
        const double *qj = Q + (size_t)j * n;
        memcpy(v, qj, sizeof(double) * n);
        if (L) ltsolve(n, L, v);
        for (int i = 0; i < n; i++) { double t = 0; for (int c = 0; c < n; c++) t += D[i + (size_t)c * n] * v[c]; w[i] = t; }
        if (L) lsolve(n, L, w);
        const double a = dot(n, qj, w);
        al[j] = a;
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i <= j; i++) {
                const double *qi = Q + (size_t)i * n;
                const double c = dot(n, qi, w);
                for (int r = 0; r < n; r++) w[r] -= c * qi[r];
            }
        const double b = sqrt(dot(n, w, w));
        be[j] = b;
        kk = j + 1;
        if (fabs(a) > scale) scale = fabs(a);
        if (b <= 1e-12 * (scale + 1e-300)) { be[j] = 0; break; }
        if (kk >= 6 && kk % 3 == 0) {
            tri_smallest(kk, al, be, &theta, &zl);
            res = fabs(be[kk - 1] * zl);
            if (res <= tol * fabs(theta) + 1e-14 * scale) break;
        }
        if (j + 1 < K) { double *qn = Q + (size_t)(j + 1) * n; for (int i = 0; i < n; i++) qn[i] = w[i] / b; }
    }
    int bad = 0;
    for (int t = 0; t < kk; t++) if (!isfinite(al[t]) || !isfinite(be[t])) bad = 1;
    double out = -1e300;
    if (!bad) {
        tri_smallest(kk, al, be, &theta, &zl);
        res = fabs(be[kk - 1] * zl);
        out = theta - res;
    }
    if (info3) { info3[0] = kk; info3[1] = theta; info3[2] = res; }
    free(Q); free(w); free(v);
    return out;
}
/* the reference's bound of lambda_min(L^{-1} D L^{-T}) for one block (sdp_maxstep_lam,
 * solver.c:6083-6090): the exact value for blocks of order <= 32 (small_lammin), else the
 * 30-step Lanczos bound with Ritz accuracy 0.02 */
double oracle_step_lambda(int n, const double *L, const double *D) {
    if (n > 32) return oracle_lanczos(n, L, D, 30, 0.02, NULL);
    double *T = malloc(sizeof(double) * (size_t)n * n);
    memcpy(T, D, sizeof(double) * (size_t)n * n);
    if (L) {
        for (int j = 0; j < n; j++) lsolve(n, L, T + (size_t)j * n);
        for (int j = 0; j < n; j++) for (int i = j + 1; i < n; i++) { const double t = T[i + (size_t)j * n]; T[i + (size_t)j * n] = T[j + (size_t)i * n]; T[j + (size_t)i * n] = t; }
        for (int j = 0; j < n; j++) lsolve(n, L, T + (size_t)j * n);
        symmetrize(n, T);
    }
    const double lam = oracle_lammin(n, T);
    free(T);
    return lam;
}

/* ---- step lengths (solver.c:6300-6388 with the exact lambda_min) ---- */
static double blk_maxstep(const OBlk *B, const double *V, const double *L, const double *dV, double *work) {
    const int n = B->n;
    if (B->lp) {
        double a = 1e30;
        for (int i = 0; i < n; i++) { if (!isfinite(dV[i])) return 0; if (dV[i] < 0) a = fmin(a, -V[i] / dV[i]); }
        return a;
    }
    /* T = L^{-1} dV L^{-T} */
    double *T = work;
    memcpy(T, dV, sizeof(double) * (size_t)n * n);
    for (int j = 0; j < n; j++) lsolve(n, L, T + (size_t)j * n);
    for (int j = 0; j < n; j++) for (int i = j + 1; i < n; i++) { double s = T[i + (size_t)j * n]; T[i + (size_t)j * n] = T[j + (size_t)i * n]; T[j + (size_t)i * n] = s; }
    for (int j = 0; j < n; j++) lsolve(n, L, T + (size_t)j * n);
    symmetrize(n, T);
    const double lam = oracle_lammin(n, T);
    if (!isfinite(lam)) return 0;
    return lam >= 0 ? 1e30 : -1.0 / lam;
}
static double maxstep(const OProb *P, const double *V, const double *L, const double *dV, double *work) {
    double a = 1e30;
    for (int k = 0; k < P->nb; k++) {
        const OBlk *B = &P->B[k];
        const double ak = blk_maxstep(B, V + B->off, L + B->off, dV + B->off, work);
        if (ak < a) a = ak;
    }
    return a;
}
/* V += (a - stepped) dV, L = chol(V); x0.8 on failure (solver.c:6393-6446) */
static int try_step(const OProb *P, double *V, double *L, const double *dV, double *alpha) {
    double a = *alpha, stepped = 0;
    for (int tries = 0; tries < 40; tries++) {
        const double inc = a - stepped;
        stepped = a;
        int ok = 1;
        for (size_t q = 0; q < P->tot; q++) V[q] += inc * dV[q];
        for (int k = 0; k < P->nb; k++) {
            const OBlk *B = &P->B[k];
            if (B->lp) { for (int i = 0; i < B->n; i++) { L[B->off + i] = V[B->off + i]; if (!(V[B->off + i] > 0)) ok = 0; } continue; }
            memcpy(L + B->off, V + B->off, sizeof(double) * (size_t)B->n * B->n);
            if (chol(B->n, L + B->off)) ok = 0;
        }
        if (ok) { *alpha = a; return 1; }
        a *= 0.8;
    }
    return 0;
}
/* G = Zi F R per block (LP: elementwise) */
static void prod3(const OProb *P, const double *Zi, const double *F, const double *R, double *G, double *W) {
    for (int k = 0; k < P->nb; k++) {
        const OBlk *B = &P->B[k];
        if (B->lp) { for (int i = 0; i < B->n; i++) G[B->off + i] = Zi[B->off + i] * F[B->off + i] * R[B->off + i]; continue; }
        gemm_nn(B->n, Zi + B->off, F + B->off, W);
        gemm_nn(B->n, W, R + B->off, G + B->off);
    }
}
static void sym_all(const OProb *P, double *G) { for (int k = 0; k < P->nb; k++) if (!P->B[k].lp) symmetrize(P->B[k].n, G + P->B[k].off); }
static void add_identity(const OProb *P, double *V, double s) {
    for (int k = 0; k < P->nb; k++) { const OBlk *B = &P->B[k]; for (int i = 0; i < B->n; i++) V[B->off + (B->lp ? (size_t)i : (size_t)i * B->n + i)] += s; }
}
static double lammin_all(const OProb *P, const double *V, double *work) {
    double lam = INFINITY;
    for (int k = 0; k < P->nb; k++) {
        const OBlk *B = &P->B[k];
        if (B->lp) { for (int i = 0; i < B->n; i++) lam = fmin(lam, V[B->off + i]); continue; }
        memcpy(work, V + B->off, sizeof(double) * (size_t)B->n * B->n);
        lam = fmin(lam, oracle_lammin(B->n, work));
    }
    return lam;
}

/* lambda_min for the DIMACS cone errors as the reference's ps_lammin_tol (postsolve.c:668-690):
 * 0 when the block has a Cholesky factor, -delta when only V + delta I has one, else the
 * eigenvalue itself; LP blocks by their smallest entry. */
static double lammin_tol_all(const OProb *P, const double *V, double *work, double delta) {
    double lam = 0;
    for (int k = 0; k < P->nb; k++) {
        const OBlk *B = &P->B[k];
        const int n = B->n;
        if (B->lp) { for (int i = 0; i < n; i++) lam = fmin(lam, V[B->off + i]); continue; }
        memcpy(work, V + B->off, sizeof(double) * (size_t)n * n);
        if (chol(n, work) == 0) continue;
        memcpy(work, V + B->off, sizeof(double) * (size_t)n * n);
        for (int i = 0; i < n; i++) work[i + (size_t)i * n] += delta;
        if (chol(n, work) == 0) { lam = fmin(lam, -delta); continue; }
        memcpy(work, V + B->off, sizeof(double) * (size_t)n * n);
        lam = fmin(lam, oracle_lammin(n, work));
    }
    return lam;
}

/* ---- DIMACS errors on the data as read (the measure of postsolve, postsolve.c:2744-2897) ---- */
void oracle_dimacs_packed(const OProb *P, const double *Xo, const double *yo, double *err, double *pobj, double *dobj);
static void dimacs(const OProb *P, const double *Xo, const double *yo, double *err, double *po, double *dob, double *work) {
    const int m = P->m;
    /* unscaled data: A_i = Ascaled_i / d_i, C = Cscaled * cs */
    double *Zo = malloc(sizeof(double) * (P->tot + 1)), *ys = malloc(sizeof(double) * (m + 1)), *rp = calloc(m + 1, sizeof(double));
    for (int i = 0; i < m; i++) ys[i] = yo[i] / P->d[i];          /* sum yo_i A_i = sum (yo_i / d_i) Ascaled_i */
    ATy(P, ys, Zo);
    for (size_t q = 0; q < P->tot; q++) Zo[q] = P->C[q] * P->cs - Zo[q];
    Aop(P, Xo, rp, 1.0);
    double pobj = 0, dobj = 0, xz = 0, r2 = 0;
    for (size_t q = 0; q < P->tot; q++) { pobj += P->C[q] * P->cs * Xo[q]; xz += Xo[q] * Zo[q]; }
    for (int i = 0; i < m; i++) { dobj += P->b0[i] * yo[i]; const double v = rp[i] / P->d[i] - P->b0[i]; r2 += v * v; }
    const double lamx = lammin_tol_all(P, Xo, work, 1e-12 * (1 + P->normb1));
    const double lamz = lammin_tol_all(P, Zo, work, 1e-12 * (1 + P->normC1));
    const double den = 1 + fabs(pobj) + fabs(dobj);
    err[0] = sqrt(r2) / (1 + P->normb1);
    err[1] = fmax(0.0, -fmin(lamx, 0.0)) / (1 + P->normb1);
    err[2] = 0;
    err[3] = fmax(0.0, -fmin(lamz, 0.0)) / (1 + P->normC1);
    err[4] = (pobj - dobj) / den;
    err[5] = xz / den;
    *po = pobj; *dob = dobj;
    free(Zo); free(ys); free(rp);
}

/* ------------------------------------------------------------------------------------- */
static int oracle_solve_hsd(const OracleData *Dd, const OracleOptions *opt, OracleResult *res);
int oracle_solve(const OracleData *Dd, const OracleOptions *opt, OracleResult *res) {
    if (opt && opt->method == 1) return oracle_solve_hsd(Dd, opt, res);
    memset(res, 0, sizeof(*res));
    res->status = -1;
    OProb P;
    int rc = prob_build(Dd, &P);
    if (rc) return rc;
    const int m = P.m;
    const size_t T = P.tot;
    const double tol = opt && opt->tol > 0 ? opt->tol : 1e-8;
    const int maxit = opt && opt->maxit > 0 ? opt->maxit : 100;
    int maxn = 1;
    for (int k = 0; k < P.nb; k++) if (P.B[k].n > maxn) maxn = P.B[k].n;
    double *X = calloc(T + 1, sizeof(double)), *Z = calloc(T + 1, sizeof(double)), *Zi = calloc(T + 1, sizeof(double));
    double *LX = calloc(T + 1, sizeof(double)), *LZ = calloc(T + 1, sizeof(double)), *dX = calloc(T + 1, sizeof(double));
    double *dZ = calloc(T + 1, sizeof(double)), *F = calloc(T + 1, sizeof(double)), *G = calloc(T + 1, sizeof(double));
    double *Q = calloc(T + 1, sizeof(double)), *R0 = calloc(T + 1, sizeof(double)), *Xb = calloc(T + 1, sizeof(double));
    double *W = malloc(sizeof(double) * (size_t)maxn * maxn + 8), *work = malloc(sizeof(double) * (size_t)maxn * maxn + 8);
    double *y = calloc(m + 1, sizeof(double)), *dy = calloc(m + 1, sizeof(double)), *rp = calloc(m + 1, sizeof(double));
    double *rhs = calloc(m + 1, sizeof(double)), *a0 = calloc(m + 1, sizeof(double)), *h = calloc(m + 1, sizeof(double));
    double *qv = calloc(m + 1, sizeof(double)), *ev = calloc(m + 1, sizeof(double)), *wv = calloc(m + 1, sizeof(double)), *yb = calloc(m + 1, sizeof(double));
    OSchur S = { m, calloc((size_t)m * m + 1, sizeof(double)), calloc((size_t)m * m + 1, sizeof(double)), calloc(m + 1, sizeof(double)), 0, 0, 0 };
    int status = 3, it = 0, stall = 0, it_best = 0, nfail = 0;
    double best_score = 1e300;
    double best_err[7] = { 0 };

    /* ---- start: least-norm X0 = A'(AA')^{-1} b, least-squares y0 = (AA')^{-1} A(C) ---- */
    memset(X, 0, sizeof(double) * T);
    add_identity(&P, X, 1.0);
    memcpy(Zi, X, sizeof(double) * T);
    memset(rhs, 0, sizeof(double) * m);
    Aop(&P, P.C, rhs, 1.0);
    schur(&P, X, Zi, S.M);
    if (sc_factor(&S, &nfail)) { status = 4; goto done; }
    sc_solve(&S, P.b, dy, 1e-8);
    sc_solve(&S, rhs, y, 1e-8);
    ATy(&P, dy, X);
    ATy(&P, y, F);
    for (size_t q = 0; q < T; q++) Z[q] = P.C[q] - F[q];
    {
        const double ax = -lammin_all(&P, X, work), az = -lammin_all(&P, Z, work);
        const double shx = ax >= -1e-8 ? 1.0 + ax : 0.0;
        const double shz = fmax(0.0, 1.0 + az);
        add_identity(&P, X, shx);
        add_identity(&P, Z, shz);
        for (int k = 0; k < P.nb; k++) {
            const OBlk *B = &P.B[k];
            const size_t len = B->lp ? (size_t)B->n : (size_t)B->n * B->n;
            if (B->lp) { memcpy(LX + B->off, X + B->off, sizeof(double) * len); memcpy(LZ + B->off, Z + B->off, sizeof(double) * len); continue; }
            for (double extra = 1e-8; ; extra *= 10) {
                memcpy(LX + B->off, X + B->off, sizeof(double) * len);
                if (chol(B->n, LX + B->off) == 0) break;
                for (int i = 0; i < B->n; i++) X[B->off + i + (size_t)i * B->n] += extra;
            }
            for (double extra = 1e-8; ; extra *= 10) {
                memcpy(LZ + B->off, Z + B->off, sizeof(double) * len);
                if (chol(B->n, LZ + B->off) == 0) break;
                for (int i = 0; i < B->n; i++) Z[B->off + i + (size_t)i * B->n] += extra;
            }
        }
    }
    /* ---- R0, theta ---- */
    double theta = 1.0, r0norm = 0;
    ATy(&P, y, F);
    for (size_t q = 0; q < T; q++) { R0[q] = P.C[q] - Z[q] - F[q]; r0norm += R0[q] * R0[q]; }
    r0norm = sqrt(r0norm);
    const double nbs = sqrt(dot(m, P.b, P.b));
    double alp = 0, ald = 0;
    memcpy(yb, y, sizeof(double) * m); memcpy(Xb, X, sizeof(double) * T);

    for (it = 0; it <= maxit; it++) {
        /* Zi */
        for (int k = 0; k < P.nb; k++) {
            const OBlk *B = &P.B[k];
            if (B->lp) { for (int i = 0; i < B->n; i++) Zi[B->off + i] = 1.0 / Z[B->off + i]; }
            else inv_from_chol(B->n, LZ + B->off, Zi + B->off);
        }
        /* residuals and objectives (solver.c:14426-14482) */
        memcpy(rp, P.b, sizeof(double) * m);
        Aop(&P, X, rp, -1.0);
        ATy(&P, y, F);
        double ncr2 = 0, pobj = 0, xz = 0, rdn2 = 0, dev2 = 0;
        for (size_t q = 0; q < T; q++) {
            const double f = F[q] + Z[q];
            ncr2 += f * f;
            F[q] = f - P.C[q];
            pobj += P.C[q] * X[q]; xz += X[q] * Z[q]; rdn2 += F[q] * F[q];
            const double w = F[q] + theta * R0[q]; dev2 += w * w;
        }
        const double dobj = dot(m, P.b, y), mu = xz / P.Ntot;
        double rpo = 0;
        for (int i = 0; i < m; i++) { const double v = rp[i] * P.du[i]; rpo += v * v; }
        rpo = sqrt(rpo);
        const double rdo = P.cs * sqrt(rdn2), sc = P.bs * P.cs;
        const double po = sc * pobj, dob = sc * dobj;
        const double pinf = rpo / (1 + P.normb2), dinf = rdo / (1 + P.normC2);
        const double oden = 1 + fabs(po) + fabs(dob);
        const double relgap = fabs(po - dob) / oden, relcomp = sc * xz / oden;
        double score = fmax(fmax(relgap, relcomp), fmax(pinf, dinf));
        if (!isfinite(score)) score = INFINITY;
        if (it < ORACLE_TRACE) { res->trace_pobj[it] = po; res->trace_dobj[it] = dob; res->trace_mu[it] = sc * mu; res->trace_score[it] = score; }
        res->iters = it;
        if (!isfinite(po) || !isfinite(dob)) { status = 4; break; }
        if (score < best_score) {
            best_score = score; it_best = it;
            memcpy(yb, y, sizeof(double) * m); memcpy(Xb, X, sizeof(double) * T);
            best_err[1] = pinf; best_err[3] = dinf;
        }
        if (score < tol) { status = 0; break; }
        const double nCR = sqrt(ncr2);
        if (dobj > 0 && nCR / dobj < 1e-8 && it > 5) { status = 1; break; }
        double nAX = 0;
        for (int i = 0; i < m; i++) { const double v = P.b[i] - rp[i]; nAX += v * v; }
        nAX = sqrt(nAX);
        if (pobj < 0 && nAX / (-pobj) < 1e-8 && it > 5) { status = 2; break; }
        if (it == maxit) { status = 3; break; }
        if (stall >= 3) { status = 4; break; }
        if (it - it_best >= 5 && best_score < 1e-6) { status = 4; break; }
        if (it - it_best >= 8 && best_score < 1e-4) { status = 4; break; }
        if (it - it_best >= 12) { status = 4; break; }
        /* re-base the dual residual (solver.c:14614-14636) */
        {
            const double dev = sqrt(dev2);
            if (dev > 1e-10 * (1.0 + theta * r0norm) && dev > 1e-12) {
                theta = 1.0; r0norm = sqrt(rdn2);
                for (size_t q = 0; q < T; q++) R0[q] = -F[q];
            }
            if (theta * r0norm <= 1e-13) theta = 0.0;
        }
        /* Schur complement, factor */
        schur(&P, X, Zi, S.M);
        if (sc_factor(&S, &nfail)) { status = 4; break; }
        /* a0 = A(Zi), h = A(Zi theta R0 X) */
        memset(a0, 0, sizeof(double) * m); memset(h, 0, sizeof(double) * m);
        Aop(&P, Zi, a0, 1.0);
        const int hasH = theta > 0;
        if (hasH) {
            for (size_t q = 0; q < T; q++) dZ[q] = theta * R0[q];
            prod3(&P, Zi, dZ, X, G, W);
            Aop(&P, G, h, 1.0);
        }
        /* predictor */
        for (int i = 0; i < m; i++) rhs[i] = P.b[i] + h[i];
        sc_solve(&S, rhs, dy, 1e-7);
        ATy(&P, dy, F);
        for (size_t q = 0; q < T; q++) dZ[q] = -F[q] + theta * R0[q];
        prod3(&P, Zi, dZ, X, G, W);
        sym_all(&P, G);
        for (size_t q = 0; q < T; q++) dX[q] = -G[q] - X[q];
        double ap = fmin(1.0, maxstep(&P, X, LX, dX, work));
        double ad = fmin(1.0, maxstep(&P, Z, LZ, dZ, work));
        const double xdz = dot(T, X, dZ), dxz = dot(T, dX, Z), dxdz = dot(T, dX, dZ);
        const double mu_a = (xz + ap * dxz + ad * xdz + ap * ad * dxdz) / P.Ntot;
        const double frac = fmax(0.0, mu_a / mu), mina = fmin(ap, ad);
        int anysdp = 0;
        for (int k = 0; k < P.nb; k++) if (!P.B[k].lp) anysdp = 1;
        const double expon = !anysdp ? 3.0 : (relgap > 1e-6 ? fmax(1.0, 3 * mina * mina) : 3.0);
        const double sigma = fmin(1.0, pow(frac, expon));
        const double gamma = 0.9 + (0.99 - 0.9) * mina;
        const double smu = sigma * mu;
        /* corrector (solver.c:14997-15055) */
        prod3(&P, Zi, dZ, dX, Q, W);
        memset(qv, 0, sizeof(double) * m);
        Aop(&P, Q, qv, 1.0);
        for (int i = 0; i < m; i++) rhs[i] = P.b[i] - smu * a0[i] + qv[i] + h[i];
        sc_solve(&S, rhs, dy, 0);
        ATy(&P, dy, F);
        for (size_t q = 0; q < T; q++) dZ[q] = -F[q] + theta * R0[q];
        prod3(&P, Zi, dZ, X, G, W);
        for (int k = 0; k < P.nb; k++) {
            const OBlk *B = &P.B[k];
            const size_t len = B->lp ? (size_t)B->n : (size_t)B->n * B->n;
            double *dXk = dX + B->off;
            const double *Gk = G + B->off, *Qk = Q + B->off, *Zik = Zi + B->off, *Xk = X + B->off;
            if (B->lp) { for (int i = 0; i < B->n; i++) dXk[i] = smu * Zik[i] - Zik[i] * dZ[B->off + i] * Xk[i] - Qk[i] - Xk[i]; continue; }
            for (size_t q = 0; q < len; q++) dXk[q] = -Gk[q] - Qk[q];
            symmetrize(B->n, dXk);
            for (size_t q = 0; q < len; q++) dXk[q] += smu * Zik[q] - Xk[q];
        }
        /* refinement enforcing A(dX) = rp (solver.c:15058-15105) */
        for (int pass = 0; pass < 3; pass++) {
            memcpy(ev, rp, sizeof(double) * m);
            Aop(&P, dX, ev, -1.0);
            const double ne = sqrt(dot(m, ev, ev)), nrp = sqrt(dot(m, rp, rp));
            if (!(ne > 1e-2 * nrp && ne > 1e-3 * tol * (1 + nbs))) break;
            sc_solve(&S, ev, wv, 0);
            memcpy(qv, ev, sizeof(double) * m);
            ATy(&P, wv, Q);
            prod3(&P, Zi, Q, X, G, W);
            sym_all(&P, G);
            Aop(&P, G, qv, -1.0);
            const double ne2 = sqrt(dot(m, qv, qv));
            if (!(ne2 < 0.5 * ne)) break;
            for (int i = 0; i < m; i++) dy[i] += wv[i];
            for (size_t q = 0; q < T; q++) { dZ[q] -= Q[q]; dX[q] += G[q]; }
        }
        const double apr = fmin(1.0, maxstep(&P, X, LX, dX, work));
        const double adr = fmin(1.0, maxstep(&P, Z, LZ, dZ, work));
        ap = fmin(1.0, gamma * apr);
        ad = fmin(1.0, gamma * adr);
        if (pinf < 1e-6) {   /* catastrophe guard (solver.c:15143-15160) */
            memcpy(ev, rp, sizeof(double) * m);
            Aop(&P, dX, ev, -1.0);
            const double nrp = sqrt(dot(m, rp, rp)), nev = sqrt(dot(m, ev, ev));
            const double g = fmax(1e3, 1e-6 / fmax(pinf, 1e-300)), lim = g * nrp;
            if (nev > lim && nrp > 0) { const double cap = lim / nev; if (cap < ap) ap = cap; }
        }
        if (!try_step(&P, X, LX, dX, &ap) || !try_step(&P, Z, LZ, dZ, &ad)) { status = 4; break; }
        for (int i = 0; i < m; i++) y[i] += ad * dy[i];
        theta *= (1.0 - ad);
        alp = ap; ald = ad;
        res->trace_ap[it] = ap; res->trace_ad[it] = ad;
        stall = (ap < 1e-6 && ad < 1e-6) ? stall + 1 : 0;
    }
    (void)alp; (void)ald; (void)best_err;
    if (status != 0) { memcpy(y, yb, sizeof(double) * m); memcpy(X, Xb, sizeof(double) * T); }
done:
    /* unscale: X = bs X~, y_i = cs d_i y~_i (Z follows from y on the data as read) */
    res->m = m; res->nblk = P.nb;
    res->y = malloc(sizeof(double) * (m + 1));
    res->X = malloc(sizeof(double) * (T + 1));
    for (int i = 0; i < m; i++) res->y[i] = P.cs * P.d[i] * y[i];
    for (size_t q = 0; q < T; q++) res->X[q] = P.bs * X[q];
    res->xlen = T;
    {
        double pobj, dobj;
        dimacs(&P, res->X, res->y, res->err, &pobj, &dobj, work);
        res->pobj = pobj; res->dobj = dobj;
        double acc = fmax(fmax(res->err[0], res->err[1]), fmax(res->err[3], fmax(fabs(res->err[4]), fabs(res->err[5]))));
        if (!isfinite(acc)) acc = INFINITY;
        if (status == 0 || status == 3 || status == 4 || status == 5) {
            if (acc <= tol) status = 0;
            else if (acc <= 1e-6) status = 5;
            else if (status == 0) status = 4;
        }
        res->acc = acc;
    }
    res->status = status;
    free(X); free(Z); free(Zi); free(LX); free(LZ); free(dX); free(dZ); free(F); free(G); free(Q); free(R0); free(Xb);
    free(W); free(work); free(y); free(dy); free(rp); free(rhs); free(a0); free(h); free(qv); free(ev); free(wv); free(yb);
    free(S.M); free(S.L); free(S.D);
    prob_free(&P);
    return 0;
}

void oracle_result_free(OracleResult *r) { if (r) { free(r->y); free(r->X); r->y = NULL; r->X = NULL; } }

/* DIMACS errors of a given (X, y) on the data as read - the same measurement the solve ends
 * with, exposed so a test can measure the GPU path's returned point independently */
int oracle_dimacs(const OracleData *Dd, const double *X, const double *y, double *err, double *pobj, double *dobj) {
    OProb P;
    int rc = prob_build(Dd, &P);
    if (rc) return rc;
    int maxn = 1;
    for (int k = 0; k < P.nb; k++) if (P.B[k].n > maxn) maxn = P.B[k].n;
    double *work = malloc(sizeof(double) * (size_t)maxn * maxn + 8);
    dimacs(&P, X, y, err, pobj, dobj, work);
    free(work);
    prob_free(&P);
    return 0;
}

/* the route-0 Schur complement of the SCALED problem at given packed X, Zi (tests of the
 * assembly kernel): M_ab = <A~_a, X A~_b Zi> */
int oracle_schur(const OracleData *Dd, const double *X, const double *Zi, double *M) {
    OProb P;
    int rc = prob_build(Dd, &P);
    if (rc) return rc;
    schur(&P, X, Zi, M);
    prob_free(&P);
    return 0;
}

/* ---- the end phase's metric corrections (reference sqrt_restore, solver.c:7551-7655, and
 * polish_xmetric, solver.c:7657-7760) ---------------------------------------------------------- */
/* A = V diag(w) V' by the row-cyclic Jacobi method (Golub and Van Loan, Matrix Computations,
 * 4th ed., algorithm 8.5.3): sweeps over (p, q), p < q, each rotation zeroing A(p, q), until the
 * off-diagonal norm is below 1e-15 of the Frobenius norm. A is destroyed. */
static void jacobi_eig(int n, double *A, double *V, double *w) {
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) V[i + (size_t)j * n] = i == j;
    double fro = 0;
    for (size_t e = 0; e < (size_t)n * n; e++) fro += A[e] * A[e];
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0;
        for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) if (i != j) off += A[i + (size_t)j * n] * A[i + (size_t)j * n];
        if (off <= 1e-30 * fro) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[p + (size_t)q * n];
                if (apq == 0) continue;
                const double tau = (A[q + (size_t)q * n] - A[p + (size_t)p * n]) / (2 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
                const double c = 1 / sqrt(1 + t * t), s = t * c;
                for (int k = 0; k < n; k++) {       /* A <- A J (columns p, q) */
                    const double akp = A[k + (size_t)p * n], akq = A[k + (size_t)q * n];
                    A[k + (size_t)p * n] = c * akp - s * akq; A[k + (size_t)q * n] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {       /* A <- J' A (rows p, q) */
                    const double apk = A[p + (size_t)k * n], aqk = A[q + (size_t)k * n];
                    A[p + (size_t)k * n] = c * apk - s * aqk; A[q + (size_t)k * n] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {       /* V <- V J */
                    const double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
                    V[k + (size_t)p * n] = c * vkp - s * vkq; V[k + (size_t)q * n] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = A[i + (size_t)i * n];
}
/* X <- X + a dX with M w = b - A(X), W = A*(w) (scaled problem) and
 *   half = 1: dX = D W D, D = X^{1/2} (eigenvalues below 0 taken as 0), M = A (D (x) D) A';
 *             a, from 1 and halved up to eight times, keeps (1 - 0.01) I + a W positive definite
 *             in every block; at most three passes, ending when the residual falls to 1e-6 of
 *             its first value;
 *   half = 0: dX = sym(Zi W X), Zi = (C - A*(y))^-1, M = A (X (x) Zi) A'; a, halved up to six
 *             times, keeps X + a dX positive definite; at most two passes.
 * One step per pass for all blocks (the GPU path's rule; the reference steps each block by its
 * own and takes the same steps when every block accepts a = 1). M is factored and solved by the
 * oracle's restatement of the FP64 factor and PCG, to 1e-14. X and y on the data as read (X
 * packed); Xout the corrected X, steps[pass] the step of each pass taken (0: none). Returns 0, or
 * 3 when a pass found no factor of M, no step, or (half = 0) no factor of Z or X. */
int oracle_correct(const OracleData *Dd, const double *Xin, const double *yin, int half, int maxpass, double *Xout, double *steps) {
    OProb P;
    int rc = prob_build(Dd, &P);
    if (rc) return rc;
    const int m = P.m;
    const size_t T = P.tot;
    int maxn = 1;
    for (int k = 0; k < P.nb; k++) if (P.B[k].n > maxn) maxn = P.B[k].n;
    const size_t nn = (size_t)maxn * maxn;
    double *X = malloc(sizeof(double) * (T + 1)), *Dm = malloc(sizeof(double) * (T + 1)), *Zi = malloc(sizeof(double) * (T + 1));
    double *W = malloc(sizeof(double) * (T + 1)), *dX = malloc(sizeof(double) * (T + 1)), *tmp = malloc(sizeof(double) * (T + 1));
    double *w1 = malloc(sizeof(double) * (nn + 1)), *w2 = malloc(sizeof(double) * (nn + 1)), *ev = malloc(sizeof(double) * (maxn + 1));
    double *r = malloc(sizeof(double) * (m + 1)), *wv = malloc(sizeof(double) * (m + 1)), *ys = malloc(sizeof(double) * (m + 1));
    OSchur S = { m, malloc(sizeof(double) * (size_t)m * m + 8), malloc(sizeof(double) * (size_t)m * m + 8), malloc(sizeof(double) * (m + 1)), 0, 0, 0 };
    for (size_t q = 0; q < T; q++) X[q] = Xin[q] / P.bs;
    for (int i = 0; i < m; i++) ys[i] = yin[i] / (P.cs * P.d[i]);
    for (int p = 0; p < maxpass; p++) steps[p] = 0;
    rc = 3;
    if (!half) {
        /* Zi = (C - A*(y))^-1 per block */
        ATy(&P, ys, Zi);
        for (size_t q = 0; q < T; q++) Zi[q] = P.C[q] - Zi[q];
        for (int k = 0; k < P.nb; k++) {
            const OBlk *B = &P.B[k];
            double *Zk = Zi + B->off;
            if (B->lp) { for (int i = 0; i < B->n; i++) { if (!(Zk[i] > 0)) goto out; Zk[i] = 1 / Zk[i]; } continue; }
            memcpy(w1, Zk, sizeof(double) * B->n * B->n);
            if (chol(B->n, w1)) goto out;
            inv_from_chol(B->n, w1, Zk);
        }
    }
    double nr0 = 0;
    for (int pass = 0; pass < maxpass; pass++) {
        memcpy(r, P.b, sizeof(double) * m);
        Aop(&P, X, r, -1.0);
        const double nr = sqrt(dot(m, r, r));
        if (pass == 0) nr0 = nr; else if (nr <= 1e-6 * nr0) break;
        if (half) {
            for (int k = 0; k < P.nb; k++) {
                const OBlk *B = &P.B[k];
                const int n = B->n;
                double *Dk = Dm + B->off;
                if (B->lp) { for (int i = 0; i < n; i++) Dk[i] = X[B->off + i] > 0 ? sqrt(X[B->off + i]) : 0.0; continue; }
                memcpy(w1, X + B->off, sizeof(double) * n * n);
                jacobi_eig(n, w1, w2, ev);
                for (int j = 0; j < n; j++) {
                    const double f = ev[j] > 0 && ev[j] < INFINITY ? sqrt(sqrt(ev[j])) : 0.0;
                    for (int i = 0; i < n; i++) w2[i + (size_t)j * n] *= f;
                }
                for (int j = 0; j < n; j++)
                    for (int i = 0; i < n; i++) {
                        double s_ = 0;
                        for (int t = 0; t < n; t++) s_ += w2[i + (size_t)t * n] * w2[j + (size_t)t * n];
                        Dk[i + (size_t)j * n] = s_;
                    }
                symmetrize(n, Dk);
            }
            schur(&P, Dm, Dm, S.M);
        } else {
            for (int k = 0; k < P.nb; k++) {
                const OBlk *B = &P.B[k];
                if (B->lp) { for (int i = 0; i < B->n; i++) if (!(X[B->off + i] > 0)) goto out; continue; }
                memcpy(w1, X + B->off, sizeof(double) * B->n * B->n);
                if (chol(B->n, w1)) goto out;
            }
            schur(&P, X, Zi, S.M);
        }
        if (sc_factor(&S, NULL)) goto out;
        memset(wv, 0, sizeof(double) * m);
        sc_solve(&S, r, wv, 1e-14);
        ATy(&P, wv, W);
        for (int k = 0; k < P.nb; k++) {
            const OBlk *B = &P.B[k];
            const int n = B->n;
            const size_t o = B->off;
            if (B->lp) {
                for (int i = 0; i < n; i++) dX[o + i] = half ? Dm[o + i] * W[o + i] * Dm[o + i] : Zi[o + i] * W[o + i] * X[o + i];
                continue;
            }
            if (half) { gemm_nn(n, W + o, Dm + o, w1); gemm_nn(n, Dm + o, w1, dX + o); }
            else { gemm_nn(n, Zi + o, W + o, w1); gemm_nn(n, w1, X + o, dX + o); }
            symmetrize(n, dX + o);
        }
        double a = 1;
        int ok = 0;
        for (int h = 0; h < (half ? 8 : 6) && !ok; h++) {
            ok = 1;
            for (int k = 0; k < P.nb && ok; k++) {
                const OBlk *B = &P.B[k];
                const int n = B->n;
                const size_t o = B->off, len = B->lp ? (size_t)n : (size_t)n * n;
                for (size_t e = 0; e < len; e++) tmp[o + e] = half ? a * W[o + e] : X[o + e] + a * dX[o + e];
                if (half) for (int i = 0; i < n; i++) tmp[o + (B->lp ? (size_t)i : i + (size_t)i * n)] += 1.0 - 0.01;
                if (B->lp) { for (int i = 0; i < n; i++) if (!(tmp[o + i] > 0)) ok = 0; }
                else if (chol(n, tmp + o)) ok = 0;
            }
            if (!ok) a *= 0.5;
        }
        if (!ok) goto out;
        for (size_t q = 0; q < T; q++) X[q] += a * dX[q];
        steps[pass] = a;
    }
    rc = 0;
out:
    for (size_t q = 0; q < T; q++) Xout[q] = X[q] * P.bs;
    free(X); free(Dm); free(Zi); free(W); free(dX); free(tmp); free(w1); free(w2); free(ev); free(r); free(wv); free(ys);
    free(S.M); free(S.L); free(S.D);
    prob_free(&P);
    return rc;
}

/* ---- the homogeneous self-dual embedding (reference hsd_run, solver.c:9803-12211) ----------
 * Variables X, y, Z, tau, kappa of the scaled problem with residuals
 *   rp = tau b - A(X),  Rd = tau C - A'(y) - Z,  rg = kappa - b'y + <C, X>,
 *   mu = (<X, Z> + tau kappa) / (nu + 1)                                   (solver.c:10357)
 * and the Newton system, every residual reduced by eta = 1 - sigma:
 *   A(dX) - b dtau = eta rp;  A'(dy) + dZ - C dtau = eta Rd;  b'dy - <C, dX> - dkappa = eta rg;
 *   dX = sigma mu Zi - X - sym(Zi dZ X) (+ corrector);  kappa dtau + tau dkappa = sigma mu - tau kappa
 * which gives dy = u + dtau v with M u = w1 and M v = w2 = b + A(X C Zi) (two right-hand sides of
 * the same Schur complement M), dtau from the gap row, dkappa = (sigma mu - tau kappa - kappa
 * dtau) / tau (solver.c:11127), one step length for all variables, and tau normalised back to 1
 * when it grows past 10 (solver.c:12054-12069). Mehrotra predictor-corrector with the standard
 * loop's sigma and gamma rules; the solution is (X, y, Z) / tau. */
/* one endpoint of the embedding's direction family (solver.c:11356-11375, 11410-11440): with
 * u = M^-1 w1(sigma) given, dtau from the gap row, dkappa from the tau-kappa row, dy = u + dtau v,
 * dZ = eta Rd + dtau C - A'(dy), dX = sigma mu Zi - X - sym(Zi dZ X) - Q (Q the second-order
 * term, NULL for none; cK = -<C, Q> and ctk = -dtau_a dkappa_a its scalar parts) */
typedef struct { double tau, kap, mu, rg, cx, c1, cz, den; } HsdScal;
static void hsd_endpoint(const OProb *P, const HsdScal *h, double sg, const double *u, const double *v, const double *hg,
                         const double *Rd, const double *Zi, const double *X, const double *Q, double cK, double ctk,
                         double *dy, double *dZ, double *dX, double *dtau_o, double *dkap_o, double *F, double *G, double *W) {
    const int m = P->m;
    const size_t T = P->tot;
    const double eta = 1.0 - sg, smu = sg * h->mu, tau = h->tau, kap = h->kap;
    const double c0 = smu * h->cz - h->cx;
    const double num = eta * h->rg - dot(m, P->b, u) + c0 - eta * h->c1 + dot(m, hg, u) + cK + (smu - tau * kap + ctk) / tau;
    const double dtau = num / h->den, dkap = (smu - tau * kap + ctk - kap * dtau) / tau;
    for (int i = 0; i < m; i++) dy[i] = u[i] + dtau * v[i];
    ATy(P, dy, F);
    for (size_t q = 0; q < T; q++) dZ[q] = eta * Rd[q] - F[q] + P->C[q] * dtau;
    prod3(P, Zi, dZ, X, G, W);
    for (int k = 0; k < P->nb; k++) {
        const OBlk *B = &P->B[k];
        const size_t len = B->lp ? (size_t)B->n : (size_t)B->n * B->n;
        double *dXk = dX + B->off;
        const double *Gk = G + B->off, *Zik = Zi + B->off, *Xk = X + B->off;
        for (size_t q = 0; q < len; q++) dXk[q] = -Gk[q];
        if (!B->lp) symmetrize(B->n, dXk);
        for (size_t q = 0; q < len; q++) dXk[q] += smu * Zik[q] - Xk[q] - (Q ? Q[B->off + q] : 0.0);
    }
    *dtau_o = dtau; *dkap_o = dkap;
}
/* the step to the boundary of the cone and of tau, kappa >= 0 along (dX, dZ, dtau, dkappa) */
static double hsd_amax(const OProb *P, const double *X, const double *LX, const double *dX, const double *Z, const double *LZ,
                       const double *dZ, double tau, double dtau, double kap, double dkap, double *work) {
    double am = fmin(maxstep(P, X, LX, dX, work), maxstep(P, Z, LZ, dZ, work));
    if (dtau < 0) am = fmin(am, -tau / dtau);
    if (dkap < 0) am = fmin(am, -kap / dkap);
    return am;
}
/* lambda_min((X + a dX)(Z + a dZ)) >= thr on every block: the neighbourhood test of the long step
 * (reference sdp_nb_ok, solver.c:6470-6504): L = chol(X + a dX), then L'(Z + a dZ) L - thr I must
 * factor; LP variables: the products themselves */
static int hsd_nb_ok(const OProb *P, const double *X, const double *dX, const double *Z, const double *dZ, double a, double thr,
                     double *L, double *T, double *W) {
    for (int k = 0; k < P->nb; k++) {
        const OBlk *B = &P->B[k];
        const int n = B->n;
        const double *Xk = X + B->off, *dXk = dX + B->off, *Zk = Z + B->off, *dZk = dZ + B->off;
        if (B->lp) { for (int i = 0; i < n; i++) if (!((Xk[i] + a * dXk[i]) * (Zk[i] + a * dZk[i]) >= thr)) return 0; continue; }
        const size_t n2 = (size_t)n * n;
        for (size_t q = 0; q < n2; q++) { L[q] = Xk[q] + a * dXk[q]; T[q] = Zk[q] + a * dZk[q]; }
        if (chol(n, L)) return 0;
        gemm_nn(n, T, L, W);                                     /* W = (Z + a dZ) L */
        for (int j = 0; j < n; j++)                              /* T = L' W - thr I */
            for (int i = 0; i < n; i++) {
                double s = 0;
                for (int r = i; r < n; r++) s += L[r + (size_t)i * n] * W[r + (size_t)j * n];
                T[i + (size_t)j * n] = s - (i == j ? thr : 0.0);
            }
        symmetrize(n, T);
        if (chol(n, T)) return 0;
    }
    return 1;
}
/* sum of <X + a dX, Z + a dZ> over the blocks */
static double hsd_xz_at(size_t T, const double *X, const double *dX, const double *Z, const double *dZ, double a) {
    double s = 0;
    for (size_t q = 0; q < T; q++) s += (X[q] + a * dX[q]) * (Z[q] + a * dZ[q]);
    return s;
}

static int oracle_solve_hsd(const OracleData *Dd, const OracleOptions *opt, OracleResult *res) {
    memset(res, 0, sizeof(*res));
    res->status = -1;
    OProb P;
    int rc = prob_build(Dd, &P);
    if (rc) return rc;
    const int m = P.m;
    const size_t T = P.tot;
    const double tol = opt && opt->tol > 0 ? opt->tol : 1e-8;
    const int maxit = opt && opt->maxit > 0 ? opt->maxit : 100;
    int maxn = 1;
    for (int k = 0; k < P.nb; k++) if (P.B[k].n > maxn) maxn = P.B[k].n;
    double *X = calloc(T + 1, sizeof(double)), *Z = calloc(T + 1, sizeof(double)), *Zi = calloc(T + 1, sizeof(double));
    double *LX = calloc(T + 1, sizeof(double)), *LZ = calloc(T + 1, sizeof(double)), *dX = calloc(T + 1, sizeof(double));
    double *dZ = calloc(T + 1, sizeof(double)), *F = calloc(T + 1, sizeof(double)), *G = calloc(T + 1, sizeof(double));
    double *Q = calloc(T + 1, sizeof(double)), *Rd = calloc(T + 1, sizeof(double)), *Xb = calloc(T + 1, sizeof(double));
    double *CX = calloc(T + 1, sizeof(double)), *dXa = calloc(T + 1, sizeof(double)), *dZa = calloc(T + 1, sizeof(double));
    double *W = malloc(sizeof(double) * (size_t)maxn * maxn + 8), *work = malloc(sizeof(double) * (size_t)maxn * maxn + 8);
    double *y = calloc(m + 1, sizeof(double)), *dy = calloc(m + 1, sizeof(double)), *rp = calloc(m + 1, sizeof(double));
    double *w1 = calloc(m + 1, sizeof(double)), *w2 = calloc(m + 1, sizeof(double)), *u = calloc(m + 1, sizeof(double));
    double *v = calloc(m + 1, sizeof(double)), *gv = calloc(m + 1, sizeof(double)), *a0 = calloc(m + 1, sizeof(double));
    double *t1 = calloc(m + 1, sizeof(double)), *yb = calloc(m + 1, sizeof(double)), *aq = calloc(m + 1, sizeof(double));
    double *ev0 = calloc(m + 1, sizeof(double)), *ev1 = calloc(m + 1, sizeof(double));
    double *dX0 = calloc(T + 1, sizeof(double)), *dX1 = calloc(T + 1, sizeof(double)), *dZ0 = calloc(T + 1, sizeof(double)), *dZ1 = calloc(T + 1, sizeof(double));
    double *hg = calloc(m + 1, sizeof(double)), *u1 = calloc(m + 1, sizeof(double)), *ev0s = calloc(m + 1, sizeof(double)), *ev1s = calloc(m + 1, sizeof(double));
    double *Qp = calloc(T + 1, sizeof(double)), *dX0s = calloc(T + 1, sizeof(double)), *dX1s = calloc(T + 1, sizeof(double)), *dZ0s = calloc(T + 1, sizeof(double)), *dZ1s = calloc(T + 1, sizeof(double));
    double *nbL = malloc(sizeof(double) * (size_t)maxn * maxn + 8);
    int hoc_nfail = 0;
    OSchur S = { m, calloc((size_t)m * m + 1, sizeof(double)), calloc((size_t)m * m + 1, sizeof(double)), calloc(m + 1, sizeof(double)), 0, 0, 0 };
    int status = 3, it = 0, it_best = 0, nfail = 0, small = 0;
    double best_score = 1e300;
    /* start: the standard loop's point (least-norm X0, least-squares y0, shifts into the cone),
     * tau = 1 and kappa = <X0, Z0> / nu so that tau kappa starts at mu */
    add_identity(&P, X, 1.0);
    memcpy(Zi, X, sizeof(double) * T);
    Aop(&P, P.C, w1, 1.0);
    schur(&P, X, Zi, S.M);
    if (sc_factor(&S, &nfail)) { status = 4; goto done; }
    sc_solve(&S, P.b, dy, 1e-8);
    sc_solve(&S, w1, y, 1e-8);
    memset(X, 0, sizeof(double) * T);
    ATy(&P, dy, X);
    ATy(&P, y, F);
    for (size_t q = 0; q < T; q++) Z[q] = P.C[q] - F[q];
    {
        const double ax = -lammin_all(&P, X, work), az = -lammin_all(&P, Z, work);
        add_identity(&P, X, ax >= -1e-8 ? 1.0 + ax : 0.0);
        add_identity(&P, Z, fmax(0.0, 1.0 + az));
    }
    for (int k = 0; k < P.nb; k++) {
        const OBlk *B = &P.B[k];
        const size_t len = B->lp ? (size_t)B->n : (size_t)B->n * B->n;
        if (B->lp) { memcpy(LX + B->off, X + B->off, sizeof(double) * len); memcpy(LZ + B->off, Z + B->off, sizeof(double) * len); continue; }
        for (double extra = 1e-8; ; extra *= 10) {
            memcpy(LX + B->off, X + B->off, sizeof(double) * len);
            if (chol(B->n, LX + B->off) == 0) break;
            for (int i = 0; i < B->n; i++) X[B->off + i + (size_t)i * B->n] += extra;
        }
        for (double extra = 1e-8; ; extra *= 10) {
            memcpy(LZ + B->off, Z + B->off, sizeof(double) * len);
            if (chol(B->n, LZ + B->off) == 0) break;
            for (int i = 0; i < B->n; i++) Z[B->off + i + (size_t)i * B->n] += extra;
        }
    }
    double tau = 1.0, kap = fmax(dot(T, X, Z) / P.Ntot, 1e-12);
    const double nu1 = P.Ntot + 1.0;
    int anysdp = 0;
    for (int k = 0; k < P.nb; k++) if (!P.B[k].lp) anysdp = 1;
    memcpy(yb, y, sizeof(double) * m); memcpy(Xb, X, sizeof(double) * T);

    for (it = 0; it <= maxit; it++) {
        for (int k = 0; k < P.nb; k++) {
            const OBlk *B = &P.B[k];
            if (B->lp) { for (int i = 0; i < B->n; i++) Zi[B->off + i] = 1.0 / Z[B->off + i]; }
            else inv_from_chol(B->n, LZ + B->off, Zi + B->off);
        }
        /* residuals */
        for (int i = 0; i < m; i++) rp[i] = tau * P.b[i];
        Aop(&P, X, rp, -1.0);
        ATy(&P, y, F);
        double xz = 0, cx = 0, rdn2 = 0;
        for (size_t q = 0; q < T; q++) {
            Rd[q] = tau * P.C[q] - F[q] - Z[q];
            xz += X[q] * Z[q]; cx += P.C[q] * X[q]; rdn2 += Rd[q] * Rd[q];
        }
        const double by = dot(m, P.b, y);
        const double rg = kap - by + cx;
        const double mu = (xz + tau * kap) / nu1;
        /* the normalised iterate's measures (as the standard loop's, solver.c:14426-14482) */
        double rpo = 0;
        for (int i = 0; i < m; i++) { const double t = rp[i] / tau * P.du[i]; rpo += t * t; }
        rpo = sqrt(rpo);
        const double sc = P.bs * P.cs, po = sc * cx / tau, dob = sc * by / tau;
        const double pinf = rpo / (1 + P.normb2), dinf = P.cs * sqrt(rdn2) / tau / (1 + P.normC2);
        const double oden = 1 + fabs(po) + fabs(dob);
        const double relgap = fabs(po - dob) / oden, relcomp = sc * xz / (tau * tau) / oden;
        double score = fmax(fmax(relgap, relcomp), fmax(pinf, dinf));
        if (!isfinite(score)) score = INFINITY;
        if (it < ORACLE_TRACE) { res->trace_pobj[it] = po; res->trace_dobj[it] = dob; res->trace_mu[it] = sc * mu; res->trace_score[it] = score; }
        res->iters = it;
        if (!isfinite(po) || !isfinite(dob)) { status = 4; break; }
        if (score < best_score) {
            best_score = score; it_best = it;
            for (int i = 0; i < m; i++) yb[i] = y[i] / tau;
            for (size_t q = 0; q < T; q++) Xb[q] = X[q] / tau;
        }
        if (score < tol) { status = 0; break; }
        if (it == maxit) { status = 3; break; }
        if (small >= 3) { status = 4; break; }
        if (it - it_best >= 5 && best_score < 1e-6) { status = 4; break; }
        if (it - it_best >= 12) { status = 4; break; }
        /* the Schur complement at (X, Zi); v = M^-1 (b + A(X C Zi)) serves both directions */
        schur(&P, X, Zi, S.M);
        if (sc_factor(&S, &nfail)) { status = 4; break; }
        prod3(&P, Zi, P.C, X, CX, W);                      /* Zi C X */
        memset(gv, 0, sizeof(double) * m);
        Aop(&P, CX, gv, 1.0);                              /* A(X C Zi) */
        for (int i = 0; i < m; i++) w2[i] = P.b[i] + gv[i];
        sc_solve(&S, w2, v, 0);
        const double c2 = dot(T, P.C, CX);                 /* <C, X C Zi> */
        prod3(&P, Zi, Rd, X, G, W);                        /* Zi Rd X */
        const double c1 = dot(T, P.C, G);
        memset(t1, 0, sizeof(double) * m);
        Aop(&P, G, t1, 1.0);                               /* A(X Rd Zi) */
        memset(a0, 0, sizeof(double) * m);
        Aop(&P, Zi, a0, 1.0);
        const double cz = dot(T, P.C, Zi);
        /* The direction family of the embedding (solver.c:11240-11480). With w1(sigma) = [tau b +
         * A(X Rd Zi) + A(Q)] + sigma [-rp - mu A(Zi) - A(X Rd Zi)], M u = w1 has the solution u0 +
         * sigma u1 with u1 independent of the second-order term Q, and every part of the direction is
         * affine in sigma: two endpoints (sigma 0 and 1) give it for any sigma.
         * 1. The affine direction (sigma 0, no Q): its step aa and the Mehrotra sigma_m.
         * 2. Second-order passes: Q = sym(Zi dZ dX) from the current direction (the affine one
         *    first, then the family's member at the sigma the first search chose), u0 solved again,
         *    the endpoints formed and sigma searched among sigma_m, 0.2 sigma_m and 3 sigma_m by the
         *    largest predicted reduction alpha (1 - sigma) of mu (solver.c:11195-11228). A pass
         *    stands while Q contracts (|Q_k - Q_k-1| <= 0.9 |Q_k-1 - Q_k-2|) and its predicted
         *    reduction is at least the first family's, else the first family stands; at most four
         *    passes, none within a factor 100 of the tolerance and none for the rest of the solve
         *    after three iterations whose passes fell behind the first family (solver.c:11236-11247). */
        HsdScal hs = { tau, kap, mu, rg, cx, c1, cz, 0 };
        memset(hg, 0, sizeof(double) * m);
        Aop(&P, CX, hg, 1.0);                              /* A(X C Zi) */
        hs.den = dot(m, P.b, v) + c2 - dot(m, hg, v) + kap / tau;
        for (int i = 0; i < m; i++) w1[i] = P.b[i] * tau + t1[i];
        sc_solve(&S, w1, u, 0);
        double dta = 0, dka = 0;
        hsd_endpoint(&P, &hs, 0.0, u, v, hg, Rd, Zi, X, NULL, 0, 0, dy, dZ, dX, &dta, &dka, F, G, W);
        const double aa = fmin(1.0, hsd_amax(&P, X, LX, dX, Z, LZ, dZ, tau, dta, kap, dka, work));
        double sigma_m;
        {
            const double xdz = dot(T, X, dZ), dxz = dot(T, dX, Z), dxdz = dot(T, dX, dZ);
            const double mua = (xz + aa * (dxz + xdz) + aa * aa * dxdz + (tau + aa * dta) * (kap + aa * dka)) / nu1;
            const double expon = relgap > 1e-6 ? fmax(1.0, 3.0 * aa * aa) : 3.0;
            sigma_m = fmin(1.0, pow(fmax(mua / mu, 0.0), expon));
        }
        for (int i = 0; i < m; i++) w1[i] = -rp[i] - mu * a0[i] - t1[i];
        sc_solve(&S, w1, u1, 0);
        const int late = fmax(pinf, fmax(relgap, dinf)) <= 100.0 * tol;
        const int nhoc = (hoc_nfail >= 3 || late) ? 0 : 4;
        double dtc = dta, dkc = dka, qd_prev = 0, tk2p = 0;
        double m0 = -1, best_m = -1, best_s = sigma_m, best_a = 0, best_am = 0, hsig = sigma_m;
        double dtv[2] = { 0, 0 }, dkv[2] = { 0, 0 }, dtv0[2] = { 0, 0 }, dkv0[2] = { 0, 0 };
        int kept = 0;
        for (int hoc = 0; ; hoc++) {
            prod3(&P, Zi, dZ, dX, Q, W);                   /* Q = sym(Zi dZ dX) at the current direction */
            sym_all(&P, Q);
            const double tk2 = dtc * dkc;
            if (nhoc > 0) {
                double d = 0;
                if (hoc == 0) d = dot(T, Q, Q) + tk2 * tk2;
                else { for (size_t q = 0; q < T; q++) d += (Q[q] - Qp[q]) * (Q[q] - Qp[q]); d += (tk2 - tk2p) * (tk2 - tk2p); }
                d = sqrt(d);
                if (hoc > 0 && !(d <= 0.9 * qd_prev)) break;   /* the last pass that stood keeps its family */
                qd_prev = d; tk2p = tk2;
                memcpy(Qp, Q, sizeof(double) * T);
            }
            memset(aq, 0, sizeof(double) * m);
            Aop(&P, Q, aq, 1.0);
            const double cK = -dot(T, P.C, Q), ctk = -tk2;
            for (int i = 0; i < m; i++) w1[i] = P.b[i] * tau + t1[i] + aq[i];
            sc_solve(&S, w1, u, 0);
            for (int i = 0; i < m; i++) w2[i] = u[i] + u1[i];
            hsd_endpoint(&P, &hs, 0.0, u, v, hg, Rd, Zi, X, Q, cK, ctk, ev0, dZ0, dX0, &dtv[0], &dkv[0], F, G, W);
            hsd_endpoint(&P, &hs, 1.0, w2, v, hg, Rd, Zi, X, Q, cK, ctk, ev1, dZ1, dX1, &dtv[1], &dkv[1], F, G, W);
            double mer = -1, sgb = sigma_m, alb = 0, amb = 0;
            {
                const double cand[3] = { sigma_m, fmin(0.95, fmax(1e-4, 0.2 * sigma_m)), fmin(0.95, fmax(1e-3, 3.0 * sigma_m)) };
                for (int c = 0; c < 3; c++) {
                    const double sg = cand[c];
                    for (size_t q = 0; q < T; q++) { dX[q] = dX0[q] + sg * (dX1[q] - dX0[q]); dZ[q] = dZ0[q] + sg * (dZ1[q] - dZ0[q]); }
                    const double dtau = dtv[0] + sg * (dtv[1] - dtv[0]), dkap = dkv[0] + sg * (dkv[1] - dkv[0]);
                    const double am = hsd_amax(&P, X, LX, dX, Z, LZ, dZ, tau, dtau, kap, dkap, work);
                    const double gam = 0.9 + (0.99 - 0.9) * fmin(1.0, am);
                    const double al = fmin(1.0, gam * am), merit = al * (1.0 - sg);
                    if (merit > mer) { mer = merit; sgb = sg; alb = al; amb = am; }
                }
            }
            if (nhoc == 0) { best_m = mer; best_s = sgb; best_a = alb; best_am = amb; break; }
            if (hoc == 0) {
                m0 = mer; hsig = sgb;
                memcpy(ev0s, ev0, sizeof(double) * m); memcpy(ev1s, ev1, sizeof(double) * m);
                memcpy(dX0s, dX0, sizeof(double) * T); memcpy(dX1s, dX1, sizeof(double) * T);
                memcpy(dZ0s, dZ0, sizeof(double) * T); memcpy(dZ1s, dZ1, sizeof(double) * T);
                dtv0[0] = dtv[0]; dtv0[1] = dtv[1]; dkv0[0] = dkv[0]; dkv0[1] = dkv[1];
            }
            if (hoc == 0 || mer >= m0) { best_m = mer; best_s = sgb; best_a = alb; best_am = amb; kept = hoc; }
            else {
                /* behind the first family: back to it */
                memcpy(ev0, ev0s, sizeof(double) * m); memcpy(ev1, ev1s, sizeof(double) * m);
                memcpy(dX0, dX0s, sizeof(double) * T); memcpy(dX1, dX1s, sizeof(double) * T);
                memcpy(dZ0, dZ0s, sizeof(double) * T); memcpy(dZ1, dZ1s, sizeof(double) * T);
                dtv[0] = dtv0[0]; dtv[1] = dtv0[1]; dkv[0] = dkv0[0]; dkv[1] = dkv0[1];
                kept = 0; hoc_nfail++;
                {   /* the first family's search result */
                    const double cand[3] = { sigma_m, fmin(0.95, fmax(1e-4, 0.2 * sigma_m)), fmin(0.95, fmax(1e-3, 3.0 * sigma_m)) };
                    best_m = -1;
                    for (int c = 0; c < 3; c++) {
                        const double sg = cand[c];
                        for (size_t q = 0; q < T; q++) { dX[q] = dX0[q] + sg * (dX1[q] - dX0[q]); dZ[q] = dZ0[q] + sg * (dZ1[q] - dZ0[q]); }
                        const double dtau = dtv[0] + sg * (dtv[1] - dtv[0]), dkap = dkv[0] + sg * (dkv[1] - dkv[0]);
                        const double am = hsd_amax(&P, X, LX, dX, Z, LZ, dZ, tau, dtau, kap, dkap, work);
                        const double gam = 0.9 + (0.99 - 0.9) * fmin(1.0, am);
                        const double al = fmin(1.0, gam * am), merit = al * (1.0 - sg);
                        if (merit > best_m) { best_m = merit; best_s = sg; best_a = al; best_am = am; }
                    }
                }
                break;
            }
            if (hoc >= nhoc) break;
            for (size_t q = 0; q < T; q++) { dX[q] = dX0[q] + hsig * (dX1[q] - dX0[q]); dZ[q] = dZ0[q] + hsig * (dZ1[q] - dZ0[q]); }
            dtc = dtv[0] + hsig * (dtv[1] - dtv[0]); dkc = dkv[0] + hsig * (dkv[1] - dkv[0]);
        }
        (void)best_m;
        {
            const double sg = best_s;
            for (size_t q = 0; q < T; q++) { dX[q] = dX0[q] + sg * (dX1[q] - dX0[q]); dZ[q] = dZ0[q] + sg * (dZ1[q] - dZ0[q]); }
            for (int i = 0; i < m; i++) dy[i] = ev0[i] + sg * (ev1[i] - ev0[i]);
            double dtau = dtv[0] + sg * (dtv[1] - dtv[0]), dkap = dkv[0] + sg * (dkv[1] - dkv[0]);
            double a = best_a;
            /* refinement (solver.c:11599-11695): two passes enforcing A(dX) - b dtau = eta rp, each the
             * embedding's Newton step for the miss e alone (M w = e, dtau_e = -(b - hg)'w / den, dy_e =
             * w + dtau_e v, dZ_e = dtau_e C - A'(dy_e), dX_e = -sym(Zi dZ_e X)), taken when it at
             * least halves the miss; none when the miss is below 1e-2 of eta |rp| */
            for (int pass = 0; pass < 2; pass++) {
                const double eta = 1.0 - sg;
                for (int i = 0; i < m; i++) w1[i] = eta * rp[i] + P.b[i] * dtau;
                Aop(&P, dX, w1, -1.0);
                const double ne = sqrt(dot(m, w1, w1)), nrp = sqrt(dot(m, rp, rp));
                if (!(ne > 1e-2 * eta * nrp && ne > 0)) break;
                sc_solve(&S, w1, u, 0);
                const double dte = -(dot(m, P.b, u) - dot(m, hg, u)) / hs.den;
                for (int i = 0; i < m; i++) u[i] += dte * v[i];
                ATy(&P, u, F);
                for (size_t q = 0; q < T; q++) F[q] = P.C[q] * dte - F[q];
                prod3(&P, Zi, F, X, G, W);
                sym_all(&P, G);
                for (size_t q = 0; q < T; q++) Qp[q] = dX[q] - G[q];
                for (int i = 0; i < m; i++) w2[i] = eta * rp[i] + P.b[i] * (dtau + dte);
                Aop(&P, Qp, w2, -1.0);
                if (!(sqrt(dot(m, w2, w2)) < 0.5 * ne)) break;
                memcpy(dX, Qp, sizeof(double) * T);
                for (size_t q = 0; q < T; q++) dZ[q] += F[q];
                for (int i = 0; i < m; i++) dy[i] += u[i];
                dtau += dte; dkap -= kap * dte / tau;
            }
            /* the long step after passes that stood (solver.c:11818-11855): 0.992, then 0.9995 and
             * 0.998 of the way to the boundary when the new point stays in the neighbourhood
             * lambda_min(XZ) >= 0.3 mu_new (and tau kappa >= 0.3 mu_new) */
            if (kept > 0 && best_am >= 0.9) {
                const double gs[3] = { 0.9995, 0.998, 0.992 };
                int short_ok = 1, short_tested = 0;
                {
                    const double al = fmin(1.0, gs[2] * best_am);
                    if (al > a && tau + al * dtau > 0 && kap + al * dkap > 0) {
                        const double mun = mu * (1.0 - al * (1.0 - sg));
                        short_ok = (tau + al * dtau) * (kap + al * dkap) >= 0.3 * mun && hsd_nb_ok(&P, X, dX, Z, dZ, al, 0.3 * mun, nbL, work, W);
                        short_tested = 1;
                    }
                }
                for (int c = 0; c < 3 && short_ok; c++) {
                    const double al = fmin(1.0, gs[c] * best_am);
                    if (short_tested && c == 2) { a = al; break; }
                    if (al <= a) break;
                    if (tau + al * dtau <= 0 || kap + al * dkap <= 0) continue;
                    const double mun = mu * (1.0 - al * (1.0 - sg));
                    if ((tau + al * dtau) * (kap + al * dkap) >= 0.3 * mun && hsd_nb_ok(&P, X, dX, Z, dZ, al, 0.3 * mun, nbL, work, W)) { a = al; break; }
                }
            }
            /* neighbourhood safeguard (solver.c:11856-11882): tau kappa and the LP products stay at
             * least 1e-3 of the new mu (x0.7, at most ten times) */
            for (int t2 = 0; t2 < 10; t2++) {
                const double tk = (tau + a * dtau) * (kap + a * dkap);
                const double munew = (hsd_xz_at(T, X, dX, Z, dZ, a) + tk) / nu1;
                double lo = tk;
                for (int k = 0; k < P.nb; k++) {
                    const OBlk *B = &P.B[k];
                    if (!B->lp) continue;
                    for (int i = 0; i < B->n; i++) {
                        const size_t o = B->off + i;
                        const double w = (X[o] + a * dX[o]) * (Z[o] + a * dZ[o]);
                        if (w < lo) lo = w;
                    }
                }
                if (!(munew > 0) || lo >= 1e-3 * munew || a < 1e-3) break;
                a *= 0.7;
            }
            /* X and Z positive definite at the common step (x0.8 on failure, solver.c:6393-6446) */
            double ax = a, az = a;
            if (!try_step(&P, X, LX, dX, &ax)) { status = 4; goto fin; }
            if (ax < a) a = ax;
            if (!try_step(&P, Z, LZ, dZ, &az)) { status = 4; goto fin; }
            if (az < a) {
                const double back = az - a;
                for (size_t q = 0; q < T; q++) X[q] += back * dX[q];
                double one = 1.0;
                memset(Q, 0, sizeof(double) * T);
                if (!try_step(&P, X, LX, Q, &one)) { status = 4; goto fin; }
                a = az;
            } else if (a < az) {
                const double back = a - az;
                for (size_t q = 0; q < T; q++) Z[q] += back * dZ[q];
                double one = 1.0;
                memset(Q, 0, sizeof(double) * T);
                if (!try_step(&P, Z, LZ, Q, &one)) { status = 4; goto fin; }
            }
            for (int i = 0; i < m; i++) y[i] += a * dy[i];
            tau += a * dtau;
            kap += a * dkap;
            res->trace_ap[it] = a; res->trace_ad[it] = sg;
            small = a < 1e-6 ? small + 1 : 0;
            if (!(tau > 0) || !(kap > 0)) { status = 4; goto fin; }
        }
        /* normalise tau to 1 when it has grown (solver.c:12054-12069) */
        if (tau > 10.0) {
            const double f = 1.0 / tau, sq = sqrt(f);
            for (size_t q = 0; q < T; q++) { X[q] *= f; Z[q] *= f; }
            for (int k = 0; k < P.nb; k++) {
                const OBlk *B = &P.B[k];
                const size_t len = B->lp ? (size_t)B->n : (size_t)B->n * B->n;
                for (size_t q = 0; q < len; q++) { LX[B->off + q] *= B->lp ? f : sq; LZ[B->off + q] *= B->lp ? f : sq; }
            }
            for (int i = 0; i < m; i++) y[i] *= f;
            kap *= f;
            tau = 1.0;
        }
    }
fin:
    if (status == 0) {
        for (int i = 0; i < m; i++) y[i] /= tau;
        for (size_t q = 0; q < T; q++) X[q] /= tau;
    } else { memcpy(y, yb, sizeof(double) * m); memcpy(X, Xb, sizeof(double) * T); }
done:
    res->m = m; res->nblk = P.nb;
    res->y = malloc(sizeof(double) * (m + 1));
    res->X = malloc(sizeof(double) * (T + 1));
    for (int i = 0; i < m; i++) res->y[i] = P.cs * P.d[i] * y[i];
    for (size_t q = 0; q < T; q++) res->X[q] = P.bs * X[q];
    res->xlen = T;
    {
        double pobj, dobj;
        dimacs(&P, res->X, res->y, res->err, &pobj, &dobj, work);
        res->pobj = pobj; res->dobj = dobj;
        double acc = fmax(fmax(res->err[0], res->err[1]), fmax(res->err[3], fmax(fabs(res->err[4]), fabs(res->err[5]))));
        if (!isfinite(acc)) acc = INFINITY;
        if (status == 0 || status == 3 || status == 4 || status == 5) {
            if (acc <= tol) status = 0;
            else if (acc <= 1e-6) status = 5;
            else if (status == 0) status = 4;
        }
        res->acc = acc;
    }
    res->status = status;
    free(X); free(Z); free(Zi); free(LX); free(LZ); free(dX); free(dZ); free(F); free(G); free(Q); free(Rd); free(Xb);
    free(CX); free(dXa); free(dZa); free(W); free(work); free(y); free(dy); free(rp); free(w1); free(w2); free(u); free(v);
    free(gv); free(a0); free(t1); free(yb); free(aq); free(ev0); free(ev1); free(dX0); free(dX1); free(dZ0); free(dZ1);
    free(hg); free(u1); free(ev0s); free(ev1s); free(Qp); free(dX0s); free(dX1s); free(dZ0s); free(dZ1s); free(nbL);
    free(S.M); free(S.L); free(S.D);
    prob_free(&P);
    return 0;
}
