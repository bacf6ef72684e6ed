/* solver.h - host-side structures of the GPU interior-point solver (C).
 *
 * A solve owns a problem (host copy of the scaled data, constraint entries and position lists)
 * and device arrays carved from the one device heap. The iteration runs on a batch of solves at
 * once: every kernel launch covers all solves of the batch, and each solve keeps its own scalars
 * (step lengths, sigma, theta, convergence state) on the host. */
#ifndef BRISK_SOLVER_H
#define BRISK_SOLVER_H
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include "brisk_dev.h"

#define BK_LZ_K 30          /* Lanczos steps (reference Params.lanczos_k, solver.c:119) */

/* max and min in which a NaN operand wins as +infinity (max) or -infinity (min): a NaN error or
 * score never passes a test (reference brisk_nmax / brisk_nmin, brisk.h:182-184) */
static inline double nmax(double a, double b) { return (a != a || b != b) ? INFINITY : (b > a ? b : a); }
static inline double nmin(double a, double b) { return (a != a || b != b) ? -INFINITY : (a < b ? a : b); }
/* the error a result is accepted by: primal equations, X's cone, Z's cone, gap, complementarity
 * (err[0], err[1], err[3], |err[4]|, |err[5]|, 0-based; reference brisk_acc_err, brisk.h:187-189) */
static inline double acc_err(const double *e) {
    return nmax(nmax(e[0], e[1]), nmax(e[3], nmax(fabs(e[4]), fabs(e[5]))));
}
/* the status of a returned point from its measured DIMACS errors (reference main.c:2349-2359):
 * a run that ended optimal, at a limit or in difficulties is OPTIMAL when acc <= tol, REDUCED
 * ACCURACY (5) when acc <= 1e-6, NUMERICAL DIFFICULTIES (4) when it had claimed optimal or
 * reduced accuracy; infeasibility verdicts (1, 2) stand */
static inline int status_from_errors(int st, const double *err, double tol) {
    const double acc = acc_err(err);
    if (st == 0 || st == 3 || st == 4 || st == 5 || st == 6) {
        if (acc <= tol) return 0;
        if (acc <= 1e-6) return 5;
        if (st == 0 || st == 5) return 4;
    }
    return st;
}
/* the termination score of an iterate: the largest of relative gap, complementarity and the
 * two infeasibilities (reference res_score, solver.c:8089-8103) */
static inline double iterate_score(double relgap, double relcomp, double pinf, double dinf) {
    return nmax(nmax(relgap, relcomp), nmax(pinf, dinf));
}
/* fault injection for the tests of failure paths (the reference's brisk_fault, brisk.h:169-173):
 * BRISK_FAULT=<names> makes each named check fail ("setup": a solve's device setup) */
int brisk_fault(const char *name);

typedef struct {
    /* the problem as read */
    int m, nblk_in;
    int *bs_in;
    int *in_first;          /* internal block of input block k (LP: its first variable) */
    /* internal blocks: every SDP block, and every LP variable as a 1 x 1 block */
    int nb;
    int *bn;
    int64_t *boff;          /* element offset of block k in a block arena */
    int64_t T;              /* arena length: sum of n^2 */
    int nmax, any_sdp;
    double Ntot;            /* sum of block orders (the barrier's dimension) */
    /* scaled data (problem.c:1866-1909) */
    double *b, *b0, *d, *du, *C;
    /* C given by its nonzeros instead (C NULL): positions in the block arena, ascending, and values;
     * the device arena is zeroed and these scattered into it (sparse C: max-cut's Laplacian) */
    int64_t nc, *cpos;
    double *cval;
    double normb1, normb2, normC1, normC2, bs, cs, nbs;
    /* constraint entries, upper (r <= c), CSR by constraint, grouped by block */
    int *cptr;
    SEnt *ent;
    int nent;
    /* groups: the runs of one constraint's entries in one block (gptr: m+1); heavy groups keep
     * their product X A Zi (gp_tot doubles in all, poff within that arena) */
    int ngrp;
    int *gptr;
    GrpRef *grp;
    int64_t gp_tot;
    /* positions touched by constraints, each with its (constraint, value) list */
    int npos;
    int *pptr, *ppos, *pcon;
    double *pval;
    int diag;               /* every constraint is one diagonal entry: Hadamard Schur form */
    int *bdiag;             /* per block: every constraint entry in it is diagonal (NULL: none is known to be) */
    int *pairs;             /* the pairs (a, b), a >= b, of constraints sharing a block, by column (npair of them), or NULL */
    int64_t npair;
    /* dense route: one SDP block with dense constraints; Ad (host) holds the A_i as read in full
     * symmetric storage, column i = vec(A_i) (n*n x m), scaled by d_i on the device after the
     * upload; the entry and position lists are empty */
    int dense;
    double *Ad;
    size_t ad_cap;          /* doubles allocated at Ad */
    int want_hsd;           /* the solve will run the self-dual embedding (its arrays are counted) */
    int want_nt;            /* the solve will take the Nesterov-Todd direction (its arrays are counted) */
} HProb;

enum { PREC_F64 = 0, PREC_MIXED = 1 };

typedef struct {
    double tol;             /* 1e-8 */
    int maxit;              /* 100 */
    int prec;               /* PREC_F64 or PREC_MIXED */
    int verbose;
    double mixed_until;     /* mixed: f32 factor while max(relgap, pinf, dinf) exceeds this (1e-6) */
    double time_limit;      /* s, <= 0 none */
    int hsd;                /* 1: self-dual embedding, 0: infeasible start, -1: chosen per problem (hsd_first_rule) */
    int hsd_first;          /* automatic method: the embedding first where its extra dense work is small (1), the infeasible start first (0) */
    double hsd_first_c, hsd_first_asm;   /* the rule's constants: dense work c sum n^3 against m^3/3, or against asm (m^3/3 + the assembly) */
    int apply;              /* solves with the Schur factor: -1 a cost model on the solve's own order and precision (default), 0 explicit inverse, 1 block triangular solves */
    int tput;               /* 1: the plan choices for throughput in large batches (256-column panels from order 1024, blocked solves for f64 factors from order 1500); 0: for latency */
    int dir;                /* direction of the infeasible-start method: 0 HKM, 1 Nesterov-Todd, -1 chosen per problem (nt_rule) */
    int polish;             /* 1: the FP64 solves' returned point polished (projected onto A(X) = b) when that improves it (reference Params.polish, -nopolish) */
    int polish_r, polish_x; /* the square-root-metric restoration and the X-metric polish before it (reference Params.polish_r, polish_x: -nopolishr, -nopolishx) */
} SolveOpts;

typedef struct Solve {
    HProb P;
    SolveOpts o;
    int id;
    /* device heap offsets in doubles (block arena arrays have P.T elements) */
    int64_t hbase, hsize;   /* the solve's heap region (bytes) */
    int64_t X, Z, Zi, LX, LZ, LXi, LZi, LXiT, LZiT, dX, dZ, G, Q, R0, F, W1, W2, C, DinvX, DinvZ, Xb;
    int64_t lzQ, lzw, lzt1, lzt2;
    int64_t M, MF, MLi, MLiT, MDinv, Mwork, Dm;
    int64_t y, dy, rp, rhs, a0, h, qv, ev, wv, bvec, duv, pr, pz, pp, pq, ps, pxb, t1, t2, yb;
    int64_t sc;             /* 64 scalars */
    int64_t cptr, ent, blk, pptr, ppos, pcon, pval, dinvoff; /* problem data on the device */
    int64_t gptr, grp, Pg, Tg;      /* constraint groups (bytes), heavy-group products P and T (doubles) */
    int64_t pairs;                  /* the constraint pairs sharing a block (bytes) */
    int64_t Rd, CX, dX0, dZ0, dX1, dZ1, hv, hg, ht, hy0, hy1;   /* self-dual embedding: residual, Zi C X, the direction family (base, difference), vectors */
    int64_t Qp, dX0s, dZ0s, hy0s;   /* self-dual embedding: the last second-order term that stood, the first family's base */
    int64_t SX, SZi, SLX, SLZi;     /* the scaling pair of the Schur complement and the direction's products: (X, Zi) and
                                       their factors for HKM, (W, W) and (G, G') for the Nesterov-Todd direction */
    int nt;                         /* the solve takes the Nesterov-Todd direction */
    int64_t Wn, Gn, GnT, Gi, ntV, ntA, lam;   /* Nesterov-Todd: W = G G', G, G', G^-1, eigenvectors, L_X' Z L_X, the eigenvalue vectors (4 per block order) */
    int64_t rhs2, rhs3, xvec, xom;  /* self-dual embedding: right-hand sides and the PCG vectors (8 m each) and operator scratch (2 T each) of two more systems */
    int hoc_nfail;          /* self-dual embedding: iterations whose second-order passes fell behind the first family */
    int nretry;             /* step-length factorizations that failed (x0.8 retries), all iterations */
    int64_t Ad, Af, Pm, Bm, Mpart, LXf, LZif, Om1, Om2;  /* dense route: A f64 / f32 (n^2 x m), P, B, split-K partials, f32 factors, M v scratch */
    int nsplit, kchunk;     /* dense route: split-K of M = P'P */
    int m_exact;            /* the stored M is exact in f64 (else M v uses the operator form) */
    /* factor state of M */
    int f32;                /* current factor is f32 */
    int trsv;               /* current factor is applied by block triangular solves (no explicit inverse) */
    int want_double;        /* sticky: f32 no longer used */
    double reg, reg_ok;
    int reg_hint;
    /* iteration state */
    int it, status, done, stall, it_best;
    double theta, r0norm, best_score, alp, ald;
    int hsd;                /* 1: the homogeneous self-dual embedding (tau, kappa) */
    double tau, kap;
    double pobj, dobj, pinf, dinf, relgap, relcomp, mu, score;
    double err_it[7];
    double t_start, t_solve;
    int pcg_total, nfactor_f32, nfactor_f64;
    /* results (data as read) */
    double res_pobj, res_dobj, res_err[6];
    double *res_y, *res_X, *res_Z;  /* y (m), X and Z packed per input block */
    /* trace */
    double tr_pobj[128], tr_dobj[128], tr_mu[128];
} Solve;

/* problem.c */
/* the checks of hprob_build on the data alone (0, or 2: bad data), without building anything */
int hprob_check(int m, int nblk, const int *bs, const double *c, int64_t nnz, const int *mat, const int *blk, const int *ii,
                const int *jj, const double *v);
int hprob_build(HProb *P, int m, int nblk, const int *bs, const double *c, int64_t nnz, const int *mat, const int *blk,
                const int *ii, const int *jj, const double *v);
void hprob_free(HProb *P);
int64_t solve_heap_bytes(const HProb *P);   /* device memory a solve needs */
int hsd_first_rule(const HProb *P, const SolveOpts *o);   /* the method of an automatic choice: 1 the embedding, 0 the infeasible start */
int nt_rule(const HProb *P);                /* the direction of an automatic choice: 1 Nesterov-Todd, 0 HKM */

/* heap allocator (host-managed, first fit) */
int heap_setup(size_t bytes);
int64_t heap_alloc(size_t bytes);           /* byte offset, 256-aligned; -1 when full */
void heap_free(int64_t off);
size_t heap_largest_free(void);

/* ipm.c */
int solve_setup(Solve *S);                  /* allocate device arrays and upload the problem */
void solve_release(Solve *S);
int ipm_run_batch(Solve **S, int ns);       /* runs every solve to its end */
/* stepwise interface for the serving loop: start, one iteration for a batch, finish */
int ipm_start_batch(Solve **S, int ns);
int ipm_iterate_batch(Solve **S, int ns);   /* one iteration; sets done on solves that end */
int ipm_iterate_hsd(Solve **S, int ns);     /* the same for the self-dual embedding */
int ipm_dimacs(Solve **S, int ns);          /* DIMACS errors and status of the returned points (device) */
int ipm_end_corrections(Solve **S, int ns, const int *st0);  /* after ipm_dimacs: the square-root-metric restoration and the X-metric polish of points that missed the tolerance */
int ipm_polish(Solve **S, int ns, const int *st0);  /* after ipm_dimacs: the returned points polished where that helps (st0: the runs' statuses before it) */
int ipm_end_batch(Solve **S, int ns);       /* DIMACS on the device, polish, status, results to the host */
int ipm_finish_batch(Solve **S, int ns);   /* results of the solves to the host, one batch of reads */
int ipm_finish(Solve *S);                   /* results of one solve to the host (after ipm_end_batch's DIMACS) */
void solve_log(const char *fmt, ...);
void ipm_release_scratch(void);           /* frees the batch scratch (before the heap is re-sized) */

#endif
