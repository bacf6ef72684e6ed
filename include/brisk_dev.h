/* brisk_dev.h - the thin C ABI between the host solver (C) and the HIP side (rt.hip, k_dense.hip,
 * k_ipm.hip). Plain pointers and sizes only. Every kernel wrapper is asynchronous on the
 * library's one stream; the host reads scalars back with bk_d2h (which synchronises).
 *
 * Matrix conventions: column-major, lower triangle significant for symmetric and triangular
 * matrices. "Items" describe one independent piece of a batched launch; the host builds item
 * arrays in host memory and the wrappers stage them to the device through a pinned ring that is
 * recycled at every synchronisation point. */
#ifndef BRISK_DEV_H
#define BRISK_DEV_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ---- */
int bk_init(void);                               /* idempotent; 0 on success */
const char *bk_error(void);                      /* last HIP error text ("" if none) */
int bk_failed(void);                             /* a HIP call has failed since bk_init */
void *bk_malloc(size_t bytes);                   /* device memory (NULL on failure) */
void bk_free(void *p);
void *bk_host_alloc(size_t bytes);               /* pinned host memory (direct DMA copies) */
void bk_host_free(void *p);
int bk_h2d(void *dst, const void *src, size_t bytes);        /* synchronous copy */
int bk_h2d_async(void *dst, const void *src, size_t bytes);  /* copy ordered on the stream; src may be reused at once */
int bk_d2h(void *dst, const void *src, size_t bytes);        /* synchronous (drains the stream) */
/* n device-to-host reads (dst[i] <- src[i], bytes[i]) with one wait per 32 MB of small reads */
int bk_d2h_many(int n, void *const *dst, const void *const *src, const size_t *bytes);
int bk_d2d(void *dst, const void *src, size_t bytes);        /* async device copy */
int bk_zero(void *dst, size_t bytes);                         /* async fill kernel (graph safe) */
int bk_sync(void);
/* a second stream for independent parts of a batch: bk_fork (it waits for the work queued so
 * far), bk_side(1) / bk_side(0) direct launches to it / back, bk_join (the library stream waits
 * for it). bk_sync waits for both while forked. */
int bk_fork(void);
void bk_side(int on);
int bk_join(void);
void *bk_stage(const void *host, size_t bytes);  /* device copy of host data, valid until next sync */
double bk_now(void);                             /* host wall clock, s */
size_t bk_mem_free(void);                        /* free device memory, bytes */
/* events on the library stream (timing of kernels in bench): create once, record, elapsed ms */
void *bk_event_new(void);
void bk_event_record(void *e);
float bk_event_ms(void *a, void *b);
void bk_event_free(void *e);
const char *bk_device_name(void);
/* per-class kernel timing (events around instrumented launches on the library stream):
 * out4 = { total ms, launches, algorithmic flops, algorithmic bytes } of the class */
void bk_prof_enable(int on);
void bk_prof_reset(void);
int bk_prof_begin(void);
void bk_prof_end(int tok, int cls, double flops, double bytes);
void bk_prof_read(int cls, double *out4);
int bk_prof_on(void);                            /* profiling enabled */
void bk_prof_host(int cls, double seconds);      /* add a host-side interval to class cls (40-59) */
/* classes: 0-3 k_gemm<double, ta, tb> (2 ta + tb), 4-7 k_gemm<float, ta, tb>, 8 k_potrf_diag,
 * 9 k_gemv, 10 Schur assembly, 11 constraint operators, 12 Lanczos steps, 13 dot products,
 * 14 linear combinations, 15 k_symv (tiles and reduction), 16 k_trsv (both directions), 17 k_syev */

/* ---- batched GEMM: C_it = alpha op(A_it) op(B_it) + beta C_it ---- */
typedef struct {
    int M, N, K;             /* C is M x N, op(A) M x K, op(B) K x N */
    int lda, ldb, ldc;
    int64_t a, b, c;         /* element offsets from the base pointers */
    int guard;               /* >= 0: skip this item when info[guard] != 0 */
    int tri;                 /* 1: write C's lower triangle only (tiles above the diagonal skipped);
                                2: op(A) upper triangular, 4: op(A) lower triangular,
                                8: op(B) lower triangular, 16: op(B) upper triangular - the K loop
                                is trimmed to the range where both operands can be nonzero (the
                                other triangle must hold explicit zeros) */
} GemmItem;
int bk_gemm_f64(int ta, int tb, double alpha, double beta, const double *A, const double *B, double *C,
                const GemmItem *items, int nitems, const int *info);
int bk_gemm_f32(int ta, int tb, float alpha, float beta, const float *A, const float *B, float *C,
                const GemmItem *items, int nitems, const int *info);

/* ---- matrix sets (square n x n at base + off) ---- */
typedef struct {
    int n, slot;             /* order; index into per-item arrays (info, pmin, per-matrix scalars) */
    int64_t off;             /* the matrix at base + off */
    int64_t off2;            /* potrf/trtri: diagonal-block inverses (64*64 per 64-wide panel) */
    int64_t off3;            /* potrf/trtri: scratch of n*n elements */
    int64_t off4;            /* trtri: the inverse; transpose: the destination */
    int opt, pad;            /* potrf: the panel width, 128 or 256 (0: 256 from order 3685, else 128), plus
                                BK_POTRF_SPLIT for the diagonal block and the panel below it in two launches */
} MItem;
/* potrf: a matrix whose 64-column steps factor the diagonal block in one workgroup and multiply
 * the panel by its inverse in a GEMM (the least work, for batches that fill the GPU), instead of
 * the one launch in which every panel tile's workgroup factors the block itself (the shortest
 * chain, for latency) */
#define BK_POTRF_SPLIT 1024

/* Cholesky of a set of matrices, blocked right-looking (panels of 128 or 256 columns, chosen per
 * matrix by its item's opt; each factored in 64-column steps), in place on A (its
 * lower triangle is read; on return A holds L with the strict upper triangle zeroed).
 * Diagonal-block inverses go to Dinv + off2 (64*64 per 64-column step); W + off3 is scratch
 * (n*64 used).
 * info[slot] = first failing column + 1 (0: factored); pmin[slot] = smallest squared pivot. */
int bk_potrf_f64(const MItem *it, int n_it, double *A, double *Dinv, double *W, int *info, double *pmin);
int bk_potrf_f32(const MItem *it, int n_it, float *A, float *Dinv, float *W, int *info, float *pmin);
/* Linv + off4 = L^{-1} (lower, strict upper zero) from the factor L + off and its diagonal-block
 * inverses (Dinv + off2); W + off3 scratch (n*n). Items with info[slot] != 0 are skipped. */
int bk_trtri_f64(const MItem *it, int n_it, const double *L, const double *Dinv, double *Linv, double *W, const int *info);
int bk_trtri_f32(const MItem *it, int n_it, const float *L, const float *Dinv, float *Linv, float *W, const int *info);
/* the same, limited to the 256 x 256 diagonal blocks: Linv + off4 holds the inverse of every
 * diagonal block of L (rows and columns 256 b .. 256 b + 255) and zeros elsewhere */
int bk_trtri_diag_f64(const MItem *it, int n_it, const double *L, const double *Dinv, double *Linv, double *W, const int *info);
int bk_trtri_diag_f32(const MItem *it, int n_it, const float *L, const float *Dinv, float *Linv, float *W, const int *info);
/* ---- blocked triangular solves with a factor: L v1 = v0, then L' v0 = v1 ----
 * (the preconditioner application z = L^{-T} L^{-1} r of the PCG without an explicit inverse,
 * reference schur_solve_ solver.c:5483-5622 applies its factor by the same two solves).
 * L (order m, leading dimension m) at offset L; Di at offset Di holds the inverses of its 256 x
 * 256 diagonal blocks (bk_trtri_diag); DiT at offset DiT receives their transposes from
 * bk_trsv_prep, once per factor. v0, v1: f64 vectors (offsets in doubles); v0 is the right-hand
 * side and receives the solution, v1 is overwritten. Matrix offsets in the factor's type; pre
 * is set by the library. One launch per 256-wide block column and direction. */
typedef struct { int m, c; int64_t pre, L, Di, DiT, v0, v1; } TrsvItem;
int bk_trsv_prep_f64(const TrsvItem *it, int n_it, double *A);
int bk_trsv_prep_f32(const TrsvItem *it, int n_it, float *A);
int bk_trsv_f64(const TrsvItem *it, int n_it, const double *A, double *V);
int bk_trsv_f32(const TrsvItem *it, int n_it, const float *A, double *V);

/* ---- batched matrix-vector: y_it = alpha op(A_it) x_it + beta y_it ---- */
typedef struct { int M, N, lda; int64_t a, x, y; int tri; } GemvItem;   /* tri (trans = 1 only): 1 A lower (rows >= j), 2 A upper (rows <= j) */
int bk_gemv_f64(int trans, double alpha, double beta, const double *A, const double *x, double *y, const GemvItem *it, int n_it);
/* nstage products y = A' x (alpha 1, beta 0) in sequence from one staging; stage s takes the
 * items it[s n_it .. s n_it + n_it) (offsets into the heap H) */
int bk_gemv_chain_f64(double *H, const GemvItem *it, int n_it, int nstage);
int bk_gemv_f32(int trans, float alpha, float beta, const float *A, const float *x, float *y, const GemvItem *it, int n_it);
/* mixed: y (f64) = alpha op(A) x + beta y with A in f32, x in f64 (products and sums in f64) */
int bk_gemv_mixed(int trans, double alpha, double beta, const float *A, const double *x, double *y, const GemvItem *it, int n_it);
/* symmetric: y = alpha A x + beta y for symmetric n x n A of which only the lower triangle is
 * read (half the traffic of bk_gemv_f64). Each 64 x 64 lower tile (I, J) writes its row part
 * A_IJ x_J and, for I > J, its column part A_IJ' x_I to the partial slots (I, J) and (J, I) of
 * p (T x T slots of 64 doubles, T = ceil(n / 64)); each row block then sums its T slots in a
 * fixed order, so the result depends on the item's own data only. Offsets in doubles. */
typedef struct { int n, ld; int64_t a, x, y, p; } SymvItem;
int bk_symv_f64(double alpha, double beta, const double *A, const double *x, double *y, double *p, const SymvItem *it, int n_it);
/* The same product with no device allocation and no workspace argument. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The partial-sum slots live in a
 * file-scope device array of 32 MiB, one per GPU, held until the process ends; the items' p
 * fields are ignored. At most 16 items per call and sum over the items of ceil(n/64)^2 * 64
 * slots of 4 Mi doubles (a single item of order up to 16384): larger calls return -1 with the
 * reason in bk_error() and launch nothing. Bit-identical to bk_symv_f64 (the same kernels'
 * arithmetic). Calls that overlap on one GPU (other streams, other threads) race on the array. */
int bk_symv_f64_noalloc(double alpha, double beta, const double *A, const double *x, double *y, const SymvItem *it, int n_it);
/* looks up the device address of that array for the current device (bk_init calls it) */
int bk_symv_scratch_init(void);

/* ---- the device heap: one allocation; every array is an offset from its base ---- */
int bk_heap_init(size_t bytes);                  /* (re)allocate; 0 on success */
void *bk_heap(void);                             /* base pointer */
size_t bk_heap_size(void);

/* ---- segmented deterministic reductions ---- */
/* out[s] = sum_q x[xo + q] * y[yo + q], q < len, for each item; chunked partial sums (4096
 * elements per chunk, fixed-order tree in a chunk), the chunks summed by 64 lanes (lane l the
 * chunks l, l + 64, ... in order) and a butterfly: the result depends on the data and len only.
 * All offsets are elements of double from the heap base. */
typedef struct { int64_t len, xo, yo, out; } DotItem;
int bk_dots(const DotItem *it, int n_it);
/* dot results whose output offset lies in [off, off + len) (doubles) are written to host[out - off]
 * (pinned, device-visible memory) instead of the heap: read them after bk_sync, no copy */
void bk_dot_mirror(int64_t off, int64_t len, double *host);

/* out[q] = c0 in0[q] + c1 in1[q] + c2 in2[q] + c3 in3[q] for q < len (an offset < 0 means the
 * term is absent; out may alias any input) */
typedef struct { int64_t len, out, in[4]; double c[4]; } LinItem;
/* out[pos[i]] = val[i] for i < n (device arrays; positions in elements): a sparse array into a
 * dense one that the caller zeroed */
int bk_scatter_f64(const int64_t *pos, const double *val, int64_t n, double *out);
int bk_lincomb(const LinItem *it, int n_it);
/* out[q] = c0 in0[q] in1[q] (+ c1 in2[q] when in[2] >= 0) */
int bk_vmul(const LinItem *it, int n_it);

/* per square matrix at heap offset off (n = order): */
int bk_symmetrize(const MItem *it, int n_it);                       /* G = (G + G')/2; opt 1: upper := lower' */
int bk_transpose_f64(const MItem *it, int n_it);                    /* matrix at off4 = (matrix at off)' */
int bk_transpose_f32(const MItem *it, int n_it);
/* matrix at off3 = Diag(matrix at off) (matrix at off2): each row r scaled by the first matrix's
 * diagonal entry (r, r); offsets in doubles from the heap base */
int bk_rowscale(const MItem *it, int n_it);
/* matrix at off4, its diagonal only: G(r, r) = sum_t Zi(t, r) V(t, t) X(t, r), with Zi at off, V
 * (diagonal) at off2, X at off3, Zi symmetric; the rest of G is not written */
int bk_zvx_diag(const MItem *it, int n_it);
/* A += s[slot] I, with s read from the heap at offset s_off (doubles) */
int bk_add_diag(const MItem *it, int n_it, int64_t s_off);
/* equilibration of M (full f64 at off): D[off2 + i] = 1/sqrt(M_ii) (1 when M_ii <= 0);
 * F (at off3, lower triangle, strict upper zero) = D M D + reg[slot] I in f32 (prec 0) or f64 (1);
 * reg read from the heap at reg_off (doubles, per slot); entries below 1e-30 flushed to 0 */
int bk_equil(const MItem *it, int n_it, int prec, int64_t reg_off);

/* ---- dense route ---- */
/* M (f64 full, at M_off) = sum of nchunk partial products (lower triangles, f32 or f64 by prec,
 * chunk c at part_off + c * stride elements), summed in chunk order */
int bk_splitk_reduce(int prec, int64_t part_off, int nchunk, int64_t stride, int m, int64_t M_off);
/* f32 copy: floats at dst_off_f32 = doubles at src_off (n elements) */
int bk_d2f(int64_t n, int64_t src_off, int64_t dst_off_f32);
/* column scaling with f32 copy: column c (rows elements, c < cols) of the doubles at off is
 * multiplied by the double at d_off + c in place, and the result copied to the floats at
 * dst_off_f32 (the dense route's constraint columns scaled by 1 / ||A_i||) */
int bk_colscale_d2f(int64_t rows, int cols, int64_t off, int64_t dst_off_f32, int64_t d_off);

/* ---- constraint operators (route 0: sparse entries) ---- */
typedef struct { int k, r, c, pad; double v; } SEnt;   /* upper entry (r <= c) of block k */
typedef struct { int n; int pad; int64_t off; } BlkRef; /* block k: order and offset in the solve's block arena */
/* the entries of one constraint in one block (a run of its SEnt list): block k, first entry beg,
 * count cnt; poff >= 0 (doubles) where the product P = X_k A Zi_k of a heavy group is kept */
typedef struct { int k, beg, cnt, pad; int64_t poff; } GrpRef;
/* one item per solve: m constraints; CSR cptr (m+1 ints) into SEnt entries (ent); block table
 * (blk); cptr, ent, blk, pptr, ppos, pcon are BYTE offsets from the heap base, the rest doubles;
 * aop: out[i] = beta out[i] + alpha <A_i, D> (D read as sym(D): both triangles);
 * aty: D[pos] += alpha sum_i y_i A_i over the positions of the position CSR (pptr, pcon, pval);
 * schur: M (m x m full) = sum over common blocks of sum_{p,q} v_p v_q X[c_p, r_q] Zi[c_q, r_p]
 * (full symmetric expansion of each upper entry); where one of the two constraints is heavy in
 * the block (its product P = X A Zi stored, see bk_grp_t), <A_light, P_heavy> instead;
 * gptr (m+1 ints) and grp (GrpRef) are the constraints' groups (BYTE offsets) */
typedef struct {
    int m, nblk;
    int64_t cptr, ent, blk;          /* constraint CSR (int), SEnt, BlkRef */
    int64_t pptr, ppos, pcon, pval;  /* position CSR: npos+1 ints, positions (2 x int: o1, o2 or -1), con (int), val (double) */
    int npos, diag;                  /* diag: every constraint one diagonal entry of one block (Hadamard Schur) */
    int64_t dpos, dval;              /* diag: per constraint its position in the arena (int64) and value */
    int64_t D, out, y, X, Zi, M;     /* operand offsets (doubles) for this call */
    double alpha, beta;
    int64_t gptr, grp;               /* constraint groups (BYTE offsets) */
    int64_t pairs, npair;            /* schur: npair > 0 lists the pairs (a, b), a >= b, sharing a block (2 ints each,
                                        BYTE offset; by column b, rows a ascending): only those are assembled, the
                                        other entries of M keep the zeros they were given once */
} ConItem;
int bk_aop(const ConItem *it, int n_it);
int bk_aty(const ConItem *it, int n_it);
int bk_schur(const ConItem *it, int n_it);
/* T = A Zi_k for heavy groups: the n x n product of a group's sparse symmetric entries (SEnt
 * list at byte offset ent, cnt entries) with the block of Zi at zi; stored at t (doubles) */
typedef struct { int n, cnt; int64_t ent, zi, t; } GrpItem;
int bk_grp_t(const GrpItem *it, int n_it);

/* ---- Lanczos on explicit symmetric T (per block) ---- */
/* step j for every item: given w = T q_j (at w), two passes of classical Gram-Schmidt against
 * q_0..q_j (all coefficients of a pass in one reduction; the first pass's coefficient of q_j is
 * alpha_j); beta_j = |w|; q_{j+1} = w / beta_j. Q holds the basis (K x n per item), ab the
 * coefficients (alpha at ab, beta at ab + 32). Items with j >= K are skipped. */
typedef struct { int n, K; int64_t Q, w, ab; } LzItem;
int bk_lanczos_step(const LzItem *it, int n_it, int j);
/* q_0 = the fixed start vector of the reference (1 + 0.5 sin(1.7 i + 0.3), normalised) */
int bk_lanczos_start(const LzItem *it, int n_it);
/* the whole Lanczos bound in one workgroup per block (order n <= 256): T = V (mode 0) or
 * T = Li V Li' (mode 1, Li lower and LiT = Li' stored); K steps with CGS2 reorthogonalisation,
 * every 3 steps from the 6th (and at K) the smallest Ritz value theta of the tridiagonal and its
 * residual res = |beta z_last|; stops when res <= tol |theta| (+1e-14 scale), at a breakdown or
 * at K. Writes theta - res to out (doubles), -1e300 when the coefficients are not finite. */
typedef struct { int n, K, mode, pad; int64_t Li, LiT, V, out; double tol; } LzfItem;
int bk_lanczos_fused(const LzfItem *it, int n_it);

/* ---- symmetric eigendecomposition (parallel cyclic Jacobi, one workgroup per matrix) ----
 * A (n x n at a, both triangles) = V diag(w) V': V (n x n, column-major) to v, the eigenvalues
 * (unsorted, in the order of V's columns) to w; offsets in doubles from H. The matrix at a is
 * overwritten for n > 64 (transformed in place); n <= 4096. */
typedef struct { int n, pad; int64_t a, v, w; } EigItem;
int bk_syev_f64(const EigItem *it, int n_it, double *H);
/* the Nesterov-Todd frame's scalings of n x n matrices at a (offsets in doubles from H): mode 0
 * A(i, j) *= u(j) (columns), 1 A(i, j) *= u(i) (rows), 2 A(i, j) *= 2 / (u(i) + u(j)); mode 3 on the
 * n eigenvalues w at a: u(i) = sqrt(w(i)), u(n + i) = w(i)^(-1/4), u(2n + i) = w(i)^(1/4) (NaN where
 * w(i) <= 0); mode 4: u(i) = w(i)^(1/4), 0 where w(i) <= 0 */
typedef struct { int n, mode; int64_t a, u; } ScaleItem;
int bk_nt_scale(const ScaleItem *it, int n_it, double *H);

#ifdef __cplusplus
}
#endif
#endif
