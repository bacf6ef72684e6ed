/* dev.h - the thin C ABI between the host solver (C) and the HIP side (rt.hip, k_dense.hip,
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
int bk_h2d(void *dst, const void *src, size_t bytes);        /* synchronous copy */
int bk_d2h(void *dst, const void *src, size_t bytes);        /* synchronous (drains the stream) */
int bk_d2d(void *dst, const void *src, size_t bytes);        /* async device copy */
int bk_zero(void *dst, size_t bytes);                         /* async fill kernel (graph safe) */
int bk_sync(void);
void *bk_stage(const void *host, size_t bytes);  /* device copy of host data, valid until next sync */
double bk_now(void);                             /* host wall clock, s */
size_t bk_mem_free(void);                        /* free device memory, bytes */
/* events on the library stream (timing of kernels in bench): create once, record, elapsed ms */
void *bk_event_new(void);
void bk_event_record(void *e);
float bk_event_ms(void *a, void *b);
void bk_event_free(void *e);
const char *bk_device_name(void);

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
} MItem;

/* Cholesky of a set of matrices, blocked right-looking with 64-wide panels, in place on A (its
 * lower triangle is read; on return A holds L with the strict upper triangle zeroed). Diagonal-
 * block inverses go to Dinv + off2 (64*64 per panel); W + off3 is scratch (n*64 used).
 * info[slot] = first failing column + 1 (0: factored); pmin[slot] = smallest squared pivot. */
int bk_potrf_f64(const MItem *it, int n_it, double *A, double *Dinv, double *W, int *info, double *pmin);
int bk_potrf_f32(const MItem *it, int n_it, float *A, float *Dinv, float *W, int *info, float *pmin);
/* Linv + off = L^{-1} (lower, strict upper zero) from the factor L + off and its diagonal-block
 * inverses (Dinv + off2); W + off3 scratch (n*n). Items with info[slot] != 0 are skipped. */
int bk_trtri_f64(const MItem *it, int n_it, const double *L, const double *Dinv, double *Linv, double *W, const int *info);
int bk_trtri_f32(const MItem *it, int n_it, const float *L, const float *Dinv, float *Linv, float *W, const int *info);

/* ---- batched matrix-vector: y_it = alpha op(A_it) x_it + beta y_it ---- */
typedef struct { int M, N, lda; int64_t a, x, y; } GemvItem;
int bk_gemv_f64(int trans, double alpha, double beta, const double *A, const double *x, double *y, const GemvItem *it, int n_it);
int bk_gemv_f32(int trans, float alpha, float beta, const float *A, const float *x, float *y, const GemvItem *it, int n_it);
/* mixed: y (f64) = alpha op(A) x + beta y with A in f32, x in f64 (products and sums in f64) */
int bk_gemv_mixed(int trans, double alpha, double beta, const float *A, const double *x, double *y, const GemvItem *it, int n_it);

/* ---- segmented deterministic reductions ---- */
typedef struct { int64_t off, len; } Seg;
/* out[s] = sum_q a[off+q] * b[off+q] over each segment (b == NULL: sum of a^2 ... a == b allowed);
 * the order of summation is fixed by the segment length alone */
int bk_seg_dot(const Seg *segs, int nseg, const double *a, const double *b, double *out);
/* out[s] = sum of a[off+q] */
int bk_seg_sum(const Seg *segs, int nseg, const double *a, double *out);

/* ---- elementwise over a flat range ---- */
/* y = a*x + b*y + c*z  (any of x, z may be NULL) */
int bk_axpbypcz(int64_t n, double a, const double *x, double b, double *y, double c, const double *z);
/* per-segment scalars: y[q] = x[q] + s[seg] * d[q] (y may alias x); V update of try_step */
int bk_seg_axpy(const Seg *segs, int nseg, const double *s, const double *x, const double *d, double *y);
/* per-segment scalar a[s]: y[q] += a[s] * x[q] */
int bk_seg_addscal(const Seg *segs, int nseg, const double *a, const double *x, double *y);
/* symmetrise square blocks: G = (G + G') / 2 */
int bk_symmetrize(const MItem *it, int n_it, double *G);
/* copy the lower triangle of src into dst and zero dst's strict upper triangle */
int bk_copy_lower(const MItem *it, int n_it, const double *src, double *dst);
/* add s * I to every matrix */
int bk_add_diag(const MItem *it, int n_it, const double *s_per_slot, double *A);
/* equilibrate and convert: F = D A D (lower from A's full storage), D = 1/sqrt(diag A) (or 1) */
int bk_equil_f32(const MItem *it, int n_it, const double *A, float *F, double *D, const int64_t *doff);
int bk_equil_f64(const MItem *it, int n_it, const double *A, double *F, double *D, const int64_t *doff, const double *reg_per_slot);
/* v[q] = v[q] * s[q] (f64) and conversions */
int bk_vmul(int64_t n, const double *a, const double *b, double *out);
int bk_d2f(int64_t n, const double *a, float *out);
int bk_f2d(int64_t n, const float *a, double *out);

#ifdef __cplusplus
}
#endif
#endif
