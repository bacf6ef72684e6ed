/* brisk_amd.h - the C ABI of the MI355X SDP interior-point solver (libbrisk_amd.so).
 *
 * The first group is a drop-in for the reference library's flat interface (capi.c of
 * MilanKorda/brisk 1.7): same names, argument meaning, return codes and result accessors, so a
 * ctypes / ccall / MEX binding written for libbrisk binds this library unchanged. The solve
 * runs on the GPU; options select its precision ("-prec double" or "-prec mixed").
 *
 * Conventions (reference brisk.h:1-9, capi.c:15-17): (P) min <C,X> s.t. <A_i,X> = b_i, X in K;
 * (D) max b'y s.t. C - sum y_i A_i = Z in K, with C = -F0, A_i = F_i, b = c of an SDPA file.
 * X and Z blocks are n x n column-major (SDP) or n (LP). Not reentrant: serialize calls (the
 * reference's rule, capi.c:19-20); the batch and serving entries below are how several
 * problems share the GPU. */
#ifndef BRISK_AMD_H
#define BRISK_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct BriskResult BriskResult;

/* replaces brisk_set_output (capi.c:60-63): 0 stdout/stderr, 1 errors only, 2 callback */
void brisk_set_output(int mode, int (*cb)(const char *s, int is_err));
/* replaces brisk_version (capi.c:65) */
const char *brisk_version(void);
/* replace brisk_result_new / brisk_result_delete (capi.c:68-77) */
BriskResult *brisk_result_new(void);
void brisk_result_delete(BriskResult *r);
/* replace the accessors of capi.c:78-139 */
int brisk_result_status(const BriskResult *r);            /* 0..6, -1 not solved */
const char *brisk_result_status_str(const BriskResult *r);
int brisk_result_exit_code(const BriskResult *r);
int brisk_result_iterations(const BriskResult *r);
int brisk_result_m(const BriskResult *r);
int brisk_result_nblk(const BriskResult *r);
int brisk_result_blocksize(const BriskResult *r, int k);
double brisk_result_pobj(const BriskResult *r);
double brisk_result_dobj(const BriskResult *r);
double brisk_result_time(const BriskResult *r);
int brisk_result_have_x(const BriskResult *r);
void brisk_result_dimacs(const BriskResult *r, double *err);   /* err[0..5] = DIMACS errors 1..6 */
const double *brisk_result_y(const BriskResult *r);
const double *brisk_result_X(const BriskResult *r, int k);
const double *brisk_result_Z(const BriskResult *r, int k);
const char *brisk_result_cause(const BriskResult *r);
int brisk_result_resolves(const BriskResult *r, double *t);
/* replace brisk_interrupt / brisk_interrupted / brisk_clear_interrupt / brisk_set_interrupt_owner
 * (capi.c:156-169): the running solve stops at its next iteration with status TIME LIMIT */
void brisk_interrupt(void);
int brisk_interrupted(void);
void brisk_clear_interrupt(void);
void brisk_set_interrupt_owner(int caller);
/* replaces brisk_solve_data (capi.c:175-192): the numbers of an SDPA sparse file. Returns the
 * reference's exit code: 0 optimal, 10 reduced accuracy, 11/12 primal/dual infeasible,
 * 13 iteration limit, 14 numerical difficulties, 15 time limit, 1 bad option, 2 bad data,
 * < 0 aborted (no GPU, out of device memory). */
int brisk_solve_data(int m, int nblk, const int *bs, const double *c, int64_t nnz, const int *mat, const int *blk,
                     const int *i, const int *j, const double *v, int nopt, const char *const *opts, BriskResult *res);
/* replaces brisk_solve_file (capi.c:520-533) for SDPA sparse files (.dat-s) */
int brisk_solve_file(const char *fname, int nopt, const char *const *opts, BriskResult *res);

/* ---- entries with no counterpart in the reference -------------------------------------- */
/* Options understood by every solve entry:
 *   -prec double|mixed   FP64 throughout, or FP32-MFMA Schur factor refined in FP64 (default:
 *                        mixed when m >= 1000, the reference's own threshold, solver.c:13814)
 *   -tol t, -maxit k, -timelimit s, -v, -q; -threads k and -bm k are accepted and ignored;
 *   -schurapply auto|inverse|solves   how the PCG applies the Schur factor (default auto: a
 *                        cost model on the solve's own order and precision);
 *   -throughput 0|1      the plan choices for latency (0, the default of the solve entries) or
 *                        for throughput in large batches (1, the serving loop's default):
 *                        256-column Cholesky panels from order 1024 and blocked solves for FP64
 *                        factors from order 1500. A solve's arithmetic depends on its own
 *                        options and data only, never on the problems batched with it. */
/* time split of the solve (s): out[0] host setup and upload, out[1] start point and iterations,
 * out[2] DIMACS evaluation on the device, out[3] result download */
void brisk_result_phases(const BriskResult *r, double *out);
/* iterations that used the f32 / f64 factor, PCG steps */
void brisk_result_counts(const BriskResult *r, int *out3);
/* trace of the run: objectives and mu per iteration (n <= 128 entries each); returns the count */
int brisk_result_trace(const BriskResult *r, double *pobj, double *dobj, double *mu, int n);

/* A problem for the batch and serving entries (the arguments of brisk_solve_data). */
typedef struct {
    int m, nblk;
    const int *bs;
    const double *c;
    int64_t nnz;
    const int *mat, *blk, *i, *j;
    const double *v;
} BriskProblem;
/* replaces the reader of problem.c:426 (read_sdpa) that brisk_solve_file uses: the numbers of
 * an SDPA sparse file into p, whose arrays live in one allocation returned in *owned (release
 * it with brisk_sdpa_free). Returns 0, or -1 when the file cannot be read or parsed. */
int brisk_read_sdpa(const char *fname, BriskProblem *p, void **owned);
void brisk_sdpa_free(void *owned);

/* Solve n problems together: every iteration is one set of batched launches over all of them
 * (their results equal their single solves'). res[k] receives problem k; returns 0 when every
 * problem was read and solved (statuses in the results), else the first error code. */
int brisk_solve_batch(int n, const BriskProblem *probs, int nopt, const char *const *opts, BriskResult **res);

/* Serving: a persistent request loop that batches solves of uneven sizes as they arrive.
 * brisk_serve_new: options as above plus -slots k (most solves in flight, default 64) and
 * -heapgb g (device memory for solves in flight; admission by each solve's projected bytes).
 * brisk_serve_submit copies the problem and returns a request id (>= 0) or < 0 (bad data).
 * brisk_serve_step admits waiting requests while slots and memory allow, starts them, runs one
 * iteration of every solve in flight, and finishes those that ended; returns the number of
 * requests still waiting or in flight. brisk_serve_take moves a finished request's result into
 * res (1 when finished, 0 not yet, < 0 unknown id); latency = finish - submit time (s). */
typedef struct BriskServe BriskServe;
BriskServe *brisk_serve_new(int nopt, const char *const *opts);
int brisk_serve_submit(BriskServe *sv, const BriskProblem *p);
int brisk_serve_step(BriskServe *sv);
int brisk_serve_take(BriskServe *sv, int id, BriskResult *res, double *latency);
int brisk_serve_inflight(const BriskServe *sv);
/* ids of the requests finished since the last call (at most cap), in finishing order */
int brisk_serve_done(BriskServe *sv, int *ids, int cap);
void brisk_serve_delete(BriskServe *sv);

#ifdef __cplusplus
}
#endif
#endif
