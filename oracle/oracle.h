/* oracle.h - TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference's SDP interior-point
 * iteration (ipm_oracle.c). Loaded by tests/, the smoke entry and bench.py's CPU-baseline leg;
 * never by the product library. */
#ifndef ORACLE_H
#define ORACLE_H
#include <stddef.h>
#include <stdint.h>
#define ORACLE_TRACE 128
typedef struct {                /* the numbers of an SDPA sparse file (capi.c:175-192) */
    int m, nblk;
    const int *bs;
    const double *c;
    int64_t nnz;
    const int *mat, *blk, *i, *j;
    const double *v;
} OracleData;
typedef struct { double tol; int maxit; int method; } OracleOptions;   /* method 0: infeasible start, 1: self-dual embedding */
typedef struct {
    int status, iters, m, nblk;
    double pobj, dobj, acc;
    double err[6];              /* DIMACS errors 1..6 */
    double *y, *X;              /* y (m); X packed over blocks (n*n column-major SDP, n LP) */
    size_t xlen;
    double trace_pobj[ORACLE_TRACE], trace_dobj[ORACLE_TRACE], trace_mu[ORACLE_TRACE], trace_score[ORACLE_TRACE];
    double trace_ap[ORACLE_TRACE], trace_ad[ORACLE_TRACE];
} OracleResult;
int oracle_solve(const OracleData *d, const OracleOptions *opt, OracleResult *res);
void oracle_result_free(OracleResult *r);
int oracle_dimacs(const OracleData *d, const double *X, const double *y, double *err, double *pobj, double *dobj);
int oracle_schur(const OracleData *d, const double *X, const double *Zi, double *M);
double oracle_lammin(int n, double *B);
double oracle_lanczos(int n, const double *L, const double *D, int K, double tol, double *info3);
double oracle_step_lambda(int n, const double *L, const double *D);
int oracle_correct(const OracleData *d, const double *X, const double *y, int half, int maxpass, double *Xout, double *steps);
#endif
