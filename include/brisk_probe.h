/* brisk_probe.h - kernel-level entry points of libbrisk_amd.so for parity tests.
 *
 * No counterpart in the reference: they expose single kernel families of the GPU iteration so
 * that each can be checked against an independent computation (the oracle, dense numpy) on
 * caller data, instead of only through a whole solve. Every call builds the problem as a solve
 * does (scaling, constraint lists, device layout), runs the kernels on the GPU and returns
 * results on the data as read. Problem data as for brisk_solve_data (BriskProblem, brisk_amd.h).
 * Block arrays are packed per input block: n x n column-major (SDP), n numbers (LP); the
 * packed length is brisk_probe_packed_len. Returns 0, 2 (bad data) or < 0 (GPU failure). */
#ifndef BRISK_PROBE_H
#define BRISK_PROBE_H
#include <stdint.h>
#include "brisk_amd.h"
#ifdef __cplusplus
extern "C" {
#endif

/* M (m x m, column-major) = the HKM Schur complement M_ij = <A_i, X A_j Z^-1> at SPD X and Z:
 * Z^-1 through the Cholesky factor and its inverse, then the route the solve takes (sparse
 * gathers, heavy-group products, Hadamard form, or the dense route's P'P; prec 1 FP64, 0 the
 * mixed path's FP32 products). The reference's assembly: solver.c:2339-3017, 4931-5058. */
int brisk_probe_schur(const BriskProblem *p, const double *X, const double *Z, int prec, double *M);
/* out_i = <A_i, D> for symmetric D (reference blk_Aop, solver.c:884-992) */
int brisk_probe_aop(const BriskProblem *p, const double *D, double *out);
/* out = sum_i y_i A_i, packed blocks (reference blk_ATy, solver.c:995-1077) */
int brisk_probe_aty(const BriskProblem *p, const double *y, double *out);
/* the Lanczos bound theta - res of the step-length test, the smallest over the blocks: mode 0
 * of lambda_min(X), mode 1 of lambda_min(L^-1 dX L^-T) with X = L L' (reference
 * sdp_maxstep_lam, solver.c:6083-6225; 30 steps, Ritz accuracy 0.02) */
int brisk_probe_lammin(const BriskProblem *p, const double *X, const double *dX, int mode, double *out);
/* the six DIMACS errors and the objectives (obj[0] = <C,X>, obj[1] = b'y) of the point (X, y),
 * as the solve measures a returned point (reference ps_measure_xc, postsolve.c:3037-3112) */
int brisk_probe_dimacs(const BriskProblem *p, const double *X, const double *y, double *err, double *obj);
/* the end phase's metric correction of the point (X, y) into Xout (packed like X), with the step
 * of each pass in steps[3] (0: none taken): half 1 the restoration of A(X) = b in the square-root
 * metric, X <- X + a D A*(w) D with D = X^{1/2} (reference sqrt_restore, solver.c:7551-7655);
 * half 0 the polish in the metric of X, X <- X + a sym(X A*(w) Z^-1) with Z = C - A*(y)
 * (reference polish_xmetric, solver.c:7657-7760) */
int brisk_probe_correct(const BriskProblem *p, const double *X, const double *y, int half, double *Xout, double *steps);
/* the packed length of a problem's block arrays */
int64_t brisk_probe_packed_len(const BriskProblem *p);

#ifdef __cplusplus
}
#endif
#endif
