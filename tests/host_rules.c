/* host_rules.c - a driver for the host-side acceptance rules of the solver (brisk_amd/csrc/solver.h):
 * the NaN-propagating max and min, the acceptance error, the termination score and the status
 * from measured errors. Built and run by tests/test_host.py; prints one line per case:
 * "<name> <value>". */
#include <stdio.h>
#include "solver.h"

int brisk_fault(const char *name) { (void)name; return 0; }

static void e6(double *e, double a, double b, double c, double d, double f, double g) {
    e[0] = a; e[1] = b; e[2] = c; e[3] = d; e[4] = f; e[5] = g;
}

int main(void) {
    const double nan = NAN;
    double e[6];
    printf("nmax_nan_first %.17g\n", nmax(nan, 1.0));
    printf("nmax_nan_second %.17g\n", nmax(1.0, nan));
    printf("nmax_plain %.17g\n", nmax(2.0, 3.0));
    printf("nmin_nan %.17g\n", nmin(nan, -1.0));
    printf("nmin_plain %.17g\n", nmin(2.0, 3.0));
    e6(e, 1e-9, 0, 0, 0, 2e-9, -3e-9);
    printf("acc_plain %.17g\n", acc_err(e));
    printf("status_plain %d\n", status_from_errors(0, e, 1e-8));
    e6(e, 1e-9, nan, 0, 0, 2e-9, 3e-9);
    printf("acc_nan_cone %.17g\n", acc_err(e));
    printf("status_nan_cone %d\n", status_from_errors(0, e, 1e-8));
    e6(e, 1e-9, 0, 0, 0, 2e-9, nan);
    printf("status_nan_compl %d\n", status_from_errors(0, e, 1e-8));
    /* err[2] (the dual equations, exact by construction) is not part of the acceptance error */
    e6(e, 1e-9, 0, nan, 0, 2e-9, 3e-9);
    printf("status_nan_err3 %d\n", status_from_errors(0, e, 1e-8));
    e6(e, 1e-9, 0, 0, 0, 5e-7, 3e-9);
    printf("status_reduced %d\n", status_from_errors(0, e, 1e-8));
    printf("status_limit_reduced %d\n", status_from_errors(3, e, 1e-8));
    e6(e, 1e-9, 0, 0, 0, 5e-5, 3e-9);
    printf("status_numeric %d\n", status_from_errors(0, e, 1e-8));
    printf("status_limit_kept %d\n", status_from_errors(3, e, 1e-8));
    printf("status_infeasible_kept %d\n", status_from_errors(1, e, 1e-8));
    printf("score_nan %.17g\n", iterate_score(1e-9, nan, 1e-9, 1e-9));
    printf("score_plain %.17g\n", iterate_score(1e-9, 2e-9, 3e-9, 4e-10));
    return 0;
}
