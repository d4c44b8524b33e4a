/*
 * Acceptance program: the reference's example call sequence (test/ex.c:95,
 * :100-101) and timing-harness globals (test/runtime_analysis.c) against
 * libspike_pstrsv.so, in plain C.
 *
 * Matrix: lower triangle of the 2D 5-point Laplacian on a g x g grid (diag
 * 4, off-diagonals -1), and its upper-triangular transpose.  The right-hand
 * side is b = A x_real with the harness's manufactured solution
 * x_real[i] = (rand() % 71 + 1) + (i % 1000) / 10.0 (runtime_analysis.c:315-316),
 * and the check is ||x - x_real||_2 <= 1e-4 (runtime_analysis.c:434-435).
 * Prints PASSED and exits 0 when every check holds.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spike_pstrsv.h"
#include "spike_pstrsv_globals.h"

static int lap2d(int g, int upper, double **a, MKL_INT **ia, MKL_INT **ja)
{
    int n = g * g, nnz = 0;
    *ia = malloc(sizeof(MKL_INT) * (n + 1));
    *ja = malloc(sizeof(MKL_INT) * 3 * n);
    *a = malloc(sizeof(double) * 3 * n);
    (*ia)[0] = 0;
    for (int i = 0; i < n; i++) {
        int x = i % g, y = i / g;
        if (!upper) {
            if (y > 0) { (*ja)[nnz] = i - g; (*a)[nnz++] = -1.0; }
            if (x > 0) { (*ja)[nnz] = i - 1; (*a)[nnz++] = -1.0; }
            (*ja)[nnz] = i; (*a)[nnz++] = 4.0;
        } else {
            (*ja)[nnz] = i; (*a)[nnz++] = 4.0;
            if (x < g - 1) { (*ja)[nnz] = i + 1; (*a)[nnz++] = -1.0; }
            if (y < g - 1) { (*ja)[nnz] = i + g; (*a)[nnz++] = -1.0; }
        }
        (*ia)[i + 1] = nnz;
    }
    return n;
}

static int check(char uplo, int g, int nthreads, int profiled)
{
    double *a;
    MKL_INT *ia, *ja;
    int n = lap2d(g, uplo == 'U', &a, &ia, &ja);
    double *xr = malloc(sizeof(double) * n), *b = malloc(sizeof(double) * n), *x = calloc(n, sizeof(double));
    srand(1);
    for (int i = 0; i < n; i++) xr[i] = (rand() % 71 + 1) + (i % 1000) / 10.0;
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int e = ia[i]; e < ia[i + 1]; e++) s += a[e] * xr[ja[e]];
        b[i] = s;
    }
    if (profiled) {
        _spike_pstrsv_preproc(a, ia, ja, n, nthreads, uplo);
        _spike_pstrsv_solve(uplo, n, b, x, nthreads);
    } else {
        spike_pstrsv_preproc(a, ia, ja, n, nthreads, uplo);
        spike_pstrsv_solve(uplo, n, b, x, nthreads);
    }
    double err = 0.0;
    for (int i = 0; i < n; i++) err += (x[i] - xr[i]) * (x[i] - xr[i]);
    err = sqrt(err);
    int ok = err <= 1e-4;
    /* one-shot form as well */
    double *x2 = calloc(n, sizeof(double));
    spike_pstrsv(uplo, n, a, ia, ja, b, x2, nthreads);
    double d = 0.0;
    for (int i = 0; i < n; i++) d = fmax(d, fabs(x2[i] - x[i]));
    ok = ok && d <= 1e-10;
    printf("uplo %c g %d P %d profiled %d: ||x - x_real||_2 = %.3e, one-shot diff %.3e -> %s\n", uplo, g, nthreads,
           profiled, err, d, ok ? "ok" : "FAIL");
    free(a); free(ia); free(ja); free(xr); free(b); free(x); free(x2);
    return ok;
}

int main(void)
{
    int ok = 1;
    PSTRSV_VERBOSE = 1;
    ok &= check('L', 96, 8, 0);
    PSTRSV_VERBOSE = 0;
    ok &= check('U', 96, 8, 0);
    ok &= check('L', 128, 16, 1);
    printf("timers: preprocessD %.3e s, produceS %.3e s, produceReducedSystem %.3e s, reduced solve %.3e s, "
           "computeY %.3e s, subtract_Rx %.3e s, solveD2 %.3e s\n",
           preprocessD_timer_pstrsv3, produceS_timer_pstrsv3, produceReducedSystem_timer_pstrsv3,
           reduced_solver_timer_pstrsv3, computeY_timer[0], subtract_Rx_timer[0], solveD2_timer[0]);
    ok &= preprocessD_timer_pstrsv3 > 0 && computeY_timer[0] > 0 && solveD2_timer[0] > 0;
    printf("%s\n", ok ? "PASSED" : "FAILED");
    return ok ? 0 : 1;
}
