/*
 * CPU ORACLE for the spike_pstrsv solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's algorithm, used by tests/,
 * the smoke check and the benchmark's CPU baseline as the checker.  The product
 * library never links, loads or calls it.
 *
 * Reference files restated (paths relative to the reference repository):
 *   src/spike_pstrsv.c          api.c   (solve phase order, :149-239)
 *   src/spike_pstrsv_ops.c      ops.c   (dependents :129-182, spike sizing
 *                                        :214-373, spikes :1021-1069 + :1196-1228,
 *                                        reduced system :1071-1150, kernels
 *                                        :1152-1279)
 *   src/matrix_partitioning.c   mp.c    (partitioning :18-36, D extraction :38-76)
 *   test/runtime_analysis.c     ra.c    (udsol_/ldsol_ :37-65, manufactured
 *                                        solution :315-318)
 *
 * How each function is pinned (tests/test_oracle.py):
 *   oracle_ldsol / oracle_udsol    -- brute-force dense substitution and
 *                                     residual ||Ax-b|| on random systems; the
 *                                     ra.c manufactured-solution check.
 *   oracle_spmv                    -- dense matrix-vector product.
 *   oracle_partition               -- closed form of mp.c:18-33.
 *   oracle_flip                    -- U = J L J element-wise on tiny inputs.
 *   oracle_manufactured_x          -- glibc rand() sequence (known values).
 *   oracle_spike_* ('L' and 'U')   -- x equals sequential substitution
 *                                     (rel. err <= 1e-13) for many P;
 *                                     dependents / K_p / intervals / reduced
 *                                     pattern against brute force from the
 *                                     definitions on tiny CSR; spike invariant
 *                                     D0 * W = C[:,K]; 2D-Laplacian spike
 *                                     entries against the closed-form Green's
 *                                     function; 'U' == flipped 'L'.
 */
#ifndef SPIKE_ORACLE_H
#define SPIKE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sequential forward substitution, diagonal = last entry of each row
 * (ops.c:1264-1279, ra.c:52-65). */
void oracle_ldsol(int n, const double *a, const int *ia, const int *ja,
                  const double *b, double *x);

/* Sequential back substitution, diagonal = first entry of each row
 * (ops.c:1179-1194, ra.c:37-50). */
void oracle_udsol(int n, const double *a, const int *ia, const int *ja,
                  const double *b, double *x);

/* y = A x on zero-based CSR (the mkl_dcsrmv('N', alpha=1, beta=0) call of
 * ra.c:27-35 used to build b = A x_real). */
void oracle_spmv(int n, const double *a, const int *ia, const int *ja,
                 const double *x, double *y);

/* Equal row blocks, the last one takes the remainder (mp.c:18-33).
 * starts[P+1]: partition t is rows [starts[t], starts[t+1]). */
void oracle_partition(int n, int P, int *starts);

/* x_real[i] = (rand() % 71 + 1) + (i % 1000) / 10.0 after srand(seed)
 * (ra.c:315-316; the reference never seeds, i.e. seed 1). */
void oracle_manufactured_x(int n, unsigned seed, double *x);

/* Index reversal U = J L J (and back): row i of the output is row n-1-i of
 * the input with columns c -> n-1-c and the entry order reversed, so the
 * diagonal moves from last to first position (or vice versa). */
void oracle_flip(int n, const double *a, const int *ia, const int *ja,
                 double *a2, int *ia2, int *ja2);

/* ---------------- Spike restatement ---------------- */

typedef struct oracle_spike oracle_spike;

typedef struct {
    int n, P;
    int64_t nnz;
    int n_reduced;             /* |dependents| (numRowsReducedSystem)          */
    int64_t nnz_reduced;       /* kept off-diagonals + n_reduced (ops.c:1149)  */
    int64_t spike_area;        /* sum over partitions of rows_S * cols_S       */
    double dropped_rowsum_max; /* max over reduced rows of sum |dropped w|     */
} oracle_stats_t;

/* Per-partition information, mirrors logPSTRSV3 (ops.c:1311-1333). */
typedef struct {
    int start, end;            /* rows [start, end)                            */
    int rows_S, cols_S;        /* spike block size (numRows_S, numCols_S)      */
    int first_coupling;        /* local row of first coupling entry, or k      */
    int last_dependent;        /* local row of last dependent, or -1           */
    int has_dependence;        /* partition has coupling entries              */
    int has_reflection;        /* partition holds dependent rows              */
} oracle_part_info;

/* Preprocess (symbolic + spikes + reduced system).  tau: spike entries with
 * |w| <= tau are dropped from the reduced system (tau = 0 keeps every nonzero,
 * as ops.c:1127 does).  nthreads <= 0 uses the OpenMP default.
 * Returns NULL on invalid input (error text via oracle_last_error). */
oracle_spike *oracle_spike_preproc(char uplo, int n, const double *a, const int *ia,
                                   const int *ja, int P, double tau, int nthreads);
/* Solve A x = b (b is not modified). */
void oracle_spike_solve(oracle_spike *s, const double *b, double *x);
void oracle_spike_stats(const oracle_spike *s, oracle_stats_t *st);
int  oracle_spike_part_info(const oracle_spike *s, int p, oracle_part_info *info);
/* Dependent row list (sorted) -> dep[n_reduced]. */
void oracle_spike_dependents(const oracle_spike *s, int *dep);
/* K_p (sorted global columns) of partition p -> cols[cols_S]. */
int  oracle_spike_kcols(const oracle_spike *s, int p, int *cols);
/* Dense spike block of partition p, row-major rows_S x cols_S. */
int  oracle_spike_block(const oracle_spike *s, int p, double *w);
/* Reduced system in CSR over reduced indices (off-diagonals only). */
void oracle_spike_reduced(const oracle_spike *s, int *rp, int *ci, double *v);
void oracle_spike_free(oracle_spike *s);
const char *oracle_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
