/*
 * spike_pstrsv for AMD Instinct MI355X (gfx950) -- public C ABI.
 *
 * Drop-in replacement for the reference library's public interface
 * (src/spike_pstrsv.h in the reference repository).  The five legacy entry
 * points keep their exact C signatures; MKL_INT is a 32-bit integer, as in the
 * reference's LP64 MKL link (Makefile.sample:15).
 *
 * Matrices are zero-based 3-array CSR (ia[0] == 0, columns ascending, no empty
 * rows), with the diagonal stored as the first entry of each row for
 * uplo 'U'/'u' and as the last entry for 'L'/'l' (reference
 * src/spike_pstrsv_ops.c:1186 and :1277).  A zero or non-finite diagonal
 * value is rejected (SPIKE_ERR_MATRIX); the reference divides by it.
 *
 * Differences in behaviour from the reference, all compatible with its callers:
 *   - 'L' is solved (the reference leaves x untouched for 'L');
 *   - b is never modified (the reference uses it as scratch);
 *   - the number of Spike partitions is independent of nthreads; by default
 *     it equals nthreads, and SPIKE_PSTRSV_PARTITIONS or
 *     spike_pstrsv_set_partitions() override it;
 *   - errors never exit(): the legacy void functions print one line to stderr
 *     and leave x untouched; the extension API returns status codes;
 *   - host and device pointers are both accepted for b and x.
 */
#ifndef SPIKE_PSTRSV_H
#define SPIKE_PSTRSV_H

#include <stddef.h>
#include <stdint.h>

#ifndef MKL_INT
#define MKL_INT int
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Legacy entry points (one implicit, library-owned current factor)    */
/* ------------------------------------------------------------------ */

/* Replaces spike_pstrsv_preproc, reference src/spike_pstrsv.h:7
 * (implementation src/spike_pstrsv.c:40-81).  Note: uplo is the LAST argument. */
void spike_pstrsv_preproc(double *a, MKL_INT *ia, MKL_INT *ja, int n, int nthreads, const char uplo);

/* Replaces _spike_pstrsv_preproc, reference src/spike_pstrsv.h:14
 * (src/spike_pstrsv.c:83-140): same as spike_pstrsv_preproc, and records
 * per-phase wall times (see spike_pstrsv_get_timers). */
void _spike_pstrsv_preproc(double *a, MKL_INT *ia, MKL_INT *ja, int n, int nthreads, const char uplo);

/* Replaces spike_pstrsv, reference src/spike_pstrsv.h:21
 * (src/spike_pstrsv.c:142-147): preprocess, then solve. */
void spike_pstrsv(const char uplo, int m, double *a, MKL_INT *ia, MKL_INT *ja, double *b, double *x, int nthreads);

/* Replaces spike_pstrsv_solve, reference src/spike_pstrsv.h:31
 * (src/spike_pstrsv.c:149-239): solves with the current factor. */
void spike_pstrsv_solve(const char uplo, int m, double *b, double *x, int nthreads);

/* Replaces _spike_pstrsv_solve, reference src/spike_pstrsv.h:38
 * (src/spike_pstrsv.c:241-355): same as spike_pstrsv_solve, and records
 * per-phase device times (see spike_pstrsv_get_timers). */
void _spike_pstrsv_solve(const char uplo, int m, double *b, double *x, int nthreads);

/* ------------------------------------------------------------------ */
/* Extension API (no counterpart in the reference)                      */
/* ------------------------------------------------------------------ */

/* Status codes returned by the extension API. */
enum {
    SPIKE_OK = 0,
    SPIKE_ERR_ARG = 1,        /* invalid argument (uplo, n, nrhs, ld, pointers) */
    SPIKE_ERR_MATRIX = 2,     /* CSR violates the conventions above            */
    SPIKE_ERR_DEVICE = 3,     /* HIP runtime error / no usable gfx950 device   */
    SPIKE_ERR_NOMEM = 4,      /* device or host allocation failed              */
    SPIKE_ERR_STATE = 5,      /* e.g. solve before preprocess, uplo mismatch   */
    SPIKE_ERR_UNSUPPORTED = 6 /* configuration outside the implemented range   */
};

/* Precision modes of a factor. */
enum {
    SPIKE_PREC_FP64 = 0,      /* FP64 values and arithmetic throughout         */
    SPIKE_PREC_FP32_REFINED = 1 /* block solves and coupling updates with FP32 */
                              /* values and FP32 accumulation; the residual    */
                              /* b - A x, the reduced system and the update    */
                              /* x += d stay FP64 (iterative refinement)       */
};

typedef struct spike_factor_s *spike_factor_t;

typedef struct {
    int partitions;           /* Spike partition count P; 0 = automatic        */
    double drop_tol;          /* reduced-system drop tolerance on the spike    */
                              /* row sum (sum of dropped |w| <= drop_tol per   */
                              /* row); < 0 = default (1e-15), 0 = keep all     */
                              /* nonzeros (as the reference)                   */
    int precision;            /* SPIKE_PREC_*                                  */
    int lanes_per_row;        /* block-solve lanes per row slot (1, 2, 4, 8   */
                              /* or 16; longer rows run as chunks of that many */
                              /* entries in successive steps); 0 = automatic   */
    int verbose;              /* print the partition log after preprocessing   */
    int reduced_solver;       /* SPIKE_REDUCED_*                               */
    int refine_steps;         /* SPIKE_PREC_FP32_REFINED: refinement steps per */
                              /* solve; < 0 = calibrated at preprocessing on a */
                              /* random right-hand side until the predicted    */
                              /* relative error is <= refine_tol               */
    double refine_tol;        /* target of the calibration; <= 0 = 1e-13       */
    int method;               /* SPIKE_METHOD_*                                */
    int use_graph;            /* 1: a device-pointer solve of many kernel      */
                              /* launches (>= 24: level schedules, refinement) */
                              /* whose (b, x, ldb, ldx, nrhs) repeats the      */
                              /* previous call's is captured into a hipGraph   */
                              /* once and replayed (up to 16 cached per        */
                              /* factor); shorter solves launch directly,      */
                              /* which costs the device less per solve;        */
                              /* 2: every such repeated solve is captured      */
                              /* (least host time per solve; the serving loop  */
                              /* uses it); 0: never                            */
    int autotune;             /* 1: time the RHS tile candidates at creation;  */
                              /* 2: also build and time half and twice the     */
                              /* automatic partition count (partitions == 0)   */
                              /* and lanes per row (lanes_per_row == 0) over   */
                              /* one-RHS solves, keeping the fastest           */
} spike_options_t;

/* Solve methods. */
enum {
    SPIKE_METHOD_AUTO = 0,    /* Spike unless its reduced system degenerates   */
                              /* (dependents > n/8) or the level schedule is   */
                              /* cheaper by the cost model                     */
    SPIKE_METHOD_SPIKE = 1,   /* Spike partitions + reduced system             */
    SPIKE_METHOD_LEVELS = 2   /* level-scheduled substitution over the whole   */
                              /* GPU, one launch per dependency level (for     */
                              /* couplings wider than a partition, e.g. 3D     */
                              /* grids)                                        */
};

/* Reduced-system solvers. */
enum {
    SPIKE_REDUCED_AUTO = 0,   /* cheaper of the two by a cost model            */
    SPIKE_REDUCED_EXACT = 1,  /* level-scheduled forward substitution, one     */
                              /* workgroup                                     */
    SPIKE_REDUCED_JACOBI = 2  /* S Jacobi sweeps x = y - W x over the whole    */
                              /* GPU; S is fixed at preprocessing so that the  */
                              /* bound sum_{j>S} max(|W|^j 1) <= 2e-15 holds   */
};

/* Fills opts with the defaults. */
void spike_options_default(spike_options_t *opts);

/* Analyses and factors the CSR matrix (host pointers).  The matrix is copied;
 * the caller keeps ownership of a/ia/ja. */
int spike_factor_create(spike_factor_t *f, char uplo, int n, const double *a,
                        const int *ia, const int *ja, const spike_options_t *opts);

/* Same as spike_factor_create with the caller's partition boundaries:
 * nparts partitions of rows [starts[t], starts[t+1]) in the caller's row
 * order (0 = starts[0] < ... < starts[nparts] = n), each no longer than the
 * device's LDS limit (16383 rows in FP64).  Replaces the reference's
 * IS_MATRIX_PARTITIONED path (src/matrix_partitioning.c:23, :34-35), where
 * partition_starts/partition_ends come from the caller (e.g. a reordering);
 * the factor always uses the Spike method. */
int spike_factor_create_partitioned(spike_factor_t *f, char uplo, int n, const double *a, const int *ia,
                                    const int *ja, int nparts, const int *starts, const spike_options_t *opts);

/* Solves A X = B for nrhs right-hand sides (column-major, leading dimensions
 * ldb/ldx >= n).  b and x may be host or device pointers.  stream is a
 * hipStream_t (NULL = default stream); for device pointers the call is
 * asynchronous with respect to the host, for host pointers it synchronises. */
int spike_factor_solve(spike_factor_t f, int nrhs, const double *b, int ldb,
                       double *x, int ldx, void *stream);

/* Same solve as spike_factor_solve, for callers that cannot allocate memory:
 * code under hipGraph stream capture, tight serving loops, harnesses that
 * forbid hipMalloc.  No counterpart in the reference.  The call allocates
 * nothing (no hipMalloc / hipFree, no staging, no graph instantiation): b and
 * x must be device pointers (SPIKE_ERR_ARG otherwise), and the solve's
 * workspaces (reduced vectors, coupling right-hand sides, residuals,
 * interleaved vectors) come from a static device array of 64 MiB that every
 * device the library runs on holds until the process ends.  A solve whose
 * workspaces exceed it fails with SPIKE_ERR_UNSUPPORTED before launching
 * anything.  The array is shared by all calls on one device and keeps the
 * last call's contents (every kernel writes it before reading it): calls
 * that overlap on one device race and must be ordered by the caller (one
 * stream, or events).  Batch groups are not supported. */
int spike_factor_solve_noalloc(spike_factor_t f, int nrhs, const double *b, int ldb,
                               double *x, int ldx, void *stream);

/* Same as spike_factor_solve, and adds the device time of each phase, in
 * milliseconds, to phase_ms[0..3]: pass-1 block solve, reduced solve, coupling
 * update, pass-2 block solve (hipEvents recorded on stream).  Synchronises. */
int spike_factor_solve_profiled(spike_factor_t f, int nrhs, const double *b, int ldb,
                                double *x, int ldx, void *stream, double *phase_ms);

int spike_factor_destroy(spike_factor_t f);

/* Device waits of a solve are bounded (the sync-free multi-RHS level solve
 * and the single-launch Spike solve poll values that other workgroups
 * publish); a solve whose wait gives up writes NaN where rows could not be
 * computed.  That outcome is returned as SPIKE_ERR_DEVICE by the call itself
 * when it synchronises (host-pointer and profiled solves), otherwise by the
 * next spike_factor_solve* call on the factor or by spike_factor_status,
 * which the caller may call after synchronising the solve's stream.  Each
 * give-up is reported once.  No counterpart in the reference, whose solve
 * has no device waits. */
int spike_factor_status(spike_factor_t f);

/* Writes a factor (partition maps, block-solve schedules, spikes' reduced
 * system, coupling data, solver choices) to a file, so that a later process
 * can skip preprocessing; spike_factor_load re-creates it in device memory
 * on the current device.  The file is native-endian and tied to this
 * library's version; batch groups are not saved. */
int spike_factor_save(spike_factor_t f, const char *path);
int spike_factor_load(spike_factor_t *f, const char *path);

typedef struct {
    int n;
    int partitions;           /* effective P                                   */
    int64_t nnz;
    int n_reduced;            /* dependents = reduced-system size              */
    int64_t nnz_reduced;      /* kept off-diagonals + n_reduced                */
    int reduced_levels;       /* dependency levels of the reduced system       */
    int64_t spike_area;       /* sum of spike block sizes (rows x cols)        */
    double dropped_rowsum_max;/* largest sum of dropped |w| over reduced rows  */
    int steps_pass1;          /* longest block-solve step chain, pass 1        */
    int steps_pass2;          /* longest block-solve step chain, pass 2        */
    int lanes_per_row;        /* block-solve lanes per row in use              */
    int64_t device_bytes;     /* device memory held by the factor              */
    double preproc_seconds;   /* wall time of spike_factor_create              */
    int reduced_sweeps;       /* Jacobi sweeps per solve, 0 = exact solver     */
    double jacobi_bound;      /* bound on the reduced solve's relative error   */
    int rhs_per_launch;       /* right-hand sides per solve launch sequence    */
    int precision;            /* SPIKE_PREC_* of the factor                    */
    int refine_steps;         /* refinement steps per solve (FP32_REFINED)     */
    double refine_rate;       /* calibrated contraction per refinement step    */
    double refine_est_err;    /* calibrated predicted relative error           */
    int method;               /* SPIKE_METHOD_SPIKE or SPIKE_METHOD_LEVELS     */
    int levels;               /* dependency levels of the whole matrix (0 if   */
                              /* not computed)                                 */
    int rhs_tile[4];          /* right-hand sides per block-solve wave at 1,   */
                              /* 2, 4, 8 slots (Spike method)                  */
    int single_launch;        /* 1: a Spike solve of up to 8 right-hand sides  */
                              /* is one kernel launch (pass 1, the Jacobi      */
                              /* sweeps and pass 2 as items of one persistent  */
                              /* grid; FP64 factors, opt-in with the           */
                              /* environment SPIKE_PSTRSV_FUSED=1); 0: a       */
                              /* launch sequence                               */
} spike_stats_t;

int spike_factor_stats(spike_factor_t f, spike_stats_t *st);

/* Partition log in the reference's logPSTRSV3 field order
 * (src/spike_pstrsv_ops.c:1311-1333), written to stdout. */
int spike_factor_log(spike_factor_t f);

/* ------------------------------------------------------------------ */
/* Batches: many independent matrices solved together                  */
/* ------------------------------------------------------------------ */

/* One matrix of a batch (host CSR, same conventions as above). */
typedef struct {
    int n;
    const double *a;
    const int *ia;
    const int *ja;
    int nrhs;                 /* right-hand sides of this matrix, 1..8         */
} spike_batch_matrix_t;

typedef struct {
    int rows_per_partition;   /* Spike partition size; 0 = autotuned           */
    int lanes_per_row;        /* 0 = per matrix from its row lengths           */
    int precision;            /* SPIKE_PREC_*                                  */
    int use_graph;            /* 1 = solves replay a captured hipGraph         */
    int autotune_sample;      /* most matrices the autotuner times per RHS-slot */
                              /* class; 0 = enough to fill the GPU             */
    double drop_tol;          /* as in spike_options_t                         */
    double refine_tol;        /* as in spike_options_t                         */
} spike_batch_options_t;

typedef struct spike_batch_s *spike_batch_t;

typedef struct {
    int count;                /* matrices                                      */
    int groups;               /* sub-factors (RHS slots x lanes per row)       */
    int64_t rows;             /* sum of n                                      */
    int64_t nnz;
    int64_t rows_rhs;         /* sum of n * nrhs                               */
    int64_t partitions;
    int64_t n_reduced;
    int rows_per_partition[4];/* in use at 1, 2, 4, 8 RHS slots (autotuned    */
                              /* per slot class when 0 was given)              */
    int refine_steps;
    double refine_est_err;
    int64_t device_bytes;
    double preproc_seconds;
    double autotune_seconds;
    int graph_nodes;          /* kernel launches per solve                     */
    int lane_shift[4];        /* lanes per row at 1, 2, 4, 8 RHS slots: each   */
                              /* matrix's row-length rule x 2^lane_shift       */
                              /* (autotuned when lanes_per_row was 0)          */
} spike_batch_stats_t;

void spike_batch_options_default(spike_batch_options_t *o);

/* Preprocesses count matrices (all with the same uplo) into one batch held in
 * device memory.  Matrices are grouped by right-hand-side slots (1, 2, 4, 8)
 * and lanes per row; each group is one block-diagonal Spike factor whose
 * partitions never straddle two matrices. */
int spike_batch_create(spike_batch_t *bt, char uplo, int count, const spike_batch_matrix_t *m,
                       const spike_batch_options_t *opts);

/* Packed layout of a batch's right-hand sides and solutions: matrix i's nrhs
 * columns start at element spike_batch_offset(bt, i), column-major with
 * leading dimension n_i; spike_batch_length(bt) elements in total. */
int64_t spike_batch_offset(spike_batch_t bt, int i);
int64_t spike_batch_length(spike_batch_t bt);

/* Solves every matrix of the batch for its right-hand sides (packed layout;
 * host or device pointers).  With use_graph, the launch sequence of every
 * group is captured once per (b, x) pair into a hipGraph, the groups on
 * parallel branches, and replayed on stream. */
int spike_batch_solve(spike_batch_t bt, const double *b, double *x, void *stream);

int spike_batch_stats(spike_batch_t bt, spike_batch_stats_t *st);
int spike_batch_destroy(spike_batch_t bt);

/* ------------------------------------------------------------------ */
/* Serving: resident factors, batched solve requests                   */
/* ------------------------------------------------------------------ */

typedef struct {
    size_t hbm_budget;        /* bytes for resident matrices (factor + staging)*/
                              /* 0 = 80% of the free device memory             */
    int slots;                /* most solves in flight; 0 = 64                 */
    int streams;              /* stream pool; 0 = 8                            */
    double window_us;         /* batching window after the oldest waiting      */
                              /* request; < 0 = 200                            */
    int max_batch_rhs;        /* right-hand sides merged into one solve, and   */
                              /* the most one request may carry; 0 = 16        */
    double admit_timeout_ms;  /* spike_server_add of a matrix whose footprint  */
                              /* does not fit the remaining budget waits this  */
                              /* long for spike_server_remove to free memory;  */
                              /* < 0 = without limit, 0 = refuse at once       */
} spike_server_options_t;

typedef struct {
    int matrices;
    int64_t resident_bytes;
    int64_t completed;        /* requests                                      */
    int64_t solves;           /* merged solves launched                        */
    int64_t windows;
    int64_t merged_rhs;       /* right-hand sides in solves of > 1 column      */
    int max_inflight;
    int64_t rows_rhs;         /* sum of n x nrhs over completed requests       */
    double p50_us, p99_us, max_us;   /* latency submit -> completion           */
    double seconds;           /* first submit -> last completion               */
    double rows_rhs_per_s;
} spike_server_stats_t;

typedef struct spike_server_s *spike_server_t;

void spike_server_options_default(spike_server_options_t *o);
int spike_server_create(spike_server_t *s, const spike_server_options_t *o);
/* Preprocesses a matrix and keeps it resident; *id names it.  Its footprint
 * (factor + staging) is admitted against the budget: when it does not fit
 * what is left, the call waits up to admit_timeout_ms for removals to free
 * memory, then fails with SPIKE_ERR_NOMEM (at once when the footprint alone
 * exceeds the budget).  May be called while requests are being served. */
int spike_server_add(spike_server_t s, char uplo, int n, const double *a, const int *ia, const int *ja,
                     const spike_options_t *opts, int *id);
/* Stops serving matrix id: new submits for it fail with SPIKE_ERR_ARG, the
 * call waits until its queued and in-flight requests have completed, then
 * frees its factor and staging and returns its footprint to the budget
 * (waking adds that wait for memory). */
int spike_server_remove(spike_server_t s, int id);
/* Queues a solve of matrix id for nrhs right-hand sides (device pointers,
 * column-major; host pointers are refused with SPIKE_ERR_ARG); b and x must
 * stay valid until the ticket completes. */
int spike_server_submit(spike_server_t s, int id, int nrhs, const double *b, int ldb, double *x, int ldx,
                        uint64_t *ticket);
/* Blocks until the ticket's solve has completed (x written). */
int spike_server_wait(spike_server_t s, uint64_t ticket);
/* Blocks until every queued request has completed. */
int spike_server_drain(spike_server_t s);
int spike_server_stats(spike_server_t s, spike_server_stats_t *st);
int spike_server_reset_stats(spike_server_t s);
int spike_server_destroy(spike_server_t s);
const char *spike_server_last_error(void);

/* ------------------------------------------------------------------ */
/* MatrixMarket ingest                                                   */
/* ------------------------------------------------------------------ */

/* Replaces matrixMarketToCSR, reference src/misc.c:109-245 (called from
 * test/ex.c and test/runtime_analysis.c): reads a real/integer coordinate
 * MatrixMarket file into zero-based CSR (columns ascending).  uplo 'L' keeps
 * the lower triangle, 'U' the upper, 'N' everything; symm 'Y' mirrors every
 * off-diagonal entry of a symmetric file, 'S'/'N' read entries as stored, 'A'
 * takes 'Y' when the banner says symmetric.  *a, *ia, *ja are allocated by
 * the library; release them with spike_free.  Errors: SPIKE_ERR_ARG (cannot
 * open), SPIKE_ERR_MATRIX (malformed), SPIKE_ERR_UNSUPPORTED (pattern or
 * complex files); spike_mm_last_error() describes the last error. */
int spike_read_matrix_market(const char *path, char uplo, char symm, int *n, double **a, int **ia, int **ja);
void spike_free(void *p);
const char *spike_mm_last_error(void);

/* Host-only symbolic analysis (no device needed): fills st with the sizes the
 * factorisation would have, except fields that need the spikes (nnz_reduced,
 * reduced_levels, dropped_rowsum_max, device_bytes are 0). */
int spike_analyze(char uplo, int n, const double *a, const int *ia, const int *ja,
                  const spike_options_t *opts, spike_stats_t *st);

/* Partition count used by the legacy API when > 0 (overrides nthreads and the
 * SPIKE_PSTRSV_PARTITIONS environment variable); 0 restores the default. */
void spike_pstrsv_set_partitions(int P);

/* Factor behind the legacy API (NULL before the first preprocess). */
spike_factor_t spike_pstrsv_current_factor(void);

/* Last error message of the calling thread ("" if none). */
const char *spike_last_error(void);

/* Per-phase timers of the profiling twins, in seconds:
 * t[0] preprocess total, t[1] symbolic analysis, t[2] device spike build,
 * t[3] reduced-system assembly, t[4] solve total (device),
 * t[5] pass 1 block solve, t[6] reduced solve, t[7] coupling update,
 * t[8] pass 2 block solve.  Returns the number of timers written (<= cap). */
int spike_pstrsv_get_timers(double *t, int cap);

#ifdef __cplusplus
}
#endif
#endif /* SPIKE_PSTRSV_H */
