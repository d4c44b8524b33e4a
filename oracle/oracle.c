/*
 * CPU ORACLE for the spike_pstrsv solve path -- TEST INFRASTRUCTURE ONLY.
 * Used by tests/, the smoke check and the benchmark's CPU baseline as the
 * checker; never linked into or called by the product library.
 * Function-by-function reference citations and the checks that pin each
 * function are listed in oracle.h.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static char g_err[256];
const char *oracle_last_error(void) { return g_err; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }

/* ------------------------------------------------------------------ */
/* Sequential kernels                                                  */
/* ------------------------------------------------------------------ */

/* ops.c:1264-1279 / ra.c:52-65: L x = b, pivot A[IA[k+1]-1]. */
void oracle_ldsol(int n, const double *a, const int *ia, const int *ja,
                  const double *b, double *x)
{
    for (int k = 0; k < n; k++) {
        double t = b[k];
        for (int j = ia[k]; j < ia[k + 1] - 1; j++) t -= a[j] * x[ja[j]];
        x[k] = t / a[ia[k + 1] - 1];
    }
}

/* ops.c:1179-1194 / ra.c:37-50: U x = b, pivot A[IA[k]]. */
void oracle_udsol(int n, const double *a, const int *ia, const int *ja,
                  const double *b, double *x)
{
    for (int k = n - 1; k >= 0; k--) {
        double t = b[k];
        for (int j = ia[k] + 1; j < ia[k + 1]; j++) t -= a[j] * x[ja[j]];
        x[k] = t / a[ia[k]];
    }
}

void oracle_spmv(int n, const double *a, const int *ia, const int *ja,
                 const double *x, double *y)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = ia[i]; j < ia[i + 1]; j++) s += a[j] * x[ja[j]];
        y[i] = s;
    }
}

/* mp.c:18-33 */
void oracle_partition(int n, int P, int *starts)
{
    int k = n / P;
    for (int t = 0; t < P; t++) starts[t] = k * t;
    starts[P] = n;   /* last partition: k + n % P rows */
}

/* ra.c:315-316 */
void oracle_manufactured_x(int n, unsigned seed, double *x)
{
    srand(seed);
    for (int i = 0; i < n; i++)
        x[i] = ((rand() % 71) + 1) + ((double)(i % 1000) / 10.0);
}

void oracle_flip(int n, const double *a, const int *ia, const int *ja,
                 double *a2, int *ia2, int *ja2)
{
    int p = 0;
    ia2[0] = 0;
    for (int i = 0; i < n; i++) {
        int src = n - 1 - i;
        for (int j = ia[src + 1] - 1; j >= ia[src]; j--) {
            a2[p] = a[j];
            ja2[p] = n - 1 - ja[j];
            p++;
        }
        ia2[i + 1] = p;
    }
}

/* ------------------------------------------------------------------ */
/* Spike restatement                                                   */
/* ------------------------------------------------------------------ */

struct oracle_spike {
    char uplo;
    int n, P, nthreads;
    int64_t nnz;
    double *a; int *ia, *ja;     /* private copy of the matrix               */
    int *starts;                 /* P+1                                       */
    int *ridx;                   /* n: reduced index + 1 of a dependent, or 0 */
    int *deplist; int nred;      /* dependents, ascending                    */
    oracle_part_info *info;      /* P                                         */
    int *kstart; int *kcols;     /* K_p lists, CSR over partitions           */
    int64_t *wstart; double *W;  /* dense spike blocks, row-major            */
    int *rrp, *rci; double *rv;  /* reduced system, off-diagonals           */
    double tau, dropmax;
    /* L: spike rows [first_coupling, last_dependent] (local);
       U: spike rows [first dependent, last coupling] (local). */
    int *slo, *shi;
};

static int validate(char uplo, int n, const int *ia, const int *ja, int P)
{
    if (n <= 0) { set_err("n must be positive"); return 0; }
    if (P <= 0 || P > n) { set_err("partition count must be in [1, n]"); return 0; }
    if (ia[0] != 0) { set_err("ia[0] must be 0"); return 0; }
    for (int i = 0; i < n; i++) {
        if (ia[i + 1] <= ia[i]) { set_err("empty row (diagonal missing)"); return 0; }
        for (int j = ia[i] + 1; j < ia[i + 1]; j++)
            if (ja[j] <= ja[j - 1]) { set_err("columns not strictly ascending"); return 0; }
        if (uplo == 'L') {
            if (ja[ia[i + 1] - 1] != i) { set_err("diagonal must be the last entry of each row"); return 0; }
        } else {
            if (ja[ia[i]] != i) { set_err("diagonal must be the first entry of each row"); return 0; }
        }
    }
    return 1;
}

static int find_col(const int *cols, int m, int c)
{
    int lo = 0, hi = m - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (cols[mid] == c) return mid;
        if (cols[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

oracle_spike *oracle_spike_preproc(char uplo, int n, const double *a, const int *ia,
                                   const int *ja, int P, double tau, int nthreads)
{
    if (uplo == 'l') uplo = 'L';
    if (uplo == 'u') uplo = 'U';
    if (uplo != 'L' && uplo != 'U') { set_err("uplo must be L or U"); return NULL; }
    if (!validate(uplo, n, ia, ja, P)) return NULL;

    oracle_spike *s = (oracle_spike *)calloc(1, sizeof *s);
    s->uplo = uplo; s->n = n; s->P = P; s->tau = tau;
#ifdef _OPENMP
    s->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    s->nthreads = 1;
#endif
    s->nnz = ia[n];
    s->a = (double *)malloc(sizeof(double) * s->nnz);
    s->ja = (int *)malloc(sizeof(int) * s->nnz);
    s->ia = (int *)malloc(sizeof(int) * (n + 1));
    memcpy(s->a, a, sizeof(double) * s->nnz);
    memcpy(s->ja, ja, sizeof(int) * s->nnz);
    memcpy(s->ia, ia, sizeof(int) * (n + 1));
    s->starts = (int *)malloc(sizeof(int) * (P + 1));
    oracle_partition(n, P, s->starts);
    s->info = (oracle_part_info *)calloc(P, sizeof(oracle_part_info));
    s->slo = (int *)malloc(sizeof(int) * P);
    s->shi = (int *)malloc(sizeof(int) * P);
    s->ridx = (int *)calloc(n, sizeof(int));

    /* dependents: columns referenced outside their own block
       (ops.c:129-182; U marks col >= end for partitions 0..P-2, L marks
       col < start for every partition). */
    for (int t = 0; t < P; t++) {
        int st = s->starts[t], en = s->starts[t + 1];
        oracle_part_info *inf = &s->info[t];
        inf->start = st; inf->end = en;
        for (int i = st; i < en; i++)
            for (int j = ia[i]; j < ia[i + 1]; j++) {
                int c = ja[j];
                if ((uplo == 'L' && c < st) || (uplo == 'U' && c >= en)) {
                    s->ridx[c] = 1;
                    inf->has_dependence = 1;
                }
            }
    }
    int nred = 0;
    for (int i = 0; i < n; i++) if (s->ridx[i]) nred++;
    s->nred = nred;
    s->deplist = (int *)malloc(sizeof(int) * (nred > 0 ? nred : 1));
    nred = 0;
    for (int i = 0; i < n; i++) if (s->ridx[i]) { s->deplist[nred] = i; s->ridx[i] = ++nred; }

    /* per-partition intervals (ops.c:214-373, L mirrored):
       L: first coupling row f, last dependent row g, spike rows [f, g];
       U: first dependent row (minDependentRow, ops.c:256-264), last coupling
          row (getIntervalEnd, ops.c:184-212), spike rows between them. */
    for (int t = 0; t < P; t++) {
        oracle_part_info *inf = &s->info[t];
        int st = inf->start, en = inf->end, k = en - st;
        int first_cpl = k, last_cpl = -1, first_dep = k, last_dep = -1;
        for (int r = 0; r < k; r++) {
            int i = st + r;
            for (int j = ia[i]; j < ia[i + 1]; j++) {
                int c = ja[j];
                if ((uplo == 'L' && c < st) || (uplo == 'U' && c >= en)) {
                    if (first_cpl == k) first_cpl = r;
                    last_cpl = r;
                }
            }
            if (s->ridx[i]) { if (first_dep == k) first_dep = r; last_dep = r; }
        }
        if (uplo == 'L') {
            inf->first_coupling = first_cpl; inf->last_dependent = last_dep;
            inf->has_reflection = last_dep >= 0;
            s->slo[t] = first_cpl; s->shi[t] = last_dep;
        } else {
            inf->first_coupling = last_cpl;  /* U: last coupling row */
            inf->last_dependent = first_dep; /* U: first dependent row */
            inf->has_reflection = first_dep < k && t > 0;
            s->slo[t] = first_dep; s->shi[t] = last_cpl;
        }
        if (!inf->has_reflection || s->slo[t] > s->shi[t]) { s->slo[t] = 0; s->shi[t] = -1; }
        inf->rows_S = s->shi[t] - s->slo[t] + 1;
    }

    /* K_p: distinct coupling columns of the spike rows (ops.c:266-280) */
    s->kstart = (int *)calloc(P + 1, sizeof(int));
    int *mark = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) mark[i] = -1;
    int64_t ktot = 0;
    for (int t = 0; t < P; t++) {
        oracle_part_info *inf = &s->info[t];
        int cnt = 0;
        for (int r = s->slo[t]; r <= s->shi[t]; r++) {
            int i = inf->start + r;
            for (int j = ia[i]; j < ia[i + 1]; j++) {
                int c = ja[j];
                if (((uplo == 'L' && c < inf->start) || (uplo == 'U' && c >= inf->end)) && mark[c] != t) {
                    mark[c] = t; cnt++;
                }
            }
        }
        inf->cols_S = inf->rows_S > 0 ? cnt : 0;
        s->kstart[t + 1] = s->kstart[t] + inf->cols_S;
        ktot += inf->cols_S;
    }
    s->kcols = (int *)malloc(sizeof(int) * (ktot > 0 ? ktot : 1));
    for (int i = 0; i < n; i++) mark[i] = -1;
    for (int t = 0; t < P; t++) {
        oracle_part_info *inf = &s->info[t];
        if (inf->cols_S == 0) continue;
        int *kc = s->kcols + s->kstart[t]; int cnt = 0;
        for (int r = s->slo[t]; r <= s->shi[t]; r++) {
            int i = inf->start + r;
            for (int j = ia[i]; j < ia[i + 1]; j++) {
                int c = ja[j];
                if (((uplo == 'L' && c < inf->start) || (uplo == 'U' && c >= inf->end)) && mark[c] != t) {
                    mark[c] = t; kc[cnt++] = c;
                }
            }
        }
        /* sort ascending */
        for (int u = 1; u < cnt; u++) { int v = kc[u], w = u - 1; while (w >= 0 && kc[w] > v) { kc[w + 1] = kc[w]; w--; } kc[w + 1] = v; }
    }
    free(mark);

    /* spikes: W_p = D0^-1 C[:,K_p] on the spike rows (ops.c:1021-1069 with
       __udsol2 ops.c:1196-1228; L uses forward substitution) */
    s->wstart = (int64_t *)calloc(P + 1, sizeof(int64_t));
    for (int t = 0; t < P; t++) {
        oracle_part_info *inf = &s->info[t];
        int64_t area = inf->cols_S > 0 ? (int64_t)inf->rows_S * inf->cols_S : 0;
        s->wstart[t + 1] = s->wstart[t] + area;
    }
    s->W = (double *)calloc(s->wstart[P] > 0 ? s->wstart[P] : 1, sizeof(double));
    #pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
    for (int t = 0; t < P; t++) {
        oracle_part_info *inf = &s->info[t];
        int m = inf->cols_S;
        if (m == 0) continue;
        int st = inf->start, lo = s->slo[t], hi = s->shi[t];
        double *W = s->W + s->wstart[t];
        const int *kc = s->kcols + s->kstart[t];
        /* dense coupling block (produceCompressedR, ops.c:888-959) */
        for (int r = lo; r <= hi; r++) {
            int i = st + r;
            for (int j = ia[i]; j < ia[i + 1]; j++) {
                int c = ja[j];
                if ((uplo == 'L' && c < st) || (uplo == 'U' && c >= inf->end))
                    W[(int64_t)(r - lo) * m + find_col(kc, m, c)] = a[j];
            }
        }
        if (uplo == 'L') {
            for (int r = lo; r <= hi; r++) {
                int i = st + r;
                double *wr = W + (int64_t)(r - lo) * m;
                for (int j = ia[i]; j < ia[i + 1] - 1; j++) {
                    int c = ja[j];
                    if (c < st + lo) continue;           /* coupling, or spike row known 0 */
                    const double *wc = W + (int64_t)(c - st - lo) * m;
                    double av = a[j];
                    for (int l = 0; l < m; l++) wr[l] -= av * wc[l];
                }
                double d = a[ia[i + 1] - 1];
                for (int l = 0; l < m; l++) wr[l] /= d;
            }
        } else {
            for (int r = hi; r >= lo; r--) {
                int i = st + r;
                double *wr = W + (int64_t)(r - lo) * m;
                for (int j = ia[i] + 1; j < ia[i + 1]; j++) {
                    int c = ja[j];
                    if (c > st + hi) continue;           /* coupling, or spike row known 0 */
                    const double *wc = W + (int64_t)(c - st - lo) * m;
                    double av = a[j];
                    for (int l = 0; l < m; l++) wr[l] -= av * wc[l];
                }
                double d = a[ia[i]];
                for (int l = 0; l < m; l++) wr[l] /= d;
            }
        }
    }

    /* reduced system over the dependents (ops.c:1071-1150): row of dependent
       r holds the spike row of r on the reduced indices of K_p, keeping
       entries with |w| > tau (tau = 0 keeps != 0 as ops.c:1127) */
    s->rrp = (int *)calloc(s->nred + 1, sizeof(int));
    int64_t cap = 0;
    for (int q = 0; q < s->nred; q++) {
        int i = s->deplist[q];
        int t = 0;
        { int lo = 0, hi = P - 1; while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (s->starts[mid] <= i) lo = mid; else hi = mid - 1; } t = lo; }
        oracle_part_info *inf = &s->info[t];
        int r = i - inf->start;
        if (inf->cols_S > 0 && r >= s->slo[t] && r <= s->shi[t]) cap += inf->cols_S;
    }
    s->rci = (int *)malloc(sizeof(int) * (cap > 0 ? cap : 1));
    s->rv = (double *)malloc(sizeof(double) * (cap > 0 ? cap : 1));
    int64_t pos = 0; double dropmax = 0.0;
    for (int q = 0; q < s->nred; q++) {
        int i = s->deplist[q];
        int t;
        { int lo = 0, hi = P - 1; while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (s->starts[mid] <= i) lo = mid; else hi = mid - 1; } t = lo; }
        oracle_part_info *inf = &s->info[t];
        int r = i - inf->start;
        double dropped = 0.0;
        if (inf->cols_S > 0 && r >= s->slo[t] && r <= s->shi[t]) {
            const double *wr = s->W + s->wstart[t] + (int64_t)(r - s->slo[t]) * inf->cols_S;
            const int *kc = s->kcols + s->kstart[t];
            for (int l = 0; l < inf->cols_S; l++) {
                double w = wr[l];
                if (fabs(w) > tau || (tau == 0.0 && w != 0.0)) {
                    s->rci[pos] = s->ridx[kc[l]] - 1; s->rv[pos] = w; pos++;
                } else dropped += fabs(w);
            }
        }
        if (dropped > dropmax) dropmax = dropped;
        s->rrp[q + 1] = (int)pos;
    }
    s->dropmax = dropmax;
    return s;
}

void oracle_spike_solve(oracle_spike *s, const double *b, double *x)
{
    const int P = s->P;
    const double *a = s->a; const int *ia = s->ia, *ja = s->ja;
    double *xr = (double *)malloc(sizeof(double) * (s->nred > 0 ? s->nred : 1));

    if (s->uplo == 'L') {
        /* 1. y_p = D_p^-1 b_p on rows [0, max(g, f-1)] (coupling ignored);
              rows before the first coupling row are final */
        #pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
        for (int t = 0; t < P; t++) {
            oracle_part_info *inf = &s->info[t];
            int st = inf->start, k = inf->end - st;
            int hi = inf->last_dependent > inf->first_coupling - 1 ? inf->last_dependent : inf->first_coupling - 1;
            if (hi > k - 1) hi = k - 1;
            for (int r = 0; r <= hi; r++) {
                int i = st + r;
                double v = b[i];
                for (int j = ia[i]; j < ia[i + 1] - 1; j++) if (ja[j] >= st) v -= a[j] * x[ja[j]];
                x[i] = v / a[ia[i + 1] - 1];
            }
        }
        /* 2. gather y at the dependents (ops.c:1152-1164) */
        for (int q = 0; q < s->nred; q++) xr[q] = x[s->deplist[q]];
        /* 3. unit lower reduced solve (L mirror of __usol2, ops.c:1247-1262) */
        for (int q = 0; q < s->nred; q++) {
            double v = 0.0;
            for (int j = s->rrp[q]; j < s->rrp[q + 1]; j++) v += s->rv[j] * xr[s->rci[j]];
            xr[q] -= v;
        }
        /* 4. scatter (updateX, ops.c:1166-1177) */
        for (int q = 0; q < s->nred; q++) x[s->deplist[q]] = xr[q];
        /* 5. b_p -= C_p x_K and re-solve rows [f, k) (api.c:213-229 mirrored);
              coupling reads the reduced solution, never another partition's x */
        #pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
        for (int t = 0; t < P; t++) {
            oracle_part_info *inf = &s->info[t];
            int st = inf->start, k = inf->end - st;
            for (int r = inf->first_coupling; r < k; r++) {
                int i = st + r;
                double v = b[i];
                for (int j = ia[i]; j < ia[i + 1] - 1; j++) {
                    int c = ja[j];
                    v -= a[j] * (c < st ? xr[s->ridx[c] - 1] : x[c]);
                }
                x[i] = v / a[ia[i + 1] - 1];
            }
        }
    } else {
        /* 1. D1 solve (api.c:193-199): rows [min(si, ei), k) by back
              substitution, R ignored; rows after the last coupling row are final */
        #pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
        for (int t = 0; t < P; t++) {
            oracle_part_info *inf = &s->info[t];
            int st = inf->start, en = inf->end, k = en - st;
            int lo = inf->last_dependent < inf->first_coupling + 1 ? inf->last_dependent : inf->first_coupling + 1;
            if (!inf->has_reflection) lo = inf->first_coupling + 1;
            if (lo < 0) lo = 0;
            for (int r = k - 1; r >= lo; r--) {
                int i = st + r;
                double v = b[i];
                for (int j = ia[i] + 1; j < ia[i + 1]; j++) if (ja[j] < en) v -= a[j] * x[ja[j]];
                x[i] = v / a[ia[i]];
            }
        }
        /* 2. gather (ops.c:1152-1164) */
        for (int q = 0; q < s->nred; q++) xr[q] = x[s->deplist[q]];
        /* 3. unit upper reduced solve, backward (__usol2, ops.c:1247-1262) */
        for (int q = s->nred - 1; q >= 0; q--) {
            double v = 0.0;
            for (int j = s->rrp[q]; j < s->rrp[q + 1]; j++) v += s->rv[j] * xr[s->rci[j]];
            xr[q] -= v;
        }
        /* 4. scatter (ops.c:1166-1177) */
        for (int q = 0; q < s->nred; q++) x[s->deplist[q]] = xr[q];
        /* 5. b -= R_new x, then D2 solve on rows [0, ei) (api.c:212-226) */
        #pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
        for (int t = 0; t < P; t++) {
            oracle_part_info *inf = &s->info[t];
            int st = inf->start, en = inf->end;
            for (int r = inf->first_coupling; r >= 0; r--) {
                int i = st + r;
                double v = b[i];
                for (int j = ia[i] + 1; j < ia[i + 1]; j++) {
                    int c = ja[j];
                    v -= a[j] * (c >= en ? xr[s->ridx[c] - 1] : x[c]);
                }
                x[i] = v / a[ia[i]];
            }
        }
    }
    free(xr);
}

void oracle_spike_stats(const oracle_spike *s, oracle_stats_t *st)
{
    st->n = s->n; st->P = s->P; st->nnz = s->nnz;
    st->n_reduced = s->nred;
    st->nnz_reduced = (int64_t)s->rrp[s->nred] + s->nred;
    st->spike_area = s->wstart[s->P];
    st->dropped_rowsum_max = s->dropmax;
}

int oracle_spike_part_info(const oracle_spike *s, int p, oracle_part_info *info)
{
    if (p < 0 || p >= s->P) return -1;
    *info = s->info[p];
    return 0;
}

void oracle_spike_dependents(const oracle_spike *s, int *dep)
{
    memcpy(dep, s->deplist, sizeof(int) * s->nred);
}

int oracle_spike_kcols(const oracle_spike *s, int p, int *cols)
{
    if (p < 0 || p >= s->P) return -1;
    int m = s->kstart[p + 1] - s->kstart[p];
    memcpy(cols, s->kcols + s->kstart[p], sizeof(int) * m);
    return m;
}

int oracle_spike_block(const oracle_spike *s, int p, double *w)
{
    if (p < 0 || p >= s->P) return -1;
    int64_t m = s->wstart[p + 1] - s->wstart[p];
    memcpy(w, s->W + s->wstart[p], sizeof(double) * m);
    return (int)m;
}

void oracle_spike_reduced(const oracle_spike *s, int *rp, int *ci, double *v)
{
    memcpy(rp, s->rrp, sizeof(int) * (s->nred + 1));
    memcpy(ci, s->rci, sizeof(int) * s->rrp[s->nred]);
    memcpy(v, s->rv, sizeof(double) * s->rrp[s->nred]);
}

void oracle_spike_free(oracle_spike *s)
{
    if (!s) return;
    free(s->a); free(s->ia); free(s->ja); free(s->starts); free(s->ridx);
    free(s->deplist); free(s->info); free(s->kstart); free(s->kcols);
    free(s->wstart); free(s->W); free(s->rrp); free(s->rci); free(s->rv);
    free(s->slo); free(s->shi); free(s);
}
