/*
 * Deterministic synthetic inputs for the sparse triangular solve workloads.
 *
 * Every matrix is a zero-based 3-array CSR lower triangle (uplo 'L') with
 * ascending column indices and the diagonal as the LAST entry of each row,
 * i.e. the layout the reference's lower-triangular solver expects
 * (src/spike_pstrsv_ops.c:1264-1279, test/runtime_analysis.c:52-65: the pivot
 * of row k is A[IA[k+1]-1]).  Generation is counter-based (SplitMix64, Steele,
 * Lea & Flood 2014): each row draws from its own stream keyed by (seed, row),
 * so rows can be generated in parallel and the result never depends on the
 * thread count.
 *
 * Workloads:
 *   - 2D 5-point Laplacian lower triangle on a g x g grid (diag 4, off -1);
 *   - 3D 7-point Laplacian lower triangle on a g^3 grid (diag 6, off -1);
 *   - random banded lower triangle: row i holds min(q, i-lo) distinct columns
 *     drawn uniformly from [lo, i), lo = max(0, i-bw), plus, with probability
 *     plr, one long-range column uniform in [0, i) (redrawn on collision);
 *     off-diagonals U(-1,1)/16, diagonal 1 + sum |off-diagonals|;
 *   - right-hand sides with entries U(-1,1).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define GEN_API __attribute__((visibility("default")))

static inline uint64_t sm_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s; } sm_rng;

static inline uint64_t sm_next(sm_rng *r)
{
    r->s += 0x9E3779B97F4A7C15ULL;
    return sm_mix(r->s);
}

/* Stream for (seed, stream id, index): independent of evaluation order. */
static inline sm_rng sm_stream(uint64_t seed, uint64_t stream, uint64_t idx)
{
    sm_rng r;
    r.s = sm_mix(seed ^ sm_mix(stream * 0xD1B54A32D192ED03ULL + idx + 1));
    return r;
}

/* uniform double in [0,1) with 53 random bits */
static inline double sm_u01(sm_rng *r) { return (double)(sm_next(r) >> 11) * 0x1.0p-53; }
/* uniform double in [-1,1) */
static inline double sm_u11(sm_rng *r) { return 2.0 * sm_u01(r) - 1.0; }
/* uniform integer in [0, m) for m >= 1 (multiply-shift, bias < 2^-40) */
static inline int64_t sm_below(sm_rng *r, int64_t m)
{
    return (int64_t)(((unsigned __int128)sm_next(r) * (uint64_t)m) >> 64);
}

enum { STREAM_BAND = 1, STREAM_RHS = 2, STREAM_PARAM = 3 };

/* ------------------------------------------------------------------ */
/* Laplacian lower triangles                                           */
/* ------------------------------------------------------------------ */

GEN_API int64_t spike_gen_lap2d_nnz(int g)
{
    int64_t n = (int64_t)g * g;
    return n + 2 * (int64_t)g * (g - 1);
}

/* Fills ia[n+1], ja[nnz], a[nnz]; returns nnz. */
GEN_API int64_t spike_gen_lap2d(int g, int *ia, int *ja, double *a)
{
    int64_t n = (int64_t)g * g;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t gy = i / g, gx = i % g;
        /* entries of rows 0..i-1: one diagonal each, plus one per row with
           gy>0 and one per row with gx>0 */
        int64_t p = i + (i > g ? i - g : 0) + gy * (g - 1) + (gx > 0 ? gx - 1 : 0);
        ia[i] = (int)p;
        if (gy > 0) { ja[p] = (int)(i - g); a[p] = -1.0; p++; }
        if (gx > 0) { ja[p] = (int)(i - 1); a[p] = -1.0; p++; }
        ja[p] = (int)i; a[p] = 4.0;
    }
    int64_t nnz = spike_gen_lap2d_nnz(g);
    ia[n] = (int)nnz;
    return nnz;
}

GEN_API int64_t spike_gen_lap3d_nnz(int g)
{
    int64_t g2 = (int64_t)g * g, n = g2 * g;
    return n + 3 * g2 * (g - 1);
}

GEN_API int64_t spike_gen_lap3d(int g, int *ia, int *ja, double *a)
{
    int64_t g2 = (int64_t)g * g, n = g2 * g;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t z = i / g2, rem = i % g2, y = rem / g, x = rem % g;
        /* entries of rows 0..i-1: one diagonal each, plus one per row with
           z>0, y>0 and x>0 respectively */
        int64_t zpos = i > g2 ? i - g2 : 0;
        int64_t ypos = z * (g2 - g) + (rem > g ? rem - g : 0);
        int64_t xpos = (z * g + y) * (g - 1) + (x > 0 ? x - 1 : 0);
        int64_t p = i + zpos + ypos + xpos;
        ia[i] = (int)p;
        if (z > 0) { ja[p] = (int)(i - g2); a[p] = -1.0; p++; }
        if (y > 0) { ja[p] = (int)(i - g); a[p] = -1.0; p++; }
        if (x > 0) { ja[p] = (int)(i - 1); a[p] = -1.0; p++; }
        ja[p] = (int)i; a[p] = 6.0;
    }
    int64_t nnz = spike_gen_lap3d_nnz(g);
    ia[n] = (int)nnz;
    return nnz;
}

/* ------------------------------------------------------------------ */
/* Random banded lower triangle with long-range entries                */
/* ------------------------------------------------------------------ */

#define BAND_MAX_BW 4096

/* Draws the off-diagonal columns of row i into cols (ascending), returns count. */
static int band_row_cols(uint64_t seed, int64_t i, int bw, int q, double plr,
                         int64_t *cols, sm_rng *rng_out)
{
    sm_rng r = sm_stream(seed, STREAM_BAND, (uint64_t)i);
    int64_t lo = i - bw; if (lo < 0) lo = 0;
    int w = (int)(i - lo);
    int m = q < w ? q : w;
    /* partial Fisher-Yates over the window offsets 0..w-1 */
    int idx[BAND_MAX_BW];
    for (int t = 0; t < w; t++) idx[t] = t;
    for (int t = 0; t < m; t++) {
        int j = t + (int)sm_below(&r, w - t);
        int tmp = idx[t]; idx[t] = idx[j]; idx[j] = tmp;
    }
    int c = 0;
    for (int t = 0; t < m; t++) cols[c++] = lo + idx[t];
    /* long-range entry */
    if (i >= 1 && sm_u01(&r) < plr && (int64_t)m < i) {
        for (;;) {
            int64_t cc = sm_below(&r, i);
            int hit = 0;
            for (int t = 0; t < m; t++) if (cols[t] == cc) { hit = 1; break; }
            if (!hit) { cols[c++] = cc; break; }
        }
    }
    /* insertion sort (c <= q+1) */
    for (int s = 1; s < c; s++) {
        int64_t v = cols[s]; int t = s - 1;
        while (t >= 0 && cols[t] > v) { cols[t + 1] = cols[t]; t--; }
        cols[t + 1] = v;
    }
    *rng_out = r;
    return c;
}

/* Row lengths (including diagonal) into rowlen[n]; returns nnz. */
GEN_API int64_t spike_gen_band_count(int64_t n, int bw, int q, double plr, uint64_t seed, int *rowlen)
{
    if (bw > BAND_MAX_BW || bw < 1 || q < 0) return -1;
    int64_t total = 0;
    #pragma omp parallel for schedule(static) reduction(+:total)
    for (int64_t i = 0; i < n; i++) {
        int64_t cols[BAND_MAX_BW + 1]; sm_rng r;
        int c = band_row_cols(seed, i, bw, q, plr, cols, &r);
        rowlen[i] = c + 1;
        total += c + 1;
    }
    return total;
}

/* Fills ia[n+1], ja, a given rowlen from spike_gen_band_count. */
GEN_API int64_t spike_gen_band_fill(int64_t n, int bw, int q, double plr, uint64_t seed,
                                    const int *rowlen, int *ia, int *ja, double *a)
{
    int64_t acc = 0;
    for (int64_t i = 0; i < n; i++) { ia[i] = (int)acc; acc += rowlen[i]; }
    ia[n] = (int)acc;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t cols[BAND_MAX_BW + 1]; sm_rng r;
        int c = band_row_cols(seed, i, bw, q, plr, cols, &r);
        int64_t p = ia[i];
        double s = 0.0;
        for (int t = 0; t < c; t++) {
            double v = sm_u11(&r) / 16.0;
            ja[p + t] = (int)cols[t];
            a[p + t] = v;
            s += fabs(v);
        }
        ja[p + c] = (int)i;
        a[p + c] = 1.0 + s;
    }
    return acc;
}

/* ------------------------------------------------------------------ */
/* Right-hand sides and workload parameters                            */
/* ------------------------------------------------------------------ */

/* b[j*ld + i] = U(-1,1) for i < n, j < nrhs (column-major). */
GEN_API void spike_gen_rhs(int64_t n, int nrhs, int64_t ld, uint64_t seed, double *b)
{
    for (int j = 0; j < nrhs; j++) {
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            sm_rng r = sm_stream(seed, STREAM_RHS + 16 * (uint64_t)j, (uint64_t)i);
            b[(int64_t)j * ld + i] = sm_u11(&r);
        }
    }
}

/* Uniform draws in [0,1) from the parameter stream (for per-matrix parameters
 * of batched / serving workloads): out[t] for t < count. */
GEN_API void spike_gen_params(uint64_t seed, uint64_t stream, int64_t count, double *out)
{
    for (int64_t t = 0; t < count; t++) {
        sm_rng r = sm_stream(seed, STREAM_PARAM + 16 * stream, (uint64_t)t);
        out[t] = sm_u01(&r);
    }
}
