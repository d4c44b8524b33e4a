// Host-side symbolic analysis for the Spike solve (no device code).
//
// Everything is expressed for a LOWER triangular matrix: an upper triangular
// input is converted once by index reversal U = J L J (row i of L is row n-1-i
// of U, columns c -> n-1-c, entry order reversed), and the device kernels map
// b/x indices back on load/store.  The Spike quantities follow the reference's
// construction for 'U' (src/spike_pstrsv_ops.c:129-373, :1071-1150) mirrored
// to 'L':
//   partition p      rows [s_p, s_p + k_p)  (src/matrix_partitioning.c:18-33)
//   coupling C_p     entries of partition p with column < s_p
//   dependents       columns referenced by any coupling entry (ops.c:162-181)
//   f_p              first local row holding a coupling entry (k_p if none)
//   g_p              last local row that is a dependent (-1 if none)
//   K_p              distinct coupling columns of rows [f_p, g_p]
//   spike W_p        D_p^-1 C_p[:, K_p] restricted to rows [f_p, g_p]
#pragma once

#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace spike {

// Host threads for preprocessing: the hardware threads, at most 32.
inline int host_threads()
{
    const int nt = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(nt, 32));
}

// fn(i) for every i in [0, count), indices taken from a shared counter by up
// to host_threads() threads.
template <class F>
void parallel_for(int count, F &&fn)
{
    int nt = host_threads();
    if (count < 2 * nt) nt = std::max(1, std::min(nt, count));
    if (nt == 1) {
        for (int i = 0; i < count; i++) fn(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&]() {
            for (;;) {
                int i = next.fetch_add(1);
                if (i >= count) break;
                fn(i);
            }
        });
    for (auto &x : th) x.join();
}

// Chunks parallel_chunks(n, grain, ...) cuts [0, n) into.
inline int chunk_count(int64_t n, int64_t grain)
{
    if (n <= 0) return 0;
    const int64_t g = std::max<int64_t>(1, grain);
    return (int)std::min<int64_t>((n + g - 1) / g, 1 << 20);
}

// fn(c, lo, hi) over [0, n) cut into chunk_count(n, grain) chunks of about
// grain indices (c is the chunk number, in index order); returns the count.
template <class F>
int parallel_chunks(int64_t n, int64_t grain, F &&fn)
{
    const int nc = chunk_count(n, grain);
    if (nc == 0) return 0;
    const int64_t step = (n + nc - 1) / nc;
    parallel_for(nc, [&](int c) {
        const int64_t lo = (int64_t)c * step, hi = std::min(n, lo + step);
        if (lo < hi) fn(c, lo, hi);
    });
    return nc;
}

// v resized to n value-initialised elements; an empty v of at least 4 MiB is
// first reserved and advised for transparent huge pages, which cuts the
// first-touch page faults of large host arrays (a no-op where the kernel
// offers none).
template <class T>
void resize_large(std::vector<T> &v, size_t n)
{
    if (v.empty() && n * sizeof(T) >= ((size_t)4 << 20)) {
        v.reserve(n);
        const uintptr_t b = ((uintptr_t)v.data() + 4095) & ~(uintptr_t)4095;
        const uintptr_t e = ((uintptr_t)v.data() + n * sizeof(T)) & ~(uintptr_t)4095;
        if (e > b) (void)madvise((void *)b, e - b, MADV_HUGEPAGE);
    }
    v.resize(n);
}

// Block solves laid out as fixed-shape "steps" executed by one 64-lane
// wavefront per partition; both solve passes run the same steps.  A step
// holds R = 64/G row slots; slot t is served by lanes [t*G, (t+1)*G) and each
// lane holds M entries ("rounds").
// Within a partition, steps run in order and every row (chunk) in a step only
// depends on rows finished in earlier steps.  Every value is pre-scaled by
// its row's 1/diag and the right-hand side is scaled on load, so a step slot
// computes x[row] -= sum val * x[col]; rows longer than G*M entries are split
// into chunks that subtract their partial sums in turn.
//
// Forwarding layout (fwd = true, G = 16, R = 4): every value is pre-scaled by
// the row's 1/diag (the right-hand side is scaled on load), so a final chunk
// computes x = acc - sum(val * x_col) with no division.  Entries whose column
// was produced by the immediately preceding step ("late" entries) are not in
// the lane layout; they are a dense R x R block late[step][slot][prev_slot]
// applied to the previous step's results forwarded through registers.  Rows
// of empty slots are k+1 (a scratch LDS slot).  Consecutive chunks of a long
// row are at least two steps apart.
struct StepStream {
    int G = 16, M = 1, R = 4;
    bool fwd = false;
    int64_t nsteps = 0;
    int max_steps = 0;                  // longest per-partition step chain
    std::vector<int64_t> part_begin;    // P+1 offsets into the step arrays
    std::vector<double> val;            // nsteps * M * 64  [step][round][lane]
    std::vector<uint16_t> col;          // nsteps * M * 64  local column, pad = k_p
    std::vector<uint16_t> row;          // nsteps * R       local row, k+1 = empty
    std::vector<double> late;           // nsteps * R * R   fwd only
};

struct Part {
    int s = 0, k = 0;       // rows [s, s+k)
    int f = 0;              // first coupling row (local), k if none
    int g = -1;             // last dependent row (local), -1 if none
    int hi1 = -1;           // pass 1 needs rows [0, hi1] (y at the dependents)
    int kbeg = 0, kcnt = 0; // K_p = kcols[kbeg .. kbeg+kcnt)
    int rows_S = 0;         // spike rows [f, g] when kcnt > 0
    int dbeg = 0, dcnt = 0; // dependents of this partition: dep_rows[dbeg..)
    int cbeg = 0, ccnt = 0; // coupling rows of this partition: cpl_rows[cbeg..)
};

struct Plan {
    char uplo = 'L';
    bool flip = false;          // input was 'U'
    int n = 0, P = 0;
    int64_t nnz = 0;
    int kmax = 0;               // largest partition
    // lower-triangular CSR (after the flip for 'U'); diagonal last
    std::vector<int> ia, ja;
    std::vector<double> a;
    std::vector<int> starts;    // P+1
    std::vector<int> ridx;      // n: reduced index, -1 if not a dependent
    std::vector<int> dep_rows;  // dependents in ascending order (global rows)
    std::vector<Part> parts;
    std::vector<int> kcols;     // concatenated K_p, global columns ascending
    // coupling rows for the pass-2 right-hand-side update b' = b - C x_K
    std::vector<int> cpl_rows;  // global rows, grouped by partition, ascending
    std::vector<int> cpl_ptr;   // CSR over cpl_rows
    std::vector<int> cpl_col;   // reduced index of the coupling column
    std::vector<double> cpl_val;
    StepStream steps;           // block solve of every partition, rows [0, k)
    std::vector<double> inv_diag;  // n: 1/diag of each row (lower form)
    int64_t spike_area = 0;
};

struct PlanOptions {
    int P = 1;
    int lanes_per_row = 0;      // 0 = automatic
    int kmax_limit = 16384;     // largest partition a workgroup can hold in LDS
    bool forward = false;       // use the forwarding layout when G == 16
    // FP32 step values: automatic G covers whole rows (choose_lanes), since
    // every chunk rounds the row's partial sum to FP32 once more (measured on
    // a 128k-row banded matrix: FP32 solve error 1.8e-7 at G = 16, 3.7e-7 at
    // G = 1; on the 4096-matrix batch the smaller G needs a second
    // refinement step)
    bool fp32 = false;
    // caller partition boundaries (P+1 values, 0 = starts[0] < ... <
    // starts[P] = n); empty: the reference's formula (k = n/P, remainder last)
    std::vector<int> starts;
};

// Builds the plan.  Returns an empty string on success, else an error message.
// The input CSR is validated (ia[0] == 0, no empty rows, strictly ascending
// columns, diagonal last for 'L' / first for 'U', n >= 1, 1 <= P <= n).
std::string build_plan(char uplo, int n, const double *a, const int *ia, const int *ja,
                       const PlanOptions &opt, Plan &plan);

// Spike build layout.  Every spike row (rows f .. g of a partition with a
// spike, numbered from slot0[t]) gets rinfo = (W slot, far slot, coupling
// entries, 4-entry chunks of its other off-diagonal entries): the W slot
// numbers the rows that are stored in W, the far slot the LDS row that keeps
// the row for a row R or more further on (ring build; live far rows never
// share a slot; when those LDS rows do not fit, gfar: the far-read rows are
// stored in W and read from there, far slot -1).  pia: offsets of the rows in the ring build's staged entry
// layout (a row: its reciprocal diagonal, its coupling entries, padding to a
// multiple of four, then its other entries in chunks of four).  Ring build: only the dependents are stored, R ring
// rows (a power of two) and the far rows fit lds_limit with sb spike rows
// per stage of at most cap entries.  Direct build (ring = false): every
// spike row is stored (W slot = its row).
struct SpikeLayout {
    bool ring = false, gfar = false;   // gfar: far rows read from W, not LDS
    int R = 0, nlr = 0, kstride = 0, sb = 0, cap = 0, block_threads = 64;
    int max_kcnt = 0, max_rows = 0, nsp = 0;
    std::vector<int> slot0;     // P, -1 for a partition without a spike
    std::vector<int> rinfo;     // 4 per spike row
    std::vector<int> pia;       // spike rows + 1
    std::vector<int64_t> woff;  // P, -1 for a partition without a spike
    int64_t area = 0;           // W entries
};
// lds(nlr, kstride, sb, cap): LDS bytes of the ring build; stage_entries:
// most staged entries per stage for block_threads threads.
SpikeLayout plan_spike_layout(const Plan &pl, size_t lds_limit, size_t (*lds)(int, int, int, int),
                              int (*stage_entries)(int), bool allow_gfar);

// Lanes per row of the level kernel: the smallest power of two (<= 16) that
// covers the 90th percentile of the row lengths.
int choose_lanes(const std::vector<int> &row_len);
// Lanes per row slot of the block solve, whose rows are split into chunks of
// G entries scheduled as separate steps: the largest power of two (<= 16)
// with 2G <= p90 / 3, i.e. rows of about three to five chunks.  Measured on
// gfx950 (one RHS): 16 entries per row runs fastest at G = 4 (153 us per
// solve on a 1M-row banded matrix, against 157 at G = 8 and 165 at G = 16);
// two- and three-entry rows (2D / 3D Laplacians) at G = 1.
int choose_block_lanes(const std::vector<int> &row_len);

// Level (longest dependency chain, 0-based) of each row of a lower triangular
// CSR given as (rp, ci) over rows [0, m); entries must point to earlier rows.
std::vector<int> row_levels(int m, const int *rp, const int *ci);

}  // namespace spike
