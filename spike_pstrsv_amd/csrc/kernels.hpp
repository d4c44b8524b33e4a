// Launch wrappers of the gfx950 kernels (implemented in kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace spike {

// LDS slots of a block-solve x segment after the partition's k rows: slot k
// holds 0 (target of padding entries), slot k+1 receives writes of empty row
// slots, slots k+2 .. k+65 receive the writes of non-leader lanes.
constexpr int XS_PAD = 66;

// Per-partition record read by the block-solve kernels.
struct DevPart {
    int s, k;       // rows [s, s+k) of the lower-triangular form
    int f;          // first coupling row (local), k if none
    int hi1;        // < 0: the partition needs no pass 1
    int dbeg, dcnt; // dependents (reduced indices dbeg .. dbeg+dcnt)
    int cbeg, ccnt; // coupling rows (indices into the coupling arrays)
    // Right-hand-side addressing of the matrix the partition belongs to (a
    // factor may hold many matrices as one block-diagonal system).  Row g of
    // the lower form is entry l = g - row0 of that matrix, stored at caller
    // index (mflip ? mn-1-l : l); its right-hand sides and solutions start at
    // b + boff / x + boff with leading dimension ld and mnr columns (ld == 0 /
    // mnr == 0: the launch's ldb, ldx and nr).
    int row0, mn, mflip, ld, mnr;
    int cpe0, cpe1; // coupling entries of the partition (cpl_ptr[cbeg], cpl_ptr[cbeg+ccnt])
    int pad_;
    int64_t boff;
};

// Device view of a StepStream.
struct DevSteps {
    const double *val;
    const uint16_t *col;
    const uint16_t *row;
    const double *late;          // fwd: R x R block per step
    const float *val32;          // FP32 factors: values
    const int64_t *part_begin;   // P+1
    int G, M, R;
    int fwd;
};

struct BlockArgs {
    const DevPart *parts;
    DevSteps st;
    const double *b;        // right-hand side (caller's index order)
    double *x;              // solution (caller's index order)
    double *xred;           // reduced vector (pass 1 writes y at dependents)
    const double *bc;       // pass 2: b - C x_K at coupling rows
    const int *dep_rows;    // global (lower-form) rows of the dependents
    const int *cpl_rows;    // global (lower-form) rows of the coupling rows
    int n;
    int flip;               // 1: caller's index = n-1-row
    int kmax;               // largest partition (LDS x segment = kmax+1 doubles)
    int ring;               // step records kept in flight per wavefront (8 or 16)
    const double *inv_diag; // fwd: 1/diag per row (lower form)
    // pass 2 computes b - C x_K itself when cpl_fused (the FP64 coupling
    // products of a partition fit an LDS scratch of cmax doubles per RHS after
    // the x segment); otherwise bc is used
    int cpl_fused;
    int cmax;
    const int *cpl_ptr;
    const int *cpl_col;
    const double *cpl_val;
    const double *xr;       // pass 2: reduced solution
    // right-hand sides of this launch: nr active (<= nr_slots) columns with
    // leading dimensions ldb / ldx; reduced vectors (xred, xr, bc) hold
    // nr_slots interleaved values per index
    int nr, nr_slots;
    int ldb, ldx;
    // FP32 factors: step values and multipliers are FP32 and the block solve
    // accumulates in FP32 (b, xred, xr, x and the coupling update stay FP64)
    int fp32;
    // 1: the solve's result is added to x (refinement update x += d)
    int accum;
    // right-hand sides per block-solve wave (divides nr_slots): a launch runs
    // nr_slots / tile waves per partition
    int tile;
};

// Pass 1 (y = D^-1 b, written at the dependents; also the solution of
// partitions without coupling) or pass 2 (re-solve with the coupling-corrected
// right-hand side), one wavefront per partition.
hipError_t launch_block_solve(int pass, int P, const BlockArgs &a, hipStream_t s);

// One launch for a whole Spike solve (pass 1, S Jacobi sweeps of the reduced
// system, pass 2), as a persistent grid that takes work items from an
// in-order queue: first the pass-1 items (one per partition and RHS tile),
// then for every sweep the sweep items (contiguous reduced rows of about
// equal entry counts), then the pass-2 items.  An item only waits for
// values that items taken before it publish, so the grid makes progress
// whatever number of its workgroups is resident.  Published values carry
// their readiness: X holds two buffers (solve parity) of S+1 vectors of
// nred x nr_slots values, x^(m)[q*nr_slots + r] (x^(0) = y at the
// dependents); the buffer of the current parity starts filled with a
// not-ready pattern, every value is published by one agent-scope store and
// read by polling, and each solve refills the other parity's buffer for the
// next.  Waits are bounded (spin_limit polls): an item that gives up publishes
// NaN, raises *abort_word and sets *fail_word (host memory).  The queue
// counters, the parity and the exit count live in ctr[0], ctr[32], ctr[64]
// (separate cache lines); the last workgroup to leave resets the queue and
// flips the parity, so the launch can be replayed from a graph.
struct FusedArgs {
    double *X;
    unsigned *ctr;
    const int *chunks;      // int4 per sweep item: first row, end row, first entry, end entry
    int nchunks;
    int nred, S;
    const int *rp, *ci;     // reduced system (unit lower, diagonal implicit): CSR
    const double *v;        // over reduced rows, ci = reduced column
    int spin_limit;
    int *abort_word, *fail_word;
};
// Bytes of dynamic LDS the fused solve needs for these arguments, or 0 when
// it cannot run (LDS or occupancy).
size_t spike_fused_lds(const BlockArgs &a);
hipError_t launch_spike_fused(int P, const BlockArgs &a, const FusedArgs &f, int ncu, hipStream_t s);
// X[0 .. len) = not-ready pattern and the counters (96 words) cleared; run
// once when a fused workspace is allocated, never inside a captured graph.
hipError_t launch_fused_reset(double *X, size_t len, unsigned *ctr, int *abort_word, hipStream_t s);
// Sweep items rows per item at most (4 per lane).
constexpr int kFusedRowsPerItem = 256;

// Single-workgroup forward solve of the unit lower reduced system, in place on
// xred.  rlist holds the rows with off-diagonals, grouped by level;
// levptr[nlev+1] (device) delimits the levels.
hipError_t launch_reduced_solve(int nred, int nlev, const int *levptr, const int *rlist,
                                const int *rp, const int *ci, const double *v,
                                double *xred, bool in_lds, int nr_slots, hipStream_t s);

// One Jacobi sweep xout = y - W xin of the unit lower reduced system over a
// row list of int4 {row, first entry, end entry, 0}.  A row with the top bit
// set only copies xin to xout.
hipError_t launch_jacobi_sweep(int nrows, const int *rows4, const int *ci, const double *v, const double *y,
                               const double *xin, double *xout, int nr_slots, hipStream_t s);

// Jacobi sweeps over items: runs of consecutive dependents of one partition
// (reduced rows q0 .. q1).  An item's W entries (reduced CSR rp / v) read
// only the columns kcol[kb .. ke) (reduced indices, ascending), addressed by
// their position lci; row q runs need[q] sweeps.  Sweep m launches the items
// of list (top bit: some row of the item computes, so the item gathers xin
// at its columns into LDS once).  Same results, bit for bit, as
// launch_jacobi_sweep over the same sweep's row list.
struct JacobiItems {
    const int *items;       // int4 per item: q0, q1, kb, ke
    const int *kcol;
    const int *rp;
    const uint16_t *lci;
    const double *v;
    const int *need;
    const double *y;
    int kmax;               // most columns of an item
};
size_t jacobi_items_lds(int kmax, int nr_slots);
hipError_t launch_jacobi_items(const JacobiItems &j, int m, const int *list, int nitems, const double *xin,
                               double *xout, int nr_slots, hipStream_t s);

// bc[c] = b[row_c] - sum_j val_j * xred[col_j] for every coupling row c and
// right-hand side (bc, xred interleaved with a.nr_slots values per index); b
// is addressed per partition as in the block solve (a.b, a.ldb, a.nr).  The
// update is FP64 for FP32 factors too; products are rounded and summed in
// entry order, exactly as pass 2 does when it fuses the coupling update, so
// both forms give the same bits.
hipError_t launch_coupling(int P, const DevPart *parts, const int *cpl_rows, const int *cpl_ptr, const int *cpl_col,
                           const double *cpl_val, const double *xred, const BlockArgs &a, double *bc, hipStream_t s);

// Spike columns of one partition: spike rows r = 0 .. rows-1 are matrix rows
// s + f + r; K_p is kcols[kbeg .. kbeg + kcnt) (ascending); slot0 indexes
// the partition's first spike row in per-row arrays; its W rows start at woff.
struct SpikePart {
    int s, f, rows, kbeg, kcnt, slot0;
    int64_t woff;
};
// Ring spike build (layout from plan_spike_layout).  rinfo holds four ints
// per spike row (W slot or -1, far slot or -1, coupling entries, 4-entry
// chunks); a row's spike is stored at W[woff + wslot * kcnt] when its W slot
// is set.  launch_spike_codes writes the staged entries (pcode, pval) at the
// offsets pia (spike rows + 1 of them) from the lower-form CSR; R (a power of
// two): ring rows; nlr = R + 1 + far slots; kstride: LDS row stride (>= max
// kcnt when it fits one workgroup, else block_threads); sb (< block_threads):
// spike rows per stage; cap (a multiple of 4, at most
// spike_ring_stage_entries(block_threads)): the most staged entries of sb
// consecutive spike rows.  gfar: far rows are read from W (their W slots
// set), nlr = R + 1.
size_t spike_ring_lds(int nlr, int kstride, int sb, int cap);
int spike_ring_stage_entries(int block_threads);
hipError_t launch_spike_codes(int nsp, int max_rows, const SpikePart *sp, const int *kcols, const int *ia,
                              const int *ja, const double *a, const int *rinfo, const int *pia, int R, int kstride,
                              bool gfar, int *pcode, double *pval, hipStream_t s);
hipError_t launch_spike_ring(int nsp, int max_kcnt, int block_threads, const SpikePart *sp, const int *pia,
                             const int *pcode, const double *pval, const int *rinfo, int R, int nlr, int kstride,
                             int sb, int cap, bool gfar, double *W, hipStream_t s);
// Direct spike build: every spike row stored, row r at W[woff + r * kcnt].
hipError_t launch_spike_direct(int nsp, int max_kcnt, const SpikePart *sp, const int *kcols, const int *ia,
                               const int *ja, const double *a, double *W, hipStream_t s);

// Reduced-system assembly.  For reduced row q: its spike row is
// W[wrow[q] .. wrow[q]+kcnt[q]) (wrow < 0: identity row); entry l is kept when
// |w| > tol[q] (tol < 0: keep w != 0).  kridx[kbeg[q] + l] is the reduced
// column.  count pass: cnt[q], dropped[q] = sum of dropped |w|.
hipError_t launch_reduced_count(int nred, const int64_t *wrow, const int *kcnt, const double *tol,
                                const double *W, int *cnt, double *dropped, hipStream_t s);
hipError_t launch_reduced_fill(int nred, const int64_t *wrow, const int *kbeg, const int *kcnt,
                               const double *tol, const double *W, const int *kridx,
                               const int *rp, int *ci, double *v, hipStream_t s);

// r = b - A x in FP64 over the lower-form CSR (ia, ja, a_) for the
// right-hand sides of a.b / a.x (partition addressing as in the block solve;
// nr_max = most columns of any partition); r has b's layout, with ldr where
// a partition has no leading dimension of its own.
hipError_t launch_residual(int P, const DevPart *parts, const int *ia, const int *ja, const double *a_,
                           const BlockArgs &a, int nr_max, double *r, int ldr, hipStream_t s);

// Level-scheduled substitution over the lower-form CSR (diagonal last).
struct LevelArgs {
    // per level-ordered position q (levels are contiguous ranges of q): int4
    // {row, first entry, end entry, CSR start} into the strictly lower
    // entries, stored in level order (columns lja, values lval / lval32);
    // CSR start is the row's first entry in the lower-form CSR (a64)
    const int *rec;
    const int *lja;
    const double *lval;     // FP64 values
    const double *invd;     // 1/diag per row
    const float *lval32;    // FP32 values / 1/diag (fp32)
    const float *invd32;
    const double *a64;      // lower-form CSR values, diagonal last (fp32 factors)
    const double *b;
    double *x;
    int ldb, ldx, nr, nr_slots, n, flip, G, fp32;
    int spin_limit;         // sync-free: polls before a row gives up
    int *fail_word;         // sync-free: set to 1 (system scope) when a wait
                            // gives up; host memory mapped for the device
};
// Default sync-free poll limit: about a second of polling.
constexpr int kSpinLimit = 1 << 20;
// Positions r0 .. r0+cnt (one level): x = (b - strict-lower part . x) / diag.
hipError_t launch_level(int r0, int cnt, const LevelArgs &a, hipStream_t s);
// Single-RHS level step of an FP32 level factor with one FP64 refinement
// step (a.b, a.x of one column, a.fp32, a.lval32, a.invd32, a.a64): x0 into
// a.x and the correction into dl (n values, caller-indexed); after the last
// level x += dl finishes the solve.
hipError_t launch_level_ref1(int r0, int cnt, const LevelArgs &a, double *dl, hipStream_t s);
// Level step of nrp interleaved right-hand sides (v[i*nrp + r], nrp in
// {2, 4, 8, 16, 32}): one lane per row and right-hand side.
hipError_t launch_level_il(int r0, int cnt, int nrp, const LevelArgs &a, const double *bi, double *xi, hipStream_t s);
// Sync-free solve of every row in one launch, interleaved right-hand sides
// (bi / xi with nrp in {2, ..., 32} values per row, as launch_level_il),
// after a launch that fills xi with a not-ready pattern: the npos
// level-ordered records (levels padded to multiples of 64 positions) are
// taken in order, and each row polls the x values it reads until they are
// published.  A row whose wait gives up (after a.spin_limit polls) writes
// NaN and sets *abort_word.  Measured on gfx950 (lower 3D Laplacian, 128^3):
// 1.18 ms at 8 RHS and 3.14 ms at 32 RHS against 1.72 and 4.49 ms for one
// launch per level (whole solves, with the layout changes of that time); with one RHS one launch per level stays faster (1.10 ms
// against 1.59), so the single-RHS path keeps it.
hipError_t launch_sync_free(int npos, int n, int nrp, const LevelArgs &a, const double *bi, double *xi,
                            int *abort_word, hipStream_t s);
// FP32 level solve with one FP64 refinement step of nrp interleaved
// right-hand sides in one sync-free launch (after the not-ready fill of pub):
// per row x0 = M^-1 b, r = b - A x0 and dx = M^-1 r, with the arithmetic of
// launch_sync_free, launch_residual_il and launch_sync_free again, in one
// pass over the levels.  pub (n * nrp words) receives each row's x0 and dx
// as one packed 64-bit word; launch_deinterleave_pairs forms x0 + dx.
// Needs a.fp32, a.lval32, a.invd32 and a.a64.
hipError_t launch_sync_free_refined1(int npos, int n, int nrp, const LevelArgs &a, const double *bi, double *pub,
                                     int *abort_word, hipStream_t s);
// Column-major caller layout <-> interleaved lower-form layout.
hipError_t launch_interleave(int n, int nr, int nrp, int flip, const double *src, int ld, double *dst, hipStream_t s);
hipError_t launch_deinterleave(int n, int nr, int nrp, int flip, const double *v, double *x, int ld, int accum,
                               hipStream_t s);
// x (caller layout) = v + v2 for two interleaved vectors (solution plus
// refinement correction), first nr columns.
hipError_t launch_deinterleave_sum(int n, int nr, int nrp, int flip, const double *v, const double *v2, double *x,
                                   int ld, hipStream_t s);
// x (caller layout) = x0 + dx from launch_sync_free_refined1's packed words.
hipError_t launch_deinterleave_pairs(int n, int nr, int nrp, int flip, const double *v, double *x, int ld,
                                     hipStream_t s);
// r = b - A x in FP64 on interleaved vectors (nrp values per row, a power of
// two <= 32) over the lower-form CSR (ia, ja, a; diagonal included).
hipError_t launch_residual_il(int n, int nrp, const int *ia, const int *ja, const double *a, const double *bi,
                              const double *xi, double *r, hipStream_t s);
// x[:, t] += d[:, t], t < nr.
hipError_t launch_axpy(int n, int nr, const double *d, int ldd, double *x, int ldx, hipStream_t s);

// Copies ncols (<= kMaxCopyCols) columns of n doubles, dst[c][i] = src[c][i],
// in one launch (the columns may come from different allocations).
constexpr int kMaxCopyCols = 32;
struct ColCopy {
    const double *src[kMaxCopyCols];
    double *dst[kMaxCopyCols];
};
hipError_t launch_copy_cols(int n, int ncols, const ColCopy &cc, hipStream_t s);

// Largest dynamic LDS a block-solve workgroup may use, in bytes.
size_t block_solve_lds_limit();

// Device scratch of the solve entry point that allocates nothing: a
// file-scope __device__ array of kNoAllocScratchBytes on every device the
// library runs on (held until the process ends).  noalloc_scratch returns its
// address on the current device, resolved with hipGetSymbolAddress once per
// device (later calls return the kept address).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
hipError_t noalloc_scratch(double **ptr);

}  // namespace spike
