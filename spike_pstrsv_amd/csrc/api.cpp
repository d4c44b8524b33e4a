// C ABI of spike_pstrsv for MI355X: the reference's five legacy entry points
// (reference src/spike_pstrsv.h:7-38) over a device-resident factor object,
// plus the extension API declared in include/spike_pstrsv.h.
#include "spike_pstrsv.h"
#include "spike_pstrsv_globals.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"

using spike::DevPart;
using spike::Plan;

// Globals read by the reference's timing harness (include/spike_pstrsv_globals.h).
extern "C" {
char PSTRSV_VERBOSE = 0;
char FORCE_LOAD_BALANCE = 0;
char IS_MATRIX_PARTITIONED = 0;
double memoryAlloc_timer_pstrsv3 = 0;
double preprocessD_timer_pstrsv3 = 0;
double preprocessR_timer_pstrsv3 = 0;
double setDependents_timer_pstrsv3 = 0;
double preprocessS_timer_pstrsv3 = 0;
double preprocessMappings_timer_pstrsv3 = 0;
double produceCompressedR_timer_pstrsv3 = 0;
double produceS_timer_pstrsv3 = 0;
double produceReducedSystem_timer_pstrsv3 = 0;
double init_vars_pstrsv3 = 0;
double reduced_solver_timer_pstrsv3 = 0;
double temp_timer[MAXTHREADCOUNT];
double waitBefore_timer_pstrsv3[MAXTHREADCOUNT];
double computeY_timer[MAXTHREADCOUNT];
double solveIndependent_timer[MAXTHREADCOUNT];
double subtract_Rx_timer[MAXTHREADCOUNT];
double waitAfter_timer_pstrsv3[MAXTHREADCOUNT];
double solveD2_timer[MAXTHREADCOUNT];
double solveD_timer[MAXTHREADCOUNT];
}

namespace {

thread_local std::string t_err;

int fail(int code, const std::string &msg)
{
    t_err = msg;
    return code;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? SPIKE_ERR_NOMEM : SPIKE_ERR_DEVICE,        \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                    \
    } while (0)

// Global timers of the profiling twins (seconds, accumulated like the
// reference's timers in src/spike_pstrsv.c:18-37).
double g_timers[9];

}  // namespace

struct spike_factor_s {
    Plan plan;
    spike_options_t opts{};
    spike_stats_t stats{};
    int device = 0;
    std::vector<void *> allocs;
    std::vector<size_t> alloc_bytes;   // size of allocs[i]
    int64_t device_bytes = 0;
    // SPIKE_PSTRSV_FAIL_ALLOC=<k> (tests): the factor's k-th device
    // allocation fails as if the device were out of memory
    int fail_alloc_at = 0, nalloc = 0;

    DevPart *d_parts = nullptr;
    int *d_dep_rows = nullptr;
    int *d_cpl_rows = nullptr, *d_cpl_ptr = nullptr, *d_cpl_col = nullptr;
    double *d_cpl_val = nullptr;
    spike::DevSteps steps{};
    int *d_rrp = nullptr, *d_rci = nullptr, *d_rlist = nullptr, *d_levptr = nullptr;
    double *d_rv = nullptr;
    int nlev = 0;
    bool red_in_lds = false;
    int sweeps = 0;             // > 0: Jacobi sweeps, 0: exact level solve
    // sweeps an FP32 factor's solves run: the inner solves of refinement
    // need the reduced solve only to about FP32 accuracy (tail bound 1e-8)
    int sweeps_fp32 = 0;
    std::vector<int> sweep_off; // sweep m: int4 row records off[m-1] .. off[m]
    int *d_sweep_rows = nullptr;
    // Jacobi sweeps over items (spike::launch_jacobi_items): the items, their
    // column lists, each entry's local column and each row's sweep count;
    // sweep m launches the items ji_lists[ji_off[m-1] .. ji_off[m]).  Used at
    // RHS slot counts >= ji_min_slots (0: never).
    int ji_min_slots = 0, ji_kmax = 0;
    int64_t ji_rnnz = 0;        // reduced-system entries (the item path's work gate)
    bool ji_force = false;      // SPIKE_PSTRSV_JACOBI_ITEMS set: no work gate
    int *d_ji_items = nullptr, *d_ji_kcol = nullptr, *d_ji_need = nullptr, *d_ji_lists = nullptr;
    uint16_t *d_ji_lci = nullptr;
    std::vector<int> ji_off;
    int ring = 16;              // block-solve step records in flight (tiles 1, 2; 8 above)
    bool cpl_fused = false;     // pass 2 can compute its coupling update itself
    bool fuse_at[4] = {};       // ... and does so at 1, 2, 4, 8 RHS slots
    int tile_at[4] = {1, 1, 1, 1};  // RHS per wave at 1, 2, 4, 8 slots
    int ncu = 1, lds_cu = 65536;
    int max_nr = 1;             // right-hand sides per block-solve launch
    int cmax = 0;               // most coupling entries of one partition
    double *d_y = nullptr, *d_xa = nullptr, *d_xb = nullptr, *d_bc = nullptr;
    double *d_inv_diag = nullptr;
    // staging for host right-hand sides / solutions
    double *d_stage_b = nullptr, *d_stage_x = nullptr;
    size_t stage_cap = 0;
    // SPIKE_PREC_FP32_REFINED: FP32 step and coupling values, the FP64
    // lower-form CSR for the residual, a residual workspace (n x max_nr)
    bool fp32 = false;
    const double *d_a = nullptr;
    const int *d_ia = nullptr, *d_ja = nullptr;
    double *d_r = nullptr;
    int refine_steps = 0;
    double refine_rate = 0.0, refine_est = 0.0;
    // Matrices held by the factor.  Empty: one matrix, addressed by the
    // solve call's leading dimensions and RHS count.  Otherwise a
    // block-diagonal batch: matrix i owns lower-form rows [row0, row0+n) and
    // its nrhs right-hand sides start at element boff of the packed b and x
    // (column-major, leading dimension n); batch_flip: inputs were 'U'.
    struct Mat {
        int row0, n, nrhs;
        int64_t boff;
    };
    std::vector<Mat> mats;
    bool batch_flip = false;
    int batch_slots = 0;        // RHS slots of every launch of a batch factor
    int batch_max_nrhs = 0;
    int64_t vec_len = 0;        // batch: packed length of b and x
    // SPIKE_METHOD_LEVELS: rows grouped by dependency level (lev_ptr on the
    // host delimits the launches), their records and strictly lower entries
    // in level order and 1/diag on the device (plus the FP64 lower-form CSR
    // for the residual of FP32 factors), a correction workspace for
    // refinement
    int method = SPIKE_METHOD_SPIKE;
    bool caller_parts = false;  // partitions given by the caller: Spike only
    int nlev_full = 0;
    std::vector<int> lev_ptr;
    int *d_lrec = nullptr;      // int4 per level-ordered position: row, entries
    int *d_lja = nullptr;       // strictly lower entries in level order
    double *d_lval = nullptr;
    float *d_lval32 = nullptr, *d_invd32 = nullptr;
    bool sync_free = true;      // multi-RHS: one sync-free launch instead of one per level
    int spin_limit = spike::kSpinLimit;
    int *d_abort = nullptr;     // its abort word
    // Failure word of the bounded waits: pinned host memory that a kernel
    // sets (system scope) when a wait gives up; the library reads it after
    // synchronising solves, at the start of the next call on the factor and
    // in spike_factor_status.  d_fail is its device address.
    int *h_fail = nullptr, *d_fail = nullptr;
    // Single-launch solve (spike::launch_spike_fused): its sweep items (int4
    // {first row, end row, first entry, end entry} of the reduced system)
    // and, per RHS-slot class (1, 2, 4, 8), its published vectors and queue
    // counters, allocated on first use outside any graph capture
    // (prepare_fused); fused_fit caches per slot class and tile whether the
    // launch fits the device (-1 unknown).
    bool fused = false;
    bool spike_ring = false;   // spikes built by the LDS ring kernel (else direct)
    int *d_chunks = nullptr;
    int nchunks = 0;
    double *d_fx[4] = {};
    unsigned *d_fctr[4] = {};
    signed char fused_fit[4][4] = {{-1, -1, -1, -1}, {-1, -1, -1, -1}, {-1, -1, -1, -1}, {-1, -1, -1, -1}};
    double *d_dlt = nullptr;
    int lev_G = 1;
    // interleaved right-hand sides / solutions of multi-RHS level solves
    double *d_bi = nullptr, *d_xi = nullptr;
    size_t il_cap = 0;
    // row blocks of the residual kernel (at most 64 rows of one partition)
    int nres = 0;
    DevPart *d_res = nullptr;

    template <class T>
    int dalloc(T **p, size_t count)
    {
        *p = nullptr;
        if (count == 0) count = 1;
        if (fail_alloc_at > 0 && ++nalloc == fail_alloc_at)
            return fail(SPIKE_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                             " bytes failed (simulated by SPIKE_PSTRSV_FAIL_ALLOC)");
        hipError_t e = hipMalloc((void **)p, count * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SPIKE_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        allocs.push_back(*p);
        alloc_bytes.push_back(count * sizeof(T));
        device_bytes += (int64_t)(count * sizeof(T));
        return SPIKE_OK;
    }
    template <class T>
    int upload(T **p, const std::vector<T> &v)
    {
        int rc = dalloc(p, v.size());
        if (rc) return rc;
        if (!v.empty()) HIP_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return SPIKE_OK;
    }
    // Solve graphs for device-pointer solves: a key (b, x, ldb, ldx, nrhs)
    // seen on two consecutive calls is captured once and replayed.
    struct GraphKey {
        const double *b;
        double *x;
        int ldb, ldx, nrhs;
        bool operator==(const GraphKey &o) const
        {
            return b == o.b && x == o.x && ldb == o.ldb && ldx == o.ldx && nrhs == o.nrhs;
        }
    };
    struct GraphEntry {
        GraphKey key;
        hipGraphExec_t exec;
        hipGraph_t graph;
        uint64_t used;
    };
    std::vector<GraphEntry> graphs;
    GraphKey last_key{nullptr, nullptr, 0, 0, 0};
    uint64_t graph_clock = 0;
    hipStream_t cap_stream = nullptr;

    void drop_graphs()
    {
        for (GraphEntry &g : graphs) {
            (void)hipGraphExecDestroy(g.exec);
            (void)hipGraphDestroy(g.graph);
        }
        graphs.clear();
    }
    void release()
    {
        drop_graphs();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        cap_stream = nullptr;
        for (void *p : allocs) (void)hipFree(p);
        allocs.clear();
        alloc_bytes.clear();
        device_bytes = 0;
        if (h_fail) (void)hipHostFree(h_fail);
        h_fail = d_fail = nullptr;
    }
};

namespace {

// n timing events, all or none (on failure the created ones are destroyed).
int create_events(hipEvent_t *ev, int n)
{
    for (int i = 0; i < n; i++) {
        ev[i] = nullptr;
        if (hipEventCreate(&ev[i]) != hipSuccess) {
            for (int j = 0; j < i; j++) (void)hipEventDestroy(ev[j]);
            (void)hipGetLastError();
            return fail(SPIKE_ERR_DEVICE, "event creation failed");
        }
    }
    return SPIKE_OK;
}

std::vector<float> to_f32(const std::vector<double> &v)
{
    return std::vector<float>(v.begin(), v.end());
}

// Step stream to the device: FP64 values, or (fp32) their FP32 roundings only.
int upload_steps(spike_factor_s *f, const spike::StepStream &ss, spike::DevSteps &ds, bool fp32)
{
    double *val = nullptr, *late = nullptr;
    float *val32 = nullptr;
    uint16_t *col, *row;
    int64_t *pb;
    int rc;
    if ((rc = f->upload(&col, ss.col)) || (rc = f->upload(&row, ss.row)) || (rc = f->upload(&pb, ss.part_begin)))
        return rc;
    if (fp32) {
        if (ss.fwd) return fail(SPIKE_ERR_UNSUPPORTED, "the forwarding block solve has no FP32 form");
        if ((rc = f->upload(&val32, to_f32(ss.val)))) return rc;
    } else {
        if ((rc = f->upload(&val, ss.val))) return rc;
        if (ss.fwd && (rc = f->upload(&late, ss.late))) return rc;
    }
    ds.val = val; ds.col = col; ds.row = row; ds.late = late; ds.part_begin = pb;
    ds.val32 = val32;
    ds.G = ss.G; ds.M = ss.M; ds.R = ss.R; ds.fwd = ss.fwd ? 1 : 0;
    return SPIKE_OK;
}

// Frees a temporary device buffer that was registered with dalloc.
void dfree(spike_factor_s *f, void *p, size_t bytes)
{
    auto it = std::find(f->allocs.begin(), f->allocs.end(), p);
    if (it != f->allocs.end()) {
        f->alloc_bytes.erase(f->alloc_bytes.begin() + (it - f->allocs.begin()));
        f->allocs.erase(it);
        f->device_bytes -= (int64_t)bytes;
    }
    (void)hipFree(p);
}

int kmax_limit_for_device()
{
    // x segment of kmax + XS_PAD doubles (spare slots after the rows)
    size_t lds = spike::block_solve_lds_limit();
    int lim = (int)(lds / sizeof(double)) - spike::XS_PAD;
    return std::min(lim, 16383);
}

// Launch rounds of a block solve with slots columns in tiles of t per wave,
// each wave holding `bytes` of LDS.
long tile_rounds(const spike_factor_s *f, int slots, int t, size_t bytes)
{
    long per_cu = std::max<long>(1, (long)((size_t)f->lds_cu / std::max<size_t>(bytes, 1)));
    per_cu = std::min<long>(per_cu, 32);
    const long waves = (long)f->plan.P * (slots / t);
    return (waves + (long)f->ncu * per_cu - 1) / ((long)f->ncu * per_cu);
}

// Whether pass 2 fuses its coupling update at this slot count and tile: the
// coupling products must fit the LDS next to the x segment without costing a
// launch round.
bool fuse_ok(const spike_factor_s *f, int slots, int t)
{
    const size_t lim = spike::block_solve_lds_limit();
    const size_t vb = f->fp32 ? sizeof(float) : sizeof(double);
    const size_t x_per = (size_t)(f->plan.kmax + spike::XS_PAD) * vb + 8, c_per = (size_t)f->cmax * sizeof(double);
    if (!f->cpl_fused || (x_per + c_per) * t > lim) return false;
    return tile_rounds(f, slots, t, (x_per + c_per) * t) == tile_rounds(f, slots, t, x_per * t);
}

// Right-hand sides per block-solve launch (chunk of up to 8 columns) and per
// wave (the tile).  Every partition is solved by nr_slots / tile waves, each
// keeping its x segment (kmax+XS_PAD values per RHS of its tile, plus cmax coupling
// products per RHS when pass 2 fuses its coupling update) in LDS, so the tile
// trades work per step against resident waves per CU.  Model: relative step
// cost at tile 1/2/4/8 (measured on gfx950 with 1024-row partitions) 1.0,
// 1.22, 2.1, 4.2; a launch costs rounds x step cost, rounds = ceil(waves /
// (CUs x waves resident per CU)).  The model gives the default tile of each
// slot count; the autotuner (tune_tiles) replaces it by measurement.  Pass 2
// fuses the coupling update only where that keeps its round count.
void choose_rhs_slots(spike_factor_s *f)
{
    const Plan &pl = f->plan;
    const size_t lim = spike::block_solve_lds_limit();
    int dev = f->device, ncu = 0, lds_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 1;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        lds_cu <= 0)
        lds_cu = (int)lim;
    (void)hipGetLastError();
    f->ncu = ncu;
    f->lds_cu = lds_cu;
    const size_t vb = f->fp32 ? sizeof(float) : sizeof(double);
    const size_t x_per = (size_t)(pl.kmax + spike::XS_PAD) * vb;
    const size_t c_per = (size_t)f->cmax * sizeof(double);
    f->cpl_fused = x_per + 8 + c_per <= lim;
    static const double step_cost[4] = {1.0, 1.22, 2.1, 4.2};
    for (int li = 0; li < 4; li++) {
        f->tile_at[li] = 0;
        f->fuse_at[li] = false;
    }
    if (f->batch_slots > 0) {
        const int li = __builtin_ctz((unsigned)f->batch_slots);
        if (x_per * f->batch_slots > lim) {
            f->max_nr = 0;
            return;
        }
        // a batch group: waves of at most 4 of its RHS slots (8-slot groups
        // run two waves per partition, which halves the LDS x segment of a
        // wave; on the 4096-matrix batch 19.7 ms per solve against 20.2 with
        // 8-RHS waves, 25.2 with 2-RHS waves, and 20.1 with 2-RHS waves for
        // the 4-slot groups only); SPIKE_PSTRSV_BATCH_TILE=<t>
        // (1, 2, 4 or 8) sets another bound
        int t = std::min(f->batch_slots, 4);
        if (const char *e = std::getenv("SPIKE_PSTRSV_BATCH_TILE")) {
            const int w = std::atoi(e);
            if (w == 1 || w == 2 || w == 4 || w == 8) t = std::min(f->batch_slots, w);
        }
        f->tile_at[li] = t;
        f->fuse_at[li] = fuse_ok(f, f->batch_slots, t);
        f->max_nr = f->batch_slots;
        return;
    }
    f->max_nr = pl.steps.fwd ? 1 : 8;
    for (int li = 0; li < 4 && (1 << li) <= f->max_nr; li++) {
        const int c = 1 << li;
        double best = -1.0;
        for (int ti = 0; ti <= li; ti++) {
            const int t = 1 << ti;
            if (x_per * t > lim) break;
            const double thr = c / (step_cost[ti] * (double)tile_rounds(f, c, t, x_per * t));
            if (thr > best * 1.02) {
                best = thr;
                f->tile_at[li] = t;
            }
        }
        if (f->tile_at[li] == 0) f->tile_at[li] = 1;
        f->fuse_at[li] = fuse_ok(f, c, f->tile_at[li]);
    }
}

int calibrate_refinement(spike_factor_s *f);
int solve_pass(spike_factor_s *f, const double *b, int ldb, double *x, int ldx, int nr, int accum, hipStream_t s,
               hipEvent_t *ev, double *ph);
int prepare_fused(spike_factor_s *f, int nr);

// Autotuner of the RHS tile: for 2, 4 and 8 slots, every tile that fits the
// LDS is timed on zero right-hand sides (two warm-up solves, then the fastest
// of three rounds of two solves); the fastest tile replaces the model's
// choice (ties within 3% keep the smaller tile).
int tune_tiles(spike_factor_s *f)
{
    const int n = f->plan.n;
    const size_t lim = spike::block_solve_lds_limit();
    const size_t vb = f->fp32 ? sizeof(float) : sizeof(double);
    const size_t x_per = (size_t)(f->plan.kmax + spike::XS_PAD) * vb;
    double *db = nullptr, *dx = nullptr;
    int rc;
    if ((rc = f->dalloc(&db, (size_t)n * 8)) || (rc = f->dalloc(&dx, (size_t)n * 8))) return rc;
    HIP_TRY(hipMemset(db, 0, sizeof(double) * (size_t)n * 8));
    hipEvent_t ee[2];
    if ((rc = create_events(ee, 2))) return rc;
    hipEvent_t e0 = ee[0], e1 = ee[1];
    for (int li = 1; li < 4 && (1 << li) <= f->max_nr && !rc; li++) {
        const int c = 1 << li;
        if ((rc = prepare_fused(f, c))) break;
        int best_t = f->tile_at[li];
        float best_ms = 1e30f;
        for (int ti = 0; ti <= li && !rc; ti++) {
            const int t = 1 << ti;
            if (x_per * t > lim) break;
            f->tile_at[li] = t;
            f->fuse_at[li] = fuse_ok(f, c, t);
            for (int r = 0; r < 2 && !rc; r++) rc = solve_pass(f, db, n, dx, n, c, 0, nullptr, nullptr, nullptr);
            if (rc) break;
            // the fastest of three rounds of two solves (single rounds were
            // seen to flip the choice between runs)
            float ms = 1e30f;
            for (int round = 0; round < 3 && !rc; round++) {
                (void)hipEventRecord(e0, nullptr);
                for (int r = 0; r < 2 && !rc; r++) rc = solve_pass(f, db, n, dx, n, c, 0, nullptr, nullptr, nullptr);
                (void)hipEventRecord(e1, nullptr);
                (void)hipEventSynchronize(e1);
                float t_ms = 0.f;
                (void)hipEventElapsedTime(&t_ms, e0, e1);
                ms = std::min(ms, t_ms);
            }
            if (ms < best_ms * 0.97f) {
                best_ms = ms;
                best_t = t;
            }
        }
        f->tile_at[li] = best_t;
        f->fuse_at[li] = fuse_ok(f, c, best_t);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(f, db, sizeof(double) * (size_t)n * 8);
    dfree(f, dx, sizeof(double) * (size_t)n * 8);
    return rc;
}

// Automatic partition count: partitions of about 16x the coupling bandwidth
// (95th percentile of the distance from a row to its farthest column, so rare
// long-range entries do not count), at least 1024 rows and at most kmax_limit.
// Partitions much longer than the bandwidth keep the dependents (and so the
// reduced system) a small fraction of n.
int default_partitions(char uplo, int n, const int *ia, const int *ja, int kmax_limit)
{
    std::vector<int> reach;
    const int stride = std::max(1, n / (1 << 16));
    const bool up = uplo == 'U' || uplo == 'u';
    for (int i = 0; i < n; i += stride) {
        if (ia[i + 1] <= ia[i]) continue;
        reach.push_back(up ? ja[ia[i + 1] - 1] - i : i - ja[ia[i]]);
    }
    long bw = 1;
    if (!reach.empty()) {
        const size_t q = (reach.size() * 95) / 100;
        std::nth_element(reach.begin(), reach.begin() + q, reach.end());
        bw = std::max(1, reach[q]);
    }
    const long k = std::min<long>(std::max<long>(1024, 16 * bw), std::max(kmax_limit, 1));
    return (int)std::min<long>(n, std::max<long>(1, (n + k / 2) / k));
}

// Row blocks of the residual kernel: 64 rows each, cut from the given
// partition records (or from one record covering the whole matrix), so that
// every block keeps its matrix's right-hand-side addressing.
int upload_row_blocks(spike_factor_s *f, const std::vector<DevPart> &parts)
{
    std::vector<DevPart> rb;
    for (const DevPart &p : parts)
        for (int s0 = 0; s0 < p.k; s0 += 64) {
            DevPart d = p;
            d.s = p.s + s0;
            d.k = std::min(64, p.k - s0);
            rb.push_back(d);
        }
    f->nres = (int)rb.size();
    return f->upload(&f->d_res, rb);
}

// The pinned host failure word of the bounded waits (see spike_factor_s).
int setup_fail_word(spike_factor_s *f)
{
    if (f->h_fail) return SPIKE_OK;
    void *h = nullptr;
    HIP_TRY(hipHostMalloc(&h, sizeof(int), hipHostMallocMapped));
    f->h_fail = static_cast<int *>(h);
    *f->h_fail = 0;
    void *d = nullptr;
    HIP_TRY(hipHostGetDevicePointer(&d, h, 0));
    f->d_fail = static_cast<int *>(d);
    return SPIKE_OK;
}

// SPIKE_ERR_DEVICE (once) when a bounded wait of an earlier solve on f gave
// up, i.e. that solve's x holds NaN where rows could not be computed.
int take_failure(spike_factor_s *f)
{
    if (!f->h_fail || __atomic_load_n(f->h_fail, __ATOMIC_ACQUIRE) == 0) return SPIKE_OK;
    __atomic_store_n(f->h_fail, 0, __ATOMIC_RELEASE);
    return fail(SPIKE_ERR_DEVICE, "a solve on this factor gave up a bounded device wait (its x holds NaN); "
                                  "the device did not run the solve's workgroups in dispatch order");
}

// Poll limit of the bounded device waits; SPIKE_PSTRSV_SPIN_LIMIT (a test
// hook: a small limit forces the give-up path) overrides the default.
void read_spin_limit(spike_factor_s *f)
{
    if (const char *l = std::getenv("SPIKE_PSTRSV_SPIN_LIMIT")) f->spin_limit = std::max(0, std::atoi(l));
}

// Sync-free switch (SPIKE_PSTRSV_SYNC_FREE=0 selects one launch per level)
// and the abort word it needs (every solve clears it with a kernel).
int setup_sync_free(spike_factor_s *f)
{
    const char *e = std::getenv("SPIKE_PSTRSV_SYNC_FREE");
    f->sync_free = !(e && std::atoi(e) == 0);
    read_spin_limit(f);
    if (!f->sync_free || f->d_abort) return SPIKE_OK;
    int rc;
    if ((rc = f->dalloc(&f->d_abort, 1)) || (rc = setup_fail_word(f))) return rc;
    HIP_TRY(hipMemset(f->d_abort, 0, sizeof(int)));
    return SPIKE_OK;
}

// Items of the Jacobi sweeps (spike::launch_jacobi_items): the dependents of
// each partition cut into runs of consecutive rows of at most about 2048
// entries and 256 distinct columns (the column values in LDS stay small
// enough for eight workgroups per CU at 8 RHS slots); an item's column list
// holds the distinct reduced columns its entries read.  Sweep m lists the
// items with a row that computes (need >= m; flagged: the item gathers) or
// copies (need == m-1, or need == 0 at m == 2).
//
// The item path serves solves of at least SPIKE_PSTRSV_JACOBI_ITEMS RHS
// slots (default 4; 0 turns it off; a value set there also drops the size
// gate) whose reduced system times the slot count reaches 2^25 entries: with
// several right-hand sides one gather per column replaces one per entry, and
// small reduced systems keep the row sweeps (a factor none of whose solves
// can pass both gates builds no items).  Measured on the 4096-matrix
// batch (512-row partitions, kernel time per solve): 8 slots 14.8 -> 10.9 ms,
// 4 slots 5.7 -> 4.4 ms, 2 slots 2.7 -> 3.4 ms; one RHS on the 1M-row banded
// matrix 47 -> 60 us.
int build_jacobi_items(spike_factor_s *f, const std::vector<int> &hrp, const std::vector<int> &hci,
                       const std::vector<int> &row_sweeps)
{
    const Plan &pl = f->plan;
    f->ji_min_slots = 0;
    int min_slots = 4;
    f->ji_force = false;
    if (const char *e = std::getenv("SPIKE_PSTRSV_JACOBI_ITEMS")) {
        min_slots = std::max(0, std::atoi(e));
        f->ji_force = true;
    }
    if (min_slots == 0 || f->sweeps <= 0) return SPIKE_OK;
    constexpr int kTarget = 2048, kColCap = 256;
    const int nred = (int)pl.dep_rows.size();
    const size_t rnnz = (size_t)hrp[nred];
    // no solve of this factor can take the item path: skip the build
    if (f->max_nr < min_slots || (!f->ji_force && (int64_t)rnnz * f->max_nr < ((int64_t)1 << 25))) return SPIKE_OK;
    std::vector<uint16_t> lci(std::max<size_t>(rnnz, 1), 0);
    // partitions on host threads; a run's distinct columns live in an
    // open-addressing set (generation-stamped, so a new run clears it free)
    struct PartItems {
        std::vector<int> items, kcol;
        int kmax = 0;
        bool too_wide = false;
    };
    std::vector<PartItems> per(pl.P);
    spike::parallel_for(pl.P, [&](int t) {
        const spike::Part &pt = pl.parts[t];
        if (pt.dcnt == 0) return;
        thread_local std::vector<int> hkey, hval, hgen, cols;
        thread_local int gen = 0;
        auto slot_of = [&](int c) {
            const size_t mask = hkey.size() - 1;
            size_t h = ((uint32_t)c * 2654435761u) & mask;
            while (hgen[h] == gen && hkey[h] != c) h = (h + 1) & mask;
            return h;
        };
        auto reserve = [&](size_t need) {   // room for `need` keys at load <= 1/2
            size_t cap = std::max<size_t>(hkey.size(), 1024);
            while (cap < 2 * need) cap *= 2;
            if (cap == hkey.size()) return;
            hkey.assign(cap, -1);
            hval.assign(cap, 0);
            hgen.assign(cap, -1);
            gen++;
            for (int c : cols) {   // re-insert the current run's columns
                const size_t h = slot_of(c);
                hkey[h] = c;
                hgen[h] = gen;
            }
        };
        PartItems &out = per[t];
        for (int q0 = pt.dbeg, end = pt.dbeg + pt.dcnt; q0 < end;) {
            // extend the run while it stays within kTarget entries and
            // kColCap distinct columns (a single row may exceed either)
            gen++;
            cols.clear();
            reserve(0);
            int q1 = q0, ncol = 0;
            while (q1 < end) {
                reserve((size_t)ncol + (size_t)(hrp[q1 + 1] - hrp[q1]));
                int fresh = 0;
                for (int j = hrp[q1]; j < hrp[q1 + 1]; j++) {
                    const size_t h = slot_of(hci[j]);
                    if (hgen[h] != gen) {
                        hkey[h] = hci[j];
                        hgen[h] = gen;
                        cols.push_back(hci[j]);
                        fresh++;
                    }
                }
                if (q1 > q0 && (hrp[q1 + 1] - hrp[q0] > kTarget || ncol + fresh > kColCap)) {
                    cols.resize(ncol);   // row q1 opens the next run
                    break;
                }
                ncol += fresh;
                q1++;
            }
            if (cols.size() > 65535) {
                out.too_wide = true;
                return;
            }
            std::sort(cols.begin(), cols.end());
            const int kb = (int)out.kcol.size();
            for (size_t i = 0; i < cols.size(); i++) {
                hval[slot_of(cols[i])] = (int)i;
                out.kcol.push_back(cols[i]);
            }
            for (int j = hrp[q0]; j < hrp[q1]; j++) lci[j] = (uint16_t)hval[slot_of(hci[j])];
            out.kmax = std::max(out.kmax, (int)cols.size());
            out.items.insert(out.items.end(), {q0, q1, kb, kb + (int)cols.size()});
            q0 = q1;
        }
    });
    std::vector<int> items, kcol;
    int kmax = 0;
    for (const PartItems &pi : per) {
        if (pi.too_wide) return SPIKE_OK;
        const int off = (int)kcol.size();
        for (size_t i = 0; i < pi.items.size(); i += 4)
            items.insert(items.end(), {pi.items[i], pi.items[i + 1], pi.items[i + 2] + off, pi.items[i + 3] + off});
        kcol.insert(kcol.end(), pi.kcol.begin(), pi.kcol.end());
        kmax = std::max(kmax, pi.kmax);
    }
    const int nit = (int)(items.size() / 4);
    if (nit == 0) return SPIKE_OK;
    // per item: the largest need and which needs below 64 occur
    std::vector<int> maxneed(nit, 0);
    std::vector<uint64_t> needs(nit, 0);
    spike::parallel_chunks(nit, 4096, [&](int, int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi; t++)
            for (int q = items[4 * t]; q < items[4 * t + 1]; q++) {
                const int nd = row_sweeps[q];
                maxneed[t] = std::max(maxneed[t], nd);
                if (nd < 64) needs[t] |= (uint64_t)1 << nd;
            }
    });
    std::vector<int> lists;
    f->ji_off.assign(f->sweeps + 1, 0);
    for (int m = 1; m <= f->sweeps; m++) {
        for (int t = 0; t < nit; t++) {
            const bool comp = maxneed[t] >= m;
            const bool copy = ((needs[t] >> (m - 1)) & 1) || (m == 2 && (needs[t] & 1));
            if (comp) lists.push_back(t | (int)0x80000000);
            else if (copy) lists.push_back(t);
        }
        f->ji_off[m] = (int)lists.size();
    }
    int rc;
    if ((rc = f->upload(&f->d_ji_items, items)) || (rc = f->upload(&f->d_ji_kcol, kcol)) ||
        (rc = f->upload(&f->d_ji_need, row_sweeps)) || (rc = f->upload(&f->d_ji_lci, lci)) ||
        (rc = f->upload(&f->d_ji_lists, lists)))
        return rc;
    f->ji_kmax = kmax;
    f->ji_rnnz = (int64_t)rnnz;
    f->ji_min_slots = min_slots;
    return SPIKE_OK;
}

// Device data of the level-scheduled method.
int build_levels(spike_factor_s *f, const std::vector<int> &rows, double t0)
{
    Plan &pl = f->plan;
    const int n = pl.n;
    int rc;
    // Row records {row, first entry, end entry, the row's first entry in the
    // lower-form CSR} and strictly lower entries in level order.  Every level
    // starts at a multiple of 64 positions (padding records have row -1), so
    // no wavefront holds rows of two levels: in the sync-free solve a row may
    // wait for another wavefront's rows, never for its own wavefront's.
    std::vector<int> rec, lja;
    std::vector<double> lval;
    lja.reserve((size_t)(pl.nnz - n));
    lval.reserve((size_t)(pl.nnz - n));
    std::vector<int> padded_ptr(1, 0);
    for (int l = 0; l + 1 < (int)f->lev_ptr.size(); l++) {
        for (int q = f->lev_ptr[l]; q < f->lev_ptr[l + 1]; q++) {
            const int row = rows[q];
            const int eb = (int)lja.size();
            for (int e = pl.ia[row]; e < pl.ia[row + 1] - 1; e++) {   // the diagonal is last
                lja.push_back(pl.ja[e]);
                lval.push_back(pl.a[e]);
            }
            rec.insert(rec.end(), {row, eb, (int)lja.size(), pl.ia[row]});
        }
        while ((rec.size() / 4) % 64) rec.insert(rec.end(), {-1, 0, 0, 0});
        padded_ptr.push_back((int)(rec.size() / 4));
    }
    f->lev_ptr = padded_ptr;
    if ((rc = f->upload(&f->d_lrec, rec)) || (rc = f->upload(&f->d_lja, lja)) ||
        (rc = f->upload(&f->d_inv_diag, pl.inv_diag)))
        return rc;
    if (f->fp32) {
        // FP32 values for the solves; the lower-form CSR in FP64 for the residual
        int *d_ia, *d_ja;
        double *d_a;
        if ((rc = f->upload(&f->d_lval32, to_f32(lval))) || (rc = f->upload(&f->d_invd32, to_f32(pl.inv_diag))) ||
            (rc = f->upload(&d_ia, pl.ia)) || (rc = f->upload(&d_ja, pl.ja)) || (rc = f->upload(&d_a, pl.a)))
            return rc;
        f->d_a = d_a;
        f->d_ia = d_ia;
        f->d_ja = d_ja;
    } else if ((rc = f->upload(&f->d_lval, lval))) {
        return rc;
    }
    std::vector<int> len(n);
    for (int i = 0; i < n; i++) len[i] = pl.ia[i + 1] - pl.ia[i] - 1;
    f->lev_G = f->opts.lanes_per_row > 0 ? std::min(f->opts.lanes_per_row, 16) : spike::choose_lanes(len);
    f->max_nr = 32;
    if ((rc = setup_sync_free(f))) return rc;
    {
        DevPart whole{};
        whole.s = 0;
        whole.k = n;
        whole.mn = n;
        whole.mflip = pl.flip ? 1 : 0;
        if ((rc = upload_row_blocks(f, std::vector<DevPart>{whole}))) return rc;
    }
    if (f->fp32 && ((rc = f->dalloc(&f->d_r, (size_t)n * 32)) || (rc = f->dalloc(&f->d_dlt, (size_t)n)))) return rc;
    pl.a.clear(); pl.a.shrink_to_fit();
    pl.ja.clear(); pl.ja.shrink_to_fit();
    pl.steps = spike::StepStream();
    spike_stats_t &st = f->stats;
    st.n = n;
    st.partitions = 0;
    st.nnz = pl.nnz;
    st.lanes_per_row = f->lev_G;
    st.rhs_per_launch = f->max_nr;
    st.precision = f->opts.precision;
    st.method = SPIKE_METHOD_LEVELS;
    st.levels = f->nlev_full;
    if (f->fp32 && (rc = calibrate_refinement(f))) return rc;
    st.refine_steps = f->refine_steps;
    st.refine_rate = f->refine_rate;
    st.refine_est_err = f->refine_est;
    st.device_bytes = f->device_bytes;
    st.preproc_seconds = now_s() - t0;
    g_timers[0] += st.preproc_seconds;
    return SPIKE_OK;
}

int factor_build(spike_factor_s *f, char uplo, int n, const double *a, const int *ia, const int *ja,
                 const std::vector<int> *starts = nullptr)
{
    const double t0 = now_s();
    int rc;
    HIP_TRY(hipGetDevice(&f->device));
    if (const char *e = std::getenv("SPIKE_PSTRSV_FAIL_ALLOC")) f->fail_alloc_at = std::max(0, std::atoi(e));
    spike::PlanOptions po;
    po.kmax_limit = kmax_limit_for_device();
    int P = f->opts.partitions;
    if (P <= 0) P = default_partitions(uplo, n, ia, ja, po.kmax_limit);
    P = std::min(P, n);
    // raise P until the largest partition fits the LDS x segment
    while (P < n && n - (n / P) * (P - 1) > po.kmax_limit) P++;
    if (starts) {
        po.starts = *starts;
        P = (int)starts->size() - 1;
    }
    f->fp32 = f->opts.precision == SPIKE_PREC_FP32_REFINED;
    if (const char *e = std::getenv("SPIKE_PSTRSV_FORWARD")) po.forward = std::atoi(e) != 0 && !f->fp32;
    po.P = P;
    po.lanes_per_row = f->opts.lanes_per_row;
    po.fp32 = f->fp32;
    if (const char *e = std::getenv("SPIKE_PSTRSV_RING")) f->ring = std::atoi(e) >= 16 ? 16 : 8;
    std::string err = spike::build_plan(uplo, n, a, ia, ja, po, f->plan);
    if (!err.empty()) return fail(SPIKE_ERR_MATRIX, err);
    Plan &pl = f->plan;
    const double t1 = now_s();
    g_timers[1] += t1 - t0;
    {
        // dependency levels of the whole matrix (diagonal last)
        std::vector<int> lev(n, 0);
        int nl = 0;
        for (int i = 0; i < n; i++) {
            int l = 0;
            for (int e = pl.ia[i]; e < pl.ia[i + 1] - 1; e++) l = std::max(l, lev[pl.ja[e]] + 1);
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
        f->nlev_full = nl;
        // cost model (microseconds): a level launch ~3 plus streaming the
        // matrix at ~4 TB/s; Spike ~0.19 per block-solve step and pass plus
        // ~7 Jacobi sweeps over ~25 entries per dependent
        const double nd = (double)pl.dep_rows.size();
        const double t_lev = nl * 3.0 + (double)pl.nnz * 12.0 / 4e6;
        const double t_spk = 2.0 * pl.steps.max_steps * 0.19 + (nd > 0 ? 7.0 * (2.5 + nd * 25.0 * 12.0 / 4e6) : 0.0);
        int method = f->opts.method;
        if (!f->mats.empty() || f->caller_parts) method = SPIKE_METHOD_SPIKE;
        if (method == SPIKE_METHOD_AUTO)
            method = (nd > n / 8.0 || t_lev < t_spk) ? SPIKE_METHOD_LEVELS : SPIKE_METHOD_SPIKE;
        f->method = method;
        if (method == SPIKE_METHOD_LEVELS) {
            std::vector<int> ptr(nl + 1, 0), rows(n);
            for (int i = 0; i < n; i++) ptr[lev[i] + 1]++;
            for (int l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
            std::vector<int> fill(ptr.begin(), ptr.end() - 1);
            for (int i = 0; i < n; i++) rows[fill[lev[i]]++] = i;
            f->lev_ptr = ptr;
            return build_levels(f, rows, t0);
        }
    }

    // ---- per-partition records, dependents, coupling rows
    std::vector<DevPart> hp(P);
    for (int t = 0, mi = 0; t < P; t++) {
        const spike::Part &pt = pl.parts[t];
        const int ce0 = pt.ccnt > 0 ? pl.cpl_ptr[pt.cbeg] : 0, ce1 = pt.ccnt > 0 ? pl.cpl_ptr[pt.cbeg + pt.ccnt] : 0;
        hp[t] = DevPart{pt.s, pt.k, pt.f, pt.hi1, pt.dbeg, pt.dcnt, pt.cbeg, pt.ccnt,
                        0, n, pl.flip ? 1 : 0, 0, 0, ce0, ce1, 0, 0};
        if (!f->mats.empty()) {
            while (mi + 1 < (int)f->mats.size() && f->mats[mi].row0 + f->mats[mi].n <= pt.s) mi++;
            const spike_factor_s::Mat &m = f->mats[mi];
            if (pt.s < m.row0 || pt.s + pt.k > m.row0 + m.n)
                return fail(SPIKE_ERR_ARG, "a partition straddles two matrices of the batch");
            hp[t].row0 = m.row0;
            hp[t].mn = m.n;
            hp[t].mflip = f->batch_flip ? 1 : 0;
            hp[t].ld = m.n;
            hp[t].mnr = m.nrhs;
            hp[t].boff = m.boff;
        }
    }
    if ((rc = f->upload(&f->d_parts, hp)) || (rc = f->upload(&f->d_dep_rows, pl.dep_rows)) ||
        (rc = f->upload(&f->d_cpl_rows, pl.cpl_rows)) || (rc = f->upload(&f->d_cpl_ptr, pl.cpl_ptr)) ||
        (rc = f->upload(&f->d_cpl_col, pl.cpl_col)) || (rc = f->upload(&f->d_cpl_val, pl.cpl_val)))
        return rc;
    if ((rc = upload_steps(f, pl.steps, f->steps, f->fp32)) || (rc = f->upload(&f->d_inv_diag, pl.inv_diag)))
        return rc;

    const int nred = (int)pl.dep_rows.size();
    for (int t = 0; t < P; t++) {
        const spike::Part &pt = pl.parts[t];
        f->cmax = std::max(f->cmax, pl.cpl_ptr[pt.cbeg + pt.ccnt] - pl.cpl_ptr[pt.cbeg]);
    }
    choose_rhs_slots(f);
    if (f->batch_slots > 0 && f->max_nr != f->batch_slots)
        return fail(SPIKE_ERR_UNSUPPORTED, "partitions too large for " + std::to_string(f->batch_slots) +
                                               " right-hand sides in LDS; use fewer rows per partition");
    if ((rc = f->dalloc(&f->d_y, (size_t)nred * f->max_nr)) || (rc = f->dalloc(&f->d_xa, (size_t)nred * f->max_nr)) ||
        (rc = f->dalloc(&f->d_xb, (size_t)nred * f->max_nr)) ||
        (rc = f->dalloc(&f->d_bc, pl.cpl_rows.size() * f->max_nr)))
        return rc;
    // (a batch shares one residual workspace across its groups)
    if (f->fp32 && f->mats.empty() && (rc = f->dalloc(&f->d_r, (size_t)n * f->max_nr))) return rc;

    // lower-form CSR on the device: the spike build reads it, the FP32 mode
    // keeps it for the FP64 residual
    int *d_ia = nullptr, *d_ja = nullptr;
    double *d_a = nullptr;
    if (nred > 0 || f->fp32) {
        if ((rc = f->upload(&d_ia, pl.ia)) || (rc = f->upload(&d_ja, pl.ja)) || (rc = f->upload(&d_a, pl.a)))
            return rc;
        if (f->fp32) {
            f->d_a = d_a;
            f->d_ia = d_ia;
            f->d_ja = d_ja;
        }
    }

    // ---- spikes on the device (layout: plan_spike_layout)
    const char *gfe = std::getenv("SPIKE_PSTRSV_SPIKE_GFAR");
    const spike::SpikeLayout sl = spike::plan_spike_layout(pl, spike::block_solve_lds_limit(), spike::spike_ring_lds,
                                                           spike::spike_ring_stage_entries, !(gfe && std::atoi(gfe) == 0));
    std::vector<spike::SpikePart> sp;
    for (int t = 0; t < P; t++) {
        const spike::Part &pt = pl.parts[t];
        if (sl.slot0[t] >= 0) sp.push_back(spike::SpikePart{pt.s, pt.f, pt.rows_S, pt.kbeg, pt.kcnt, sl.slot0[t], sl.woff[t]});
    }
    f->spike_ring = sl.ring;
    std::vector<int> hrp(nred + 1, 0);
    std::vector<int> hci;
    std::vector<double> hv;
    double dropmax = 0.0;
    int64_t rnnz = 0;
    double t2 = now_s(), t3 = t2;
    if (nred > 0) {
        spike::SpikePart *d_sp;
        int *d_kcols, *d_rinfo, *d_pia = nullptr, *d_pcode = nullptr;
        double *d_W, *d_pval = nullptr;
        const size_t npad = sl.ring ? (size_t)sl.pia.back() : 0;
        if ((rc = f->upload(&d_sp, sp)) || (rc = f->upload(&d_kcols, pl.kcols)) ||
            (rc = f->upload(&d_rinfo, sl.rinfo)) || (rc = f->dalloc(&d_W, (size_t)std::max<int64_t>(sl.area, 1))))
            return rc;
        if (sl.ring) {
            if ((rc = f->upload(&d_pia, sl.pia)) || (rc = f->dalloc(&d_pcode, npad)) || (rc = f->dalloc(&d_pval, npad)))
                return rc;
            HIP_TRY(spike::launch_spike_codes((int)sp.size(), sl.max_rows, d_sp, d_kcols, d_ia, d_ja, d_a, d_rinfo, d_pia,
                                              sl.R, sl.kstride, sl.gfar, d_pcode, d_pval, nullptr));
            HIP_TRY(spike::launch_spike_ring((int)sp.size(), sl.max_kcnt, sl.block_threads, d_sp, d_pia, d_pcode, d_pval,
                                             d_rinfo, sl.R, sl.nlr, sl.kstride, sl.sb, sl.cap, sl.gfar, d_W, nullptr));
        } else {
            HIP_TRY(spike::launch_spike_direct((int)sp.size(), sl.max_kcnt, d_sp, d_kcols, d_ia, d_ja, d_a, d_W, nullptr));
        }
        HIP_TRY(hipDeviceSynchronize());
        t2 = now_s();
        g_timers[2] += t2 - t1;

        // ---- reduced-system assembly on the device
        std::vector<int64_t> wrow(nred, -1);
        std::vector<int> qk(nred, 0), qkb(nred, 0);
        std::vector<double> tol(nred, -1.0);
        std::vector<int> kridx(pl.kcols.size());
        for (size_t j = 0; j < pl.kcols.size(); j++) kridx[j] = pl.ridx[pl.kcols[j]];
        const double dt = f->opts.drop_tol < 0 ? 1e-15 : f->opts.drop_tol;
        for (int t = 0; t < P; t++) {
            const spike::Part &pt = pl.parts[t];
            for (int q = pt.dbeg; q < pt.dbeg + pt.dcnt; q++) {
                int r = pl.dep_rows[q] - pt.s;
                if (sl.woff[t] >= 0 && r >= pt.f && r <= pt.g) {
                    wrow[q] = sl.woff[t] + (int64_t)sl.rinfo[4 * ((size_t)sl.slot0[t] + r - pt.f)] * pt.kcnt;
                    qk[q] = pt.kcnt;
                    qkb[q] = pt.kbeg;
                    tol[q] = dt > 0 ? dt / pt.kcnt : -1.0;
                }
            }
        }
        int64_t *d_wrow;
        int *d_qk, *d_qkb, *d_kridx, *d_cnt;
        double *d_tol, *d_drop;
        if ((rc = f->upload(&d_wrow, wrow)) || (rc = f->upload(&d_qk, qk)) || (rc = f->upload(&d_qkb, qkb)) ||
            (rc = f->upload(&d_tol, tol)) || (rc = f->upload(&d_kridx, kridx)) || (rc = f->dalloc(&d_cnt, nred)) ||
            (rc = f->dalloc(&d_drop, nred)))
            return rc;
        HIP_TRY(spike::launch_reduced_count(nred, d_wrow, d_qk, d_tol, d_W, d_cnt, d_drop, nullptr));
        std::vector<int> cnt(nred);
        std::vector<double> drop(nred);
        HIP_TRY(hipMemcpy(cnt.data(), d_cnt, sizeof(int) * nred, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(drop.data(), d_drop, sizeof(double) * nred, hipMemcpyDeviceToHost));
        for (int q = 0; q < nred; q++) {
            rnnz += cnt[q];
            if (rnnz > INT32_MAX) return fail(SPIKE_ERR_UNSUPPORTED, "reduced system exceeds 2^31-1 entries");
            hrp[q + 1] = (int)rnnz;
            dropmax = std::max(dropmax, drop[q]);
        }
        if ((rc = f->upload(&f->d_rrp, hrp)) || (rc = f->dalloc(&f->d_rci, (size_t)rnnz)) ||
            (rc = f->dalloc(&f->d_rv, (size_t)rnnz)))
            return rc;
        HIP_TRY(spike::launch_reduced_fill(nred, d_wrow, d_qkb, d_qk, d_tol, d_W, d_kridx, f->d_rrp, f->d_rci,
                                           f->d_rv, nullptr));
        hci.resize(rnnz);
        hv.resize(rnnz);
        if (rnnz) {
            HIP_TRY(hipMemcpy(hci.data(), f->d_rci, sizeof(int) * rnnz, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(hv.data(), f->d_rv, sizeof(double) * rnnz, hipMemcpyDeviceToHost));
        }
        HIP_TRY(hipDeviceSynchronize());
        dfree(f, d_W, sizeof(double) * std::max<int64_t>(sl.area, 1));
        if (d_pcode) {
            dfree(f, d_pval, sizeof(double) * npad);
            dfree(f, d_pcode, sizeof(int) * npad);
            dfree(f, d_pia, sizeof(int) * sl.pia.size());
        }
        dfree(f, d_rinfo, sizeof(int) * std::max<size_t>(sl.rinfo.size(), 1));
        dfree(f, d_kcols, sizeof(int) * std::max<size_t>(pl.kcols.size(), 1));
        dfree(f, d_sp, sizeof(spike::SpikePart) * std::max<size_t>(sp.size(), 1));
        if (!f->fp32) {
            dfree(f, d_a, sizeof(double) * pl.a.size());
            dfree(f, d_ja, sizeof(int) * pl.ja.size());
            dfree(f, d_ia, sizeof(int) * pl.ia.size());
        }
        dfree(f, d_wrow, sizeof(int64_t) * nred);
        dfree(f, d_cnt, sizeof(int) * nred);
        dfree(f, d_drop, sizeof(double) * nred);
        t3 = now_s();
        g_timers[3] += t3 - t2;
    }

    // ---- level schedule of the reduced system (rows with off-diagonals)
    std::vector<int> lev = spike::row_levels(nred, hrp.data(), hci.data());
    int maxlev = 0;
    for (int q = 0; q < nred; q++) maxlev = std::max(maxlev, lev[q]);
    std::vector<int> levptr(maxlev + 1, 0), rlist;
    for (int q = 0; q < nred; q++)
        if (lev[q] > 0) levptr[lev[q]]++;
    for (int L = 1; L <= maxlev; L++) levptr[L] += levptr[L - 1];
    rlist.resize(levptr[maxlev]);
    {
        std::vector<int> fill(levptr.begin(), levptr.end() - 1);
        for (int q = 0; q < nred; q++)
            if (lev[q] > 0) rlist[fill[lev[q] - 1]++] = q;
    }
    f->nlev = maxlev;
    if ((rc = f->upload(&f->d_levptr, levptr)) || (rc = f->upload(&f->d_rlist, rlist))) return rc;
    f->red_in_lds = (size_t)nred * sizeof(double) <= spike::block_solve_lds_limit();

    // Jacobi sweep count: after S sweeps from x = y the error is
    // sum_{j>S} (-W)^j y, bounded entry-wise by ||y||_inf sum_{j>S} mx_j with
    // mx_j = max_i (|W|^j 1)_i.  mx is submultiplicative (mx_{a+b} <=
    // mx_a mx_b for the nonnegative |W|), so with J computed terms and
    // mx_J < 1 the terms beyond J sum to at most
    // mx_J / (1 - mx_J) * (1 + mx_1 + ... + mx_{J-1}).  W is nilpotent, so
    // terms past its level count vanish exactly.
    int sweeps = -1, sweeps32 = -1;
    double jbound = 0.0;
    std::vector<int> row_sweeps(nred, 0);    // sweeps each row needs
    if (rnnz > 0) {
        std::vector<std::vector<double>> T;  // T[j-1][q] = (|W|^j 1)_q
        std::vector<double> r(nred, 1.0);
        std::vector<double> mx;              // mx[j-1] = max_q T[j-1][q]
        bool exact_end = false;
        // row chunks of about 2^16 entries on the host threads
        const int64_t grain = std::max<int64_t>(256, (int64_t)nred * 65536 / std::max<int64_t>(rnnz, 1));
        std::vector<double> cmax((size_t)spike::chunk_count(nred, grain));
        for (int j = 1; j <= 64; j++) {
            std::vector<double> r2;
            spike::resize_large(r2, nred);
            std::fill(cmax.begin(), cmax.end(), 0.0);
            spike::parallel_chunks(nred, grain, [&](int c, int64_t lo, int64_t hi) {
                double cm = 0.0;
                for (int64_t q = lo; q < hi; q++) {
                    double acc = 0.0;
                    for (int e = hrp[q]; e < hrp[q + 1]; e++) acc += std::fabs(hv[e]) * r[hci[e]];
                    r2[q] = acc;
                    cm = std::max(cm, acc);
                }
                cmax[c] = cm;
            });
            double m = 0.0;
            for (double v : cmax) m = std::max(m, v);
            mx.push_back(m);
            T.push_back(r2);
            r.swap(r2);
            if (m == 0.0 || j >= maxlev) { exact_end = true; break; }
            if (m < 1e-20) break;
        }
        const int J = (int)mx.size();
        double beyond = 0.0, mass = 0.0;
        for (int j = 0; j < J; j++) mass += mx[j];
        if (!exact_end) {
            double partial = 1.0;
            for (int j = 0; j < J - 1; j++) partial += mx[j];
            beyond = mx[J - 1] < 1.0 ? mx[J - 1] / (1.0 - mx[J - 1]) * partial : 1e300;
            mass += beyond;
        }
        // smallest S with sum_{j>S} mx_j <= kTol: the truncation of the
        // reduced system already allows 1e-15 (row sum of dropped |w|), so
        // iterating to a tail bound far below that gains nothing; 2e-15
        // keeps the reduced solve within about 3e-15 of ||y|| (on the 1M-row
        // banded matrix it drops a seventh sweep of 320 rows, 4.5 us, and
        // leaves the solution's error at 9.3e-16 unchanged)
        constexpr double kTol = 2e-15;
        {
            // FP32 factors: the smallest S whose tail bound is <= 1e-8
            double t = beyond;
            for (int j = 0; j < J; j++) t += mx[j];
            sweeps32 = J;
            for (int S = 0; S < J; S++) {
                if (t <= 1e-8) {
                    sweeps32 = S;
                    break;
                }
                t -= mx[S];
            }
            if (beyond > 1e-8) sweeps32 = -1;
        }
        double tail = beyond;
        if (tail <= kTol) {
            sweeps = J;
            jbound = tail;
            for (int S = J - 1; S >= 0; S--) {
                tail += mx[S];           // mx[S] is the term j = S+1
                if (tail > kTol) break;
                sweeps = S;
                jbound = tail;
            }
            // per row: terms past J are bounded by T_J[q] * mass
            spike::parallel_chunks(nred, 1 << 16, [&](int, int64_t lo, int64_t hi) {
                for (int64_t q = lo; q < hi; q++) {
                    double t = exact_end ? 0.0 : T[J - 1][q] * mass;
                    int m = J;
                    for (int S = J - 1; S >= 0; S--) {
                        t += T[S][q];
                        if (t > kTol) break;
                        m = S;
                    }
                    row_sweeps[q] = std::min(m, sweeps);
                }
            });
        }
    }
    // cost model (microseconds): the level solve runs in one workgroup,
    // ~0.15 per level in LDS or ~1 in L2, plus ~4 ns per entry and RHS slot
    // (one workgroup streams the reduced system, measured); a sweep is ~2.5
    // launch + the reduced system streamed by the whole GPU at ~4 TB/s
    const double slots_avg = f->batch_slots > 0 ? f->batch_slots : 1.0;
    const double t_exact = maxlev * (f->red_in_lds ? 0.15 : 1.0) + (double)rnnz * slots_avg * 4e-3;
    const double t_jacobi = sweeps < 0 ? 1e30 : sweeps * (2.5 + (double)rnnz * 12.0 / 4e6);
    int mode = f->opts.reduced_solver;
    if (mode == SPIKE_REDUCED_JACOBI && sweeps < 0)
        return fail(SPIKE_ERR_UNSUPPORTED, "Jacobi sweeps do not reach the reduced-system tolerance in 64 terms");
    if (mode == SPIKE_REDUCED_AUTO) mode = (rnnz > 0 && t_jacobi < t_exact) ? SPIKE_REDUCED_JACOBI : SPIKE_REDUCED_EXACT;
    f->sweeps = (mode == SPIKE_REDUCED_JACOBI && rnnz > 0) ? std::max(sweeps, 1) : 0;
    f->sweeps_fp32 = f->sweeps;
    if (f->fp32 && f->sweeps > 0 && sweeps32 >= 0) f->sweeps_fp32 = std::max(1, std::min(sweeps32, f->sweeps));
    if (f->sweeps > 0) {
        // sweep m computes rows that need >= m sweeps and copies rows that
        // finished at m-1 (rows finished at 0 are copied in sweeps 1 and 2),
        // so both buffers hold every finished row's value
        std::vector<std::vector<int>> per(f->sweeps + 1);
        spike::parallel_for(f->sweeps, [&](int mi) {
            const int m = mi + 1;
            std::vector<int> &l = per[m];
            for (int q = 0; q < nred; q++) {
                const int need = row_sweeps[q];
                int code = -1;
                if (need >= m) code = q;
                else if (need == m - 1 || (need == 0 && m == 2)) code = q | (int)0x80000000;
                if (code == -1) continue;
                l.insert(l.end(), {code, hrp[q], hrp[q + 1], 0});
            }
        });
        std::vector<int> lists;
        f->sweep_off.assign(f->sweeps + 1, 0);
        for (int m = 1; m <= f->sweeps; m++) {
            lists.insert(lists.end(), per[m].begin(), per[m].end());
            f->sweep_off[m] = (int)(lists.size() / 4);
        }
        if ((rc = f->upload(&f->d_sweep_rows, lists))) return rc;
        if ((rc = build_jacobi_items(f, hrp, hci, row_sweeps))) return rc;
    }

    // Single-launch solve: with Jacobi sweeps (or a reduced system without
    // off-diagonals) and the pass-2 coupling products in LDS.  Sweep items are
    // runs of consecutive reduced rows of about rnnz / (4 x CUs) entries (at
    // most 2048, one batch of a wave at one RHS) and kFusedRowsPerItem rows.
    // Opt-in (SPIKE_PSTRSV_FUSED=1): measured slower than the launch sequence
    // on a 1M-row banded matrix (0.29 against 0.13 ms per solve), and one run
    // of an FP32 factor ended in a device memory fault whose cause is not
    // identified; the launch sequence stays the default.
    {
        const char *e = std::getenv("SPIKE_PSTRSV_FUSED");
        f->fused = e && std::atoi(e) == 1 && nred > 0 && (f->sweeps > 0 || rnnz == 0) && f->cpl_fused &&
                   !pl.steps.fwd && !f->fp32;
        if (f->fused) {
            const int64_t target =
                std::min<int64_t>(2048, std::max<int64_t>(256, (rnnz + 4LL * f->ncu - 1) / (4LL * f->ncu)));
            std::vector<int> ch;
            for (int q = 0; q < nred;) {
                const int q0 = q;
                int64_t ent = 0;
                while (q < nred && q - q0 < spike::kFusedRowsPerItem &&
                       (q == q0 || ent + (hrp[q + 1] - hrp[q]) <= target)) {
                    ent += hrp[q + 1] - hrp[q];
                    q++;
                }
                ch.insert(ch.end(), {q0, q, hrp[q0], hrp[q]});
            }
            f->nchunks = (int)(ch.size() / 4);
            read_spin_limit(f);
            if ((rc = f->upload(&f->d_chunks, ch)) || (rc = setup_fail_word(f))) return rc;
            if (!f->d_abort) {
                if ((rc = f->dalloc(&f->d_abort, 1))) return rc;
                HIP_TRY(hipMemset(f->d_abort, 0, sizeof(int)));
            }
        }
    }

    // host copies no longer needed
    pl.a.clear(); pl.a.shrink_to_fit();
    pl.ja.clear(); pl.ja.shrink_to_fit();
    pl.steps.val.clear(); pl.steps.val.shrink_to_fit(); pl.steps.col.clear(); pl.steps.col.shrink_to_fit();
    pl.steps.late.clear(); pl.steps.late.shrink_to_fit(); pl.inv_diag.clear(); pl.inv_diag.shrink_to_fit();
    pl.cpl_val.clear(); pl.cpl_val.shrink_to_fit();

    spike_stats_t &st = f->stats;
    st.n = n;
    st.partitions = P;
    st.nnz = pl.nnz;
    st.n_reduced = nred;
    st.nnz_reduced = rnnz + nred;
    st.reduced_levels = maxlev;
    st.spike_area = pl.spike_area;
    st.dropped_rowsum_max = dropmax;
    st.steps_pass1 = pl.steps.max_steps;
    st.steps_pass2 = pl.steps.max_steps;
    st.lanes_per_row = pl.steps.G;
    st.reduced_sweeps = f->sweeps;
    st.jacobi_bound = f->sweeps > 0 ? jbound : 0.0;
    st.rhs_per_launch = f->max_nr;
    st.precision = f->opts.precision;
    st.method = SPIKE_METHOD_SPIKE;
    st.levels = f->nlev_full;
    if (f->batch_slots == 0 && f->opts.autotune && !f->steps.fwd && (rc = tune_tiles(f))) return rc;
    for (int li = 0; li < 4; li++) st.rhs_tile[li] = f->tile_at[li];
    st.single_launch = f->fused ? 1 : 0;
    if (f->fp32) {
        if ((rc = upload_row_blocks(f, hp))) return rc;
        if ((rc = calibrate_refinement(f))) return rc;
    }
    st.refine_steps = f->refine_steps;
    st.refine_rate = f->refine_rate;
    st.refine_est_err = f->refine_est;
    st.device_bytes = f->device_bytes;
    st.preproc_seconds = now_s() - t0;
    g_timers[0] += st.preproc_seconds;
    if (f->opts.verbose) spike_factor_log(f);
    return SPIKE_OK;
}

// ---------------------------------------------------------------- serialization

// File layout: magic, a block of scalars, then tagged records
// {tag (4 chars), byte count (8), bytes}.  Host vectors and device arrays are
// stored by content; workspaces only by size (re-allocated on load).
constexpr char kMagic[8] = {'S', 'P', 'K', 'F', 'A', 'C', '0', '8'};

struct FactorScalars {
    spike_options_t opts;
    spike_stats_t stats;
    char uplo;
    int flip, n, P, kmax, nlev, red_in_lds, sweeps, ring, cpl_fused, fuse_at[4], tile_at[4], max_nr, cmax;
    int fp32, refine_steps, method, caller_parts, nlev_full, lev_G, nres, G, M, R, fwd, fused, nchunks;
    int ji_min_slots, ji_kmax, sweeps_fp32;
    int64_t ji_rnnz;
    int ji_force;
    int64_t nnz, spike_area;
    double refine_rate, refine_est;
    uint64_t lds_needed;
};

// The device arrays of a factor, by tag.
std::vector<std::pair<const char *, void **>> device_fields(spike_factor_s *f)
{
    return {{"part", (void **)&f->d_parts},          {"deps", (void **)&f->d_dep_rows},
            {"crow", (void **)&f->d_cpl_rows},       {"cptr", (void **)&f->d_cpl_ptr},
            {"ccol", (void **)&f->d_cpl_col},        {"cval", (void **)&f->d_cpl_val},
            {"sval", (void **)&f->steps.val},
            {"scol", (void **)&f->steps.col},        {"srow", (void **)&f->steps.row},
            {"slat", (void **)&f->steps.late},
            {"spb_", (void **)&f->steps.part_begin}, {"sv32", (void **)&f->steps.val32},
            {"invd", (void **)&f->d_inv_diag},
            {"rrp_", (void **)&f->d_rrp},            {"rci_", (void **)&f->d_rci},
            {"rv__", (void **)&f->d_rv},             {"rlst", (void **)&f->d_rlist},
            {"rlev", (void **)&f->d_levptr},         {"swrw", (void **)&f->d_sweep_rows},
            {"a___", (void **)&f->d_a},              {"ia__", (void **)&f->d_ia},
            {"ja__", (void **)&f->d_ja},             {"lrec", (void **)&f->d_lrec},
            {"lja_", (void **)&f->d_lja},            {"lval", (void **)&f->d_lval},
            {"lv32", (void **)&f->d_lval32},         {"id32", (void **)&f->d_invd32},
            {"res_", (void **)&f->d_res},            {"fchk", (void **)&f->d_chunks},
            {"jiit", (void **)&f->d_ji_items},       {"jikc", (void **)&f->d_ji_kcol},
            {"jind", (void **)&f->d_ji_need},        {"jilc", (void **)&f->d_ji_lci},
            {"jils", (void **)&f->d_ji_lists}};
}
std::vector<std::pair<const char *, double **>> workspace_fields(spike_factor_s *f)
{
    return {{"w_y_", &f->d_y}, {"w_xa", &f->d_xa}, {"w_xb", &f->d_xb},
            {"w_bc", &f->d_bc}, {"w_r_", &f->d_r},  {"w_dl", &f->d_dlt}};
}

size_t alloc_size(const spike_factor_s *f, const void *p)
{
    for (size_t i = 0; i < f->allocs.size(); i++)
        if (f->allocs[i] == p) return f->alloc_bytes[i];
    return 0;
}

int write_rec(FILE *fp, const char *tag, const void *data, uint64_t bytes)
{
    if (std::fwrite(tag, 1, 4, fp) != 4 || std::fwrite(&bytes, sizeof bytes, 1, fp) != 1 ||
        (bytes && std::fwrite(data, 1, bytes, fp) != bytes))
        return fail(SPIKE_ERR_ARG, "write failed");
    return SPIKE_OK;
}

template <class T>
int write_vec(FILE *fp, const char *tag, const std::vector<T> &v)
{
    return write_rec(fp, tag, v.data(), v.size() * sizeof(T));
}

int factor_save(spike_factor_s *f, const char *path)
{
    if (!f->mats.empty()) return fail(SPIKE_ERR_UNSUPPORTED, "batch groups are not saved");
    const Plan &pl = f->plan;
    FactorScalars sc{};
    sc.opts = f->opts;
    sc.stats = f->stats;
    sc.uplo = pl.uplo;
    sc.flip = pl.flip;
    sc.n = pl.n;
    sc.P = pl.P;
    sc.kmax = pl.kmax;
    sc.nlev = f->nlev;
    sc.red_in_lds = f->red_in_lds;
    sc.sweeps = f->sweeps;
    sc.ring = f->ring;
    sc.cpl_fused = f->cpl_fused;
    for (int i = 0; i < 4; i++) {
        sc.fuse_at[i] = f->fuse_at[i];
        sc.tile_at[i] = f->tile_at[i];
    }
    sc.max_nr = f->max_nr;
    sc.cmax = f->cmax;
    sc.fp32 = f->fp32;
    sc.refine_steps = f->refine_steps;
    sc.method = f->method;
    sc.caller_parts = f->caller_parts;
    sc.nlev_full = f->nlev_full;
    sc.lev_G = f->lev_G;
    sc.nres = f->nres;
    sc.G = f->steps.G;
    sc.M = f->steps.M;
    sc.R = f->steps.R;
    sc.fwd = f->steps.fwd;
    sc.fused = f->fused;
    sc.nchunks = f->nchunks;
    sc.ji_min_slots = f->ji_min_slots;
    sc.ji_kmax = f->ji_kmax;
    sc.sweeps_fp32 = f->sweeps_fp32;
    sc.ji_rnnz = f->ji_rnnz;
    sc.ji_force = f->ji_force;
    sc.nnz = pl.nnz;
    sc.spike_area = pl.spike_area;
    sc.refine_rate = f->refine_rate;
    sc.refine_est = f->refine_est;
    {
        const size_t vb = f->fp32 ? sizeof(float) : sizeof(double);
        int maxt = 1;
        for (int i = 0; i < 4; i++) maxt = std::max(maxt, f->tile_at[i]);
        sc.lds_needed = f->method == SPIKE_METHOD_LEVELS ? 0 : (uint64_t)(pl.kmax + spike::XS_PAD) * vb * maxt;
    }
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(SPIKE_ERR_ARG, std::string("cannot open ") + path + " for writing");
    int rc = SPIKE_OK;
    if (std::fwrite(kMagic, 1, 8, fp) != 8 || std::fwrite(&sc, sizeof sc, 1, fp) != 1) rc = fail(SPIKE_ERR_ARG, "write failed");
    if (!rc) rc = write_vec(fp, "hprt", pl.parts);
    if (!rc) rc = write_vec(fp, "hdep", pl.dep_rows);
    if (!rc) rc = write_vec(fp, "hcpl", pl.cpl_rows);
    if (!rc) rc = write_vec(fp, "hsta", pl.starts);
    if (!rc) rc = write_vec(fp, "hswo", f->sweep_off);
    if (!rc) rc = write_vec(fp, "hjio", f->ji_off);
    if (!rc) rc = write_vec(fp, "hlvp", f->lev_ptr);
    std::vector<char> buf;
    for (auto &fld : device_fields(f)) {
        if (rc) break;
        const void *p = *fld.second;
        if (!p) continue;
        const size_t bytes = alloc_size(f, p);
        if (!bytes) {
            rc = fail(SPIKE_ERR_STATE, std::string("unknown size of device array ") + fld.first);
            break;
        }
        buf.resize(bytes);
        if (hipMemcpy(buf.data(), p, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = fail(SPIKE_ERR_DEVICE, "device read failed");
            break;
        }
        rc = write_rec(fp, fld.first, buf.data(), bytes);
    }
    for (auto &fld : workspace_fields(f)) {
        if (rc) break;
        const uint64_t bytes = *fld.second ? alloc_size(f, *fld.second) : 0;
        if (bytes) rc = write_rec(fp, fld.first, &bytes, sizeof bytes);
    }
    if (!rc) rc = write_rec(fp, "end_", nullptr, 0);
    if (std::fclose(fp) != 0 && !rc) rc = fail(SPIKE_ERR_ARG, "close failed");
    return rc;
}

int factor_load(spike_factor_s *f, const char *path)
{
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(SPIKE_ERR_ARG, std::string("cannot open ") + path);
    struct Closer {
        FILE *fp;
        ~Closer() { std::fclose(fp); }
    } closer{fp};
    char mg[8];
    FactorScalars sc{};
    if (std::fread(mg, 1, 8, fp) != 8 || std::memcmp(mg, kMagic, 8) != 0)
        return fail(SPIKE_ERR_ARG, "not a saved spike_pstrsv factor (bad magic)");
    if (std::fread(&sc, sizeof sc, 1, fp) != 1) return fail(SPIKE_ERR_ARG, "truncated factor file");
    if (sc.lds_needed > spike::block_solve_lds_limit())
        return fail(SPIKE_ERR_UNSUPPORTED, "factor needs more LDS per workgroup than this device has");
    HIP_TRY(hipGetDevice(&f->device));
    Plan &pl = f->plan;
    f->opts = sc.opts;
    f->stats = sc.stats;
    pl.uplo = sc.uplo;
    pl.flip = sc.flip != 0;
    pl.n = sc.n;
    pl.P = sc.P;
    pl.kmax = sc.kmax;
    pl.nnz = sc.nnz;
    pl.spike_area = sc.spike_area;
    f->nlev = sc.nlev;
    f->red_in_lds = sc.red_in_lds != 0;
    f->sweeps = sc.sweeps;
    f->ring = sc.ring;
    f->cpl_fused = sc.cpl_fused != 0;
    for (int i = 0; i < 4; i++) {
        f->fuse_at[i] = sc.fuse_at[i] != 0;
        f->tile_at[i] = sc.tile_at[i];
    }
    f->max_nr = sc.max_nr;
    f->cmax = sc.cmax;
    f->fp32 = sc.fp32 != 0;
    f->refine_steps = sc.refine_steps;
    f->refine_rate = sc.refine_rate;
    f->refine_est = sc.refine_est;
    f->method = sc.method;
    f->caller_parts = sc.caller_parts != 0;
    f->nlev_full = sc.nlev_full;
    f->lev_G = sc.lev_G;
    f->nres = sc.nres;
    f->steps.G = sc.G;
    f->steps.M = sc.M;
    f->steps.R = sc.R;
    f->steps.fwd = sc.fwd;
    f->fused = sc.fused != 0;
    f->nchunks = sc.nchunks;
    f->ji_min_slots = sc.ji_min_slots;
    f->ji_kmax = sc.ji_kmax;
    f->sweeps_fp32 = sc.sweeps_fp32;
    f->ji_rnnz = sc.ji_rnnz;
    f->ji_force = sc.ji_force != 0;
    (void)hipDeviceGetAttribute(&f->ncu, hipDeviceAttributeMultiprocessorCount, f->device);
    (void)hipDeviceGetAttribute(&f->lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, f->device);
    (void)hipGetLastError();
    auto dev = device_fields(f);
    auto ws = workspace_fields(f);
    for (;;) {
        char tag[5] = {0};
        uint64_t bytes = 0;
        if (std::fread(tag, 1, 4, fp) != 4 || std::fread(&bytes, sizeof bytes, 1, fp) != 1)
            return fail(SPIKE_ERR_ARG, "truncated factor file");
        if (std::strcmp(tag, "end_") == 0) break;
        if (bytes > ((uint64_t)1 << 40)) return fail(SPIKE_ERR_ARG, "corrupt factor file");
        std::vector<char> buf(bytes);
        if (bytes && std::fread(buf.data(), 1, bytes, fp) != bytes) return fail(SPIKE_ERR_ARG, "truncated factor file");
        auto host = [&](auto &vec) {
            using T = typename std::decay_t<decltype(vec)>::value_type;
            vec.resize(bytes / sizeof(T));
            if (bytes) std::memcpy(vec.data(), buf.data(), bytes);
        };
        if (!std::strcmp(tag, "hprt")) { host(pl.parts); continue; }
        if (!std::strcmp(tag, "hdep")) { host(pl.dep_rows); continue; }
        if (!std::strcmp(tag, "hcpl")) { host(pl.cpl_rows); continue; }
        if (!std::strcmp(tag, "hsta")) { host(pl.starts); continue; }
        if (!std::strcmp(tag, "hswo")) { host(f->sweep_off); continue; }
        if (!std::strcmp(tag, "hjio")) { host(f->ji_off); continue; }
        if (!std::strcmp(tag, "hlvp")) { host(f->lev_ptr); continue; }
        bool done = false;
        for (auto &fld : dev)
            if (!std::strcmp(tag, fld.first)) {
                char *p = nullptr;
                int rc = f->dalloc(&p, (size_t)bytes);
                if (rc) return rc;
                HIP_TRY(hipMemcpy(p, buf.data(), bytes, hipMemcpyHostToDevice));
                *fld.second = p;
                done = true;
                break;
            }
        for (auto &fld : ws)
            if (!done && !std::strcmp(tag, fld.first)) {
                uint64_t wb = 0;
                std::memcpy(&wb, buf.data(), sizeof wb);
                int rc = f->dalloc(fld.second, (size_t)(wb / sizeof(double)));
                if (rc) return rc;
                done = true;
                break;
            }
        if (!done) return fail(SPIKE_ERR_ARG, std::string("unknown record ") + tag);
    }
    if (f->method == SPIKE_METHOD_LEVELS)
        if (int rc = setup_sync_free(f)) return rc;
    if (f->fused) {
        read_spin_limit(f);
        if (int rc = setup_fail_word(f)) return rc;
        if (int rc = f->dalloc(&f->d_abort, 1)) return rc;
        HIP_TRY(hipMemset(f->d_abort, 0, sizeof(int)));
    }
    f->stats.device_bytes = f->device_bytes;
    return SPIKE_OK;
}

bool is_device_ptr(const void *p)
{
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.isManaged;
}

// Launch arguments of one solve of nr right-hand sides (<= f->max_nr).
spike::BlockArgs make_args(spike_factor_s *f, const double *b, int ldb, double *x, int ldx, int nr, int accum)
{
    const Plan &pl = f->plan;
    int slots = f->batch_slots > 0 ? f->batch_slots : 1;
    while (slots < nr) slots <<= 1;
    spike::BlockArgs ba{};
    ba.parts = f->d_parts;
    ba.st = f->steps;
    ba.b = b;
    ba.x = x;
    ba.xred = f->d_y;
    ba.bc = f->d_bc;
    ba.dep_rows = f->d_dep_rows;
    ba.cpl_rows = f->d_cpl_rows;
    ba.n = pl.n;
    ba.flip = pl.flip ? 1 : 0;
    ba.kmax = pl.kmax;
    ba.ring = f->ring;
    ba.inv_diag = f->d_inv_diag;
    ba.cpl_fused = f->fuse_at[__builtin_ctz((unsigned)slots)] ? 1 : 0;
    ba.tile = std::max(1, f->tile_at[__builtin_ctz((unsigned)slots)]);
    ba.cmax = f->cmax;
    ba.cpl_ptr = f->d_cpl_ptr;
    ba.cpl_col = f->d_cpl_col;
    ba.cpl_val = f->d_cpl_val;
    ba.xr = f->d_y;
    ba.nr = nr;
    ba.nr_slots = slots;
    ba.ldb = ldb;
    ba.ldx = ldx;
    ba.fp32 = f->fp32 ? 1 : 0;
    ba.accum = accum;
    return ba;
}

// One Spike solve of nr right-hand sides (<= f->max_nr), all pointers on the
// device: x = M^-1 b, or x += M^-1 b with accum (refinement update).  ev
// (optional, 5 events) brackets the phases; their device times are added to
// ph[0..3].
// Interleaved workspaces for level solves of up to nrhs right-hand sides
// (allocated before any launch, so never inside a graph capture).
int ensure_level_ws(spike_factor_s *f, int nrhs)
{
    if (f->method != SPIKE_METHOD_LEVELS || nrhs < 2) return SPIKE_OK;
    int nrp = 2;
    while (nrp < std::min(nrhs, 32)) nrp <<= 1;
    const size_t need = (size_t)f->plan.n * nrp;
    if (f->il_cap >= need) return SPIKE_OK;
    if (f->d_bi) dfree(f, f->d_bi, sizeof(double) * f->il_cap);
    if (f->d_xi) dfree(f, f->d_xi, sizeof(double) * f->il_cap);
    f->d_bi = f->d_xi = nullptr;
    f->il_cap = 0;
    f->drop_graphs();
    int rc;
    if ((rc = f->dalloc(&f->d_bi, need)) || (rc = f->dalloc(&f->d_xi, need))) return rc;
    f->il_cap = need;
    return SPIKE_OK;
}

// Level arguments of the interleaved (multi-RHS) launches.
spike::LevelArgs level_il_args(const spike_factor_s *f)
{
    spike::LevelArgs la{};
    la.rec = f->d_lrec;
    la.lja = f->d_lja;
    la.lval = f->d_lval;
    la.invd = f->d_inv_diag;
    la.lval32 = f->d_lval32;
    la.invd32 = f->d_invd32;
    la.a64 = f->d_a;
    la.fp32 = f->fp32 ? 1 : 0;
    la.spin_limit = f->spin_limit;
    la.fail_word = f->d_fail;
    return la;
}

// xi = M^-1 bi on interleaved vectors of nrp values per row.
hipError_t level_il_solve(spike_factor_s *f, const spike::LevelArgs &la, int nrp, const double *bi, double *xi,
                          hipStream_t s)
{
    if (f->sync_free) return spike::launch_sync_free(f->lev_ptr.back(), f->plan.n, nrp, la, bi, xi, f->d_abort, s);
    for (int l = 0; l + 1 < (int)f->lev_ptr.size(); l++) {
        hipError_t e = spike::launch_level_il(f->lev_ptr[l], f->lev_ptr[l + 1] - f->lev_ptr[l], nrp, la, bi, xi, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int solve_levels(spike_factor_s *f, const double *b, int ldb, double *x, int ldx, int nr, int accum, hipStream_t s,
                 hipEvent_t *ev, double *ph)
{
    const int n = f->plan.n;
    if (nr >= 2) {
        // one lane per row and right-hand side on interleaved vectors
        int nrp = 2;
        while (nrp < nr) nrp <<= 1;
        if ((size_t)n * nrp > f->il_cap) return fail(SPIKE_ERR_STATE, "level workspaces not allocated");
        const spike::LevelArgs la = level_il_args(f);
        const int flip = f->plan.flip ? 1 : 0;
        if (ev) HIP_TRY(hipEventRecord(ev[0], s));
        HIP_TRY(spike::launch_interleave(n, nr, nrp, flip, b, ldb, f->d_bi, s));
        HIP_TRY(level_il_solve(f, la, nrp, f->d_bi, f->d_xi, s));
        HIP_TRY(spike::launch_deinterleave(n, nr, nrp, flip, f->d_xi, x, ldx, accum, s));
        if (ev) {
            HIP_TRY(hipEventRecord(ev[1], s));
            HIP_TRY(hipEventSynchronize(ev[1]));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
            ph[0] += ms;
        }
        return SPIKE_OK;
    }
    spike::LevelArgs la{};
    la.rec = f->d_lrec;
    la.lja = f->d_lja;
    la.lval = f->d_lval;
    la.invd = f->d_inv_diag;
    la.lval32 = f->d_lval32;
    la.invd32 = f->d_invd32;
    la.b = b;
    la.ldb = ldb;
    la.x = accum ? f->d_dlt : x;
    la.ldx = accum ? n : ldx;
    la.nr = nr;
    int slots = 1;
    while (slots < nr) slots <<= 1;
    la.nr_slots = slots;
    la.n = n;
    la.flip = f->plan.flip ? 1 : 0;
    la.G = f->lev_G;
    la.fp32 = f->fp32 ? 1 : 0;
    if (ev) HIP_TRY(hipEventRecord(ev[0], s));
    for (int l = 0; l + 1 < (int)f->lev_ptr.size(); l++)
        HIP_TRY(spike::launch_level(f->lev_ptr[l], f->lev_ptr[l + 1] - f->lev_ptr[l], la, s));
    if (accum) HIP_TRY(spike::launch_axpy(n, nr, f->d_dlt, n, x, ldx, s));
    if (ev) {
        HIP_TRY(hipEventRecord(ev[1], s));
        HIP_TRY(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        ph[0] += ms;
    }
    return SPIKE_OK;
}

// Workspace of the single-launch solve for solves of nr right-hand sides:
// the slot class's published vectors (two parities of S+1 vectors of
// nred x slots values) filled with the not-ready pattern, and its counters.
// Called before any launch that may be captured into a graph.
int prepare_fused(spike_factor_s *f, int nr)
{
    if (!f->fused || nr < 1) return SPIKE_OK;
    int slots = f->batch_slots > 0 ? f->batch_slots : 1;
    while (slots < nr) slots <<= 1;
    const int li = __builtin_ctz((unsigned)slots);
    if (li > 3 || f->d_fx[li]) return SPIKE_OK;
    const size_t len = 2 * (size_t)(f->sweeps + 1) * f->plan.dep_rows.size() * (size_t)slots;
    int rc;
    if ((rc = f->dalloc(&f->d_fx[li], len)) || (rc = f->dalloc(&f->d_fctr[li], 96))) return rc;
    HIP_TRY(spike::launch_fused_reset(f->d_fx[li], len, f->d_fctr[li], f->d_abort, nullptr));
    HIP_TRY(hipStreamSynchronize(nullptr));
    return SPIKE_OK;
}

// Whether the single-launch solve serves these launch arguments (workspace
// prepared, and the kernel fits the device at this tile).
bool use_fused(spike_factor_s *f, const spike::BlockArgs &ba)
{
    if (!f->fused || f->plan.dep_rows.empty()) return false;
    const int li = __builtin_ctz((unsigned)ba.nr_slots), ti = __builtin_ctz((unsigned)ba.tile);
    if (li > 3 || ti > 3 || !f->d_fx[li]) return false;
    if (f->fused_fit[li][ti] < 0) {
        spike::BlockArgs q = ba;
        q.cpl_fused = 1;
        f->fused_fit[li][ti] = spike::spike_fused_lds(q) > 0 ? 1 : 0;
    }
    return f->fused_fit[li][ti] == 1;
}

int solve_pass(spike_factor_s *f, const double *b, int ldb, double *x, int ldx, int nr, int accum, hipStream_t s,
               hipEvent_t *ev, double *ph)
{
    if (f->method == SPIKE_METHOD_LEVELS) return solve_levels(f, b, ldb, x, ldx, nr, accum, s, ev, ph);
    const Plan &pl = f->plan;
    spike::BlockArgs ba = make_args(f, b, ldb, x, ldx, nr, accum);
    const int slots = ba.nr_slots;
    if (!ev && use_fused(f, ba)) {
        // one launch: pass 1, the Jacobi sweeps and pass 2 with its coupling
        // update (profiled solves keep the launch sequence, whose phases can
        // be timed apart)
        const int li = __builtin_ctz((unsigned)slots);
        ba.cpl_fused = 1;
        spike::FusedArgs fa{};
        fa.X = f->d_fx[li];
        fa.ctr = f->d_fctr[li];
        fa.chunks = f->d_chunks;
        fa.nchunks = f->nchunks;
        fa.nred = (int)pl.dep_rows.size();
        fa.S = f->sweeps;
        fa.rp = f->d_rrp;
        fa.ci = f->d_rci;
        fa.v = f->d_rv;
        fa.spin_limit = f->spin_limit;
        fa.abort_word = f->d_abort;
        fa.fail_word = f->d_fail;
        HIP_TRY(spike::launch_spike_fused(pl.P, ba, fa, f->ncu, s));
        return SPIKE_OK;
    }
    if (ev) HIP_TRY(hipEventRecord(ev[0], s));
    HIP_TRY(spike::launch_block_solve(1, pl.P, ba, s));
    if (ev) HIP_TRY(hipEventRecord(ev[1], s));
    const int nred = (int)pl.dep_rows.size();
    if (nred > 0) {
        const double *xr = f->d_y;
        if (f->sweeps > 0) {
            const double *xin = f->d_y;
            // items where the sweeps are bandwidth-bound (at least 2^25
            // entry x RHS-slot products per sweep); small reduced systems
            // keep the row sweeps, whose launches have one round trip less
            const bool items = f->ji_min_slots > 0 && slots >= f->ji_min_slots &&
                               (f->ji_force || f->ji_rnnz * slots >= ((int64_t)1 << 25)) &&
                               spike::jacobi_items_lds(f->ji_kmax, slots) <= spike::block_solve_lds_limit();
            spike::JacobiItems ji{f->d_ji_items, f->d_ji_kcol, f->d_rrp, f->d_ji_lci, f->d_rv, f->d_ji_need, f->d_y,
                                  f->ji_kmax};
            const int S = f->fp32 ? f->sweeps_fp32 : f->sweeps;
            for (int m = 1; m <= S; m++) {
                double *xout = (m & 1) ? f->d_xa : f->d_xb;
                if (items) {
                    const int o0 = f->ji_off[m - 1], o1 = f->ji_off[m];
                    HIP_TRY(spike::launch_jacobi_items(ji, m, f->d_ji_lists + o0, o1 - o0, xin, xout, slots, s));
                } else {
                    const int o0 = f->sweep_off[m - 1], o1 = f->sweep_off[m];
                    HIP_TRY(spike::launch_jacobi_sweep(o1 - o0, f->d_sweep_rows + 4 * (size_t)o0, f->d_rci, f->d_rv,
                                                       f->d_y, xin, xout, slots, s));
                }
                xin = xout;
            }
            xr = xin;
        } else {
            const bool in_lds = (size_t)nred * slots * sizeof(double) <= spike::block_solve_lds_limit();
            HIP_TRY(spike::launch_reduced_solve(nred, f->nlev, f->d_levptr, f->d_rlist, f->d_rrp, f->d_rci,
                                                f->d_rv, f->d_y, in_lds, slots, s));
        }
        if (ev) HIP_TRY(hipEventRecord(ev[2], s));
        ba.xr = xr;
        if (!ba.cpl_fused)
            HIP_TRY(spike::launch_coupling(pl.P, f->d_parts, f->d_cpl_rows, f->d_cpl_ptr, f->d_cpl_col,
                                           f->d_cpl_val, xr, ba, f->d_bc, s));
        if (ev) HIP_TRY(hipEventRecord(ev[3], s));
        HIP_TRY(spike::launch_block_solve(2, pl.P, ba, s));
    } else if (ev) {
        HIP_TRY(hipEventRecord(ev[2], s));
        HIP_TRY(hipEventRecord(ev[3], s));
    }
    if (ev) {
        HIP_TRY(hipEventRecord(ev[4], s));
        HIP_TRY(hipEventSynchronize(ev[4]));
        for (int i = 0; i < 4; i++) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            ph[i] += ms;
        }
    }
    return SPIKE_OK;
}

// r = b - A x (FP64) into the factor's residual workspace.
int residual(spike_factor_s *f, const double *b, int ldb, const double *x, int ldx, int nr, hipStream_t s)
{
    spike::BlockArgs ra = make_args(f, b, ldb, const_cast<double *>(x), ldx, nr, 0);
    const int nr_max = f->mats.empty() ? nr : f->batch_max_nrhs;
    HIP_TRY(spike::launch_residual(f->nres, f->d_res, f->d_ia, f->d_ja, f->d_a, ra, nr_max, f->d_r, f->plan.n, s));
    return SPIKE_OK;
}

// Solve of nr right-hand sides with the factor's refinement: x_0 = M^-1 b,
// then refine_steps times r = b - A x (FP64), x += M^-1 r.  The residual's
// device time is added to ph[2].
// FP32 level solve of 2..32 right-hand sides with one refinement step, kept
// interleaved between its launches.  Sync-free: one launch computes x0, the
// FP64 residual and the correction row by row (packed into d_xi), and one
// de-interleave writes x0 + d.  One launch per level: x0 = M^-1 b into
// d_xi, r = b - A x0 (FP64) into d_r, d = M^-1 r into d_bi (b is not read
// after the residual), and x = x0 + d in one de-interleave.  Both give the
// same bits.
int solve_levels_refined1(spike_factor_s *f, const double *b, int ldb, double *x, int ldx, int nr, hipStream_t s)
{
    const int n = f->plan.n;
    int nrp = 2;
    while (nrp < nr) nrp <<= 1;
    if ((size_t)n * nrp > f->il_cap) return fail(SPIKE_ERR_STATE, "level workspaces not allocated");
    const spike::LevelArgs la = level_il_args(f);
    const int flip = f->plan.flip ? 1 : 0;
    HIP_TRY(spike::launch_interleave(n, nr, nrp, flip, b, ldb, f->d_bi, s));
    if (f->sync_free) {
        HIP_TRY(spike::launch_sync_free_refined1(f->lev_ptr.back(), n, nrp, la, f->d_bi, f->d_xi, f->d_abort, s));
        HIP_TRY(spike::launch_deinterleave_pairs(n, nr, nrp, flip, f->d_xi, x, ldx, s));
        return SPIKE_OK;
    }
    HIP_TRY(level_il_solve(f, la, nrp, f->d_bi, f->d_xi, s));
    HIP_TRY(spike::launch_residual_il(n, nrp, f->d_ia, f->d_ja, f->d_a, f->d_bi, f->d_xi, f->d_r, s));
    HIP_TRY(level_il_solve(f, la, nrp, f->d_r, f->d_bi, s));
    HIP_TRY(spike::launch_deinterleave_sum(n, nr, nrp, flip, f->d_xi, f->d_bi, x, ldx, s));
    return SPIKE_OK;
}

// FP32 level solve of one right-hand side with one refinement step: one
// launch per level computes x0 (into x), the FP64 residual and the
// correction (into d_dlt) row by row, then x += d.
int solve_levels_refined1_single(spike_factor_s *f, const double *b, double *x, hipStream_t s)
{
    spike::LevelArgs la = level_il_args(f);
    la.b = b;
    la.ldb = f->plan.n;
    la.x = x;
    la.ldx = f->plan.n;
    la.nr = 1;
    la.nr_slots = 1;
    la.n = f->plan.n;
    la.flip = f->plan.flip ? 1 : 0;
    la.G = f->lev_G;
    for (int l = 0; l + 1 < (int)f->lev_ptr.size(); l++)
        HIP_TRY(spike::launch_level_ref1(f->lev_ptr[l], f->lev_ptr[l + 1] - f->lev_ptr[l], la, f->d_dlt, s));
    HIP_TRY(spike::launch_axpy(f->plan.n, 1, f->d_dlt, f->plan.n, x, f->plan.n, s));
    return SPIKE_OK;
}

int solve_chunk(spike_factor_s *f, const double *b, int ldb, double *x, int ldx, int nr, hipStream_t s,
                hipEvent_t *ev, double *ph)
{
    if (f->method == SPIKE_METHOD_LEVELS && f->fp32 && f->refine_steps == 1 && nr == 1 && !ev && f->mats.empty() &&
        f->d_dlt)
        return solve_levels_refined1_single(f, b, x, s);
    // d_r holds n x 32 values on level factors with FP32 values
    if (f->method == SPIKE_METHOD_LEVELS && f->fp32 && f->refine_steps == 1 && nr >= 2 && nr <= 32 && !ev &&
        f->mats.empty() && f->d_r)
        return solve_levels_refined1(f, b, ldb, x, ldx, nr, s);
    int rc = solve_pass(f, b, ldb, x, ldx, nr, 0, s, ev, ph);
    if (rc) return rc;
    const int n = f->plan.n;
    for (int it = 0; it < f->refine_steps; it++) {
        if (ev) HIP_TRY(hipEventRecord(ev[0], s));
        if ((rc = residual(f, b, ldb, x, ldx, nr, s))) return rc;
        if (ev) {
            HIP_TRY(hipEventRecord(ev[1], s));
            HIP_TRY(hipEventSynchronize(ev[1]));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
            ph[2] += ms;
        }
        if ((rc = solve_pass(f, f->d_r, n, x, ldx, nr, 1, s, ev, ph))) return rc;
    }
    return SPIKE_OK;
}

// Refinement steps of an FP32 factor.  A random right-hand side (entries
// uniform in (-1, 1), SplitMix64; every column of every matrix of a batch) is
// solved with refinement steps added one at a time; c_j = max over matrices
// of max|x_j - x_{j-1}| / max|x_j| is the size of the j-th correction, which
// estimates the relative error left after step j-1.  The step count S is the
// smallest with c_{S+1} <= refine_tol.
int calibrate_refinement(spike_factor_s *f)
{
    const int n = f->plan.n;
    if (f->opts.refine_steps >= 0) {
        f->refine_steps = f->opts.refine_steps;
        return SPIKE_OK;
    }
    const double tol = f->opts.refine_tol > 0 ? f->opts.refine_tol : 1e-13;
    constexpr int kMax = 8;
    const bool bat = !f->mats.empty();
    const size_t len = bat ? (size_t)f->vec_len : (size_t)n;
    const int nr = bat ? f->batch_slots : 1;
    std::vector<std::pair<size_t, size_t>> seg;   // per matrix [begin, end) of its columns
    if (bat)
        for (const auto &m : f->mats) seg.emplace_back((size_t)m.boff, (size_t)m.boff + (size_t)m.n * m.nrhs);
    else
        seg.emplace_back(0, (size_t)n);
    std::vector<double> hb(len), hx(len), hprev(len);
    uint64_t st = 0x5350494B45434131ull;
    for (size_t i = 0; i < len; i++) {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        hb[i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
    }
    double *db = nullptr, *dx = nullptr;
    int rc;
    if ((rc = f->dalloc(&db, len)) || (rc = f->dalloc(&dx, len))) return rc;
    HIP_TRY(hipMemcpy(db, hb.data(), sizeof(double) * len, hipMemcpyHostToDevice));
    f->refine_steps = 0;
    if ((rc = prepare_fused(f, nr))) return rc;
    if ((rc = solve_pass(f, db, n, dx, n, nr, 0, nullptr, nullptr, nullptr))) return rc;
    HIP_TRY(hipMemcpy(hprev.data(), dx, sizeof(double) * len, hipMemcpyDeviceToHost));
    std::vector<double> c;
    int S = -1;
    for (int j = 1; j <= kMax && S < 0; j++) {
        if ((rc = residual(f, db, n, dx, n, nr, nullptr))) return rc;
        if ((rc = solve_pass(f, f->d_r, n, dx, n, nr, 1, nullptr, nullptr, nullptr))) return rc;
        HIP_TRY(hipMemcpy(hx.data(), dx, sizeof(double) * len, hipMemcpyDeviceToHost));
        double worst = 0.0;
        for (const auto &sg : seg) {
            double dm = 0.0, xm = 0.0;
            for (size_t i = sg.first; i < sg.second; i++) {
                dm = std::max(dm, std::fabs(hx[i] - hprev[i]));
                xm = std::max(xm, std::fabs(hx[i]));
            }
            const double cj = xm > 0 ? dm / xm : (dm > 0 ? HUGE_VAL : 0.0);
            worst = std::max(worst, cj);
        }
        c.push_back(worst);
        if (!std::isfinite(worst)) break;
        if (worst <= tol) S = j - 1;
        hprev.swap(hx);
    }
    dfree(f, db, sizeof(double) * len);
    dfree(f, dx, sizeof(double) * len);
    if (S < 0) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "FP32 refinement does not reach %.1e in %d steps (last correction %.2e)", tol,
                      kMax, c.empty() ? 0.0 : c.back());
        return fail(SPIKE_ERR_UNSUPPORTED, msg);
    }
    f->refine_steps = S;
    f->refine_est = c[S];
    double rate = 0.0;
    for (int j = 1; j <= S; j++)
        if (c[j - 1] > 0) rate = std::max(rate, c[j] / c[j - 1]);
    f->refine_rate = S >= 1 ? rate : c[0];
    return SPIKE_OK;
}

// Every chunk of a device-pointer solve, launched on s.
int solve_all(spike_factor_s *f, int nrhs, const double *b, int ldb, double *x, int ldx, hipStream_t s)
{
    for (int j = 0; j < nrhs; j += f->max_nr) {
        const int nr = std::min(f->max_nr, nrhs - j);
        int rc = solve_chunk(f, b + (size_t)j * ldb, ldb, x + (size_t)j * ldx, ldx, nr, s, nullptr, nullptr);
        if (rc) return rc;
    }
    return SPIKE_OK;
}

// Device-pointer solve through the graph cache.  Returns 1 when no graph is
// used for this call (the caller launches eagerly).
// Kernel launches of one solve (estimate).  A replayed graph costs the
// device several microseconds more than the same kernels launched one by one
// (measured on gfx950: about 6 us per solve on a 1M-row Spike solve of ten
// launches), so the graph cache only serves solves of many launches, where
// host launch overhead would exceed the device time.
long solve_launches(const spike_factor_s *f, int nrhs)
{
    const int per_call = std::max(f->max_nr, 1);
    const long chunks = (nrhs + per_call - 1) / per_call;
    long per = f->method != SPIKE_METHOD_LEVELS ? 3L + std::max(f->sweeps, 1)
               : (f->sync_free && nrhs >= 2)    ? 4L
                                                : (long)f->nlev_full + 2;
    if (f->fp32) per = per * (1 + f->refine_steps) + 2L * f->refine_steps;
    return chunks * per;
}
constexpr long kGraphMinLaunches = 24;

int solve_graph(spike_factor_s *f, int nrhs, const double *b, int ldb, double *x, int ldx, hipStream_t s)
{
    const spike_factor_s::GraphKey key{b, x, ldb, ldx, nrhs};
    f->graph_clock++;
    for (auto &g : f->graphs)
        if (g.key == key) {
            g.used = f->graph_clock;
            HIP_TRY(hipGraphLaunch(g.exec, s));
            return SPIKE_OK;
        }
    const bool repeat = key == f->last_key;
    f->last_key = key;
    if (!repeat) return 1;
    if (!f->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&f->cap_stream, hipStreamNonBlocking));
    // the capture stream must start after work already queued on s
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipStreamBeginCapture(f->cap_stream, hipStreamCaptureModeRelaxed));
    int rc = solve_all(f, nrhs, b, ldb, x, ldx, f->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(f->cap_stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc ? rc : fail(SPIKE_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(e));
    }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(SPIKE_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    constexpr size_t kMaxGraphs = 16;
    if (f->graphs.size() >= kMaxGraphs) {
        auto lru = std::min_element(f->graphs.begin(), f->graphs.end(),
                                    [](const auto &a, const auto &c) { return a.used < c.used; });
        (void)hipGraphExecDestroy(lru->exec);
        (void)hipGraphDestroy(lru->graph);
        f->graphs.erase(lru);
    }
    f->graphs.push_back(spike_factor_s::GraphEntry{key, ex, g, f->graph_clock});
    HIP_TRY(hipGraphLaunch(ex, s));
    return SPIKE_OK;
}

int factor_solve(spike_factor_s *f, int nrhs, const double *b, int ldb, double *x, int ldx, hipStream_t s,
                 bool timed, double *phase_ms = nullptr)
{
    const int n = f->plan.n;
    if (nrhs < 0 || ldb < n || ldx < n) return fail(SPIKE_ERR_ARG, "nrhs must be >= 0 and ldb, ldx >= n");
    if (nrhs == 0) return SPIKE_OK;
    if (!b || !x) return fail(SPIKE_ERR_ARG, "b and x must be non-NULL");
    // an asynchronous solve before this one that gave up is reported here
    if (int rc = take_failure(f)) return rc;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != f->device) HIP_TRY(hipSetDevice(f->device));
    const bool bdev = is_device_ptr(b), xdev = is_device_ptr(x);
    {
        int rc = ensure_level_ws(f, nrhs);
        if (!rc) rc = prepare_fused(f, std::min(nrhs, f->max_nr));
        if (!rc && nrhs > f->max_nr && nrhs % f->max_nr) rc = prepare_fused(f, nrhs % f->max_nr);
        if (rc) return rc;
    }
    if (bdev && xdev && !timed && f->opts.use_graph > 0 &&
        (f->opts.use_graph >= 2 || solve_launches(f, nrhs) >= kGraphMinLaunches)) {
        int rc = solve_graph(f, nrhs, b, ldb, x, ldx, s);
        if (rc != 1) return rc;
    }
    const double *db = b;
    double *dx = x;
    int ldb_d = ldb, ldx_d = ldx;
    const size_t need = (size_t)n * nrhs;
    if (!bdev || !xdev) {
        if (f->stage_cap < need) {
            if (f->d_stage_b) dfree(f, f->d_stage_b, sizeof(double) * f->stage_cap);
            if (f->d_stage_x) dfree(f, f->d_stage_x, sizeof(double) * f->stage_cap);
            f->d_stage_b = f->d_stage_x = nullptr;
            f->stage_cap = 0;
            int rc;
            if ((rc = f->dalloc(&f->d_stage_b, need)) || (rc = f->dalloc(&f->d_stage_x, need))) return rc;
            f->stage_cap = need;
        }
    }
    if (!bdev) {
        HIP_TRY(hipMemcpy2DAsync(f->d_stage_b, sizeof(double) * n, b, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                                 hipMemcpyHostToDevice, s));
        db = f->d_stage_b;
        ldb_d = n;
    }
    if (!xdev) {
        dx = f->d_stage_x;
        ldx_d = n;
    }
    hipEvent_t ev[5];
    if (timed)
        if (int erc = create_events(ev, 5)) return erc;
    for (int j = 0; j < nrhs; j += f->max_nr) {
        const int nr = std::min(f->max_nr, nrhs - j);
        double ms[4] = {0.0, 0.0, 0.0, 0.0};
        int rc = solve_chunk(f, db + (size_t)j * ldb_d, ldb_d, dx + (size_t)j * ldx_d, ldx_d, nr, s,
                             timed ? ev : nullptr, ms);
        if (rc) return rc;
        if (timed) {
            if (phase_ms)
                for (int i = 0; i < 4; i++) phase_ms[i] += ms[i];
            g_timers[5] += ms[0] * 1e-3;
            g_timers[6] += ms[1] * 1e-3;
            g_timers[7] += ms[2] * 1e-3;
            g_timers[8] += ms[3] * 1e-3;
            g_timers[4] += (ms[0] + ms[1] + ms[2] + ms[3]) * 1e-3;
        }
    }
    if (timed)
        for (int i = 0; i < 5; i++) (void)hipEventDestroy(ev[i]);
    if (!xdev) {
        HIP_TRY(hipMemcpy2DAsync(x, sizeof(double) * ldx, dx, sizeof(double) * n, sizeof(double) * n, nrhs,
                                 hipMemcpyDeviceToHost, s));
    }
    if (!bdev || !xdev || timed) {
        HIP_TRY(hipStreamSynchronize(s));
        return take_failure(f);
    }
    return SPIKE_OK;
}

// ---------------------------------------------------------------- legacy state

std::mutex g_mu;
spike_factor_t g_cur = nullptr;
int g_partitions_override = 0;

int legacy_partitions(int n, int nthreads)
{
    int P = g_partitions_override;
    if (P <= 0) {
        const char *e = std::getenv("SPIKE_PSTRSV_PARTITIONS");
        if (e && *e) P = std::atoi(e);
    }
    if (P <= 0) P = nthreads;
    if (P <= 0) P = 1;
    return std::min(P, std::max(n, 1));
}

// Copies the library's phase timers into the harness globals.
void sync_harness_timers()
{
    preprocessD_timer_pstrsv3 = g_timers[1];
    produceS_timer_pstrsv3 = g_timers[2];
    produceReducedSystem_timer_pstrsv3 = g_timers[3];
    reduced_solver_timer_pstrsv3 = g_timers[6];
    computeY_timer[0] = g_timers[5];
    subtract_Rx_timer[0] = g_timers[7];
    solveD2_timer[0] = g_timers[8];
}

void reset_harness_timers()
{
    for (double &v : g_timers) v = 0.0;
    memoryAlloc_timer_pstrsv3 = preprocessD_timer_pstrsv3 = preprocessR_timer_pstrsv3 = 0.0;
    setDependents_timer_pstrsv3 = preprocessS_timer_pstrsv3 = preprocessMappings_timer_pstrsv3 = 0.0;
    produceCompressedR_timer_pstrsv3 = produceS_timer_pstrsv3 = produceReducedSystem_timer_pstrsv3 = 0.0;
    init_vars_pstrsv3 = reduced_solver_timer_pstrsv3 = 0.0;
    for (int i = 0; i < MAXTHREADCOUNT; i++)
        temp_timer[i] = waitBefore_timer_pstrsv3[i] = computeY_timer[i] = solveIndependent_timer[i] =
            subtract_Rx_timer[i] = waitAfter_timer_pstrsv3[i] = solveD2_timer[i] = solveD_timer[i] = 0.0;
}

void legacy_preproc(double *a, int *ia, int *ja, int n, int nthreads, char uplo, const char *who)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_cur) {
        spike_factor_destroy(g_cur);
        g_cur = nullptr;
    }
    spike_options_t o;
    spike_options_default(&o);
    o.partitions = legacy_partitions(n, nthreads);
    spike_factor_t f = nullptr;
    int rc = spike_factor_create(&f, uplo, n, a, ia, ja, &o);
    if (rc != SPIKE_OK) {
        std::fprintf(stderr, "%s: %s\n", who, t_err.c_str());
        return;
    }
    g_cur = f;
    if (PSTRSV_VERBOSE) spike_factor_log(f);
}

void legacy_solve(char uplo, int m, double *b, double *x, const char *who, bool timed)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_cur) {
        std::fprintf(stderr, "%s: no preprocessed matrix (call spike_pstrsv_preproc first)\n", who);
        return;
    }
    char u = (char)(uplo == 'l' ? 'L' : uplo == 'u' ? 'U' : uplo);
    if (u != g_cur->plan.uplo) {
        std::fprintf(stderr, "%s: uplo '%c' does not match the preprocessed matrix ('%c')\n", who, uplo,
                     g_cur->plan.uplo);
        return;
    }
    if (m != g_cur->plan.n) {
        std::fprintf(stderr, "%s: m = %d does not match the preprocessed matrix (n = %d)\n", who, m, g_cur->plan.n);
        return;
    }
    int rc = factor_solve(g_cur, 1, b, m, x, m, nullptr, timed);
    if (rc != SPIKE_OK) std::fprintf(stderr, "%s: %s\n", who, t_err.c_str());
}

}  // namespace

// ---------------------------------------------------------------- batches

struct spike_batch_s {
    char uplo = 'L';
    int count = 0;
    spike_batch_options_t opts{};
    std::vector<int64_t> off;              // count+1 packed offsets
    std::vector<spike_factor_s *> groups;  // block-diagonal factors
    int device = 0;
    int rpp[4] = {0, 0, 0, 0};             // rows per partition at 1, 2, 4, 8 RHS slots
    // lanes per row at 1, 2, 4, 8 RHS slots, relative to each matrix's row
    // length rule: G = rule x 2^lane_shift (within 1 .. 16); autotuned
    int lane_shift[4] = {0, 0, 0, 0};
    double autotune_s = 0.0, preproc_s = 0.0;
    // captured solve graph for one (b, x) pair
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double *gb = nullptr;
    double *gx = nullptr;
    int graph_nodes = 0;
    // milliseconds of each group's solve run alone (measured before the
    // first capture; they order the groups over the graph's branches)
    std::vector<float> group_ms;
    // host-pointer staging
    double *d_sb = nullptr, *d_sx = nullptr;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        gb = nullptr;
        gx = nullptr;
    }
    void release()
    {
        drop_graph();
        for (hipStream_t s : streams) (void)hipStreamDestroy(s);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        streams.clear();
        events.clear();
        if (d_sb) (void)hipFree(d_sb);
        if (d_sx) (void)hipFree(d_sx);
        d_sb = d_sx = nullptr;
        for (spike_factor_s *f : groups) {
            f->release();
            delete f;
        }
        groups.clear();
    }
};

namespace {

struct GroupSpec {
    int slots = 1, G = 1;
    std::vector<int> mats;
};

// Basic per-matrix CSR checks needed before the matrix is copied into a
// block-diagonal system (the full conventions are checked on that system).
std::string check_batch_matrix(const spike_batch_matrix_t &m, int i)
{
    const std::string who = "matrix " + std::to_string(i) + ": ";
    if (m.n < 1) return who + "n must be >= 1";
    if (!m.a || !m.ia || !m.ja) return who + "a, ia and ja must be non-NULL";
    if (m.nrhs < 1 || m.nrhs > 8) return who + "nrhs must be in [1, 8]";
    if (m.ia[0] != 0) return who + "ia[0] must be 0";
    for (int r = 0; r < m.n; r++)
        if (m.ia[r + 1] <= m.ia[r]) return who + "row " + std::to_string(r) + " is empty (diagonal missing)";
    for (int e = 0; e < m.ia[m.n]; e++)
        if (m.ja[e] < 0 || m.ja[e] >= m.n) return who + "column index out of range";
    return "";
}

// Builds one group: its matrices as one block-diagonal lower-form system
// ('U' matrices reversed one by one), partitioned matrix by matrix with about
// rpp rows per partition (the reference's formula within each matrix).
int build_group(spike_batch_s *bt, const spike_batch_matrix_t *m, const GroupSpec &g, int rpp, int precision,
                int refine_steps, spike_factor_s **out)
{
    *out = nullptr;
    const bool flip = bt->uplo == 'U';
    int64_t rows = 0, nnz = 0;
    for (int i : g.mats) {
        rows += m[i].n;
        nnz += m[i].ia[m[i].n];
    }
    if (rows >= INT32_MAX || nnz > INT32_MAX) return fail(SPIKE_ERR_UNSUPPORTED, "batch group exceeds 2^31-1 entries");
    std::vector<int> ia, ja, starts{0};
    std::vector<double> a;
    spike::resize_large(ia, (size_t)rows + 1);
    spike::resize_large(ja, (size_t)nnz);
    spike::resize_large(a, (size_t)nnz);
    spike_factor_s *f = new spike_factor_s();
    spike_options_default(&f->opts);
    f->opts.precision = precision;
    f->opts.lanes_per_row = g.G;
    f->opts.drop_tol = bt->opts.drop_tol;
    f->opts.refine_steps = refine_steps;
    f->opts.refine_tol = bt->opts.refine_tol;
    f->batch_flip = flip;
    f->batch_slots = g.slots;
    f->vec_len = bt->off[bt->count];
    const size_t vb = precision == SPIKE_PREC_FP32_REFINED ? sizeof(float) : sizeof(double);
    const int kcap = std::min(kmax_limit_for_device(), (int)(spike::block_solve_lds_limit() / (vb * g.slots)) - spike::XS_PAD);
    // each matrix's rows and entries start at the sums over the ones before
    // it, so the matrices are copied in parallel
    std::vector<int64_t> row0(g.mats.size() + 1, 0), ent0(g.mats.size() + 1, 0);
    for (size_t k = 0; k < g.mats.size(); k++) {
        row0[k + 1] = row0[k] + m[g.mats[k]].n;
        ent0[k + 1] = ent0[k] + m[g.mats[k]].ia[m[g.mats[k]].n];
    }
    ia[0] = 0;
    spike::parallel_for((int)g.mats.size(), [&](int k) {
        const spike_batch_matrix_t &M = m[g.mats[k]];
        const int n = M.n;
        const int64_t r0 = row0[k];
        int64_t p = ent0[k];
        for (int r = 0; r < n; r++) {
            const int src = flip ? n - 1 - r : r;
            const int b0 = M.ia[src], b1 = M.ia[src + 1];
            if (!flip)
                for (int e = b0; e < b1; e++, p++) {
                    ja[p] = (int)r0 + M.ja[e];
                    a[p] = M.a[e];
                }
            else
                for (int e = b1 - 1; e >= b0; e--, p++) {
                    ja[p] = (int)r0 + (n - 1 - M.ja[e]);
                    a[p] = M.a[e];
                }
            ia[r0 + r + 1] = (int)p;
        }
    });
    int64_t r0 = 0;
    for (int i : g.mats) {
        const spike_batch_matrix_t &M = m[i];
        const int n = M.n;
        f->mats.push_back(spike_factor_s::Mat{(int)r0, n, M.nrhs, bt->off[i]});
        f->batch_max_nrhs = std::max(f->batch_max_nrhs, M.nrhs);
        int P = std::max(1, (n + rpp / 2) / std::max(rpp, 1));
        P = std::min(P, n);
        while (P < n && n - (n / P) * (P - 1) > kcap) P++;
        for (int t = 1; t < P; t++) starts.push_back((int)r0 + (n / P) * t);
        starts.push_back((int)(r0 + n));
        r0 += n;
    }
    int rc = factor_build(f, 'L', (int)rows, a.data(), ia.data(), ja.data(), &starts);
    if (rc != SPIKE_OK) {
        std::string keep = t_err;
        f->release();
        delete f;
        t_err = keep;
        return rc;
    }
    *out = f;
    return SPIKE_OK;
}

// Groups: right-hand-side slots (next power of two >= nrhs) x lanes per row
// (from the p90 off-diagonal row length, or the option); a group is split
// where it would pass 1.5e9 entries.
std::vector<GroupSpec> plan_groups(const spike_batch_matrix_t *m, const std::vector<int> &idx, int lanes, bool fp32,
                                   const int *lane_shift)
{
    std::vector<GroupSpec> out;
    std::vector<std::pair<int, int>> key(idx.size());
    for (size_t j = 0; j < idx.size(); j++) {
        const spike_batch_matrix_t &M = m[idx[j]];
        int slots = 1;
        while (slots < M.nrhs) slots <<= 1;
        int G = lanes;
        if (G <= 0) {
            std::vector<int> len(M.n);
            for (int r = 0; r < M.n; r++) len[r] = M.ia[r + 1] - M.ia[r] - 1;
            G = fp32 ? spike::choose_lanes(len) : spike::choose_block_lanes(len);
            const int sh = lane_shift[__builtin_ctz((unsigned)slots)];
            G = std::min(16, std::max(1, sh >= 0 ? G << sh : G >> -sh));
        }
        key[j] = {slots, G};
    }
    std::vector<std::pair<int, int>> keys(key);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (const auto &k : keys) {
        GroupSpec cur;
        cur.slots = k.first;
        cur.G = k.second;
        int64_t nnz = 0;
        for (size_t j = 0; j < idx.size(); j++) {
            if (key[j] != k) continue;
            const int64_t e = m[idx[j]].ia[m[idx[j]].n];
            if (!cur.mats.empty() && nnz + e > 1500000000) {
                out.push_back(cur);
                cur.mats.clear();
                nnz = 0;
            }
            cur.mats.push_back(idx[j]);
            nnz += e;
        }
        if (!cur.mats.empty()) out.push_back(cur);
    }
    return out;
}

int batch_run(spike_batch_s *bt, const double *b, double *x, hipStream_t s)
{
    for (spike_factor_s *f : bt->groups) {
        int rc = solve_chunk(f, b, f->plan.n, x, f->plan.n, f->batch_slots, s, nullptr, nullptr);
        if (rc) return rc;
    }
    return SPIKE_OK;
}

// Builds the groups of matrices idx (packed offsets already in bt->off).
int batch_build(spike_batch_s *bt, const spike_batch_matrix_t *m, const std::vector<int> &idx, const int *rpp,
                int refine_steps)
{
    std::vector<GroupSpec> specs =
        plan_groups(m, idx, bt->opts.lanes_per_row, bt->opts.precision == SPIKE_PREC_FP32_REFINED, bt->lane_shift);
    for (const GroupSpec &g : specs) {
        spike_factor_s *f = nullptr;
        int rc = build_group(bt, m, g, rpp[__builtin_ctz((unsigned)g.slots)], bt->opts.precision, refine_steps,
                             &f);
        if (rc) return rc;
        bt->groups.push_back(f);
        if ((rc = prepare_fused(f, g.slots))) return rc;
    }
    if (bt->opts.precision == SPIKE_PREC_FP32_REFINED && !bt->groups.empty()) {
        // one residual workspace for all groups: their matrices own disjoint
        // ranges of the packed layout
        spike_factor_s *f0 = bt->groups[0];
        int rc = f0->dalloc(&f0->d_r, (size_t)bt->off[bt->count]);
        if (rc) return rc;
        for (spike_factor_s *f : bt->groups) f->d_r = f0->d_r;
    }
    return SPIKE_OK;
}

// Refinement steps of an FP32 batch: calibrate_refinement's procedure over
// every matrix of every group at once.
int batch_calibrate(spike_batch_s *bt)
{
    const double tol = bt->opts.refine_tol > 0 ? bt->opts.refine_tol : 1e-13;
    constexpr int kMax = 8;
    const size_t len = (size_t)bt->off[bt->count];
    std::vector<double> hb(len), hx(len), hprev(len);
    uint64_t st = 0x5350494B45434131ull;
    for (size_t i = 0; i < len; i++) {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        hb[i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
    }
    double *db = nullptr, *dx = nullptr;
    HIP_TRY(hipMalloc((void **)&db, sizeof(double) * len));
    hipError_t e = hipMalloc((void **)&dx, sizeof(double) * len);
    if (e != hipSuccess) {
        (void)hipFree(db);
        return fail(SPIKE_ERR_NOMEM, "calibration buffers");
    }
    auto done = [&](int rc) {
        (void)hipFree(db);
        (void)hipFree(dx);
        return rc;
    };
    if (hipMemcpy(db, hb.data(), sizeof(double) * len, hipMemcpyHostToDevice) != hipSuccess)
        return done(fail(SPIKE_ERR_DEVICE, "calibration upload"));
    for (spike_factor_s *f : bt->groups) f->refine_steps = 0;
    int rc = batch_run(bt, db, dx, nullptr);
    if (rc) return done(rc);
    if (hipMemcpy(hprev.data(), dx, sizeof(double) * len, hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(SPIKE_ERR_DEVICE, "calibration download"));
    std::vector<double> c;
    int S = -1;
    for (int j = 1; j <= kMax && S < 0; j++) {
        for (spike_factor_s *f : bt->groups) {
            if ((rc = residual(f, db, f->plan.n, dx, f->plan.n, f->batch_slots, nullptr))) return done(rc);
            if ((rc = solve_pass(f, f->d_r, f->plan.n, dx, f->plan.n, f->batch_slots, 1, nullptr, nullptr, nullptr)))
                return done(rc);
        }
        if (hipMemcpy(hx.data(), dx, sizeof(double) * len, hipMemcpyDeviceToHost) != hipSuccess)
            return done(fail(SPIKE_ERR_DEVICE, "calibration download"));
        double worst = 0.0;
        for (int i = 0; i < bt->count; i++) {
            double dm = 0.0, xm = 0.0;
            for (int64_t q = bt->off[i]; q < bt->off[i + 1]; q++) {
                dm = std::max(dm, std::fabs(hx[q] - hprev[q]));
                xm = std::max(xm, std::fabs(hx[q]));
            }
            worst = std::max(worst, xm > 0 ? dm / xm : (dm > 0 ? HUGE_VAL : 0.0));
        }
        c.push_back(worst);
        if (!std::isfinite(worst)) break;
        if (worst <= tol) S = j - 1;
        hprev.swap(hx);
    }
    if (S < 0) return done(fail(SPIKE_ERR_UNSUPPORTED, "FP32 refinement of the batch does not converge"));
    for (spike_factor_s *f : bt->groups) {
        f->refine_steps = S;
        f->refine_est = c[S];
        f->stats.refine_steps = S;
        f->stats.refine_est_err = c[S];
    }
    return done(SPIKE_OK);
}

// Autotuner for the partition size, per right-hand-side slot class (1, 2,
// 4, 8 slots).  Block solves are latency-bound, so their speed depends on how
// many partitions each CU keeps resident, which the LDS x segments (rows x
// slots x value size) limit; the reduced system grows as partitions shrink.
// For every class present, a sample of the class's matrices (every j-th,
// about 1024 x the largest candidate rows, enough to fill the GPU) is built
// with each candidate size and timed over three whole solves (best of three
// rounds); the largest candidate within 10% of the most rows x RHS per second
// wins.
int autotune_rpp(spike_batch_s *bt, const spike_batch_matrix_t *m, int *best)
{
    // FP64 candidates bracket the sizes the 4096-matrix banded batch settles
    // on (2048, 1024, 512 and 512 rows at 1-8 slots: each the middle one, the
    // larger and smaller neighbours measured slower).  FP32 keeps its own
    // range: allowed 2048 rows at 8 slots, it picked 1024 there (within the
    // 10% margin on the sample), and the full batch then calibrated two
    // refinement steps instead of one (51.9 against 34.9 ms per solve).
    static const int cand64[4][3] = {{1024, 2048, 4096}, {512, 1024, 2048}, {256, 512, 1024}, {256, 512, 1024}};
    static const int cand32[4][3] = {{1024, 2048, 4096}, {512, 1024, 2048}, {512, 1024, 2048}, {256, 512, 1024}};
    const int(*cand)[3] = bt->opts.precision == SPIKE_PREC_FP32_REFINED ? cand32 : cand64;
    // The sample holds about 4096 x the largest candidate's rows
    // (SPIKE_PSTRSV_TUNE_SAMPLE: a multiple of 1024 x, default 4), so that it
    // is bandwidth-bound like the full batch: with 1024 x the sample favoured
    // small partitions, [1024, 512, 512, 256] rows at 1-8 slots, and the
    // 4096-matrix batch took 29.3 ms per solve; with 4096 x it picks [2048,
    // 1024, 512, 512] and half the lanes per row, 20.7 ms (FP64).
    // SPIKE_PSTRSV_TUNE_LOG=1 prints the timings to stderr.
    int sample_scale = 4;
    if (const char *e = std::getenv("SPIKE_PSTRSV_TUNE_SAMPLE")) sample_scale = std::max(1, std::atoi(e));
    const bool tlog = std::getenv("SPIKE_PSTRSV_TUNE_LOG") && std::atoi(std::getenv("SPIKE_PSTRSV_TUNE_LOG")) == 1;
    hipEvent_t ee[2];
    if (int erc = create_events(ee, 2)) return erc;
    hipEvent_t e0 = ee[0], e1 = ee[1];
    int rc = SPIKE_OK;
    for (int li = 0; li < 4 && !rc; li++) {
        best[li] = cand[li][1];
        std::vector<int> members;
        int64_t rows = 0;
        for (int i = 0; i < bt->count; i++) {
            int slots = 1;
            while (slots < m[i].nrhs) slots <<= 1;
            if (slots == (1 << li)) {
                members.push_back(i);
                rows += m[i].n;
            }
        }
        if (members.empty()) continue;
        const int64_t target = 1024LL * cand[li][2] * sample_scale;
        const int64_t want = bt->opts.autotune_sample > 0 ? bt->opts.autotune_sample : INT64_MAX;
        const size_t step = rows > target ? (size_t)(rows / target) : 1;
        std::vector<spike_batch_matrix_t> sm;
        for (size_t j = 0; j < members.size() && (int64_t)sm.size() < want; j += std::max<size_t>(step, 1))
            sm.push_back(m[members[j]]);
        std::vector<int> sidx(sm.size());
        spike_batch_s tmp;
        tmp.uplo = bt->uplo;
        tmp.count = (int)sm.size();
        tmp.opts = bt->opts;
        tmp.off.assign(sm.size() + 1, 0);
        for (size_t j = 0; j < sm.size(); j++) {
            sidx[j] = (int)j;
            tmp.off[j + 1] = tmp.off[j] + (int64_t)sm[j].n * sm[j].nrhs;
        }
        const size_t len = (size_t)tmp.off[sm.size()];
        double *db = nullptr, *dx = nullptr;
        if (hipMalloc((void **)&db, sizeof(double) * len) != hipSuccess ||
            hipMalloc((void **)&dx, sizeof(double) * len) != hipSuccess) {
            (void)hipGetLastError();
            if (db) (void)hipFree(db);
            rc = fail(SPIKE_ERR_NOMEM, "autotune buffers");
            break;
        }
        std::vector<double> hb(len);
        for (size_t i = 0; i < len; i++) hb[i] = std::sin(0.37 * (double)i);
        (void)hipMemcpy(db, hb.data(), sizeof(double) * len, hipMemcpyHostToDevice);
        double rate_of[3] = {-1.0, -1.0, -1.0};
        for (int ci = 0; ci < 3; ci++) {
            const int c = cand[li][ci];
            int r4[4] = {c, c, c, c};
            tmp.release();
            if (batch_build(&tmp, sm.data(), sidx, r4, bt->opts.precision == SPIKE_PREC_FP32_REFINED ? 1 : 0)) {
                t_err.clear();
                continue;   // candidate does not fit: skip it
            }
            if ((rc = batch_run(&tmp, db, dx, nullptr))) break;
            // best of three rounds of three solves
            for (int round = 0; round < 3 && !rc; round++) {
                (void)hipEventRecord(e0, nullptr);
                for (int r = 0; r < 3 && !rc; r++) rc = batch_run(&tmp, db, dx, nullptr);
                (void)hipEventRecord(e1, nullptr);
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms > 0) rate_of[ci] = std::max(rate_of[ci], 3.0 * (double)len / (ms * 1e-3));
            }
            if (rc) break;
        }
        // The largest candidate within 10% of the best sampled rate wins: the
        // sample runs one class alone, while the full batch runs the classes
        // concurrently (the GPU is kept busy by the other groups, and larger
        // partitions mean a smaller reduced system); on a 4096-matrix batch
        // the sample's pick of 256 rows for the 4- and 8-RHS classes measured
        // 38.5 ms per solve (34.8 with 256 for 8 RHS only) against 24.5 with
        // 512 for both.
        const double top = std::max(rate_of[0], std::max(rate_of[1], rate_of[2]));
        if (tlog)
            std::fprintf(stderr, "autotune: %d slots, %zu sample matrices, rows per partition %d / %d / %d: %.3e %.3e %.3e rows*RHS/s\n",
                         1 << li, sm.size(), cand[li][0], cand[li][1], cand[li][2],
                         rate_of[0], rate_of[1], rate_of[2]);
        double base = -1.0;
        for (int ci = 0; ci < 3; ci++)
            if (top > 0 && rate_of[ci] >= 0.90 * top) {
                best[li] = cand[li][ci];
                base = rate_of[ci];
            }
        // lanes per row (rows per wavefront step): half and twice each
        // matrix's rule at the chosen partition size, timed the same way;
        // a shift must beat the rule by 3%
        bt->lane_shift[li] = 0;
        if (bt->opts.lanes_per_row <= 0 && base > 0) {
            double best_rate = base * 1.03;
            for (int sh : {-1, 1}) {
                tmp.release();
                tmp.lane_shift[li] = sh;
                int r4[4] = {best[li], best[li], best[li], best[li]};
                if (batch_build(&tmp, sm.data(), sidx, r4, bt->opts.precision == SPIKE_PREC_FP32_REFINED ? 1 : 0)) {
                    t_err.clear();
                    continue;
                }
                if ((rc = batch_run(&tmp, db, dx, nullptr))) break;
                double rate = -1.0;
                for (int round = 0; round < 3 && !rc; round++) {
                    (void)hipEventRecord(e0, nullptr);
                    for (int r = 0; r < 3 && !rc; r++) rc = batch_run(&tmp, db, dx, nullptr);
                    (void)hipEventRecord(e1, nullptr);
                    (void)hipEventSynchronize(e1);
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (ms > 0) rate = std::max(rate, 3.0 * (double)len / (ms * 1e-3));
                }
                if (rc) break;
                if (tlog)
                    std::fprintf(stderr, "autotune: %d slots, lanes x 2^%d: %.3e rows*RHS/s (rule %.3e)\n", 1 << li, sh,
                                 rate, base);
                if (rate > best_rate) {
                    best_rate = rate;
                    bt->lane_shift[li] = sh;
                }
            }
            tmp.lane_shift[li] = 0;
        }
        tmp.release();
        (void)hipFree(db);
        (void)hipFree(dx);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Captures every group's launch sequence into one graph: group i on branch
// i % B of B capture streams forked from and joined back to the first.
int batch_capture(spike_batch_s *bt, const double *b, double *x)
{
    bt->drop_graph();
    // four branches (4096-matrix batch: 19.7 ms per solve, against 20.7 with
    // two and 20.1 with eight)
    const int B = std::max(1, std::min(4, (int)bt->groups.size()));
    while ((int)bt->streams.size() < B) {
        hipStream_t s;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        bt->streams.push_back(s);
    }
    while ((int)bt->events.size() < B) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        bt->events.push_back(e);
    }
    hipStream_t s0 = bt->streams[0];
    const int G = (int)bt->groups.size();
    if ((int)bt->group_ms.size() != G) {
        // each group's solve alone, once (it writes the x this solve writes)
        bt->group_ms.assign(G, 0.0f);
        hipEvent_t tt[2];
        if (int erc = create_events(tt, 2)) {
            bt->group_ms.clear();
            return erc;
        }
        hipEvent_t t0 = tt[0], t1 = tt[1];
        int rc = SPIKE_OK;
        for (int g = 0; g < G && !rc; g++) {
            spike_factor_s *f = bt->groups[g];
            (void)hipEventRecord(t0, s0);
            rc = solve_chunk(f, b, f->plan.n, x, f->plan.n, f->batch_slots, s0, nullptr, nullptr);
            (void)hipEventRecord(t1, s0);
            (void)hipEventSynchronize(t1);
            (void)hipEventElapsedTime(&bt->group_ms[g], t0, t1);
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        if (rc) {
            bt->group_ms.clear();
            return rc;
        }
    }
    // longest first, each to the least loaded branch: the long chains start
    // together and the short groups fill the other branches behind them
    std::vector<int> order(G), branch(G, 0);
    for (int g = 0; g < G; g++) order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return bt->group_ms[p] > bt->group_ms[q]; });
    {
        std::vector<double> load(B, 0.0);
        for (int g : order) {
            const int br = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            branch[g] = br;
            load[br] += bt->group_ms[g];
        }
    }
    HIP_TRY(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
    int rc = SPIKE_OK;
    hipError_t e = hipEventRecord(bt->events[0], s0);
    for (int i = 1; i < B && e == hipSuccess; i++) e = hipStreamWaitEvent(bt->streams[i], bt->events[0], 0);
    for (int k = 0; k < G && e == hipSuccess && !rc; k++) {
        spike_factor_s *f = bt->groups[order[k]];
        rc = solve_chunk(f, b, f->plan.n, x, f->plan.n, f->batch_slots, bt->streams[branch[order[k]]], nullptr,
                         nullptr);
    }
    for (int i = 1; i < B && e == hipSuccess && !rc; i++) {
        e = hipEventRecord(bt->events[i], bt->streams[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s0, bt->events[i], 0);
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s0, &g);
    if (rc || e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        if (rc) return rc;
        return fail(SPIKE_ERR_DEVICE, std::string("graph capture failed: ") +
                                          hipGetErrorString(e != hipSuccess ? e : e2));
    }
    bt->graph = g;
    size_t nodes = 0;
    (void)hipGraphGetNodes(g, nullptr, &nodes);
    bt->graph_nodes = (int)nodes;
    HIP_TRY(hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0));
    bt->gb = b;
    bt->gx = x;
    return SPIKE_OK;
}

// Wall-clock milliseconds of one 1-RHS device solve of f: two warm-up
// solves, then the fastest of three rounds of three.
int time_factor(spike_factor_s *f, const double *db, double *dx, double *ms_out)
{
    hipEvent_t ee[2];
    if (int erc = create_events(ee, 2)) return erc;
    hipEvent_t e0 = ee[0], e1 = ee[1];
    const int n = f->plan.n;
    int rc = SPIKE_OK;
    for (int r = 0; r < 2 && !rc; r++) rc = factor_solve(f, 1, db, n, dx, n, nullptr, false);
    float best = 1e30f;
    for (int round = 0; round < 3 && !rc; round++) {
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < 3 && !rc; r++) rc = factor_solve(f, 1, db, n, dx, n, nullptr, false);
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        best = std::min(best, ms / 3.0f);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = best;
    return rc;
}

// spike_factor_create with autotune >= 2: the rule's factor, then half and
// twice its partition count (when partitions is 0) and, at the faster count,
// half and twice its lanes per row (when lanes_per_row is 0), each built and
// timed over one-RHS solves; a candidate replaces the current choice when
// it is 3% faster.  Level-method factors are kept as built.
int factor_create_tuned(spike_factor_s **out, char uplo, int n, const double *a, const int *ia, const int *ja,
                        const spike_options_t &o)
{
    auto build = [&](int P, int G, spike_factor_s **res) -> int {
        spike_factor_s *f = new spike_factor_s();
        f->opts = o;
        f->opts.autotune = 1;
        f->opts.partitions = P;
        f->opts.lanes_per_row = G;
        int rc = factor_build(f, uplo, n, a, ia, ja);
        if (rc != SPIKE_OK) {
            f->release();
            delete f;
            *res = nullptr;
            return rc;
        }
        *res = f;
        return SPIKE_OK;
    };
    spike_factor_s *best = nullptr;
    int rc = build(o.partitions, o.lanes_per_row, &best);
    if (rc != SPIKE_OK) return rc;
    if (best->method == SPIKE_METHOD_LEVELS) {
        *out = best;
        return SPIKE_OK;
    }
    double *db = nullptr, *dx = nullptr;
    if (hipMalloc((void **)&db, sizeof(double) * n) != hipSuccess ||
        hipMalloc((void **)&dx, sizeof(double) * n) != hipSuccess) {
        (void)hipGetLastError();
        if (db) (void)hipFree(db);
        *out = best;   // untuned, but valid
        return SPIKE_OK;
    }
    {
        std::vector<double> hb(n);
        for (int i = 0; i < n; i++) hb[i] = std::sin(0.37 * (double)i);
        (void)hipMemcpy(db, hb.data(), sizeof(double) * n, hipMemcpyHostToDevice);
    }
    double best_ms = 0.0;
    rc = time_factor(best, db, dx, &best_ms);
    auto consider = [&](int P, int G) {
        spike_factor_s *c = nullptr;
        if (rc != SPIKE_OK || build(P, G, &c) != SPIKE_OK) {
            t_err.clear();   // a candidate that does not build is skipped
            return;
        }
        double ms = 0.0;
        if (c->method == SPIKE_METHOD_SPIKE && time_factor(c, db, dx, &ms) == SPIKE_OK && ms < best_ms * 0.97) {
            best->release();
            delete best;
            best = c;
            best_ms = ms;
            return;
        }
        t_err.clear();
        c->release();
        delete c;
    };
    if (rc == SPIKE_OK && o.partitions == 0) {
        const int P0 = best->plan.P, G0 = best->steps.G;
        if (P0 / 2 >= 1) consider(P0 / 2, o.lanes_per_row > 0 ? o.lanes_per_row : G0);
        if (P0 * 2 <= n) consider(P0 * 2, o.lanes_per_row > 0 ? o.lanes_per_row : G0);
    }
    if (rc == SPIKE_OK && o.lanes_per_row == 0 && !best->steps.fwd) {
        const int P1 = best->plan.P, G1 = best->steps.G;
        if (G1 / 2 >= 1) consider(P1, G1 / 2);
        if (G1 * 2 <= 16) consider(P1, G1 * 2);
    }
    (void)hipFree(db);
    (void)hipFree(dx);
    if (rc != SPIKE_OK) {
        std::string keep = t_err;
        best->release();
        delete best;
        t_err = keep;
        return rc;
    }
    *out = best;
    return SPIKE_OK;
}

}  // namespace

extern "C" {

void spike_options_default(spike_options_t *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->partitions = 0;
    o->drop_tol = -1.0;
    o->precision = SPIKE_PREC_FP64;
    o->lanes_per_row = 0;
    o->verbose = 0;
    o->reduced_solver = SPIKE_REDUCED_AUTO;
    o->refine_steps = -1;
    o->refine_tol = 0.0;
    o->method = SPIKE_METHOD_AUTO;
    o->use_graph = 1;
    o->autotune = 1;
}

int spike_factor_create(spike_factor_t *out, char uplo, int n, const double *a, const int *ia, const int *ja,
                        const spike_options_t *opts)
{
    if (!out) return fail(SPIKE_ERR_ARG, "factor pointer is NULL");
    *out = nullptr;
    if (uplo != 'L' && uplo != 'l' && uplo != 'U' && uplo != 'u') return fail(SPIKE_ERR_ARG, "uplo must be L or U");
    if (n < 1) return fail(SPIKE_ERR_ARG, "n must be >= 1");
    if (!a || !ia || !ja) return fail(SPIKE_ERR_ARG, "a, ia and ja must be non-NULL");
    spike_options_t o;
    spike_options_default(&o);
    if (opts) o = *opts;
    if (o.precision != SPIKE_PREC_FP64 && o.precision != SPIKE_PREC_FP32_REFINED)
        return fail(SPIKE_ERR_ARG, "unknown precision mode");
    if (o.autotune >= 2 && (o.partitions == 0 || o.lanes_per_row == 0)) {
        spike_factor_s *t = nullptr;
        const int rc = factor_create_tuned(&t, uplo, n, a, ia, ja, o);
        if (rc != SPIKE_OK) return rc;
        *out = t;
        t_err.clear();
        return SPIKE_OK;
    }
    spike_factor_s *f = new spike_factor_s();
    f->opts = o;
    int rc = factor_build(f, uplo, n, a, ia, ja);
    if (rc != SPIKE_OK) {
        std::string keep = t_err;
        f->release();
        delete f;
        t_err = keep;
        return rc;
    }
    *out = f;
    t_err.clear();
    return SPIKE_OK;
}

int spike_factor_create_partitioned(spike_factor_t *out, char uplo, int n, const double *a, const int *ia,
                                    const int *ja, int nparts, const int *starts, const spike_options_t *opts)
{
    if (!out) return fail(SPIKE_ERR_ARG, "factor pointer is NULL");
    *out = nullptr;
    if (uplo == 'l') uplo = 'L';
    if (uplo == 'u') uplo = 'U';
    if (uplo != 'L' && uplo != 'U') return fail(SPIKE_ERR_ARG, "uplo must be L or U");
    if (n < 1) return fail(SPIKE_ERR_ARG, "n must be >= 1");
    if (!a || !ia || !ja || !starts) return fail(SPIKE_ERR_ARG, "a, ia, ja and starts must be non-NULL");
    if (nparts < 1 || nparts > n) return fail(SPIKE_ERR_ARG, "nparts must be in [1, n]");
    if (starts[0] != 0 || starts[nparts] != n) return fail(SPIKE_ERR_ARG, "starts must begin at 0 and end at n");
    for (int t = 0; t < nparts; t++)
        if (starts[t + 1] <= starts[t]) return fail(SPIKE_ERR_ARG, "starts must be strictly increasing");
    spike_options_t o;
    spike_options_default(&o);
    if (opts) o = *opts;
    if (o.precision != SPIKE_PREC_FP64 && o.precision != SPIKE_PREC_FP32_REFINED)
        return fail(SPIKE_ERR_ARG, "unknown precision mode");
    // boundaries of the lower form: 'U' rows are reversed, so partition t of
    // the caller is partition nparts-1-t there
    std::vector<int> sl(nparts + 1);
    for (int t = 0; t <= nparts; t++) sl[t] = uplo == 'U' ? n - starts[nparts - t] : starts[t];
    spike_factor_s *f = new spike_factor_s();
    f->opts = o;
    f->caller_parts = true;
    int rc = factor_build(f, uplo, n, a, ia, ja, &sl);
    if (rc != SPIKE_OK) {
        std::string keep = t_err;
        f->release();
        delete f;
        t_err = keep;
        return rc;
    }
    *out = f;
    t_err.clear();
    return SPIKE_OK;
}

int spike_factor_solve(spike_factor_t f, int nrhs, const double *b, int ldb, double *x, int ldx, void *stream)
{
    if (!f) return fail(SPIKE_ERR_STATE, "factor is NULL");
    return factor_solve(f, nrhs, b, ldb, x, ldx, (hipStream_t)stream, false);
}

int spike_factor_solve_profiled(spike_factor_t f, int nrhs, const double *b, int ldb, double *x, int ldx,
                                void *stream, double *phase_ms)
{
    if (!f) return fail(SPIKE_ERR_STATE, "factor is NULL");
    if (!phase_ms) return fail(SPIKE_ERR_ARG, "phase_ms is NULL");
    return factor_solve(f, nrhs, b, ldb, x, ldx, (hipStream_t)stream, true, phase_ms);
}

int spike_factor_solve_noalloc(spike_factor_t f, int nrhs, const double *b, int ldb, double *x, int ldx,
                               void *stream)
{
    if (!f) return fail(SPIKE_ERR_STATE, "factor is NULL");
    const int n = f->plan.n;
    if (nrhs < 0 || ldb < n || ldx < n) return fail(SPIKE_ERR_ARG, "nrhs must be >= 0 and ldb, ldx >= n");
    if (nrhs == 0) return SPIKE_OK;
    if (!b || !x) return fail(SPIKE_ERR_ARG, "b and x must be non-NULL");
    if (!f->mats.empty()) return fail(SPIKE_ERR_UNSUPPORTED, "batch groups have no no-alloc solve");
    if (int rc = take_failure(f)) return rc;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != f->device) HIP_TRY(hipSetDevice(f->device));
    if (!is_device_ptr(b) || !is_device_ptr(x))
        return fail(SPIKE_ERR_ARG, "the no-alloc solve takes device pointers only (host data would need staging)");
    double *scr = nullptr;
    HIP_TRY(spike::noalloc_scratch(&scr));
    // Workspaces of the largest chunk, carved from the scratch: the Spike
    // reduced vectors (y and the two Jacobi buffers), coupling right-hand
    // sides and the FP32 residual; the level method's interleaved vectors,
    // residual and correction.  Every kernel writes these before it reads
    // them, so nothing needs clearing between calls.
    const int nr0 = std::min(nrhs, std::max(f->max_nr, 1));
    const size_t nred = f->plan.dep_rows.size(), ncpl = f->plan.cpl_rows.size();
    size_t need = 0, slots = 1;
    if (f->method == SPIKE_METHOD_LEVELS) {
        if (nr0 >= 2) {
            while ((int)slots < nr0) slots <<= 1;
            slots = std::max<size_t>(slots, 2);
            need = 2 * (size_t)n * slots + (f->fp32 ? (size_t)n * slots : 0);
        } else if (f->fp32) {
            need = 2 * (size_t)n;   // correction and residual
        }
    } else {
        while ((int)slots < nr0) slots <<= 1;
        need = (3 * nred + ncpl) * slots + (f->fp32 ? (size_t)n * slots : 0);
    }
    const size_t cap = spike::kNoAllocScratchBytes / sizeof(double);
    if (need > cap) {
        char msg[200];
        std::snprintf(msg, sizeof msg, "this solve needs %zu bytes of workspace; the no-alloc solve's scratch holds %zu",
                      need * sizeof(double), spike::kNoAllocScratchBytes);
        return fail(SPIKE_ERR_UNSUPPORTED, msg);
    }
    // the factor's workspace pointers point into the scratch for this call
    struct Saved {
        double *y, *xa, *xb, *bc, *r, *bi, *xi, *dlt;
        size_t il_cap;
    } sv{f->d_y, f->d_xa, f->d_xb, f->d_bc, f->d_r, f->d_bi, f->d_xi, f->d_dlt, f->il_cap};
    double *p = scr;
    auto take = [&](size_t count) {
        double *q = p;
        p += count;
        return q;
    };
    if (f->method == SPIKE_METHOD_LEVELS) {
        if (nr0 >= 2) {
            f->d_bi = take((size_t)n * slots);
            f->d_xi = take((size_t)n * slots);
            f->il_cap = (size_t)n * slots;
            if (f->fp32) f->d_r = take((size_t)n * slots);
        } else if (f->fp32) {
            f->d_dlt = take((size_t)n);
            f->d_r = take((size_t)n);
        }
    } else {
        f->d_y = take(nred * slots);
        f->d_xa = take(nred * slots);
        f->d_xb = take(nred * slots);
        f->d_bc = take(ncpl * slots);
        if (f->fp32) f->d_r = take((size_t)n * slots);
    }
    // the factor's launch sequence, chunk by chunk (no graphs, no staging)
    int rc = SPIKE_OK;
    for (int j = 0; j < nrhs && rc == SPIKE_OK; j += f->max_nr) {
        const int nr = std::min(f->max_nr, nrhs - j);
        rc = solve_chunk(f, b + (size_t)j * ldb, ldb, x + (size_t)j * ldx, ldx, nr, (hipStream_t)stream, nullptr,
                         nullptr);
    }
    f->d_y = sv.y;
    f->d_xa = sv.xa;
    f->d_xb = sv.xb;
    f->d_bc = sv.bc;
    f->d_r = sv.r;
    f->d_bi = sv.bi;
    f->d_xi = sv.xi;
    f->d_dlt = sv.dlt;
    f->il_cap = sv.il_cap;
    return rc;
}

int spike_factor_status(spike_factor_t f)
{
    if (!f) return fail(SPIKE_ERR_ARG, "factor is NULL");
    return take_failure(f);
}

int spike_factor_save(spike_factor_t f, const char *path)
{
    if (!f || !path) return fail(SPIKE_ERR_ARG, "NULL argument");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != f->device) HIP_TRY(hipSetDevice(f->device));
    HIP_TRY(hipDeviceSynchronize());
    return factor_save(f, path);
}

int spike_factor_load(spike_factor_t *out, const char *path)
{
    if (!out || !path) return fail(SPIKE_ERR_ARG, "NULL argument");
    *out = nullptr;
    spike_factor_s *f = new spike_factor_s();
    int rc = factor_load(f, path);
    if (rc != SPIKE_OK) {
        std::string keep = t_err;
        f->release();
        delete f;
        t_err = keep;
        return rc;
    }
    *out = f;
    t_err.clear();
    return SPIKE_OK;
}

int spike_factor_destroy(spike_factor_t f)
{
    if (!f) return SPIKE_OK;
    f->release();
    delete f;
    return SPIKE_OK;
}

int spike_factor_stats(spike_factor_t f, spike_stats_t *st)
{
    if (!f || !st) return fail(SPIKE_ERR_ARG, "NULL argument");
    *st = f->stats;
    st->device_bytes = f->device_bytes;
    return SPIKE_OK;
}

int spike_factor_log(spike_factor_t f)
{
    if (!f) return fail(SPIKE_ERR_ARG, "factor is NULL");
    const Plan &pl = f->plan;
    // The reference's text byte for byte (ops.c:1317-1332): the marker lines
    // end in a space and every field line in tabs, which log parsers match
    // (the marker string) and split on (the tab after each value).
    std::printf("* * * * * * * * * * * * * * * * * * * * * * \n");
    std::printf("* Matrix Information\t\t\t\t\t\t\n");
    std::printf("* ------------------\t\t\t\t\t\t\n");
    std::printf("* Reduced system size: %d x %d\t\t\t\t\n", f->stats.n_reduced, f->stats.n_reduced);
    std::printf("* Reduced system nnzs: %lld \t\t\t\t\t\n", (long long)f->stats.nnz_reduced);
    for (int t = 0; t < pl.P; t++) {
        const spike::Part &pt = pl.parts[t];
        std::printf("* S[%d] size: %d x %d\t\t\t\t\t\n", t, pt.rows_S, pt.kcnt);
        std::printf("* isOptimized[%d]: %d\t\t\t\t\t\n", t, 0);
        std::printf("* hasReflection[%d]: %d\t\t\t\t\t\n", t, pt.dcnt > 0 ? 1 : 0);
        std::printf("* hasDependences[%d]: %d\t\t\t\t\n", t, pt.ccnt > 0 ? 1 : 0);
        std::printf("* hasToCalculateS[%d]: %d\t\t\t\t\n", t, pt.rows_S > 0 && pt.kcnt > 0 ? 1 : 0);
        std::printf("* Intervals start[%d]: %d, end[%d]: %d\t\n", t, pt.f, t, pt.g + 1);
        std::printf("* Partition start[%d]: %d, end[%d]: %d\t\n", t, pt.s, t, pt.s + pt.k);
    }
    std::printf("* * * * * * * * * * * * * * * * * * * * * * \n\n");
    std::fflush(stdout);
    return SPIKE_OK;
}

int spike_analyze(char uplo, int n, const double *a, const int *ia, const int *ja, const spike_options_t *opts,
                  spike_stats_t *st)
{
    if (!st) return fail(SPIKE_ERR_ARG, "stats pointer is NULL");
    if (uplo != 'L' && uplo != 'l' && uplo != 'U' && uplo != 'u') return fail(SPIKE_ERR_ARG, "uplo must be L or U");
    if (n < 1) return fail(SPIKE_ERR_ARG, "n must be >= 1");
    spike_options_t o;
    spike_options_default(&o);
    if (opts) o = *opts;
    spike::PlanOptions po;
    if (!ia || !ja) return fail(SPIKE_ERR_ARG, "ia and ja must be non-NULL");
    int P = o.partitions > 0 ? o.partitions : default_partitions(uplo, n, ia, ja, kmax_limit_for_device());
    po.P = std::min(P, n);
    po.lanes_per_row = o.lanes_per_row;
    po.fp32 = o.precision == SPIKE_PREC_FP32_REFINED;
    Plan pl;
    std::string err = spike::build_plan(uplo, n, a, ia, ja, po, pl);
    if (!err.empty()) return fail(SPIKE_ERR_MATRIX, err);
    std::memset(st, 0, sizeof *st);
    st->n = n;
    st->partitions = pl.P;
    st->nnz = pl.nnz;
    st->n_reduced = (int)pl.dep_rows.size();
    st->spike_area = pl.spike_area;
    st->steps_pass1 = pl.steps.max_steps;
    st->steps_pass2 = pl.steps.max_steps;
    st->lanes_per_row = pl.steps.G;
    return SPIKE_OK;
}

void spike_pstrsv_set_partitions(int P)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_partitions_override = P > 0 ? P : 0;
}

spike_factor_t spike_pstrsv_current_factor(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_cur;
}

const char *spike_last_error(void) { return t_err.c_str(); }

int spike_pstrsv_get_timers(double *t, int cap)
{
    int m = std::min(cap, 9);
    for (int i = 0; i < m; i++) t[i] = g_timers[i];
    return m;
}

void spike_pstrsv_preproc(double *a, MKL_INT *ia, MKL_INT *ja, int n, int nthreads, const char uplo)
{
    legacy_preproc(a, ia, ja, n, nthreads, uplo, "spike_pstrsv_preproc");
}

void _spike_pstrsv_preproc(double *a, MKL_INT *ia, MKL_INT *ja, int n, int nthreads, const char uplo)
{
    // timers restart per preprocess, like initTimers_PSTRSV3 (ops.c:108-120)
    reset_harness_timers();
    legacy_preproc(a, ia, ja, n, nthreads, uplo, "_spike_pstrsv_preproc");
    sync_harness_timers();
}

void spike_pstrsv(const char uplo, int m, double *a, MKL_INT *ia, MKL_INT *ja, double *b, double *x, int nthreads)
{
    legacy_preproc(a, ia, ja, m, nthreads, uplo, "spike_pstrsv");
    legacy_solve(uplo, m, b, x, "spike_pstrsv", false);
}

void spike_pstrsv_solve(const char uplo, int m, double *b, double *x, int nthreads)
{
    (void)nthreads;
    legacy_solve(uplo, m, b, x, "spike_pstrsv_solve", false);
}

void _spike_pstrsv_solve(const char uplo, int m, double *b, double *x, int nthreads)
{
    (void)nthreads;
    legacy_solve(uplo, m, b, x, "_spike_pstrsv_solve", true);
    sync_harness_timers();
}

// ---------------------------------------------------------------- batch API

void spike_batch_options_default(spike_batch_options_t *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->rows_per_partition = 0;
    o->lanes_per_row = 0;
    o->precision = SPIKE_PREC_FP64;
    o->use_graph = 1;
    o->autotune_sample = 0;
    o->drop_tol = -1.0;
    o->refine_tol = 0.0;
}

int spike_batch_create(spike_batch_t *out, char uplo, int count, const spike_batch_matrix_t *m,
                       const spike_batch_options_t *opts)
{
    if (!out) return fail(SPIKE_ERR_ARG, "batch pointer is NULL");
    *out = nullptr;
    if (uplo == 'l') uplo = 'L';
    if (uplo == 'u') uplo = 'U';
    if (uplo != 'L' && uplo != 'U') return fail(SPIKE_ERR_ARG, "uplo must be L or U");
    if (count < 1 || !m) return fail(SPIKE_ERR_ARG, "count must be >= 1 and m non-NULL");
    spike_batch_options_t o;
    spike_batch_options_default(&o);
    if (opts) o = *opts;
    if (o.precision != SPIKE_PREC_FP64 && o.precision != SPIKE_PREC_FP32_REFINED)
        return fail(SPIKE_ERR_ARG, "unknown precision mode");
    for (int i = 0; i < count; i++) {
        std::string err = check_batch_matrix(m[i], i);
        if (!err.empty()) return fail(SPIKE_ERR_MATRIX, err);
    }
    const double t0 = now_s();
    spike_batch_s *bt = new spike_batch_s();
    bt->uplo = uplo;
    bt->count = count;
    bt->opts = o;
    (void)hipGetDevice(&bt->device);
    bt->off.assign(count + 1, 0);
    for (int i = 0; i < count; i++) bt->off[i + 1] = bt->off[i] + (int64_t)m[i].n * m[i].nrhs;
    auto bail = [&](int rc) {
        std::string keep = t_err;
        bt->release();
        delete bt;
        t_err = keep;
        return rc;
    };
    int rc = SPIKE_OK;
    for (int li = 0; li < 4; li++) bt->rpp[li] = o.rows_per_partition;
    if (o.rows_per_partition <= 0) {
        const double ta = now_s();
        if ((rc = autotune_rpp(bt, m, bt->rpp))) return bail(rc);
        bt->autotune_s = now_s() - ta;
    }
    std::vector<int> idx(count);
    for (int i = 0; i < count; i++) idx[i] = i;
    if ((rc = batch_build(bt, m, idx, bt->rpp, 0))) return bail(rc);
    if (o.precision == SPIKE_PREC_FP32_REFINED && (rc = batch_calibrate(bt))) return bail(rc);
    bt->preproc_s = now_s() - t0;
    *out = bt;
    t_err.clear();
    return SPIKE_OK;
}

int64_t spike_batch_offset(spike_batch_t bt, int i)
{
    if (!bt || i < 0 || i > bt->count) return -1;
    return bt->off[i];
}

int64_t spike_batch_length(spike_batch_t bt) { return bt ? bt->off[bt->count] : -1; }

int spike_batch_solve(spike_batch_t bt, const double *b, double *x, void *stream)
{
    if (!bt) return fail(SPIKE_ERR_STATE, "batch is NULL");
    if (!b || !x) return fail(SPIKE_ERR_ARG, "b and x must be non-NULL");
    hipStream_t s = (hipStream_t)stream;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != bt->device) HIP_TRY(hipSetDevice(bt->device));
    const size_t len = (size_t)bt->off[bt->count];
    const bool bdev = is_device_ptr(b), xdev = is_device_ptr(x);
    const double *db = b;
    double *dx = x;
    if (!bdev || !xdev) {
        if (!bt->d_sb) {
            HIP_TRY(hipMalloc((void **)&bt->d_sb, sizeof(double) * len));
            HIP_TRY(hipMalloc((void **)&bt->d_sx, sizeof(double) * len));
        }
        if (!bdev) {
            HIP_TRY(hipMemcpyAsync(bt->d_sb, b, sizeof(double) * len, hipMemcpyHostToDevice, s));
            db = bt->d_sb;
        }
        if (!xdev) dx = bt->d_sx;
    }
    int rc;
    if (bt->opts.use_graph) {
        if (!bt->exec || bt->gb != db || bt->gx != dx) {
            // the capture streams must not overlap earlier work on s
            HIP_TRY(hipStreamSynchronize(s));
            if ((rc = batch_capture(bt, db, dx))) return rc;
        }
        HIP_TRY(hipGraphLaunch(bt->exec, s));
    } else if ((rc = batch_run(bt, db, dx, s))) {
        return rc;
    }
    if (!xdev) HIP_TRY(hipMemcpyAsync(x, dx, sizeof(double) * len, hipMemcpyDeviceToHost, s));
    if (!bdev || !xdev) HIP_TRY(hipStreamSynchronize(s));
    return SPIKE_OK;
}

int spike_batch_stats(spike_batch_t bt, spike_batch_stats_t *st)
{
    if (!bt || !st) return fail(SPIKE_ERR_ARG, "NULL argument");
    std::memset(st, 0, sizeof *st);
    st->count = bt->count;
    st->groups = (int)bt->groups.size();
    st->rows_rhs = bt->off[bt->count];
    for (spike_factor_s *f : bt->groups) {
        st->rows += f->plan.n;
        st->nnz += f->stats.nnz;
        st->partitions += f->plan.P;
        st->n_reduced += f->stats.n_reduced;
        st->device_bytes += f->device_bytes;
        st->refine_steps = std::max(st->refine_steps, f->refine_steps);
        st->refine_est_err = std::max(st->refine_est_err, f->refine_est);
    }
    for (int li = 0; li < 4; li++) {
        st->rows_per_partition[li] = bt->rpp[li];
        st->lane_shift[li] = bt->lane_shift[li];
    }
    st->preproc_seconds = bt->preproc_s;
    st->autotune_seconds = bt->autotune_s;
    st->graph_nodes = bt->graph_nodes;
    return SPIKE_OK;
}

int spike_batch_destroy(spike_batch_t bt)
{
    if (!bt) return SPIKE_OK;
    bt->release();
    delete bt;
    return SPIKE_OK;
}

}  // extern "C"
