// Host-side symbolic analysis for the Spike solve.  See plan.hpp.
#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace spike {

namespace {

// First defect of rows [lo, hi), or "".
std::string validate_rows(char uplo, int n, const int *ia, const int *ja, int lo, int hi)
{
    for (int i = lo; i < hi; i++) {
        if (ia[i + 1] <= ia[i]) return "row " + std::to_string(i) + " is empty (diagonal missing)";
        for (int j = ia[i]; j < ia[i + 1]; j++) {
            if (ja[j] < 0 || ja[j] >= n) return "column index out of range in row " + std::to_string(i);
            if (j > ia[i] && ja[j] <= ja[j - 1])
                return "columns of row " + std::to_string(i) + " are not strictly ascending";
        }
        if (uplo == 'L') {
            if (ja[ia[i + 1] - 1] != i) return "row " + std::to_string(i) + ": diagonal must be the last entry ('L')";
        } else {
            if (ja[ia[i]] != i) return "row " + std::to_string(i) + ": diagonal must be the first entry ('U')";
        }
    }
    return "";
}

// The first defect in row order (the message a row-by-row scan gives).  Row
// pointers are checked first, so no row's entries are read through a
// pointer that decreases or passes ia[n].
std::string validate(char uplo, int n, const int *ia, const int *ja)
{
    if (ia[0] != 0) return "ia[0] must be 0";
    const int nnz = ia[n];
    std::vector<int> bad((size_t)chunk_count(n, 1 << 18), n);
    parallel_chunks(n, 1 << 18, [&](int c, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++)
            if (ia[i + 1] <= ia[i] || ia[i + 1] > nnz) {
                bad[c] = (int)i;
                break;
            }
    });
    const int rows_ok = *std::min_element(bad.begin(), bad.end());
    std::vector<std::string> err((size_t)chunk_count(rows_ok, 65536));
    const int nc = parallel_chunks(rows_ok, 65536, [&](int c, int64_t lo, int64_t hi) {
        err[c] = validate_rows(uplo, n, ia, ja, (int)lo, (int)hi);
    });
    for (int c = 0; c < nc; c++)
        if (!err[c].empty()) return err[c];
    if (rows_ok < n) {
        if (ia[rows_ok + 1] <= ia[rows_ok]) return "row " + std::to_string(rows_ok) + " is empty (diagonal missing)";
        return "row pointer ia[" + std::to_string(rows_ok + 1) + "] exceeds ia[n]";
    }
    return "";
}

// A schedulable unit: one chunk of one row.
struct Task {
    int row;        // local row
    int eb, ee;     // entry range in the partition-local entry list
    int level;
    bool final_chunk;
};

// Builds the steps of one partition over local rows [lo, hi] (inclusive).
// Entries of row r: local columns ecol[rb[r]..rb[r+1]) (ascending, < r).
void schedule_partition(int k, int lo, int hi, const std::vector<int> &rb,
                        const std::vector<int> &ecol, const std::vector<double> &evals,
                        const std::vector<double> &diag, int G, int M, bool fwd,
                        std::vector<double> &val, std::vector<uint16_t> &col,
                        std::vector<uint16_t> &row, std::vector<double> &late, int &nsteps)
{
    const int R = 64 / G, cap = G * M;
    std::vector<Task> tasks;
    std::vector<int> final_task(k, -1);      // index of the final task of a row
    std::vector<int> task_level;
    for (int r = lo; r <= hi; r++) {
        int len = rb[r + 1] - rb[r];
        int nch = len == 0 ? 1 : (len + cap - 1) / cap;
        int prev_level = -1;
        for (int c = 0; c < nch; c++) {
            Task t;
            t.row = r;
            t.eb = rb[r] + c * cap;
            t.ee = std::min(rb[r + 1], t.eb + cap);
            int lev = prev_level + 1;  // after the previous chunk of the row
            for (int e = t.eb; e < t.ee; e++) {
                int d = ecol[e];
                if (d >= lo) lev = std::max(lev, tasks[final_task[d]].level + 1);
            }
            t.level = lev;
            t.final_chunk = (c == nch - 1);
            prev_level = lev;
            tasks.push_back(t);
        }
        final_task[r] = (int)tasks.size() - 1;
    }
    const int ntask = (int)tasks.size();
    std::vector<int> order(ntask);
    for (int i = 0; i < ntask; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return tasks[x].level < tasks[y].level;
    });
    // greedy packing: a task may not share a step with a task it depends on
    std::vector<int> step_of(ntask, -1);
    std::vector<int> prev_chunk(ntask, -1);
    for (int i = 1; i < ntask; i++)
        if (tasks[i].row == tasks[i - 1].row) prev_chunk[i] = i - 1;
    std::vector<std::vector<int>> steps;
    int cur = -1, used = R;
    for (int oi = 0; oi < ntask; oi++) {
        int ti = order[oi];
        const Task &t = tasks[ti];
        // earliest step: after every dependency, and after the previous chunk
        // of the row by >= 1 step (>= 2 in the forwarding layout, whose
        // accumulator reads run one step ahead)
        int min_step = 0;
        if (prev_chunk[ti] >= 0) min_step = step_of[prev_chunk[ti]] + (fwd ? 2 : 1);
        for (int e = t.eb; e < t.ee; e++) {
            int d = ecol[e];
            if (d >= lo) min_step = std::max(min_step, step_of[final_task[d]] + 1);
        }
        if (used >= R || cur < min_step) {
            do {
                steps.emplace_back();
                cur = (int)steps.size() - 1;
            } while (cur < min_step);
            used = 0;
        }
        steps[cur].push_back(ti);
        step_of[ti] = cur;
        used++;
    }
    nsteps = (int)steps.size();
    val.assign((size_t)nsteps * M * 64, 0.0);
    col.assign((size_t)nsteps * M * 64, (uint16_t)k);
    if (!fwd) {
        // values pre-scaled by the row's 1/diag (the right-hand side is
        // scaled on load), so every chunk of a row computes x -= sum val * x
        row.assign((size_t)nsteps * R, (uint16_t)(k + 1));
        for (int s = 0; s < nsteps; s++) {
            for (int slot = 0; slot < (int)steps[s].size(); slot++) {
                const Task &t = tasks[steps[s][slot]];
                const double inv = 1.0 / diag[t.row];
                row[(size_t)s * R + slot] = (uint16_t)t.row;
                for (int e = t.eb; e < t.ee; e++) {
                    int q = e - t.eb;
                    int m = q / G, li = q % G;
                    size_t idx = ((size_t)s * M + m) * 64 + slot * G + li;
                    val[idx] = evals[e] * inv;
                    col[idx] = (uint16_t)ecol[e];
                }
            }
        }
        return;
    }
    // forwarding layout
    std::vector<int> pos_in_step(ntask, -1);
    for (int s = 0; s < nsteps; s++)
        for (int slot = 0; slot < (int)steps[s].size(); slot++) pos_in_step[steps[s][slot]] = slot;
    row.assign((size_t)nsteps * R, (uint16_t)(k + 1));
    late.assign((size_t)nsteps * R * R, 0.0);
    for (int s = 0; s < nsteps; s++) {
        for (int slot = 0; slot < (int)steps[s].size(); slot++) {
            const Task &t = tasks[steps[s][slot]];
            const double inv = 1.0 / diag[t.row];
            row[(size_t)s * R + slot] = (uint16_t)t.row;
            int q = 0;
            for (int e = t.eb; e < t.ee; e++) {
                const int d = ecol[e];
                const int dt = d >= lo ? final_task[d] : -1;
                if (dt >= 0 && step_of[dt] == s - 1) {
                    late[((size_t)s * R + slot) * R + pos_in_step[dt]] = evals[e] * inv;
                    continue;
                }
                int m = q / G, li = q % G;
                size_t idx = ((size_t)s * M + m) * 64 + slot * G + li;
                val[idx] = evals[e] * inv;
                col[idx] = (uint16_t)d;
                q++;
            }
        }
    }
}

}  // namespace

std::vector<int> row_levels(int m, const int *rp, const int *ci)
{
    std::vector<int> lev(m, 0);
    for (int i = 0; i < m; i++) {
        int l = 0;
        for (int j = rp[i]; j < rp[i + 1]; j++) l = std::max(l, lev[ci[j]] + 1);
        lev[i] = l;
    }
    return lev;
}

int choose_lanes(const std::vector<int> &row_len)
{
    if (row_len.empty()) return 1;
    std::vector<int> v(row_len);
    size_t q = (v.size() * 9) / 10;
    std::nth_element(v.begin(), v.begin() + q, v.end());
    int p90 = v[q];
    int G = 1;
    while (G < p90 && G < 16) G <<= 1;
    return G;
}

int choose_block_lanes(const std::vector<int> &row_len)
{
    if (row_len.empty()) return 1;
    std::vector<int> v(row_len);
    size_t q = (v.size() * 9) / 10;
    std::nth_element(v.begin(), v.begin() + q, v.end());
    const int p90 = v[q];
    int G = 1;
    while (G * 2 * 3 <= p90 && G < 16) G <<= 1;
    return G;
}

std::string build_plan(char uplo, int n, const double *a, const int *ia, const int *ja,
                       const PlanOptions &opt, Plan &pl)
{
    if (uplo == 'l') uplo = 'L';
    if (uplo == 'u') uplo = 'U';
    if (uplo != 'L' && uplo != 'U') return "uplo must be 'L' or 'U'";
    if (n < 1) return "n must be >= 1";
    if (!a || !ia || !ja) return "a, ia and ja must be non-NULL";
    int P = opt.starts.empty() ? opt.P : (int)opt.starts.size() - 1;
    if (P < 1 || P > n) return "partition count must be in [1, n]";
    if (!opt.starts.empty()) {
        if (opt.starts[0] != 0 || opt.starts[P] != n) return "partition starts must begin at 0 and end at n";
        for (int t = 0; t < P; t++)
            if (opt.starts[t + 1] <= opt.starts[t]) return "partition starts must be strictly increasing";
    }
    std::string err = validate(uplo, n, ia, ja);
    if (!err.empty()) return err;
    const int64_t nnz = ia[n];
    if (nnz > INT32_MAX) return "nnz exceeds 2^31-1";

    pl = Plan();
    pl.uplo = uplo;
    pl.flip = (uplo == 'U');
    pl.n = n;
    pl.P = P;
    pl.nnz = nnz;
    resize_large(pl.ia, n + 1);
    resize_large(pl.ja, nnz);
    resize_large(pl.a, nnz);
    if (!pl.flip) {
        std::copy(ia, ia + n + 1, pl.ia.begin());
        parallel_chunks(nnz, 1 << 20, [&](int, int64_t lo, int64_t hi) {
            std::copy(ja + lo, ja + hi, pl.ja.begin() + lo);
            std::copy(a + lo, a + hi, pl.a.begin() + lo);
        });
    } else {
        // U = J L J: row i of L is row n-1-i of U reversed, columns mirrored
        pl.ia[0] = 0;
        for (int i = 0; i < n; i++) {
            int src = n - 1 - i;
            pl.ia[i + 1] = pl.ia[i] + (ia[src + 1] - ia[src]);
        }
        parallel_chunks(n, 1 << 14, [&](int, int64_t lo, int64_t hi) {
            for (int i = (int)lo; i < (int)hi; i++) {
                const int src = n - 1 - i;
                int p = pl.ia[i];
                for (int j = ia[src + 1] - 1; j >= ia[src]; j--, p++) {
                    pl.a[p] = a[j];
                    pl.ja[p] = n - 1 - ja[j];
                }
            }
        });
    }
    const int *LA = pl.ia.data();
    const int *LJ = pl.ja.data();
    const double *LV = pl.a.data();

    // a zero or non-finite diagonal would be divided by: rejected, naming the
    // first such row in the caller's order (the last in lower-form order for
    // 'U')
    {
        const int nc = chunk_count(n, 1 << 16);
        std::vector<int> first(nc, n), last(nc, -1);
        parallel_chunks(n, 1 << 16, [&](int c, int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; i++) {
                const double d = LV[LA[i + 1] - 1];
                if (d == 0.0 || !std::isfinite(d)) {
                    first[c] = std::min(first[c], (int)i);
                    last[c] = (int)i;
                }
            }
        });
        const int lo = *std::min_element(first.begin(), first.end());
        const int hi = *std::max_element(last.begin(), last.end());
        if (lo < n)
            return "row " + std::to_string(pl.flip ? n - 1 - hi : lo) + ": the diagonal is zero or not finite";
    }

    // partitions (matrix_partitioning.c:18-33): equal blocks, remainder last;
    // or the caller's boundaries (the reference's IS_MATRIX_PARTITIONED path,
    // matrix_partitioning.c:23, :34-35)
    pl.starts.resize(P + 1);
    if (opt.starts.empty()) {
        int kb = n / P;
        for (int t = 0; t < P; t++) pl.starts[t] = kb * t;
        pl.starts[P] = n;
    } else {
        pl.starts = opt.starts;
    }
    pl.kmax = 0;
    for (int t = 0; t < P; t++) pl.kmax = std::max(pl.kmax, pl.starts[t + 1] - pl.starts[t]);
    if (pl.kmax > opt.kmax_limit)
        return "partition of " + std::to_string(pl.kmax) + " rows exceeds the device limit of " +
               std::to_string(opt.kmax_limit) + " rows; use more partitions";

    // dependents (ops.c:162-181, lower branch)
    std::vector<char> isdep(n, 0);
    pl.parts.resize(P);
    for (int t = 0; t < P; t++) {
        Part &pt = pl.parts[t];
        pt.s = pl.starts[t];
        pt.k = pl.starts[t + 1] - pt.s;
    }
    for (int t = 0; t < P; t++) {
        // coupling entries come first in a row (ascending columns)
        const int s = pl.starts[t];
        for (int i = s; i < pl.starts[t + 1]; i++)
            for (int j = LA[i]; j < LA[i + 1] && LJ[j] < s; j++) isdep[LJ[j]] = 1;
    }
    pl.ridx.assign(n, -1);
    for (int i = 0; i < n; i++)
        if (isdep[i]) {
            pl.ridx[i] = (int)pl.dep_rows.size();
            pl.dep_rows.push_back(i);
        }

    // intervals, K_p, dependents and coupling rows per partition
    std::vector<std::vector<int>> kc(P);
    parallel_for(P, [&](int t) {
        Part &pt = pl.parts[t];
        int s = pt.s, k = pt.k;
        pt.f = k;
        pt.g = -1;
        for (int r = 0; r < k; r++) {
            int i = s + r;
            if (pt.f == k && LJ[LA[i]] < s) pt.f = r;
            if (isdep[i]) pt.g = r;
        }
        pt.hi1 = (pt.f == k) ? k - 1 : std::max(pt.g, pt.f - 1);
        if (pt.hi1 > k - 1) pt.hi1 = k - 1;
        std::vector<int> &K = kc[t];
        for (int r = pt.f; r <= pt.g; r++) {
            int i = s + r;
            for (int j = LA[i]; j < LA[i + 1] && LJ[j] < s; j++) K.push_back(LJ[j]);
        }
        std::sort(K.begin(), K.end());
        K.erase(std::unique(K.begin(), K.end()), K.end());
        pt.kcnt = (int)K.size();
        pt.rows_S = pt.kcnt > 0 ? pt.g - pt.f + 1 : 0;
    });
    {
        int acc = 0;
        for (int t = 0; t < P; t++) {
            pl.parts[t].kbeg = acc;
            acc += pl.parts[t].kcnt;
            pl.spike_area += (int64_t)pl.parts[t].rows_S * pl.parts[t].kcnt;
        }
        pl.kcols.resize(acc);
        for (int t = 0; t < P; t++) std::copy(kc[t].begin(), kc[t].end(), pl.kcols.begin() + pl.parts[t].kbeg);
        // dependents are ascending, so each partition's are contiguous
        size_t d = 0;
        for (int t = 0; t < P; t++) {
            Part &pt = pl.parts[t];
            pt.dbeg = (int)d;
            while (d < pl.dep_rows.size() && pl.dep_rows[d] < pt.s + pt.k) d++;
            pt.dcnt = (int)d - pt.dbeg;
        }
        // coupling rows (b' = b - C x_K in pass 2)
        pl.cpl_ptr.push_back(0);
        for (int t = 0; t < P; t++) {
            Part &pt = pl.parts[t];
            pt.cbeg = (int)pl.cpl_rows.size();
            for (int r = pt.f; r < pt.k; r++) {
                int i = pt.s + r;
                if (LJ[LA[i]] >= pt.s) continue;
                pl.cpl_rows.push_back(i);
                for (int j = LA[i]; j < LA[i + 1] && LJ[j] < pt.s; j++) {
                    pl.cpl_col.push_back(pl.ridx[LJ[j]]);
                    pl.cpl_val.push_back(LV[j]);
                }
                pl.cpl_ptr.push_back((int)pl.cpl_col.size());
            }
            pt.ccnt = (int)pl.cpl_rows.size() - pt.cbeg;
        }
    }

    // block-solve step streams
    std::vector<int> inblock_len(n);
    parallel_for(P, [&](int t) {
        const Part &pt = pl.parts[t];
        for (int r = 0; r < pt.k; r++) {
            int i = pt.s + r, c = 0;
            for (int j = LA[i]; j < LA[i + 1] - 1; j++) c += (LJ[j] >= pt.s);
            inblock_len[i] = c;
        }
    });
    int G = opt.lanes_per_row > 0 ? opt.lanes_per_row
                                  : (opt.fp32 ? choose_lanes(inblock_len) : choose_block_lanes(inblock_len));
    if (G != 1 && G != 2 && G != 4 && G != 8 && G != 16) return "lanes_per_row must be 1, 2, 4, 8 or 16";
    const int M = 1;
    const bool fwd = opt.forward && G == 16;
    StepStream &ss = pl.steps;
    ss.G = G; ss.M = M; ss.R = 64 / G; ss.fwd = fwd;
    resize_large(pl.inv_diag, n);
    parallel_chunks(n, 1 << 16, [&](int, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) pl.inv_diag[i] = 1.0 / LV[LA[i + 1] - 1];
    });
    std::vector<std::vector<double>> pv(P), pl_late(P);
    std::vector<std::vector<uint16_t>> pc(P), pr(P);
    std::vector<int> pn(P, 0);
    parallel_for(P, [&](int t) {
        const Part &pt = pl.parts[t];
        std::vector<int> rb(pt.k + 1, 0), ecol;
        std::vector<double> ev, dg(pt.k);
        for (int r = 0; r < pt.k; r++) {
            int i = pt.s + r;
            for (int j = LA[i]; j < LA[i + 1] - 1; j++)
                if (LJ[j] >= pt.s) { ecol.push_back(LJ[j] - pt.s); ev.push_back(LV[j]); }
            rb[r + 1] = (int)ecol.size();
            dg[r] = LV[LA[i + 1] - 1];
        }
        schedule_partition(pt.k, 0, pt.k - 1, rb, ecol, ev, dg, G, M, fwd, pv[t], pc[t], pr[t], pl_late[t],
                           pn[t]);
    });
    ss.part_begin.assign(P + 1, 0);
    for (int t = 0; t < P; t++) {
        ss.part_begin[t + 1] = ss.part_begin[t] + pn[t];
        ss.max_steps = std::max(ss.max_steps, pn[t]);
    }
    ss.nsteps = ss.part_begin[P];
    resize_large(ss.val, (size_t)ss.nsteps * M * 64);
    resize_large(ss.col, (size_t)ss.nsteps * M * 64);
    resize_large(ss.row, (size_t)ss.nsteps * ss.R);
    if (fwd) ss.late.resize((size_t)ss.nsteps * ss.R * ss.R);
    parallel_for(P, [&](int t) {
        size_t s0 = (size_t)ss.part_begin[t];
        std::copy(pv[t].begin(), pv[t].end(), ss.val.begin() + s0 * M * 64);
        std::copy(pc[t].begin(), pc[t].end(), ss.col.begin() + s0 * M * 64);
        std::copy(pr[t].begin(), pr[t].end(), ss.row.begin() + s0 * ss.R);
        if (fwd) std::copy(pl_late[t].begin(), pl_late[t].end(), ss.late.begin() + s0 * ss.R * ss.R);
    });
    return "";
}

namespace {

// Far slots for one partition: last[c] is the last row that reads spike row
// c from R or more rows back (-1: none).  Row r first reads, then writes its
// own far row, so a slot whose row is last read at r is free for r itself.
// Returns the number of slots; slot (if not null) receives each row's slot.
int color_far(const int *last, int rows, int *slot)
{
    std::vector<int> head(rows, -1), nxt(rows, -1), mine(rows, -1), freel;
    for (int c = 0; c < rows; c++)
        if (last[c] >= 0) {
            nxt[c] = head[last[c]];
            head[last[c]] = c;
        }
    int F = 0;
    for (int r = 0; r < rows; r++) {
        for (int c = head[r]; c >= 0; c = nxt[c]) freel.push_back(mine[c]);
        if (last[r] >= 0) {
            if (freel.empty()) mine[r] = F++;
            else {
                mine[r] = freel.back();
                freel.pop_back();
            }
        }
    }
    if (slot) std::copy(mine.begin(), mine.end(), slot);
    return F;
}

}  // namespace

SpikeLayout plan_spike_layout(const Plan &pl, size_t lds_limit, size_t (*lds)(int, int, int, int),
                              int (*stage_entries)(int), bool allow_gfar)
{
    SpikeLayout L;
    const int P = pl.P;
    L.slot0.assign(P, -1);
    L.woff.assign(P, -1);
    int nrows = 0;
    for (int t = 0; t < P; t++) {
        const Part &pt = pl.parts[t];
        if (pt.kcnt <= 0 || pt.rows_S <= 0) continue;
        L.slot0[t] = nrows;
        nrows += pt.rows_S;
        L.nsp++;
        L.max_kcnt = std::max(L.max_kcnt, pt.kcnt);
        L.max_rows = std::max(L.max_rows, pt.rows_S);
    }
    if (L.nsp == 0) return L;
    const int bt = std::max(64, std::min(256, (L.max_kcnt + 63) / 64 * 64));
    L.block_threads = bt;
    L.kstride = L.max_kcnt > bt ? bt : L.max_kcnt;
    L.rinfo.assign(4 * (size_t)nrows, -1);

    // per spike row: coupling entries (rinfo.z), 4-entry chunks of the other
    // off-diagonal entries (rinfo.w); last far readers per ring candidate
    constexpr int NC = 4;
    const int cand[NC] = {16, 32, 64, 128};
    std::vector<int> fmax(P * NC, 0);
    std::vector<int64_t> nfar(P * NC, 0);   // entries read from R or more rows back
    parallel_for(P, [&](int t) {
        const Part &pt = pl.parts[t];
        if (L.slot0[t] < 0) return;
        const int rows = pt.rows_S, base = pt.s + pt.f;
        int *ri = L.rinfo.data() + 4 * (size_t)L.slot0[t];
        std::vector<int> last((size_t)NC * rows, -1);
        for (int r = 0; r < rows; r++) {
            int ncp = 0, nin = 0;
            for (int e = pl.ia[base + r]; e < pl.ia[base + r + 1] - 1; e++) {
                if (pl.ja[e] < pt.s) ncp++;
                else nin++;
                const int c = pl.ja[e] - base;
                if (c < 0) continue;
                for (int q = 0; q < NC; q++)
                    if (r - c >= cand[q]) {
                        last[(size_t)q * rows + c] = r;
                        nfar[t * NC + q]++;
                    }
            }
            ri[4 * r + 2] = ncp;
            ri[4 * r + 3] = (nin + 3) / 4;
        }
        for (int q = 0; q < NC; q++) fmax[t * NC + q] = color_far(last.data() + (size_t)q * rows, rows, nullptr);
    });
    // staged row layout: [1/diagonal, coupling entries, pad to 4 | chunks]
    L.pia.assign((size_t)nrows + 1, 0);
    bool pia_ok = true;
    for (int i = 0; i < nrows && pia_ok; i++) {
        const int64_t len = (int64_t)(1 + L.rinfo[4 * (size_t)i + 2] + 3) / 4 * 4 + 4 * (int64_t)L.rinfo[4 * (size_t)i + 3];
        const int64_t nx = L.pia[i] + len;
        pia_ok = nx <= INT32_MAX;
        L.pia[i + 1] = pia_ok ? (int)nx : 0;
    }
    int best = -1;
    int64_t best_rows = 0;
    for (int q = 0; q < NC; q++) {
        int F = 0;
        for (int t = 0; t < P; t++) F = std::max(F, fmax[t * NC + q]);
        const int64_t nl = cand[q] + 1 + F;
        if (best < 0 || nl < best_rows || (nl == best_rows && cand[q] > cand[best])) {
            best = q;
            best_rows = nl;
        }
    }
    // rows per stage: the most (<= 32) whose staged entries fit a stage
    int sb = 0, cap = 0;
    for (int cs = 32; cs >= 1 && sb == 0 && pia_ok; cs >>= 1) {
        int m = 0;
        for (int t = 0; t < P; t++) {
            if (L.slot0[t] < 0) continue;
            const int *pr = L.pia.data() + L.slot0[t], rows = pl.parts[t].rows_S;
            for (int r0 = 0; r0 < rows; r0 += cs) m = std::max(m, pr[std::min(r0 + cs, rows)] - pr[r0]);
        }
        if (m <= stage_entries(bt)) {
            sb = cs;
            cap = m;
        }
    }
    const bool stage_ok = sb > 0 && sb < bt;
    L.ring = stage_ok && best_rows <= INT32_MAX && lds((int)best_rows, L.kstride, sb, cap) <= lds_limit;
    if (!L.ring && stage_ok && allow_gfar) {
        // far rows in W instead of LDS: the largest ring that fits, when it
        // leaves at most two far entries per row (each a memory round trip
        // on the row's chain; with more, the direct build is faster: a
        // 4096-matrix batch of bandwidths up to 128 built in 274 ms of
        // kernel time this way against 189 direct, while 2D Laplacians with
        // one far entry per row took 136 ms against 204 for 64 serving
        // matrices)
        for (int q = NC - 1; q >= 0 && !L.ring; q--) {
            int64_t far = 0;
            for (int t = 0; t < P; t++) far += nfar[t * NC + q];
            if (far <= 2LL * nrows && lds(cand[q] + 1, L.kstride, sb, cap) <= lds_limit) {
                L.ring = L.gfar = true;
                best = q;
                best_rows = cand[q] + 1;
            }
        }
    }
    if (L.ring) {
        L.R = cand[best];
        L.nlr = (int)best_rows;
        L.sb = sb;
        L.cap = cap;
        parallel_for(P, [&](int t) {
            const Part &pt = pl.parts[t];
            if (L.slot0[t] < 0) return;
            const int rows = pt.rows_S, base = pt.s + pt.f;
            std::vector<int> last(rows, -1), slot(rows, -1);
            for (int r = 0; r < rows; r++)
                for (int e = pl.ia[base + r]; e < pl.ia[base + r + 1] - 1; e++) {
                    const int c = pl.ja[e] - base;
                    if (c >= 0 && r - c >= L.R) last[c] = r;
                }
            int *ri = L.rinfo.data() + 4 * (size_t)L.slot0[t];
            if (L.gfar) {
                // far-read rows get W slots (numbered below with the dependents)
                for (int r = 0; r < rows; r++) ri[4 * r] = last[r] >= 0 ? 1 : -1;
                return;
            }
            color_far(last.data(), rows, slot.data());
            for (int r = 0; r < rows; r++) ri[4 * r + 1] = slot[r];
        });
    }
    // W slots: the dependents (and, far rows in W, the far-read rows) for the
    // ring build, every spike row for the direct build
    for (int t = 0; t < P; t++) {
        const Part &pt = pl.parts[t];
        if (L.slot0[t] < 0) continue;
        int *ri = L.rinfo.data() + 4 * (size_t)L.slot0[t];
        int kept = 0;
        if (L.ring) {
            for (int q = pt.dbeg; q < pt.dbeg + pt.dcnt; q++) {
                const int r = pl.dep_rows[q] - pt.s;
                if (r >= pt.f && r <= pt.g) ri[4 * (r - pt.f)] = 1;
            }
            for (int r = 0; r < pt.rows_S; r++) ri[4 * r] = ri[4 * r] > 0 ? kept++ : -1;
        } else {
            for (int r = 0; r < pt.rows_S; r++) ri[4 * r] = kept++;
        }
        L.woff[t] = L.area;
        L.area += (int64_t)kept * pt.kcnt;
    }
    return L;
}

}  // namespace spike
