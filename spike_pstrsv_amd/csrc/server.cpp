// Serving loop over resident factors (extension API, include/spike_pstrsv.h).
//
// Factors stay preprocessed in device memory; a matrix is admitted when the
// resident footprint (factor + its staging buffers) fits the HBM budget, and
// an add that does not fit waits for removals to free memory.  Per-matrix
// state lives in heap records that never move, and a launched solve carries
// copies of what it needs, so matrices can be added and removed while
// requests are served.
// Solve requests (device pointers) are queued by spike_server_submit.  A
// dispatcher thread closes a batching window window_us after the oldest
// waiting request arrived, merges the window's requests of each matrix into
// solves of up to max_batch_rhs right-hand sides, and launches them on a pool
// of streams (a matrix with solves in flight keeps their stream, so the
// solves of one factor stay ordered and never share its workspaces; an idle
// matrix takes the least busy stream).  Right-hand sides are
// gathered into the matrix's fixed staging buffers and solutions scattered
// back, so every solve of a matrix has the same pointers and replays the
// factor's captured hipGraph.  At most `slots` solves are in flight; requests
// that arrive meanwhile wait and are admitted as solves complete (continuous
// batching).  Completion is observed by polling events; latency = completion
// time - submit time.
#include "spike_pstrsv.h"
#include "kernels.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Request {
    uint64_t ticket;
    int id, nrhs, ldb, ldx;
    const double *b;
    double *x;
    clk::time_point t_submit;
};

struct InFlight {
    hipEvent_t done;
    std::vector<Request> reqs;
    int64_t rows_rhs;
    int id, sidx;
};

// One resident matrix.  Records are heap objects owned by the server and
// never move or disappear while the server lives (removal frees the factor
// and staging and marks the record), so a pointer to one stays valid.
struct MatEntry {
    spike_factor_t f = nullptr;
    int n = 0;
    int64_t bytes = 0;
    double *stage = nullptr;  // b then x, n x max_batch_rhs each
    int sidx = -1, busy = 0;  // stream of its solves in flight, their count
    int64_t open = 0;         // requests queued or in flight
    bool removed = false;
};

// A merged solve: everything launch() reads is copied here under the lock.
struct Job {
    int id, cols, sidx;
    spike_factor_t f;
    int n;
    double *stage;
    std::vector<Request> reqs;
    hipEvent_t ev;
};

}  // namespace

struct spike_server_s {
    spike_server_options_t opts{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done, cv_admit;
    std::vector<std::unique_ptr<MatEntry>> mats;
    int64_t resident = 0;
    int admitting = 0;                    // adds waiting for memory
    std::deque<Request> pending;
    std::vector<InFlight> inflight;
    std::vector<hipStream_t> streams;
    // Stream choice: a matrix with solves in flight (or being launched) keeps
    // their stream, so its solves stay ordered; otherwise it takes the stream
    // with the fewest solves in flight.
    std::vector<int> stream_busy;
    std::vector<hipEvent_t> spare_events;
    uint64_t next_ticket = 1;
    std::map<uint64_t, int> finished;     // ticket -> status (until collected)
    std::vector<double> lat;              // seconds, per completed request
    int64_t rows_rhs_done = 0;
    int64_t solves = 0, windows = 0, merged_rhs = 0;
    int max_inflight = 0;
    clk::time_point t_first, t_last;
    bool have_first = false;
    int inflight_reserved = 0;
    int error = SPIKE_OK;
    std::string error_msg;
    std::atomic<bool> stop{false};
    std::thread worker;

    void run();
    void dispatch_window(std::unique_lock<std::mutex> &lk);
    void poll_done(std::unique_lock<std::mutex> &lk);
    int launch(const Job &j);
    int pick_stream(int id);
    void release_stream(int id, int sidx);
};

namespace {

thread_local std::string s_err;

int sfail(int code, const std::string &m)
{
    s_err = m;
    return code;
}

bool device_ptr(const void *p)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.isManaged;
}

}  // namespace

// Polls in-flight solves and retires the finished ones (mu held).
void spike_server_s::poll_done(std::unique_lock<std::mutex> &lk)
{
    (void)lk;
    for (size_t i = 0; i < inflight.size();) {
        hipError_t q = hipEventQuery(inflight[i].done);
        if (q == hipErrorNotReady) {
            i++;
            continue;
        }
        const clk::time_point now = clk::now();
        const int st = q == hipSuccess ? SPIKE_OK : SPIKE_ERR_DEVICE;
        for (const Request &r : inflight[i].reqs) {
            finished[r.ticket] = st;
            lat.push_back(secs(r.t_submit, now));
        }
        mats[inflight[i].id]->open -= (int64_t)inflight[i].reqs.size();
        rows_rhs_done += inflight[i].rows_rhs;
        release_stream(inflight[i].id, inflight[i].sidx);
        t_last = now;
        spare_events.push_back(inflight[i].done);
        inflight[i] = std::move(inflight.back());
        inflight.pop_back();
        cv_done.notify_all();
    }
}

// Stream of a new solve of matrix id (mu held); counted until released.
int spike_server_s::pick_stream(int id)
{
    MatEntry &m = *mats[id];
    int sidx = m.busy > 0 ? m.sidx : -1;
    if (sidx < 0) {
        sidx = 0;
        for (int k = 1; k < (int)streams.size(); k++)
            if (stream_busy[k] < stream_busy[sidx]) sidx = k;
    }
    m.sidx = sidx;
    m.busy++;
    stream_busy[sidx]++;
    return sidx;
}

void spike_server_s::release_stream(int id, int sidx)
{
    mats[id]->busy--;
    stream_busy[sidx]--;
}

// Launches one merged solve (mu not held): gather, solve, scatter, event.
// Reads only the job and the immutable options and stream pool.
int spike_server_s::launch(const Job &j)
{
    const int n = j.n;
    hipStream_t s = streams[j.sidx];
    double *mb = j.stage, *mx = j.stage + (size_t)n * opts.max_batch_rhs;
    // every column of the merged requests moves in one copy launch (per
    // spike::kMaxCopyCols columns) each way
    auto copy_cols = [&](bool in) {
        spike::ColCopy cc{};
        int c = 0, k = 0;
        for (const Request &r : j.reqs)
            for (int t = 0; t < r.nrhs; t++, c++) {
                double *st = (in ? mb : mx) + (size_t)c * n;
                cc.src[k] = in ? r.b + (size_t)t * r.ldb : st;
                cc.dst[k] = in ? st : r.x + (size_t)t * r.ldx;
                if (++k == spike::kMaxCopyCols) {
                    if (spike::launch_copy_cols(n, k, cc, s) != hipSuccess) return false;
                    k = 0;
                }
            }
        return spike::launch_copy_cols(n, k, cc, s) == hipSuccess;
    };
    if (!copy_cols(true)) return SPIKE_ERR_DEVICE;
    // solve a power-of-two column count (the kernels' slot count): columns
    // past j.cols hold earlier right-hand sides and their results are unused
    int padded = 1;
    while (padded < j.cols) padded <<= 1;
    padded = std::min(padded, opts.max_batch_rhs);
    int rc = spike_factor_solve(j.f, padded, mb, n, mx, n, s);
    if (rc != SPIKE_OK) return rc;
    if (!copy_cols(false)) return SPIKE_ERR_DEVICE;
    return hipEventRecord(j.ev, s) == hipSuccess ? SPIKE_OK : SPIKE_ERR_DEVICE;
}

// Dispatches every request whose batching window has closed: the oldest
// window's requests and, when solves are backed up, every older-than-a-window
// request behind them, so a backlog is merged into multi-RHS solves (mu held
// on entry and exit; released while launching).  Matrices are served in the
// order of their oldest request; solves beyond the free slots stay pending.
void spike_server_s::dispatch_window(std::unique_lock<std::mutex> &lk)
{
    windows++;
    const auto win = std::chrono::nanoseconds((int64_t)(opts.window_us * 1e3));
    const clk::time_point cutoff = std::max(pending.front().t_submit + win, clk::now() - win);
    std::vector<std::pair<int, std::vector<Request>>> by_mat;
    std::map<int, size_t> slot_of;
    std::deque<Request> rest;
    while (!pending.empty()) {
        Request r = pending.front();
        pending.pop_front();
        if (r.t_submit > cutoff) {
            rest.push_back(r);
            continue;
        }
        auto it = slot_of.find(r.id);
        if (it == slot_of.end()) {
            slot_of[r.id] = by_mat.size();
            by_mat.emplace_back(r.id, std::vector<Request>{r});
        } else {
            by_mat[it->second].second.push_back(r);
        }
    }
    std::vector<Job> jobs;
    std::vector<Request> keep;
    int free_slots = opts.slots - (int)inflight.size();
    for (auto &kv : by_mat) {
        std::vector<Request> &rs = kv.second;
        size_t i = 0;
        while (i < rs.size()) {
            if (free_slots <= 0) {
                keep.insert(keep.end(), rs.begin() + i, rs.end());
                break;
            }
            Job j;
            j.id = kv.first;
            j.cols = 0;
            j.f = mats[j.id]->f;
            j.n = mats[j.id]->n;
            j.stage = mats[j.id]->stage;
            while (i < rs.size() && j.cols + rs[i].nrhs <= opts.max_batch_rhs) {
                j.cols += rs[i].nrhs;
                j.reqs.push_back(rs[i++]);
            }
            if (!spare_events.empty()) {
                j.ev = spare_events.back();
                spare_events.pop_back();
            } else if (hipEventCreateWithFlags(&j.ev, hipEventDisableTiming) != hipSuccess) {
                keep.insert(keep.end(), j.reqs.begin(), j.reqs.end());
                keep.insert(keep.end(), rs.begin() + i, rs.end());
                break;
            }
            j.sidx = pick_stream(j.id);
            jobs.push_back(std::move(j));
            free_slots--;
        }
    }
    // waiting requests go back to the front, in arrival order
    std::vector<Request> back(keep.begin(), keep.end());
    back.insert(back.end(), rest.begin(), rest.end());
    std::sort(back.begin(), back.end(), [](const Request &a, const Request &b) { return a.t_submit < b.t_submit; });
    for (auto it = back.rbegin(); it != back.rend(); ++it) pending.push_front(*it);
    if (jobs.empty()) return;
    // reserve the slots, then launch without the lock
    std::vector<int> rcs(jobs.size(), SPIKE_OK);
    inflight_reserved += (int)jobs.size();
    lk.unlock();
    for (size_t k = 0; k < jobs.size(); k++) rcs[k] = launch(jobs[k]);
    lk.lock();
    inflight_reserved -= (int)jobs.size();
    for (size_t k = 0; k < jobs.size(); k++) {
        Job &j = jobs[k];
        if (rcs[k] != SPIKE_OK) {
            error = rcs[k];
            error_msg = spike_last_error();
            (void)hipStreamSynchronize(streams[j.sidx]);
            release_stream(j.id, j.sidx);
            for (const Request &r : j.reqs) finished[r.ticket] = rcs[k];
            mats[j.id]->open -= (int64_t)j.reqs.size();
            spare_events.push_back(j.ev);
            cv_done.notify_all();
            continue;
        }
        InFlight fl;
        fl.done = j.ev;
        fl.reqs = std::move(j.reqs);
        fl.rows_rhs = (int64_t)j.n * j.cols;
        fl.id = j.id;
        fl.sidx = j.sidx;
        inflight.push_back(std::move(fl));
        solves++;
        if (j.cols > 1) merged_rhs += j.cols;
    }
    max_inflight = std::max(max_inflight, (int)inflight.size());
}

void spike_server_s::run()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (!stop.load()) {
        poll_done(lk);
        if (pending.empty()) {
            if (inflight.empty())
                cv_submit.wait_for(lk, std::chrono::milliseconds(5));
            else
                cv_submit.wait_for(lk, std::chrono::microseconds(20));
            continue;
        }
        const clk::time_point close =
            pending.front().t_submit + std::chrono::nanoseconds((int64_t)(opts.window_us * 1e3));
        const clk::time_point now = clk::now();
        if (now < close) {
            // keep retiring solves while the window fills
            const auto wait = std::min<clk::duration>(close - now, std::chrono::microseconds(20));
            cv_submit.wait_for(lk, wait);
            continue;
        }
        if ((int)inflight.size() >= opts.slots) {
            // no free slot: wait for a completion
            cv_submit.wait_for(lk, std::chrono::microseconds(20));
            continue;
        }
        dispatch_window(lk);
    }
    // drain
    while (!inflight.empty()) {
        poll_done(lk);
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
    }
}

extern "C" {

void spike_server_options_default(spike_server_options_t *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->hbm_budget = 0;
    o->slots = 64;
    o->streams = 8;
    o->window_us = 200.0;
    // 16: the serving workload (64 matrices, 20k req/s offered) served
    // 19.0-19.1k req/s with merges of up to 16 columns, against 17.8-17.9k
    // with 8 and 18.6k with 32 (a level solve costs about the same from 2 to
    // 16 right-hand sides)
    o->max_batch_rhs = 16;
    o->admit_timeout_ms = -1.0;
}

int spike_server_create(spike_server_t *out, const spike_server_options_t *opts)
{
    if (!out) return sfail(SPIKE_ERR_ARG, "server pointer is NULL");
    *out = nullptr;
    spike_server_options_t o;
    spike_server_options_default(&o);
    if (opts) o = *opts;
    if (o.streams <= 0) o.streams = 8;
    if (o.window_us < 0) o.window_us = 200.0;
    if (o.max_batch_rhs <= 0) o.max_batch_rhs = 16;
    if (o.slots <= 0) o.slots = 64;
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return sfail(SPIKE_ERR_DEVICE, "hipMemGetInfo failed");
    if (o.hbm_budget == 0) o.hbm_budget = (size_t)(freeb * 0.8);
    spike_server_s *s = new spike_server_s();
    s->opts = o;
    (void)hipGetDevice(&s->device);
    for (int i = 0; i < o.streams; i++) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t t : s->streams) (void)hipStreamDestroy(t);
            delete s;
            return sfail(SPIKE_ERR_DEVICE, "hipStreamCreate failed");
        }
        s->streams.push_back(st);
    }
    s->stream_busy.assign(s->streams.size(), 0);
    s->worker = std::thread([s]() { s->run(); });
    *out = s;
    return SPIKE_OK;
}

int spike_server_add(spike_server_t s, char uplo, int n, const double *a, const int *ia, const int *ja,
                     const spike_options_t *fo, int *id)
{
    if (!s || !id) return sfail(SPIKE_ERR_ARG, "NULL argument");
    *id = -1;
    // a served factor replays captured graphs for every solve (one launch
    // per solve keeps the dispatcher thread's host time small) unless the
    // caller turned graphs off
    spike_options_t fopt;
    if (fo) fopt = *fo;
    else spike_options_default(&fopt);
    if (fopt.use_graph == 1) fopt.use_graph = 2;
    const int64_t stage_bytes = (int64_t)n * s->opts.max_batch_rhs * 2 * (int64_t)sizeof(double);
    spike_factor_t f = nullptr;
    int rc = spike_factor_create(&f, uplo, n, a, ia, ja, &fopt);
    if (rc != SPIKE_OK) return sfail(rc, spike_last_error());
    spike_stats_t st;
    spike_factor_stats(f, &st);
    int64_t need = st.device_bytes + stage_bytes;
    if (need > (int64_t)s->opts.hbm_budget) {
        spike_factor_destroy(f);
        return sfail(SPIKE_ERR_NOMEM, "matrix footprint of " + std::to_string(need) +
                                          " bytes exceeds the whole HBM budget (" +
                                          std::to_string((int64_t)s->opts.hbm_budget) + " bytes)");
    }
    // Staging and the warm-up solves (which capture the factor's solve graph
    // for every padded column count) run before the matrix is visible to the
    // dispatcher, on a stream of their own, without the server lock.
    double *stg = nullptr;
    if (hipMalloc((void **)&stg, (size_t)stage_bytes) != hipSuccess) {
        (void)hipGetLastError();
        spike_factor_destroy(f);
        return sfail(SPIKE_ERR_NOMEM, "staging allocation failed");
    }
    hipStream_t ws = nullptr;
    if (hipMemset(stg, 0, (size_t)stage_bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(stg);
        spike_factor_destroy(f);
        return sfail(SPIKE_ERR_DEVICE, "staging set-up failed");
    }
    {
        double *mb = stg, *mx = stg + (size_t)n * s->opts.max_batch_rhs;
        for (int c = 1; c <= s->opts.max_batch_rhs && rc == SPIKE_OK; c <<= 1)
            for (int rep = 0; rep < 2 && rc == SPIKE_OK; rep++) rc = spike_factor_solve(f, c, mb, n, mx, n, ws);
        if (rc == SPIKE_OK && hipStreamSynchronize(ws) != hipSuccess) rc = SPIKE_ERR_DEVICE;
        if (rc == SPIKE_OK) rc = spike_factor_status(f);
        (void)hipStreamDestroy(ws);
        // the footprint includes workspaces the warm-up solves allocated
        spike_factor_stats(f, &st);
        need = st.device_bytes + stage_bytes;
        if (rc != SPIKE_OK) {
            (void)hipFree(stg);
            spike_factor_destroy(f);
            return sfail(rc, "warm-up solve failed");
        }
    }
    // admission: wait for the budget (removals notify cv_admit)
    std::unique_lock<std::mutex> lk(s->mu);
    const double tmo = s->opts.admit_timeout_ms;
    const clk::time_point deadline =
        clk::now() + std::chrono::microseconds((int64_t)(std::max(tmo, 0.0) * 1e3));
    s->admitting++;
    while (s->resident + need > (int64_t)s->opts.hbm_budget && !s->stop.load()) {
        if (tmo == 0.0 || (tmo > 0.0 && clk::now() >= deadline)) break;
        s->cv_admit.wait_for(lk, std::chrono::milliseconds(2));
    }
    s->admitting--;
    if (s->resident + need > (int64_t)s->opts.hbm_budget || s->stop.load()) {
        const int64_t left = (int64_t)s->opts.hbm_budget - s->resident;
        lk.unlock();
        (void)hipFree(stg);
        spike_factor_destroy(f);
        return sfail(SPIKE_ERR_NOMEM, "matrix footprint of " + std::to_string(need) +
                                          " bytes exceeds the HBM budget left (" + std::to_string(left) +
                                          " bytes)");
    }
    auto m = std::make_unique<MatEntry>();
    m->f = f;
    m->n = n;
    m->bytes = need;
    m->stage = stg;
    s->resident += need;
    s->mats.push_back(std::move(m));
    *id = (int)s->mats.size() - 1;
    return SPIKE_OK;
}

int spike_server_remove(spike_server_t s, int id)
{
    if (!s) return sfail(SPIKE_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int)s->mats.size() || s->mats[id]->removed) return sfail(SPIKE_ERR_ARG, "unknown matrix id");
    MatEntry &m = *s->mats[id];
    m.removed = true;   // no new requests
    while (m.open > 0) s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
    spike_factor_t f = m.f;
    double *stg = m.stage;
    m.f = nullptr;
    m.stage = nullptr;
    lk.unlock();
    spike_factor_destroy(f);
    (void)hipFree(stg);
    lk.lock();
    s->resident -= m.bytes;
    m.bytes = 0;
    s->cv_admit.notify_all();
    return SPIKE_OK;
}

int spike_server_submit(spike_server_t s, int id, int nrhs, const double *b, int ldb, double *x, int ldx,
                        uint64_t *ticket)
{
    if (!s || !ticket || !b || !x) return sfail(SPIKE_ERR_ARG, "NULL argument");
    // the copy kernels read b and write x directly: host memory would fault
    if (!device_ptr(b) || !device_ptr(x)) return sfail(SPIKE_ERR_ARG, "b and x must be device pointers");
    Request r;
    r.t_submit = clk::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int)s->mats.size() || s->mats[id]->removed)
        return sfail(SPIKE_ERR_ARG, "unknown matrix id");
    const int n = s->mats[id]->n;
    if (nrhs < 1 || nrhs > s->opts.max_batch_rhs || ldb < n || ldx < n)
        return sfail(SPIKE_ERR_ARG, "nrhs must be in [1, max_batch_rhs] and ldb, ldx >= n");
    if (s->error) return sfail(s->error, "server failed: " + s->error_msg);
    r.ticket = s->next_ticket++;
    r.id = id;
    r.nrhs = nrhs;
    r.b = b;
    r.ldb = ldb;
    r.x = x;
    r.ldx = ldx;
    if (!s->have_first) {
        s->t_first = r.t_submit;
        s->have_first = true;
    }
    s->pending.push_back(r);
    s->mats[id]->open++;
    *ticket = r.ticket;
    s->cv_submit.notify_one();
    return SPIKE_OK;
}

int spike_server_wait(spike_server_t s, uint64_t ticket)
{
    if (!s) return sfail(SPIKE_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket == 0 || ticket >= s->next_ticket) return sfail(SPIKE_ERR_ARG, "unknown ticket");
    for (;;) {
        auto it = s->finished.find(ticket);
        if (it != s->finished.end()) {
            const int st = it->second;
            s->finished.erase(it);
            return st == SPIKE_OK ? SPIKE_OK : sfail(st, s->error_msg);
        }
        s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
    }
}

int spike_server_drain(spike_server_t s)
{
    if (!s) return sfail(SPIKE_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->pending.empty() || !s->inflight.empty() || s->inflight_reserved > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
    s->finished.clear();
    return s->error ? sfail(s->error, s->error_msg) : SPIKE_OK;
}

int spike_server_stats(spike_server_t s, spike_server_stats_t *st)
{
    if (!s || !st) return sfail(SPIKE_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof *st);
    for (const auto &m : s->mats) st->matrices += m->removed ? 0 : 1;
    st->resident_bytes = s->resident;
    st->completed = (int64_t)s->lat.size();
    st->solves = s->solves;
    st->windows = s->windows;
    st->merged_rhs = s->merged_rhs;
    st->max_inflight = s->max_inflight;
    st->rows_rhs = s->rows_rhs_done;
    if (!s->lat.empty()) {
        std::vector<double> v(s->lat);
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p * (double)v.size()))]; };
        st->p50_us = pct(0.50) * 1e6;
        st->p99_us = pct(0.99) * 1e6;
        st->max_us = v.back() * 1e6;
        const double span = s->have_first ? secs(s->t_first, s->t_last) : 0.0;
        st->seconds = span;
        st->rows_rhs_per_s = span > 0 ? (double)s->rows_rhs_done / span : 0.0;
    }
    return SPIKE_OK;
}

int spike_server_reset_stats(spike_server_t s)
{
    if (!s) return sfail(SPIKE_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat.clear();
    s->rows_rhs_done = 0;
    s->solves = s->windows = s->merged_rhs = 0;
    s->max_inflight = 0;
    s->have_first = false;
    return SPIKE_OK;
}

int spike_server_destroy(spike_server_t s)
{
    if (!s) return SPIKE_OK;
    s->stop.store(true);
    s->cv_submit.notify_all();
    s->cv_admit.notify_all();
    if (s->worker.joinable()) s->worker.join();
    {
        // adds still waiting for memory see stop and fail
        std::unique_lock<std::mutex> lk(s->mu);
        while (s->admitting > 0) s->cv_admit.wait_for(lk, std::chrono::milliseconds(1));
    }
    for (hipStream_t t : s->streams) (void)hipStreamDestroy(t);
    for (hipEvent_t e : s->spare_events) (void)hipEventDestroy(e);
    for (auto &m : s->mats) {
        if (m->f) spike_factor_destroy(m->f);
        if (m->stage) (void)hipFree(m->stage);
    }
    delete s;
    return SPIKE_OK;
}

const char *spike_server_last_error(void) { return s_err.c_str(); }

}  // extern "C"
