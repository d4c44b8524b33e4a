#!/usr/bin/env python3
"""Benchmark of the Spike sparse triangular solve on one MI355X.

Workload (default): random banded lower triangle, n = 1,048,576, 16
off-diagonals per row drawn from the previous 64 columns plus a long-range
entry in 2% of the rows, FP64 CSR with int32 indices, one right-hand side,
1024 Spike partitions, plain FP64 path.  A "step" is one complete solve
(pass-1 block solve, reduced solve, coupling update, pass-2 block solve) of one
right-hand side; b and x are resident in HBM when the timed region starts.

Prints one JSON line.  `value` = rows*RHS solved per second over the timed
steps.  Other workloads (--workload): lap2d_256 (P=8), lap3d_128 (32 RHS,
FP64 and FP32 + FP64 refinement), batched (4096 uneven banded matrices with
1-8 RHS each, one graph launch per step), serving (64 resident matrices,
Poisson arrivals of 1-4 RHS requests, 200 us batching windows).  `roofline` is for the dominant kernel (the block solve, two launches
per step): algorithmic bytes per launch = 12 B per in-block nonzero (FP64
value + int32 column) + 4 B per row pointer + 8 B read of b + 8 B write per
row, divided by the kernel's average device time measured with HIP events on
the launch stream.  `cpu_baseline` times the CPU restatement of the
reference's Spike algorithm (OpenMP, same partitions) on a bounded sample,
and beside it (`competitors`) the reference harness's MKL triangular solves
on the same system when MKL's runtime library loads.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "SpTRSV rows·RHS/s"
UNIT = "rows·RHS/s"


def workload(name: str):
    from spike_pstrsv_amd import gen

    if name == "banded_1m":
        m = gen.banded(1 << 20, 64, 16, 0.02, gen.SEED_BASE + 2)
        return m, 1024, "random banded lower triangle n=1048576, bw 64, 16 nnz/row, 2% long-range, P=1024"
    if name == "lap2d_256":
        m = gen.lap2d(256)
        return m, 8, "2D 5-point Laplacian lower triangle 256x256 (n=65536), P=8"
    raise SystemExit(f"unknown workload {name}")


def workload_rhs(n: int):
    """The single right-hand side of the banded_1m / lap2d_256 lines."""
    from spike_pstrsv_amd import gen

    return gen.rhs(n, 1, seed=gen.SEED_BASE + 2)


def inblock_nnz(m, P: int) -> int:
    import numpy as np

    n = m.n
    k = n // P
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(m.ia))
    part = np.minimum(rows // k, P - 1)
    return int(np.sum(m.ja >= part * k))


def _allowed_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


# Read before any OpenMP runtime starts: with OMP_PROC_BIND set, the runtime
# binds the initial thread to one core, and the affinity mask read afterwards
# holds that core only.
ALLOWED_CPUS = _allowed_cpus()


def host_cpu():
    """(threads to use, nproc, CPU model): min(nproc, 128) threads, capped by
    the CPUs this process may run on (read at start-up) and by
    OMP_NUM_THREADS when it is set."""
    nproc = os.cpu_count() or 1
    allowed = ALLOWED_CPUS
    threads = min(nproc, 128, allowed)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(1, threads), nproc, model


def cpu_baseline(workload_name: str, m, b, x_gpu, seconds: float):
    """The CPU baseline of `workload_name`, measured in a child process (the
    only one whose OpenMP threads are pinned: a binding set in this process
    would pin its main thread to one core once torch loads the OpenMP
    runtime, and every host thread it starts, the library's preprocessing
    threads among them, would share that core).  Returns the child's record
    and the GPU solution's error against the oracle's sequential
    substitution."""
    import subprocess

    import numpy as np

    from oracle import pyoracle as orc

    env = dict(os.environ)
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    cmd = [sys.executable, str(Path(__file__).resolve()), "--workload", workload_name, "--cpu-seconds", str(seconds),
           "--cpu-baseline-child"]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline child failed ({r.returncode}): {r.stderr[-2000:]}")
    cpu = json.loads(r.stdout.strip().splitlines()[-1])
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    rel = float(np.max(np.abs(x_gpu - ref)) / np.max(np.abs(ref)))
    return cpu, rel


def cpu_baseline_measure(m, P: int, b, seconds: float):
    """Times the oracle's Spike restatement (OpenMP, threads pinned close to
    cores) on the host; runs in the child process of cpu_baseline."""
    from oracle import pyoracle as orc

    threads, nproc, model = host_cpu()
    # at most one thread per partition (the reference runs one partition per
    # OpenMP thread, src/spike_pstrsv.c:168-170): 8 threads for the 2D Laplacian's P = 8
    threads = min(threads, P)
    t0 = time.perf_counter()
    sp = orc.Spike("L", m.a, m.ia, m.ja, P, tau=0.0, nthreads=threads)
    preproc_s = time.perf_counter() - t0
    x = sp.solve(b)
    x = sp.solve(b)
    count, per, t0 = 0, [], time.perf_counter()
    while True:
        ts = time.perf_counter()
        sp.solve(b, x)
        per.append(time.perf_counter() - ts)
        count += 1
        el = time.perf_counter() - t0
        if el >= seconds or count >= 5000:
            break
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    t1 = time.perf_counter()
    orc.ldsol(m.a, m.ia, m.ja, b)
    t_seq = time.perf_counter() - t1
    cpu = {
        "value": m.n * count / el,
        "unit": UNIT,
        "cores": threads,
        "kind": "port",
        "sample": f"{count} solves in {el:.2f} s of the OpenMP Spike restatement (P={P}, same matrix and RHS) "
                  f"after 2 warm-up solves; sequential FP64 substitution alone: {m.n / t_seq:.3e} rows/s on 1 core",
        "nproc": nproc,
        "cpu_model": model,
        "omp": {"OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "OMP_PLACES": os.environ.get("OMP_PLACES")},
        "preproc_s": preproc_s,
        "solve_ms": solve_stats(per),
        "competitors": mkl_competitors(m, b, ref, threads, min(seconds / 2.0, 5.0)),
    }
    return cpu


def solve_stats(per):
    """min / mean / p50 / p99 of per-solve seconds, in milliseconds (the
    statistics test/runtime_analysis.c reports for its timed solves)."""
    v = sorted(per)
    return {"min": v[0] * 1e3, "mean": statistics.fmean(v) * 1e3, "p50": statistics.median(v) * 1e3,
            "p99": v[min(len(v) - 1, int(0.99 * len(v)))] * 1e3, "solves": len(v)}


MKL_CANDIDATES = ("/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so", "libmkl_rt.so.1")


def mkl_competitors(m, b, ref, threads: int, seconds: float):
    """Times the reference harness's CPU competitors (test/runtime_analysis.c:
    67-87): MKL's inspector-executor `mkl_sparse_d_trsv` (analysis with
    `mkl_sparse_set_sv_hint` + `mkl_sparse_optimize` outside the timed region,
    as test/runtime_analysis.c:358-362) and the zero-based
    `mkl_cspblas_dcsrtrsv`, on the same matrix and right-hand side, when
    MKL's runtime library loads.  Each gets a bounded sample of `seconds`."""
    import ctypes

    import numpy as np

    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")   # share the restatement's libgomp
    lib, path = None, None
    for cand in MKL_CANDIDATES:
        try:
            lib, path = ctypes.CDLL(cand), cand
            break
        except OSError:
            continue
    if lib is None:
        return {"available": False, "reason": "libmkl_rt not found"}
    try:
        lib.MKL_Set_Num_Threads(ctypes.c_int(threads))
    except AttributeError:
        pass
    n = m.n
    a = np.ascontiguousarray(m.a, dtype=np.float64)
    ia = np.ascontiguousarray(m.ia, dtype=np.int32)
    ja = np.ascontiguousarray(m.ja, dtype=np.int32)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    ptr = lambda v: v.ctypes.data_as(ctypes.c_void_p)   # noqa: E731

    def sample(call):
        x = np.zeros(n)
        call(x)                                           # warm-up
        count, per, t0 = 0, [], time.perf_counter()
        while True:
            ts = time.perf_counter()
            call(x)
            per.append(time.perf_counter() - ts)
            count += 1
            el = time.perf_counter() - t0
            if el >= seconds or count >= 5000:
                break
        return {"value": n * count / el, "unit": UNIT, "solves": count, "seconds": round(el, 3),
                "solve_ms": solve_stats(per), "rel_err_vs_oracle": reldiff(x, ref)}

    out = {"available": True, "library": path, "threads": int(lib.MKL_Get_Max_Threads()),
           "threading_layer": os.environ.get("MKL_THREADING_LAYER")}

    class Descr(ctypes.Structure):
        _fields_ = [("type", ctypes.c_int), ("mode", ctypes.c_int), ("diag", ctypes.c_int)]

    # sparse_matrix_type TRIANGULAR = 23, fill LOWER = 40, diag NON_UNIT = 50,
    # operation NON_TRANSPOSE = 10, index base ZERO = 0 (mkl_spblas.h)
    descr = Descr(23, 40, 50)
    A = ctypes.c_void_p()
    rs, re_ = ia[:-1].copy(), ia[1:].copy()
    st = lib.mkl_sparse_d_create_csr(ctypes.byref(A), 0, n, n, ptr(rs), ptr(re_), ptr(ja), ptr(a))
    if st == 0:
        lib.mkl_sparse_set_sv_hint.argtypes = [ctypes.c_void_p, ctypes.c_int, Descr, ctypes.c_int]
        lib.mkl_sparse_d_trsv.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p, Descr, ctypes.c_void_p,
                                          ctypes.c_void_p]
        st = lib.mkl_sparse_set_sv_hint(A, 10, descr, 1000) or lib.mkl_sparse_optimize(A)
    if st == 0:
        out["mkl_sparse_d_trsv"] = sample(lambda x: lib.mkl_sparse_d_trsv(10, 1.0, A, descr, ptr(bb), ptr(x)))
    else:
        out["mkl_sparse_d_trsv"] = {"error": f"sparse_status_t {st}"}
    if A:
        lib.mkl_sparse_destroy(A)
    nn = ctypes.c_int(n)
    lib.mkl_cspblas_dcsrtrsv.restype = None
    out["mkl_cspblas_dcsrtrsv"] = sample(
        lambda x: lib.mkl_cspblas_dcsrtrsv(b"L", b"N", b"N", ctypes.byref(nn), ptr(a), ptr(ia), ptr(ja), ptr(bb),
                                           ptr(x)))
    return out


def progress(msg: str):
    """One line on stderr, flushed (long workloads report their phases)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def oracle_err(m, B, X, uplo="L"):
    """Largest relative error of the columns X[j] against the oracle's
    sequential FP64 substitution of B[j] (host threads)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import pyoracle as orc

    B = B.reshape(-1, m.n)
    X = X.reshape(-1, m.n)
    with ThreadPoolExecutor(host_cpu()[0]) as ex:
        errs = ex.map(lambda j: reldiff(X[j], orc.trsv(uplo, m.a, m.ia, m.ja, B[j])), range(B.shape[0]))
        return max(errs)


def reldiff(x, ref):
    import numpy as np

    return float(np.max(np.abs(x - ref)) / max(float(np.max(np.abs(ref))), 1e-300))


def timed_solves(f, B, X, nr, stream, reps):
    """Average device milliseconds of one solve (b, x resident; graph replays after the first two)."""
    import torch

    for _ in range(3):
        f.solve(B.data_ptr(), X.data_ptr(), nrhs=nr, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        f.solve(B.data_ptr(), X.data_ptr(), nrhs=nr, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def precisions(args):
    return ("fp64", "fp32_refined") if args.precision == "both" else (args.precision,)


def emit(out):
    print(json.dumps(out), flush=True)
    return 0


def base_line(args, value, ms_step, dtype, config):
    return {"metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms_step, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic (deterministic SplitMix64 generator)", "config": config}


def bench_lap3d(args):
    """3D 7-point Laplacian lower triangle on 128^3, 32 RHS: FP64 and FP32 + FP64 refinement."""
    import numpy as np
    import torch

    from spike_pstrsv_amd import gen, native

    m = gen.lap3d(128)
    n, nr = m.n, 32
    stream = torch.cuda.current_stream()
    B = torch.from_numpy(gen.rhs(n, nr, seed=gen.SEED_BASE + 3).reshape(nr, n)).cuda()
    res = {}
    x64 = None
    for prec in precisions(args):
        f = native.Factor("L", m.a, m.ia, m.ja, precision=prec)
        st = f.stats()
        X = torch.zeros_like(B)
        ms = timed_solves(f, B, X, nr, stream, max(3, args.steps // 20))
        Xh = X.cpu().numpy()
        if prec == "fp64":
            x64 = Xh
        res[prec] = {"ms_per_solve": ms, "rows_rhs_per_s": n * nr / (ms * 1e-3),
                     "rel_err_vs_oracle": oracle_err(m, B.cpu().numpy(), Xh),
                     "rel_diff_vs_fp64_gpu": reldiff(Xh, x64) if x64 is not None and prec != "fp64" else None,
                     "method": "levels" if st["method"] == 2 else "spike", "levels": st["levels"],
                     "refine_steps": st["refine_steps"], "preproc_s": st["preproc_seconds"]}
        f.close()
    best = max(res.values(), key=lambda r: r["rows_rhs_per_s"])
    b_alg = 12 * m.nnz + 4 * (n + 1) + 16 * n * nr
    out = base_line(args, best["rows_rhs_per_s"], best["ms_per_solve"], "f64",
                    {"workload": "lap3d_128", "n": n, "nnz": m.nnz, "nrhs": nr, "uplo": "L"})
    out["precisions"] = res
    out["solve_roofline_frac"] = b_alg / (best["ms_per_solve"] * 1e-3) / (HBM_PEAK_GBS * 1e9)
    return emit(out)


def bench_batched(args):
    """4096 banded matrices, n log-uniform in [1000, 65536], 1-8 RHS each, one graph launch per solve."""
    import numpy as np
    import torch

    from spike_pstrsv_amd import gen, native

    progress("batched: generating 4096 matrices")
    specs, mats = gen.batched(4096)
    ms_in = [(mm.a, mm.ia, mm.ja, s["nrhs"]) for s, mm in zip(specs, mats)]
    stream = torch.cuda.current_stream()
    res = {}
    x64 = None
    for prec in precisions(args):
        progress(f"batched: creating the {prec} batch (autotuned)")
        t0 = time.perf_counter()
        bt = native.Batch("L", ms_in, precision=prec, use_graph=not args.no_graph)
        create_s = time.perf_counter() - t0
        progress(f"batched: {prec} batch created in {create_s:.1f} s; timing solves")
        L = bt.length
        b = np.empty(L)
        for i, s in enumerate(specs):
            b[bt.offsets[i]:bt.offsets[i + 1]] = gen.rhs(s["n"], s["nrhs"], seed=gen.SEED_BASE + 300 + i).ravel()
        db = torch.from_numpy(b).cuda()
        dx = torch.zeros_like(db)
        for _ in range(2):
            bt.solve(db.data_ptr(), dx.data_ptr(), stream=stream.cuda_stream)
        torch.cuda.synchronize()
        reps = max(3, args.steps // 20)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            bt.solve(db.data_ptr(), dx.data_ptr(), stream=stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        x = dx.cpu().numpy()
        worst = None
        if prec == "fp64":
            x64 = x
        elif x64 is not None:
            worst = max(reldiff(x[bt.offsets[i]:bt.offsets[i + 1]], x64[bt.offsets[i]:bt.offsets[i + 1]])
                        for i in range(len(specs)))
        progress(f"batched: {prec} {ms:.2f} ms per solve; checking every column against the oracle")
        err = batch_oracle_err(specs, mats, bt.offsets, b, x)
        st = bt.stats()
        res[prec] = {"ms_per_solve": ms, "rows_rhs_per_s": L / (ms * 1e-3), "rel_err_vs_oracle": err,
                     "rel_diff_vs_fp64_gpu": worst,
                     "create_s": create_s, "rows_per_partition": st["rows_per_partition"],
                     "refine_steps": st["refine_steps"], "groups": st["groups"], "graph_nodes": st["graph_nodes"]}
        bt.close()
        del db, dx
        torch.cuda.empty_cache()
    best = max(res.values(), key=lambda r: r["rows_rhs_per_s"])
    nnz = sum(mm.nnz for mm in mats)
    rows = sum(s["n"] for s in specs)
    rows_rhs = sum(s["n"] * s["nrhs"] for s in specs)
    b_alg = 12 * nnz + 4 * (rows + len(specs)) + 16 * rows_rhs
    out = base_line(args, best["rows_rhs_per_s"], best["ms_per_solve"], "f64",
                    {"workload": "batched_4096", "matrices": len(specs), "rows": rows, "nnz": nnz,
                     "rows_rhs": rows_rhs, "graph": not args.no_graph})
    out["precisions"] = res
    out["solve_roofline_frac"] = b_alg / (best["ms_per_solve"] * 1e-3) / (HBM_PEAK_GBS * 1e9)
    # the target's ratio: batched FP32 + refinement over the plain FP64 path
    if len(res) == 2:
        out["fp32_refined_over_fp64"] = res["fp32_refined"]["rows_rhs_per_s"] / res["fp64"]["rows_rhs_per_s"]
    return emit(out)


def batch_oracle_err(specs, mats, offsets, b, x):
    """Largest relative error over every column of every matrix of a batch
    against the oracle (host threads)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import pyoracle as orc

    jobs = []
    for i, (s, m) in enumerate(zip(specs, mats)):
        n, r = s["n"], s["nrhs"]
        Bi = b[offsets[i]:offsets[i + 1]].reshape(r, n)
        Xi = x[offsets[i]:offsets[i + 1]].reshape(r, n)
        jobs += [(m, Bi[j], Xi[j]) for j in range(r)]
    with ThreadPoolExecutor(host_cpu()[0]) as ex:
        return max(ex.map(lambda t: reldiff(t[2], orc.ldsol(t[0].a, t[0].ia, t[0].ja, t[1])), jobs))


def bench_serving(args):
    """64 resident matrices; Poisson arrivals at --rate req/s for --duration s; 1-4 RHS per request."""
    import numpy as np
    import torch

    from spike_pstrsv_amd import gen, native

    specs = gen.serving_specs()
    srv = native.Server(window_us=200.0)
    ids, ns = [], []
    t0 = time.perf_counter()
    for s in specs:
        mm = gen.serving_matrix(s)
        ids.append(srv.add("L", mm.a, mm.ia, mm.ja))
        ns.append(mm.n)
    setup_s = time.perf_counter() - t0
    B = [torch.from_numpy(gen.rhs(nn, 4, seed=gen.SEED_BASE + 600 + i).reshape(4, nn)).cuda()
         for i, nn in enumerate(ns)]
    R = 8
    X = [[torch.zeros(4, nn, dtype=torch.float64, device="cuda") for _ in range(R)] for nn in ns]
    rng = np.random.default_rng(11)
    srv.reset_stats()
    for _ in range(200):                                   # requests served alone (closed loop)
        i = int(rng.integers(len(ns)))
        srv.wait(srv.submit(ids[i], int(rng.integers(1, 5)), B[i], X[i][0]))
    alone = srv.stats()
    srv.reset_stats()
    nreq = int(args.rate * args.duration)
    arr = np.cumsum(rng.exponential(1.0 / args.rate, nreq))
    mids = rng.integers(len(ns), size=nreq)
    nrs = rng.integers(1, 5, size=nreq)
    t0 = time.perf_counter()
    for k in range(nreq):
        while time.perf_counter() - t0 < arr[k]:
            pass
        i = int(mids[k])
        srv.submit(ids[i], int(nrs[k]), B[i], X[i][k % R])
    submit_s = time.perf_counter() - t0
    srv.drain()
    st = srv.stats()
    srv.close()
    out = base_line(args, st["rows_rhs_per_s"], 1e3 * st["seconds"] / max(st["solves"], 1), "f64",
                    {"workload": "serving_64", "matrices": len(ns), "rows": int(sum(ns)),
                     "offered_req_per_s": args.rate, "duration_s": args.duration, "window_us": 200.0})
    out["scaling"] = "weak"
    out["latency_us"] = {"p50": st["p50_us"], "p99": st["p99_us"], "max": st["max_us"]}
    out["alone_latency_us"] = {"p50": alone["p50_us"], "p99": alone["p99_us"]}
    out["served_req_per_s"] = st["completed"] / st["seconds"] if st["seconds"] > 0 else None
    out["serving"] = {k: st[k] for k in ("completed", "solves", "windows", "merged_rhs", "max_inflight",
                                         "resident_bytes", "seconds")}
    out["submit_seconds"] = submit_s
    out["setup_seconds"] = setup_s
    return emit(out)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="banded_1m",
                    choices=["banded_1m", "lap2d_256", "lap3d_128", "batched", "serving"])
    ap.add_argument("--rate", type=float, default=20000.0, help="serving: offered requests per second")
    ap.add_argument("--duration", type=float, default=0.5, help="serving: seconds of arrivals")
    ap.add_argument("--precision", choices=["both", "fp64", "fp32_refined"], default="both",
                    help="lap3d_128 / batched: precision modes to run (profiling runs take one)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-file", default=None)
    ap.add_argument("--no-graph", action="store_true",
                    help="batched: launch the groups eagerly on one stream instead of the captured graph")
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy three times (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale factors)")
    args = ap.parse_args()
    if args.cpu_baseline_child:
        m, P, _ = workload(args.workload)
        print(json.dumps(cpu_baseline_measure(m, P, workload_rhs(m.n), args.cpu_seconds)), flush=True)
        return 0
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8, "copies": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (the Spike path would shard by contiguous partition "
              f"ranges across GPUs, exchanging only the reduced right-hand side); --gpus {args.gpus} is not "
              f"supported", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    from spike_pstrsv_amd import gen, native

    if args.workload == "lap3d_128":
        return bench_lap3d(args)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "serving":
        return bench_serving(args)
    m, P, desc = workload(args.workload)
    n = m.n
    b = workload_rhs(n)
    t_before = native.timers()
    t0 = time.perf_counter()
    fac = native.Factor("L", m.a, m.ia, m.ja, partitions=P)
    create_s = time.perf_counter() - t0
    t_after = native.timers()
    st = fac.stats()
    # preprocessing phases (wall seconds of the library's phase timers):
    # plan (partitions, dependents, step streams), spike build (device,
    # synchronised), reduced-system assembly (device)
    prep = {"factor_create_s": create_s, "preproc_s": st["preproc_seconds"],
            "plan_s": t_after[1] - t_before[1], "spike_build_ms": 1e3 * (t_after[2] - t_before[2]),
            "reduced_assembly_ms": 1e3 * (t_after[3] - t_before[3])}
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)

    for _ in range(args.warmup):
        fac.solve(db.data_ptr(), dx.data_ptr(), stream=sh)
    torch.cuda.synchronize()

    # timed region: K solves back to back
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        fac.solve(db.data_ptr(), dx.data_ptr(), stream=sh)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_total = e0.elapsed_time(e1)
    ms_step = ms_total / args.steps
    value = n * args.steps / (ms_total * 1e-3)
    x_gpu = dx.cpu().numpy()

    # per-solve latency (each solve bracketed by events)
    lat = []
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, c in evs:
        a.record(stream)
        fac.solve(db.data_ptr(), dx.data_ptr(), stream=sh)
        c.record(stream)
    torch.cuda.synchronize()
    lat = sorted(a.elapsed_time(c) for a, c in evs)

    # per-phase device times (events on the launch stream)
    nprof = max(10, args.steps // 4)
    acc = [0.0, 0.0, 0.0, 0.0]
    for _ in range(nprof):
        ph = fac.solve_profiled(db.data_ptr(), dx.data_ptr(), stream=sh)
        acc = [u + v for u, v in zip(acc, ph)]
    ph = [v / nprof for v in acc]
    blk_ms = (ph[0] + ph[3]) / 2.0

    # latency with the Infinity Cache flushed (a 1 GiB write between solves)
    flush = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    fl = []
    for _ in range(min(args.steps, 50)):
        flush.fill_(1.0)
        a = torch.cuda.Event(enable_timing=True)
        c = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fac.solve(db.data_ptr(), dx.data_ptr(), stream=sh)
        c.record(stream)
        torch.cuda.synchronize()
        fl.append(a.elapsed_time(c))
    del flush

    nnz_in = inblock_nnz(m, P)
    launch_bytes = 12 * nnz_in + 4 * (n + 1) + 16 * n
    achieved = launch_bytes / (blk_ms * 1e-3) / 1e9
    b_alg = 12 * m.nnz + 4 * (n + 1) + 16 * n
    # HBM traffic per block-solve launch is not measured in this run: it is
    # read from the committed rocprofv3 counter summary (FETCH_SIZE and
    # WRITE_SIZE runs, scaled by a plain device copy's readings), and the
    # line names that file
    traffic, traffic_from = None, None
    tf = Path(args.traffic_file) if args.traffic_file else ROOT / "profiles" / f"traffic_block_solve_{args.workload}.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
            traffic_from = str(tf.relative_to(ROOT)) if tf.is_relative_to(ROOT) else str(tf)
        except Exception:
            traffic = None
    solve_frac = b_alg / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9)
    flushed_frac = b_alg / (statistics.median(fl) * 1e-3) / (HBM_PEAK_GBS * 1e9)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (deterministic SplitMix64 generator)",
        "config": {"workload": args.workload, "description": desc, "n": n, "nnz": m.nnz, "nrhs": 1,
                   "partitions": st["partitions"], "uplo": "L", "precision": "fp64",
                   "lanes_per_row": st["lanes_per_row"], "reduced_solver":
                   "jacobi" if st["reduced_sweeps"] > 0 else "exact", "reduced_sweeps": st["reduced_sweeps"]},
        "roofline": {"bound": "hbm", "kernel": "k_block_solve (pass 1 and pass 2)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_from": traffic_from, "bytes_per_launch": launch_bytes, "avg_launch_ms": blk_ms,
                     "solve_bytes": b_alg, "solve_frac": solve_frac, "solve_frac_mall_flushed": flushed_frac},
        "phases_ms": {"pass1": ph[0], "reduced": ph[1], "coupling": ph[2], "pass2": ph[3]},
        "solve_roofline_frac": solve_frac,
        "latency_ms": {"p50": statistics.median(lat), "p99": lat[min(len(lat) - 1, int(0.99 * len(lat)))],
                       "mall_flushed_p50": statistics.median(fl)},
        "wall_s_timed": wall,
        "preprocess": prep,
    }
    # variants of the same matrix (not the metric line): 4 RHS, FP32 +
    # refinement; every column is checked against the oracle
    var = {}
    Bv = torch.from_numpy(gen.rhs(n, 4, seed=gen.SEED_BASE + 7).reshape(4, n)).cuda()
    x64 = None
    for prec, nr in (("fp64", 4), ("fp32_refined", 1), ("fp32_refined", 4)):
        fv = fac if prec == "fp64" else native.Factor("L", m.a, m.ia, m.ja, partitions=P, precision=prec)
        B1 = Bv[:nr].contiguous()
        Xv = torch.zeros_like(B1)
        ms_v = timed_solves(fv, B1, Xv, nr, stream, max(5, args.steps // 10))
        xh = Xv.cpu().numpy()
        if x64 is None:
            x64 = xh
        var[f"{prec}_nrhs{nr}"] = {"rows_rhs_per_s": n * nr / (ms_v * 1e-3), "ms": ms_v,
                                   "rel_err_vs_oracle": oracle_err(m, B1.cpu().numpy(), xh),
                                   "rel_diff_vs_fp64_gpu": reldiff(xh, x64[:nr]),
                                   "refine_steps": fv.stats()["refine_steps"]}
        if fv is not fac:
            fv.close()
    out["variants"] = var
    if not args.no_cpu_baseline:
        cpu, rel = cpu_baseline(args.workload, m, b, x_gpu, args.cpu_seconds)
        out["cpu_baseline"] = cpu
        out["rel_err_vs_sequential_fp64"] = rel
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
