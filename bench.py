#!/usr/bin/env python3
"""Benchmark of the Spike sparse triangular solve on one MI355X.

Workload (default): random banded lower triangle, n = 1,048,576, 16
off-diagonals per row drawn from the previous 64 columns plus a long-range
entry in 2% of the rows, FP64 CSR with int32 indices, one right-hand side,
1024 Spike partitions, plain FP64 path.  A "step" is one complete solve
(pass-1 block solve, reduced solve, coupling update, pass-2 block solve) of one
right-hand side; b and x are resident in HBM when the timed region starts.

Prints one JSON line.  `value` = rows*RHS solved per second over the timed
steps.  `roofline` is for the dominant kernel (the block solve, two launches
per step): algorithmic bytes per launch = 12 B per in-block nonzero (FP64
value + int32 column) + 4 B per row pointer + 8 B read of b + 8 B write per
row, divided by the kernel's average device time measured with HIP events on
the launch stream.  `cpu_baseline` times the CPU restatement of the
reference's Spike algorithm (OpenMP, same partitions) on a bounded sample.
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


def inblock_nnz(m, P: int) -> int:
    import numpy as np

    n = m.n
    k = n // P
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(m.ia))
    part = np.minimum(rows // k, P - 1)
    return int(np.sum(m.ja >= part * k))


def cpu_baseline(m, P: int, b, x_gpu, seconds: float):
    """Times the oracle's Spike restatement (OpenMP) on the host."""
    import numpy as np

    from oracle import pyoracle as orc

    threads = max(1, min(16, os.cpu_count() or 1))
    sp = orc.Spike("L", m.a, m.ia, m.ja, P, tau=0.0, nthreads=threads)
    x = sp.solve(b)
    x = sp.solve(b)
    count, t0 = 0, time.perf_counter()
    while True:
        sp.solve(b, x)
        count += 1
        el = time.perf_counter() - t0
        if el >= seconds or count >= 5000:
            break
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    t1 = time.perf_counter()
    orc.ldsol(m.a, m.ia, m.ja, b)
    t_seq = time.perf_counter() - t1
    rel = float(np.max(np.abs(x_gpu - ref)) / np.max(np.abs(ref)))
    cpu = {
        "value": m.n * count / el,
        "unit": UNIT,
        "cores": threads,
        "kind": "port",
        "sample": f"{count} solves in {el:.2f} s of the OpenMP Spike restatement (P={P}, same matrix and RHS) "
                  f"after 2 warm-up solves; sequential FP64 substitution alone: {m.n / t_seq:.3e} rows/s on 1 core",
    }
    return cpu, rel


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="banded_1m")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-file", default=None)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy three times (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale factors)")
    args = ap.parse_args()
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
        print(f"bench.py: this build runs on one GPU only (the Spike path would shard by partition ranges "
              f"across GPUs, see DESIGN.md); --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    from spike_pstrsv_amd import gen, native

    m, P, desc = workload(args.workload)
    n = m.n
    b = gen.rhs(n, 1, seed=gen.SEED_BASE + 2)
    fac = native.Factor("L", m.a, m.ia, m.ja, partitions=P)
    st = fac.stats()
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
    traffic = None
    tf = Path(args.traffic_file) if args.traffic_file else ROOT / "profiles" / f"traffic_block_solve_{args.workload}.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None

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
                     "bytes_per_launch": launch_bytes, "avg_launch_ms": blk_ms},
        "phases_ms": {"pass1": ph[0], "reduced": ph[1], "coupling": ph[2], "pass2": ph[3]},
        "solve_roofline_frac": b_alg / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "latency_ms": {"p50": statistics.median(lat), "p99": lat[min(len(lat) - 1, int(0.99 * len(lat)))],
                       "mall_flushed_p50": statistics.median(fl)},
        "wall_s_timed": wall,
    }
    if not args.no_cpu_baseline:
        cpu, rel = cpu_baseline(m, P, b, x_gpu, args.cpu_seconds)
        out["cpu_baseline"] = cpu
        out["rel_err_vs_sequential_fp64"] = rel
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
