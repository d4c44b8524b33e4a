"""GPU parity at the full sizes of the workloads that are not the metric line.

- 3D 7-point Laplacian lower triangle on 128^3 (n = 2,097,152), 32 right-hand
  sides, FP64 and FP32 values with FP64 refinement;
- 4096 uneven banded matrices (n log-uniform in [1000, 65536], bandwidth
  8-128, 4-24 off-diagonals per row, 2% long-range rows, 1-8 right-hand sides
  each) in one autotuned batch, FP64 and FP32 + refinement, solved through
  the batch's captured graph;
- serving: 64 resident matrices (n log-uniform in [1e4, 2e6], half 2D / 3D
  Laplacians, half banded), a 0.25 s Poisson stream at 20k requests/s of 1-4
  right-hand sides.

Bar: every column (batch: of every matrix; serving: of every sampled request,
at least one per matrix) within max|x - x_ref| / max|x_ref| <= 1e-12 of
sequential FP64 substitution (the oracle's ldsol / udsol), in both precision
modes; serving results equal to the same request served alone within 1e-13.
The oracle runs on host threads (its C calls release the interpreter lock).
"""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import pyoracle as orc
from spike_pstrsv_amd import gen

pytestmark = pytest.mark.gpu
TOL = 1e-12
THREADS = max(1, min(16, os.cpu_count() or 1))


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="module")
def native():
    import torch

    assert torch.cuda.is_available(), "no GPU visible"
    from spike_pstrsv_amd import native as nat

    return nat


def oracle_errors(jobs):
    """jobs: (uplo, csr, b, x) -> relative errors of x against the oracle."""
    def one(j):
        uplo, m, b, x = j
        return relerr(x, orc.trsv(uplo, m.a, m.ia, m.ja, b))

    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(one, jobs))


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_lap3d_128_32rhs(native, precision):
    import torch

    m = gen.lap3d(128)
    n, nr = m.n, 32
    B = gen.rhs(n, nr, seed=gen.SEED_BASE + 3).reshape(nr, n)
    f = native.Factor("L", m.a, m.ia, m.ja, precision=precision)
    st = f.stats()
    dB = torch.from_numpy(B).cuda()
    dX = torch.zeros_like(dB)
    for _ in range(2):   # the second solve replays the captured graph
        f.solve(dB.data_ptr(), dX.data_ptr(), nrhs=nr)
    f.status()
    torch.cuda.synchronize()
    X = dX.cpu().numpy()
    errs = oracle_errors([("L", m, B[j], X[j]) for j in range(nr)])
    print(f"lap3d_128 32 RHS {precision}: method {st['method']}, refine steps {st['refine_steps']}, "
          f"max rel. err vs oracle {max(errs):.3e}")
    assert max(errs) <= TOL


@pytest.fixture(scope="module")
def batch4096():
    specs, mats = gen.batched(4096)
    return specs, mats


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_batched_4096(native, batch4096, precision):
    import torch

    specs, mats = batch4096
    t0 = time.perf_counter()
    bt = native.Batch("L", [(m.a, m.ia, m.ja, s["nrhs"]) for s, m in zip(specs, mats)], precision=precision)
    create_s = time.perf_counter() - t0
    off = bt.offsets
    b = np.empty(bt.length)
    for i, s in enumerate(specs):
        b[off[i]:off[i + 1]] = gen.rhs(s["n"], s["nrhs"], seed=gen.SEED_BASE + 300 + i).ravel()
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    bt.solve(db.data_ptr(), dx.data_ptr())           # captures the graph
    torch.cuda.synchronize()
    x1 = dx.cpu().numpy()
    dx.zero_()
    torch.cuda.synchronize()
    bt.solve(db.data_ptr(), dx.data_ptr())           # replays it
    torch.cuda.synchronize()
    x = dx.cpu().numpy()
    assert np.array_equal(x, x1), "graph replay differs from the captured solve"
    jobs = []
    for i, (s, m) in enumerate(zip(specs, mats)):
        n, r = s["n"], s["nrhs"]
        Bi = b[off[i]:off[i + 1]].reshape(r, n)
        Xi = x[off[i]:off[i + 1]].reshape(r, n)
        jobs += [("L", m, Bi[j], Xi[j]) for j in range(r)]
    errs = oracle_errors(jobs)
    st = bt.stats()
    print(f"batched 4096 {precision}: {len(jobs)} columns, created in {create_s:.1f} s, groups {st['groups']}, "
          f"rows/partition {st['rows_per_partition']}, refine steps {st['refine_steps']}, "
          f"max rel. err vs oracle {max(errs):.3e}")
    bt.close()
    assert max(errs) <= TOL


def test_serving_64_poisson_20k(native):
    import torch

    specs = gen.serving_specs()
    mats = [gen.serving_matrix(s) for s in specs]
    srv = native.Server(window_us=200.0)
    t0 = time.perf_counter()
    ids = [srv.add("L", m.a, m.ia, m.ja) for m in mats]
    setup_s = time.perf_counter() - t0
    B = [torch.from_numpy(gen.rhs(m.n, 4, seed=gen.SEED_BASE + 600 + i).reshape(4, m.n)).cuda()
         for i, m in enumerate(mats)]
    scratch = [torch.zeros(4, m.n, dtype=torch.float64, device="cuda") for m in mats]
    rng = np.random.default_rng(23)
    rate, dur = 20000.0, 0.25
    nreq = int(rate * dur)
    arr = np.cumsum(rng.exponential(1.0 / rate, nreq))
    mids = rng.integers(len(mats), size=nreq)
    nrs = rng.integers(1, 5, size=nreq)
    # sampled requests (every 25th and the first of every matrix) write to
    # outputs of their own, the others to a shared buffer per matrix
    first = {}
    for k in range(nreq):
        first.setdefault(int(mids[k]), k)
    sampled = sorted(set(range(0, nreq, 25)) | set(first.values()))
    outs = {k: torch.zeros(int(nrs[k]), mats[int(mids[k])].n, dtype=torch.float64, device="cuda")
            for k in sampled}
    srv.reset_stats()
    t0 = time.perf_counter()
    for k in range(nreq):
        while time.perf_counter() - t0 < arr[k]:
            pass
        i, r = int(mids[k]), int(nrs[k])
        srv.submit(ids[i], r, B[i], outs[k] if k in outs else scratch[i])
    srv.drain()
    torch.cuda.synchronize()
    st = srv.stats()
    mixed = {k: outs[k].cpu().numpy() for k in sampled}
    # the same requests served alone
    worst_alone = 0.0
    for k in sampled:
        i, r = int(mids[k]), int(nrs[k])
        xa = torch.zeros_like(outs[k])
        srv.wait(srv.submit(ids[i], r, B[i], xa))
        torch.cuda.synchronize()
        worst_alone = max(worst_alone, relerr(mixed[k], xa.cpu().numpy()))
    jobs = []
    for k in sampled:
        i, r = int(mids[k]), int(nrs[k])
        Bh = B[i][:r].cpu().numpy()
        jobs += [("L", mats[i], Bh[j], mixed[k][j]) for j in range(r)]
    errs = oracle_errors(jobs)
    print(f"serving 64: setup {setup_s:.1f} s, {st['completed']} requests in {st['seconds']:.3f} s "
          f"({st['completed'] / max(st['seconds'], 1e-9):.0f} req/s), p50 {st['p50_us']:.0f} us, "
          f"p99 {st['p99_us']:.0f} us; {len(sampled)} sampled requests covering {len(first)} matrices: "
          f"max rel. err vs oracle {max(errs):.3e}, alone vs mixed {worst_alone:.3e}")
    srv.close()
    assert len(first) == len(mats)
    assert st["completed"] == nreq
    assert max(errs) <= TOL
    assert worst_alone <= 1e-13
