"""GPU tests of the serving loop (resident factors, batched requests).

Each request's result in a mixed burst (requests of several matrices
arriving together, merged per matrix into multi-RHS solves and run
concurrently on several streams) must equal the same request served alone
to within 1e-13 relative error (merging changes the RHS slot count, which
may change where the coupling update is fused), and both must be within
1e-12 of sequential FP64 substitution.
"""
import numpy as np
import pytest

from oracle import pyoracle as orc
from spike_pstrsv_amd import gen

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="module")
def native():
    import torch

    assert torch.cuda.is_available(), "no GPU visible"
    from spike_pstrsv_amd import native as nat

    return nat


def test_serving_alone_equals_mixed(native):
    import torch

    specs = [{"kind": "banded", "n": 20000, "seed": gen.SEED_BASE + 400},
             {"kind": "lap2d", "side": 120, "n": 14400},
             {"kind": "lap3d", "side": 24, "n": 13824},
             {"kind": "banded", "n": 70000, "seed": gen.SEED_BASE + 401},
             {"kind": "lap2d", "side": 300, "n": 90000}]
    mats = [gen.serving_matrix(s) for s in specs]
    srv = native.Server(window_us=500.0, streams=4, slots=16)
    ids = [srv.add("L", m.a, m.ia, m.ja) for m in mats]
    rng = np.random.default_rng(7)
    reqs = []
    for k in range(40):
        mid = int(rng.integers(len(mats)))
        nrhs = int(rng.integers(1, 5))
        n = mats[mid].n
        b = torch.from_numpy(gen.rhs(n, nrhs, seed=gen.SEED_BASE + 500 + k).reshape(nrhs, n)).cuda()
        reqs.append((mid, nrhs, b))
    # alone: one request at a time
    alone = []
    for mid, nrhs, b in reqs:
        x = torch.zeros_like(b)
        srv.wait(srv.submit(ids[mid], nrhs, b, x))
        torch.cuda.synchronize()
        alone.append(x.cpu().numpy())
    srv.reset_stats()
    # mixed: everything submitted back to back
    outs = [torch.zeros_like(b) for _, _, b in reqs]
    tickets = [srv.submit(ids[mid], nrhs, b, x) for (mid, nrhs, b), x in zip(reqs, outs)]
    srv.drain()
    torch.cuda.synchronize()
    st = srv.stats()
    print("serving mix", st)
    assert st["completed"] == len(reqs) and st["merged_rhs"] > 0
    bitwise = 0
    for (mid, nrhs, b), x, xa in zip(reqs, outs, alone):
        xm = x.cpu().numpy()
        bitwise += int(np.array_equal(xm, xa))
        assert relerr(xm, xa) <= 1e-13
    print("bitwise identical alone vs mixed:", bitwise, "of", len(reqs))
    # spot-check against the oracle
    for k in range(0, len(reqs), 7):
        mid, nrhs, b = reqs[k]
        m = mats[mid]
        B = b.cpu().numpy().reshape(nrhs, m.n)
        X = alone[k].reshape(nrhs, m.n)
        for j in range(nrhs):
            assert relerr(X[j], orc.ldsol(m.a, m.ia, m.ja, B[j])) <= TOL
    assert len(tickets) == len(set(tickets))
    srv.close()


def test_serving_admission_by_footprint(native):
    small = gen.banded(20000, 64, 16, 0.02, gen.SEED_BASE + 410)
    big = gen.banded(400000, 64, 16, 0.02, gen.SEED_BASE + 411)
    probe = native.Factor("L", big.a, big.ia, big.ja)
    big_bytes = probe.stats()["device_bytes"]
    probe.close()
    srv = native.Server(hbm_budget=int(big_bytes * 0.8))
    srv.add("L", small.a, small.ia, small.ja)
    with pytest.raises(native.SpikeError):
        srv.add("L", big.a, big.ia, big.ja)
    assert srv.stats()["matrices"] == 1
    srv.close()


def _served_bytes(native, m, max_batch_rhs=16):
    """Footprint the server admits for matrix m: the factor after its
    warm-up solves (every padded column count) plus the staging buffers."""
    import torch

    f = native.Factor("L", m.a, m.ia, m.ja, use_graph=2)
    b = torch.zeros(max_batch_rhs * m.n, dtype=torch.float64, device="cuda:0")
    x = torch.zeros_like(b)
    c = 1
    while c <= max_batch_rhs:
        for _ in range(2):
            f.solve(b.data_ptr(), x.data_ptr(), nrhs=c)
        c *= 2
    torch.cuda.synchronize()
    nb = f.stats()["device_bytes"]
    f.close()
    return nb + m.n * max_batch_rhs * 2 * 8


def test_serving_admission_waits_for_memory(native):
    """An add whose footprint does not fit what is left of the budget waits;
    removing a resident matrix (after its requests complete) frees its
    memory and admits the waiting matrix, which then serves correctly."""
    import threading
    import time

    import torch

    mats = [gen.banded(30000, 64, 16, 0.02, gen.SEED_BASE + 420 + i) for i in range(3)]
    need = [_served_bytes(native, m) for m in mats]
    budget = need[0] + need[1] + need[2] // 2      # two fit, the third does not
    srv = native.Server(hbm_budget=budget, window_us=100.0)
    ida = srv.add("L", mats[0].a, mats[0].ia, mats[0].ja)
    srv.add("L", mats[1].a, mats[1].ia, mats[1].ja)
    out = {}

    def add_third():
        try:
            out["id"] = srv.add("L", mats[2].a, mats[2].ia, mats[2].ja)
        except Exception as e:      # reported below
            out["err"] = e

    th = threading.Thread(target=add_third)
    th.start()
    time.sleep(0.5)
    assert th.is_alive() and "id" not in out, "the third matrix was admitted over budget"
    assert srv.stats()["matrices"] == 2
    # a request in flight on the matrix being removed completes first
    b = torch.from_numpy(gen.rhs(mats[0].n, 2, seed=gen.SEED_BASE + 425).reshape(2, -1)).cuda()
    x = torch.zeros_like(b)
    t = srv.submit(ida, 2, b, x)
    srv.remove(ida)
    srv.wait(t)
    th.join(timeout=60)
    assert not th.is_alive() and "err" not in out, out
    torch.cuda.synchronize()
    for j in range(2):
        bj = b[j].cpu().numpy()
        assert relerr(x[j].cpu().numpy(), orc.ldsol(mats[0].a, mats[0].ia, mats[0].ja, bj)) <= TOL
    with pytest.raises(native.SpikeError):
        srv.submit(ida, 1, b, x)
    st = srv.stats()
    assert st["matrices"] == 2 and st["resident_bytes"] <= budget
    idc = out["id"]
    bc = torch.from_numpy(gen.rhs(mats[2].n, 1, seed=gen.SEED_BASE + 426)).cuda()
    xc = torch.zeros_like(bc)
    srv.wait(srv.submit(idc, 1, bc, xc))
    torch.cuda.synchronize()
    ref = orc.ldsol(mats[2].a, mats[2].ia, mats[2].ja, bc.cpu().numpy())
    assert relerr(xc.cpu().numpy(), ref) <= TOL
    srv.close()


def test_serving_add_while_serving(native):
    """Matrices added while another thread submits requests: every request
    equals its result served alone."""
    import threading

    import torch

    base = gen.banded(40000, 64, 16, 0.02, gen.SEED_BASE + 430)
    extra = [gen.lap2d(150), gen.banded(25000, 64, 16, 0.02, gen.SEED_BASE + 431), gen.lap3d(20)]
    srv = native.Server(window_us=100.0, streams=4, slots=16)
    mid = srv.add("L", base.a, base.ia, base.ja)
    B = [torch.from_numpy(gen.rhs(base.n, 1, seed=gen.SEED_BASE + 440 + k)).cuda() for k in range(60)]
    X = [torch.zeros_like(b) for b in B]
    done = {}

    def client():
        try:
            for b, x in zip(B, X):
                srv.wait(srv.submit(mid, 1, b, x))
            done["ok"] = True
        except Exception as e:
            done["err"] = e

    th = threading.Thread(target=client)
    th.start()
    ids = [srv.add("L", m.a, m.ia, m.ja) for m in extra]
    th.join(timeout=120)
    assert done.get("ok"), done
    torch.cuda.synchronize()
    for k in range(0, len(B), 6):
        ref = orc.ldsol(base.a, base.ia, base.ja, B[k].cpu().numpy())
        assert relerr(X[k].cpu().numpy(), ref) <= TOL
    for m, i in zip(extra, ids):
        b = torch.from_numpy(gen.rhs(m.n, 1, seed=gen.SEED_BASE + 450)).cuda()
        x = torch.zeros_like(b)
        srv.wait(srv.submit(i, 1, b, x))
        torch.cuda.synchronize()
        assert relerr(x.cpu().numpy(), orc.ldsol(m.a, m.ia, m.ja, b.cpu().numpy())) <= TOL
    srv.close()


def test_serving_submit_refuses_host_pointers(native):
    """The copy kernels address b and x directly, so host memory is refused
    at submit (SPIKE_ERR_ARG) and nothing is launched."""
    m = gen.banded(10000, 64, 16, 0.02, gen.SEED_BASE + 460)
    srv = native.Server()
    mid = srv.add("L", m.a, m.ia, m.ja)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 461)
    x = np.zeros_like(b)
    with pytest.raises(native.SpikeError) as ei:
        srv.submit(mid, 1, b, x)
    assert ei.value.code == 1
    srv.drain()
    assert srv.stats()["completed"] == 0
    assert not x.any()
    srv.close()
