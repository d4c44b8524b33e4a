"""GPU parity of batched solves (many independent matrices in one batch).

Every matrix of a batch is checked column by column against sequential FP64
substitution (the oracle's ldsol / udsol) at relative error <= 1e-12.  The
graph-replayed solve is checked bit for bit against the eagerly launched one,
over several replays with eager work and synchronisation in between.
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


def small_batch(count, uplo="L", seed=gen.SEED_BASE + 200):
    specs = gen.batch_specs(count, n_lo=300, n_hi=6000, seed=seed)
    mats = []
    for s in specs:
        m = gen.banded(s["n"], s["bw"], s["q"], 0.02, s["seed"])
        mats.append(gen.to_upper(m) if uplo == "U" else m)
    return specs, mats


def packed_rhs(specs, offsets, seed):
    b = np.zeros(offsets[-1])
    for i, s in enumerate(specs):
        b[offsets[i]:offsets[i + 1]] = gen.rhs(s["n"], s["nrhs"], seed=seed + i).ravel()
    return b


def check_batch(specs, mats, uplo, offsets, b, x):
    worst = 0.0
    for i, (s, m) in enumerate(zip(specs, mats)):
        n, r = s["n"], s["nrhs"]
        B = b[offsets[i]:offsets[i + 1]].reshape(r, n)
        X = x[offsets[i]:offsets[i + 1]].reshape(r, n)
        for j in range(r):
            e = relerr(X[j], orc.trsv(uplo, m.a, m.ia, m.ja, B[j]))
            worst = max(worst, e)
            assert e <= TOL, (i, j, e)
    return worst


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_batch_fp64(native, uplo):
    specs, mats = small_batch(40, uplo)
    bt = native.Batch(uplo, [(m.a, m.ia, m.ja, s["nrhs"]) for s, m in zip(specs, mats)], rows_per_partition=512)
    b = packed_rhs(specs, bt.offsets, gen.SEED_BASE + 210)
    x = bt.solve(b)
    print("batch", uplo, bt.stats(), "worst", check_batch(specs, mats, uplo, bt.offsets, b, x))


def test_batch_graph_replays_match_eager(native):
    import torch

    specs, mats = small_batch(24, seed=gen.SEED_BASE + 220)
    ms = [(m.a, m.ia, m.ja, s["nrhs"]) for s, m in zip(specs, mats)]
    eager = native.Batch("L", ms, rows_per_partition=256, use_graph=False)
    graph = native.Batch("L", ms, rows_per_partition=256, use_graph=True)
    b = packed_rhs(specs, eager.offsets, gen.SEED_BASE + 221)
    x_ref = eager.solve(b)
    check_batch(specs, mats, "L", eager.offsets, b, x_ref)
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    s = torch.cuda.current_stream().cuda_stream
    for rep in range(4):
        dx.zero_()
        graph.solve(db.data_ptr(), dx.data_ptr(), stream=s)
        torch.cuda.synchronize()
        got = dx.cpu().numpy()
        assert np.array_equal(got, x_ref), f"graph replay {rep} differs from the eager solve"
        eager.solve(b)                      # eager work between replays
    assert graph.stats()["graph_nodes"] > 0


def test_batch_fp32_refined(native):
    specs, mats = small_batch(32, seed=gen.SEED_BASE + 230)
    bt = native.Batch("L", [(m.a, m.ia, m.ja, s["nrhs"]) for s, m in zip(specs, mats)], rows_per_partition=512,
                      precision="fp32_refined")
    st = bt.stats()
    print("batch fp32_refined", st)
    assert st["refine_steps"] >= 1
    b = packed_rhs(specs, bt.offsets, gen.SEED_BASE + 231)
    check_batch(specs, mats, "L", bt.offsets, b, bt.solve(b))


def test_batch_autotuned(native):
    specs, mats = small_batch(64, seed=gen.SEED_BASE + 240)
    bt = native.Batch("L", [(m.a, m.ia, m.ja, s["nrhs"]) for s, m in zip(specs, mats)], autotune_sample=16)
    st = bt.stats()
    print("batch autotuned", st)
    # candidates: 1024-4096 rows at 1 slot, 512-2048 at 2, 256-1024 at 4 and 8
    lo_hi = [(1024, 4096), (512, 2048), (256, 1024), (256, 1024)]
    for (lo, hi), r in zip(lo_hi, st["rows_per_partition"]):
        assert r in (lo, 2 * lo, hi), st["rows_per_partition"]
    assert all(sh in (-1, 0, 1) for sh in st["lane_shift"])
    b = packed_rhs(specs, bt.offsets, gen.SEED_BASE + 241)
    check_batch(specs, mats, "L", bt.offsets, b, bt.solve(b))


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16])
def test_batch_every_lane_count(native, lanes):
    """Every lanes-per-row choice the autotuner may make solves to the bar."""
    specs, mats = small_batch(24, seed=gen.SEED_BASE + 260)
    bt = native.Batch("L", [(m.a, m.ia, m.ja, s["nrhs"]) for s, m in zip(specs, mats)], rows_per_partition=512,
                      lanes_per_row=lanes)
    b = packed_rhs(specs, bt.offsets, gen.SEED_BASE + 261)
    check_batch(specs, mats, "L", bt.offsets, b, bt.solve(b))


def test_batch_matches_single_factors(native):
    """Each matrix solved inside the batch equals its own factor's solve."""
    specs, mats = small_batch(12, seed=gen.SEED_BASE + 250)
    bt = native.Batch("L", [(m.a, m.ia, m.ja, s["nrhs"]) for s, m in zip(specs, mats)], rows_per_partition=1024)
    b = packed_rhs(specs, bt.offsets, gen.SEED_BASE + 251)
    x = bt.solve(b)
    for i, (s, m) in enumerate(zip(specs, mats)):
        n, r = s["n"], s["nrhs"]
        f = native.Factor("L", m.a, m.ia, m.ja, partitions=max(1, (n + 512) // 1024))
        X1 = f.solve(b[bt.offsets[i]:bt.offsets[i + 1]].copy(), nrhs=r)
        assert relerr(x[bt.offsets[i]:bt.offsets[i + 1]], np.asarray(X1).ravel()) <= TOL


def test_batch_rejects_bad_input(native):
    m = gen.banded(100, 8, 4, 0.0, 1)
    with pytest.raises(native.SpikeError):
        native.Batch("L", [(m.a, m.ia, m.ja, 9)])          # nrhs > 8
    ja = m.ja.copy()
    ja[5] = 500                                          # column out of range
    with pytest.raises(native.SpikeError):
        native.Batch("L", [(m.a, m.ia, ja, 1)])


def test_tiles_give_the_same_bits(native):
    """One matrix in groups of 1, 2, 4 and 8 right-hand sides (block-solve
    tiles of that many RHS per wave): a column shared by all of them comes
    out bit for bit the same, as the step arithmetic is per RHS."""
    m = gen.banded(3000, 40, 10, 0.02, gen.SEED_BASE + 260)
    counts = (1, 2, 4, 8)
    bt = native.Batch("L", [(m.a, m.ia, m.ja, r) for r in counts], rows_per_partition=512)
    col = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 261).ravel()
    b = np.zeros(bt.offsets[-1])
    for i, r in enumerate(counts):
        B = gen.rhs(m.n, r, seed=gen.SEED_BASE + 262 + i).reshape(r, m.n)
        B[r - 1] = col
        b[bt.offsets[i]:bt.offsets[i + 1]] = B.ravel()
    x = bt.solve(b)
    ref = x[bt.offsets[0]:bt.offsets[1]]
    assert relerr(ref, orc.ldsol(m.a, m.ia, m.ja, col)) <= TOL
    for i, r in enumerate(counts[1:], 1):
        X = x[bt.offsets[i]:bt.offsets[i + 1]].reshape(r, m.n)
        assert np.array_equal(X[r - 1], ref), r
