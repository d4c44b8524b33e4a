"""GPU parity of the FP32-valued solve with FP64 iterative refinement.

Bar: relative error max|x - x_ref| / max|x_ref| <= 1e-12 against sequential
FP64 substitution (the oracle's ldsol / udsol), the same bar as the FP64
path.  The unrefined FP32 solve is checked to be an FP32-accurate result
(error well above FP64 rounding), so a test cannot pass on an FP64 path that
ignores the precision option.
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


def test_refined_lap2d_256(native):
    m = gen.lap2d(256)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 101)
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=8, precision="fp32_refined")
    st = f.stats()
    print("lap2d_256 fp32_refined", st)
    assert st["precision"] == 1 and st["refine_steps"] >= 1
    assert st["refine_est_err"] <= 1e-13
    assert relerr(f.solve(b), ref) <= TOL


def test_unrefined_fp32_is_fp32_accurate(native):
    m = gen.banded(1 << 16, seed=gen.SEED_BASE + 102)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 103)
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    f0 = native.Factor("L", m.a, m.ia, m.ja, partitions=64, precision="fp32_refined", refine_steps=0)
    e0 = relerr(f0.solve(b), ref)
    print("fp32 without refinement: relerr", e0)
    assert 1e-10 < e0 < 1e-5
    f1 = native.Factor("L", m.a, m.ia, m.ja, partitions=64, precision="fp32_refined", refine_steps=1)
    e1 = relerr(f1.solve(b), ref)
    print("fp32 + 1 refinement step: relerr", e1)
    assert e1 < e0 * 1e-4


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("nrhs", [1, 5, 8])
def test_refined_banded_multi_rhs(native, uplo, nrhs):
    m = gen.banded(1 << 16, seed=gen.SEED_BASE + 110 + nrhs)
    if uplo == "U":
        m = gen.to_upper(m)
    B = gen.rhs(m.n, nrhs, seed=gen.SEED_BASE + 111)
    f = native.Factor(uplo, m.a, m.ia, m.ja, partitions=64, precision="fp32_refined")
    X = f.solve(B, nrhs=nrhs).reshape(nrhs, m.n)
    for j in range(nrhs):
        assert relerr(X[j], orc.trsv(uplo, m.a, m.ia, m.ja, B.reshape(nrhs, m.n)[j])) <= TOL


def test_refined_lap3d_32_multi_rhs(native):
    """3D 7-point Laplacian lower triangle (diag 6, off -1): plane coupling."""
    m = gen.lap3d(32)
    B = gen.rhs(m.n, 4, seed=gen.SEED_BASE + 120)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=16, precision="fp32_refined")
    print("lap3d_32 fp32_refined", f.stats())
    X = f.solve(B, nrhs=4)
    for j in range(4):
        assert relerr(X[j], orc.ldsol(m.a, m.ia, m.ja, B[j])) <= TOL


def test_refined_matches_fp64_path(native):
    m = gen.banded(1 << 15, seed=gen.SEED_BASE + 130)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 131)
    x64 = native.Factor("L", m.a, m.ia, m.ja, partitions=32).solve(b)
    x32 = native.Factor("L", m.a, m.ia, m.ja, partitions=32, precision="fp32_refined").solve(b)
    assert relerr(x32, x64) <= TOL


def test_refined_device_pointers_and_profiled(native):
    import torch

    m = gen.banded(1 << 15, seed=gen.SEED_BASE + 140)
    B = gen.rhs(m.n, 3, seed=gen.SEED_BASE + 141)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=32, precision="fp32_refined")
    X = f.solve(B, nrhs=3)
    db = torch.from_numpy(B).cuda()
    dx = torch.zeros_like(db)
    s = torch.cuda.current_stream().cuda_stream
    f.solve(db.data_ptr(), dx.data_ptr(), nrhs=3, stream=s)
    torch.cuda.synchronize()
    assert np.array_equal(dx.cpu().numpy(), X)
    ph = f.solve_profiled(db.data_ptr(), dx.data_ptr(), nrhs=3, stream=s)
    assert len(ph) == 4 and all(v > 0 for v in ph)
    assert np.array_equal(dx.cpu().numpy(), X)


# ---------------------------------------------------------------- level method

@pytest.mark.parametrize("uplo", ["L", "U"])
def test_levels_chosen_for_lap3d(native, uplo):
    """Plane coupling wider than a partition: the cost model picks the level
    schedule (the Spike reduced system would hold nearly every row)."""
    m = gen.lap3d(64)
    if uplo == "U":
        m = gen.to_upper(m)
    B = gen.rhs(m.n, 3, seed=gen.SEED_BASE + 150)
    f = native.Factor(uplo, m.a, m.ia, m.ja)
    st = f.stats()
    print("lap3d_64", uplo, st)
    assert st["method"] == 2 and st["levels"] == 3 * 63 + 1
    X = f.solve(B, nrhs=3)
    for j in range(3):
        assert relerr(X[j], orc.trsv(uplo, m.a, m.ia, m.ja, B[j])) <= TOL


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("nrhs", [2, 3, 17, 32])
def test_levels_refined_interleaved(native, uplo, nrhs):
    """FP32 level solve with one refinement step of several right-hand sides:
    the residual and the correction stay interleaved between launches."""
    m = gen.lap3d(64)
    if uplo == "U":
        m = gen.to_upper(m)
    B = gen.rhs(m.n, nrhs, seed=gen.SEED_BASE + 160 + nrhs)
    f = native.Factor(uplo, m.a, m.ia, m.ja, precision="fp32_refined")
    st = f.stats()
    print("lap3d_64 levels fp32_refined", uplo, nrhs, st)
    assert st["method"] == 2 and st["refine_steps"] == 1
    X = f.solve(B, nrhs=nrhs).reshape(nrhs, m.n)
    for j in range(nrhs):
        assert relerr(X[j], orc.trsv(uplo, m.a, m.ia, m.ja, B.reshape(nrhs, m.n)[j])) <= TOL


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_levels_refined_single_rhs(native, uplo):
    """One right-hand side of an FP32 level factor with one refinement step:
    one launch per level does x0, the FP64 residual and the correction.  The
    profiled solve runs the separate launches (x0 solve, residual launch,
    correction solve); both meet the oracle bar and agree to rounding."""
    import torch

    m = gen.lap3d(48)
    if uplo == "U":
        m = gen.to_upper(m)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 185)
    f = native.Factor(uplo, m.a, m.ia, m.ja, method="levels", precision="fp32_refined", refine_steps=1)
    x = f.solve(b)
    want = orc.trsv(uplo, m.a, m.ia, m.ja, b)
    print("lap3d_48 levels fp32_refined 1 RHS", uplo, relerr(x, want))
    assert relerr(x, want) <= TOL
    db = torch.from_numpy(np.ascontiguousarray(b)).to("cuda:0")
    dx = torch.zeros_like(db)
    f.solve_profiled(db.data_ptr(), dx.data_ptr())
    xp = dx.cpu().numpy()
    assert relerr(xp, want) <= TOL
    assert relerr(x, xp) <= 1e-13


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
@pytest.mark.parametrize("nrhs", [1, 5, 11])
def test_levels_forced_banded(native, precision, nrhs):
    m = gen.banded(1 << 13, 32, 8, 0.05, gen.SEED_BASE + 160)
    B = gen.rhs(m.n, nrhs, seed=gen.SEED_BASE + 161).reshape(nrhs, m.n)
    f = native.Factor("L", m.a, m.ia, m.ja, method="levels", precision=precision)
    assert f.stats()["method"] == 2
    X = f.solve(B, nrhs=nrhs).reshape(nrhs, m.n)
    for j in range(nrhs):
        assert relerr(X[j], orc.ldsol(m.a, m.ia, m.ja, B[j])) <= TOL


def test_levels_lap3d_64_refined_multi_rhs(native):
    m = gen.lap3d(64)
    B = gen.rhs(m.n, 4, seed=gen.SEED_BASE + 170)
    f = native.Factor("L", m.a, m.ia, m.ja, precision="fp32_refined")
    st = f.stats()
    assert st["method"] == 2 and st["refine_steps"] >= 1
    X = f.solve(B, nrhs=4)
    for j in range(4):
        assert relerr(X[j], orc.ldsol(m.a, m.ia, m.ja, B[j])) <= TOL


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_sync_free_matches_level_launches(native, monkeypatch, uplo, precision):
    """Multi-RHS level solves run as one sync-free launch (rows poll the x
    values they read) or as one launch per level: same arithmetic per row,
    so the same bits, and both at the parity bar."""
    m = gen.lap3d(32)
    if uplo == "U":
        m = gen.to_upper(m)
    B = gen.rhs(m.n, 8, seed=gen.SEED_BASE + 180)
    monkeypatch.setenv("SPIKE_PSTRSV_SYNC_FREE", "0")
    f0 = native.Factor(uplo, m.a, m.ia, m.ja, method="levels", precision=precision)
    monkeypatch.setenv("SPIKE_PSTRSV_SYNC_FREE", "1")
    f1 = native.Factor(uplo, m.a, m.ia, m.ja, method="levels", precision=precision)
    X0 = f0.solve(B, nrhs=8)
    X1 = f1.solve(B, nrhs=8)
    assert np.array_equal(X0, X1)
    for j in range(8):
        assert relerr(X1[j], orc.trsv(uplo, m.a, m.ia, m.ja, B[j])) <= TOL


def test_sync_free_refined_give_up_is_reported(native, monkeypatch):
    """The FP32 level solve with one refinement step runs as one sync-free
    launch (x0, the FP64 residual and the correction per row, published as
    one packed word).  With a poll limit of 0 (forced from the host) its
    give-up is reported as SPIKE_ERR_DEVICE, and every finite value is the
    result of the full solve."""
    m = gen.lap3d(32)
    B = gen.rhs(m.n, 8, seed=gen.SEED_BASE + 182)
    want = native.Factor("L", m.a, m.ia, m.ja, method="levels", precision="fp32_refined",
                         refine_steps=1).solve(B, nrhs=8)
    monkeypatch.setenv("SPIKE_PSTRSV_SPIN_LIMIT", "0")
    f = native.Factor("L", m.a, m.ia, m.ja, method="levels", precision="fp32_refined", refine_steps=1)
    X = np.zeros((8, m.n))
    with pytest.raises(native.SpikeError) as ei:
        f.solve(B, X, nrhs=8)
    assert ei.value.code == 3, ei.value
    fin = np.isfinite(X)
    assert not fin.all(), "the give-up path did not run"
    assert np.array_equal(X[fin], want[fin])
    f.status()


def test_sync_free_give_up_is_reported(native, monkeypatch):
    """With a poll limit of 0 (forced from the host) rows that find an
    operand unpublished give up at once.  The solve still ends, every finite
    value is the exact result, and the give-up is returned as
    SPIKE_ERR_DEVICE: by the call itself when it synchronises (host
    pointers), by spike_factor_status after an asynchronous (device-pointer)
    solve; each give-up is reported once."""
    import torch

    m = gen.lap3d(32)
    B = gen.rhs(m.n, 8, seed=gen.SEED_BASE + 181)
    want = native.Factor("L", m.a, m.ia, m.ja, method="levels").solve(B, nrhs=8)
    monkeypatch.setenv("SPIKE_PSTRSV_SPIN_LIMIT", "0")
    f = native.Factor("L", m.a, m.ia, m.ja, method="levels")
    X = np.zeros((8, m.n))
    with pytest.raises(native.SpikeError) as ei:
        f.solve(B, X, nrhs=8)
    assert ei.value.code == 3, ei.value
    fin = np.isfinite(X)
    assert not fin.all(), "the give-up path did not run"
    assert np.array_equal(X[fin], want[fin])
    f.status()                                   # reported once
    db = torch.from_numpy(np.ascontiguousarray(B)).to("cuda:0")
    dx = torch.zeros_like(db)
    f.solve(db.data_ptr(), dx.data_ptr(), nrhs=8)   # asynchronous: returns
    torch.cuda.synchronize()
    with pytest.raises(native.SpikeError) as ei:
        f.status()
    assert ei.value.code == 3, ei.value
    f.status()
