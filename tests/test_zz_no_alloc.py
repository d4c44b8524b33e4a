"""GPU test of spike_factor_solve_noalloc (the entry point that allocates
nothing): through the C ABI, its output equals spike_factor_solve's output
bit for bit (same kernels), on two calls in a row where the second has the
smaller input, for the Spike method (FP64 and FP32 + refinement) and the
level method (one and several right-hand sides); sizes beyond its static
scratch and host pointers are refused before anything is launched.
"""
import numpy as np
import pytest

from oracle import pyoracle as orc
from spike_pstrsv_amd import gen

pytestmark = pytest.mark.gpu


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="module")
def native():
    import torch

    assert torch.cuda.is_available(), "no GPU visible"
    from spike_pstrsv_amd import native as nat

    return nat


def _both(native, f, B, nrhs):
    """(existing entry point's x, no-alloc entry point's x) for device copies of B."""
    import torch

    dB = torch.from_numpy(np.ascontiguousarray(B)).cuda()
    x_ref = torch.zeros_like(dB)
    x_na = torch.full_like(dB, np.nan)
    f.solve(dB.data_ptr(), x_ref.data_ptr(), nrhs=nrhs)
    f.solve_noalloc(dB.data_ptr(), x_na.data_ptr(), nrhs=nrhs)
    torch.cuda.synchronize()
    f.status()
    return x_ref.cpu().numpy(), x_na.cpu().numpy()


@pytest.mark.parametrize("case", ["spike_fp64", "spike_fp32", "levels_fp64", "levels_fp32"])
def test_noalloc_matches_solve(native, case):
    # first call: the larger input; second call: a smaller matrix and fewer
    # right-hand sides on the same (shared) scratch
    if case.startswith("spike"):
        big, small = gen.banded(1 << 18, seed=gen.SEED_BASE + 700), gen.banded(1 << 14, seed=gen.SEED_BASE + 701)
        kw = {"method": "spike"}
    else:
        big, small = gen.lap3d(48), gen.lap3d(20)
        kw = {"method": "levels"}
    prec = "fp32_refined" if case.endswith("fp32") else "fp64"
    for m, nrhs in ((big, 5), (small, 1)):
        f = native.Factor("L", m.a, m.ia, m.ja, precision=prec, **kw)
        B = gen.rhs(m.n, nrhs, seed=gen.SEED_BASE + 702 + nrhs).reshape(nrhs, m.n)
        x_ref, x_na = _both(native, f, B, nrhs)
        assert np.array_equal(x_ref, x_na), (case, m.n, nrhs)
        assert relerr(x_na[0], orc.ldsol(m.a, m.ia, m.ja, B[0])) <= 1e-12
        f.close()


def test_noalloc_refuses(native):
    import torch

    m = gen.lap3d(128)
    f = native.Factor("L", m.a, m.ia, m.ja, method="levels")
    dB = torch.zeros(32 * m.n, dtype=torch.float64, device="cuda:0")
    dX = torch.full_like(dB, 7.0)
    with pytest.raises(native.SpikeError) as ei:      # 2 x n x 32 values: beyond the scratch
        f.solve_noalloc(dB.data_ptr(), dX.data_ptr(), nrhs=32)
    assert ei.value.code == 6
    b = np.zeros(m.n)
    x = np.zeros(m.n)
    with pytest.raises(native.SpikeError) as ei:      # host pointers would need staging
        f.solve_noalloc(b, x)
    assert ei.value.code == 1
    torch.cuda.synchronize()
    assert bool(torch.all(dX == 7.0)), "a refused call launched work"
    f.close()
