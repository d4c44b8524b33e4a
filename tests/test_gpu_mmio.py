"""GPU tests of the reference example program's flow (example_usage.c:71,
:85-101): a MatrixMarket file is read (matrixMarketToCSR, misc.c:109-245, here
spike_read_matrix_market), b = A x_real is formed from the triangle that was
read, and the solve runs through the legacy entry points, in the single-call
form (spike_pstrsv) and the split form (spike_pstrsv_preproc +
spike_pstrsv_solve).

Bar: the reference's own check ||x - x_real||_2 <= 1e-4 (ex.c:104) and
max|x - x_ref| / max|x_ref| <= 1e-12 against sequential FP64 substitution of
the same triangle (the oracle's ldsol / udsol).
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


def _write_mm(path, n, rows, cols, vals, symmetric):
    order = np.lexsort((rows, cols))          # column-major file order, as SuiteSparse writes
    with open(path, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real " + ("symmetric" if symmetric else "general") + "\n")
        fh.write("% written by the spike_pstrsv test suite\n")
        fh.write(f"{n} {n} {len(rows)}\n")
        fh.writelines(f"{int(rows[k]) + 1} {int(cols[k]) + 1} {float(vals[k])!r}\n" for k in order)


@pytest.fixture(scope="module")
def sym_coo():
    """S = L + L^T - D for a diagonally dominant banded lower triangle L
    (n = 6000, bandwidth 64, 16 entries per row, 2% long-range entries)."""
    m = gen.banded(6000, 64, 16, 0.02, gen.SEED_BASE + 470)
    rows = np.repeat(np.arange(m.n), np.diff(m.ia))
    cols = m.ja.astype(np.int64)
    return m, rows, cols, m.a


@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_matrix_market_to_solve(native, tmp_path, sym_coo, symmetric, uplo):
    m, r, c, v = sym_coo
    if symmetric:
        # the lower triangle with a symmetric banner, expanded by the reader
        _write_mm(tmp_path / "A.mtx", m.n, r, c, v, True)
        symm = "Y"
    else:
        off = r != c
        rows = np.concatenate([r, c[off]])
        cols = np.concatenate([c, r[off]])
        vals = np.concatenate([v, v[off]])
        _write_mm(tmp_path / "A.mtx", m.n, rows, cols, vals, False)
        symm = "N"
    a, ia, ja, n = native.read_matrix_market(tmp_path / "A.mtx", uplo, symm)
    assert n == m.n
    if uplo == "L":
        assert np.array_equal(ia, m.ia) and np.array_equal(ja, m.ja) and np.array_equal(a, m.a)
    else:
        assert ia[-1] == m.ia[-1] and np.all(ja[ia[:-1]] == np.arange(n)), "diagonal first in every row"
    x_real = orc.manufactured_x(n)
    b = orc.spmv(a, ia, ja, x_real)
    ref = orc.trsv(uplo, a, ia, ja, b)
    x = np.zeros(n)
    native.spike_pstrsv(uplo, n, a, ia, ja, b, x, 8)
    err = relerr(x, ref)
    print(f"mtx {'symmetric' if symmetric else 'general'} uplo={uplo} n={n}: rel.err {err:.3e}, "
          f"||x - x_real|| = {np.linalg.norm(x - x_real):.3e}")
    assert np.linalg.norm(x - x_real) <= 1e-4
    assert err <= TOL
    x2 = np.zeros(n)
    native.spike_pstrsv_preproc(a, ia, ja, n, 8, uplo)
    native.spike_pstrsv_solve(uplo, n, b, x2, 8)
    assert relerr(x2, ref) <= TOL
