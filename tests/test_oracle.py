"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

The reference repository ships no fixtures or golden vectors for this path;
its only known-answer check is the manufactured solution of
test/runtime_analysis.c:315-318 / :434-435.  Each oracle function is pinned
here by an independent check chosen so that a plausible mistake (a dropped
term, a wrong sign or index, a transposed operand) fails:

  ldsol/udsol      dense numpy substitution and the residual ||Ax - b||
  spmv             dense product
  partition        closed form of src/matrix_partitioning.c:18-33
  flip             element-wise U = J L J
  manufactured_x   glibc rand() sequence (values for srand(1) are fixed)
  Spike ('L','U')  equals sequential substitution for many P; dependents,
                   K_p, intervals against brute force from the definitions;
                   spike invariant D0 W = C[:, K]; the 2D-Laplacian spike
                   against its closed-form Green's function; 'U' == flipped 'L'
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as orc
from spike_pstrsv_amd import gen


def dense(m):
    n = m.n
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(m.ia[i], m.ia[i + 1]):
            A[i, m.ja[j]] = m.a[j]
    return A


def rel(x, r):
    return float(np.max(np.abs(x - r)) / max(np.max(np.abs(r)), 1e-300))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ldsol_udsol_against_dense(seed):
    m = gen.banded(200, bw=20, q=6, plr=0.2, seed=seed)
    A = dense(m)
    b = gen.rhs(m.n, 1, seed=seed + 10)
    x = orc.ldsol(m.a, m.ia, m.ja, b)
    assert rel(x, np.linalg.solve(A, b)) < 1e-12
    assert np.max(np.abs(A @ x - b)) < 1e-12
    u = gen.to_upper(m)
    xu = orc.udsol(u.a, u.ia, u.ja, b)
    assert rel(xu, np.linalg.solve(dense(u), b)) < 1e-12


def test_spmv_dense():
    m = gen.banded(150, bw=10, q=5, plr=0.3, seed=7)
    x = gen.rhs(m.n, 1, seed=8)
    assert np.allclose(orc.spmv(m.a, m.ia, m.ja, x), dense(m) @ x, rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,P", [(10, 3), (65536, 8), (1 << 20, 1024), (1000, 7), (5, 5)])
def test_partition_closed_form(n, P):
    s = orc.partition(n, P)
    k = n // P
    assert list(s[:P]) == [k * t for t in range(P)]
    assert s[P] == n and s[P] - s[P - 1] == k + n % P


def test_flip_elementwise():
    m = gen.banded(40, bw=6, q=3, plr=0.5, seed=3)
    a2, ia2, ja2 = orc.flip(m.a, m.ia, m.ja)
    L = dense(m)
    U = dense(gen.CSR(a2, ia2, ja2, "U"))
    n = m.n
    J = np.eye(n)[::-1]
    assert np.array_equal(U, J @ L @ J)
    assert all(ja2[ia2[i]] == i for i in range(n))          # diagonal first
    # the generator's own flip agrees
    u = gen.to_upper(m)
    assert np.array_equal(u.a, a2) and np.array_equal(u.ia, ia2) and np.array_equal(u.ja, ja2)


def test_manufactured_x_glibc_sequence():
    x = orc.manufactured_x(5, 1)
    # glibc rand() after srand(1): 1804289383, 846930886, 1681692777, 1714636915, 1957747793
    first = [1804289383, 846930886, 1681692777, 1714636915, 1957747793]
    expect = [(r % 71 + 1) + (i % 1000) / 10.0 for i, r in enumerate(first)]
    assert list(x) == expect


def test_manufactured_solution_check():
    """ra.c:315-318 and :434-435: b = A x_real, ||x - x_real||_2 <= 1e-4."""
    m = gen.lap2d(64)
    xr = orc.manufactured_x(m.n)
    b = orc.spmv(m.a, m.ia, m.ja, xr)
    assert np.linalg.norm(orc.ldsol(m.a, m.ia, m.ja, b) - xr) <= 1e-4
    s = orc.Spike("L", m.a, m.ia, m.ja, 8)
    assert np.linalg.norm(s.solve(b) - xr) <= 1e-4


@pytest.mark.parametrize("P", [1, 2, 3, 8, 33, 200])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_spike_equals_substitution(P, uplo):
    m = gen.banded(600, bw=24, q=8, plr=0.1, seed=P)
    if uplo == "U":
        m = gen.to_upper(m)
    b = gen.rhs(m.n, 1, seed=P + 1)
    ref = orc.trsv(uplo, m.a, m.ia, m.ja, b)
    s = orc.Spike(uplo, m.a, m.ia, m.ja, P)
    assert rel(s.solve(b), ref) <= 1e-13
    assert s.stats()["dropped_rowsum_max"] == 0.0


def brute_force_structure(m, P, uplo):
    n = m.n
    starts = orc.partition(n, P)
    dep = set()
    info = []
    for t in range(P):
        st, en = starts[t], starts[t + 1]
        cpl_rows = []
        for i in range(st, en):
            cols = m.ja[m.ia[i]:m.ia[i + 1]]
            c = [int(x) for x in cols if (x < st if uplo == "L" else x >= en)]
            if c:
                cpl_rows.append(i - st)
                dep.update(c)
        info.append((st, en, cpl_rows))
    return starts, sorted(dep), info


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_spike_structure_brute_force(uplo):
    m = gen.banded(64, bw=9, q=3, plr=0.3, seed=11)
    if uplo == "U":
        m = gen.to_upper(m)
    P = 5
    s = orc.Spike(uplo, m.a, m.ia, m.ja, P)
    starts, dep, info = brute_force_structure(m, P, uplo)
    assert list(s.dependents()) == dep
    depset = set(dep)
    for t in range(P):
        st, en, cpl = info[t]
        pi = s.part(t)
        assert (pi["start"], pi["end"]) == (st, en)
        k = en - st
        local_deps = [r for r in range(k) if st + r in depset]
        if uplo == "L":
            f = cpl[0] if cpl else k
            g = local_deps[-1] if local_deps else -1
            assert pi["first_coupling"] == f and pi["last_dependent"] == g
            lo, hi = f, g
        else:
            f = cpl[-1] if cpl else -1                      # last coupling row
            g = local_deps[0] if local_deps else k          # first dependent row
            assert pi["first_coupling"] == f and pi["last_dependent"] == g
            lo, hi = g, f
            if t == 0:
                lo, hi = 0, -1
        rows = range(lo, hi + 1) if (local_deps and hi >= lo) else []
        K = sorted({int(c) for r in rows for c in m.ja[m.ia[st + r]:m.ia[st + r + 1]]
                    if (c < st if uplo == "L" else c >= en)})
        assert list(s.kcols(t)) == K
        assert pi["cols_S"] == len(K)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_spike_invariant(uplo):
    """D0 W = C[:, K] on the spike rows (ops.c:1021-1069 restated)."""
    m = gen.banded(400, bw=16, q=6, plr=0.1, seed=21)
    if uplo == "U":
        m = gen.to_upper(m)
    P = 4
    s = orc.Spike(uplo, m.a, m.ia, m.ja, P)
    A = dense(m)
    checked = 0
    for t in range(P):
        pi = s.part(t)
        if pi["rows_S"] == 0 or pi["cols_S"] == 0:
            continue
        st = pi["start"]
        lo = pi["first_coupling"] if uplo == "L" else pi["last_dependent"]
        rows = np.arange(st + lo, st + lo + pi["rows_S"])
        K = s.kcols(t)
        W = s.block(t)
        D0 = A[np.ix_(rows, rows)]
        assert np.max(np.abs(D0 @ W - A[np.ix_(rows, K)])) < 1e-13
        checked += 1
    assert checked >= 2


def test_spike_lap2d_green_function():
    """2D 5-point lower triangle (diag 4, off -1): the spike of the last grid
    line of partition p-1 on the rows of partition p is the Green's function
    G(a, b) = C(a+b, a) / 4^(a+b+1) of the unit source; row norms decay as
    3^-(b+1)."""
    g = 16
    m = gen.lap2d(g)
    P = 4                                   # 4 grid lines per partition
    s = orc.Spike("L", m.a, m.ia, m.ja, P)
    W = s.block(1)
    K = s.kcols(1)
    pi = s.part(1)
    st = pi["start"]
    lo = pi["first_coupling"]
    assert lo == 0 and list(K) == list(range(st - g, st))
    # W[r, c]: row r = st + b*g + gx, column c = st - g + cx (source above
    # line 0 at x = cx); coupling value -1, so W = -G(gx - cx, b) for gx >= cx
    for b in range(min(4, pi["rows_S"] // g)):
        for gx in range(g):
            for cx in range(g):
                w = W[b * g + gx, cx]
                a = gx - cx
                expect = 0.0 if a < 0 else -math.comb(a + b, a) / 4.0 ** (a + b + 1)
                assert abs(w - expect) <= 1e-15, (b, gx, cx, w, expect)


@pytest.mark.parametrize("P", [2, 6, 17])
def test_spike_upper_equals_flipped_lower(P):
    m = gen.banded(500, bw=12, q=5, plr=0.15, seed=31)
    u = gen.to_upper(m)
    b = gen.rhs(m.n, 1, seed=32)
    xl = orc.Spike("L", m.a, m.ia, m.ja, P).solve(b)
    xu = orc.Spike("U", u.a, u.ia, u.ja, P).solve(b[::-1].copy())
    assert rel(xu[::-1], xl) <= 1e-13


def test_spike_truncation_bound():
    m = gen.banded(1 << 14, seed=41)
    b = gen.rhs(m.n, 1, seed=42)
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    s = orc.Spike("L", m.a, m.ia, m.ja, 16, tau=1e-17)
    st = s.stats()
    assert 0 < st["dropped_rowsum_max"] <= 16 * 80 * 1e-17
    assert rel(s.solve(b), ref) <= 1e-13


def test_oracle_rejects_bad_input():
    m = gen.banded(50, bw=5, q=2, plr=0.0, seed=1)
    with pytest.raises(ValueError):
        orc.Spike("U", m.a, m.ia, m.ja, 2)        # diagonal is last, not first
    with pytest.raises(ValueError):
        orc.Spike("L", m.a, m.ia, m.ja, 0)
    ja = m.ja.copy()
    ja[m.ia[10]], ja[m.ia[10] + 1] = ja[m.ia[10] + 1], ja[m.ia[10]]
    with pytest.raises(ValueError):
        orc.Spike("L", m.a, m.ia, ja, 2)
