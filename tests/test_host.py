"""Host-side checks (no GPU): the C-ABI library loads and exports every symbol
of include/spike_pstrsv.h, the symbolic analysis agrees with the oracle and
with brute force, and invalid input is rejected with an error code."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as orc
from spike_pstrsv_amd import gen

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "spike_pstrsv.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_library_exports_every_declared_symbol():
    from spike_pstrsv_amd import native

    lib = ctypes.CDLL(str(native.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/spike_pstrsv.h but not exported"
    assert sorted(native.EXPORTED) == names


def test_legacy_signatures_match_reference_header():
    """The five legacy prototypes keep the reference's argument lists
    (reference src/spike_pstrsv.h:7, 14, 21, 31, 38)."""
    text = (ROOT / "include" / "spike_pstrsv.h").read_text()
    for proto in [
        "void spike_pstrsv_preproc(double *a, MKL_INT *ia, MKL_INT *ja, int n, int nthreads, const char uplo);",
        "void _spike_pstrsv_preproc(double *a, MKL_INT *ia, MKL_INT *ja, int n, int nthreads, const char uplo);",
        "void spike_pstrsv(const char uplo, int m, double *a, MKL_INT *ia, MKL_INT *ja, double *b, double *x, int nthreads);",
        "void spike_pstrsv_solve(const char uplo, int m, double *b, double *x, int nthreads);",
        "void _spike_pstrsv_solve(const char uplo, int m, double *b, double *x, int nthreads);",
    ]:
        assert proto in text


@pytest.mark.parametrize("P", [1, 3, 8, 64])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_analyze_matches_oracle(P, uplo):
    from spike_pstrsv_amd import native

    m = gen.banded(4096, bw=32, q=8, plr=0.05, seed=P)
    if uplo == "U":
        m = gen.to_upper(m)
    st = native.analyze(uplo, m.a, m.ia, m.ja, partitions=P)
    s = orc.Spike(uplo, m.a, m.ia, m.ja, P)
    os_ = s.stats()
    assert st["partitions"] == P
    assert st["n_reduced"] == os_["n_reduced"]
    if uplo == "L":
        assert st["spike_area"] == os_["spike_area"]
    assert st["steps_pass1"] > 0


def test_analyze_lap2d_depth():
    """2D 5-point lower triangle, 32 grid lines per partition: the block
    solve's dependency depth is 256 + 32 - 1 levels (anti-diagonals)."""
    from spike_pstrsv_amd import native

    m = gen.lap2d(256)
    st = native.analyze("L", m.a, m.ia, m.ja, partitions=8, lanes_per_row=2)
    assert st["steps_pass1"] == 256 + 32 - 1
    assert st["n_reduced"] == 7 * 256


@pytest.mark.parametrize("case,precision,lanes", [
    ("banded", "fp64", 4), ("lap2d", "fp64", 1), ("lap3d", "fp64", 1),
    ("banded", "fp32_refined", 16), ("lap2d", "fp32_refined", 2)])
def test_block_lanes_choice(case, precision, lanes):
    """Automatic lanes per row slot: FP64 factors split rows into chunks of
    about p90/3..p90/6 entries (16-entry banded rows -> 4 lanes, the 2- and
    3-entry Laplacian rows -> 1); FP32 factors cover whole rows (the smallest
    power of two >= p90), so each row's FP32 partial sum is rounded once."""
    from spike_pstrsv_amd import native

    m = {"banded": lambda: gen.banded(1 << 14), "lap2d": lambda: gen.lap2d(64),
         "lap3d": lambda: gen.lap3d(16)}[case]()
    st = native.analyze("L", m.a, m.ia, m.ja, partitions=8, precision=precision)
    assert st["lanes_per_row"] == lanes
    # a row of len entries spans ceil(len / lanes) chunks, each >= 1 step
    # after the previous one: the schedule is at least that deep
    lens = np.diff(m.ia) - 1
    assert st["steps_pass1"] >= int(np.max((lens + lanes - 1) // lanes))


@pytest.mark.parametrize("bad", ["unsorted", "nodiag", "emptyrow", "ia0", "range", "zerodiag", "nandiag"])
def test_invalid_csr_rejected(bad):
    from spike_pstrsv_amd import native

    m = gen.banded(100, bw=8, q=4, plr=0.0, seed=5)
    a, ia, ja = m.a.copy(), m.ia.copy(), m.ja.copy()
    if bad == "unsorted":
        j = ia[50]
        ja[j], ja[j + 1] = ja[j + 1], ja[j]
    elif bad == "nodiag":
        ja[ia[50 + 1] - 1] = ja[ia[50 + 1] - 1] - 1 if ia[51] - ia[50] == 1 else 40
    elif bad == "emptyrow":
        ia[51:] -= ia[51] - ia[50]
        ia[51] = ia[50]
        ia = ia.copy()
    elif bad == "ia0":
        ia = ia + 1
    elif bad == "range":
        ja[3] = 10_000
    elif bad in ("zerodiag", "nandiag"):
        a[ia[51] - 1] = 0.0 if bad == "zerodiag" else np.nan   # row 50's diagonal (last in 'L')
    with pytest.raises(native.SpikeError) as e:
        native.analyze("L", a, ia, ja, partitions=4)
    assert e.value.code == 2
    if bad in ("zerodiag", "nandiag"):
        assert "row 50:" in str(e.value)


def test_invalid_csr_first_defect_and_pointer_jump():
    """Validation reports the first defect in row order, and a row pointer
    past ia[n] is reported without reading entries through it."""
    from spike_pstrsv_amd import native

    m = gen.banded(100, bw=8, q=4, plr=0.0, seed=5)
    a, ia, ja = m.a.copy(), m.ia.copy(), m.ja.copy()
    j = ia[20]
    ja[j], ja[j + 1] = ja[j + 1], ja[j]         # row 20 unsorted
    ia[61:] -= ia[61] - ia[60]                  # row 60 empty
    with pytest.raises(native.SpikeError) as e:
        native.analyze("L", a, ia.copy(), ja, partitions=4)
    assert "row 20 " in str(e.value)
    ia2 = m.ia.copy()
    ia2[50] = 10**9                             # row 49 ends past ia[n]
    with pytest.raises(native.SpikeError) as e:
        native.analyze("L", m.a, ia2, m.ja, partitions=4)
    assert "ia[50] exceeds ia[n]" in str(e.value)


def test_invalid_arguments():
    from spike_pstrsv_amd import native

    m = gen.banded(100, seed=6)
    with pytest.raises(native.SpikeError):
        native.analyze("X", m.a, m.ia, m.ja, partitions=4)
    # more partitions than rows: clamped to one row per partition
    assert native.analyze("L", m.a, m.ia, m.ja, partitions=101)["partitions"] == 100


def test_generators_deterministic_and_well_formed():
    a = gen.banded(5000, seed=9)
    b = gen.banded(5000, seed=9)
    assert np.array_equal(a.a, b.a) and np.array_equal(a.ja, b.ja)
    n = a.n
    for i in range(0, n, 97):
        cols = a.ja[a.ia[i]:a.ia[i + 1]]
        assert cols[-1] == i and np.all(np.diff(cols) > 0)
        vals = a.a[a.ia[i]:a.ia[i + 1]]
        assert abs(vals[-1] - (1 + np.abs(vals[:-1]).sum())) < 1e-15
        band = cols[:-1][cols[:-1] >= max(0, i - 64)]
        assert len(band) in (min(16, i), min(16, i) - 0) or len(cols) - 1 <= min(16, i) + 1
    l2 = gen.lap2d(256)
    assert l2.nnz == 196096 and l2.n == 65536
    l3 = gen.lap3d(16)
    assert l3.nnz == 16 ** 3 + 3 * 16 * 16 * 15
    for m in (l2, l3):
        assert np.all(m.ja[m.ia[1:] - 1] == np.arange(m.n))


def test_banded_1m_expected_size():
    m = gen.banded(1 << 20)
    # E[nnz] = 17,846,628 by the generator's definition; realised within 0.1%
    assert abs(m.nnz - 17_846_628) / 17_846_628 < 1e-3


def globals_declared():
    text = (ROOT / "include" / "spike_pstrsv_globals.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"extern\s+(char|double)\s+(\w+)(\[MAXTHREADCOUNT\])?\s*;", text)


def test_harness_globals_exported():
    """Every global of include/spike_pstrsv_globals.h is a data symbol of the
    library (the names the reference's harness reads)."""
    from spike_pstrsv_amd import native

    lib = ctypes.CDLL(str(native.LIB_PATH))
    decl = globals_declared()
    assert len(decl) >= 21
    for ctype, name, arr in decl:
        t = ctypes.c_char if ctype == "char" else ctypes.c_double
        if arr:
            t = t * 128
        v = t.in_dll(lib, name)
        assert v is not None


def test_acceptance_program_compiles_and_links(tmp_path):
    """The C acceptance program (reference call sequence + harness globals)
    builds against include/ and links against libspike_pstrsv.so."""
    import subprocess

    from spike_pstrsv_amd import native

    exe = tmp_path / "example_usage"
    libdir = str(native.LIB_PATH.parent)
    r = subprocess.run(["gcc", "-O2", "-Wall", "-Werror", "-std=c99", "-I", str(ROOT / "include"),
                        str(ROOT / "tests" / "c" / "example_usage.c"), "-L", libdir, "-lspike_pstrsv",
                        f"-Wl,-rpath,{libdir}", "-lm", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()


def write_mm(path, n, rows, cols, vals, symmetric=False, comments=("% test matrix",)):
    with open(path, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real " + ("symmetric" if symmetric else "general") + "\n")
        for c in comments:
            fh.write(c + "\n")
        fh.write(f"{n} {n} {len(rows)}\n")
        for r, c, v in zip(rows, cols, vals):
            fh.write(f"{int(r) + 1} {int(c) + 1} {float(v)!r}\n")


def test_matrix_market_general_and_triangles(tmp_path):
    """Reader semantics of the reference's matrixMarketToCSR (misc.c:109-245),
    checked against a dense matrix built from the same coordinate list."""
    from spike_pstrsv_amd import native

    rng = np.random.default_rng(5)
    n = 40
    D = np.zeros((n, n))
    mask = rng.random((n, n)) < 0.15
    D[mask] = rng.uniform(-1, 1, mask.sum())
    np.fill_diagonal(D, 2.0 + rng.random(n))
    rows, cols = np.nonzero(D)
    perm = rng.permutation(len(rows))                      # file order is arbitrary
    write_mm(tmp_path / "g.mtx", n, rows[perm], cols[perm], D[rows[perm], cols[perm]],
             comments=("% comment one", "%  comment two"))
    for uplo, ref in (("N", D), ("L", np.tril(D)), ("U", np.triu(D))):
        a, ia, ja, nn = native.read_matrix_market(tmp_path / "g.mtx", uplo, "N")
        assert nn == n
        M = np.zeros((n, n))
        for i in range(n):
            cols_i = ja[ia[i]:ia[i + 1]]
            assert np.all(np.diff(cols_i) > 0), "columns ascending"
            M[i, cols_i] = a[ia[i]:ia[i + 1]]
        assert np.array_equal(M, ref), uplo


def test_matrix_market_symmetric_expansion(tmp_path):
    from spike_pstrsv_amd import native

    m = gen.banded(60, 8, 3, 0.1, 9)
    n = m.n
    L = np.zeros((n, n))
    for i in range(n):
        L[i, m.ja[m.ia[i]:m.ia[i + 1]]] = m.a[m.ia[i]:m.ia[i + 1]]
    S = L + np.tril(L, -1).T                                # symmetric, lower triangle stored
    r, c = np.nonzero(np.tril(S))
    write_mm(tmp_path / "s.mtx", n, r, c, S[r, c], symmetric=True)
    for symm in ("Y", "A"):
        a, ia, ja, nn = native.read_matrix_market(tmp_path / "s.mtx", "U", symm)
        U = np.zeros((n, n))
        for i in range(n):
            U[i, ja[ia[i]:ia[i + 1]]] = a[ia[i]:ia[i + 1]]
        assert np.array_equal(U, np.triu(S))
    a, ia, ja, nn = native.read_matrix_market(tmp_path / "s.mtx", "L", "Y")
    assert np.array_equal(a, m.a) and np.array_equal(ia, m.ia) and np.array_equal(ja, m.ja)


def test_matrix_market_errors(tmp_path):
    from spike_pstrsv_amd import native

    with pytest.raises(native.SpikeError):
        native.read_matrix_market(tmp_path / "missing.mtx")
    (tmp_path / "bad.mtx").write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n")
    with pytest.raises(native.SpikeError):
        native.read_matrix_market(tmp_path / "bad.mtx")
    (tmp_path / "pat.mtx").write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 1\n")
    with pytest.raises(native.SpikeError):
        native.read_matrix_market(tmp_path / "pat.mtx")


def test_factor_load_rejects_foreign_file(tmp_path):
    from spike_pstrsv_amd import native

    p = tmp_path / "x.spk"
    p.write_bytes(b"NOTAFACTORFILE" * 10)
    with pytest.raises(native.SpikeError):
        native.Factor.load(p)
    with pytest.raises(native.SpikeError):
        native.Factor.load(tmp_path / "missing.spk")


def test_bench_mkl_competitors_match_oracle():
    """The CPU competitors bench.py reports (MKL's inspector-executor and
    cspblas triangular solves) solve the same system as the oracle's
    sequential substitution; skipped where MKL's runtime does not load."""
    import sys

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench

    m = gen.banded(20000, 64, 16, 0.02, gen.SEED_BASE + 2)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 3).ravel().copy()
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    out = bench.mkl_competitors(m, b, ref, 2, 0.05)
    if not out["available"]:
        pytest.skip(out["reason"])
    for name in ("mkl_sparse_d_trsv", "mkl_cspblas_dcsrtrsv"):
        assert out[name]["rel_err_vs_oracle"] <= 1e-12, (name, out[name])
        assert out[name]["solves"] >= 1


def test_bench_cpu_baseline_runs_in_a_pinned_child():
    """bench.py's CPU baseline leg runs in a child process (pinned OpenMP
    threads there only) and reports the restatement and the MKL competitors;
    the solution handed in is compared with the oracle."""
    import os
    import sys

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench

    m, P, _ = bench.workload("lap2d_256")
    b = bench.workload_rhs(m.n)
    x = orc.ldsol(m.a, m.ia, m.ja, b)
    cpu, rel = bench.cpu_baseline("lap2d_256", m, b, x, 0.2)
    assert rel == 0.0
    assert cpu["value"] > 0 and cpu["kind"] == "port" and "competitors" in cpu
    assert cpu["omp"]["OMP_PROC_BIND"] == os.environ.get("OMP_PROC_BIND", "close")


def test_legacy_solve_before_preprocess_leaves_x(capfd):
    """The legacy solve without a preprocessed matrix prints one line and
    leaves x untouched (the reference would read uninitialised globals)."""
    import ctypes as C

    from spike_pstrsv_amd import native

    if native.lib.spike_pstrsv_current_factor():
        pytest.skip("a legacy factor is current in this process")
    n = 16
    b = np.ones(n)
    x = np.full(n, 7.0)
    native.lib.spike_pstrsv_solve(b"L", n, C.c_void_p(b.ctypes.data), C.c_void_p(x.ctypes.data), 1)
    err = capfd.readouterr().err
    assert "no preprocessed matrix" in err
    assert np.all(x == 7.0) and np.all(b == 1.0)


def test_zero_diagonal_named_in_callers_order_for_upper():
    """A zero diagonal in an upper triangle is reported at its own row (the
    first such row of the caller's matrix), not at its lower-form index."""
    from spike_pstrsv_amd import native

    u = gen.to_upper(gen.banded(100, bw=8, q=4, plr=0.0, seed=5))
    a = u.a.copy()
    for r in (30, 70):
        a[u.ia[r]] = 0.0                        # the diagonal is first in 'U'
    with pytest.raises(native.SpikeError) as e:
        native.analyze("U", a, u.ia, u.ja, partitions=4)
    assert e.value.code == 2 and "row 30:" in str(e.value)
