"""GPU parity: the HIP Spike solve (through the C ABI) against the CPU oracle.

Bar: relative error max|x - x_ref| / max|x_ref| <= 1e-12 against sequential
FP64 substitution (oracle_ldsol / oracle_udsol, i.e. ops.c:1264-1279 and
ops.c:1179-1194 of the reference).
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


@pytest.fixture(scope="module")
def lap2d():
    return gen.lap2d(256)


@pytest.fixture(scope="module")
def band1m():
    return gen.banded(1 << 20)


def test_lap2d_256_P8_lower(native, lap2d):
    m = lap2d
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 1)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=8)
    x = f.solve(b)
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    st = f.stats()
    print("lap2d_256 P=8", st, "relerr", relerr(x, ref))
    assert st["partitions"] == 8
    assert relerr(x, ref) <= TOL


def test_lap2d_256_P8_upper(native, lap2d):
    u = gen.to_upper(lap2d)
    b = gen.rhs(u.n, 1, seed=gen.SEED_BASE + 11)
    f = native.Factor("U", u.a, u.ia, u.ja, partitions=8)
    x = f.solve(b)
    ref = orc.udsol(u.a, u.ia, u.ja, b)
    assert relerr(x, ref) <= TOL


@pytest.mark.parametrize("P", [1, 2, 7, 64, 256])
def test_banded_64k(native, P):
    m = gen.banded(1 << 16, seed=gen.SEED_BASE + 20 + P)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 21)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=P)
    x = f.solve(b)
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    assert relerr(x, ref) <= TOL


def test_banded_1m_P1024(native, band1m):
    m = band1m
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 2)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=1024)
    x = f.solve(b)
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    st = f.stats()
    print("banded_1m P=1024", st, "relerr", relerr(x, ref))
    assert relerr(x, ref) <= TOL


def test_legacy_api_manufactured(native, lap2d):
    """The reference harness's check (ra.c:315-318, :434-435): b = A x_real,
    ||x - x_real||_2 <= 1e-4, through the legacy entry points."""
    m = lap2d
    x_real = orc.manufactured_x(m.n)
    b = orc.spmv(m.a, m.ia, m.ja, x_real)
    b_copy = b.copy()
    x = np.zeros(m.n)
    native.spike_pstrsv("L", m.n, m.a, m.ia, m.ja, b, x, 8)
    assert np.array_equal(b, b_copy), "b must not be modified"
    assert np.linalg.norm(x - x_real) <= 1e-4
    assert relerr(x, orc.ldsol(m.a, m.ia, m.ja, b)) <= TOL
    # split form (ex.c:100-101)
    x2 = np.zeros(m.n)
    native.spike_pstrsv_preproc(m.a, m.ia, m.ja, m.n, 8, "L")
    native.spike_pstrsv_solve("L", m.n, b, x2, 8)
    assert relerr(x2, x) <= TOL
    assert native.current_stats()["partitions"] == 8


@pytest.mark.parametrize("nrhs", [2, 3, 5, 8, 13])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_multi_rhs(native, nrhs, uplo):
    """Right-hand sides share each block-solve launch (up to 8 per chunk)."""
    m = gen.banded(1 << 15, seed=gen.SEED_BASE + 70 + nrhs)
    if uplo == "U":
        m = gen.to_upper(m)
    B = gen.rhs(m.n, nrhs, seed=gen.SEED_BASE + 71)
    f = native.Factor(uplo, m.a, m.ia, m.ja, partitions=32)
    X = f.solve(B, nrhs=nrhs)
    for j in range(nrhs):
        assert relerr(X[j], orc.trsv(uplo, m.a, m.ia, m.ja, B[j])) <= TOL
    # column j solved alone gives the same bits
    xj = f.solve(B[nrhs - 1].copy())
    assert np.array_equal(xj, X[nrhs - 1])


def test_multi_rhs_leading_dimension(native):
    m = gen.lap2d(128)
    n = m.n
    ld = n + 37
    B = np.zeros((4, ld))
    B[:, :n] = gen.rhs(n, 4, seed=gen.SEED_BASE + 75)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=8)
    X = np.full((4, ld), 7.0)
    f.solve(B, X, nrhs=4, ldb=ld, ldx=ld)
    for j in range(4):
        assert relerr(X[j, :n], orc.ldsol(m.a, m.ia, m.ja, B[j, :n])) <= TOL
        assert np.all(X[j, n:] == 7.0), "entries past n must not be written"


def test_multi_rhs_and_device_pointers(native):
    import torch

    m = gen.banded(1 << 15, seed=gen.SEED_BASE + 30)
    B = gen.rhs(m.n, 3, seed=gen.SEED_BASE + 31)  # (3, n) column-major
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=32)
    X = f.solve(B, nrhs=3)
    for j in range(3):
        assert relerr(X[j], orc.ldsol(m.a, m.ia, m.ja, B[j])) <= TOL
    db = torch.from_numpy(B).cuda()
    dx = torch.zeros_like(db)
    f.solve(db.data_ptr(), dx.data_ptr(), nrhs=3, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(dx.cpu().numpy(), X), "device-pointer path must match the host path bit for bit"


@pytest.mark.parametrize("n,P", [(1, 1), (2, 2), (5, 5), (97, 3), (300, 299)])
def test_tiny_and_edge(native, n, P):
    m = gen.banded(n, bw=8, q=4, plr=0.3, seed=gen.SEED_BASE + 40 + n)
    b = gen.rhs(n, 1, seed=gen.SEED_BASE + 41)
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    for lanes in (1, 4, 16):
        f = native.Factor("L", m.a, m.ia, m.ja, partitions=P, lanes_per_row=lanes)
        assert relerr(f.solve(b), ref) <= TOL


def test_long_rows_are_chunked(native):
    """Rows longer than the lane group are split into partial chunks."""
    m = gen.banded(4096, bw=200, q=90, plr=0.5, seed=gen.SEED_BASE + 50)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 51)
    ref = orc.ldsol(m.a, m.ia, m.ja, b)
    for lanes in (2, 16):
        f = native.Factor("L", m.a, m.ia, m.ja, partitions=4, lanes_per_row=lanes)
        assert relerr(f.solve(b), ref) <= TOL


def test_deterministic(native):
    m = gen.banded(1 << 16, seed=gen.SEED_BASE + 60)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 61)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=64)
    x1 = f.solve(b)
    x2 = f.solve(b)
    assert np.array_equal(x1, x2)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_caller_partitions(native, uplo):
    """Caller-provided partition boundaries (uneven, in the caller's row order)."""
    m = gen.banded(1 << 15, seed=gen.SEED_BASE + 80)
    if uplo == "U":
        m = gen.to_upper(m)
    rng = np.random.default_rng(3)
    cuts = np.sort(rng.choice(np.arange(1, m.n), size=23, replace=False))
    starts = np.concatenate([[0], cuts, [m.n]]).astype(np.int32)
    assert np.max(np.diff(starts)) <= 16000
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 81)
    f = native.Factor(uplo, m.a, m.ia, m.ja, starts=starts)
    assert f.stats()["partitions"] == 24 and f.stats()["method"] == 1
    assert relerr(f.solve(b), orc.trsv(uplo, m.a, m.ia, m.ja, b)) <= TOL
    # the formula's boundaries reproduce the default factor bit for bit
    P = 32
    k = m.n // P
    fs = native.Factor(uplo, m.a, m.ia, m.ja, starts=[k * t for t in range(P)] + [m.n])
    fd = native.Factor(uplo, m.a, m.ia, m.ja, partitions=P)
    if uplo == "L":
        assert np.array_equal(fs.solve(b), fd.solve(b))
    else:
        assert relerr(fs.solve(b), fd.solve(b)) <= TOL


def test_acceptance_program(tmp_path):
    """The C acceptance program (reference call sequence, harness globals)."""
    import subprocess
    from pathlib import Path

    import torch

    assert torch.cuda.is_available()
    from spike_pstrsv_amd import native

    root = Path(__file__).resolve().parent.parent
    exe = tmp_path / "example_usage"
    libdir = str(native.LIB_PATH.parent)
    r = subprocess.run(["gcc", "-O2", "-std=c99", "-I", str(root / "include"),
                        str(root / "tests" / "c" / "example_usage.c"), "-L", libdir, "-lspike_pstrsv",
                        f"-Wl,-rpath,{libdir}", "-lm", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(run.stdout[-3000:])
    assert run.returncode == 0 and "PASSED" in run.stdout, run.stdout + run.stderr
    assert "PSTRSV" in run.stdout or "partition" in run.stdout.lower()


@pytest.mark.parametrize("case", ["banded_L", "banded_U_fp32", "lap2d_exact", "lap3d_levels"])
def test_factor_save_load(native, tmp_path, case):
    """A saved and reloaded factor solves bit for bit like the original."""
    if case == "banded_L":
        m, kw = gen.banded(1 << 15, seed=gen.SEED_BASE + 90), {"partitions": 32}
        uplo = "L"
    elif case == "banded_U_fp32":
        m, kw = gen.to_upper(gen.banded(1 << 15, seed=gen.SEED_BASE + 91)), {"precision": "fp32_refined"}
        uplo = "U"
    elif case == "lap2d_exact":
        m, kw = gen.lap2d(128), {"partitions": 8, "reduced_solver": "exact"}
        uplo = "L"
    else:
        m, kw = gen.lap3d(40), {"method": "levels"}
        uplo = "L"
    B = gen.rhs(m.n, 3, seed=gen.SEED_BASE + 92)
    f = native.Factor(uplo, m.a, m.ia, m.ja, **kw)
    X0 = f.solve(B, nrhs=3)
    path = tmp_path / "factor.spk"
    f.save(path)
    s0 = f.stats()
    f.close()
    g = native.Factor.load(path)
    s1 = g.stats()
    for k in ("n", "partitions", "n_reduced", "method", "refine_steps", "rhs_tile", "reduced_sweeps"):
        assert s0[k] == s1[k], k
    X1 = g.solve(B, nrhs=3)
    assert np.array_equal(X0, X1)
    assert relerr(X1[0], orc.trsv(uplo, m.a, m.ia, m.ja, B[0])) <= TOL


@pytest.mark.parametrize("method,use_graph", [("levels", 1), ("spike", 1), ("spike", 2)])
def test_graph_replays_match_direct_launches(native, method, use_graph):
    """Repeated device-pointer solves: the level schedule (many launches) is
    captured into a hipGraph on the second call and replayed after that; the
    Spike solve (ten launches) keeps launching directly unless use_graph = 2.
    Every call, with eager device work and synchronisations between them,
    must give the bits of a solve made without the graph cache."""
    import torch

    m = gen.lap3d(32)
    B = gen.rhs(m.n, 3, seed=gen.SEED_BASE + 260)
    f0 = native.Factor("L", m.a, m.ia, m.ja, partitions=16, method=method, use_graph=False)
    want = f0.solve(B, nrhs=3)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=16, method=method, use_graph=use_graph)
    db = torch.from_numpy(B).cuda()
    dx = torch.zeros_like(db)
    s = torch.cuda.current_stream().cuda_stream
    junk = torch.ones(1 << 20, device="cuda", dtype=torch.float64)
    for call in range(5):
        dx.fill_(-1.0)
        f.solve(db.data_ptr(), dx.data_ptr(), nrhs=3, stream=s)
        junk.mul_(1.0001)
        torch.cuda.synchronize()
        assert np.array_equal(dx.cpu().numpy(), want), f"call {call}"
    for j in range(3):
        assert relerr(want[j], orc.ldsol(m.a, m.ia, m.ja, B[j])) <= TOL


def _parse_partition_log(text, nthreads):
    """The reference's log parser, restated: tools/plotSpeedUp.py:186-198
    finds the block between two marker lines by the marker string (with its
    trailing space), and :68-84 reads the fields, cutting each at its tab."""
    marker = "* * * * * * * * * * * * * * * * * * * * * * "
    block, inside = [], None
    for line in text.splitlines(keepends=True):
        if marker in line:
            if inside is None:
                inside = True
            else:
                break
        elif inside:
            block.append(line)
    assert inside, "no marker line"
    t = 0
    while "Matrix Information" not in block[t]:
        t += 1
    content = block[t + 2:]
    k = int(content[0][23:].split(" x ")[0])
    nnz = int(content[1].split(": ")[1])
    per = []
    for th in range(nthreads):
        s = 1 + th * 7
        field = lambda i: content[s + i].split("\t")[0].split(": ")[1]  # noqa: E731
        iv = content[s + 6].split(", ")
        per.append({"S": field(1), "isOptimized": field(2) == "1", "hasReflection": field(3) == "1",
                    "hasDependence": field(4) == "1", "hasToCalculateS": field(5) == "1",
                    "start": int(iv[0].split(": ")[-1]), "end": int(iv[1].split("\t")[0].split(": ")[1])})
    return k, nnz, per


def test_partition_log_parses_like_the_reference_tool(native, lap2d, capfd):
    """spike_factor_log prints the reference's logPSTRSV3 text byte for byte
    (ops.c:1317-1332: marker lines ending in a space, field lines ending in
    tabs), so the reference's parser reads it."""
    m = lap2d
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=8)
    capfd.readouterr()
    f.log()
    out = capfd.readouterr().out
    lines = out.splitlines(keepends=True)
    assert lines[0] == "* * * * * * * * * * * * * * * * * * * * * * \n"
    assert lines[1] == "* Matrix Information\t\t\t\t\t\t\n"
    assert lines[3].endswith("\t\t\t\t\n") and lines[4].endswith(" \t\t\t\t\t\n")
    st = f.stats()
    k, nnz, per = _parse_partition_log(out, st["partitions"])
    assert k == st["n_reduced"] and nnz == st["nnz_reduced"]
    assert len(per) == 8
    assert per[0]["hasDependence"] is False and all(p["hasDependence"] for p in per[1:])
    assert all(p["hasReflection"] for p in per[:-1]) and not per[-1]["hasReflection"]
    for p in per:
        r, c = (int(v) for v in p["S"].split(" x "))
        assert p["hasToCalculateS"] == (r > 0 and c > 0)


@pytest.mark.parametrize("nrhs,precision", [(1, "fp64"), (3, "fp64"), (8, "fp64"), (4, "fp32_refined")])
def test_jacobi_items_match_row_sweeps(native, monkeypatch, nrhs, precision):
    """Jacobi sweeps over items (one gather per distinct column into LDS) give
    the bits of the row-list sweeps, at every RHS count."""
    m = gen.banded(1 << 18, seed=gen.SEED_BASE + 43)
    B = gen.rhs(m.n, nrhs, seed=gen.SEED_BASE + 44)
    monkeypatch.setenv("SPIKE_PSTRSV_JACOBI_ITEMS", "0")
    f0 = native.Factor("L", m.a, m.ia, m.ja, partitions=256, reduced_solver="jacobi", precision=precision)
    monkeypatch.setenv("SPIKE_PSTRSV_JACOBI_ITEMS", "1")
    f1 = native.Factor("L", m.a, m.ia, m.ja, partitions=256, reduced_solver="jacobi", precision=precision)
    assert f1.stats()["reduced_sweeps"] > 0
    x0 = f0.solve(B, nrhs=nrhs)
    x1 = f1.solve(B, nrhs=nrhs)
    refs = [orc.ldsol(m.a, m.ia, m.ja, bj) for bj in B.reshape(nrhs, m.n)]
    err = max(relerr(xj, r) for xj, r in zip(x1.reshape(nrhs, m.n), refs))
    print(f"nrhs={nrhs} {precision} sweeps={f1.stats()['reduced_sweeps']} relerr {err:.3e}", flush=True)
    assert err <= TOL
    assert np.array_equal(x0, x1)


def test_autotuned_partitions_and_lanes(native):
    """autotune = 2 builds half and twice the automatic partition count and
    lanes per row, times them and keeps the fastest; the result solves to
    the bar."""
    m = gen.banded(1 << 18, seed=gen.SEED_BASE + 45)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 46)
    s0 = native.Factor("L", m.a, m.ia, m.ja).stats()
    f2 = native.Factor("L", m.a, m.ia, m.ja, autotune=2)
    s2 = f2.stats()
    print("rule", s0["partitions"], s0["lanes_per_row"], "tuned", s2["partitions"], s2["lanes_per_row"], flush=True)
    assert s2["partitions"] in (s0["partitions"] // 2, s0["partitions"], 2 * s0["partitions"])
    assert s2["lanes_per_row"] in (max(1, s0["lanes_per_row"] // 2), s0["lanes_per_row"],
                                   min(16, 2 * s0["lanes_per_row"]))
    x = f2.solve(b)
    assert relerr(x, orc.ldsol(m.a, m.ia, m.ja, b)) <= TOL


def test_allocation_failure_is_reported_and_releases_memory(native, monkeypatch):
    """A device allocation that fails during factor creation (forced from the
    host: SPIKE_PSTRSV_FAIL_ALLOC=k fails the factor's k-th) returns
    SPIKE_ERR_NOMEM, frees what the factor had allocated, and leaves the
    library usable."""
    import torch

    m = gen.banded(1 << 16, 64, 16, 0.02, gen.SEED_BASE + 90)
    native.Factor("L", m.a, m.ia, m.ja, partitions=64).close()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    failures = 0
    for k in range(1, 400, 3):
        monkeypatch.setenv("SPIKE_PSTRSV_FAIL_ALLOC", str(k))
        try:
            f = native.Factor("L", m.a, m.ia, m.ja, partitions=64)
        except native.SpikeError as e:
            assert e.code == 4, (k, str(e))
            failures += 1
            continue
        f.close()
        break                                   # k is past the last allocation
    monkeypatch.delenv("SPIKE_PSTRSV_FAIL_ALLOC")
    print("forced allocation failures:", failures)
    assert failures >= 5
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 < (64 << 20), (free0, free1)
    f = native.Factor("L", m.a, m.ia, m.ja, partitions=64)
    b = gen.rhs(m.n, 1, seed=gen.SEED_BASE + 91).ravel()
    assert relerr(np.asarray(f.solve(b)).ravel(), orc.ldsol(m.a, m.ia, m.ja, b)) <= TOL
    f.close()
