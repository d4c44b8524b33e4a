"""Deterministic synthetic workloads (binding of libspike_gen.so, csrc/gen.c).

All matrices are lower-triangular CSR (diagonal last) with int32 indices and
float64 values; `to_upper` gives the index-reversed upper-triangular form
(diagonal first) of the same system.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libspike_gen.so"
if not _LIB_PATH.exists():
    raise ImportError(f"{_LIB_PATH} is not built; run spike_pstrsv_amd/build.py")
_g = C.CDLL(str(_LIB_PATH))
_vp, _i64, _u64, _d, _i = C.c_void_p, C.c_int64, C.c_uint64, C.c_double, C.c_int
_g.spike_gen_lap2d_nnz.argtypes = [_i]; _g.spike_gen_lap2d_nnz.restype = _i64
_g.spike_gen_lap2d.argtypes = [_i, _vp, _vp, _vp]; _g.spike_gen_lap2d.restype = _i64
_g.spike_gen_lap3d_nnz.argtypes = [_i]; _g.spike_gen_lap3d_nnz.restype = _i64
_g.spike_gen_lap3d.argtypes = [_i, _vp, _vp, _vp]; _g.spike_gen_lap3d.restype = _i64
_g.spike_gen_band_count.argtypes = [_i64, _i, _i, _d, _u64, _vp]; _g.spike_gen_band_count.restype = _i64
_g.spike_gen_band_fill.argtypes = [_i64, _i, _i, _d, _u64, _vp, _vp, _vp, _vp]; _g.spike_gen_band_fill.restype = _i64
_g.spike_gen_rhs.argtypes = [_i64, _i, _i64, _u64, _vp]; _g.spike_gen_rhs.restype = None
_g.spike_gen_params.argtypes = [_u64, _u64, _i64, _vp]; _g.spike_gen_params.restype = None

# seed family: 0x5350494B45000000 + workload number
SEED_BASE = 0x5350494B45000000


@dataclass
class CSR:
    a: np.ndarray
    ia: np.ndarray
    ja: np.ndarray
    uplo: str = "L"

    @property
    def n(self) -> int:
        return len(self.ia) - 1

    @property
    def nnz(self) -> int:
        return int(self.ia[-1])


def lap2d(g: int) -> CSR:
    n = g * g
    nnz = _g.spike_gen_lap2d_nnz(g)
    ia = np.empty(n + 1, np.int32); ja = np.empty(nnz, np.int32); a = np.empty(nnz, np.float64)
    _g.spike_gen_lap2d(g, ia.ctypes.data, ja.ctypes.data, a.ctypes.data)
    return CSR(a, ia, ja)


def lap3d(g: int) -> CSR:
    n = g * g * g
    nnz = _g.spike_gen_lap3d_nnz(g)
    ia = np.empty(n + 1, np.int32); ja = np.empty(nnz, np.int32); a = np.empty(nnz, np.float64)
    _g.spike_gen_lap3d(g, ia.ctypes.data, ja.ctypes.data, a.ctypes.data)
    return CSR(a, ia, ja)


def banded(n: int, bw: int = 64, q: int = 16, plr: float = 0.02, seed: int = SEED_BASE + 2) -> CSR:
    rowlen = np.empty(n, np.int32)
    nnz = _g.spike_gen_band_count(n, bw, q, plr, seed, rowlen.ctypes.data)
    if nnz < 0:
        raise ValueError("invalid banded parameters")
    ia = np.empty(n + 1, np.int32); ja = np.empty(nnz, np.int32); a = np.empty(nnz, np.float64)
    _g.spike_gen_band_fill(n, bw, q, plr, seed, rowlen.ctypes.data, ia.ctypes.data, ja.ctypes.data, a.ctypes.data)
    return CSR(a, ia, ja)


def rhs(n: int, nrhs: int = 1, seed: int = SEED_BASE + 100) -> np.ndarray:
    """Column-major right-hand sides: shape (nrhs, n) in memory order, entries U(-1,1)."""
    b = np.empty((nrhs, n), np.float64)
    _g.spike_gen_rhs(n, nrhs, n, seed, b.ctypes.data)
    return b if nrhs > 1 else b[0]


def params(seed: int, stream: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float64)
    _g.spike_gen_params(seed, stream, count, out.ctypes.data)
    return out


def batch_specs(count: int, n_lo: int = 1000, n_hi: int = 65536, seed: int = SEED_BASE + 4) -> list[dict]:
    """Per-matrix parameters of a batch of banded matrices: n log-uniform in
    [n_lo, n_hi], bandwidth uniform in {8..128}, off-diagonals per row uniform
    in {4..24} (at most the bandwidth), 1-8 right-hand sides; each drawn from
    the matrix's own counter stream."""
    out = []
    lo, hi = np.log(n_lo), np.log(n_hi)
    for i in range(count):
        u = params(seed, i, 4)
        n = int(np.floor(np.exp(lo + (hi - lo) * u[0])))
        n = min(max(n, n_lo), n_hi)
        bw = 8 + min(int(u[1] * 121), 120)
        q = min(4 + min(int(u[2] * 21), 20), bw)
        nrhs = 1 + min(int(u[3] * 8), 7)
        out.append({"n": n, "bw": bw, "q": q, "nrhs": nrhs, "seed": seed + 0x1000 * (i + 1)})
    return out


def batched(count: int, n_lo: int = 1000, n_hi: int = 65536, seed: int = SEED_BASE + 4):
    """(specs, matrices): the banded matrices of batch_specs (2% long-range rows)."""
    specs = batch_specs(count, n_lo, n_hi, seed)
    mats = [banded(s["n"], s["bw"], s["q"], 0.02, s["seed"]) for s in specs]
    return specs, mats


def serving_specs(count: int = 64, n_lo: int = 10_000, n_hi: int = 2_000_000, seed: int = SEED_BASE + 5) -> list[dict]:
    """Resident matrices of the serving workload: n log-uniform in [n_lo, n_hi];
    even ids are Laplacian lower triangles (alternately 2D with side
    round(sqrt(n)) and 3D with side round(cbrt(n))), odd ids random banded
    matrices with long-range rows as in the 1M-row banded workload."""
    out = []
    lo, hi = np.log(n_lo), np.log(n_hi)
    for i in range(count):
        u = params(seed, i, 1)
        n = int(np.floor(np.exp(lo + (hi - lo) * u[0])))
        if i % 2 == 0:
            if (i // 2) % 2 == 0:
                side = max(2, int(round(n ** 0.5)))
                out.append({"kind": "lap2d", "side": side, "n": side * side})
            else:
                side = max(2, int(round(n ** (1.0 / 3.0))))
                out.append({"kind": "lap3d", "side": side, "n": side ** 3})
        else:
            out.append({"kind": "banded", "n": n, "seed": seed + 0x1000 * (i + 1)})
    return out


def serving_matrix(spec: dict) -> CSR:
    if spec["kind"] == "lap2d":
        return lap2d(spec["side"])
    if spec["kind"] == "lap3d":
        return lap3d(spec["side"])
    return banded(spec["n"], 64, 16, 0.02, spec["seed"])


def to_upper(m: CSR) -> CSR:
    """Index reversal U = J L J: the upper-triangular system with U (J x) = J b."""
    n = m.n
    lens = np.diff(m.ia)[::-1]
    ia = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=ia[1:])
    # rows reversed, entries within each row reversed == the whole entry array reversed
    ja = (n - 1 - m.ja[::-1]).astype(np.int32)
    a = m.a[::-1].copy()
    return CSR(a, ia, ja, "U")


def spmv(m: CSR, x: np.ndarray) -> np.ndarray:
    rows = np.repeat(np.arange(m.n), np.diff(m.ia))
    y = np.zeros(m.n)
    np.add.at(y, rows, m.a * x[m.ja])
    return y


def workload(name: str) -> tuple[CSR, int]:
    """Named workloads with their partition counts."""
    if name == "lap2d_256":
        return lap2d(256), 8
    if name == "banded_1m":
        return banded(1 << 20), 1024
    if name == "lap3d_128":
        return lap3d(128), 0
    raise KeyError(name)
