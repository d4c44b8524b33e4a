"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Imported only by tests/, the smoke check and the benchmark's CPU-baseline leg,
as the checker.  See oracle.h for what each function restates and how it is
pinned.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_PATH = Path(__file__).resolve().parent / "liboracle.so"
if not _PATH.exists():
    raise ImportError(f"{_PATH} is not built; run `make -C oracle` or spike_pstrsv_amd/build.py")
_o = C.CDLL(str(_PATH))
_vp, _i, _d, _c = C.c_void_p, C.c_int, C.c_double, C.c_char


class _Stats(C.Structure):
    _fields_ = [("n", _i), ("P", _i), ("nnz", C.c_int64), ("n_reduced", _i), ("nnz_reduced", C.c_int64),
                ("spike_area", C.c_int64), ("dropped_rowsum_max", _d)]


class PartInfo(C.Structure):
    _fields_ = [("start", _i), ("end", _i), ("rows_S", _i), ("cols_S", _i), ("first_coupling", _i),
                ("last_dependent", _i), ("has_dependence", _i), ("has_reflection", _i)]


for name in ("oracle_ldsol", "oracle_udsol", "oracle_spmv"):
    getattr(_o, name).argtypes = [_i, _vp, _vp, _vp, _vp, _vp]
    getattr(_o, name).restype = None
_o.oracle_partition.argtypes = [_i, _i, _vp]
_o.oracle_manufactured_x.argtypes = [_i, C.c_uint, _vp]
_o.oracle_flip.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _vp]
_o.oracle_spike_preproc.argtypes = [_c, _i, _vp, _vp, _vp, _i, _d, _i]
_o.oracle_spike_preproc.restype = _vp
_o.oracle_spike_solve.argtypes = [_vp, _vp, _vp]
_o.oracle_spike_stats.argtypes = [_vp, C.POINTER(_Stats)]
_o.oracle_spike_part_info.argtypes = [_vp, _i, C.POINTER(PartInfo)]
_o.oracle_spike_part_info.restype = _i
_o.oracle_spike_dependents.argtypes = [_vp, _vp]
_o.oracle_spike_kcols.argtypes = [_vp, _i, _vp]
_o.oracle_spike_kcols.restype = _i
_o.oracle_spike_block.argtypes = [_vp, _i, _vp]
_o.oracle_spike_block.restype = _i
_o.oracle_spike_reduced.argtypes = [_vp, _vp, _vp, _vp]
_o.oracle_spike_free.argtypes = [_vp]
_o.oracle_last_error.restype = C.c_char_p


def _arrs(a, ia, ja):
    return (np.ascontiguousarray(a, np.float64), np.ascontiguousarray(ia, np.int32),
            np.ascontiguousarray(ja, np.int32))


def ldsol(a, ia, ja, b):
    a, ia, ja = _arrs(a, ia, ja)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(ia) - 1)
    _o.oracle_ldsol(len(ia) - 1, a.ctypes.data, ia.ctypes.data, ja.ctypes.data, b.ctypes.data, x.ctypes.data)
    return x


def udsol(a, ia, ja, b):
    a, ia, ja = _arrs(a, ia, ja)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(ia) - 1)
    _o.oracle_udsol(len(ia) - 1, a.ctypes.data, ia.ctypes.data, ja.ctypes.data, b.ctypes.data, x.ctypes.data)
    return x


def trsv(uplo, a, ia, ja, b):
    return ldsol(a, ia, ja, b) if uplo.upper() == "L" else udsol(a, ia, ja, b)


def spmv(a, ia, ja, x):
    a, ia, ja = _arrs(a, ia, ja)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(len(ia) - 1)
    _o.oracle_spmv(len(ia) - 1, a.ctypes.data, ia.ctypes.data, ja.ctypes.data, x.ctypes.data, y.ctypes.data)
    return y


def partition(n, P):
    s = np.zeros(P + 1, np.int32)
    _o.oracle_partition(n, P, s.ctypes.data)
    return s


def manufactured_x(n, seed=1):
    x = np.zeros(n)
    _o.oracle_manufactured_x(n, seed, x.ctypes.data)
    return x


def flip(a, ia, ja):
    a, ia, ja = _arrs(a, ia, ja)
    n = len(ia) - 1
    a2 = np.zeros_like(a); ia2 = np.zeros_like(ia); ja2 = np.zeros_like(ja)
    _o.oracle_flip(n, a.ctypes.data, ia.ctypes.data, ja.ctypes.data, a2.ctypes.data, ia2.ctypes.data, ja2.ctypes.data)
    return a2, ia2, ja2


class Spike:
    """CPU Spike restatement (preprocess once, solve many)."""

    def __init__(self, uplo, a, ia, ja, P, tau=0.0, nthreads=0):
        self._keep = _arrs(a, ia, ja)
        a, ia, ja = self._keep
        self.n = len(ia) - 1
        self.P = P
        h = _o.oracle_spike_preproc(uplo.encode()[:1], self.n, a.ctypes.data, ia.ctypes.data, ja.ctypes.data, P,
                                    tau, nthreads)
        if not h:
            raise ValueError(_o.oracle_last_error().decode())
        self.h = h

    def solve(self, b, x=None):
        b = np.ascontiguousarray(b, np.float64)
        if x is None:
            x = np.zeros(self.n)
        _o.oracle_spike_solve(self.h, b.ctypes.data, x.ctypes.data)
        return x

    def stats(self):
        st = _Stats()
        _o.oracle_spike_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def part(self, p):
        pi = PartInfo()
        _o.oracle_spike_part_info(self.h, p, C.byref(pi))
        return {k: getattr(pi, k) for k, _ in pi._fields_}

    def dependents(self):
        d = np.zeros(max(1, self.stats()["n_reduced"]), np.int32)
        _o.oracle_spike_dependents(self.h, d.ctypes.data)
        return d[: self.stats()["n_reduced"]]

    def kcols(self, p):
        m = self.part(p)["cols_S"]
        c = np.zeros(max(1, self.n), np.int32)
        k = _o.oracle_spike_kcols(self.h, p, c.ctypes.data)
        return c[:k]

    def block(self, p):
        info = self.part(p)
        w = np.zeros(max(1, info["rows_S"] * info["cols_S"]))
        m = _o.oracle_spike_block(self.h, p, w.ctypes.data)
        return w[:m].reshape(info["rows_S"], info["cols_S"]) if m else np.zeros((0, 0))

    def reduced(self):
        st = self.stats()
        nr = st["n_reduced"]
        nnz = st["nnz_reduced"] - nr
        rp = np.zeros(nr + 1, np.int32); ci = np.zeros(max(1, nnz), np.int32); v = np.zeros(max(1, nnz))
        _o.oracle_spike_reduced(self.h, rp.ctypes.data, ci.ctypes.data, v.ctypes.data)
        return rp, ci[:nnz], v[:nnz]

    def close(self):
        if self.h:
            _o.oracle_spike_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
