"""ctypes binding of libspike_pstrsv.so (the C ABI in include/spike_pstrsv.h).

Mirrors the reference's C interface (src/spike_pstrsv.h:7-38: spike_pstrsv,
spike_pstrsv_preproc, spike_pstrsv_solve and the profiling twins) plus the
extension API (factor handles, multi-RHS, device pointers, stats).  Arrays are
numpy arrays (host) or integer device addresses (e.g. torch.Tensor.data_ptr()).
There is no fallback: if the native library is missing, importing this module
raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libspike_pstrsv.so"

SPIKE_OK = 0
ERRORS = {1: "SPIKE_ERR_ARG", 2: "SPIKE_ERR_MATRIX", 3: "SPIKE_ERR_DEVICE", 4: "SPIKE_ERR_NOMEM",
          5: "SPIKE_ERR_STATE", 6: "SPIKE_ERR_UNSUPPORTED"}

# Every symbol declared in include/spike_pstrsv.h.
EXPORTED = [
    "spike_pstrsv_preproc", "_spike_pstrsv_preproc", "spike_pstrsv", "spike_pstrsv_solve", "_spike_pstrsv_solve",
    "spike_options_default", "spike_factor_create", "spike_factor_create_partitioned", "spike_factor_solve",
    "spike_factor_solve_profiled", "spike_factor_status", "spike_factor_solve_noalloc",
    "spike_factor_destroy",
    "spike_factor_stats", "spike_factor_log", "spike_analyze", "spike_pstrsv_set_partitions",
    "spike_pstrsv_current_factor", "spike_last_error", "spike_pstrsv_get_timers",
    "spike_batch_options_default", "spike_batch_create", "spike_batch_offset", "spike_batch_length",
    "spike_batch_solve", "spike_batch_stats", "spike_batch_destroy",
    "spike_server_options_default", "spike_server_create", "spike_server_add", "spike_server_remove",
    "spike_server_submit",
    "spike_server_wait", "spike_server_drain", "spike_server_stats", "spike_server_reset_stats",
    "spike_server_destroy", "spike_server_last_error",
    "spike_read_matrix_market", "spike_free", "spike_mm_last_error",
    "spike_factor_save", "spike_factor_load",
]


class Options(C.Structure):
    _fields_ = [("partitions", C.c_int), ("drop_tol", C.c_double), ("precision", C.c_int),
                ("lanes_per_row", C.c_int), ("verbose", C.c_int), ("reduced_solver", C.c_int),
                ("refine_steps", C.c_int), ("refine_tol", C.c_double), ("method", C.c_int),
                ("use_graph", C.c_int), ("autotune", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("n", C.c_int), ("partitions", C.c_int), ("nnz", C.c_int64), ("n_reduced", C.c_int),
                ("nnz_reduced", C.c_int64), ("reduced_levels", C.c_int), ("spike_area", C.c_int64),
                ("dropped_rowsum_max", C.c_double), ("steps_pass1", C.c_int), ("steps_pass2", C.c_int),
                ("lanes_per_row", C.c_int), ("device_bytes", C.c_int64), ("preproc_seconds", C.c_double),
                ("reduced_sweeps", C.c_int), ("jacobi_bound", C.c_double),
                ("rhs_per_launch", C.c_int), ("precision", C.c_int), ("refine_steps", C.c_int),
                ("refine_rate", C.c_double), ("refine_est_err", C.c_double), ("method", C.c_int),
                ("levels", C.c_int), ("rhs_tile", C.c_int * 4), ("single_launch", C.c_int)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["rhs_tile"] = list(self.rhs_tile)
        return d


class BatchMatrix(C.Structure):
    _fields_ = [("n", C.c_int), ("a", C.c_void_p), ("ia", C.c_void_p), ("ja", C.c_void_p), ("nrhs", C.c_int)]


class BatchOptions(C.Structure):
    _fields_ = [("rows_per_partition", C.c_int), ("lanes_per_row", C.c_int), ("precision", C.c_int),
                ("use_graph", C.c_int), ("autotune_sample", C.c_int), ("drop_tol", C.c_double),
                ("refine_tol", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("count", C.c_int), ("groups", C.c_int), ("rows", C.c_int64), ("nnz", C.c_int64),
                ("rows_rhs", C.c_int64), ("partitions", C.c_int64), ("n_reduced", C.c_int64),
                ("rows_per_partition", C.c_int * 4), ("refine_steps", C.c_int), ("refine_est_err", C.c_double),
                ("device_bytes", C.c_int64), ("preproc_seconds", C.c_double), ("autotune_seconds", C.c_double),
                ("graph_nodes", C.c_int), ("lane_shift", C.c_int * 4)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["rows_per_partition"] = list(self.rows_per_partition)
        d["lane_shift"] = list(self.lane_shift)
        return d


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("slots", C.c_int), ("streams", C.c_int), ("window_us", C.c_double),
                ("max_batch_rhs", C.c_int), ("admit_timeout_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("matrices", C.c_int), ("resident_bytes", C.c_int64), ("completed", C.c_int64),
                ("solves", C.c_int64), ("windows", C.c_int64), ("merged_rhs", C.c_int64),
                ("max_inflight", C.c_int), ("rows_rhs", C.c_int64), ("p50_us", C.c_double), ("p99_us", C.c_double),
                ("max_us", C.c_double), ("seconds", C.c_double), ("rows_rhs_per_s", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SpikeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built; run spike_pstrsv_amd/build.py (or __graft_entry__.build())")

lib = C.CDLL(str(LIB_PATH))
_vp, _i, _d, _c = C.c_void_p, C.c_int, C.c_double, C.c_char
lib.spike_pstrsv_preproc.argtypes = [_vp, _vp, _vp, _i, _i, _c]
lib._spike_pstrsv_preproc.argtypes = [_vp, _vp, _vp, _i, _i, _c]
lib.spike_pstrsv.argtypes = [_c, _i, _vp, _vp, _vp, _vp, _vp, _i]
lib.spike_pstrsv_solve.argtypes = [_c, _i, _vp, _vp, _i]
lib._spike_pstrsv_solve.argtypes = [_c, _i, _vp, _vp, _i]
for _f in ("spike_pstrsv_preproc", "_spike_pstrsv_preproc", "spike_pstrsv", "spike_pstrsv_solve",
           "_spike_pstrsv_solve"):
    getattr(lib, _f).restype = None
lib.spike_options_default.argtypes = [C.POINTER(Options)]
lib.spike_options_default.restype = None
lib.spike_factor_create.argtypes = [C.POINTER(_vp), _c, _i, _vp, _vp, _vp, C.POINTER(Options)]
lib.spike_factor_create.restype = _i
lib.spike_factor_create_partitioned.argtypes = [C.POINTER(_vp), _c, _i, _vp, _vp, _vp, _i, _vp, C.POINTER(Options)]
lib.spike_factor_create_partitioned.restype = _i
lib.spike_factor_solve.argtypes = [_vp, _i, _vp, _i, _vp, _i, _vp]
lib.spike_factor_solve.restype = _i
lib.spike_factor_solve_profiled.argtypes = [_vp, _i, _vp, _i, _vp, _i, _vp, C.POINTER(_d)]
lib.spike_factor_solve_profiled.restype = _i
lib.spike_factor_solve_noalloc.argtypes = [_vp, _i, _vp, _i, _vp, _i, _vp]
lib.spike_factor_solve_noalloc.restype = _i
lib.spike_factor_status.argtypes = [_vp]
lib.spike_factor_status.restype = _i
lib.spike_factor_destroy.argtypes = [_vp]
lib.spike_factor_destroy.restype = _i
lib.spike_factor_stats.argtypes = [_vp, C.POINTER(Stats)]
lib.spike_factor_stats.restype = _i
lib.spike_factor_log.argtypes = [_vp]
lib.spike_factor_log.restype = _i
lib.spike_analyze.argtypes = [_c, _i, _vp, _vp, _vp, C.POINTER(Options), C.POINTER(Stats)]
lib.spike_analyze.restype = _i
lib.spike_pstrsv_set_partitions.argtypes = [_i]
lib.spike_pstrsv_set_partitions.restype = None
lib.spike_pstrsv_current_factor.argtypes = []
lib.spike_pstrsv_current_factor.restype = _vp
lib.spike_last_error.argtypes = []
lib.spike_last_error.restype = C.c_char_p
lib.spike_pstrsv_get_timers.argtypes = [C.POINTER(_d), _i]
lib.spike_batch_options_default.argtypes = [C.POINTER(BatchOptions)]
lib.spike_batch_options_default.restype = None
lib.spike_batch_create.argtypes = [C.POINTER(_vp), _c, _i, C.POINTER(BatchMatrix), C.POINTER(BatchOptions)]
lib.spike_batch_create.restype = _i
lib.spike_batch_offset.argtypes = [_vp, _i]
lib.spike_batch_offset.restype = C.c_int64
lib.spike_batch_length.argtypes = [_vp]
lib.spike_batch_length.restype = C.c_int64
lib.spike_batch_solve.argtypes = [_vp, _vp, _vp, _vp]
lib.spike_batch_solve.restype = _i
lib.spike_batch_stats.argtypes = [_vp, C.POINTER(BatchStats)]
lib.spike_batch_stats.restype = _i
lib.spike_batch_destroy.argtypes = [_vp]
lib.spike_batch_destroy.restype = _i
lib.spike_server_options_default.argtypes = [C.POINTER(ServerOptions)]
lib.spike_server_options_default.restype = None
lib.spike_server_create.argtypes = [C.POINTER(_vp), C.POINTER(ServerOptions)]
lib.spike_server_create.restype = _i
lib.spike_server_add.argtypes = [_vp, _c, _i, _vp, _vp, _vp, C.POINTER(Options), C.POINTER(_i)]
lib.spike_server_add.restype = _i
lib.spike_server_remove.argtypes = [_vp, _i]
lib.spike_server_remove.restype = _i
lib.spike_server_submit.argtypes = [_vp, _i, _i, _vp, _i, _vp, _i, C.POINTER(C.c_uint64)]
lib.spike_server_submit.restype = _i
lib.spike_server_wait.argtypes = [_vp, C.c_uint64]
lib.spike_server_wait.restype = _i
lib.spike_server_drain.argtypes = [_vp]
lib.spike_server_drain.restype = _i
lib.spike_server_stats.argtypes = [_vp, C.POINTER(ServerStats)]
lib.spike_server_stats.restype = _i
lib.spike_server_reset_stats.argtypes = [_vp]
lib.spike_server_reset_stats.restype = _i
lib.spike_server_destroy.argtypes = [_vp]
lib.spike_server_destroy.restype = _i
lib.spike_server_last_error.argtypes = []
lib.spike_server_last_error.restype = C.c_char_p
lib.spike_read_matrix_market.argtypes = [C.c_char_p, _c, _c, C.POINTER(_i), C.POINTER(C.POINTER(_d)),
                                         C.POINTER(C.POINTER(_i)), C.POINTER(C.POINTER(_i))]
lib.spike_read_matrix_market.restype = _i
lib.spike_factor_save.argtypes = [_vp, C.c_char_p]
lib.spike_factor_save.restype = _i
lib.spike_factor_load.argtypes = [C.POINTER(_vp), C.c_char_p]
lib.spike_factor_load.restype = _i
lib.spike_free.argtypes = [_vp]
lib.spike_free.restype = None
lib.spike_mm_last_error.argtypes = []
lib.spike_mm_last_error.restype = C.c_char_p
lib.spike_pstrsv_get_timers.restype = _i


def last_error() -> str:
    return lib.spike_last_error().decode()


def _check(rc: int) -> None:
    if rc != SPIKE_OK:
        raise SpikeError(rc, last_error())


def _ptr(a) -> int:
    """Address of a numpy array (host) or an int device address."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    raise TypeError(f"cannot take the address of {type(a)}")


def _csr(a, ia, ja):
    a = np.ascontiguousarray(a, dtype=np.float64)
    ia = np.ascontiguousarray(ia, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    return a, ia, ja


REDUCED_SOLVERS = {"auto": 0, "exact": 1, "jacobi": 2}
PRECISIONS = {"fp64": 0, "fp32_refined": 1}
METHODS = {"auto": 0, "spike": 1, "levels": 2}


def options(partitions: int = 0, drop_tol: float = -1.0, lanes_per_row: int = 0, verbose: bool = False,
            reduced_solver: str = "auto", precision: str = "fp64", refine_steps: int = -1,
            refine_tol: float = 0.0, method: str = "auto", use_graph: int = 1, autotune: int = 1) -> Options:
    o = Options()
    lib.spike_options_default(C.byref(o))
    o.partitions = partitions
    o.drop_tol = drop_tol
    o.lanes_per_row = lanes_per_row
    o.verbose = int(verbose)
    o.reduced_solver = REDUCED_SOLVERS[reduced_solver]
    o.precision = PRECISIONS[precision]
    o.refine_steps = refine_steps
    o.refine_tol = refine_tol
    o.method = METHODS[method]
    o.use_graph = int(use_graph)
    o.autotune = int(autotune)
    return o


class Factor:
    """A preprocessed matrix resident in device memory (spike_factor_t)."""

    def __init__(self, uplo: str, a, ia, ja, partitions: int = 0, drop_tol: float = -1.0,
                 lanes_per_row: int = 0, verbose: bool = False, reduced_solver: str = "auto",
                 precision: str = "fp64", refine_steps: int = -1, refine_tol: float = 0.0,
                 method: str = "auto", starts=None, use_graph: int = 1, autotune: int = 1):
        a, ia, ja = _csr(a, ia, ja)
        self.n = len(ia) - 1
        h = _vp()
        o = options(partitions, drop_tol, lanes_per_row, verbose, reduced_solver, precision, refine_steps,
                    refine_tol, method, use_graph, autotune)
        if starts is None:
            _check(lib.spike_factor_create(C.byref(h), uplo.encode()[:1], self.n, a.ctypes.data, ia.ctypes.data,
                                           ja.ctypes.data, C.byref(o)))
        else:
            st = np.ascontiguousarray(starts, dtype=np.int32)
            _check(lib.spike_factor_create_partitioned(C.byref(h), uplo.encode()[:1], self.n, a.ctypes.data,
                                                       ia.ctypes.data, ja.ctypes.data, len(st) - 1, st.ctypes.data,
                                                       C.byref(o)))
        self.handle = h.value

    def solve(self, b, x=None, nrhs: int = 1, ldb: int | None = None, ldx: int | None = None, stream: int = 0):
        """Solves A X = B.  b/x: numpy arrays (column-major for nrhs > 1) or device addresses."""
        n = self.n
        ldb = n if ldb is None else ldb
        ldx = n if ldx is None else ldx
        if x is None:
            x = np.zeros((nrhs, ldx) if nrhs > 1 else ldx, dtype=np.float64)
        _check(lib.spike_factor_solve(self.handle, nrhs, _ptr(b), ldb, _ptr(x), ldx, stream or None))
        return x

    def solve_noalloc(self, b, x, nrhs: int = 1, ldb: int | None = None, ldx: int | None = None,
                      stream: int = 0) -> None:
        """spike_factor_solve_noalloc: device pointers only, workspaces from the library's static scratch."""
        n = self.n
        _check(lib.spike_factor_solve_noalloc(self.handle, nrhs, _ptr(b), n if ldb is None else ldb, _ptr(x),
                                              n if ldx is None else ldx, stream or None))

    def solve_profiled(self, b, x, nrhs: int = 1, ldb: int | None = None, ldx: int | None = None,
                       stream: int = 0) -> list[float]:
        """Solve and return the device time (ms) of pass 1, reduced solve, coupling, pass 2."""
        ms = (C.c_double * 4)()
        n = self.n
        _check(lib.spike_factor_solve_profiled(self.handle, nrhs, _ptr(b), n if ldb is None else ldb, _ptr(x),
                                               n if ldx is None else ldx, stream or None, ms))
        return list(ms)

    def status(self) -> None:
        """Raises SpikeError (SPIKE_ERR_DEVICE) if an earlier solve gave up a bounded device wait."""
        _check(lib.spike_factor_status(self.handle))

    def save(self, path) -> None:
        _check(lib.spike_factor_save(self.handle, str(path).encode()))

    @classmethod
    def load(cls, path) -> "Factor":
        h = _vp()
        _check(lib.spike_factor_load(C.byref(h), str(path).encode()))
        f = cls.__new__(cls)
        f.handle = h.value
        st = Stats()
        _check(lib.spike_factor_stats(f.handle, C.byref(st)))
        f.n = st.n
        return f

    def stats(self) -> dict:
        st = Stats()
        _check(lib.spike_factor_stats(self.handle, C.byref(st)))
        return st.as_dict()

    def log(self) -> None:
        _check(lib.spike_factor_log(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.spike_factor_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def analyze(uplo: str, a, ia, ja, partitions: int = 0, lanes_per_row: int = 0, precision: str = "fp64") -> dict:
    """Host-only symbolic analysis (no device needed)."""
    a, ia, ja = _csr(a, ia, ja)
    st = Stats()
    o = options(partitions, -1.0, lanes_per_row, precision=precision)
    _check(lib.spike_analyze(uplo.encode()[:1], len(ia) - 1, a.ctypes.data, ia.ctypes.data, ja.ctypes.data,
                             C.byref(o), C.byref(st)))
    return st.as_dict()


# ---- legacy entry points (reference src/spike_pstrsv.h) -------------------

def spike_pstrsv(uplo: str, m: int, a, ia, ja, b, x, nthreads: int) -> None:
    lib.spike_pstrsv(uplo.encode()[:1], m, _ptr(a), _ptr(ia), _ptr(ja), _ptr(b), _ptr(x), nthreads)


def spike_pstrsv_preproc(a, ia, ja, n: int, nthreads: int, uplo: str, profile: bool = False) -> None:
    f = lib._spike_pstrsv_preproc if profile else lib.spike_pstrsv_preproc
    f(_ptr(a), _ptr(ia), _ptr(ja), n, nthreads, uplo.encode()[:1])


def spike_pstrsv_solve(uplo: str, m: int, b, x, nthreads: int, profile: bool = False) -> None:
    f = lib._spike_pstrsv_solve if profile else lib.spike_pstrsv_solve
    f(uplo.encode()[:1], m, _ptr(b), _ptr(x), nthreads)


def set_partitions(P: int) -> None:
    lib.spike_pstrsv_set_partitions(P)


def timers() -> list[float]:
    buf = (C.c_double * 9)()
    k = lib.spike_pstrsv_get_timers(buf, 9)
    return list(buf[:k])


def current_stats() -> dict | None:
    h = lib.spike_pstrsv_current_factor()
    if not h:
        return None
    st = Stats()
    _check(lib.spike_factor_stats(h, C.byref(st)))
    return st.as_dict()


class Batch:
    """Many independent matrices preprocessed together (spike_batch_t).

    mats: sequence of (a, ia, ja, nrhs).  Right-hand sides and solutions use
    the packed layout: matrix i's nrhs columns (leading dimension n_i) start
    at element offsets[i]."""

    def __init__(self, uplo: str, mats, rows_per_partition: int = 0, lanes_per_row: int = 0,
                 precision: str = "fp64", use_graph: bool = True, autotune_sample: int = 0,
                 drop_tol: float = -1.0, refine_tol: float = 0.0):
        self._keep = []
        arr = (BatchMatrix * len(mats))()
        for i, (a, ia, ja, nrhs) in enumerate(mats):
            a, ia, ja = _csr(a, ia, ja)
            self._keep.append((a, ia, ja))
            arr[i] = BatchMatrix(len(ia) - 1, a.ctypes.data, ia.ctypes.data, ja.ctypes.data, int(nrhs))
        o = BatchOptions()
        lib.spike_batch_options_default(C.byref(o))
        o.rows_per_partition = rows_per_partition
        o.lanes_per_row = lanes_per_row
        o.precision = PRECISIONS[precision]
        o.use_graph = int(use_graph)
        o.autotune_sample = autotune_sample
        o.drop_tol = drop_tol
        o.refine_tol = refine_tol
        h = _vp()
        _check(lib.spike_batch_create(C.byref(h), uplo.encode()[:1], len(mats), arr, C.byref(o)))
        self._keep = None          # the batch holds its own copies
        self.handle = h.value
        self.count = len(mats)
        self.length = int(lib.spike_batch_length(self.handle))
        self.offsets = [int(lib.spike_batch_offset(self.handle, i)) for i in range(self.count + 1)]

    def solve(self, b, x=None, stream: int = 0):
        """b, x: packed float64 arrays (host) or device addresses."""
        if x is None:
            x = np.zeros(self.length)
        _check(lib.spike_batch_solve(self.handle, _ptr(b), _ptr(x), stream or None))
        return x

    def stats(self) -> dict:
        st = BatchStats()
        _check(lib.spike_batch_stats(self.handle, C.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.spike_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _scheck(rc: int) -> None:
    if rc != SPIKE_OK:
        raise SpikeError(rc, (lib.spike_server_last_error() or b"").decode())


class Server:
    """Resident factors plus a batching dispatcher (spike_server_t).

    Requests carry device pointers (int addresses or tensors with data_ptr())."""

    def __init__(self, hbm_budget: int = 0, slots: int = 64, streams: int = 8, window_us: float = 200.0,
                 max_batch_rhs: int = 16, admit_timeout_ms: float = -1.0):
        o = ServerOptions()
        lib.spike_server_options_default(C.byref(o))
        o.hbm_budget = hbm_budget
        o.slots = slots
        o.streams = streams
        o.window_us = window_us
        o.max_batch_rhs = max_batch_rhs
        o.admit_timeout_ms = admit_timeout_ms
        h = _vp()
        _scheck(lib.spike_server_create(C.byref(h), C.byref(o)))
        self.handle = h.value
        self.n = []

    def add(self, uplo: str, a, ia, ja, **kw) -> int:
        a, ia, ja = _csr(a, ia, ja)
        o = options(**kw)
        mid = C.c_int(-1)
        _scheck(lib.spike_server_add(self.handle, uplo.encode()[:1], len(ia) - 1, a.ctypes.data, ia.ctypes.data,
                                     ja.ctypes.data, C.byref(o), C.byref(mid)))
        while len(self.n) <= mid.value:
            self.n.append(0)
        self.n[mid.value] = len(ia) - 1
        return mid.value

    def remove(self, mid: int) -> None:
        """Stops serving matrix mid once its requests complete and frees its memory."""
        _scheck(lib.spike_server_remove(self.handle, mid))

    def submit(self, mid: int, nrhs: int, b, x, ldb: int | None = None, ldx: int | None = None) -> int:
        n = self.n[mid]
        t = C.c_uint64(0)
        _scheck(lib.spike_server_submit(self.handle, mid, nrhs, _ptr(b), n if ldb is None else ldb, _ptr(x),
                                        n if ldx is None else ldx, C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> None:
        _scheck(lib.spike_server_wait(self.handle, ticket))

    def drain(self) -> None:
        _scheck(lib.spike_server_drain(self.handle))

    def stats(self) -> dict:
        st = ServerStats()
        _scheck(lib.spike_server_stats(self.handle, C.byref(st)))
        return st.as_dict()

    def reset_stats(self) -> None:
        _scheck(lib.spike_server_reset_stats(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.spike_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_matrix_market(path, uplo: str = "L", symm: str = "A"):
    """CSR (a, ia, ja, n) of a MatrixMarket file (spike_read_matrix_market)."""
    n = C.c_int(0)
    pa = C.POINTER(_d)()
    pia = C.POINTER(_i)()
    pja = C.POINTER(_i)()
    rc = lib.spike_read_matrix_market(str(path).encode(), uplo.encode()[:1], symm.encode()[:1], C.byref(n),
                                      C.byref(pa), C.byref(pia), C.byref(pja))
    if rc != SPIKE_OK:
        raise SpikeError(rc, (lib.spike_mm_last_error() or b"").decode())
    try:
        nn = n.value
        ia = np.ctypeslib.as_array(pia, shape=(nn + 1,)).copy()
        nnz = int(ia[-1])
        ja = np.ctypeslib.as_array(pja, shape=(max(nnz, 1),))[:nnz].copy()
        a = np.ctypeslib.as_array(pa, shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        lib.spike_free(C.cast(pa, C.c_void_p))
        lib.spike_free(C.cast(pia, C.c_void_p))
        lib.spike_free(C.cast(pja, C.c_void_p))
    return a, ia, ja, nn
