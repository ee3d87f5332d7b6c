"""Python binding of libperman_hip.so (ctypes over the C ABI in
include/perman_hip.h) plus the seeded synthetic-matrix generator used by the
benchmarks and tests.

The library is the product: every function here calls into the HIP/C++
code and raises if the shared object is missing -- there is no Python or
CPU fallback for the GPU entry points.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libperman_hip.so")
CLI_PATH = os.path.join(HERE, "bin", "perman")

DTYPE_I32, DTYPE_F32, DTYPE_F64 = 0, 1, 2
_DT = {np.dtype(np.int32): DTYPE_I32, np.dtype(np.float32): DTYPE_F32,
       np.dtype(np.float64): DTYPE_F64}
_SFX = {DTYPE_I32: "i32", DTYPE_F32: "f32", DTYPE_F64: "f64"}

DENSE_VARIANTS = ("xglobal", "xlocal", "xshared", "xshared_coalescing",
                  "xshared_coalescing_mshared")          # algorithm IDs 0..4
SPARSE_VARIANTS = (None, "xlocal", "xshared", "xshared_coalescing",
                   "xshared_coalescing_mshared")         # algorithm IDs 1..4


class PermError(RuntimeError):
    pass


class PermResultC(ctypes.Structure):
    _fields_ = [("value", ctypes.c_double), ("sum_hi", ctypes.c_double),
                ("sum_lo", ctypes.c_double), ("abs_sum", ctypes.c_double),
                ("err_bound", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("subsets", ctypes.c_double), ("n", ctypes.c_int), ("status", ctypes.c_int),
                ("mode", ctypes.c_int), ("integral", ctypes.c_int), ("err_est", ctypes.c_double),
                ("nu", ctypes.c_int), ("main_ms", ctypes.c_double),
                ("main_subsets", ctypes.c_double), ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
                ("certified", ctypes.c_int), ("zero_pattern", ctypes.c_int),
                ("clock_ghz", ctypes.c_double), ("verified_units", ctypes.c_int)]


@dataclass
class PermResult:
    value: float
    sum_hi: float
    sum_lo: float
    abs_sum: float
    err_bound: float
    kernel_ms: float
    subsets: float
    n: int
    mode: int
    integral: int
    err_est: float
    nu: int
    main_ms: float
    main_subsets: float
    nu1: int = 0
    nu2: int = 0
    certified: int = 0
    zero_pattern: int = 0
    clock_ghz: float = 0.0
    verified_units: int = 0

    @property
    def stated(self) -> float:
        """The stated error of the value: err_bound for an integer matrix
        (X exact: a worst-case bound of the whole error), else the larger of
        err_bound and err_est (err_bound leaves out the X drift)."""
        return max(self.err_bound, self.err_est)

    @classmethod
    def from_c(cls, r: PermResultC) -> "PermResult":
        return cls(r.value, r.sum_hi, r.sum_lo, r.abs_sum, r.err_bound, r.kernel_ms, r.subsets,
                   r.n, r.mode, r.integral, r.err_est, r.nu, r.main_ms,
                   r.main_subsets, r.nu1, r.nu2, r.certified, r.zero_pattern, r.clock_ghz,
                   r.verified_units)


class BatchTuningC(ctypes.Structure):
    _fields_ = [("item_log2", ctypes.c_int), ("streams", ctypes.c_int),
                ("max_grid", ctypes.c_int), ("device_ms", ctypes.c_double),
                ("trials", ctypes.c_int)]


class ServerConfigC(ctypes.Structure):
    _fields_ = [("item_log2", ctypes.c_int), ("grid", ctypes.c_int),
                ("max_items", ctypes.c_int), ("idle_us", ctypes.c_double),
                ("max_resident", ctypes.c_int), ("max_queued_work", ctypes.c_double),
                ("prep_threads", ctypes.c_int), ("arena_mb", ctypes.c_double)]


class ServerStatsC(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in
                ("submitted", "admitted", "completed", "launches", "items", "pending",
                 "uploads", "stops", "laps")] + \
               [(k, ctypes.c_double) for k in
                ("admit_ms", "publish_ms", "retire_ms", "stop_ms", "gpu_ms")] + \
               [("dd_passes", ctypes.c_longlong)] + \
               [(k, ctypes.c_double) for k in ("prep_ms", "submit_ms")]


PERM_PENDING = 5

_LIB = None


def _bind(L, name: str, argtypes, restype=ctypes.c_int):
    """Declare an entry point's signature; an entry point this build of the
    library lacks stays undeclared and fails when called (AttributeError),
    the others keep working."""
    try:
        f = getattr(L, name)
    except AttributeError:
        return
    f.argtypes = argtypes
    f.restype = restype


def lib():
    """Load the HIP library; raises PermError when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise PermError(f"{LIB_PATH} not built (run __graft_entry__.build() or make)")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
        L.perm_last_error.restype = ctypes.c_char_p
        L.perm_last_status.restype = ctypes.c_int
        L.perm_max_n.restype = ctypes.c_int
        L.perm_set_device.argtypes = [ctypes.c_int]
        L.perm_dense.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(PermResultC)]
        L.perm_sparse.argtypes = [ctypes.c_int, vp, ip, ip, vp, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(PermResultC)]
        _bind(L, "perm_dense_noalloc", [ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(PermResultC)])
        L.perm_plan_dense.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.perm_plan_sparse.argtypes = [ctypes.c_int, vp, ip, ip, vp, ctypes.c_int,
                                       ctypes.POINTER(vp)]
        L.perm_plan_run.argtypes = [vp, ctypes.POINTER(PermResultC)]
        L.perm_plan_run_mode.argtypes = [vp, ctypes.c_int, ctypes.POINTER(PermResultC)]
        _bind(L, "perm_plan_geometry", [vp, ip, ip, ctypes.POINTER(ctypes.c_ulonglong)])
        _bind(L, "perm_plan_run_blocks", [vp, ctypes.c_int, ctypes.c_ulonglong,
                                          ctypes.c_ulonglong, ctypes.POINTER(PermResultC)])
        L.perm_plan_destroy.argtypes = [vp]
        L.perm_plan_destroy.restype = None
        for v in DENSE_VARIANTS:
            for dt, sfx in _SFX.items():
                f = getattr(L, f"perm_{v}_dense_{sfx}")
                f.restype = ctypes.c_double
                f.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for v in SPARSE_VARIANTS[1:]:
            for dt, sfx in _SFX.items():
                f = getattr(L, f"perm_{v}_sparse_{sfx}")
                f.restype = ctypes.c_double
                f.argtypes = [vp, ip, ip, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.perm_batch_create.argtypes = [ctypes.POINTER(vp)]
        L.perm_batch_set.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.perm_batch_add_dense.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ip]
        L.perm_batch_add_sparse.argtypes = [vp, ctypes.c_int, vp, ip, ip, vp, ctypes.c_int, ip]
        _bind(L, "perm_batch_add_dense_many", [vp, ctypes.c_int, ip,
                                               ctypes.POINTER(ctypes.c_void_p), ip,
                                               ctypes.c_int, ip])
        _bind(L, "perm_batch_compute_dense", [vp, ctypes.c_int, ip,
                                              ctypes.POINTER(ctypes.c_void_p), ip, ctypes.c_int,
                                              ctypes.POINTER(PermResultC),
                                              ctypes.POINTER(ctypes.c_double)])
        L.perm_batch_size.argtypes = [vp]
        L.perm_batch_run.argtypes = [vp, ctypes.POINTER(PermResultC),
                                     ctypes.POINTER(ctypes.c_double)]
        L.perm_batch_destroy.argtypes = [vp]
        L.perm_batch_destroy.restype = None
        L.perm_batch_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(BatchTuningC)]
        L.perm_batch_get.argtypes = [vp, ctypes.POINTER(BatchTuningC)]
        u64p, dp = ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_double)
        L.perm_server_default_config.argtypes = [ctypes.POINTER(ServerConfigC)]
        L.perm_server_default_config.restype = None
        L.perm_server_create.argtypes = [ctypes.POINTER(ServerConfigC), ctypes.POINTER(vp)]
        L.perm_server_submit_dense.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, u64p]
        L.perm_server_submit_sparse.argtypes = [vp, ctypes.c_int, vp, ip, ip, vp, ctypes.c_int,
                                                u64p]
        ipp = ctypes.POINTER(ip)
        _bind(L, "perm_server_submit_many", [vp, ctypes.c_int, ip, ctypes.POINTER(vp), ipp, ipp,
                                             ctypes.POINTER(vp), ip, u64p])
        L.perm_server_wait.argtypes = [vp, ctypes.c_ulonglong, ctypes.POINTER(PermResultC), dp]
        L.perm_server_poll.argtypes = [vp, ctypes.c_ulonglong, ctypes.POINTER(PermResultC), dp]
        L.perm_server_wait_many.argtypes = [vp, u64p, ctypes.c_int, ctypes.POINTER(PermResultC),
                                            dp]
        L.perm_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStatsC)]
        L.perm_server_destroy.argtypes = [vp]
        L.perm_server_destroy.restype = None
        _bind(L, "perm_server_drive", [vp, ctypes.c_int, ip, ctypes.POINTER(vp), ipp, ipp,
                                       ctypes.POINTER(vp), ip, dp, ctypes.POINTER(PermResultC),
                                       dp, dp])
        _bind(L, "perm_server_autotune", [ctypes.POINTER(ServerConfigC), ctypes.c_int, ip,
                                          ctypes.POINTER(vp), ipp, ipp, ctypes.POINTER(vp), ip,
                                          ctypes.POINTER(ServerConfigC), dp])
        L.perm_connect.restype = None
        L.perm_chunked.argtypes = [ctypes.c_int, vp, ip, ip, vp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.POINTER(PermResultC),
                                   ctypes.POINTER(ctypes.c_int)]
        L.perm_skip.argtypes = [ctypes.c_int, vp, ip, ip, ip, ip, vp, ctypes.c_int,
                                ctypes.POINTER(PermResultC)]
        L.perm_cpu_skip.argtypes = [ctypes.c_int, ip, ip, vp, ip, ip, vp, ctypes.c_int,
                                    ctypes.c_int, ctypes.POINTER(PermResultC)]
        for sfx in _SFX.values():
            f = getattr(L, f"perm_xshared_coalescing_mshared_skipper_{sfx}")
            f.restype = ctypes.c_double
            f.argtypes = [vp, ip, ip, ip, ip, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.perm_read_calculate_return.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 5
        L.perm_read_calculate_return.restype = ctypes.c_double
        L.perm_matlab_calculate_return.argtypes = [ip] + [ctypes.c_int] * 7
        L.perm_matlab_calculate_return.restype = ctypes.c_double
        for sfx in _SFX.values():
            f = getattr(L, f"perm_parallel_perman64_{sfx}")
            f.restype = ctypes.c_double
            f.argtypes = [vp, ctypes.c_int, ctypes.c_int]
            f = getattr(L, f"perm_parallel_perman64_sparse_{sfx}")
            f.restype = ctypes.c_double
            f.argtypes = [vp, ip, ip, vp, ctypes.c_int, ctypes.c_int]
        _LIB = L
    return _LIB


def _check(status: int):
    if status != 0:
        raise PermError(f"status {status}: {lib().perm_last_error().decode()}")


def _mat(mat) -> tuple[np.ndarray, int]:
    a = np.ascontiguousarray(mat)
    if a.dtype not in _DT:
        a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("square matrix expected")
    return a, _DT[a.dtype]


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def compress(mat) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """CCS of the entries > 0 (the reference's matrix2compressed filter)."""
    a, _ = _mat(mat)
    n = a.shape[0]
    cptrs = np.zeros(n + 1, dtype=np.int32)
    rows, vals = [], []
    for k in range(n):
        cptrs[k] = len(rows)
        nz = np.nonzero(a[:, k] > 0)[0]
        rows.extend(nz.tolist())
        vals.extend(a[nz, k].tolist())
    cptrs[n] = len(rows)
    return cptrs, np.array(rows, dtype=np.int32), np.array(vals, dtype=a.dtype)


def dense(mat, algo_id: int = 4) -> PermResult:
    """Exact permanent on the GPU through the dense path (reference IDs 0-4)."""
    a, dt = _mat(mat)
    r = PermResultC()
    _check(lib().perm_dense(dt, a.ctypes.data, a.shape[0], algo_id, ctypes.byref(r)))
    return PermResult.from_c(r)


def sparse(mat, cptrs=None, rows=None, cvals=None, algo_id: int = 4) -> PermResult:
    """Exact permanent on the GPU through the sparse (-s) path (IDs 1-4).
    x is initialised from the dense matrix, column updates from the CCS
    (built from `mat` with the > 0 filter when not given)."""
    a, dt = _mat(mat)
    if cptrs is None:
        cptrs, rows, cvals = compress(a)
    cptrs = np.ascontiguousarray(cptrs, dtype=np.int32)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cvals = np.ascontiguousarray(cvals, dtype=a.dtype)
    r = PermResultC()
    _check(lib().perm_sparse(dt, a.ctypes.data, _ip(cptrs), _ip(rows), cvals.ctypes.data,
                             a.shape[0], algo_id, ctypes.byref(r)))
    return PermResult.from_c(r)


def dropin_dense(mat, algo_id: int, grid_dim: int = 2048, block_dim: int = 256) -> float:
    """The reference-signature entry point gpu_perman64_<variant>(mat, nov, grid, block)."""
    a, dt = _mat(mat)
    f = getattr(lib(), f"perm_{DENSE_VARIANTS[algo_id]}_dense_{_SFX[dt]}")
    return f(a.ctypes.data, a.shape[0], grid_dim, block_dim)


def dropin_sparse(mat, algo_id: int, grid_dim: int = 2048, block_dim: int = 256) -> float:
    a, dt = _mat(mat)
    cptrs, rows, cvals = compress(a)
    f = getattr(lib(), f"perm_{SPARSE_VARIANTS[algo_id]}_sparse_{_SFX[dt]}")
    return f(a.ctypes.data, _ip(cptrs), _ip(rows), cvals.ctypes.data, a.shape[0], grid_dim,
             block_dim)


def cpu_dense(mat, threads: int = 8) -> float:
    """`perman -c`: the library's OpenMP FP64 Ryser (parallel_perman64 contract)."""
    a, dt = _mat(mat)
    return getattr(lib(), f"perm_parallel_perman64_{_SFX[dt]}")(a.ctypes.data, a.shape[0],
                                                                  threads)


def cpu_sparse(mat, threads: int = 8) -> float:
    a, dt = _mat(mat)
    cptrs, rows, cvals = compress(a)
    return getattr(lib(), f"perm_parallel_perman64_sparse_{_SFX[dt]}")(
        a.ctypes.data, _ip(cptrs), _ip(rows), cvals.ctypes.data, a.shape[0], threads)


class Plan:
    """A matrix resident on the device; run() times only device work."""

    def __init__(self, mat, sparse_path: bool = False):
        a, dt = _mat(mat)
        self._h = ctypes.c_void_p()
        if sparse_path:
            cptrs, rows, cvals = compress(a)
            self._keep = (cptrs, rows, cvals)
            st = lib().perm_plan_sparse(dt, a.ctypes.data, _ip(cptrs), _ip(rows),
                                        cvals.ctypes.data, a.shape[0], ctypes.byref(self._h))
        else:
            st = lib().perm_plan_dense(dt, a.ctypes.data, a.shape[0], ctypes.byref(self._h))
        _check(st)

    def run(self, mode: int = -1) -> PermResult:
        """mode -1: automatic; 0 FP64; 1 double-double products; 2 double-double X too."""
        r = PermResultC()
        if mode < 0:
            _check(lib().perm_plan_run(self._h, ctypes.byref(r)))
        else:
            _check(lib().perm_plan_run_mode(self._h, mode, ctypes.byref(r)))
        return PermResult.from_c(r)

    def geometry(self) -> tuple:
        """(walk_bits, high_bits, blocks) of the plan's walk split."""
        w, h, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_ulonglong()
        _check(lib().perm_plan_geometry(self._h, ctypes.byref(w), ctypes.byref(h),
                                        ctypes.byref(b)))
        return w.value, h.value, b.value

    def run_blocks(self, block_begin: int, block_end: int, mode: int = 0) -> PermResult:
        """The sum over workgroups [block_begin, block_end) in `mode`."""
        r = PermResultC()
        _check(lib().perm_plan_run_blocks(self._h, mode, block_begin, block_end,
                                          ctypes.byref(r)))
        return PermResult.from_c(r)

    def close(self):
        if self._h:
            lib().perm_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchInputs:
    """The C arrays of perm_batch_add_dense_many for a list of dense
    matrices, built once (the converted matrices are kept alive here)."""

    def __init__(self, mats):
        arrs = [_mat(m) for m in mats]
        k = len(arrs)
        self.mats = [a for a, _ in arrs]
        self.count = k
        self.dts = np.array([dt for _, dt in arrs], dtype=np.int32)
        self.ns = np.array([a.shape[0] for a, _ in arrs], dtype=np.int32)
        self.ptrs = (ctypes.c_void_p * k)(*[a.ctypes.data for a, _ in arrs])


class Batch:
    """Many matrices (n <= 40) resident on the device; run() replays the
    captured graph of persistent variant kernels and returns one result per
    matrix (insertion order) and the device time in ms."""

    def __init__(self, item_log2: int = 0, streams: int = 0, max_grid: int = 0):
        self._h = ctypes.c_void_p()
        _check(lib().perm_batch_create(ctypes.byref(self._h)))
        _check(lib().perm_batch_set(self._h, item_log2, streams, max_grid))
        self._keep = []

    def add(self, mat, sparse_path: bool = False) -> int:
        a, dt = _mat(mat)
        idx = ctypes.c_int()
        if sparse_path:
            cptrs, rows, cvals = compress(a)
            self._keep.append((cptrs, rows, cvals))
            st = lib().perm_batch_add_sparse(self._h, dt, a.ctypes.data, _ip(cptrs), _ip(rows),
                                             cvals.ctypes.data, a.shape[0], ctypes.byref(idx))
        else:
            st = lib().perm_batch_add_dense(self._h, dt, a.ctypes.data, a.shape[0],
                                            ctypes.byref(idx))
        _check(st)
        return idx.value

    def add_many(self, mats, threads: int = 8) -> int:
        """Add dense matrices in one call (host preparation on `threads`
        threads); returns the index of the first."""
        arrs = [_mat(m) for m in mats]  # the library copies what it keeps
        k = len(arrs)
        dts = np.array([dt for _, dt in arrs], dtype=np.int32)
        ns = np.array([a.shape[0] for a, _ in arrs], dtype=np.int32)
        ptrs = (ctypes.c_void_p * k)(*[a.ctypes.data for a, _ in arrs])
        first = ctypes.c_int()
        _check(lib().perm_batch_add_dense_many(self._h, k, _ip(dts), ptrs, _ip(ns), threads,
                                               ctypes.byref(first)))
        return first.value

    def add_prepared(self, inputs: BatchInputs, threads: int = 8) -> int:
        """add_many for prebuilt BatchInputs."""
        first = ctypes.c_int()
        _check(lib().perm_batch_add_dense_many(self._h, inputs.count, _ip(inputs.dts),
                                               inputs.ptrs, _ip(inputs.ns), threads,
                                               ctypes.byref(first)))
        return first.value

    def compute_prepared(self, inputs: BatchInputs, results, threads: int = 8) -> float:
        """perm_batch_compute_dense on an empty batch: prepare, upload and
        run pipelined; results (a ctypes array of PermResultC, caller's
        order) are filled; returns the device time in ms."""
        ms = ctypes.c_double()
        _check(lib().perm_batch_compute_dense(self._h, inputs.count, _ip(inputs.dts),
                                              inputs.ptrs, _ip(inputs.ns), threads, results,
                                              ctypes.byref(ms)))
        return ms.value

    def set(self, item_log2: int = 0, streams: int = 0, max_grid: int = 0):
        _check(lib().perm_batch_set(self._h, item_log2, streams, max_grid))

    def autotune(self, reps: int = 2) -> dict:
        """Time candidate work splits on this batch and keep the fastest."""
        t = BatchTuningC()
        _check(lib().perm_batch_autotune(self._h, reps, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in BatchTuningC._fields_}

    def tuning(self) -> dict:
        t = BatchTuningC()
        _check(lib().perm_batch_get(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in BatchTuningC._fields_}

    def run(self):
        n = lib().perm_batch_size(self._h)
        res = (PermResultC * n)()
        ms = ctypes.c_double()
        _check(lib().perm_batch_run(self._h, res, ctypes.byref(ms)))
        return [PermResult.from_c(r) for r in res], ms.value

    def close(self):
        if self._h:
            lib().perm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _crs(a: np.ndarray):
    """CRS (rptrs, cols, rvals) with the reference's > 0 filter."""
    rptrs, cols, rvals = [0], [], []
    for i in range(a.shape[0]):
        for j in range(a.shape[1]):
            if a[i, j] > 0:
                cols.append(j)
                rvals.append(a[i, j])
        rptrs.append(len(cols))
    return (np.asarray(rptrs, dtype=np.int32), np.asarray(cols, dtype=np.int32),
            np.asarray(rvals, dtype=a.dtype))


def skip(mat) -> PermResult:
    """Sparse skipper (ID 7, gpu_exact_sparse.cu:1124): zero-run skipping for
    integer matrices (n >= 16), the sparse path otherwise.  main_subsets of
    the result = subsets actually walked."""
    a, dt = _mat(mat)
    cptrs, rows, cvals = compress(a)
    rptrs, cols, _ = _crs(a)
    r = PermResultC()
    _check(lib().perm_skip(dt, a.ctypes.data, _ip(rptrs), _ip(cols), _ip(cptrs), _ip(rows),
                           cvals.ctypes.data, a.shape[0], ctypes.byref(r)))
    return PermResult.from_c(r)


def dropin_skipper(mat) -> float:
    a, dt = _mat(mat)
    cptrs, rows, cvals = compress(a)
    rptrs, cols, _ = _crs(a)
    f = getattr(lib(), f"perm_xshared_coalescing_mshared_skipper_{_SFX[dt]}")
    return f(a.ctypes.data, _ip(rptrs), _ip(cols), _ip(cptrs), _ip(rows), cvals.ctypes.data,
             a.shape[0], 2048, 256)


def cpu_skip(mat, threads: int = 8) -> PermResult:
    """CPU skip walk of `perman -c -s -p2/-p3` (parallel_skip_perman64_w and
    _w_balanced, algo.h:749, :886).  main_subsets = terms evaluated."""
    a, dt = _mat(mat)
    cptrs, rows, cvals = compress(a)
    rptrs, cols, rvals = _crs(a)
    r = PermResultC()
    _check(lib().perm_cpu_skip(dt, _ip(rptrs), _ip(cols), rvals.ctypes.data, _ip(cptrs),
                               _ip(rows), cvals.ctypes.data, a.shape[0], threads,
                               ctypes.byref(r)))
    return PermResult.from_c(r)


def chunked(mat, use_cpu: bool = False, threads: int = 8, sparse_path: bool = False):
    """The chunked GPU(+CPU) scheduler (reference IDs 6, gpu_exact_dense.cu:777):
    returns (PermResult, number of chunks the CPU computed)."""
    a, dt = _mat(mat)
    r, nc = PermResultC(), ctypes.c_int()
    if sparse_path:
        cptrs, rows, cvals = compress(a)
        st = lib().perm_chunked(dt, a.ctypes.data, _ip(cptrs), _ip(rows), cvals.ctypes.data,
                                a.shape[0], int(use_cpu), threads, ctypes.byref(r),
                                ctypes.byref(nc))
    else:
        st = lib().perm_chunked(dt, a.ctypes.data, None, None, None, a.shape[0], int(use_cpu),
                                threads, ctypes.byref(r), ctypes.byref(nc))
    _check(st)
    return PermResult.from_c(r), nc.value


def read_calculate_return(filename: str, algorithm: int, nt: int = 16, x: int = 0, y: int = 0,
                          z: int = 0) -> float:
    """The reference connector's entry (py_connector.c:63) on the GPU path:
    NaN for the approximation IDs 0-3 (see perm_last_error)."""
    return lib().perm_read_calculate_return(filename.encode(), algorithm, nt, x, y, z)


def matlab_calculate_return(mat, algorithm: int, nt: int = 16, x: int = 0, y: int = 0,
                            z: int = 0) -> float:
    """matlab_calculate_return (py_connector.c:146): int32 row-major matrix."""
    a = np.ascontiguousarray(mat, dtype=np.int32)
    return lib().perm_matlab_calculate_return(_ip(a.reshape(-1)), algorithm, nt, x, y, z,
                                              a.shape[0], int((a != 0).sum()))


def last_error() -> str:
    return lib().perm_last_error().decode()


class Request:
    """A matrix ready to hand to the library: the contiguous array, its dtype
    code and, for the sparse path, its CCS (built once, outside any timing)."""
    __slots__ = ("a", "dt", "n", "sparse", "cptrs", "rows", "cvals", "_args")

    def __init__(self, mat, sparse_path: bool = False):
        self.a, self.dt = _mat(mat)
        self.n = self.a.shape[0]
        self.sparse = bool(sparse_path)
        if self.sparse:
            self.cptrs, self.rows, self.cvals = compress(self.a)
            # the C arguments, converted once (the arrays above keep them valid)
            self._args = (self.dt, self.a.ctypes.data, _ip(self.cptrs), _ip(self.rows),
                          self.cvals.ctypes.data, self.n)
        else:
            self.cptrs = self.rows = self.cvals = None
            self._args = (self.dt, self.a.ctypes.data, self.n)


def solve(req: Request) -> PermResult:
    """One plain call for a prepared request (perm_dense / perm_sparse, ID 4)."""
    r = PermResultC()
    if req.sparse:
        st = lib().perm_sparse(*req._args, 4, ctypes.byref(r))
    else:
        st = lib().perm_dense(*req._args, 4, ctypes.byref(r))
    _check(st)
    return PermResult.from_c(r)


class RequestBatch:
    """The C arrays of perm_server_submit_many for a list of Requests, built
    once (the Requests keep the matrices alive)."""

    def __init__(self, reqs):
        k = len(reqs)
        vp = ctypes.c_void_p
        ip = ctypes.POINTER(ctypes.c_int)
        self.reqs = list(reqs)
        self.count = k
        dts = (ctypes.c_int * k)(*[q.dt for q in reqs])
        ns = (ctypes.c_int * k)(*[q.n for q in reqs])
        mats = (vp * k)(*[q.a.ctypes.data for q in reqs])
        cps = (ip * k)(*[_ip(q.cptrs) if q.sparse else None for q in reqs])
        rws = (ip * k)(*[_ip(q.rows) if q.sparse else None for q in reqs])
        cvs = (vp * k)(*[q.cvals.ctypes.data if q.sparse else None for q in reqs])
        self.args = (k, dts, mats, cps, rws, cvs, ns)


def autotune_server(reqs, **base) -> tuple:
    """perm_server_autotune on a sample of Requests: (best config as a dict,
    its closed-loop requests per second)."""
    c = ServerConfigC()
    lib().perm_server_default_config(ctypes.byref(c))
    for k, v in base.items():
        setattr(c, k, v)
    b = RequestBatch(reqs)
    best = ServerConfigC()
    rate = ctypes.c_double()
    _check(lib().perm_server_autotune(ctypes.byref(c), *b.args, ctypes.byref(best),
                                      ctypes.byref(rate)))
    return {k: getattr(best, k) for k, _ in ServerConfigC._fields_}, rate.value


class Server:
    """Continuous-batching permanent server on one GPU (perm_server_*).
    submit() returns a ticket at once; wait() blocks for its result and
    returns (PermResult, latency_ms); poll() returns None while pending.
    Keyword arguments override perm_server_config fields."""

    def __init__(self, **cfg):
        c = ServerConfigC()
        lib().perm_server_default_config(ctypes.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self._h = ctypes.c_void_p()
        _check(lib().perm_server_create(ctypes.byref(c), ctypes.byref(self._h)))

    def submit(self, mat, sparse_path: bool = False) -> int:
        a, dt = _mat(mat)
        tid = ctypes.c_ulonglong()
        if sparse_path:
            cptrs, rows, cvals = compress(a)
            st = lib().perm_server_submit_sparse(self._h, dt, a.ctypes.data, _ip(cptrs),
                                                 _ip(rows), cvals.ctypes.data, a.shape[0],
                                                 ctypes.byref(tid))
        else:
            st = lib().perm_server_submit_dense(self._h, dt, a.ctypes.data, a.shape[0],
                                                ctypes.byref(tid))
        _check(st)  # the matrix is copied before submit returns
        return tid.value

    def submit_many(self, reqs) -> list:
        """Submit prepared Requests in one call (perm_server_submit_many);
        returns their tickets."""
        return list(self.submit_batch(RequestBatch(reqs)))

    def submit_batch(self, batch: "RequestBatch"):
        """Submit a RequestBatch (its C arrays were built beforehand) in one
        call; returns the ctypes array of tickets."""
        ids = (ctypes.c_ulonglong * batch.count)()
        _check(lib().perm_server_submit_many(self._h, *batch.args, ids))
        return ids

    def drive(self, batch: "RequestBatch", arrival_s):
        """Open loop (perm_server_drive): request i arrives arrival_s[i]
        seconds after the start (requests already due are submitted
        together); returns the ctypes arrays of results and latencies (from
        each request's arrival time to its result) and the wall time in
        seconds."""
        k = batch.count
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        res = (PermResultC * k)()
        lat = (ctypes.c_double * k)()
        wall = ctypes.c_double()
        _check(lib().perm_server_drive(self._h, *batch.args,
                                       arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       res, lat, ctypes.byref(wall)))
        return res, lat, wall.value

    def wait_many_raw(self, ids):
        """Wait for a ctypes array of tickets; returns the ctypes arrays of
        results and latencies (no Python objects per result)."""
        k = len(ids)
        res = (PermResultC * k)()
        lat = (ctypes.c_double * k)()
        _check(lib().perm_server_wait_many(self._h, ids, k, res, lat))
        return res, lat

    def submit_request(self, req: Request) -> int:
        tid = ctypes.c_ulonglong()
        if req.sparse:
            st = lib().perm_server_submit_sparse(self._h, *req._args, ctypes.byref(tid))
        else:
            st = lib().perm_server_submit_dense(self._h, *req._args, ctypes.byref(tid))
        _check(st)
        return tid.value

    def wait(self, ticket: int):
        r, lat = PermResultC(), ctypes.c_double()
        _check(lib().perm_server_wait(self._h, ticket, ctypes.byref(r), ctypes.byref(lat)))
        return PermResult.from_c(r), lat.value

    def wait_many(self, tickets):
        """Results of several tickets at once: [(PermResult, latency_ms)]."""
        k = len(tickets)
        ids = (ctypes.c_ulonglong * k)(*tickets)
        res = (PermResultC * k)()
        lat = (ctypes.c_double * k)()
        _check(lib().perm_server_wait_many(self._h, ids, k, res, lat))
        return [(PermResult.from_c(res[i]), lat[i]) for i in range(k)]

    def poll(self, ticket: int):
        r, lat = PermResultC(), ctypes.c_double()
        st = lib().perm_server_poll(self._h, ticket, ctypes.byref(r), ctypes.byref(lat))
        if st == PERM_PENDING:
            return None
        _check(st)
        return PermResult.from_c(r), lat.value

    def stats(self) -> dict:
        s = ServerStatsC()
        _check(lib().perm_server_stats_get(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStatsC._fields_}

    def close(self):
        if self._h:
            lib().perm_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def served_workload(count: int, seed: int = 6, n_lo: int = 16, n_hi: int = 36,
                    decay: float = 4.0):
    """Request mix of the served configuration: n in [n_lo, n_hi] with
    P(n) proportional to 2^(-(n - n_lo)/decay), density U[0.1, 0.9], half of
    the requests 0/1 (int32) and half FP64 U(0,1], every third one through
    the sparse path.  Request i draws from stream seed*2^20 + i (sizes and
    kinds from stream seed*2^20 + 2^19)."""
    ns = np.arange(n_lo, n_hi + 1)
    w = np.power(2.0, -(ns - n_lo) / decay)
    cdf = np.cumsum(w / w.sum())
    base = np.uint64(seed) << np.uint64(20)
    u = uniforms(int(base) + (1 << 19), 3 * count)
    reqs = []
    for i in range(count):
        n = int(ns[min(np.searchsorted(cdf, u[3 * i]), len(ns) - 1)])
        density = 0.1 + 0.8 * float(u[3 * i + 1])
        binary = u[3 * i + 2] < 0.5
        m = synthetic_matrix(n, density, seed=int(base) + i, binary=bool(binary))
        reqs.append((m.astype(np.int32) if binary else m, i % 3 == 2))
    return reqs


# --------------------------------------------------------------------------
# Seeded synthetic matrices: counter-based splitmix64, identical on every
# machine.  Draw k of stream `seed` is mix(mix(seed) + (k+1)*G), i.e. the
# k-th output of the splitmix64 sequence started at state mix(seed).
_G = np.uint64(0x9E3779B97F4A7C15)


def _mix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniforms(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Draws offset..offset+count-1 of stream `seed` as doubles in [0, 1)."""
    base = _mix(np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
    k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = _mix(base + k * _G)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def synthetic_matrix(n: int, density: float, seed: int, binary: bool = False,
                     full_support: bool = False) -> np.ndarray:
    """n x n float64 matrix: entry (i, j) is nonzero when draw 2(i*n+j) of the
    stream is below `density`, its value 1 - draw 2(i*n+j)+1, in (0, 1] (1.0
    when `binary`).  With `full_support`, attempts r = 1, 2, ... use stream
    seed + r*1000003 until every row and column has a nonzero."""
    for attempt in range(10000):
        u = uniforms(seed + attempt * 1000003, 2 * n * n).reshape(n, n, 2)
        a = np.where(u[:, :, 0] < density, 1.0 if binary else 1.0 - u[:, :, 1], 0.0)
        if not full_support or ((a != 0).any(axis=0).all() and (a != 0).any(axis=1).all()):
            return a
    raise ValueError("no full-support draw found")


def mixed_batch(count: int, n_lo: int, n_hi: int, d_lo: float, d_hi: float, seed: int):
    """`count` matrices with n uniform in [n_lo, n_hi] and density uniform in
    [d_lo, d_hi] (drawn from stream `seed`), matrix i from stream seed*2^20+i."""
    u = uniforms(seed, 2 * count)
    ns = n_lo + np.minimum((u[0::2] * (n_hi - n_lo + 1)).astype(int), n_hi - n_lo)
    ds = d_lo + u[1::2] * (d_hi - d_lo)
    return [synthetic_matrix(int(n), float(d), seed * (1 << 20) + i)
            for i, (n, d) in enumerate(zip(ns, ds))]
