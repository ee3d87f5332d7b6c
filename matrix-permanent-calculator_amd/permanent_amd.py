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
                ("nu", ctypes.c_int)]


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

    @classmethod
    def from_c(cls, r: PermResultC) -> "PermResult":
        return cls(r.value, r.sum_hi, r.sum_lo, r.abs_sum, r.err_bound, r.kernel_ms, r.subsets,
                   r.n, r.mode, r.integral, r.err_est, r.nu)


_LIB = None


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
        L.perm_plan_dense.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.perm_plan_sparse.argtypes = [ctypes.c_int, vp, ip, ip, vp, ctypes.c_int,
                                       ctypes.POINTER(vp)]
        L.perm_plan_run.argtypes = [vp, ctypes.POINTER(PermResultC)]
        L.perm_plan_run_mode.argtypes = [vp, ctypes.c_int, ctypes.POINTER(PermResultC)]
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

    def close(self):
        if self._h:
            lib().perm_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# Seeded synthetic matrices (splitmix64 stream; identical on every machine)
_M64 = 0xFFFFFFFFFFFFFFFF


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _M64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1) with 53 random bits."""
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def synthetic_matrix(n: int, density: float, seed: int, binary: bool = False,
                     full_support: bool = False) -> np.ndarray:
    """n x n float64 matrix: each entry nonzero with probability `density`,
    nonzeros uniform in (0, 1] (1.0 when `binary`).  With `full_support`,
    the draw is repeated until every row and column has a nonzero."""
    rng = SplitMix64(seed)
    while True:
        a = np.zeros((n, n), dtype=np.float64)
        for i in range(n):
            for j in range(n):
                if rng.uniform() < density:
                    a[i, j] = 1.0 if binary else 1.0 - rng.uniform()  # (0, 1]
        if not full_support or ((a != 0).any(axis=0).all() and (a != 0).any(axis=1).all()):
            return a
