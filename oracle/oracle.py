"""CPU oracle for the Ryser permanent -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as a checker.  The product (the HIP library, its C-ABI
and the `perman` CLI) never links or calls anything here.

Pieces (reference = berkyaglioglu/matrix-permanent-calculator):

* ctypes bindings of ryser_oracle.c -- restatements of the reference's
  double-X evaluators (algo.h:1031-1088, gpu_exact_dense.cu:6-69,
  gpu_exact_sparse.cu:6-87), its float-X CPU baseline (algo.h:662-746,
  algo.h:568-660), the float-X per-thread semantics of its GPU kernels
  (gpu_exact_dense.cu:71-699), a mod-p Ryser and a double-double Ryser.
* exact_permanent(): exact rational permanent by multi-modular Ryser + CRT.
* brute_permanent(): sum over all n! permutations with Fractions (n <= 9).
* read_matrix()/matrix2compressed(): the reference's text format and CCS/CRS
  builder (main.cu:481-587, util.h:343-358, util.h:522-551).
* grid_graph_matrix(): the reference's grid-graph biadjacency generator
  (util.h:403-520); domino_tilings(): exact tiling count by transfer matrix,
  the known answer for its permanent.
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

c_i64 = ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def lib():
    """Load oracle/liboracle.so (built by `make -C oracle`)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orc_perman64.restype = ctypes.c_double
        L.orc_perman64.argtypes = [_dp, ctypes.c_int]
        L.orc_nw_setup_f64.restype = None
        L.orc_nw_setup_f64.argtypes = [_dp, ctypes.c_int, _dp, _dp, _dp]
        L.orc_cpu_perman64.restype = ctypes.c_double
        L.orc_cpu_perman64.argtypes = [_dp, _dp, ctypes.c_int, c_i64, c_i64, ctypes.c_int]
        L.orc_cpu_perman64_sparse.restype = ctypes.c_double
        L.orc_cpu_perman64_sparse.argtypes = [_ip, _ip, _dp, _dp, ctypes.c_int, c_i64, c_i64,
                                              ctypes.c_int]
        for sfx, ptr in (("i32", _ip), ("f32", _fp), ("f64", _dp)):
            f = getattr(L, f"orc_parallel_perman64_{sfx}")
            f.restype = ctypes.c_double
            f.argtypes = [ptr, ctypes.c_int, ctypes.c_int]
            f = getattr(L, f"orc_parallel_perman64_slice_{sfx}")
            f.restype = ctypes.c_double
            f.argtypes = [ptr, ctypes.c_int, ctypes.c_int, c_i64, c_i64, ctypes.c_int]
            f = getattr(L, f"orc_parallel_perman64_sparse_{sfx}")
            f.restype = ctypes.c_double
            f.argtypes = [ptr, _ip, _ip, ptr, ctypes.c_int, ctypes.c_int]
            f = getattr(L, f"orc_gpu_dense_floatx_{sfx}")
            f.restype = ctypes.c_double
            f.argtypes = [ptr, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_lane_range_dd.restype = None
        L.orc_lane_range_dd.argtypes = [_dp, ctypes.c_int, ctypes.c_int, c_i64, c_i64,
                                        ctypes.c_int, _dp, _dp]
        L.orc_ryser_u128.restype = None
        L.orc_ryser_u128.argtypes = [ctypes.POINTER(ctypes.c_longlong), ctypes.c_int,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_uint64)]
        L.orc_ryser_mod.restype = ctypes.c_uint64
        L.orc_ryser_mod.argtypes = [_u64p, ctypes.c_int, ctypes.c_uint64]
        L.orc_ryser_dd.restype = None
        L.orc_ryser_dd.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        _LIB = L
    return _LIB


def ref_lib():
    """Load oracle/_ref/libref_algo.so (the reference's own algo.h compiled by
    `make -C oracle ref`), or None when it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_algo.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        for sfx, ptr in (("i32", _ip), ("f32", _fp), ("f64", _dp)):
            f = getattr(R, f"ref_parallel_perman64_{sfx}")
            f.restype = ctypes.c_double
            f.argtypes = [ptr, ctypes.c_int, ctypes.c_int]
        R.ref_parallel_perman64_sparse_f64.restype = ctypes.c_double
        R.ref_parallel_perman64_sparse_f64.argtypes = [_dp, _ip, _ip, _dp, ctypes.c_int,
                                                       ctypes.c_int]
        R.ref_perman64_f64.restype = ctypes.c_double
        R.ref_perman64_f64.argtypes = [_dp, ctypes.c_int]
        R.ref_matrix2compressed_f64.restype = ctypes.c_int
        R.ref_matrix2compressed_f64.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp,
                                                _ip, _ip, _dp]
        _REF = R
    return _REF


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


_SFX = {np.dtype(np.int32): ("i32", ctypes.c_int), np.dtype(np.float32): ("f32", ctypes.c_float),
        np.dtype(np.float64): ("f64", ctypes.c_double)}


# --------------------------------------------------------------------------
# Restated reference evaluators
def perman64(mat: np.ndarray) -> float:
    """Serial double-X Ryser, algo.h:1031-1088 (any T converts exactly to
    double before every operation in that function)."""
    a = np.ascontiguousarray(mat, dtype=np.float64)
    return lib().orc_perman64(_ptr(a, ctypes.c_double), a.shape[0])


def nw_setup(mat: np.ndarray):
    """(x, p0, mat_t) of the double-X GPU wrappers, gpu_exact_dense.cu:641-662."""
    a = np.ascontiguousarray(mat, dtype=np.float64)
    n = a.shape[0]
    x = np.zeros(n)
    p0 = ctypes.c_double()
    mt = np.zeros((n, n))
    lib().orc_nw_setup_f64(_ptr(a, ctypes.c_double), n, _ptr(x, ctypes.c_double),
                           ctypes.byref(p0), _ptr(mt, ctypes.c_double))
    return x, p0.value, mt


def cpu_perman64(mat: np.ndarray, threads: int = 1, start: int = 1, end: int | None = None):
    """Full permanent through the double-X range evaluator
    gpu_exact_dense.cu:6-69 with `threads` independent chunks:
    (4(n&1)-2) * (p0 + cpu_perman64(mat_t, x, n, start, end, threads))."""
    a = np.ascontiguousarray(mat, dtype=np.float64)
    n = a.shape[0]
    x, p0, mt = nw_setup(a)
    end = (1 << (n - 1)) if end is None else end
    raw = lib().orc_cpu_perman64(_ptr(mt, ctypes.c_double), _ptr(x, ctypes.c_double), n, start,
                                 end, threads)
    return (4 * (n & 1) - 2) * (p0 + raw)


def cpu_perman64_sparse(mat: np.ndarray, threads: int = 1):
    """Full permanent through the double-X CCS evaluator gpu_exact_sparse.cu:6-87,
    x-initialised as the sparse GPU wrappers do (gpu_exact_sparse.cu:854-871)."""
    a = np.ascontiguousarray(mat, dtype=np.float64)
    n = a.shape[0]
    x, p0 = sparse_wrapper_setup(a)
    cptrs, rows, cvals, _, _, _ = matrix2compressed(a)
    raw = lib().orc_cpu_perman64_sparse(_ptr(cptrs, ctypes.c_int), _ptr(rows, ctypes.c_int),
                                        _ptr(cvals, ctypes.c_double), _ptr(x, ctypes.c_double), n,
                                        1, 1 << (n - 1), threads)
    return (4 * (n & 1) - 2) * (p0 + raw)


def sparse_wrapper_setup(mat: np.ndarray):
    """x-vector and p0 of the sparse GPU wrappers (gpu_exact_sparse.cu:854-871):
    the row sum adds only entries != 0 (so the same as the dense sum), x and
    p0 in double."""
    a = np.ascontiguousarray(mat, dtype=np.float64)
    n = a.shape[0]
    x = np.zeros(n)
    p = 1.0
    for j in range(n):
        rs = 0.0
        for k in range(n):
            if a[j, k] != 0:
                rs += float(a[j, k])
        x[j] = float(a[j, n - 1]) - rs / 2
        p *= x[j]
    return x, p


def parallel_perman64(mat: np.ndarray, threads: int = 1) -> float:
    """Float-X CPU baseline algo.h:662-746 restated; the element type of
    `mat` (int32/float32/float64) selects the template instance."""
    a = np.ascontiguousarray(mat)
    sfx, ct = _SFX[a.dtype]
    return getattr(lib(), f"orc_parallel_perman64_{sfx}")(_ptr(a, ct), a.shape[0], threads)


def parallel_perman64_slice(mat: np.ndarray, threads: int, start: int, end: int,
                            parallel: bool = True) -> float:
    """The float-X baseline restricted to the Gray slice [start, end) (used
    to time a bounded sample of a large workload)."""
    a = np.ascontiguousarray(mat)
    sfx, ct = _SFX[a.dtype]
    return getattr(lib(), f"orc_parallel_perman64_slice_{sfx}")(_ptr(a, ct), a.shape[0],
                                                                 threads, start, end,
                                                                 1 if parallel else 0)


def parallel_perman64_sparse(mat: np.ndarray, threads: int = 1) -> float:
    """Float-X sparse CPU algorithm algo.h:568-660 restated."""
    a = np.ascontiguousarray(mat)
    sfx, ct = _SFX[a.dtype]
    cptrs, rows, cvals, _, _, _ = matrix2compressed(a)
    return getattr(lib(), f"orc_parallel_perman64_sparse_{sfx}")(
        _ptr(a, ct), _ptr(cptrs, ctypes.c_int), _ptr(rows, ctypes.c_int), _ptr(cvals, ct),
        a.shape[0], threads)


def gpu_dense_floatx(mat: np.ndarray, algo_id: int, grid_dim: int = 2048,
                     block_dim: int | None = None) -> float:
    """What the reference GPU kernel `algo_id` (0-4) returns, evaluated with
    its float-X per-thread arithmetic on the CPU.  block_dim defaults to the
    reference's choice (main.cu:24-28): 128 for double, 256 otherwise."""
    a = np.ascontiguousarray(mat)
    sfx, ct = _SFX[a.dtype]
    if block_dim is None:
        block_dim = 128 if a.dtype == np.float64 else 256
    return getattr(lib(), f"orc_gpu_dense_floatx_{sfx}")(_ptr(a, ct), a.shape[0], grid_dim,
                                                         block_dim, algo_id)


def ryser_dd(mat: np.ndarray, threads: int = 8) -> tuple[float, float]:
    """Double-double Ryser (relative precision ~1e-30 per operation)."""
    a = np.ascontiguousarray(mat, dtype=np.float64)
    hi, lo = ctypes.c_double(), ctypes.c_double()
    lib().orc_ryser_dd(_ptr(a, ctypes.c_double), a.shape[0], threads, ctypes.byref(hi),
                       ctypes.byref(lo))
    return hi.value, lo.value


# --------------------------------------------------------------------------
# Exact permanent
def _is_probable_prime(n: int) -> bool:
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:  # deterministic for n < 3.3e24
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


_PRIMES: list[int] = []


def _primes(count: int) -> list[int]:
    p = (1 << 62) - 1 if not _PRIMES else _PRIMES[-1] - 2
    while len(_PRIMES) < count:
        if _is_probable_prime(p):
            _PRIMES.append(p)
        p -= 2
    return _PRIMES[:count]


def exact_integer_matrix(mat) -> tuple[list[list[int]], int]:
    """Write mat = M * 2^-e exactly with integer M (each row scaled by its
    own power of two; every finite double is a dyadic rational)."""
    rows = []
    e_total = 0
    for row in np.asarray(mat, dtype=np.float64):
        fr = [Fraction(float(v)) for v in row]
        den = 1
        for f in fr:
            den = max(den, f.denominator)  # all denominators are powers of two
        rows.append([int(f * den) for f in fr])
        e_total += den.bit_length() - 1
    return rows, e_total


def exact_permanent(mat) -> Fraction:
    """Exact permanent of a real (double) matrix: multi-modular Ryser over
    61/62-bit primes, recombined by CRT, with enough primes to exceed twice
    the bound |per(M)| <= prod_i sum_j |M_ij|."""
    M, e = exact_integer_matrix(mat)
    n = len(M)
    if n == 0:
        return Fraction(1)
    bound = 1
    for r in M:
        bound *= max(1, sum(abs(v) for v in r))
    primes = []
    prod = 1
    k = 1
    while prod <= 2 * bound:
        primes = _primes(k)
        prod = math.prod(primes)
        k += 1
    residues = []
    for p in primes:
        a = np.array([[v % p for v in r] for r in M], dtype=np.uint64)
        residues.append(int(lib().orc_ryser_mod(a.ctypes.data_as(_u64p), n, p)))
    # CRT
    x = 0
    for p, r in zip(primes, residues):
        q = prod // p
        x = (x + r * q * pow(q, -1, p)) % prod
    if x > prod // 2:
        x -= prod
    return Fraction(x, 1 << e)


def lane_range_dd(mat, walk_bits: int, lane_begin: int, lane_end: int,
                  threads: int = 8) -> tuple[float, float]:
    """Double-double sum (no scale) over the subsets of lanes
    [lane_begin, lane_end) of the GPU walk's aligned split in natural column
    order (orc_lane_range_dd)."""
    a = np.ascontiguousarray(mat, dtype=np.float64)
    hi, lo = ctypes.c_double(), ctypes.c_double()
    lib().orc_lane_range_dd(_ptr(a, ctypes.c_double), a.shape[0], walk_bits, lane_begin,
                            lane_end, threads, ctypes.byref(hi), ctypes.byref(lo))
    return hi.value, lo.value


def exact_integer_permanent(mat, threads: int = 8) -> int:
    """Exact permanent of an integer matrix (|entries| < 2^31, n <= 62) by
    Ryser's formula in 128-bit wrap-around arithmetic (orc_ryser_u128): the
    result is the permanent modulo 2^128, hence exact when the bound
    prod_i sum_j |a_ij| is below 2^127; larger bounds fall back to the
    multi-modular exact_permanent."""
    a = np.asarray(mat)
    n = a.shape[0]
    if n == 0:
        return 1
    ai = a.astype(np.int64)
    if not np.array_equal(ai, a) or np.abs(ai).max() >= 2 ** 31 or n > 62:
        raise ValueError("integer matrix with |entries| < 2^31 and n <= 62 expected")
    bound = 1
    for r in ai:
        bound *= max(1, int(np.abs(r).sum()))
    if bound >= 2 ** 127:
        f = exact_permanent(a)
        assert f.denominator == 1
        return f.numerator
    ai = np.ascontiguousarray(ai)
    lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
    lib().orc_ryser_u128(ai.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), n, threads,
                         ctypes.byref(lo), ctypes.byref(hi))
    v = (hi.value << 64) | lo.value
    return v - (1 << 128) if v >= 1 << 127 else v


def brute_permanent(mat) -> Fraction:
    """Sum over all n! permutations with exact Fractions (n <= 9)."""
    A = [[Fraction(float(v)) for v in r] for r in np.asarray(mat, dtype=np.float64)]
    n = len(A)
    if n > 9:
        raise ValueError("brute force limited to n <= 9")
    tot = Fraction(0)
    for perm in itertools.permutations(range(n)):
        p = Fraction(1)
        for i, j in enumerate(perm):
            p *= A[i][j]
            if p == 0:
                break
        tot += p
    return tot


def has_perfect_matching(mat) -> bool:
    """Whether the bipartite graph of the nonzero pattern (row i -- column j
    when mat[i, j] != 0) has a perfect matching, by augmenting paths.  Every
    term of the permanent is a product over a permutation, so without one
    the permanent is exactly 0 (checked against brute force in the tests)."""
    a = np.asarray(mat)
    n = a.shape[0]
    adj = [np.nonzero(a[i])[0].tolist() for i in range(n)]
    col_row = [-1] * n

    def augment(i, seen):
        for j in adj[i]:
            if not seen[j]:
                seen[j] = True
                if col_row[j] < 0 or augment(col_row[j], seen):
                    col_row[j] = i
                    return True
        return False

    return all(augment(i, [False] * n) for i in range(n))


# --------------------------------------------------------------------------
# Input format and sparse structures of the reference
def read_matrix(path: str, binary: bool = False):
    """Reference text format: header `n nnz type` (main.cu:484-488), then
    `i j value` lines parsed as the element type; malformed lines are
    skipped; `binary` (-b) stores 1 for every listed entry
    (util.h:343-358).  Returns (dense row-major array of the type, n, nnz,
    type)."""
    with open(path) as fh:
        header = fh.readline().split()
        n, nnz, typ = int(header[0]), int(header[1]), header[2]
        dt = {"int": np.int32, "float": np.float32, "double": np.float64}[typ]
        mat = np.zeros((n, n), dtype=dt)
        for line in fh:
            parts = line.split()
            if len(parts) < 3:
                continue
            try:
                i, j = int(parts[0]), int(parts[1])
                if typ == "int":
                    v = int(parts[2]) if parts[2].lstrip("+-").isdigit() else int(float(parts[2]))
                elif typ == "float":
                    v = _strtof(parts[2])
                else:
                    v = float(parts[2])
            except ValueError:
                continue
            mat[i, j] = 1 if binary else v
    return mat, n, nnz, typ


_LIBC = None


def _strtof(s: str) -> np.float32:
    """Correctly rounded decimal -> float32, as istream >> float does."""
    global _LIBC
    if _LIBC is None:
        _LIBC = ctypes.CDLL("libc.so.6")
        _LIBC.strtof.restype = ctypes.c_float
        _LIBC.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    return np.float32(_LIBC.strtof(s.encode(), None))


def matrix2compressed(mat: np.ndarray):
    """CCS (cptrs, rows, cvals) and CRS (rptrs, cols, rvals) keeping only
    entries > 0, rows/columns ascending (util.h:522-551)."""
    a = np.asarray(mat)
    n = a.shape[0]
    cptrs = np.zeros(n + 1, dtype=np.int32)
    rptrs = np.zeros(n + 1, dtype=np.int32)
    rows, cvals, cols, rvals = [], [], [], []
    for c in range(n):
        cptrs[c] = len(rows)
        for r in range(n):
            if a[r, c] > 0:
                rows.append(r)
                cvals.append(a[r, c])
    cptrs[n] = len(rows)
    for r in range(n):
        rptrs[r] = len(cols)
        for c in range(n):
            if a[r, c] > 0:
                cols.append(c)
                rvals.append(a[r, c])
    rptrs[n] = len(cols)
    return (cptrs, np.array(rows, dtype=np.int32), np.array(cvals, dtype=a.dtype), rptrs,
            np.array(cols, dtype=np.int32), np.array(rvals, dtype=a.dtype))


def grid_graph_matrix(m: int, n: int) -> np.ndarray:
    """Biadjacency 0/1 matrix of the m x n grid graph as built by the
    reference generator (util.h:403-520); its permanent is the number of
    domino tilings of the m x n board."""
    if m % 2 == 1 and n % 2 == 1:
        raise ValueError("one grid dimension must be even")
    if m % 2 == 0:
        R, C = n, m
    else:
        R, C = m, n
    nov = m * n // 2
    h = C // 2
    mat = np.zeros((nov, nov), dtype=np.int32)
    for i in range(R):
        for j in range(C):
            x = i * h + j // 2
            nbrs = []
            if x - h >= 0:
                nbrs.append(x - h)
            if x + h < nov:
                nbrs.append(x + h)
            if j % 2 == 0:
                nbrs += [x] if j == 0 else [x - 1, x]
            else:
                nbrs += [x] if j == C - 1 else [x, x + 1]
            first_colour = (i % 2) == (j % 2)
            for y in nbrs:
                if first_colour:
                    mat[x, y] = 1
                else:
                    mat[y, x] = 1
    return mat


def domino_tilings(m: int, n: int) -> int:
    """Exact number of domino tilings of an m x n board (column-profile
    transfer matrix)."""
    if (m * n) % 2:
        return 0
    if m > n:
        m, n = n, m
    states = {0: 1}
    for _ in range(n):
        nxt: dict[int, int] = {}
        for mask, cnt in states.items():
            def fill(row, cur, out):
                if row == m:
                    nxt[out] = nxt.get(out, 0) + cnt
                    return
                if cur >> row & 1:
                    fill(row + 1, cur, out)
                    return
                fill(row + 1, cur, out | (1 << row))           # horizontal domino
                if row + 1 < m and not (cur >> (row + 1) & 1):
                    fill(row + 2, cur, out)                    # vertical domino
            fill(0, mask, 0)
        states = nxt
    return states.get(0, 0)
