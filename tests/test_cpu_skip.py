"""CPU skip walks of `perman -c -s -p2/-p3` (parallel_skip_perman64_w and
_w_balanced, algo.h:748-1023) against the exact oracle: grid graphs (domino
tilings), random 0/1 and small-integer sparse matrices, real values within
the stated error, the drop-in names and the CLI."""
import ctypes
import re

import numpy as np
import pytest

import oracle as O
import permanent_amd as P
from conftest import cli_result
from test_capi import run_cli, write_matrix


@pytest.mark.parametrize("rows,cols", [(2, 3), (4, 4), (4, 5), (6, 6), (4, 9)])
def test_grid_graphs_exact(rows, cols):
    g = O.grid_graph_matrix(rows, cols).astype(np.int32)
    r = P.cpu_skip(g, threads=4)
    assert r.value == float(O.exact_permanent(g))
    assert r.err_bound == 0.0 and r.integral
    if g.shape[0] >= 12:
        assert r.main_subsets < 0.5 * r.subsets  # most terms skipped


@pytest.mark.parametrize("n,density,seed", [(10, 0.3, 1), (14, 0.25, 2), (17, 0.2, 3),
                                            (19, 0.3, 4)])
def test_random_binary_exact(n, density, seed):
    a = P.synthetic_matrix(n, density, seed, binary=True)
    r = P.cpu_skip(a, threads=3)
    assert r.value == float(O.exact_permanent(a))


def test_small_integers_and_zero_row():
    rng = np.random.default_rng(7)
    a = (rng.integers(0, 4, (15, 15)) * (rng.random((15, 15)) < 0.35)).astype(np.int32)
    assert P.cpu_skip(a).value == float(O.exact_permanent(a))
    a[5, :] = 0
    assert P.cpu_skip(a).value == 0.0


def test_real_values_within_stated_error():
    a = P.synthetic_matrix(18, 0.3, seed=11, full_support=True)
    r = P.cpu_skip(a, threads=4)
    ex = float(O.exact_permanent(a))
    assert abs(r.value - ex) <= r.stated
    assert r.value == pytest.approx(ex, rel=1e-12)


def test_thread_count_does_not_change_the_bits():
    a = P.synthetic_matrix(20, 0.3, seed=12, full_support=True)
    r1, r8 = P.cpu_skip(a, threads=1), P.cpu_skip(a, threads=8)
    assert (r1.sum_hi, r1.sum_lo) == (r8.sum_hi, r8.sum_lo)


@pytest.mark.parametrize("dtype,sfx", [(np.int32, "i32"), (np.float32, "f32"),
                                       (np.float64, "f64")])
def test_dropins(dtype, sfx):
    g = O.grid_graph_matrix(4, 5).astype(dtype)
    cptrs, rows, cvals = P.compress(g)
    rptrs, cols, rvals = P._crs(g)
    ip = lambda v: v.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    ex = float(O.exact_permanent(g))
    for name in ("perm_parallel_skip_perman64_w_", "perm_parallel_skip_perman64_w_balanced_"):
        f = getattr(P.lib(), name + sfx)
        f.restype = ctypes.c_double
        v = f(ip(rptrs), ip(cols), ctypes.c_void_p(rvals.ctypes.data), ip(cptrs), ip(rows),
              ctypes.c_void_p(cvals.ctypes.data), g.shape[0], 4)
        assert v == ex


def test_inconsistent_crs_is_rejected():
    g = O.grid_graph_matrix(2, 3).astype(np.float64)
    cptrs, rows, cvals = P.compress(g)
    rptrs, cols, rvals = P._crs(g)
    cols = cols.copy()
    cols[0] = (cols[0] + 1) % g.shape[0]
    while g[0, cols[0]] != 0:
        cols[0] = (cols[0] + 1) % g.shape[0]
    r = P.PermResultC()
    st = P.lib().perm_cpu_skip(P.DTYPE_F64, P._ip(rptrs), P._ip(cols), rvals.ctypes.data,
                               P._ip(cptrs), P._ip(rows), cvals.ctypes.data, g.shape[0], 2,
                               ctypes.byref(r))
    assert st == 1 and "CRS" in P.lib().perm_last_error().decode()


def test_cli_skip_ids(tmp_path):
    g = O.grid_graph_matrix(4, 6)
    f = tmp_path / "grid46"
    write_matrix(f, g, "int")
    ex = float(O.exact_permanent(g))
    for algo, name in (("2", "parallel_skip_perman64_w"),
                       ("3", "parallel_skip_perman64_w_balanced")):
        r = run_cli("-f", str(f), "-c", "-s", "-p", algo, "-t", "4")
        assert r.returncode == 0, r.stderr
        assert cli_result(r.stdout) == (name, ex), r.stdout
