"""Multi-GPU / hybrid algorithm IDs on one GPU (dense 5, 6, 66; sparse 5, 6,
8, 66): the chunk scheduler over the Gray range with and without the CPU
worker, and every drop-in name, against the one-launch path and the exact
oracle."""
import ctypes

import numpy as np
import pytest

import re

import oracle as O
import permanent_amd as P
from conftest import cli_result
from test_capi import run_cli, write_matrix

pytestmark = pytest.mark.gpu


def test_chunked_gpu_only_equals_one_launch():
    a = P.synthetic_matrix(32, 0.5, seed=32)   # 2^3 chunks
    r, cpu = P.chunked(a, use_cpu=False)
    one = P.dense(a)
    assert cpu == 0
    assert (r.value, r.sum_hi, r.sum_lo) == (one.value, one.sum_hi, one.sum_lo)


def test_chunked_with_cpu_worker():
    a = P.synthetic_matrix(31, 0.5, seed=31)   # 2^3 chunks with the CPU worker
    r, cpu = P.chunked(a, use_cpu=True, threads=8)
    one = P.dense(a)
    print(f"CPU worker computed {cpu} chunk(s)")
    assert abs(r.value - one.value) <= r.stated + one.stated
    assert abs(r.value - one.value) <= 1e-11 * abs(one.value)


def test_chunked_small_and_sparse_vs_exact():
    a = P.synthetic_matrix(18, 0.4, seed=18)
    ex = float(O.exact_permanent(a))
    for cpu in (False, True):
        r, _ = P.chunked(a, use_cpu=cpu, threads=4, sparse_path=True)
        assert r.value == pytest.approx(ex, rel=1e-12)


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def test_hybrid_dropins_every_id_and_type():
    L = P.lib()
    base = P.synthetic_matrix(16, 0.5, seed=16, binary=True)
    ex = float(O.exact_permanent(base))
    for dtype, sfx in ((np.int32, "i32"), (np.float32, "f32"), (np.float64, "f64")):
        a = np.ascontiguousarray(base.astype(dtype))
        cptrs, rows, cvals = P.compress(a)
        rptrs, cols, _ = P._crs(a)
        vals = []
        f = getattr(L, f"perm_xshared_coalescing_mshared_multigpu_dense_{sfx}")
        f.restype = ctypes.c_double
        vals.append(f(ctypes.c_void_p(a.ctypes.data), 16, 2, 2048, 256))
        f = getattr(L, f"perm_xshared_coalescing_mshared_multigpu_manual_distribution_dense_{sfx}")
        f.restype = ctypes.c_double
        vals.append(f(ctypes.c_void_p(a.ctypes.data), 16, 4, 2048, 256))
        f = getattr(L, f"perm_xshared_coalescing_mshared_multigpucpu_chunks_dense_{sfx}")
        f.restype = ctypes.c_double
        vals.append(f(ctypes.c_void_p(a.ctypes.data), 16, 2, 1, 4, 2048, 256))
        for name, extra in (("multigpu_sparse", (2, 2048, 256)),
                            ("multigpu_sparse_manual_distribution", (4, 2048, 256)),
                            ("multigpucpu_chunks_sparse", (2, 1, 4, 2048, 256))):
            f = getattr(L, f"perm_xshared_coalescing_mshared_{name}_{sfx}")
            f.restype = ctypes.c_double
            vals.append(f(ctypes.c_void_p(a.ctypes.data), _ip(cptrs), _ip(rows),
                          ctypes.c_void_p(cvals.ctypes.data), 16, *extra))
        f = getattr(L, f"perm_xshared_coalescing_mshared_multigpucpu_chunks_skipper_{sfx}")
        f.restype = ctypes.c_double
        vals.append(f(ctypes.c_void_p(a.ctypes.data), _ip(rptrs), _ip(cols), _ip(cptrs),
                      _ip(rows), ctypes.c_void_p(cvals.ctypes.data), 16, 2, 1, 4, 2048, 256))
        assert vals == [ex] * len(vals), (sfx, vals, ex)


def test_cli_every_gpu_id(tmp_path):
    """`perman -g` for every exact GPU ID of the reference's RunAlgo
    (main.cu:31-148): dense 0-6 and 66, sparse 1-8 and 66; -g -c gives the
    chunked IDs their CPU worker."""
    a = P.synthetic_matrix(17, 0.45, seed=17, binary=True) * 2
    f = tmp_path / "m17"
    write_matrix(f, a, "int")
    ex = float(O.exact_permanent(a))
    runs = [("-p", str(p)) for p in (0, 1, 2, 3, 4, 5, 6, 66)]
    runs += [("-s", "-p", str(p)) for p in (1, 2, 3, 4, 5, 6, 7, 8, 66)]
    runs += [("-c", "-p", "6", "-t", "4"), ("-c", "-s", "-p", "6", "-t", "4")]
    for args in runs:
        r = run_cli("-f", str(f), "-g", *args)
        assert r.returncode == 0, (args, r.stdout, r.stderr)
        name, v = cli_result(r.stdout)
        assert v == ex, (args, r.stdout)


def test_cli_orders_on_the_gpu(tmp_path):
    """`-r1` / `-r2` (the reference's sortOrder / skipOrder preprocessing,
    util.h:553-684, main.cu:502-507) before the GPU paths -p4, -s -p4 and
    -s -p7 (the skipper): the permanent is unchanged, exact for an integer
    matrix with zeros, ties in the column counts and rows of several
    degrees; FP64 within 1e-12 for a real matrix."""
    rng = np.random.default_rng(2024)
    a = ((rng.random((18, 18)) < 0.3) * rng.integers(1, 4, (18, 18))).astype(np.int32)
    a[np.arange(18), (np.arange(18) * 5) % 18] = 1  # every row and column nonzero
    f = tmp_path / "m18"
    write_matrix(f, a, "int")
    ex = float(O.exact_permanent(a))
    b = P.synthetic_matrix(16, 0.4, seed=161, full_support=True)
    g = tmp_path / "m16"
    write_matrix(g, b)
    exb = float(O.exact_permanent(b))
    for r_opt in ("1", "2"):
        for args in (("-p", "4"), ("-s", "-p", "4"), ("-s", "-p", "7")):
            r = run_cli("-f", str(f), "-g", "-r", r_opt, *args)
            assert r.returncode == 0, (r_opt, args, r.stdout, r.stderr)
            assert cli_result(r.stdout)[1] == ex, (r_opt, args, r.stdout)
            r = run_cli("-f", str(g), "-g", "-r", r_opt, *args)
            assert r.returncode == 0, (r_opt, args, r.stdout, r.stderr)
            assert cli_result(r.stdout)[1] == pytest.approx(exb, rel=1e-12), (r_opt, args)


def test_compiled_caller_invokes_every_gpu_template(tmp_path):
    """A main.cu-style C++ caller (RunAlgo's calls, main.cu:31-148) compiled
    with g++ against include/perman.hpp calls every GPU template -- dense
    IDs 0-6 and 66, sparse 1-8 and 66 -- for int, float and double inputs,
    and gets the exact permanent of a 0/1-pattern integer matrix."""
    import os
    import subprocess

    from conftest import ROOT

    a = (P.synthetic_matrix(15, 0.5, seed=151, binary=True) * 2).astype(np.int32)
    ex = float(O.exact_permanent(a))
    n = a.shape[0]
    cp, rows, cv = P.compress(a)
    rp, cols, rv = P._crs(a)
    arr = lambda v: ",".join(str(int(x)) for x in v)  # noqa: E731
    src = tmp_path / "caller.cpp"
    src.write_text(r"""
#include "perman.hpp"
#include <cstdio>
#include <vector>
template <class T> void run_all(const int* src, int n, int nnz, const int* cp0, const int* r0,
                                const int* rp0, const int* c0) {
  std::vector<T> m(src, src + n * n), cv, rv;
  std::vector<int> cp(cp0, cp0 + n + 1), rows(r0, r0 + nnz), rp(rp0, rp0 + n + 1), cols(c0, c0 + nnz);
  for (int k = 0; k < n; ++k) for (int e = cp[k]; e < cp[k + 1]; ++e) cv.push_back(m[rows[e] * n + k]);
  for (int i = 0; i < n; ++i) for (int e = rp[i]; e < rp[i + 1]; ++e) rv.push_back(m[i * n + cols[e]]);
  T* M = m.data(); int *CP = cp.data(), *R = rows.data(), *RP = rp.data(), *C = cols.data();
  T *CV = cv.data();
  double v[] = {
    gpu_perman64_xglobal(M, n, 2048, 256), gpu_perman64_xlocal(M, n, 2048, 256),
    gpu_perman64_xshared(M, n, 2048, 256), gpu_perman64_xshared_coalescing(M, n, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared(M, n, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_multigpu(M, n, 2, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks(M, n, 2, true, 4, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_multigpu_manual_distribution(M, n, 4, 2048, 256),
    gpu_perman64_xlocal_sparse(M, CP, R, CV, n, 2048, 256),
    gpu_perman64_xshared_sparse(M, CP, R, CV, n, 2048, 256),
    gpu_perman64_xshared_coalescing_sparse(M, CP, R, CV, n, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_sparse(M, CP, R, CV, n, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_multigpu_sparse(M, CP, R, CV, n, 2, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks_sparse(M, CP, R, CV, n, 2, true, 4, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_skipper(M, RP, C, CP, R, CV, n, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks_skipper(M, RP, C, CP, R, CV, n, 2, false, 4, 2048, 256),
    gpu_perman64_xshared_coalescing_mshared_multigpu_sparse_manual_distribution(M, CP, R, CV, n, 4, 2048, 256)};
  for (double x : v) std::printf("%.17g\n", x);
}
int main() {
  const int n = """ + str(n) + r""", nnz = """ + str(int(cp[-1])) + r""";
  const int m[] = {""" + arr(a.ravel()) + r"""};
  const int cp[] = {""" + arr(cp) + r"""}, rows[] = {""" + arr(rows) + r"""};
  const int rp[] = {""" + arr(rp) + r"""}, cols[] = {""" + arr(cols) + r"""};
  run_all<int>(m, n, nnz, cp, rows, rp, cols);
  run_all<float>(m, n, nnz, cp, rows, rp, cols);
  run_all<double>(m, n, nnz, cp, rows, rp, cols);
  return 0;
}
""")
    exe = tmp_path / "caller"
    inc = os.path.join(ROOT, "include")
    libdir = os.path.dirname(P.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{inc}", str(src), f"-L{libdir}",
                    "-lperman_hip", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, check=True)
    vals = [float(x) for x in out.stdout.split()]
    assert len(vals) == 3 * 17
    assert vals == [ex] * len(vals), (ex, vals)
