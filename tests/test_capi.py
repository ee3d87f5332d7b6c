"""C ABI and CLI checks that need no GPU: the library loads and exports
every symbol include/perman_hip.h declares, the CPU path (`perman -c`,
parallel_perman64 contract) is exact to 1e-12 on small matrices, the CLI
keeps the reference's flags and result line, and the GPU entry points fail
loudly (NaN / PERM_ERR_NODEV) rather than falling back to the CPU when no
device is present."""
import ctypes
import json
import math
import os
import re
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
import permanent_amd as P
from conftest import GOLDEN, ROOT, cli_result

HEADER = os.path.join(ROOT, "include", "perman_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(perm_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(P.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_dropin_names_cover_every_reference_id():
    names = set(declared_functions())
    for v in P.DENSE_VARIANTS:
        for s in ("i32", "f32", "f64"):
            assert f"perm_{v}_dense_{s}" in names
    for v in P.SPARSE_VARIANTS[1:]:
        for s in ("i32", "f32", "f64"):
            assert f"perm_{v}_sparse_{s}" in names


@pytest.mark.parametrize("n", [1, 2, 5, 10, 14])
def test_cpu_path_matches_exact(n):
    a = P.synthetic_matrix(n, 0.6, seed=40 + n)
    ex = float(O.exact_permanent(a))
    assert P.cpu_dense(a, 4) == pytest.approx(ex, rel=1e-12, abs=1e-300)
    assert P.cpu_dense(a.astype(np.float32), 2) == pytest.approx(
        float(O.exact_permanent(a.astype(np.float32))), rel=1e-12, abs=1e-300)
    assert P.cpu_sparse(a, 3) == pytest.approx(ex, rel=1e-12, abs=1e-300)


def test_cpu_path_c1_configuration():
    """The CPU configuration (n=20, density 0.5, U(0,1], seed 1) through
    the product CPU path, against the exact permanent."""
    c = json.load(open(os.path.join(GOLDEN, "c1_exact.json")))
    a = np.array(c["matrix"])
    ex = Fraction(c["exact"])
    v = P.cpu_dense(a, 8)
    assert abs(Fraction(v) - ex) <= Fraction(1, 10 ** 10) * abs(ex)


def test_cpu_path_integer_is_exact():
    a = (np.random.default_rng(3).random((16, 16)) < 0.5).astype(np.int32)
    ex = O.exact_permanent(a)
    assert P.cpu_dense(a, 4) == float(ex)


def test_gpu_entry_points_fail_loudly_without_device():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU may be present")
    a = np.eye(4)
    v = P.dropin_dense(a, 4)
    assert math.isnan(v)
    assert P.lib().perm_last_status() == 4
    with pytest.raises(P.PermError):
        P.dense(a)


def test_size_and_argument_errors():
    L = P.lib()
    r = P.PermResultC()
    a = np.zeros((70, 70))
    assert L.perm_dense(2, a.ctypes.data, 70, 4, ctypes.byref(r)) == 2  # PERM_ERR_SIZE
    assert L.perm_dense(2, a.ctypes.data, 4, 9, ctypes.byref(r)) == 1  # unknown ID
    assert L.perm_dense(2, None, 4, 4, ctypes.byref(r)) == 1
    assert L.perm_max_n() >= 40


# --- CLI ----------------------------------------------------------------------
def write_matrix(path, a, typ="double"):
    n = a.shape[0]
    nz = np.argwhere(a != 0)
    with open(path, "w") as fh:
        fh.write(f"{n} {len(nz)} {typ}\n")
        for i, j in nz:
            v = a[i, j]
            fh.write(f"{i} {j} {int(v) if typ == 'int' else repr(float(v))}\n")


def run_cli(*args):
    return subprocess.run([P.CLI_PATH, *args], capture_output=True, text=True, timeout=120)


def test_cli_cpu_result_line(tmp_path):
    a = P.synthetic_matrix(12, 0.5, seed=5)
    f = tmp_path / "m12"
    write_matrix(f, a)
    r = run_cli("-f", str(f), "-c", "-t", "4")
    assert r.returncode == 0, r.stderr
    name, v = cli_result(r.stdout)
    assert name == "parallel_perman64"
    assert v == pytest.approx(float(O.exact_permanent(a)), rel=1e-12)


def test_cli_binary_and_int_types(tmp_path):
    a = (np.random.default_rng(1).random((10, 10)) < 0.5).astype(float) * 3
    f = tmp_path / "m10"
    write_matrix(f, a, "int")
    r = run_cli("-f", str(f), "-c", "-b")  # -b: every listed entry becomes 1
    assert r.returncode == 0, r.stderr
    v = cli_result(r.stdout)[1]
    assert v == float(O.exact_permanent((a != 0).astype(float)))
    r = run_cli("-f", str(f), "-c", "-s", "-p", "1")
    name, v = cli_result(r.stdout)
    assert name == "parallel_perman64_sparse"
    assert v == float(O.exact_permanent(a))


def test_cli_preprocessing_orders_keep_the_permanent(tmp_path):
    a = P.synthetic_matrix(11, 0.4, seed=9, full_support=True)
    f = tmp_path / "m11"
    write_matrix(f, a)
    ex = float(O.exact_permanent(a))
    for r_opt in ("1", "2"):
        r = run_cli("-f", str(f), "-c", "-r", r_opt)
        v = cli_result(r.stdout)[1]
        assert v == pytest.approx(ex, rel=1e-12)


def test_cli_flag_errors(tmp_path):
    assert run_cli().returncode == 1  # -f is required
    assert "required" in run_cli().stderr
    f = tmp_path / "m"
    write_matrix(f, np.eye(3))
    assert run_cli("-f", str(f), "-a").returncode == 2   # approximations not in this build
    assert run_cli("-f", str(f), "-c", "-s", "-p", "4").returncode == 2
    for args in (("-p", "9"), ("-p", "7"), ("-s", "-p", "0"), ("-s", "-p", "9")):
        r = run_cli("-f", str(f), "-g", *args)  # rejected before any device call
        assert r.returncode == 2 and "Unknown Algorithm ID" in r.stdout, args
    r = run_cli("-f", str(f), "-t", "-3")
    assert r.returncode == 1 and "requires an argument" in r.stderr


def test_cli_gpu_without_device_fails_loudly(tmp_path):
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU may be present")
    f = tmp_path / "m"
    write_matrix(f, np.eye(3))
    r = run_cli("-f", str(f), "-g", "-p", "4")
    assert r.returncode == 1 and "no HIP device" in r.stderr


def test_cpp_dropin_header_compiles_and_forwards(tmp_path):
    """include/perman.hpp recreates the reference's template names; a
    main.cu-style caller compiles with g++ and gets the CPU results."""
    src = tmp_path / "caller.cpp"
    src.write_text(r"""
#include "perman.hpp"
#include <cstdio>
int main() {
  double m[9] = {1, 2, 0, 0, 3, 4, 5, 0, 6};
  int cp[4] = {0, 2, 4, 6}, rows[6] = {0, 2, 0, 1, 1, 2};
  double cv[6] = {1, 5, 2, 3, 4, 6};
  int mi[9] = {1, 2, 0, 0, 3, 4, 5, 0, 6};
  double (*g)(double*, int, int, int) = &gpu_perman64_xshared_coalescing_mshared<double>;
  double (*s)(float*, int*, int*, float*, int, int, int) = &gpu_perman64_xlocal_sparse<float>;
  int rp[4] = {0, 2, 4, 6}, cols[6] = {0, 1, 1, 2, 0, 2};
  double rv[6] = {1, 2, 3, 4, 5, 6};
  // every multi-GPU / hybrid name instantiates for the three element types
  auto h1 = &gpu_perman64_xshared_coalescing_mshared_multigpu<int>;
  auto h2 = &gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks<float>;
  auto h3 = &gpu_perman64_xshared_coalescing_mshared_multigpu_manual_distribution<double>;
  auto h4 = &gpu_perman64_xshared_coalescing_mshared_multigpu_sparse<int>;
  auto h5 = &gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks_sparse<float>;
  auto h6 = &gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks_skipper<double>;
  auto h7 = &gpu_perman64_xshared_coalescing_mshared_multigpu_sparse_manual_distribution<int>;
  bool h = h1 != nullptr && h2 != nullptr && h3 != nullptr && h4 != nullptr && h5 != nullptr &&
           h6 != nullptr && h7 != nullptr;
  std::printf("%.17g %.17g %.17g %.17g %.17g %d\n", parallel_perman64(m, 3, 2),
              parallel_perman64(mi, 3, 1), parallel_perman64_sparse(m, cp, rows, cv, 3, 1),
              parallel_skip_perman64_w(rp, cols, rv, cp, rows, cv, 3, 2),
              parallel_skip_perman64_w_balanced(rp, cols, rv, cp, rows, cv, 3, 1),
              g != nullptr && s != nullptr && h);
  return 0;
}
""")
    exe = tmp_path / "caller"
    inc = os.path.join(ROOT, "include")
    libdir = os.path.dirname(P.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{inc}", str(src), f"-L{libdir}",
                    "-lperman_hip", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    m = np.array([[1, 2, 0], [0, 3, 4], [5, 0, 6]], dtype=float)
    want = float(O.exact_permanent(m))
    assert [float(v) for v in out[:5]] == [want] * 5 and out[5] == "1"


def test_connector_entry_points_reject_without_touching_the_gpu(tmp_path):
    """The connector equivalents fail loudly (NaN + message) for the
    approximation IDs and unreadable files; these paths make no GPU call."""
    f = tmp_path / "m.txt"
    f.write_text("3 3 double\n0 0 1.5\n1 1 2.0\n2 2 0.5\n")
    for algo in (0, 1, 2, 3, 9):
        assert math.isnan(P.read_calculate_return(str(f), algo))
        assert "algorithm" in P.last_error()
    assert math.isnan(P.read_calculate_return(str(tmp_path / "missing.txt"), 5))
    assert "cannot read" in P.last_error()
    g = tmp_path / "bad.txt"
    g.write_text("3 3 complex\n")
    assert math.isnan(P.read_calculate_return(str(g), 5))
    assert math.isnan(P.matlab_calculate_return(np.eye(3, dtype=np.int32), 1))


def _order_cases():
    """Matrices for the -r1/-r2 comparison: ties in the column counts
    (small-integer entries), entries <= 0 (dropped by the > 0 filter, still
    nonzero for -r2's degrees), empty columns, dense columns."""
    rng = np.random.default_rng(77)
    cases = []
    for n in (5, 8, 13, 21, 32):
        a = rng.integers(0, 4, size=(n, n)) * (rng.random((n, n)) < 0.3)
        cases.append(("ties", a))
        b = a - 2 * (rng.random((n, n)) < 0.15)  # some negative entries
        cases.append(("negative", b))
        c = a.copy()
        c[:, n // 2] = 0  # an empty column
        cases.append(("empty_col", c))
        d = rng.integers(1, 9, size=(n, n)) * (rng.random((n, n)) < 0.7)
        cases.append(("dense", d))
    return cases


def _shim(tmp_path):
    src = os.path.join(ROOT, "tests", "permio_shim.cpp")
    so = tmp_path / "libpermio_shim.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", src, "-o", str(so)],
                   check=True)
    return ctypes.CDLL(str(so))


@pytest.mark.parametrize("typ", ["double", "float", "int"])
def test_reader_and_orders_bit_identical_to_compiled_reference(tmp_path, typ):
    """The CLI's reader and its -r1 / -r2 orderings against the reference's
    own ReadMatrix, matrix2compressed_sortOrder and _skipOrder (util.h:342-357,
    :553-684) compiled from its sources: the same matrix, the same permuted
    matrix and byte-identical CCS/CRS arrays.  Ties in -r1 follow glibc's
    qsort as the reference calls it (a stable merge sort for arrays this
    small: its comparator only ever answers 'greater').  Matrices with an
    all-zero row are left out of -r2: there the reference leaves part of its
    row permutation uninitialised."""
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("reference not compiled (oracle/_ref)")
    shim = _shim(tmp_path)
    np_t = {"double": np.float64, "float": np.float32, "int": np.int32}[typ]
    c_t = {"double": ctypes.c_double, "float": ctypes.c_float, "int": ctypes.c_int}[typ]
    sfx = {"double": "f64", "float": "f32", "int": "i32"}[typ]
    ip = ctypes.POINTER(ctypes.c_int)
    tp = ctypes.POINTER(c_t)
    for name, a in _order_cases():
        n = a.shape[0]
        vals = a.astype(np_t) if typ == "int" else (a * 0.37).astype(np_t)
        f = tmp_path / f"{name}_{n}"
        write_matrix(f, vals.astype(float), typ)
        for generic in (1, 0):
            mats = []
            for lib_, fn in ((ref, f"ref_read_matrix_{sfx}"), (shim, f"shim_read_matrix_{sfx}")):
                m = np.zeros(n * n, dtype=np_t)
                args = [str(f).encode(), generic, m.ctypes.data_as(tp), n * n]
                if lib_ is ref:
                    args.append(ctypes.byref(ctypes.c_int()))
                assert getattr(lib_, fn)(*args) == n
                mats.append(m)
            assert mats[0].tobytes() == mats[1].tobytes(), (name, n, generic)
        m0 = mats[0]
        has_zero_row = not (m0.reshape(n, n) != 0).any(axis=1).all()
        for order in (0, 1, 2):
            if order == 2 and has_zero_row:
                continue
            outs = []
            for lib_, fn in ((ref, f"ref_compressed_order_{sfx}"),
                             (shim, f"shim_compressed_order_{sfx}")):
                m = m0.copy()
                arrs = [np.zeros(n + 1, np.int32), np.zeros(n * n, np.int32),
                        np.zeros(n * n, np_t), np.zeros(n + 1, np.int32),
                        np.zeros(n * n, np.int32), np.zeros(n * n, np_t)]
                ptrs = [x.ctypes.data_as(ip if k in (0, 1, 3, 4) else tp)
                        for k, x in enumerate(arrs)]
                cargs = [order, m.ctypes.data_as(tp), n]
                if lib_ is ref:
                    cargs.append(n * n)  # capacity of the reference's arrays
                stored = getattr(lib_, fn)(*cargs, *ptrs)
                outs.append((stored, m.tobytes(), [x.tobytes() for x in arrs]))
            assert outs[0][0] == outs[1][0], (name, n, order)
            assert outs[0][1] == outs[1][1], (name, n, order, "permuted matrix")
            for k, (x, y) in enumerate(zip(outs[0][2], outs[1][2])):
                assert x == y, (name, n, order, ["cptrs", "rows", "cvals", "rptrs", "cols",
                                                 "rvals"][k])


def test_no_perfect_matching_is_exactly_zero_without_a_walk():
    """A matrix whose nonzero pattern has no perfect matching has permanent
    exactly 0: every entry point returns 0 (zero_pattern set, errors 0,
    certified) from the host check alone -- no device is touched, so this
    runs without a GPU.  The library's check agrees with the oracle's
    matching and with the exact permanent."""
    rng = np.random.default_rng(5)
    found = 0
    for trial in range(300):
        n = int(rng.integers(2, 14))
        a = (rng.random((n, n)) < 0.25) * (1.0 - rng.random((n, n)))
        if O.has_perfect_matching(a):
            continue
        found += 1
        if n <= 9:
            assert O.exact_permanent(a) == 0
        for r in (P.dense(a), P.sparse(a), P.skip(a)):
            assert r.value == 0.0 and r.zero_pattern == 1 and r.certified == 1
            assert r.err_bound == 0.0 and r.err_est == 0.0
    assert found > 30
    # a structurally singular integer matrix through the chunked entry point
    z = np.ones((12, 12), dtype=np.int32)
    z[:, 3] = 0
    r = P.PermResultC()
    assert P.lib().perm_chunked(0, z.ctypes.data, None, None, None, 12, 0, 1, ctypes.byref(r),
                                None) == 0
    assert r.value == 0.0 and r.zero_pattern == 1


def test_lib_connect_exports_the_reference_connector_names():
    """libConnect.so exports exactly the reference connector's three entry
    points (py_connector.c:59, :63, :152) and refuses the approximation IDs
    (NaN, no device call) the way the library's perm_* forms do."""
    path = os.path.join(os.path.dirname(P.LIB_PATH), "libConnect.so")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    names = sorted(line.split()[-1] for line in out.splitlines() if " T " in line)
    assert names == ["connect", "matlab_calculate_return", "read_calculate_return"]
    lib = ctypes.cdll.LoadLibrary(path)
    lib.read_calculate_return.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 5
    lib.read_calculate_return.restype = ctypes.c_double
    for algo in (0, 1, 2, 3):
        assert math.isnan(lib.read_calculate_return(b"/nonexistent", algo, 1, 1, 1, 1))


def test_bench_contract_one_gpu_and_leg_options():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says
    so (before touching a device); its options select the legs."""
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert "measures one GPU" in r.stderr
    assert r.stdout == ""
    h = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert h.returncode == 0
    for opt in ("--steps", "--warmup", "--no-cpu", "--no-mixed", "--batch-count",
                "--serve-count", "--serve-mixes", "--serve-load"):
        assert opt in h.stdout, opt

