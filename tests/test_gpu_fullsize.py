"""Full-size checks at the benchmark configurations, where no CPU oracle
finishes in seconds.  They rest on known answers and on properties that do
not depend on size:

* domino tilings of the reference's grid graphs (8x9 -> n=36, 8x10 -> n=40),
  exact integers through the 0/1 path;
* the FP64 kernel against the same sum with X and products in
  double-double (mode 2: no X drift, ~1e-30 relative per operation);
* exact doubling when one row is multiplied by 2 (every rounding scales);
* invariance under row/column permutation and transposition (different
  rounding, same permanent) within the stated error;
* the sparse path against the dense path (same column store: same bits).
Tolerance for FP64 matrices: 1e-10 relative (the target), and never more
than the stated error bound."""
import json
import os

import numpy as np
import pytest

import oracle as O
import permanent_amd as P
from conftest import GOLDEN, cli_result

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / abs(b)


@pytest.fixture(scope="module")
def c2():
    """n=40, FP64, density 0.5, U(0,1], seed 2 (the benchmark matrix)."""
    return P.synthetic_matrix(40, 0.5, seed=2)


def test_c2_fp64_vs_double_double(c2):
    plan = P.Plan(c2)
    fast = plan.run(0)
    ref = plan.run(2)
    plan.close()
    print(f"n=40 fast={fast.value!r} dd={ref.value!r} rel={rel(fast.value, ref.value):.3e} "
          f"stated={fast.stated / abs(ref.value):.3e} cond={fast.abs_sum / abs(ref.value):.3e}")
    assert rel(fast.value, ref.value) < 1e-10
    assert abs(fast.value - ref.value) <= fast.stated


def test_c2_binary_variant_vs_double_double_products():
    """The 0/1 variant of the benchmark configuration (n=40, density 0.5,
    seed 3): X is exact, so double-double products (mode 1) leave only
    ~1e-30 relative rounding per operation."""
    b = P.synthetic_matrix(40, 0.5, seed=3, binary=True).astype(np.int32)
    plan = P.Plan(b)
    fast = plan.run(0)
    ref = plan.run(1)
    plan.close()
    print(f"n=40 0/1 fast={fast.value!r} dd={ref.value!r} rel={rel(fast.value, ref.value):.3e} "
          f"shape=({fast.nu},{fast.nu1},{fast.nu2})")
    assert fast.integral == 1 and ref.value == round(ref.value)
    assert rel(fast.value, ref.value) < 1e-10
    assert abs(fast.value - ref.value) <= fast.stated


def test_grid_8x10_n40_exact_tilings():
    G = O.grid_graph_matrix(8, 10).astype(np.int32)
    assert G.shape == (40, 40)
    r = P.dense(G, 4)
    assert r.value == 1031151241.0, r


def test_grid_8x9_n36_sparse_and_dense_exact_tilings():
    G = O.grid_graph_matrix(8, 9).astype(np.int32)
    assert G.shape == (36, 36)
    want = float(O.domino_tilings(8, 9))
    assert P.sparse(G, algo_id=4).value == want
    assert P.dense(G, 1).value == want


@pytest.fixture(scope="module")
def c3():
    """n=36, density 0.2, every row/column nonzero, U(0,1], seed 4."""
    return P.synthetic_matrix(36, 0.2, seed=4, full_support=True)


def test_c3_sparse_fp64(c3):
    s = P.sparse(c3, algo_id=4)
    d = P.dense(c3, 4)
    assert s.value == d.value  # nonnegative: identical column store
    plan = P.Plan(c3, sparse_path=True)
    ref = plan.run(2)
    plan.close()
    print(f"n=36 sparse={s.value!r} dd={ref.value!r} rel={rel(s.value, ref.value):.3e}")
    assert rel(s.value, ref.value) < 1e-10
    assert abs(s.value - ref.value) <= s.stated


def test_c3_sparse_binary_exact(c3):
    b = (c3 != 0).astype(np.int32)  # the -b variant: all ones
    s = P.sparse(b, algo_id=4)
    plan = P.Plan(b)
    d = plan.run(1)  # double-double products: exact for an integer matrix
    plan.close()
    assert s.integral == 1 and d.value == round(d.value)
    assert abs(s.value - d.value) <= s.stated
    if abs(d.value) < 2 ** 53:
        assert s.value == d.value


def test_row_doubling_is_exact():
    a = P.synthetic_matrix(36, 0.5, seed=21)
    v = P.dense(a).value
    b = a.copy()
    b[7] *= 2.0
    assert P.dense(b).value == 2.0 * v


def test_permutation_and_transpose_invariance():
    a = P.synthetic_matrix(34, 0.5, seed=22)
    r0 = P.dense(a)
    g = np.random.default_rng(1)
    p = a[g.permutation(34)][:, g.permutation(34)]
    for m in (p, a.T):
        r = P.dense(np.ascontiguousarray(m))
        assert abs(r.value - r0.value) <= r0.stated + r.stated
        assert rel(r.value, r0.value) < 1e-10


def test_c1_configuration_on_gpu():
    c = json.load(open(os.path.join(GOLDEN, "c1_exact.json")))
    a = np.array(c["matrix"])
    r = P.dense(a, 4)
    assert rel(r.value, c["value"]) < 1e-12


def test_reference_data_files_n30():
    path = os.path.join(GOLDEN, "refdata.json")
    if not os.path.exists(path):
        pytest.skip("refdata fixture not generated")
    for rec in json.load(open(path)):
        tmp = os.path.join(os.environ.get("TMPDIR", "/tmp"), rec["file"].replace("/", "_"))
        with open(tmp, "w") as fh:
            fh.write(rec["text"])
        mat, n, _, typ = O.read_matrix(tmp)
        r = P.dense(mat, 4)
        want = rec["dd_hi"] + rec["dd_lo"]
        print(f"{rec['file']}: gpu={r.value!r} dd={want!r} rel={rel(r.value, want):.2e} "
              f"perman64(ref double-X)={rec['perman64']!r}")
        if typ == "int":
            want = float(int(rec["rounded"]))
        assert rel(r.value, want) < 1e-10
        assert abs(r.value - want) <= r.stated


def test_cli_on_reference_data_files(tmp_path):
    """`perman -f <reference data file>` (dense -p4 and sparse -s -p4): the
    reference's own n = 30 inputs read by the CLI's reader, the Result line
    against the double-double oracle."""
    import re

    from test_capi import run_cli

    path = os.path.join(GOLDEN, "refdata.json")
    if not os.path.exists(path):
        pytest.skip("refdata fixture not generated")
    for rec in json.load(open(path)):
        f = tmp_path / rec["file"].replace("/", "_")
        f.write_text(rec["text"])
        assert all(float(line.split()[2]) > 0 for line in rec["text"].splitlines()[1:] if line)
        want = rec["dd_hi"] + rec["dd_lo"]
        if rec["type"] == "int":
            want = float(int(rec["rounded"]))
        for args in (("-p", "4"), ("-s", "-p", "4")):
            r = run_cli("-f", str(f), "-g", *args)
            assert r.returncode == 0, (rec["file"], args, r.stderr)
            v = cli_result(r.stdout)[1]
            print(f"{rec['file']} {' '.join(args)}: {v!r} (oracle {want!r})")
            # (the sparse path keeps only entries > 0, as the reference's CCS;
            # these files list positive entries only, so both paths agree)
            assert rel(v, want) < 1e-10, (rec["file"], args, v, want)


def test_largest_sizes_exact_tilings():
    """The largest orders: the dense path at n = 42 (grid 6x14, 2^41 subsets)
    and the skipper at n = 48 = kMaxDenseN (grid 8x12), against the domino
    tiling counts (Kasteleyn's product formula)."""
    G42 = O.grid_graph_matrix(6, 14).astype(np.int32)
    r = P.dense(G42, 4)
    print(f"dense n=42: {r.value!r} mode={r.mode} kernel_ms={r.kernel_ms:.1f}")
    assert r.value == float(O.domino_tilings(6, 14))
    G48 = O.grid_graph_matrix(8, 12).astype(np.int32)
    s = P.skip(G48)
    print(f"skipper n=48: {s.value!r} walked {s.main_subsets / s.subsets:.2e} "
          f"kernel_ms={s.kernel_ms:.1f}")
    assert s.value == float(O.domino_tilings(8, 12))


def test_n40_block_ranges_against_cpu_double_double():
    """The C2 matrix (n = 40) pinned independently of the GPU's own modes:
    workgroup ranges of the natural-order walk (first, middle, last) summed
    by the GPU in FP64 and in mode 2 against a CPU double-double evaluation
    of exactly the same subsets (oracle.lane_range_dd; 2^18 subsets per
    workgroup).  Mode 2 agrees to 1e-13 of the range's sum of |terms|; FP64
    within its stated error (the range's err_est is scaled by the same
    moments)."""
    import os

    a = P.synthetic_matrix(40, 0.5, 2)
    os.environ["PERMAN_PAIR"] = "0"  # natural walk order (no row/column reordering)
    try:
        plan = P.Plan(a)
    finally:
        del os.environ["PERMAN_PAIR"]
    w, h, blocks = plan.geometry()
    assert (w, h, blocks) == (10, 29, 1 << 21)
    for b0 in (0, blocks // 2 + 12345, blocks - 1):
        fp = plan.run_blocks(b0, b0 + 1, mode=0)
        dd = plan.run_blocks(b0, b0 + 1, mode=2)
        hi, lo = O.lane_range_dd(a, w, 256 * b0, 256 * (b0 + 1), 8)
        ref = hi + lo
        a_sum = fp.abs_sum / 2.0  # |scale| = 2
        print(f"block {b0}: fp={fp.sum_hi!r} dd={dd.sum_hi!r} cpu={ref!r} |terms|={a_sum:.3e}")
        assert abs((dd.sum_hi + dd.sum_lo) - ref) <= 1e-13 * a_sum, b0
        assert abs((fp.sum_hi + fp.sum_lo) - ref) <= fp.err_est / 2.0 + 1e-13 * a_sum, b0
    plan.close()
