"""Pins of the CPU oracle (no GPU needed).

The reference ships no tests and no expected outputs, so the oracle is pinned
by mathematics (brute force over permutations, closed forms, the domino
tiling counts of the reference's own grid-graph generator), by the compiled
reference algorithms where they build (oracle/_ref), and by the committed
golden fixtures."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN


def rng_matrix(n, seed, density=0.7, signed=False):
    r = np.random.default_rng(seed)
    a = r.random((n, n)) * (r.random((n, n)) < density)
    return a - 0.5 * (a != 0) if signed else a


# --- exact oracle ----------------------------------------------------------
@pytest.mark.parametrize("n", range(1, 9))
@pytest.mark.parametrize("kind", ["dense", "sparse", "signed", "int"])
def test_exact_matches_brute_force(n, kind):
    if kind == "int":
        a = np.random.default_rng(n).integers(-3, 4, size=(n, n)).astype(float)
    else:
        a = rng_matrix(n, 10 * n + len(kind), 0.4 if kind == "sparse" else 0.9,
                       signed=(kind == "signed"))
    assert O.exact_permanent(a) == O.brute_permanent(a)


@pytest.mark.parametrize("n", [1, 5, 11, 16])
def test_exact_closed_forms(n):
    J = np.ones((n, n))
    assert O.exact_permanent(J) == math.factorial(n)
    assert O.exact_permanent(np.eye(n)) == 1
    d = round(math.factorial(n) * sum((-1) ** k / math.factorial(k) for k in range(n + 1)))
    assert O.exact_permanent(J - np.eye(n)) == d  # derangements
    t = np.triu(rng_matrix(n, n) + 0.25)
    assert O.exact_permanent(t) == math.prod(Fraction(float(v)) for v in np.diag(t))
    z = rng_matrix(n, n + 1)
    z[n // 2, :] = 0
    assert O.exact_permanent(z) == 0


def test_exact_invariants():
    a = rng_matrix(9, 3)
    p = O.exact_permanent(a)
    r = np.random.default_rng(0)
    assert O.exact_permanent(a[r.permutation(9)][:, r.permutation(9)]) == p
    assert O.exact_permanent(a.T) == p
    b = a.copy()
    b[4] *= 2.0  # power-of-two scaling of a row is exact
    assert O.exact_permanent(b) == 2 * p
    u, v = a.copy(), a.copy()  # row 4 split into two disjoint-support rows
    u[4, ::2] = 0
    v[4, 1::2] = 0
    assert O.exact_permanent(u) + O.exact_permanent(v) == p
    blk = np.zeros((9, 9))
    blk[:4, :4] = a[:4, :4]
    blk[4:, 4:] = a[4:, 4:]
    assert O.exact_permanent(blk) == O.exact_permanent(a[:4, :4]) * O.exact_permanent(a[4:, 4:])


@pytest.mark.parametrize("m,n", [(2, 2), (2, 5), (3, 4), (4, 4), (4, 5), (5, 6), (6, 6)])
def test_grid_graph_generator_counts_domino_tilings(m, n):
    """The reference's grid generator (util.h:403-520) restated: its
    permanent equals the domino tilings (transfer matrix) and Kasteleyn's
    product formula."""
    G = O.grid_graph_matrix(m, n)
    t = O.domino_tilings(m, n)
    kas = 1.0
    for j in range(1, (m + 1) // 2 + 1):
        for k in range(1, (n + 1) // 2 + 1):
            kas *= 4 * math.cos(math.pi * j / (m + 1)) ** 2 + 4 * math.cos(math.pi * k / (n + 1)) ** 2
    assert round(kas) == t
    assert O.exact_permanent(G) == t


def test_domino_known_values():
    assert O.domino_tilings(8, 8) == 12988816
    assert O.domino_tilings(8, 9) == 108435745
    assert O.domino_tilings(8, 10) == 1031151241


# --- restated reference evaluators -----------------------------------------
@pytest.mark.parametrize("n", [2, 5, 9, 13])
def test_double_x_restatements_accurate(n):
    a = rng_matrix(n, 77 + n)
    ex = float(O.exact_permanent(a))
    assert O.perman64(a) == pytest.approx(ex, rel=1e-12)
    for threads in (1, 3, 7):
        assert O.cpu_perman64(a, threads) == pytest.approx(ex, rel=1e-12)
    assert O.cpu_perman64_sparse(a, 4) == pytest.approx(ex, rel=1e-12)
    hi, lo = O.ryser_dd(a, threads=3)
    assert abs((hi + lo) - ex) <= 1e-15 * abs(ex)


@pytest.mark.parametrize("n", [3, 6, 9])
@pytest.mark.parametrize("algo", [0, 1, 2, 3, 4])
def test_floatx_gpu_semantics_exact_on_small_binary(n, algo):
    """Half-integer X and small products are exact in float/double, so the
    float-X semantics of every reference GPU kernel must give the exact
    permanent of a small 0/1 matrix (catches chunking/sign/index slips)."""
    a = (np.random.default_rng(n + algo).random((n, n)) < 0.6).astype(np.int32)
    ex = O.exact_permanent(a)
    assert O.gpu_dense_floatx(a, algo, grid_dim=4, block_dim=8) == float(ex)
    assert O.gpu_dense_floatx(a.astype(np.float64), algo, grid_dim=3, block_dim=5) == float(ex)
    assert O.parallel_perman64(a, 3) == float(ex)


def test_floatx_loses_accuracy_on_real_matrices():
    """Information: float-X (the reference's GPU kernels and CPU baseline)
    drifts far beyond 1e-10 on an FP64 matrix; FP64-X does not."""
    a = rng_matrix(18, 5)
    ex = float(O.exact_permanent(a))
    fx = O.gpu_dense_floatx(a, 4, grid_dim=16, block_dim=16)
    assert abs(fx - ex) / abs(ex) > 1e-10
    assert abs(O.perman64(a) - ex) / abs(ex) < 1e-12


def _ref():
    R = O.ref_lib()
    if R is None:
        pytest.skip("reference algorithms not built (needs the reference checkout)")
    return R


def _p(a, ct):
    import ctypes
    return a.ctypes.data_as(ctypes.POINTER(ct))


@pytest.mark.parametrize("n", [3, 8, 12, 15])
def test_restatement_matches_compiled_reference(n):
    import ctypes
    R = _ref()
    a = np.ascontiguousarray(rng_matrix(n, 200 + n, 0.6))
    assert R.ref_perman64_f64(_p(a, ctypes.c_double), n) == O.perman64(a)
    for dt, ct, sfx in ((np.float64, ctypes.c_double, "f64"), (np.float32, ctypes.c_float, "f32")):
        b = np.ascontiguousarray((a * 4).astype(dt))
        ref = getattr(R, f"ref_parallel_perman64_{sfx}")(_p(b, ct), n, 1)
        assert ref == O.parallel_perman64(b, 1)
    i = np.ascontiguousarray((a * 5).astype(np.int32))
    assert R.ref_parallel_perman64_i32(_p(i, ctypes.c_int), n, 1) == O.parallel_perman64(i, 1)
    cp, rows, cv, _, _, _ = O.matrix2compressed(a)
    ref_s = R.ref_parallel_perman64_sparse_f64(_p(a, ctypes.c_double), _p(cp, ctypes.c_int),
                                               _p(rows, ctypes.c_int), _p(cv, ctypes.c_double),
                                               n, 1)
    assert ref_s == O.parallel_perman64_sparse(a, 1)


def test_ccs_builder_matches_compiled_reference():
    import ctypes
    R = _ref()
    a = np.ascontiguousarray(rng_matrix(11, 3, 0.5, signed=True))
    n, nnz = 11, int((a > 0).sum())
    cp = np.zeros(n + 1, np.int32); rp = np.zeros(n + 1, np.int32)
    rows = np.zeros(nnz, np.int32); cols = np.zeros(nnz, np.int32)
    cv = np.zeros(nnz); rv = np.zeros(nnz)
    got = R.ref_matrix2compressed_f64(_p(a, ctypes.c_double), n, nnz, _p(cp, ctypes.c_int),
                                      _p(rows, ctypes.c_int), _p(cv, ctypes.c_double),
                                      _p(rp, ctypes.c_int), _p(cols, ctypes.c_int),
                                      _p(rv, ctypes.c_double))
    assert got == nnz
    mine = O.matrix2compressed(a)
    for x, y in zip(mine, (cp, rows, cv, rp, cols, rv)):
        np.testing.assert_array_equal(x, y)


# --- golden fixtures --------------------------------------------------------
def test_golden_small_exact_fixtures():
    cases = json.load(open(os.path.join(GOLDEN, "small_exact.json")))
    assert len(cases) > 60
    for c in cases:
        if c["n"] > 14:
            continue  # keep the CPU suite quick; larger ones run on the GPU side
        a = np.array(c["matrix"], dtype=np.float64)
        assert O.exact_permanent(a) == Fraction(c["exact"]), c["name"]


def test_golden_c1_fixture():
    c = json.load(open(os.path.join(GOLDEN, "c1_exact.json")))
    a = np.array(c["matrix"])
    assert a.shape == (20, 20)
    import permanent_amd as P
    np.testing.assert_array_equal(a, P.synthetic_matrix(20, 0.5, 1))
    # the FP64 double-X restatement lands within 1e-10 of the exact value
    assert abs(O.cpu_perman64(a, 8) - float(Fraction(c["exact"]))) <= 1e-10 * abs(c["value"])


def test_reader_on_reference_data_copy(tmp_path):
    path = os.path.join(GOLDEN, "refdata.json")
    if not os.path.exists(path):
        pytest.skip("refdata fixture not generated")
    for rec in json.load(open(path)):
        f = tmp_path / rec["file"].replace("/", "_")
        f.write_text(rec["text"])
        mat, n, nnz, typ = O.read_matrix(str(f))
        assert (n, typ) == (rec["n"], rec["type"])
        assert (mat != 0).sum() <= nnz
        if rec["type"] == "int":
            # the double-double oracle of an integer matrix is an integer
            assert abs(rec["dd_hi"] + rec["dd_lo"] - float(int(rec["rounded"]))) < 1.0


def test_oracle_cache_matches_its_matrices_and_recomputes():
    """tests/golden/oracle_cache.json: one entry per matrix of
    make_golden.cache_cases(), and the cheap entries recompute exactly."""
    import importlib.util

    from conftest import GOLDEN, matrix_key, oracle_value

    spec = importlib.util.spec_from_file_location("make_golden",
                                                  os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    cache = json.load(open(os.path.join(GOLDEN, "oracle_cache.json")))
    mats = mg.cache_cases()
    assert {matrix_key(m) for m in mats} == set(cache)
    cheap = [m for m in mats if m.shape[0] <= 20]
    assert cheap
    for m in cheap[:3]:
        assert oracle_value(m) == float(O.exact_permanent(m))


def test_perfect_matching_check_against_brute_force():
    """oracle.has_perfect_matching is False exactly when the 0/1 pattern's
    permanent (brute force over n! permutations) is 0."""
    rng = np.random.default_rng(123)
    seen = {True: 0, False: 0}
    for trial in range(400):
        n = int(rng.integers(1, 8))
        p = (rng.random((n, n)) < rng.uniform(0.15, 0.6)).astype(float)
        has = O.has_perfect_matching(p)
        assert has == (O.brute_permanent(p) != 0), p
        seen[has] += 1
    assert seen[True] > 50 and seen[False] > 50


def test_c4_mix_fixture_pins():
    """tests/golden/c4_mix.json: its keys are the members of the batched
    configuration's mix as the generator builds them today; every member
    marked zero has no perfect matching (and the rest do); a sample of the
    exact entries is recomputed with the exact oracle and the double-double
    entries of a few small members with the double-double oracle."""
    import json

    from conftest import GOLDEN, matrix_key

    import permanent_amd as P

    path = os.path.join(GOLDEN, "c4_mix.json")
    if not os.path.exists(path):
        pytest.skip("c4_mix.json not generated")
    members = json.load(open(path))["members"]
    mats = P.mixed_batch(4096, 12, 28, 0.2, 0.8, 5)
    assert len(members) == sum(m.shape[0] <= 26 for m in mats)
    checked_exact = checked_dd = 0
    for rec in members:
        m = mats[rec["i"]]
        assert matrix_key(m) == rec["key"] and m.shape[0] == rec["n"] <= 26
        assert (rec["kind"] == "zero") == (not O.has_perfect_matching(m)), rec["i"]
        if rec["kind"] == "exact" and rec["n"] <= 13 and checked_exact < 40:
            ex = O.exact_permanent(m)
            assert float(ex) == rec["hi"] and float(ex - Fraction(rec["hi"])) == rec["lo"]
            checked_exact += 1
        if rec["kind"] == "dd" and rec["n"] <= 18 and checked_dd < 10:
            hi, lo = O.ryser_dd(m, 4)
            assert abs((hi + lo) - (rec["hi"] + rec["lo"])) <= 1e-20 * abs(hi)
            checked_dd += 1
    assert checked_exact == 40 and checked_dd == 10


def test_lane_range_oracle_sums_to_the_permanent():
    """oracle.lane_range_dd (the GPU walk's lane decomposition, natural
    order): over all lanes it gives the exact permanent (times the scale
    1/(4(n&1)-2)) for every walk length, and ranges add up."""
    import permanent_amd as P

    for n in (6, 9, 12):
        a = P.synthetic_matrix(n, 0.6, seed=n)
        sc = 4 * (n & 1) - 2
        ex = float(O.exact_permanent(a))
        for w in range(1, n - 1):
            h = n - 1 - w
            hi, lo = O.lane_range_dd(a, w, 0, 1 << h, 4)
            assert abs(sc * (hi + lo) - ex) <= 1e-14 * abs(ex), (n, w)
            if h >= 2:
                m = (1 << h) // 3
                h1, l1 = O.lane_range_dd(a, w, 0, m, 4)
                h2, l2 = O.lane_range_dd(a, w, m, 1 << h, 4)
                assert abs(sc * (h1 + l1 + h2 + l2) - ex) <= 1e-14 * abs(ex), (n, w, m)
