"""Sparse GPU path (-s, reference IDs 1-4, gpu_exact_sparse.cu:673-914).

Reference semantics kept: x is initialised from the DENSE matrix (every
entry, gpu_exact_sparse.cu:861-871), the column updates come from the CCS,
which holds only entries > 0 (util.h:537-544)."""
import json
import os

import numpy as np
import pytest

import oracle as O
import permanent_amd as P
from conftest import GOLDEN, assert_target

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [2, 5, 9, 13, 17, 20])
@pytest.mark.parametrize("algo", [1, 2, 3, 4])
def test_sparse_matches_exact(n, algo):
    a = P.synthetic_matrix(n, 0.3, seed=600 + n, full_support=True)
    ex = float(O.exact_permanent(a))
    r = P.sparse(a, algo_id=algo)
    assert_target(r, ex, ctx=(n, algo))


@pytest.mark.parametrize("algo", [1, 2, 3, 4])
def test_sparse_dropins_all_types(algo):
    a = P.synthetic_matrix(12, 0.35, seed=31, full_support=True)
    ex64 = float(O.exact_permanent(a))
    assert P.dropin_sparse(a, algo) == pytest.approx(ex64, rel=1e-10)
    f = a.astype(np.float32)
    assert P.dropin_sparse(f, algo) == pytest.approx(float(O.exact_permanent(f)), rel=1e-10)
    i = (a * 4).astype(np.int32)
    assert P.dropin_sparse(i, algo) == float(O.exact_permanent(i))


@pytest.mark.parametrize("n", [8, 14])
def test_sparse_with_nonpositive_entries_follows_reference_semantics(n):
    """Entries <= 0 stay in the x-initialisation but are not in the CCS; the
    result then differs from the permanent and must match the reference's
    own double-X sparse evaluator (restated: gpu_exact_sparse.cu:6-87)."""
    a = P.synthetic_matrix(n, 0.7, seed=900 + n) - 0.3 * (np.arange(n)[:, None] % 3 == 0)
    want = O.cpu_perman64_sparse(a, threads=2)
    got = P.sparse(a, algo_id=4).value
    assert got == pytest.approx(want, rel=1e-9)
    assert got != pytest.approx(float(O.exact_permanent(a)), rel=1e-6)


def test_sparse_equals_dense_bitwise_on_nonnegative():
    """For a nonnegative matrix the CCS column store equals the dense one, so
    both paths run the same kernel on the same data: identical bits."""
    a = P.synthetic_matrix(22, 0.25, seed=5, full_support=True)
    assert P.sparse(a, algo_id=4).value == P.dense(a, 4).value


def test_sparse_binary_integer_exact():
    a = P.synthetic_matrix(20, 0.25, seed=8, binary=True, full_support=True)
    ex = O.exact_permanent(a)
    r = P.sparse(a.astype(np.int32), algo_id=1)
    assert round(r.value) == ex.numerator and abs(r.value - ex.numerator) < 0.5


def test_golden_sparse_fixtures():
    for c in json.load(open(os.path.join(GOLDEN, "small_exact.json"))):
        a = np.array(c["matrix"])
        if c["n"] < 4 or (a < 0).any():
            continue
        r = P.sparse(a, algo_id=4)
        ex = c["value"]
        assert_target(r, ex, ctx=c["name"])
