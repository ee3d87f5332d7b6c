"""Dense GPU path (reference IDs 0-4, gpu_exact_dense.cu) against the oracle.

Tolerances: results must agree with the exact permanent to 1e-10 relative
and within their stated error (conftest.assert_target); 0/1 matrices with
permanent < 2^53 must round to the exact integer."""
import math

import numpy as np
import pytest

import oracle as O
import permanent_amd as P
from conftest import assert_target

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("n", list(range(1, 21)))
def test_dense_random_vs_exact(n):
    a = P.synthetic_matrix(n, 0.6, seed=100 + n)
    ex = float(O.exact_permanent(a))
    r = P.dense(a, 4)
    assert_target(r, ex, ctx=n)


@pytest.mark.parametrize("algo", [0, 1, 2, 3, 4])
def test_dense_all_ids_same_value(algo):
    a = P.synthetic_matrix(14, 0.5, seed=7)
    ex = float(O.exact_permanent(a))
    v = P.dropin_dense(a, algo)
    assert rel(v, ex) < 1e-10


@pytest.mark.parametrize("n,density", [(8, 0.5), (12, 0.5), (16, 0.5), (20, 0.4), (24, 0.3)])
def test_binary_rounds_to_exact_integer(n, density):
    """0/1 matrices whose permanent is below 2^53 come out as the exact
    integer (automatic double-double product pass when FP64 cannot
    guarantee it)."""
    a = P.synthetic_matrix(n, density, seed=300 + n, binary=True)
    ex = O.exact_permanent(a)
    assert ex.denominator == 1 and ex.numerator < 2 ** 53
    r = P.dense(a.astype(np.int32), 1)
    assert r.value == float(ex.numerator), (r, ex)


def test_binary_above_2_53_relative():
    a = P.synthetic_matrix(24, 0.5, seed=324, binary=True)
    ex = O.exact_permanent(a)
    assert ex > 2 ** 53
    r = P.dense(a.astype(np.int32), 1)
    assert abs(r.value - float(ex)) <= max(1e-12 * float(ex), r.stated)


@pytest.mark.parametrize("n", [5, 9, 12])
def test_closed_forms(n):
    J = np.ones((n, n))
    assert P.dense(J).value == pytest.approx(math.factorial(n), rel=1e-13)
    I = np.eye(n)
    assert P.dense(I).value == pytest.approx(1.0, rel=1e-13)
    D = J - I  # derangements
    d = round(math.factorial(n) * sum((-1) ** k / math.factorial(k) for k in range(n + 1)))
    assert P.dense(D).value == pytest.approx(d, rel=1e-13)


def test_grid_graph_tilings():
    for m, k in [(4, 4), (4, 6), (6, 6)]:
        G = O.grid_graph_matrix(m, k)
        assert round(P.dense(G.astype(np.int32)).value) == O.domino_tilings(m, k)


def test_integer_exact_in_fp64_without_second_pass():
    """Small-entry integer matrices whose products and sums stay below 2^53
    (the bound of integer_sum_exact) are exact in FP64: err_bound 0, mode 0."""
    for n, d, seed in ((12, 0.4, 1), (14, 0.3, 2), (16, 0.25, 3)):
        a = P.synthetic_matrix(n, d, seed=seed, binary=True)
        r = P.dense(a)
        assert r.value == float(O.exact_permanent(a)) and r.err_bound == 0.0 and r.mode == 0
