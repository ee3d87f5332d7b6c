"""Extreme-subset cut of the FP64 walk.

The subsets with |S| <= k or |S| >= n-1-k are summed by a double-double
kernel with exact X; the FP64 walk runs in two passes (lanes that can meet
such a subset, with a per-step test, and lanes that cannot).  Checked here:

* forced cuts k = 0..4 at small n and several splits (side pass with and
  without free bits, main pass empty) against the exact oracle;
* the same matrix with and without the cut within the stated errors;
* a cut on a matrix with mixed signs (no extreme structure);
* sparse and dense paths still give the same bits under the cut;
* the statistical error estimate covers the observed error."""
import numpy as np
import pytest

import oracle as O
import permanent_amd as P

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / abs(b)


CASES = [(8, 0, 3), (8, 2, 3), (12, 1, 4), (12, 3, 5), (16, 2, 6), (16, 4, 3), (20, 3, 8),
         (20, 4, 12), (22, 2, 10), (22, 4, 19)]


@pytest.mark.parametrize("n,k,walk_bits", CASES)
def test_forced_cut_vs_exact(monkeypatch, n, k, walk_bits):
    a = P.synthetic_matrix(n, 0.7, seed=500 + 10 * n + k)
    ex = float(O.exact_permanent(a))
    monkeypatch.setenv("PERMAN_EXT_K", str(k))
    monkeypatch.setenv("PERMAN_WALK_BITS", str(walk_bits))
    r = P.dense(a, 4)
    assert r.ext_k == k and r.mode == 0
    assert rel(r.value, ex) < 1e-12, (n, k, walk_bits, r.value, ex)
    assert abs(r.value - ex) <= r.err_bound


@pytest.mark.parametrize("n", [13, 17, 21])
def test_cut_on_mixed_signs(monkeypatch, n):
    rng = np.random.default_rng(40 + n)
    a = rng.uniform(-1.0, 1.0, size=(n, n))
    ex = float(O.exact_permanent(a))
    monkeypatch.setenv("PERMAN_EXT_K", "3")
    r = P.dense(a, 4)
    assert r.ext_k == 3
    assert abs(r.value - ex) <= max(1e-11 * abs(ex), r.err_bound), (r.value, ex)


def test_auto_cut_vs_no_cut_n28(monkeypatch):
    a = P.synthetic_matrix(28, 0.5, seed=28)
    plan = P.Plan(a)
    ref = plan.run(2)
    cut = plan.run(0)
    plan.close()
    monkeypatch.setenv("PERMAN_EXT_K", "-1")
    plan = P.Plan(a)
    plain = plan.run(0)
    plan.close()
    assert cut.ext_k > 0 and plain.ext_k == -1
    assert rel(cut.value, ref.value) < 1e-12
    assert rel(plain.value, ref.value) < 1e-10
    assert abs(cut.value - ref.value) <= cut.err_bound


def test_sparse_and_dense_same_bits_under_cut():
    a = P.synthetic_matrix(26, 0.4, seed=77)
    d = P.dense(a, 4)
    s = P.sparse(a, algo_id=4)
    assert d.ext_k >= 0
    assert (d.value, d.sum_hi, d.sum_lo) == (s.value, s.sum_hi, s.sum_lo)


@pytest.mark.parametrize("n,density", [(18, 0.5), (20, 0.9), (22, 0.3), (24, 0.6), (25, 0.5)])
def test_error_estimate_covers_observed(n, density):
    a = P.synthetic_matrix(n, density, seed=900 + n)
    ex = float(O.exact_permanent(a))
    r = P.dense(a, 4)
    assert abs(r.value - ex) <= r.err_est, (n, r.value, ex, r.err_est)
    assert r.err_est <= r.err_bound
