"""Nested walk shapes of the FP64 kernel (Gray columns 0, 1 and 2 restricted
to row prefixes [0, K_l), tail products shared): every shape the host picks
must give the permanent of the exact oracle, and the same value (within the
stated errors) as the pair-only (PERMAN_NEST=0) and natural-order
(PERMAN_PAIR=0) walks."""
import numpy as np
import pytest

import oracle as O
import permanent_amd as P
from conftest import assert_target, oracle_value

pytestmark = pytest.mark.gpu


def run_with(monkeypatch, env, fn):
    with monkeypatch.context() as m:
        for k, v in env.items():
            m.setenv(k, v)
        return fn()


@pytest.mark.parametrize("n,density,seed", [(24, 0.2, 1), (24, 0.35, 2), (25, 0.5, 3),
                                            (26, 0.3, 4), (25, 0.7, 5)])
def test_nested_vs_exact(n, density, seed):
    a = P.synthetic_matrix(n, density, seed, full_support=True)
    ex = oracle_value(a)
    for sparse in (False, True):
        r = P.sparse(a) if sparse else P.dense(a)
        print(f"n={n} d={density} shape=({r.nu},{r.nu1},{r.nu2}) rel={abs(r.value - ex) / abs(ex):.2e}")
        assert_target(r, ex, ctx=(n, density, sparse))


def test_every_shape_family_appears_and_agrees(monkeypatch):
    """A sweep of densities at n = 30 and 34 reaches pair shapes and both
    nested families; each value matches the natural-order walk."""
    seen = set()
    for n in (30, 34):
        for k, density in enumerate((0.15, 0.25, 0.35, 0.5, 0.65, 0.8)):
            a = P.synthetic_matrix(n, density, 40 + k, full_support=True)
            r = P.dense(a)
            plain = run_with(monkeypatch, {"PERMAN_PAIR": "0"}, lambda: P.dense(a))
            pair = run_with(monkeypatch, {"PERMAN_NEST": "0"}, lambda: P.dense(a))
            assert plain.nu == n and pair.nu1 == n and pair.nu2 == n
            fam = "plain" if r.nu == n else "pair" if r.nu1 == n else \
                  "nest2" if r.nu2 == n else "nest3"
            seen.add(fam)
            tol = r.stated + plain.stated
            assert abs(r.value - plain.value) <= tol, (n, density, r, plain)
            assert abs(r.value - pair.value) <= r.stated + pair.stated
            assert abs(r.value - plain.value) <= 1e-11 * abs(plain.value) + r.err_est + plain.err_est
    print("families:", sorted(seen))
    assert {"pair", "nest3"} <= seen


def test_integer_matrices_exact_under_every_shape(monkeypatch):
    rng = np.random.default_rng(3)
    for n in (24, 28):
        a = ((rng.random((n, n)) < 0.3) * rng.integers(1, 3, (n, n))).astype(np.int32)
        ex = oracle_value(a)
        if ex >= 2.0 ** 53:
            continue
        for env in ({}, {"PERMAN_NEST": "0"}, {"PERMAN_PAIR": "0"}):
            r = run_with(monkeypatch, env, lambda: P.dense(a))
            assert r.value == ex, (n, env, r.value, ex)


def test_plan_and_chunked_follow_the_shape():
    a = P.synthetic_matrix(32, 0.4, seed=32)
    one = P.dense(a)
    plan = P.Plan(a)
    r = plan.run()
    plan.close()
    assert (r.value, r.nu, r.nu1, r.nu2) == (one.value, one.nu, one.nu1, one.nu2)
    c, _ = P.chunked(a, use_cpu=False)
    assert (c.sum_hi, c.sum_lo) == (one.sum_hi, one.sum_lo)


@pytest.mark.parametrize("n,density,seed", [(24, 0.3, 61), (26, 0.25, 62), (25, 0.5, 63)])
def test_double_double_products_on_nested_shapes(n, density, seed):
    """Mode 1 (double-double products) runs the nested shape too; for an
    integer matrix it is exact, and for a real one it agrees with FP64."""
    b = P.synthetic_matrix(n, density, seed, full_support=True, binary=True).astype(np.int32)
    plan = P.Plan(b)
    r = plan.run(1)
    plan.close()
    assert r.mode == 1 and r.nu < n, r
    assert r.value == float(O.exact_permanent(b))
    a = P.synthetic_matrix(n, density, seed, full_support=True)
    plan = P.Plan(a)
    f, d = plan.run(0), plan.run(1)
    plan.close()
    assert (d.nu, d.nu1, d.nu2) == (f.nu, f.nu1, f.nu2)
    assert abs(d.value - f.value) <= f.stated + d.stated
