"""Sparse skipper (ID 7): zero-run skipping on integer matrices.

* grid-graph matrices of the reference's generator (util.h:403-520):
  exact domino-tiling counts, most of the walk skipped;
* random sparse 0/1 and small-integer matrices against the exact oracle and
  the plain sparse path;
* non-integer matrices and n < 16 fall back to the sparse path (same sum);
* the drop-in entry point for each element type."""
import numpy as np
import pytest

import oracle as O
import permanent_amd as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,cols", [(6, 6), (6, 8), (8, 8), (8, 9), (8, 10)])
def test_grid_graphs_exact(rows, cols):
    G = O.grid_graph_matrix(rows, cols).astype(np.int32)
    want = float(O.domino_tilings(rows, cols))
    r = P.skip(G)
    assert r.value == want, (rows, cols, r.value, want)
    frac = r.main_subsets / r.subsets
    print(f"grid {rows}x{cols} n={G.shape[0]}: walked {frac:.4%} of the subsets, "
          f"{r.kernel_ms:.2f} ms")
    if G.shape[0] >= 32:
        assert frac < 0.1


@pytest.mark.parametrize("n,density,seed", [(16, 0.25, 1), (18, 0.2, 2), (20, 0.3, 3),
                                            (22, 0.15, 4), (24, 0.2, 5), (26, 0.12, 6)])
def test_random_binary_vs_exact(n, density, seed):
    a = P.synthetic_matrix(n, density, seed=800 + seed, binary=True,
                           full_support=True).astype(np.int32)
    ex = O.exact_permanent(a)
    r = P.skip(a)
    s = P.sparse(a, algo_id=4)
    assert r.value == s.value
    if ex < 2 ** 53:
        assert r.value == float(ex)


@pytest.mark.parametrize("n", [17, 21])
def test_small_integers_vs_exact(n):
    rng = np.random.default_rng(n)
    a = (rng.integers(0, 4, size=(n, n)) * (rng.random((n, n)) < 0.3)).astype(np.int32)
    np.fill_diagonal(a, 1)
    ex = O.exact_permanent(a)
    r = P.skip(a)
    assert r.value == float(ex) or abs(r.value - float(ex)) <= r.stated


def test_fallbacks_follow_the_sparse_path():
    real = P.synthetic_matrix(20, 0.3, seed=12, full_support=True)
    assert P.skip(real).value == P.sparse(real, algo_id=4).value
    small = P.synthetic_matrix(12, 0.4, seed=13, binary=True).astype(np.int32)
    assert P.skip(small).value == float(O.exact_permanent(small))


@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.float64])
def test_dropin_skipper(dtype):
    G = O.grid_graph_matrix(6, 6).astype(dtype)
    assert P.dropin_skipper(G) == float(O.domino_tilings(6, 6))
