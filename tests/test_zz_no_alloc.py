"""perm_dense_noalloc: the dense path with every device buffer in a static
device array (no hipMalloc/hipFree).  Called through the C ABI and checked
against perm_dense on the same inputs -- bitwise, since both run the same
split, kernels and reduction -- on two calls in a row, the second with a
smaller input (the array keeps the first call's contents), and on the size
limit."""
import ctypes

import numpy as np
import pytest

import permanent_amd as P

pytestmark = pytest.mark.gpu


def noalloc(a, algo_id=4):
    a = np.ascontiguousarray(a)
    dt = P._DT[a.dtype]
    r = P.PermResultC()
    st = P.lib().perm_dense_noalloc(dt, a.ctypes.data, a.shape[0], algo_id, ctypes.byref(r))
    return st, P.PermResult.from_c(r)


def same_bits(x, y):
    return (x.value, x.sum_hi, x.sum_lo, x.abs_sum, x.err_bound, x.err_est, x.mode) == (
        y.value, y.sum_hi, y.sum_lo, y.abs_sum, y.err_bound, y.err_est, y.mode)


def test_noalloc_matches_perm_dense_on_two_calls():
    big = P.synthetic_matrix(34, 0.5, seed=71)   # 2^33 subsets, 2^13 partial records
    small = P.synthetic_matrix(21, 0.6, seed=72)  # a smaller second call
    for a in (big, small):
        st, r = noalloc(a)
        assert st == 0, P.lib().perm_last_error()
        ref = P.dense(a, 4)
        print(f"n={a.shape[0]} noalloc={r.value!r} dense={ref.value!r} mode={r.mode}")
        assert same_bits(r, ref), (r, ref)
        assert r.certified


def test_noalloc_integer_and_zero_pattern():
    b = P.synthetic_matrix(24, 0.5, seed=324, binary=True).astype(np.int32)
    st, r = noalloc(b)
    assert st == 0 and same_bits(r, P.dense(b, 4))
    z = np.ones((20, 20))
    z[:, 5] = 0.0
    st, r = noalloc(z)
    assert st == 0 and r.value == 0.0 and r.zero_pattern == 1


def test_noalloc_rejects_sizes_beyond_its_scratch():
    a = P.synthetic_matrix(36, 0.5, seed=73)
    st, _ = noalloc(a)
    assert st == 2  # PERM_ERR_SIZE
    assert "scratch" in P.lib().perm_last_error().decode()
    st, r = noalloc(P.synthetic_matrix(12, 0.5, seed=74))  # still usable afterwards
    assert st == 0 and same_bits(r, P.dense(P.synthetic_matrix(12, 0.5, seed=74), 4))
