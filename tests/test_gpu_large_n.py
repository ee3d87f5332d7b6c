"""Orders n = 49..63, which the reference's kernels and CPU paths accept
(gpu_exact_dense.cu:197-399 keep x in 64-entry arrays; algo.h:664, :1033).

A full walk of 2^48 subsets and more does not finish in a test, so:

* the dense kernels of each order range (49-53, 54-58, 59-63) run single
  workgroups of a 2^12-step split (PERMAN_WALK_BITS; a block-range run needs
  records for its range only) against a CPU double-double evaluation of the
  same subsets (oracle.lane_range_dd), in FP64 and in the double-double X
  mode -- the first, a middle and the last workgroup, whose lanes set every
  high column;
* the skipper walks the whole sum of the 8 x 14 grid graph (n = 56, 2^55
  subsets) and must give its domino tiling count exactly.
"""
import numpy as np
import pytest

import oracle as O
import permanent_amd as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [49, 56, 63])
def test_dense_block_ranges_large_orders(monkeypatch, n):
    a = P.synthetic_matrix(n, 0.5, seed=n)
    monkeypatch.setenv("PERMAN_WALK_BITS", "12")
    plan = P.Plan(a)  # above n = 48 the walk keeps the natural column order
    w, h, blocks = plan.geometry()
    assert (w, h, blocks) == (12, n - 13, 1 << (n - 21))
    try:
        for b0 in (0, blocks // 3 + 7, blocks - 1):
            fp = plan.run_blocks(b0, b0 + 1, mode=0)
            dd = plan.run_blocks(b0, b0 + 1, mode=2)
            hi, lo = O.lane_range_dd(a, w, 256 * b0, 256 * (b0 + 1), 8)
            ref = hi + lo
            a_sum = fp.abs_sum / 2.0  # |scale| = 2
            print(f"n={n} block {b0}: fp={fp.sum_hi!r} dd={dd.sum_hi!r} cpu={ref!r} "
                  f"|terms|={a_sum:.3e}")
            assert abs((dd.sum_hi + dd.sum_lo) - ref) <= 1e-13 * a_sum, (n, b0)
            assert abs((fp.sum_hi + fp.sum_lo) - ref) <= fp.err_est / 2.0 + 1e-13 * a_sum, (n, b0)
    finally:
        plan.close()


def test_plan_geometry_large_orders():
    """Default split above n = 48: 2^30 lanes, walks of 2^(n-31) steps."""
    for n in (49, 63):
        plan = P.Plan(P.synthetic_matrix(n, 0.5, seed=7))
        try:
            assert plan.geometry() == (n - 31, 30, 1 << 22)
        finally:
            plan.close()


def test_skipper_grid_8x14_n56_exact_tilings():
    G = O.grid_graph_matrix(8, 14).astype(np.int32)
    r = P.skip(G)
    print(f"skipper n=56: {r.value!r} walked {r.main_subsets / r.subsets:.2e} "
          f"kernel_ms={r.kernel_ms:.1f}")
    assert r.value == float(O.domino_tilings(8, 14))
