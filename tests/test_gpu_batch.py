"""Batched mixed-size permanents (persistent variant kernels in one graph).

* every result against the exact oracle (n <= 20) or the single-matrix
  path (larger n, within the stated error);
* a matrix gives the same bits alone and in a mix (its items depend only on
  its own n), and across repeated graph replays;
* dense and sparse (-s semantics) entries in one batch."""
import numpy as np
import pytest

import oracle as O
import permanent_amd as P
from conftest import assert_target, oracle_value

pytestmark = pytest.mark.gpu


def test_batch_mixed_small_vs_exact():
    mats = [P.synthetic_matrix(n, d, seed=70 + n) for n in range(1, 21)
            for d in (0.35, 0.8)]
    b = P.Batch()
    for m in mats:
        b.add(m)
    res, ms = b.run()
    b.close()
    assert len(res) == len(mats) and ms > 0
    for m, r in zip(mats, res):
        ex = float(O.exact_permanent(m))
        assert_target(r, ex, ctx=m.shape[0])


def test_batch_large_vs_single_path():
    mats = P.mixed_batch(24, 22, 30, 0.2, 0.8, seed=11)
    b = P.Batch()
    for m in mats:
        b.add(m)
    res, _ = b.run()
    b.close()
    for m, r in zip(mats, res):
        s = P.dense(m)
        assert abs(r.value - s.value) <= r.stated + s.stated
        assert abs(r.value - s.value) <= 1e-10 * abs(s.value)


def test_same_alone_and_in_the_mix():
    target = P.synthetic_matrix(27, 0.45, seed=123)
    alone = P.Batch()
    alone.add(target)
    r_alone = alone.run()[0][0]
    alone.close()
    mix = P.Batch()
    others = P.mixed_batch(60, 12, 30, 0.2, 0.8, seed=3)
    for m in others[:30]:
        mix.add(m)
    k = mix.add(target)
    for m in others[30:]:
        mix.add(m, sparse_path=True)
    res1, _ = mix.run()
    res2, _ = mix.run()  # graph replay
    mix.close()
    for r in (res1[k], res2[k]):
        assert (r.value, r.sum_hi, r.sum_lo, r.abs_sum) == (
            r_alone.value, r_alone.sum_hi, r_alone.sum_lo, r_alone.abs_sum)
    assert [r.value for r in res1] == [r.value for r in res2]


def test_pipelined_compute_matches_add_and_run():
    """perm_batch_compute_dense (preparation of the smaller orders overlapped
    with the kernels of the largest) gives bitwise the results of
    add_dense_many + run, in the caller's order, zero-pattern matrices
    included; later runs of the same batch (direct, then captured and
    replayed) return them in the caller's order too."""
    import ctypes

    mats = P.mixed_batch(300, 8, 26, 0.2, 0.8, seed=21)
    z = P.synthetic_matrix(15, 0.6, seed=5).copy()
    z[:, 3] = 0.0  # a zero column: no perfect matching, exactly 0
    mats.insert(17, z)
    inputs = P.BatchInputs(mats)
    ref = P.Batch()
    ref.add_prepared(inputs, threads=4)
    want, _ = ref.run()
    ref.close()
    b = P.Batch()
    got_c = (P.PermResultC * len(mats))()
    ms = b.compute_prepared(inputs, got_c, threads=4)
    assert ms > 0
    got = [P.PermResult.from_c(r) for r in got_c]
    again1, _ = b.run()
    again2, _ = b.run()  # graph capture, then a replay
    again3, _ = b.run()
    b.close()
    assert got[17].value == 0.0
    for res in (got, again1, again2, again3):
        for r, w in zip(res, want):
            assert (r.value, r.sum_hi, r.sum_lo, r.abs_sum, r.mode) == (
                w.value, w.sum_hi, w.sum_lo, w.abs_sum, w.mode)


def test_batch_sparse_entries_follow_reference_semantics():
    a = P.synthetic_matrix(14, 0.6, seed=9) - 0.25 * (np.arange(14)[:, None] % 4 == 0)
    b = P.Batch()
    b.add(a, sparse_path=True)
    b.add(np.abs(a))
    res, _ = b.run()
    b.close()
    assert res[0].value == pytest.approx(O.cpu_perman64_sparse(a, 2), rel=1e-9)
    assert res[1].value == pytest.approx(float(O.exact_permanent(np.abs(a))), rel=1e-10)


def test_batch_item_size_parameter_keeps_results():
    mats = P.mixed_batch(16, 16, 26, 0.3, 0.7, seed=21)
    vals = []
    for item_log2 in (12, 16, 20):
        b = P.Batch(item_log2=item_log2, streams=2)
        for m in mats:
            b.add(m)
        vals.append(b.run()[0])
        b.close()
    for rs in zip(*vals):
        v0 = rs[0].value
        for r in rs[1:]:
            assert abs(r.value - v0) <= 1e-10 * abs(v0) + r.stated


def test_batch_rejects_too_large():
    b = P.Batch()
    with pytest.raises(P.PermError):
        b.add(np.ones((41, 41)))
    b.close()


def test_autotune_picks_a_split_and_keeps_results():
    mats = P.mixed_batch(48, 12, 24, 0.2, 0.8, seed=44)
    b = P.Batch()
    for m in mats:
        b.add(m)
    before, _ = b.run()
    best = b.autotune(reps=1)
    assert best["trials"] >= 10 and best["device_ms"] > 0
    assert b.tuning()["item_log2"] == best["item_log2"]
    after, _ = b.run()
    b.close()
    for x, y in zip(before, after):
        assert abs(x.value - y.value) <= x.stated + y.stated


def test_batch_nested_shapes_vs_exact_and_single():
    """Sparse-ish matrices of n >= 24 run on the three-level batch kernels:
    the results match the exact oracle (n <= 26) and the single-matrix path."""
    mats = [P.synthetic_matrix(n, d, seed=500 + 7 * n + int(10 * d), full_support=True)
            for n in (24, 26, 30, 34) for d in (0.2, 0.35)]
    b = P.Batch()
    for m in mats:
        b.add(m)
    res, _ = b.run()
    shapes = set()
    for m, r in zip(mats, res):
        n = m.shape[0]
        shapes.add((r.nu, r.nu1, r.nu2))
        one = P.dense(m)
        assert abs(r.value - one.value) <= r.stated + one.stated, (n, r, one)
        if n <= 26:
            ex = oracle_value(m)
            assert_target(r, ex, ctx=n)
    print("batch shapes:", sorted(shapes))
    assert any(r.nu2 < r.n for r in res)  # three-level shapes ran


def test_batch_integer_exact_pass_on_batch_kernels():
    """0/1 matrices whose FP64 error could reach 1/2 get the double-double
    pass inside the batch (mode 1) and come out as the exact integer, the
    same value as the single-matrix path."""
    mats = [P.synthetic_matrix(n, 0.5, seed=1300 + n, binary=True).astype(np.int32)
            for n in (20, 21, 22, 23, 24, 16)]
    b = P.Batch()
    for m in mats:
        b.add(m)
    res, _ = b.run()
    modes = [r.mode for r in res]
    print("modes:", modes)
    assert 1 in modes
    for m, r in zip(mats, res):
        ex = O.exact_permanent(m)
        one = P.dense(m)
        if ex < 2 ** 53:
            assert r.value == float(ex) == one.value and r.err_bound == 0.0, (m.shape, r, ex)
        else:  # above 2^53: the same within the stated errors
            assert abs(r.value - one.value) <= r.stated + one.stated
            assert abs(r.value - float(ex)) <= r.stated
    res2, _ = b.run()  # a replay recomputes the same values
    assert [r.value for r in res2] == [r.value for r in res]


def _c4_fixture():
    import json
    import os

    from conftest import GOLDEN

    path = os.path.join(GOLDEN, "c4_mix.json")
    if not os.path.exists(path):
        pytest.skip("c4_mix.json not generated")
    return json.load(open(path))["members"]


def test_c4_mix_every_member_exact_zero_or_within_1e10():
    """The batched configuration's whole mix (4096 matrices, n ~ U{12..28},
    density ~ U[0.2, 0.8], seed 5) through one batch and through one plain
    call per matrix: every member with n <= 26 is exactly 0 when its pattern
    has no perfect matching, else within 1e-10 of its oracle value
    (tests/golden/c4_mix.json: exact for n <= 16, double-double above) and
    certified; the batch and plain results agree to 1e-10."""
    from conftest import matrix_key

    members = _c4_fixture()
    mats = P.mixed_batch(4096, 12, 28, 0.2, 0.8, 5)
    b = P.Batch()
    for m in mats:
        b.add(m)
    res, ms = b.run()
    b.close()
    modes = {}
    for rec in members:
        m = mats[rec["i"]]
        assert matrix_key(m) == rec["key"], rec["i"]
        r = res[rec["i"]]
        ex = rec["hi"] + rec["lo"]
        if rec["kind"] == "zero":
            assert r.value == 0.0 and r.zero_pattern == 1, (rec["i"], r)
        assert_target(r, ex, ctx=("batch", rec["i"], rec["n"]))
        modes[r.mode] = modes.get(r.mode, 0) + 1
        s = P.dense(m)
        assert_target(s, ex, ctx=("plain", rec["i"], rec["n"]))
        assert abs(r.value - s.value) <= 1e-10 * abs(s.value), (rec["i"], r, s)
    print(f"C4 mix: {len(members)} members checked, batch modes {modes}, device {ms:.1f} ms")
    assert all(r.certified for r in res)


def test_replays_with_escalations_same_bits():
    """The batched configuration's mix, whose hardest members escalate (a
    sampled double-double measurement, then full double-double passes):
    every graph replay gives the first run's bits, err_est, mode and
    certificate for every member, a fresh batch computed in one call gives
    the same bits, and so do the escalated members in a batch of their own."""
    mats = P.mixed_batch(4096, 12, 28, 0.2, 0.8, 5)
    b = P.Batch(item_log2=16, streams=8, max_grid=512)
    b.add_prepared(P.BatchInputs(mats), threads=8)
    runs = [b.run() for _ in range(3)]
    b.close()
    first = runs[0][0]
    escalated = sum(r.mode != 0 for r in first)
    assert escalated >= 2, escalated  # the escalation rounds run on every replay
    for res, ms in runs[1:]:
        for i, (a, r) in enumerate(zip(first, res)):
            assert (a.value, a.sum_hi, a.sum_lo, a.err_est, a.mode, a.certified) == (
                r.value, r.sum_hi, r.sum_lo, r.err_est, r.mode, r.certified), i
    one = P.Batch(item_log2=16, streams=8, max_grid=512)
    res_c = (P.PermResultC * len(mats))()
    one.compute_prepared(P.BatchInputs(mats), res_c, threads=8)
    one.close()
    for i, (a, r) in enumerate(zip(first, res_c)):
        assert (a.value, a.sum_hi, a.sum_lo, a.mode) == (r.value, r.sum_hi, r.sum_lo, r.mode), i
    # the escalated members alone: every decision (sampling, the split of the
    # double-double passes, the rounds) depends on the matrix only
    esc = [i for i, r in enumerate(first) if r.mode != 0]
    alone = P.Batch(item_log2=16, streams=8, max_grid=512)
    for i in esc:
        alone.add(mats[i])
    res_alone, _ = alone.run()
    alone.close()
    for i, r in zip(esc, res_alone):
        a = first[i]
        assert (a.value, a.sum_hi, a.sum_lo, a.err_est, a.mode, a.certified) == (
            r.value, r.sum_hi, r.sum_lo, r.err_est, r.mode, r.certified), i
    print("escalated:", escalated, "device ms per run:", [round(ms, 2) for _, ms in runs])

