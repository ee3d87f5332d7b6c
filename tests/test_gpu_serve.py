"""Served form: continuous batching of requests of mixed n on one persistent
kernel (perm_server_*).

* results against the exact oracle (n <= 20) or the single-matrix path
  (within the stated errors), integer inputs exact, n < 8 embedded;
* a request gives the same bits alone and in a concurrent mix, whatever
  waves and launches ran its items;
* the persistent kernel ends when idle and is launched again; a small arena
  forces the ring to wrap (launches stopped for the reuse of its space);
* poll/wait/stats semantics, rejection of bad requests, shutdown drains;
* the autotuner returns a configuration from its candidate set."""
import ctypes
import threading

import numpy as np
import pytest

import oracle as O
import permanent_amd as P
from conftest import oracle_value

pytestmark = pytest.mark.gpu


def _check(m, r):
    n = m.shape[0]
    if n <= 20 or np.issubdtype(m.dtype, np.integer):
        ex = oracle_value(m)
        if np.issubdtype(m.dtype, np.integer) and ex < 2 ** 53:
            assert r.value == ex, (n, r.value, ex)
        else:
            assert abs(r.value - ex) <= max(1e-12 * abs(ex), r.stated), (n,)
    else:
        s = P.dense(m)
        assert abs(r.value - s.value) <= r.stated + s.stated
        assert abs(r.value - s.value) <= 1e-10 * abs(s.value)


def test_server_results_vs_oracle():
    reqs = []
    for i, n in enumerate(list(range(1, 21)) + [22, 24, 26, 28, 30]):
        binary = i % 3 == 0
        m = P.synthetic_matrix(n, 0.3 + 0.5 * ((i * 7) % 10) / 10, seed=1000 + i, binary=binary)
        reqs.append((m.astype(np.int32) if binary else m, i % 4 == 1))
    with P.Server(item_log2=14) as srv:
        tickets = [srv.submit(m, sparse_path=sp) for m, sp in reqs]
        results = [srv.wait(t) for t in tickets]
    for (m, sp), (r, lat) in zip(reqs, results):
        assert lat > 0
        if sp:  # nonnegative entries: the sparse path's sum is the permanent
            assert r.value == pytest.approx(oracle_value(m), rel=1e-9, abs=1e-300)
        else:
            _check(m, r)


def test_server_same_bits_alone_and_in_mix():
    target = P.synthetic_matrix(26, 0.5, seed=4242)
    with P.Server(item_log2=12, idle_us=300.0) as srv:
        alone = srv.wait(srv.submit(target))[0]
    others = P.served_workload(120, seed=9, n_lo=12, n_hi=28)
    got = {}
    with P.Server(item_log2=12, idle_us=300.0, grid=256) as srv:
        def client(part):
            for m, sp in part:
                srv.submit(m, sparse_path=sp)
        threads = [threading.Thread(target=client, args=(others[k::4],)) for k in range(4)]
        for th in threads:
            th.start()
        t1 = srv.submit(target)
        for th in threads:
            th.join()
        t2 = srv.submit(target)
        got[1] = srv.wait(t1)[0]
        got[2] = srv.wait(t2)[0]
        stats = srv.stats()
    assert stats["launches"] >= 1 and stats["items"] > 120
    for r in got.values():
        assert (r.value, r.sum_hi, r.sum_lo, r.abs_sum) == (
            alone.value, alone.sum_hi, alone.sum_lo, alone.abs_sum)


def test_server_matches_batch_path_bits():
    mats = P.mixed_batch(12, 14, 26, 0.3, 0.8, seed=31)
    b = P.Batch(item_log2=16)
    for m in mats:
        b.add(m)
    bres, _ = b.run()
    b.close()
    with P.Server(item_log2=16) as srv:
        sres = [srv.wait(srv.submit(m))[0] for m in mats]
    for x, y in zip(bres, sres):
        assert abs(x.value - y.value) <= x.stated + y.stated
        assert abs(x.value - y.value) <= 1e-12 * abs(x.value)


def test_server_poll_stats_and_errors():
    with P.Server() as srv:
        t = srv.submit(P.synthetic_matrix(30, 0.5, seed=5))
        first = srv.poll(t)
        r, lat = srv.wait(t) if first is None else first
        assert r.n == 30 and lat > 0
        with pytest.raises(P.PermError):
            srv.wait(t)  # ticket already collected
        with pytest.raises(P.PermError):
            srv.submit(np.ones((41, 41)))
        st = srv.stats()
        assert st["submitted"] == 1 and st["completed"] == 1 and st["pending"] == 0


def test_server_shutdown_drains():
    srv = P.Server(item_log2=12)
    for i in range(40):
        srv.submit(P.synthetic_matrix(20 + i % 8, 0.5, seed=i))
    srv.close()  # must finish the submitted work and return


def test_server_integer_requests_exact():
    reqs = [P.synthetic_matrix(n, d, seed=77 + n, binary=True).astype(np.int32)
            for n, d in [(18, 0.5), (22, 0.35), (24, 0.3), (26, 0.25)]]
    with P.Server() as srv:
        res = [srv.wait(srv.submit(m))[0] for m in reqs]
    for m, r in zip(reqs, res):
        ex = O.exact_permanent(m)
        if ex < 2 ** 53:
            assert r.value == float(ex), (m.shape, r, ex)


def test_server_wait_many_and_prepared_requests():
    mats = [(P.synthetic_matrix(16 + i % 12, 0.4, seed=900 + i), i % 3 == 0) for i in range(24)]
    reqs = [P.Request(m, sp) for m, sp in mats]
    with P.Server() as srv:
        one = [srv.wait(srv.submit_request(q))[0] for q in reqs]
        tickets = [srv.submit_request(q) for q in reqs]
        many = srv.wait_many(tickets)
        for a, (b, lat) in zip(one, many):
            assert (a.sum_hi, a.sum_lo) == (b.sum_hi, b.sum_lo) and lat > 0
        for q, r in zip(reqs, one):  # the plain call on the same prepared request
            v = P.solve(q)
            assert abs(v.value - r.value) <= v.stated + r.stated
        t = srv.submit_request(reqs[0])
        with pytest.raises(P.PermError):
            srv.wait_many([t, t])  # listed twice: the second is an error


def test_server_exact_pass_runs_as_batch_items():
    """Integer requests whose FP64 error could reach 1/2 get their
    double-double pass as batch items (counted in dd_passes) and come out as
    the exact integer."""
    mats = [P.synthetic_matrix(n, 0.5, seed=1200 + n, binary=True).astype(np.int32)
            for n in (20, 21, 22, 23, 24)]
    exact = [O.exact_permanent(m) for m in mats]
    with P.Server() as srv:
        res = srv.wait_many([srv.submit(m) for m in mats])
        st = srv.stats()
    for m, ex, (r, _) in zip(mats, exact, res):
        if ex < 2 ** 53:
            assert r.value == float(ex) and r.err_bound == 0.0, (m.shape, r, ex)
    print("dd passes:", st["dd_passes"], [r.mode for r, _ in res])
    assert st["dd_passes"] == sum(r.mode == 1 for r, _ in res)


def _sparse_ill_conditioned(n, seed):
    """Density 0.2, U(0,1] entries, every row and column nonzero: sums whose
    cancellation makes the modelled FP64 error estimate exceed 1e-10 at
    n >= 28 (numpy's default generator, seeded)."""
    rng = np.random.default_rng(seed)
    while True:
        mask = rng.random((n, n)) < 0.2
        if mask.any(0).all() and mask.any(1).all():
            break
    return np.where(mask, 1.0 - rng.random((n, n)), 0.0)


def test_server_two_waiters_on_one_ticket():
    """Two threads wait on the same ticket: one collects the result, the
    other is told the ticket was collected by another caller (the waiter
    looks the ticket up again after every wake-up, never a freed request)."""
    m = P.synthetic_matrix(28, 0.5, seed=4343)
    with P.Server() as srv:
        t = srv.submit(m)
        out = [None, None]

        def waiter(k):
            try:
                out[k] = srv.wait(t)[0]
            except P.PermError as e:
                out[k] = e

        ths = [threading.Thread(target=waiter, args=(k,)) for k in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
    got = [x for x in out if isinstance(x, P.PermResult)]
    err = [x for x in out if isinstance(x, P.PermError)]
    assert len(got) == 1 and len(err) == 1, out
    assert "collected" in str(err[0]) or "unknown" in str(err[0])


def test_server_submit_many_same_bits_and_errors_at_wait():
    """Bulk submission gives the same bits as one submission at a time
    (dense and sparse requests mixed); a malformed CCS passes the cheap
    submission checks and is reported by the wait of that request."""
    mats = [(P.synthetic_matrix(14 + i % 14, 0.45, seed=950 + i), i % 3 == 1) for i in range(30)]
    reqs = [P.Request(m, sp) for m, sp in mats]
    with P.Server() as srv:
        one = [srv.wait(srv.submit_request(q))[0] for q in reqs]
        many = srv.wait_many(srv.submit_many(reqs))
        for a, (b, _) in zip(one, many):
            assert (a.value, a.sum_hi, a.sum_lo) == (b.value, b.sum_hi, b.sum_lo)
        bad = P.synthetic_matrix(12, 0.5, seed=3)
        cptrs, rows, cvals = P.compress(bad)
        rows = rows.copy()
        rows[0] = 99  # row index out of range
        tid = ctypes.c_ulonglong()
        st = P.lib().perm_server_submit_sparse(srv._h, 2, bad.ctypes.data, P._ip(cptrs),
                                               P._ip(rows), cvals.ctypes.data, 12,
                                               ctypes.byref(tid))
        assert st == 0
        with pytest.raises(P.PermError, match="row index"):
            srv.wait(tid.value)


def test_server_open_loop_drive_same_bits():
    """The open-loop driver (perm_server_drive) gives every request the bits
    of a bulk submission, whether it arrives alone or in a burst of
    simultaneous arrivals handed over in one call; latencies count from the
    arrival time and the wall time covers the last arrival."""
    mats = [(P.synthetic_matrix(12 + i % 15, 0.5, seed=2300 + i), i % 4 == 2) for i in range(600)]
    reqs = [P.Request(m, sp) for m, sp in mats]
    batch = P.RequestBatch(reqs)
    # a burst of 300 at t = 0, then one every 0.2 ms
    arrivals = np.concatenate([np.zeros(300), 0.001 + 0.0002 * np.arange(300)])
    with P.Server() as srv:
        ref = srv.wait_many(srv.submit_many(reqs))
    with P.Server() as srv:
        res, lat, wall = srv.drive(batch, arrivals)
    assert wall >= arrivals[-1]
    for (a, _), r, l in zip(ref, res, lat):
        assert (a.value, a.sum_hi, a.sum_lo) == (r.value, r.sum_hi, r.sum_lo)
        assert 0.0 < l <= wall * 1e3
    lat = np.array(list(lat))
    # the burst's requests wait for each other: its last result comes after
    # its first, counted from the same arrival
    assert lat[:300].max() > lat[:300].min()


def test_server_sampled_error_measurement():
    """Non-integer requests of many items whose modelled error estimate
    misses 1e-10 have their error measured on a sample of items in the
    double-double X mode (dd_passes counts it) and come out certified and
    within 1e-10 of the single-matrix path."""
    mats = [_sparse_ill_conditioned(28, 28020), _sparse_ill_conditioned(30, 30020),
            P.synthetic_matrix(30, 0.5, seed=1530)]
    with P.Server() as srv:
        res = srv.wait_many([srv.submit(m) for m in mats])
        st = srv.stats()
    for m, (r, _) in zip(mats, res):
        one = P.dense(m)
        assert r.certified, r
        assert abs(r.value - one.value) <= 1e-10 * abs(one.value), (m.shape, r, one)
    print("dd passes:", st["dd_passes"], [(r.mode, r.err_est / abs(r.value)) for r, _ in res])


def test_server_escalated_request_same_bits_alone_and_in_mix():
    """A request whose FP64 pass does not meet the target (its error
    measured on a sample of items in the double-double X mode, and a full
    pass if that does not certify) gives the same bits, err_est, mode and
    certificate alone and among 120 concurrent requests."""
    targets = [_sparse_ill_conditioned(28, 28020), _sparse_ill_conditioned(30, 30020)]
    with P.Server() as srv:
        alone = [srv.wait(srv.submit(t))[0] for t in targets]
    others = P.served_workload(120, seed=9, n_lo=12, n_hi=28)
    with P.Server(grid=256) as srv:
        ids = [srv.submit(m, sparse_path=sp) for m, sp in others[:60]]
        tids = [srv.submit(t) for t in targets]
        ids += [srv.submit(m, sparse_path=sp) for m, sp in others[60:]]
        mixed = [srv.wait(t)[0] for t in tids]
        srv.wait_many(ids)
        st = srv.stats()
    assert st["dd_passes"] >= 2
    for a, r in zip(alone, mixed):
        assert r.mode != 0 or r.err_est > 0.0
        assert (a.value, a.sum_hi, a.sum_lo, a.err_est, a.mode, a.certified) == (
            r.value, r.sum_hi, r.sum_lo, r.err_est, r.mode, r.certified)
        assert r.certified


def test_server_relaunch_after_idle_and_arena_wrap():
    """The persistent kernel leaves after idle_us without work and the next
    submission launches it again; with a 16 MiB arena the block ring wraps
    (the live launch is stopped before its space is reused) and every result
    still matches the single-matrix path."""
    import time

    mats = [P.synthetic_matrix(22 + i % 7, 0.5, seed=5100 + i) for i in range(12)]
    with P.Server(idle_us=200.0, arena_mb=16.0, item_log2=12) as srv:
        first = srv.wait(srv.submit(mats[0]))[0]
        time.sleep(0.05)  # the launch has ended by now
        again = srv.wait(srv.submit(mats[0]))[0]
        assert (first.sum_hi, first.sum_lo) == (again.sum_hi, again.sum_lo)
        assert srv.stats()["launches"] >= 2
    # order 30 throughout: the GPU (not the host) sets the pace, requests
    # finish about in arrival order (blocks are freed in order), and one
    # launch lives through the wrap
    mats = [P.synthetic_matrix(30, 0.5, seed=5200 + i) for i in range(12)]
    with P.Server(arena_mb=16.0, item_log2=12, idle_us=20000.0) as srv:
        # (a long idle time: the launch lives across the host's turnarounds)
        # ~60 MiB of blocks submitted at once (live blocks span at most half
        # the arena): while the kernel works, space freed during its launch
        # is reused only after stopping it
        res = srv.wait_many([srv.submit(m) for _ in range(24) for m in mats])
        st = srv.stats()
    print(st)
    assert st["laps"] >= 1
    assert st["stops"] >= 1
    for k, (r, _) in enumerate(res):
        one = P.dense(mats[k % len(mats)])
        assert abs(r.value - one.value) <= r.stated + one.stated
        assert abs(r.value - one.value) <= 1e-10 * abs(one.value)


def test_server_autotune_returns_a_candidate():
    sample = [P.Request(m, sp) for m, sp in P.served_workload(200, seed=17, n_lo=12, n_hi=24)]
    cfg, rate = P.autotune_server(sample)
    print(cfg, rate)
    assert rate > 0
    assert cfg["item_log2"] in (12, 14, 16, 18, 20)
    assert cfg["grid"] in (256, 512, 1024, 2048)
