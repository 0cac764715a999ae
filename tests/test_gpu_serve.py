"""GPU: the continuous-batching server.

Each request's count must equal the graph's solo count, whatever else is in
flight; admission must respect the HBM budget.
"""
import time

import numpy as np
import pytest

import oracle_py as O
import xxtric_amd

pytestmark = pytest.mark.gpu

SEED = 2026


def _say(*a):
    print("[case]", *a, flush=True)


def _graphs(first, n):
    return [xxtric_amd.batch_graph_csr(SEED, g) for g in range(first, first + n)]


def _drain_all(srv, tickets, timeout=120.0):
    srv.drain(timeout)
    out = {}
    for t in tickets:
        r = srv.poll(t)
        assert r is not None
        out[t] = r
    return out


def test_server_counts_equal_solo(ctx):
    gs = _graphs(0, 120)
    exp = [O.count_degree_ordered(o, t) for o, t in gs]
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=1 << 30, max_batch_graphs=32)
    _say("serve 120 graphs")
    tickets = [srv.submit(o, t) for o, t in gs]
    res = _drain_all(srv, tickets)
    assert [res[t][0] for t in tickets] == exp
    st = srv.stats()
    assert st["completed"] == 120 and st["batches"] >= 4 and st["p99_us"] >= st["p50_us"] > 0
    srv.close()


def test_arrivals_while_running_overlap(ctx):
    """Requests arriving while batches run are admitted into free slots."""
    gs = _graphs(500, 200)
    exp = [O.count_degree_ordered(o, t) for o, t in gs]
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=1 << 30, max_batch_graphs=16, max_inflight=2)
    tickets = []
    rng = np.random.default_rng(1)
    _say("serve with jittered arrivals")
    for o, t in gs:
        tickets.append(srv.submit(o, t))
        time.sleep(float(rng.exponential(50e-6)))
    res = _drain_all(srv, tickets)
    assert [res[t][0] for t in tickets] == exp
    st = srv.stats()
    assert st["batches"] > 1 and st["completed"] == 200
    srv.close()


def test_alone_equals_in_mix(ctx):
    o, t = xxtric_amd.batch_graph_csr(SEED, 11)
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=1 << 30)
    _say("alone")
    alone = _drain_all(srv, [srv.submit(o, t)])
    a = list(alone.values())[0][0]
    mix = _graphs(600, 50)
    tickets = [srv.submit(*g) for g in mix[:25]] + [srv.submit(o, t)] + [srv.submit(*g) for g in mix[25:]]
    _say("in mix")
    res = _drain_all(srv, tickets)
    assert res[tickets[25]][0] == a == O.count_degree_ordered(o, t)
    srv.close()


def test_large_request_among_small_ones(ctx):
    """An R-MAT s18 request (over 4 M entries) arriving between small graphs is
    counted in a batch of its own while little is queued; every count must
    still equal the graph's solo count, and the small graphs after it are not
    held back in its batch."""
    big = ctx.rmat(18, 16, 3, True)
    bo, bt = big.download()
    big.close()
    assert len(bt) >= 1 << 22
    exp_big = O.count_degree_ordered(bo, bt)
    gs = _graphs(900, 40)
    exp = [O.count_degree_ordered(o, t) for o, t in gs]
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=4 << 30, max_batch_graphs=4096, max_inflight=2)
    _say("large request among small ones")
    srv.pause()  # all 41 queued before admission starts: the order is fixed
    tickets = [srv.submit(o, t) for o, t in gs[:20]]
    tb = srv.submit(bo, bt)
    tickets += [srv.submit(o, t) for o, t in gs[20:]]
    srv.pause(False)
    res = _drain_all(srv, tickets + [tb])
    assert [res[t][0] for t in tickets] == exp
    assert res[tb][0] == exp_big
    st = srv.stats()
    # 41 queued, far below an eighth of a slot (512 graphs): the s18 request runs
    # alone, between a batch of the 20 before it and one of the 20 after it
    assert st["isolated_batches"] == 1 and st["batches"] == 3
    srv.close()


def test_queue_full_is_rejected(ctx):
    """max_queue: with admission held, submissions beyond the queue's capacity
    fail with XT_ERR_CAPACITY (counted as rejected); the queued ones finish."""
    gs = _graphs(1200, 5)
    exp = [O.count_degree_ordered(o, t) for o, t in gs]
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=1 << 30, max_batch_graphs=16, max_queue=4)
    srv.pause()
    tickets = [srv.submit(o, t) for o, t in gs[:4]]
    assert [srv.ticket_state(t) for t in tickets] == [0, 0, 0, 0]
    for borrow in (False, True):
        with pytest.raises(xxtric_amd.XtError) as e:
            srv.submit(*gs[4], borrow=borrow)
        assert e.value.code == -5
    srv.pause(False)
    res = _drain_all(srv, tickets)
    assert [res[t][0] for t in tickets] == exp[:4]
    assert srv.ticket_state(tickets[0]) == -6  # retired once polled
    st = srv.stats()
    assert st["submitted"] == 6 and st["completed"] == 4 and st["rejected"] == 2
    srv.close()


def test_latency_statistics_from_known_arrivals(ctx):
    """p50 / p99 / max from arrivals stamped k * 10 ms in the past (k = 1..100,
    all after the server clock's start): each latency is k * 10 ms plus the time
    to count, so the order statistics are fixed to within the run's duration."""
    gs = _graphs(1300, 4)
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=1 << 30, max_batch_graphs=64)
    try:
        time.sleep(1.2)  # the clock starts at creation: room for arrivals 1 s back
        srv.pause()
        now = srv.now_us()
        t0 = time.perf_counter()
        tickets = [srv.submit(*gs[k % 4], arrival_us=now - 1e4 * (k + 1)) for k in range(100)]
        srv.pause(False)
        _drain_all(srv, tickets)
        slack = (time.perf_counter() - t0) * 1e6 + 2e3
        st = srv.stats()
        assert st["completed"] == 100
        for q, k in (("p50_us", 50), ("p99_us", 99), ("max_us", 100)):
            assert k * 1e4 <= st[q] <= k * 1e4 + slack, (q, st[q], slack)
    finally:
        srv.close()


def test_forced_allocation_failures(ctx):
    """An allocation that fails (injected) surfaces as XT_ERR_NOMEM from a count,
    a batch and a server; each object is cleaned up and the next call works."""
    o, t = xxtric_amd.batch_graph_csr(SEED, 3)
    exp = O.count_degree_ordered(o, t)
    with xxtric_amd.Context(0) as c2:  # fresh scratch: its first count allocates
        c2.fail_allocs(1)
        with pytest.raises(xxtric_amd.XtError) as e:
            c2.count_csr(o, np.asarray(t, np.int64))
        assert e.value.code == -3
        assert c2.count_csr(o, np.asarray(t, np.int64)) == exp
        c2.fail_allocs(1)
        with pytest.raises(xxtric_amd.XtError) as e:
            xxtric_amd.Batch(c2, 100_000, 1_000_000, 8)
        assert e.value.code == -3
        c2.fail_allocs(3)
        with pytest.raises(xxtric_amd.XtError) as e:
            xxtric_amd.Server(c2, hbm_budget_bytes=64 << 20, max_batch_graphs=8)
        assert e.value.code == -3
        c2.fail_allocs(0)
        srv = xxtric_amd.Server(c2, hbm_budget_bytes=64 << 20, max_batch_graphs=8)
        tk = srv.submit(o, t)
        assert _drain_all(srv, [tk])[tk][0] == exp
        srv.close()


def test_admission_budget(ctx):
    budget = 16 << 20  # 16 MiB: small graphs only, few at a time
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=budget, max_batch_graphs=64)
    big = ctx.rmat(16, 16, 1, True)
    bo, bt = big.download()
    big.close()
    with pytest.raises(xxtric_amd.XtError) as e:      # larger than the budget: rejected at submit
        srv.submit(bo, bt)
    assert e.value.code == -5
    gs = [g for g in _graphs(700, 120) if len(g[1]) <= 50_000][:60]
    exp = [O.count_degree_ordered(o, t) for o, t in gs]
    _say("serve under a 16 MiB budget")
    tickets = [srv.submit(o, t) for o, t in gs]
    res = _drain_all(srv, tickets)
    assert [res[t][0] for t in tickets] == exp
    st = srv.stats()
    assert st["rejected"] == 1 and 0 < st["peak_resident_bytes"] <= budget
    srv.close()


def test_borrowed_submissions(ctx):
    """Borrowed requests (read in place, checked in the server's thread) give
    the same counts as copied ones, mixed in the same batches; a bad borrowed
    request is reported by poll and does not disturb the others."""
    gs = _graphs(800, 80)
    exp = [O.count_degree_ordered(o, t) for o, t in gs]
    bad_off = np.array([0, 2, 3, 4], np.int64)
    bad_tails = np.array([1, 7, 0, 0], np.int32)  # id 7 outside [0, 3)
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=1 << 30, max_batch_graphs=16)
    _say("serve borrowed + copied")
    tickets = []
    for i, (o, t) in enumerate(gs):
        tickets.append(srv.submit(o, t, borrow=(i % 2 == 0)))
        if i == 40:
            bad = srv.submit(bad_off, bad_tails, borrow=True)
    srv.drain(120.0)
    with pytest.raises(xxtric_amd.XtError) as e:
        srv.poll(bad)
    assert e.value.code == -4
    res = {t: srv.poll(t) for t in tickets}
    assert [res[t][0] for t in tickets] == exp
    st = srv.stats()
    assert st["completed"] == 80 and st["rejected"] == 1
    srv.close()


def test_native_replay(ctx):
    """The native request generator: Poisson arrivals over a pool of graphs
    (one of them above the async staging threshold), every result equal to the
    oracle's count, latencies non-negative."""
    gs = _graphs(1400, 12)
    big = ctx.rmat(17, 16, 9, True)
    gs.append(big.download())
    big.close()
    exp = np.array([O.count_degree_ordered(o, t) for o, t in gs], np.uint64)
    rng = np.random.default_rng(3)
    n = 600
    order = rng.integers(0, len(gs), n)
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=2 << 30, max_batch_graphs=64, max_inflight=2)
    arrivals = srv.now_us() + 1000.0 + np.cumsum(rng.exponential(1e6 / 20000.0, n))
    _say("native replay", n)
    bad, lat = srv.replay(gs, exp, order, arrivals)
    assert bad == 0 and (lat >= 0).all()
    st = srv.stats()
    assert st["completed"] == n and st["rejected"] == 0
    # the worker's time breakdown: every part measured, the batches' device time positive
    for k in ("admit_s", "stage_s", "launch_s", "complete_s", "device_s"):
        assert st[k] >= 0.0, k
    assert st["device_s"] > 0.0 and st["launch_s"] > 0.0
    # a wrong expectation is reported, not hidden
    bad2, _ = srv.replay(gs, exp + np.uint64(1), order[:5], srv.now_us() + np.arange(5) * 10.0)
    assert bad2 == 5
    srv.close()
