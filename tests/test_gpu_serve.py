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
    exp_big = ctx.count_csr(bo, np.asarray(bt, np.int64))
    gs = _graphs(900, 40)
    exp = [O.count_degree_ordered(o, t) for o, t in gs]
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=4 << 30, max_batch_graphs=4096, max_inflight=2)
    _say("large request among small ones")
    tickets = [srv.submit(o, t) for o, t in gs[:20]]
    tb = srv.submit(bo, bt)
    tickets += [srv.submit(o, t) for o, t in gs[20:]]
    res = _drain_all(srv, tickets + [tb])
    assert [res[t][0] for t in tickets] == exp
    assert res[tb][0] == exp_big
    assert srv.stats()["batches"] >= 2
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
