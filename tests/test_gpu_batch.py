"""GPU: batched counting (persistent pipeline, hipGraph replay, per-graph baseline).

Every per-graph total must equal the oracle's count of that graph, whatever
the submission order, the kernel split and the launch mode.
"""
import numpy as np
import pytest

import oracle_py as O
import xxtric_amd
from xxtric_amd import XT_BATCH_HIPGRAPH, XT_BATCH_PER_GRAPH, XT_BATCH_PERSISTENT

pytestmark = pytest.mark.gpu

SEED = 2026


def _say(*a):
    print("[case]", *a, flush=True)


def _expected(gs, nmax=20000):
    out = []
    for g in gs:
        off, tails = xxtric_amd.batch_graph_csr(SEED, g, 200, nmax)
        out.append(O.count_degree_ordered(off, tails))
    return np.array(out, np.uint64)


def test_generated_batch_matches_oracle(ctx, golden):
    n = 48
    b = xxtric_amd.Batch(ctx, 4_000_000, 40_000_000, 4096)
    _say("batch generated", n)
    b.add_generated(SEED, 0, n)
    b.run(XT_BATCH_PERSISTENT)
    got = b.results()
    exp = _expected(range(n))
    assert (got == exp).all(), np.nonzero(got != exp)[0]
    for g in golden["batch"]:
        assert int(got[g["index"]]) == g["triangles"]
    b.close()


def test_modes_agree_and_graph_replays(ctx):
    n = 40
    b = xxtric_amd.Batch(ctx, 4_000_000, 40_000_000, 4096)
    b.add_generated(SEED, 100, n)
    _say("persistent")
    b.run(XT_BATCH_PERSISTENT)
    ref = b.results().copy()
    _say("per-graph")
    b.run(XT_BATCH_PER_GRAPH)
    assert (b.results() == ref).all()
    for rep in range(4):  # replays interleaved with eager work
        _say("hipgraph replay", rep)
        b.run(XT_BATCH_HIPGRAPH)
        assert (b.results() == ref).all()
        b.run(XT_BATCH_PERSISTENT)
        assert (b.results() == ref).all()
    assert (ref == _expected(range(100, 100 + n))).all()
    b.close()


def test_permuted_order_and_split(ctx):
    gs = list(range(200, 232))
    graphs = [xxtric_amd.batch_graph_csr(SEED, g) for g in gs]
    exp = _expected(gs)
    perm = np.random.default_rng(5).permutation(len(gs))
    for cut in (1, 16, 64, 1 << 30, xxtric_amd.DEFAULT_LIGHT_CUT):  # the last: back to the defaults
        ctx.tune(small_cut=cut)
        try:
            b = xxtric_amd.Batch(ctx, 2_000_000, 20_000_000, 1024)
            for k in perm:
                b.add(*graphs[k])
            _say("permuted batch small_cut", cut)
            b.run(XT_BATCH_PERSISTENT)
            got = b.results()
            assert (got == exp[perm]).all()
            b.close()
        finally:
            ctx.tune(small_cut=xxtric_amd.DEFAULT_LIGHT_CUT)


def test_solo_equals_in_batch(ctx):
    off, tails = xxtric_amd.batch_graph_csr(SEED, 7)
    solo = ctx.count_csr(off, tails.astype(np.int64))
    b = xxtric_amd.Batch(ctx, 2_000_000, 20_000_000, 1024)
    b.add_generated(SEED, 300, 10)
    b.add(off, tails)
    b.add_generated(SEED, 400, 10)
    _say("solo vs batch")
    b.run(XT_BATCH_PERSISTENT)
    assert int(b.results()[10]) == solo
    b.close()


def test_mixed_sizes_with_rmat(ctx):
    g = ctx.rmat(14, 16, 3, True)
    off, tails = g.download()
    solo = g.count()
    g.close()
    b = xxtric_amd.Batch(ctx, 2_000_000, 20_000_000, 1024)
    b.add_generated(SEED, 0, 5)
    b.add(off, tails)
    b.add_generated(SEED, 5, 5)
    _say("batch with rmat14")
    b.run(XT_BATCH_PERSISTENT)
    got = b.results()
    assert int(got[5]) == solo == O.count_degree_ordered(off, tails)
    assert (np.delete(got, 5) == _expected(range(10))).all()
    b.close()


def test_capacity_errors(ctx):
    b = xxtric_amd.Batch(ctx, 1000, 10000, 2)
    off, tails = xxtric_amd.batch_graph_csr(SEED, 5)   # 1139 vertices: over the vertex capacity
    with pytest.raises(xxtric_amd.XtError) as e:
        b.add(off, tails)
    assert e.value.code == -5
    small = (np.array([0, 2, 4, 6], np.int64), np.array([1, 2, 0, 2, 0, 1], np.int32))
    b.add(*small)
    b.add(*small)
    with pytest.raises(xxtric_amd.XtError):
        b.add(*small)                                   # graph slots exhausted
    b.run(XT_BATCH_PERSISTENT)
    assert list(b.results()) == [1, 1]
    b.close()


def test_staged_host_adds_survive_reuse(ctx):
    """Host adds are staged in pinned memory and flushed at launch: the caller's
    arrays may change right after add, and a cleared batch refilled with other
    graphs gives their counts in every launch mode."""
    b = xxtric_amd.Batch(ctx, 2_000_000, 20_000_000, 256)
    for rnd, (lo, hi) in enumerate([(0, 40), (40, 60)]):
        _say("staged round", rnd)
        gs = list(range(lo, hi))
        exp = _expected(gs)
        for g in gs:
            off, tails = xxtric_amd.batch_graph_csr(SEED, g, 200, 20000)
            off = off.astype(np.int64)                  # the exact buffers handed over
            tails = tails.astype(np.int32)
            b.add(off, tails)
            tails[:] = 0                                # caller reuses its buffers
            off[:] = 0
        for mode in (XT_BATCH_PERSISTENT, XT_BATCH_HIPGRAPH, XT_BATCH_PER_GRAPH):
            b.run(mode)
            assert (b.results() == exp).all()
        b.clear()
    b.close()


def _fixture_4096():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch_4096.json")) as f:
        return json.load(f)


def test_batch_4096_at_full_size(ctx, golden):
    """The 4096-graph mix itself (18 M vertices, 388 M entries), in all three
    launch modes, against the oracle's count of every graph (the fixture is
    made by the oracle's own generators and counter) and graphs 0-7 against
    the reference binary's counts."""
    fx = _fixture_4096()
    n = len(fx["graphs"])
    assert n == 4096 and fx["base_seed"] == SEED
    exp = np.array([x[3] for x in fx["graphs"]], np.uint64)
    b = xxtric_amd.Batch(ctx, 20_000_000, 400_000_000, n)
    _say("batch 4096 build")
    b.add_generated(SEED, 0, n, fx["nmin"], fx["nmax"])
    for mode, reps in ((XT_BATCH_PERSISTENT, 2), (XT_BATCH_HIPGRAPH, 3), (XT_BATCH_PER_GRAPH, 1)):
        for r in range(reps):
            _say("batch 4096 mode", mode, "run", r)
            b.run(mode)
            got = b.results()
            bad = np.nonzero(got != exp)[0]
            assert len(bad) == 0, (mode, r, bad[:10], got[bad[:10]], exp[bad[:10]])
    t = b.timing()
    assert t["n_graphs"] == n
    assert t["n_vertices"] == sum(x[0] for x in fx["graphs"])
    assert t["n_edges"] == sum(x[2] for x in fx["graphs"])
    for g in golden["batch"]:
        assert int(got[g["index"]]) == g["triangles"]
    b.close()


def _tiny_graphs(count, seed):
    """Graphs of 3-8 vertices with closed-form counts: K3 (1), K4 (4), K5 (10),
    paths (0), stars (0), a 4-cycle with one chord (2), two triangles sharing
    a vertex (2)."""
    from helpers import complete, csr_from_edges
    shapes = [(complete(3), 1), (complete(4), 4), (complete(5), 10),
              (csr_from_edges(4, [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]), 2),
              (csr_from_edges(5, [(0, 1), (1, 2), (2, 0), (0, 3), (3, 4), (4, 0)]), 2)]
    for k in range(3, 9):
        shapes.append((csr_from_edges(k, [(i, i + 1) for i in range(k - 1)]), 0))
    shapes.append((csr_from_edges(6, [(0, i) for i in range(1, 6)]), 0))
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(shapes), count)
    return [shapes[i][0] for i in pick], np.array([shapes[i][1] for i in pick], np.uint64)


def test_tiny_graphs_many_per_chunk(ctx):
    """Hundreds of graphs of 4-20 entries: one 2048-entry chunk of the light
    pass spans far more than its 64 LDS graph slots, so most totals take the
    direct per-graph atomic; every mode must still give the closed forms."""
    graphs, exp = _tiny_graphs(900, 11)
    # a K3 run long enough that one chunk holds 341 graphs
    from helpers import complete
    k3 = complete(3)
    graphs += [k3] * 400
    exp = np.concatenate([exp, np.ones(400, np.uint64)])
    b = xxtric_amd.Batch(ctx, 100_000, 1_000_000, len(graphs))
    for o, t in graphs:
        b.add(o, t.astype(np.int32))
    for mode in (XT_BATCH_PERSISTENT, XT_BATCH_HIPGRAPH, XT_BATCH_PER_GRAPH, XT_BATCH_HIPGRAPH):
        _say("tiny graphs mode", mode)
        b.run(mode)
        got = b.results()
        assert (got == exp).all(), (mode, np.nonzero(got != exp)[0][:10])
    b.close()
