"""GPU parity of wave tiles (k_count_waves): owners with short out-lists counted
one wave per tile with a per-wave LDS table, for plain counts of single graphs
and batches.  Bar: bit-exact against the oracle and closed forms.
"""
import sys

import numpy as np
import pytest

import oracle_py as O
import xxtric_amd
from helpers import complete, csr_from_edges

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _restore(ctx):
    ctx.tune(small_cut=-1, group_blocks=4, group_round=2, items_per_tile=-1, tile_run=-1, wave_owner=-1,
             group_hash=-1)


@pytest.mark.parametrize("wave_owner", [1, 16, 64, 128, 256, 512])
@pytest.mark.parametrize("small_cut", [1, 32])
def test_wave_tiles_match_oracle(ctx, golden, wave_owner, small_cut):
    """R-MAT s8-s16 and RGG fixtures: plain counts with wave tiles equal the
    golden counts in both orientations; per-vertex counts (group kernel only)
    are unchanged."""
    ctx.tune(small_cut=small_cut, wave_owner=wave_owner)
    try:
        for g in golden["rmat"]:
            off, tails = O.rmat_csr(g["scale"], g["edgefactor"], g["seed"], bool(g["permute"]))
            t64 = tails.astype(np.int64)
            for orient in ("degree", "id"):
                _say("waves", wave_owner, "cut", small_cut, "rmat", g["scale"], g["seed"], orient)
                assert ctx.count_csr(off, t64, orient=orient) == g["triangles"]
        for g in golden["rgg_lcg"][:4]:
            off, tails = xxtric_amd.rgg_generate(g["n"], seed=g["seed"])
            _say("waves", wave_owner, "cut", small_cut, "rgg", g["n"])
            assert ctx.count_csr(off, tails) == g["triangles"]
        off, tails = O.rmat_csr(14, 16, 1, True)
        exp, exp_pv = O.count_degree_ordered(off, tails, per_vertex=True)
        tot, pv = ctx.count_csr(off, tails.astype(np.int64), per_vertex=True)
        assert tot == exp and (pv == exp_pv).all()
    finally:
        _restore(ctx)


@pytest.mark.parametrize("run", [1, 5, 64])
def test_wave_tiles_tile_runs_and_grid(ctx, run):
    """Tiles drawn 1, 5 or 64 at a time (an owner's table kept or rebuilt)."""
    off, tails = O.rmat_csr(15, 16, 4, True)
    exp = O.count_degree_ordered(off, tails)
    ctx.tune(small_cut=1, wave_owner=128, tile_run=run)
    try:
        _say("waves tile_run", run)
        assert ctx.count_csr(off, tails.astype(np.int64)) == exp
    finally:
        _restore(ctx)


def _colliding_ids(count, start, bucket):
    """count ids >= start whose home bucket is `bucket` in every wave table
    (top 6, 7 and 8 bits of id * 0x9E3779B1 mod 2^32; bucket < 64 is the
    6-bit index, so the 8-bit one is 4 * bucket)."""
    ids = np.arange(start, start + 800000, dtype=np.uint64)
    home = ((ids * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)) >> np.uint64(24)
    bucket = 4 * bucket
    sel = ids[home == bucket][:count]
    assert len(sel) == count
    return [int(v) for v in sel]


@pytest.mark.parametrize("k,wave_owner", [(6, 128), (20, 128), (60, 128), (20, 256), (150, 256), (300, 512)])
def test_wave_table_spills(ctx, k, wave_owner):
    """A clique on k+1 vertices whose ids all share one home bucket of the wave
    table: buckets spill, keys sit in marked slots and further buckets, and the
    count is still C(k+1, 3); the same graph with shuffled labels gives the same."""
    ids = [0] + _colliding_ids(k, 1000, 17 if wave_owner <= 128 else 40)  # in every table size
    n = max(ids) + 1
    edges = [(a, b) for i, a in enumerate(ids) for b in ids[i + 1:]]
    off, tails = csr_from_edges(n, edges)
    exp = (k + 1) * k * (k - 1) // 6
    # per-vertex (min-id credit): the j-th smallest id of the clique closes C(k - j, 2)
    exp_pv = np.zeros(n, np.uint64)
    for j, v in enumerate(ids):
        exp_pv[v] = (k - j) * (k - j - 1) // 2
    ctx.tune(small_cut=1, wave_owner=wave_owner)
    try:
        for orient in ("degree", "id"):
            _say("spill clique", k, wave_owner, orient)
            assert ctx.count_csr(off, tails, orient=orient) == exp
            tot, pv = ctx.count_csr(off, tails, orient=orient, per_vertex=True)
            assert tot == exp and (pv == exp_pv).all()
    finally:
        _restore(ctx)


def test_book_graph_many_tiny_owners(ctx):
    """Book graph: a spine edge (0, 1) and n pages, each adjacent to both: every
    page owns one item (n wave tiles of one item each), n triangles."""
    n = 50000
    edges = [(0, 1)] + [(0, i) for i in range(2, n + 2)] + [(1, i) for i in range(2, n + 2)]
    off, tails = csr_from_edges(n + 2, edges)
    ctx.tune(small_cut=1, wave_owner=64)
    try:
        _say("book graph", n)
        assert ctx.count_csr(off, tails) == n
        tot, pv = ctx.count_csr(off, tails, per_vertex=True)   # every page's triangle belongs to vertex 0
        assert tot == n and int(pv[0]) == n and int(pv.sum()) == n
        assert ctx.count_csr(off, tails, orient="id") == n
        o2, t2 = complete(40)
        assert ctx.count_csr(o2, t2) == 40 * 39 * 38 // 6
    finally:
        _restore(ctx)


def test_wave_tiles_wide_layout(ctx):
    """The layout of graphs with 2^32+ entries (64-bit item offsets) with wave tiles."""
    off, tails = O.rmat_csr(13, 16, 2, True)
    exp = O.count_degree_ordered(off, tails)
    ctx.tune(small_cut=1, wave_owner=128)
    ctx.force_wide(True)
    try:
        _say("waves wide layout")
        assert ctx.count_csr(off, tails.astype(np.int64)) == exp
    finally:
        ctx.force_wide(False)
        _restore(ctx)


def test_wave_knob_rejects_bad_values(ctx):
    for w in (513, -3):
        with pytest.raises(xxtric_amd.XtError) as e:
            ctx.tune(wave_owner=w)
        assert e.value.code == -1


def test_batch_wave_settings_agree(ctx):
    """A 256-graph slice of the batch mix: wave tiles off, and on at several
    light cuts and owner cuts, in the persistent and hipGraph modes, give the
    oracle's per-graph counts."""
    import json
    import os
    from xxtric_amd.capi import XT_BATCH_HIPGRAPH, XT_BATCH_PERSISTENT
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch_4096.json")) as f:
        fx = json.load(f)
    n = 256
    exp = np.array([x[3] for x in fx["graphs"][:n]], np.uint64)
    b = xxtric_amd.Batch(ctx, 3_000_000, 60_000_000, n)
    b.add_generated(fx["base_seed"], 0, n, fx["nmin"], fx["nmax"])
    try:
        for cut, wown in ((64, -2), (1, 64), (8, 128), (16, 32), (1, 1)):
            ctx.tune(small_cut=cut, wave_owner=wown)
            for mode in (XT_BATCH_PERSISTENT, XT_BATCH_HIPGRAPH, XT_BATCH_HIPGRAPH):
                _say("batch waves cut", cut, "owner", wown, "mode", mode)
                b.run(mode)
                assert (b.results() == exp).all()
    finally:
        _restore(ctx)
        b.close()



@pytest.mark.parametrize("wave_owner", [-2, 16, 128, 256, 512])
@pytest.mark.parametrize("small_cut", [1, 32])
def test_wave_tiles_owned_ranges(ctx, wave_owner, small_cut):
    """Owned-range counts with wave tiles (minimum and middle vertex split by
    lane masks in k_count_waves): the owned count equals the oracle's
    per-vertex counts summed over the range, the local count a brute-force
    enumeration of (minimum, middle) pairs, for the ranges of an 8-way split."""
    from helpers import min_mid_pairs
    off, tails = O.rmat_csr(12, 16, 9, True)
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    total, pv = O.count_reference(off, tails, per_vertex=True)
    us, vs, cs = min_mid_pairs(off, tails)
    g = ctx.graph(off, tails)
    ctx.tune(small_cut=small_cut, wave_owner=wave_owner)
    try:
        for r in range(8):
            lo, hi = n * r // 8, n * (r + 1) // 8
            _say("waves owned", wave_owner, "cut", small_cut, "range", lo, hi)
            owned, local, _ = g.count_owned(lo, hi, total=False)
            assert owned == int(pv[lo:hi].sum())
            assert local == int(cs[(us >= lo) & (us < hi) & (vs < hi)].sum())
    finally:
        _restore(ctx)
        g.close()


@pytest.mark.parametrize("wave_owner", [1, 16, 64, 128, 256, 512])
def test_wave_tiles_per_vertex(ctx, golden, wave_owner):
    """Per-vertex counts with wave tiles (slot, partner and owner credits kept in
    the wave's LDS and added when the owner or tile ends) equal the oracle's,
    for every edge heavy (cut 1) and at the default cut; the batch's per-graph
    totals with per-vertex output too."""
    off, tails = O.rmat_csr(15, 16, 6, True)
    exp, exp_pv = O.count_degree_ordered(off, tails, per_vertex=True)
    try:
        for cut in (1, 32):
            ctx.tune(small_cut=cut, wave_owner=wave_owner)
            for orient in ("degree", "id"):
                _say("waves per-vertex", wave_owner, "cut", cut, orient)
                tot, pv = ctx.count_csr(off, tails.astype(np.int64), per_vertex=True, orient=orient)
                assert tot == exp and (pv == exp_pv).all()
        g = golden["rgg_lcg"][3]
        o2, t2 = xxtric_amd.rgg_generate(g["n"], seed=g["seed"])
        e2, e2_pv = O.count_reference(o2, t2, per_vertex=True)
        tot, pv = ctx.count_csr(o2, t2, per_vertex=True)
        assert tot == e2 == g["triangles"] and (pv == e2_pv).all()
    finally:
        _restore(ctx)
