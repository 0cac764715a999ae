"""One rank of a 1-D vertex split on one device: triangles whose minimum vertex
is owned (and of those, the ones whose middle vertex is owned too), counted on
the whole graph and on the rank's slice (owned rows plus ghost rows).

Pins: the owned count equals the oracle's per-vertex counts summed over the
owned range (min-vertex credit, xxtric.hpp:165); the local count equals a
brute-force enumeration of (minimum, middle) pairs; the ranges of a split add
up to the total; the slice gives the same numbers as the whole graph.
"""
import sys

import numpy as np
import pytest

import oracle_py as O
import xxtric_amd
from helpers import complete, min_mid_pairs, wheel

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _graphs():
    yield "rmat11p", O.rmat_csr(11, 16, 3, True)
    yield "rmat11", O.rmat_csr(11, 16, 5, False)
    yield "k60", complete(60)
    yield "wheel200", wheel(200)


@pytest.mark.parametrize("name,gr", list(_graphs()), ids=lambda x: x if isinstance(x, str) else "")
@pytest.mark.parametrize("parts", [1, 3, 8])
def test_owned_ranges(ctx, name, gr, parts):
    off, tails = gr
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    total, pv = O.count_reference(off, tails, per_vertex=True)
    us, vs, cs = min_mid_pairs(off, tails)
    assert cs.sum() == total
    g = ctx.graph(off, tails)
    bounds = [n * i // parts for i in range(parts + 1)]
    for orient in ("degree", "id"):
        s_owned = 0
        for lo, hi in zip(bounds[:-1], bounds[1:]):
            _say(name, orient, "owned", lo, hi)
            owned, local, tot = g.count_owned(lo, hi, orient=orient)
            exp_local = int(cs[(us >= lo) & (us < hi) & (vs < hi)].sum())
            assert tot == total
            assert owned == int(pv[lo:hi].sum()), (name, orient, lo, hi)
            assert local == exp_local, (name, orient, lo, hi, local, exp_local)
            s_owned += owned
        assert s_owned == total


@pytest.mark.parametrize("small_cut", [1, 1 << 30])
def test_owned_range_on_both_kernels(ctx, small_cut):
    """Every edge through the group kernel (cut 1) or the light pass (no cut)."""
    off, tails = O.rmat_csr(12, 16, 7, True)
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    _, pv = O.count_reference(off, tails, per_vertex=True)
    us, vs, cs = min_mid_pairs(off, tails)
    g = ctx.graph(off, tails)
    ctx.tune(small_cut=small_cut, wave_owner=-2)  # heavy edges through the group kernel
    try:
        for lo, hi in [(0, n // 8), (n // 3, n // 2), (n - 100, n)]:
            _say("cut", small_cut, lo, hi)
            owned, local, _ = g.count_owned(lo, hi)
            assert owned == int(pv[lo:hi].sum())
            assert local == int(cs[(us >= lo) & (us < hi) & (vs < hi)].sum())
    finally:
        ctx.tune(small_cut=xxtric_amd.DEFAULT_LIGHT_CUT, group_blocks=4, group_round=2, wave_owner=-1)


def test_slice_has_the_rank_view(ctx):
    off, tails = O.rmat_csr(12, 16, 9, True)
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    g = ctx.graph(off, tails)
    lo, hi = 0, n // 8
    _say("slice", lo, hi)
    s = g.slice(lo, hi)
    so, st = s.download()
    assert s.nv == n and s.ne < len(tails)
    keep = np.zeros(n, bool)
    keep[lo:hi] = True
    keep[tails[off[lo]:off[hi]]] = True
    deg = np.diff(off)
    assert (np.diff(so) == np.where(keep, deg, 0)).all()
    rows = np.nonzero(keep)[0]
    assert all((st[so[u]:so[u + 1]] == tails[off[u]:off[u + 1]]).all() for u in rows[:: max(1, len(rows) // 500)])
    for orient in ("degree", "id"):
        assert s.count_owned(lo, hi, orient=orient)[:2] == g.count_owned(lo, hi, orient=orient)[:2]


def test_owned_range_rejects_bad_arguments(ctx):
    off, tails = complete(10)
    g = ctx.graph(off, tails)
    for lo, hi in [(-1, 3), (5, 2), (0, 11)]:
        with pytest.raises(xxtric_amd.XtError):
            g.count_owned(lo, hi)
        with pytest.raises(xxtric_amd.XtError):
            g.slice(lo, hi)
