"""The pins that check the oracle must catch plausible mistakes.

A small pure-Python restatement of the reference loop (xxtric.hpp:60-208) and
of the RGG construction (graph.hpp:789-948, utils.hpp:245-430) takes mutation
switches.  Unmutated, it agrees with the C oracle; every mutation must be
caught by at least one pin the oracle is checked against (golden counts of the
reference binary, brute force, closed forms).  All of this runs on the CPU.
"""
from math import comb, log, sqrt

import numpy as np
import pytest

import oracle_py as O
from helpers import brute_force, complete, csr_from_edges, random_graph


def ref_loop(off, tails, mut=None):
    n = len(off) - 1
    er = [(0, 0)] * n
    for i in range(n):
        e0, e1 = off[i], off[i + 1]
        if e1 - e0 >= 2:
            er[i] = (tails[e0], tails[e1 - 1])
    pv = [0] * n
    for i in range(n):
        e0, e1 = off[i], off[i + 1]
        if e0 + 1 == e1:
            continue
        mend = e1 - 2 if mut == "skip_last_m" else e1 - 1
        for m in range(e0, mend):
            v = tails[m]
            if i >= v:
                continue
            lo, hi = (er[v][1], er[v][0]) if mut == "swap_prune" else er[v]
            nstart = m if mut == "n_from_m" else m + 1
            row = set(tails[off[v]:off[v + 1]])
            for k in range(nstart, e1):
                w = tails[k]
                if not w <= hi:
                    break
                if not w >= lo:
                    continue
                if w in row:
                    who = max(i, v, w) if mut == "credit_max" else i
                    pv[who] += 1
    return sum(pv), np.array(pv, np.uint64)


def pins_catch(off, tails, mut):
    tot, pv = ref_loop(off, tails, mut)
    bt, bpv = brute_force(off, tails)
    return tot != bt or not (pv == bpv).all()


GRAPHS = [complete(12),
          random_graph(40, 0.3, 1),
          csr_from_edges(8, [(0, 0), (0, 1), (1, 1), (1, 2), (0, 2), (2, 3), (3, 4), (2, 4), (4, 4)])]


def test_unmutated_matches_c_oracle():
    for off, tails in GRAPHS:
        tot, pv = ref_loop(off, tails)
        ct, cpv = O.count_reference(off, tails, per_vertex=True)
        assert tot == ct and (pv == cpv).all()


@pytest.mark.parametrize("mut", ["credit_max", "swap_prune", "n_from_m", "skip_last_m"])
def test_each_count_mutation_is_caught(mut):
    assert any(pins_catch(off, tails, mut) for off, tails in GRAPHS), mut


def test_credit_max_caught_by_complete_graph_closed_form():
    off, tails = complete(10)
    _, pv = ref_loop(off, tails, "credit_max")
    assert [int(x) for x in pv] != [comb(10 - 1 - i, 2) for i in range(10)]


def rgg_edges(n, mut=None):
    r = [O.reseeder(1)]
    for _ in range(2 * n - 1):
        r.append((r[-1] * 16807) % 2147483647)
    d = np.abs(np.array(r, np.float64)) * (1.0 / 2147483647.0)
    X = d[:n]
    Y = d[:n] if mut == "y_from_x_stream" else d[n:]
    rc = sqrt(log(n) / (3.14159 * n))
    rt = 0.0 if mut == "drop_rt" else sqrt(2.0736 / n)
    rn = (rc + rt) / 2.0
    dx = X[:, None] - X[None, :]
    dy = Y[:, None] - Y[None, :]
    ed = np.sqrt(dx * dx + dy * dy)
    iu = np.triu_indices(n, 1)
    return int((ed[iu] <= rn).sum()) * 2


@pytest.mark.parametrize("mut", [None, "drop_rt", "y_from_x_stream"])
def test_rgg_mutations_caught_by_reference_golden(golden, mut):
    g = [x for x in golden["rgg_lcg"] if x["n"] == 1024][0]
    ne = rgg_edges(1024, mut)
    if mut is None:
        assert ne == g["ne"]
    else:
        assert ne != g["ne"]


def test_degree_ordered_missing_apex_credit_is_caught():
    # crediting every triangle to min(u, v) instead of min(u, v, w) is a slip the
    # brute-force per-vertex pin must reject
    off, tails = random_graph(30, 0.4, 3)
    bt, bpv = brute_force(off, tails)
    tot, pv = O.count_degree_ordered(off, tails, per_vertex=True)
    assert tot == bt and (pv == bpv).all()
    n = len(off) - 1
    deg = np.diff(off)
    key = lambda x: (deg[x], x)  # noqa: E731
    wrong = np.zeros(n, np.uint64)
    for u in range(n):
        Nu = {x for x in tails[off[u]:off[u + 1]] if key(x) > key(u)}
        for v in Nu:
            Nv = {x for x in tails[off[v]:off[v + 1]] if key(x) > key(v)}
            wrong[min(u, v)] += len(Nu & Nv)
    assert int(wrong.sum()) == bt and not (wrong == bpv).all()


def _rgg_brute(X, Y, r, strict):
    """All pairs in numpy: distance sqrt(dx*dx + dy*dy), rounded like the C code (no FMA)."""
    n = len(X)
    i, j = np.triu_indices(n, 1)
    dx = X[i] - X[j]
    dy = Y[i] - Y[j]
    d = np.sqrt(dx * dx + dy * dy)
    keep = d < r if strict else d <= r
    return csr_from_edges(n, np.stack([i[keep], j[keep]], 1))


@pytest.mark.parametrize("strict", [False, True])
def test_rgg_strict_option_pinned_by_a_tie(strict):
    """or_rgg_lcg_opts: with the radius set to the exact distance of one pair, the
    strict test drops that pair and the reference's <= keeps it; both equal an
    all-pairs numpy enumeration of the same points."""
    n = 1500
    X, Y = O.rgg_lcg_coords(n, 5)
    dx, dy = X[10] - X, Y[10] - Y
    d = np.sqrt(dx * dx + dy * dy)
    j = int(np.argsort(d)[6])          # the 6th-nearest point to point 10: about 6 neighbours each
    r = float(d[j])
    o, t = O.rgg_lcg(n, seed=5, radius=r, strict=strict)
    bo, bt = _rgg_brute(X, Y, r, strict)
    assert (o == bo).all() and (t == bt).all()
    assert (j in set(t[o[10]:o[11]].tolist())) == (not strict)


def test_rgg_connect_radius_formula():
    import xxtric_amd
    for n in (1000, 100000):
        assert xxtric_amd.rgg_radius_connect(n) == sqrt(2 * log(n) / (np.pi * n))


@pytest.mark.parametrize("name", ["rmat11p", "rmat12", "rgg4096", "er", "k40"])
def test_vertex_min_credit_equals_reference(name):
    """or_vertex_min_credit32 (mark N(v), scan N(u) above u) equals the reference
    restatement's per-vertex counts on every vertex, and K_n's closed form."""
    if name == "rmat11p":
        off, tails = O.rmat_csr(11, 16, 3, True)
    elif name == "rmat12":
        off, tails = O.rmat_csr(12, 16, 5, False)
    elif name == "rgg4096":
        off, tails = O.rgg_lcg(4096, 1)
    elif name == "er":
        off, tails = O.er_csr(3000, 15, O.er_threshold(0.01))
    else:
        off, tails = complete(40)
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    _, pv = O.count_reference(off, tails, per_vertex=True)
    got = O.vertex_min_credit(off, tails, np.arange(n))
    assert (got == pv).all()
    if name == "k40":
        i = np.arange(n)
        assert (got == (n - 1 - i) * (n - 2 - i) // 2).all()
