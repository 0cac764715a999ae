"""CPU tests that pin the oracle (oracle/liboracle.so) before it is trusted.

Pins used (each chosen so a plausible slip -- a wrong comparison, index or
credited vertex -- fails at least one of them):
  * golden counts written by the reference binary itself (tests/golden/golden.json),
  * closed forms (K_n, K_{a,b}, grids, wheels),
  * brute force O(n^3) on random graphs,
  * scipy: per-vertex = row sums of (U U) .* U,
  * agreement between the reference restatement and the independent
    degree-ordered counter.
"""
import os
from math import comb

import numpy as np
import pytest

import oracle_py as O
from helpers import (bipartite, brute_force, complete, csr_from_edges, csr_sha256, grid,
                     random_graph, sparse_closed_form, wheel)

AMAZON_TSV = "/root/reference/cugraph/example-inputs/amazon0302_adj.tsv"


def test_reseeder_matches_seed_seq_definition():
    # std::seed_seq{1}.generate() first word; the reference LCG's x0 (utils.hpp:37-44).
    assert O.reseeder(1) == 1967017404


@pytest.mark.parametrize("idx", range(5))
def test_rgg_lcg_matches_reference_binary(golden, idx):
    g = golden["rgg_lcg"][idx]
    off, tails = O.rgg_lcg(g["n"], seed=g["seed"])
    assert off[-1] == g["ne"]
    assert csr_sha256(off, tails) == g["csr_sha256"]
    assert O.count_reference(off, tails) == g["triangles"]


def _ranks_cases(golden, nmax):
    return [i for i, g in enumerate(golden["rgg_ranks"]) if g["n"] <= nmax]


@pytest.mark.parametrize("idx", range(9))
def test_rgg_lcg_ranks_matches_reference_generator(golden, idx):
    """The P-rank RGG restated (LCG parallel prefix in wrapping int64, per-rank
    strips, cross-rank loops over j > i) equals, byte for byte, the CSR the
    reference's GenerateRGG builds under mpiexec -n P; the reference binary's
    count of it is the golden count."""
    g = golden["rgg_ranks"][idx]
    if g["n"] > 65536 and g["nprocs"] < 8:
        pytest.skip("O(n^2/P) restatement kept to the smaller fixtures")
    off, tails = O.rgg_lcg_ranks(g["n"], g["nprocs"], seed=g["seed"])
    assert off[-1] == g["ne"]
    assert csr_sha256(off, tails) == g["csr_sha256"]
    assert O.count_degree_ordered(off, tails) == g["triangles"]


def test_rgg_lcg_ranks_one_rank_is_the_serial_generator():
    o1, t1 = O.rgg_lcg_ranks(4096, 1)
    o2, t2 = O.rgg_lcg(4096)
    assert np.array_equal(o1, o2) and np.array_equal(t1, t2)


def test_rgg_lcg_ranks_first_numbers_wrap():
    """Rank r > 0 does not continue the serial stream: the reference's matrix
    powers are not reduced mod 2^31-1 and wrap in int64, so rank 1's first
    coordinate differs from the serial stream's number 2n/P (which a reduced
    power would give)."""
    n, p = 4096, 2
    Xs, _ = O.rgg_lcg_coords(n)          # serial stream: X = numbers 0..n-1
    Xr, Yr = O.rgg_lcg_coords_ranks(n, p)
    assert np.array_equal(Xr[:n // p], Xs[:n // p])
    # the serial continuation of rank 0's 2n/P numbers
    x = O.reseeder(1)
    for _ in range(2 * (n // p)):
        x = (x * 16807) % 2147483647
    assert Xr[n // p] != x / 2147483647.0
    assert 0.0 <= Xr.min() and Xr.max() < 1.0
    assert (Yr[:n // p] < 0.5).all() and (Yr[n // p:] >= 0.5).all()


def test_rgg_radius_reference_formula():
    # graph.hpp:789-791 at n = 1e5 (PI = 3.14159)
    assert abs(O.rgg_radius_ref(100000) - 5.3037e-3) < 1e-6


@pytest.mark.parametrize("idx", range(8))
def test_rmat_fixture_counts(golden, idx):
    g = golden["rmat"][idx]
    off, tails = O.rmat_csr(g["scale"], g["edgefactor"], g["seed"], bool(g["permute"]))
    assert off[-1] == g["ne"]
    assert csr_sha256(off, tails) == g["csr_sha256"]
    if g["scale"] <= 14:
        assert O.count_reference(off, tails) == g["triangles"]
    assert O.count_degree_ordered(off, tails) == g["triangles"]


@pytest.mark.parametrize("idx", range(3))
def test_er_fixture_counts(golden, idx):
    g = golden["er"][idx]
    off, tails = O.er_csr(g["n"], g["seed"], g["thr64"])
    assert off[-1] == g["ne"] and csr_sha256(off, tails) == g["csr_sha256"]
    assert O.count_reference(off, tails) == g["triangles"]


@pytest.mark.parametrize("idx", range(8))
def test_batch_fixture_counts(golden, idx):
    g = golden["batch"][idx]
    n, kind, gseed, thr = O.batch_params(g["base_seed"], g["index"])
    assert (n, kind) == (g["n"], g["kind"])
    if kind == 0:
        off, tails = O.rgg_lcg(n, seed=(gseed ^ (gseed >> 32)) & 0xFFFFFFFF)
    else:
        off, tails = O.er_csr(n, gseed + 3, thr)
    assert csr_sha256(off, tails) == g["csr_sha256"]
    assert O.count_reference(off, tails) == g["triangles"]


@pytest.mark.skipif(not os.path.exists(AMAZON_TSV), reason="reference example input not mounted")
def test_amazon0302(golden):
    import xxtric_amd
    off, tails = xxtric_amd.edgelist_to_csr(AMAZON_TSV)
    g = golden["amazon0302"]
    assert csr_sha256(off, tails) == g["csr_sha256"]
    assert O.count_reference(off, tails) == g["triangles"] == 717719


@pytest.mark.parametrize("n", [3, 4, 10, 40])
def test_complete_graph_closed_form(n):
    off, tails = complete(n)
    tot, pv = O.count_reference(off, tails, per_vertex=True)
    assert tot == comb(n, 3)
    assert [int(x) for x in pv] == [comb(n - 1 - i, 2) for i in range(n)]
    tot2, pv2 = O.count_degree_ordered(off, tails, per_vertex=True)
    assert tot2 == tot and (pv2 == pv).all()


@pytest.mark.parametrize("g", [bipartite(5, 7), grid(6, 9), csr_from_edges(6, [(0, 1), (1, 2), (2, 3), (3, 4)])])
def test_triangle_free_graphs(g):
    assert O.count_reference(*g) == 0
    assert O.count_degree_ordered(*g) == 0


@pytest.mark.parametrize("n", [4, 9, 50])
def test_wheel_credits_hub(n):
    off, tails = wheel(n)
    tot, pv = O.count_reference(off, tails, per_vertex=True)
    expect = n - 1 if n > 4 else 4  # W_4 = K_4
    assert tot == expect
    if n > 4:
        assert int(pv[0]) == n - 1 and int(pv[1:].sum()) == 0


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_random(seed):
    n = 20 + 7 * seed
    off, tails = random_graph(n, 0.25 + 0.05 * seed, seed)
    bt, bpv = brute_force(off, tails)
    tot, pv = O.count_reference(off, tails, per_vertex=True)
    assert tot == bt and (pv == bpv).all()
    tot2, pv2 = O.count_degree_ordered(off, tails, per_vertex=True)
    assert tot2 == bt and (pv2 == bpv).all()


def test_self_loops_do_not_count():
    off, tails = csr_from_edges(5, [(0, 0), (0, 1), (1, 2), (0, 2), (2, 2), (3, 3), (3, 4)])
    assert O.count_reference(off, tails) == 1
    assert O.count_degree_ordered(off, tails) == 1


@pytest.mark.parametrize("which", ["rgg4096", "rmat12", "er2000"])
def test_per_vertex_sparse_closed_form(golden, which):
    if which == "rgg4096":
        off, tails = O.rgg_lcg(4096)
    elif which == "rmat12":
        off, tails = O.rmat_csr(12, 16, 1, True)
    else:
        off, tails = O.er_csr(2000, 12, O.er_threshold(0.01))
    st, spv = sparse_closed_form(off, tails)
    tot, pv = O.count_reference(off, tails, per_vertex=True)
    assert tot == st and (pv == spv).all()
    tot2, pv2 = O.count_degree_ordered(off, tails, per_vertex=True)
    assert tot2 == st and (pv2 == spv).all()


def test_sampled_reference_count_adds_up():
    off, tails = O.rmat_csr(12, 16, 3, False)
    full, pv = O.count_reference(off, tails, per_vertex=True)
    parts = [O.count_reference(off, tails, vstart=s, vstride=4) for s in range(4)]
    assert sum(parts) == full


def test_listed_reference_count_matches_full():
    off, tails = O.rmat_csr(12, 16, 5, True)
    full, pv = O.count_reference(off, tails, per_vertex=True)
    verts = np.array([0, 5, 17, 4095, 1000, 2], np.int64)
    assert (O.count_reference_list(off, tails, verts) == pv[verts]).all()


def test_openmp_threads_agree():
    off, tails = O.rmat_csr(12, 16, 1, True)
    assert O.count_reference(off, tails, nthreads=1) == O.count_reference(off, tails, nthreads=8)


def test_vite_roundtrip_both_layouts(tmp_path):
    off, tails = O.rmat_csr(10, 16, 1, True)
    for lay in (False, True):
        p = str(tmp_path / ("g%d.bin" % lay))
        O.vite_write(p, off, tails, layout32=lay)
        o2, t2 = O.vite_read(p, layout32=lay)
        assert (o2 == off).all() and (t2 == tails).all()
        rec = 8 if lay else 16
        hdr = 4 if lay else 8
        assert os.path.getsize(p) == 2 * hdr + hdr * len(off) + rec * len(tails)


def test_rmat_generator_invariants():
    off, tails = O.rmat_csr(10, 16, 5, True)
    n = len(off) - 1
    rows = np.repeat(np.arange(n), np.diff(off))
    assert (rows != tails).all()                      # no self-loops
    d = np.diff(off)
    for u in range(n):                                # strictly ascending rows
        r = tails[off[u]:off[u + 1]]
        assert (np.diff(r) > 0).all()
    fw = set(zip(rows.tolist(), tails.tolist()))      # symmetric
    assert all((b, a) in fw for a, b in fw)
    assert off[-1] <= 2 * 16 * n and d.sum() == off[-1]
    # permutation is a relabelling: same degree multiset and count as unpermuted
    o2, t2 = O.rmat_csr(10, 16, 5, False)
    assert sorted(np.diff(o2)) == sorted(d)
    assert O.count_degree_ordered(o2, t2) == O.count_degree_ordered(off, tails)


def test_batch_4096_fixture_is_pinned(golden):
    """tests/golden/batch_4096.json (the oracle's count of every graph of the
    4096-graph mix) agrees with the reference binary on graphs 0-7 and with the
    reference-shaped count recomputed now on a random spread of graphs."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch_4096.json")) as f:
        fx = json.load(f)
    graphs = fx["graphs"]
    assert len(graphs) == 4096
    for g in golden["batch"]:
        n, kind, ne, tri = graphs[g["index"]]
        assert (n, kind, ne, tri) == (g["n"], g["kind"], g["ne"], g["triangles"])
    rng = np.random.default_rng(3)
    for gi in rng.choice(4096, 10, replace=False):
        n, kind, gseed, thr = O.batch_params(fx["base_seed"], int(gi), fx["nmin"], fx["nmax"])
        if kind == 0:
            off, tails = O.rgg_lcg(n, seed=(gseed ^ (gseed >> 32)) & 0xFFFFFFFF)
        else:
            off, tails = O.er_csr(n, gseed + 3, thr)
        assert graphs[gi] == [n, kind, int(off[-1]), O.count_reference(off, tails)], gi
