"""Vertex partitions of a balanced read (graph.hpp:397-640, flags -b and -d).

xt_partition_vertices (host code of libxtric) against:
  * the bins of the reference's own routines, compiled from its headers by
    oracle/Makefile (oracle/_ref/ref_partition) and recorded in
    tests/golden/golden.json on fixture offsets;
  * the oracle's literal restatement (or_partition_greedy / or_partition_linear)
    on random offsets, including leading isolated vertices and the cases the
    reference leaves undefined;
  * for the dynamic program, what the mathematics fixes: its largest part is
    the smallest possible largest part over every split into nparts ranges
    (brute force on small inputs, a bisection on the answer on large ones).
No GPU: this is host logic.
"""
import itertools
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import xxtric_amd
from xxtric_amd import capi

from test_capi_cpu import TRIC


def _golden_offsets(case, cache={}):
    key = (case["graph"], tuple(sorted(case["params"].items())))
    if key not in cache:
        prm = case["params"]
        if case["graph"] == "rmat":
            off, _ = O.rmat_csr(prm["scale"], 16, prm["seed"], bool(prm["permute"]))
        elif case["graph"] == "er":
            off, _ = O.er_csr(prm["n"], prm["seed"], O.er_threshold(prm["p"]))
        elif case["graph"] == "rgg_lcg":
            off, _ = O.rgg_lcg(prm["n"])
        else:
            off = None
        cache[key] = off
    return cache[key]


def _golden_cases(golden):
    return [c for c in golden.get("partitions", []) if c["graph"] != "amazon0302"]


def test_golden_has_partitions(golden):
    cases = golden.get("partitions", [])
    assert len(cases) >= 30
    assert {c["method"] for c in cases} == {"greedy", "linear"}


@pytest.mark.parametrize("method", ["greedy", "linear"])
def test_reference_bins(golden, method):
    import hashlib
    n = 0
    for case in _golden_cases(golden):
        if case["method"] != method:
            continue
        off = _golden_offsets(case)
        assert hashlib.sha256(np.ascontiguousarray(off, np.int64).tobytes()).hexdigest() == \
            case["off_sha256"]
        want = np.array(case["bins"], np.int64)
        assert np.array_equal(O.partition(off, case["nparts"], method), want), case
        got = xxtric_amd.partition_vertices(off, case["nparts"], method)
        assert np.array_equal(got, want), case
        n += 1
    assert n >= 10


def test_reference_bins_amazon(golden):
    path = "/root/reference/cugraph/example-inputs/amazon0302_adj.tsv"
    cases = [c for c in golden.get("partitions", []) if c["graph"] == "amazon0302"]
    if not cases:
        pytest.skip("no amazon0302 bins in the fixture")
    if not os.path.exists(path):
        pytest.skip("amazon0302 input not present")
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden
    off, _ = make_golden.amazon_csr()
    for case in cases:
        got = xxtric_amd.partition_vertices(off, case["nparts"], case["method"])
        assert np.array_equal(got, np.array(case["bins"], np.int64)), case


def _random_offsets(rng, n, lead_zeros=0):
    deg = rng.integers(0, 9, n) * (rng.random(n) < 0.75)
    deg[:lead_zeros] = 0
    return np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)


@pytest.mark.parametrize("seed", range(4))
def test_product_equals_oracle_on_random_offsets(seed):
    rng = np.random.default_rng(seed)
    undefined = 0
    for trial in range(150):
        n = int(rng.integers(1, 80))
        off = _random_offsets(rng, n, lead_zeros=int(rng.integers(0, 4)))
        for k in (1, 2, 3, 5, 8, 13):
            for method in ("greedy", "linear"):
                want = O.partition(off, k, method)
                if want is None:
                    undefined += 1
                    with pytest.raises(xxtric_amd.XtError):
                        xxtric_amd.partition_vertices(off, k, method)
                    continue
                got = xxtric_amd.partition_vertices(off, k, method)
                assert np.array_equal(got, want), (method, k, off)
                assert got[0] == 0 and got[-1] == n and np.all(np.diff(got) >= 0)
    assert undefined > 0  # the undefined cases were reached and rejected


def _largest_part(seq, bins):
    return max(int(seq[bins[r]:bins[r + 1]].sum()) for r in range(len(bins) - 1))


@pytest.mark.parametrize("seed", range(3))
def test_linear_partition_is_optimal_brute_force(seed):
    # The sequence the reference partitions is offsets[0..nv): its largest part
    # must be the minimum over every way to cut it into k non-empty ranges.
    rng = np.random.default_rng(100 + seed)
    checked = 0
    for trial in range(60):
        n = int(rng.integers(3, 12))
        off = _random_offsets(rng, n)
        seq = off[:n]
        for k in range(1, min(n - 1, 5) + 1):
            bins = O.partition(off, k, "linear")
            if bins is None:
                continue
            assert np.all(np.diff(bins) >= 1)  # every rank owns a vertex
            best = min(_largest_part(seq, (0,) + cuts + (n,))
                       for cuts in itertools.combinations(range(1, n), k - 1))
            assert _largest_part(seq, bins) == best
            assert _largest_part(seq, xxtric_amd.partition_vertices(off, k, "linear")) == best
            checked += 1
    assert checked > 100


def _min_largest_part(seq, k):
    # Smallest T such that a greedy left-to-right fill uses at most k ranges.
    lo, hi = int(seq.max()), int(seq.sum())
    while lo < hi:
        t = (lo + hi) // 2
        parts, run = 1, 0
        for v in seq:
            if run + v > t:
                parts, run = parts + 1, 0
            run += int(v)
        if parts <= k:
            hi = t
        else:
            lo = t + 1
    return lo


def test_linear_partition_large_input_is_optimal():
    # 2^15 vertices: the reference's table takes O(nv^2 k); the product's
    # column-at-a-time form is O(nv k) and must reach the same optimum.
    rng = np.random.default_rng(7)
    n = 1 << 15
    off = _random_offsets(rng, n)
    seq = off[:n]
    for k in (2, 8, 33):
        bins = xxtric_amd.partition_vertices(off, k, "linear")
        assert bins[0] == 0 and bins[-1] == n and np.all(np.diff(bins) >= 1)
        assert _largest_part(seq, bins) == _min_largest_part(seq, k)


def test_greedy_hand_example():
    # degrees 3,3,3,3 over 2 ranks: ne/2 = 6; vertex 2 would make 9 > 6 and
    # opens rank 1.  A vertex above the cap opens a rank of its own, and a rank
    # already above the cap passes even a degree-0 vertex on (graph.hpp:585-595).
    off = np.array([0, 3, 6, 9, 12], np.int64)
    assert list(xxtric_amd.partition_vertices(off, 2, "greedy")) == [0, 2, 4]
    off = np.array([0, 1, 11, 11, 12], np.int64)  # degrees 1, 10, 0, 1; cap 4 over 3 ranks
    assert list(O.partition(off, 3, "greedy")) == [0, 1, 2, 4]
    assert list(xxtric_amd.partition_vertices(off, 3, "greedy")) == [0, 1, 2, 4]


def test_partition_rejects_bad_input():
    L = capi.lib()
    off = np.array([0, 2, 1, 3], np.int64)  # decreasing
    bins = np.zeros(3, np.int64)
    assert L.xt_partition_vertices(3, capi.ptr(off), 2, 0, capi.ptr(bins)) == capi.XT_ERR_INPUT
    off = np.array([0, 1, 2, 3], np.int64)
    assert L.xt_partition_vertices(3, capi.ptr(off), 0, 0, capi.ptr(bins)) == capi.XT_ERR_ARG
    assert L.xt_partition_vertices(3, capi.ptr(off), 2, 7, capi.ptr(bins)) == capi.XT_ERR_ARG
    big = np.zeros(5, np.int64)
    assert L.xt_partition_vertices(3, capi.ptr(off), 3, 1, capi.ptr(big)) == capi.XT_ERR_ARG
    assert L.xt_partition_vertices(3, capi.ptr(off), 2, 1, capi.ptr(bins)) == capi.XT_OK


def test_tric_prints_balanced_bins(golden, tmp_path):
    case = next(c for c in _golden_cases(golden)
                if c["graph"] == "rmat" and c["method"] == "linear" and c["nparts"] == 4)
    prm = case["params"]
    off, tails = O.rmat_csr(prm["scale"], 16, prm["seed"], bool(prm["permute"]))
    p = str(tmp_path / "g.bin")
    O.vite_write(p, off, tails)
    for flag, method in (("-b", "greedy"), ("-d", "linear")):
        want = next(c for c in golden["partitions"] if c["graph"] == "rmat" and
                    c["params"] == prm and c["method"] == method and c["nparts"] == 4)
        r = subprocess.run([TRIC, "-f", p, flag, "-B", "4", "-x"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        assert "Trying to balance the edge distribution while reading" in r.stdout
        line = next(ln for ln in r.stdout.splitlines() if ln.startswith("Vertex bins for 4 ranks:"))
        assert [int(t) for t in line.split(":")[1].split()] == want["bins"]
