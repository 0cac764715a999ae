"""GPU: edge list -> symmetric CSR built on the device (xt_graph_from_edges,
xt_count_edges).  The device CSR must equal the host-built one exactly
(duplicates merged, both directions, self-loops dropped, rows ascending), and
counts must equal the oracle's."""
import os
import sys

import numpy as np
import pytest

import oracle_py as O
import xxtric_amd
from helpers import csr_from_edges

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _messy_edges(n, m, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    # repeat some pairs, in both directions, and add self-loops
    k = m // 5
    src = np.concatenate([src, dst[:k], src[:k], np.arange(10) % n])
    dst = np.concatenate([dst, src[:k], dst[:k], np.arange(10) % n])
    return src.astype(np.int64), dst.astype(np.int64)


@pytest.mark.parametrize("n,m,seed", [(2, 3, 1), (1000, 5000, 2), (5000, 80000, 3), (1 << 14, 200000, 4)])
def test_device_csr_equals_host_csr(ctx, n, m, seed):
    src, dst = _messy_edges(n, m, seed)
    pairs = [(int(a), int(b)) for a, b in zip(src, dst) if a != b]
    exp_off, exp_tails = csr_from_edges(n, pairs)
    _say("from_edges", n, m)
    g = ctx.graph_from_edges(n, src, dst)
    off, tails = g.download()
    assert (off == exp_off).all() and (tails.astype(np.int64) == exp_tails).all()
    exp, exp_pv = O.count_reference(exp_off, exp_tails, per_vertex=True)
    tot, pv = g.count(per_vertex=True)
    assert tot == exp and (pv == exp_pv).all()
    assert ctx.count_edges(n, src, dst) == exp
    tot2, pv2 = ctx.count_edges(n, src, dst, per_vertex=True, orient="id")
    assert tot2 == exp and (pv2 == exp_pv).all()
    g.close()


def test_rmat_directed_list_round_trip(ctx):
    off, tails = O.rmat_csr(14, 16, 6, True)
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    up = rows < tails  # one direction of each edge, shuffled
    src, dst = rows[up], tails[up]
    perm = np.random.default_rng(9).permutation(len(src))
    _say("rmat14 round trip")
    g = ctx.graph_from_edges(n, dst[perm], src[perm])
    o2, t2 = g.download()
    assert (o2 == off).all() and (t2.astype(np.int64) == tails).all()
    assert g.count() == O.count_degree_ordered(off, tails)
    g.close()


def test_empty_and_bad_edge_lists(ctx):
    g = ctx.graph_from_edges(5, np.zeros(0, np.int64), np.zeros(0, np.int64))
    assert g.ne == 0 and g.count() == 0
    g.close()
    assert ctx.count_edges(3, [0, 1, 2, 2], [0, 1, 2, 2]) == 0  # self-loops only
    with pytest.raises(xxtric_amd.XtError) as e:
        ctx.graph_from_edges(4, [0, 1], [1, 4])  # id 4 outside [0, 4)
    assert e.value.code == -4
    with pytest.raises(xxtric_amd.XtError) as e:
        ctx.graph_from_edges(0, [0], [0])
    assert e.value.code == -1


@pytest.mark.parametrize("layout32", [False, True])
def test_vite_streamed_to_device(ctx, tmp_path, layout32):
    """A vite file streamed into HBM (blocks of 4 M elements through pinned
    buffers) equals the host reader's CSR; counts equal the oracle's.  A graph
    of 5 M ids spans two id blocks."""
    off, tails = O.rmat_csr(18, 10, 8, True)
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    assert len(tails) > (1 << 22)
    path = str(tmp_path / ("g32.bin" if layout32 else "g64.bin"))
    xxtric_amd.vite_write(path, off, tails, layout32=layout32)
    _say("vite stream", layout32, len(tails))
    g = ctx.graph_from_vite(path, layout32=layout32, validate=True, symmetry=True)
    o2, t2 = g.download()
    assert (o2 == off).all() and (t2.astype(np.int64) == tails).all()
    assert g.count() == O.count_degree_ordered(off, tails)
    g.close()


def test_vite_stream_rejects_bad_files(ctx, tmp_path):
    off = np.array([0, 2, 3, 4], np.int64)
    tails = np.array([1, 2, 0, 0], np.int64)
    good = str(tmp_path / "ok.bin")
    xxtric_amd.vite_write(good, off, tails)
    assert ctx.graph_from_vite(good).ne == 4
    raw = open(good, "rb").read()
    short = str(tmp_path / "short.bin")
    open(short, "wb").write(raw[:-8])  # last record cut
    with pytest.raises(xxtric_amd.XtError) as e:
        ctx.graph_from_vite(short)
    assert e.value.code == -4
    bad = str(tmp_path / "badid.bin")
    xxtric_amd.vite_write(bad, off, np.array([1, 2, 0, 9], np.int64))  # id 9 outside [0, 3)
    with pytest.raises(xxtric_amd.XtError) as e:
        ctx.graph_from_vite(bad)
    assert e.value.code == -4
    with pytest.raises(xxtric_amd.XtError):
        ctx.graph_from_vite(str(tmp_path / "missing.bin"))


def _simple_csr(n, src, dst):
    """Sorted, unique, symmetric CSR without self-loops from directed pairs (numpy)."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    keep = src != dst
    a = np.concatenate([src[keep], dst[keep]])
    b = np.concatenate([dst[keep], src[keep]])
    key = np.unique(a * n + b)
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, key // n + 1, 1)
    return np.cumsum(off), key % n


def test_vite_clean_repairs_a_messy_file(ctx, tmp_path):
    """Rows out of order, duplicate entries, self-loops and missing reverse
    entries: XT_VALIDATE rejects the file; XT_CLEAN rebuilds it on the device as
    the simple undirected graph it describes (checked against numpy)."""
    off, tails = O.rmat_csr(12, 16, 5, True)
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    rng = np.random.default_rng(7)
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    dst = tails.copy()
    drop = rng.random(len(dst)) < 0.2                  # some entries lose their row copy
    dup = rng.random(len(dst)) < 0.1                   # some appear twice
    loops = rng.integers(0, n, 300)                    # self-loops
    s2 = np.concatenate([src[~drop], src[dup], loops])
    d2 = np.concatenate([dst[~drop], dst[dup], loops])
    perm = rng.permutation(len(s2))                    # rows keep their ids, entries shuffled
    s2, d2 = s2[perm], d2[perm]
    order = np.argsort(s2, kind="stable")              # grouped by row, unsorted inside
    s2, d2 = s2[order], d2[order]
    moff = np.zeros(n + 1, np.int64)
    np.add.at(moff, s2 + 1, 1)
    moff = np.cumsum(moff)
    path = str(tmp_path / "messy.bin")
    xxtric_amd.vite_write(path, moff, d2)
    with pytest.raises(xxtric_amd.XtError) as e:
        ctx.graph_from_vite(path, validate=True)
    assert e.value.code == -4
    eo, et = _simple_csr(n, s2, d2)
    _say("vite clean", len(d2))
    g = ctx.graph_from_vite(path, clean=True, validate=True, symmetry=True)
    o2, t2 = g.download()
    assert (o2 == eo).all() and (t2.astype(np.int64) == et).all()
    assert g.count() == O.count_degree_ordered(eo, et)
    g.close()
    # the same through an uploaded CSR
    g = ctx.graph(moff, d2, clean=True, symmetry=True)
    o3, t3 = g.download()
    assert (o3 == eo).all() and (t3.astype(np.int64) == et).all()
    g.close()


def test_vite_rmat_s22_file_at_scale(ctx, tmp_path):
    """An R-MAT s22 graph (128 M entries) written as a vite file in the
    reference's int64 layout (2 GB) streams back into HBM unchanged; the same
    file with every row reversed comes back sorted under XT_CLEAN."""
    g = ctx.rmat(22, 16, 1, True)
    off, tails = g.download()
    exp = g.count()
    g.close()
    tails = tails.astype(np.int64)
    path = str(tmp_path / "s22.bin")
    _say("vite s22 write", len(tails))
    xxtric_amd.vite_write(path, off, tails)
    _say("vite s22 stream")
    g = ctx.graph_from_vite(path, validate=True, symmetry=True)
    o2, t2 = g.download()
    assert (o2 == off).all() and np.array_equal(t2.astype(np.int64), tails)
    assert g.count() == exp
    g.close()
    os.unlink(path)
    rows = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off))
    rev = tails[off[rows] + off[rows + 1] - 1 - np.arange(len(tails), dtype=np.int64)]
    del rows
    path = str(tmp_path / "s22rev.bin")
    xxtric_amd.vite_write(path, off, rev)
    del rev
    _say("vite s22 reversed rows, clean")
    g = ctx.graph_from_vite(path, clean=True, validate=True)
    o3, t3 = g.download()
    assert (o3 == off).all() and np.array_equal(t3.astype(np.int64), tails)
    assert g.count() == exp
    g.close()
    os.unlink(path)
