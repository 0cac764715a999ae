"""GPU parity: the HIP path against the oracle and the reference's golden counts.

Bar: bit-exact global and per-vertex counts (integer work).  Each case prints
its name before the device call so a failing run names it.
"""
import sys

import numpy as np
import pytest

import oracle_py as O
import xxtric_amd
from helpers import bipartite, complete, csr_from_edges, csr_sha256, grid, random_graph, wheel

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _check_csr(ctx, off, tails, name, per_vertex=True):
    exp, exp_pv = O.count_reference(off, tails, per_vertex=True)
    for orient in ("degree", "id"):
        _say(name, orient, "nv=%d ne=%d" % (len(off) - 1, len(tails)))
        tot, pv = ctx.count_csr(off, tails, per_vertex=True, orient=orient, validate=True)
        assert tot == exp, (name, orient, tot, exp)
        assert (pv == exp_pv).all(), (name, orient, np.nonzero(pv != exp_pv)[0][:10])
        assert ctx.count_csr(off, tails, orient=orient) == exp
    return exp


@pytest.mark.parametrize("idx", range(5))
def test_rgg_lcg_golden(ctx, golden, idx):
    g = golden["rgg_lcg"][idx]
    off, tails = xxtric_amd.rgg_generate(g["n"], seed=g["seed"])
    assert csr_sha256(off, tails) == g["csr_sha256"]
    assert _check_csr(ctx, off, tails, "rgg%d" % g["n"]) == g["triangles"]


@pytest.mark.parametrize("idx", range(8))
def test_rmat_golden_host_csr(ctx, golden, idx):
    g = golden["rmat"][idx]
    off, tails = O.rmat_csr(g["scale"], g["edgefactor"], g["seed"], bool(g["permute"]))
    if g["scale"] <= 14:
        assert _check_csr(ctx, off, tails.astype(np.int64), "rmat%d" % g["scale"]) == g["triangles"]
    else:
        _say("rmat host", g["scale"])
        exp, exp_pv = O.count_degree_ordered(off, tails, per_vertex=True)
        tot, pv = ctx.count_csr(off, tails.astype(np.int64), per_vertex=True)
        assert tot == exp == g["triangles"] and (pv == exp_pv).all()


@pytest.mark.parametrize("idx", range(8))
def test_rmat_device_generator(ctx, golden, idx):
    g = golden["rmat"][idx]
    _say("rmat device gen", g["scale"], g["seed"], g["permute"])
    dg = ctx.rmat(g["scale"], g["edgefactor"], g["seed"], bool(g["permute"]))
    off, tails = dg.download()
    assert dg.ne == g["ne"]
    assert csr_sha256(off, tails) == g["csr_sha256"]
    tot, pv = dg.count(per_vertex=True)
    assert tot == g["triangles"]
    exp_tot, exp_pv = O.count_degree_ordered(off, tails, per_vertex=True)
    assert (pv == exp_pv).all()
    assert dg.count(orient="id") == g["triangles"]
    dg.close()


@pytest.mark.parametrize("idx", range(3))
def test_er_device_generator(ctx, golden, idx):
    g = golden["er"][idx]
    _say("er device gen", g["n"])
    dg = ctx.er(g["n"], g["seed"], g["thr64"])
    off, tails = dg.download()
    assert csr_sha256(off, tails) == g["csr_sha256"]
    assert dg.count() == g["triangles"]
    dg.close()


@pytest.mark.parametrize("idx", range(8))
def test_batch_graphs_golden(ctx, golden, idx):
    g = golden["batch"][idx]
    off, tails = xxtric_amd.batch_graph_csr(g["base_seed"], g["index"])
    assert csr_sha256(off, tails) == g["csr_sha256"]
    assert _check_csr(ctx, off, tails.astype(np.int64), "batch%d" % idx) == g["triangles"]


@pytest.mark.parametrize("name,graph", [
    ("K3", complete(3)), ("K40", complete(40)), ("K300", complete(300)),
    ("bip", bipartite(30, 40)), ("grid", grid(20, 30)), ("wheel", wheel(500)),
    ("path", csr_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)])),
    ("loops", csr_from_edges(6, [(0, 0), (0, 1), (1, 2), (0, 2), (2, 2), (5, 5), (3, 4)])),
    ("isolated", csr_from_edges(50, [(10, 20), (20, 30), (10, 30), (40, 41)])),
    ("single_edge", csr_from_edges(2, [(0, 1)])),
    ("no_edges", (np.zeros(8, np.int64), np.zeros(0, np.int64))),
    ("rand", random_graph(300, 0.2, 7)),
])
def test_small_graphs(ctx, name, graph):
    _check_csr(ctx, graph[0], graph[1], name)


def test_empty_graph(ctx):
    _say("empty")
    assert ctx.count_csr(np.zeros(1, np.int64), np.zeros(0, np.int64)) == 0


def test_star_hub(ctx):
    n = 200001  # one hub joined to every other vertex: a 200000-entry row, 0 triangles
    off = np.zeros(n + 1, np.int64)
    off[1] = n - 1
    off[2:] = n - 1 + np.arange(1, n)
    tails = np.concatenate([np.arange(1, n), np.zeros(n - 1, np.int64)])
    _say("star")
    assert ctx.count_csr(off, tails, validate=True) == 0
    # add a long path among the leaves: n-2 triangles, all through the hub
    e = [(0, i) for i in range(1, n)] + [(i, i + 1) for i in range(1, n - 1)]
    o2, t2 = csr_from_edges(n, e)
    tot, pv = ctx.count_csr(o2, t2, per_vertex=True)
    assert tot == n - 2 and int(pv[0]) == n - 2
    # id orientation gives the hub a 200000-entry out-list: too long for the LDS
    # hash, so its group tiles search it in global memory
    tot, pv = ctx.count_csr(o2, t2, per_vertex=True, orient="id")
    assert tot == n - 2 and int(pv[0]) == n - 2 and ctx.timing()["n_heavy"] > 0


@pytest.mark.parametrize("small_cut", [1, 4, 64, 1 << 30])
def test_kernel_split_does_not_change_counts(ctx, small_cut):
    off, tails = O.rmat_csr(14, 16, 1, True)
    exp, exp_pv = O.count_degree_ordered(off, tails, per_vertex=True)
    ctx.tune(small_cut=small_cut, wave_owner=-2)  # heavy edges through the group kernel
    try:
        _say("small_cut", small_cut)
        tot, pv = ctx.count_csr(off, tails.astype(np.int64), per_vertex=True)
        assert tot == exp and (pv == exp_pv).all()
        t = ctx.timing()
        if small_cut == 1:
            assert t["n_heavy"] > 0 and t["n_tiles"] > 0
        for gb in (1, 3):  # persistent grid size of the group kernel
            ctx.tune(group_blocks=gb)
            assert ctx.count_csr(off, tails.astype(np.int64)) == exp
        for gr in (1, 2):  # chunks in flight per group-kernel thread
            ctx.tune(group_blocks=4, group_round=gr)
            tot, pv = ctx.count_csr(off, tails.astype(np.int64), per_vertex=True)
            assert tot == exp and (pv == exp_pv).all()
    finally:
        ctx.tune(small_cut=xxtric_amd.DEFAULT_LIGHT_CUT, group_blocks=4, group_round=2, wave_owner=-1)


@pytest.mark.parametrize("ipt,run", [(64, 1), (128, 64), (256, 7), (64, 16)])
def test_tile_knobs_do_not_change_counts(ctx, ipt, run):
    """Items per group tile and tiles per queue draw: global, per-vertex and
    owned-range counts unchanged (every edge heavy, so every edge is tiled)."""
    off, tails = O.rmat_csr(14, 16, 3, True)
    exp, exp_pv = O.count_degree_ordered(off, tails, per_vertex=True)
    n = len(off) - 1
    ctx.tune(small_cut=1, items_per_tile=ipt, tile_run=run, wave_owner=-2)  # group tiles only
    try:
        _say("items_per_tile", ipt, "tile_run", run)
        tot, pv = ctx.count_csr(off, tails.astype(np.int64), per_vertex=True)
        assert tot == exp and (pv == exp_pv).all()
        t = ctx.timing()
        assert t["n_tiles"] >= (t["n_heavy"] + ipt - 1) // ipt
        g = ctx.graph(off, tails)
        owned, _, _ = g.count_owned(n // 4, n // 2, total=False)
        assert owned == int(exp_pv[n // 4: n // 2].sum())
        g.close()
    finally:
        ctx.tune(small_cut=xxtric_amd.DEFAULT_LIGHT_CUT, items_per_tile=-1, tile_run=-1, wave_owner=-1)


def test_tile_knobs_reject_bad_values(ctx):
    for ipt, run in ((100, 0), (512, 0), (0, 65), (0, -2)):
        with pytest.raises(xxtric_amd.XtError) as e:
            ctx.tune(items_per_tile=ipt, tile_run=run)
        assert e.value.code == -1


def test_autotune_search_keeps_counts(ctx, tmp_path):
    """The autotuner's coordinate descent over a small space: every setting it
    times gives the same count, and the chosen one is applied and cached."""
    import json
    from xxtric_amd import autotune
    g = ctx.rmat(13, 16, 2, True)
    exp = g.count()
    space = {"light_cut": (16, 64), "group_round": (1, 2), "group_blocks": (4,), "items_per_tile": (64, 256),
             "tile_run": (4, 16)}
    cache = str(tmp_path / "tuning.json")
    runner = lambda: autotune._time_graph(ctx, g, 1)  # noqa: E731
    try:
        best, trials = autotune.search(ctx, runner, autotune.DEFAULTS, space=space, passes=2, log=_say)
        assert len(trials) >= 5 and best["ms"] == min(t["ms"] for t in trials)
        assert g.count() == exp
        e = autotune.tune(ctx, "rmat", 13, reps=1, passes=1, graph=g, cache=cache, log=_say)
        with open(cache) as f:
            assert json.load(f)["rmat_s13"]["best"] == e["best"]
        assert autotune.apply_cached(ctx, "rmat_s13", cache=cache) is not None
        assert g.count() == exp
    finally:
        ctx.tune(small_cut=xxtric_amd.DEFAULT_LIGHT_CUT, group_blocks=4, group_round=2, items_per_tile=-1,
                 tile_run=-1)
        g.close()


@pytest.mark.parametrize("n", [1500, 4098])
def test_complete_graph_fills_the_lds_hash(ctx, n):
    """K_n: out-lists of up to n-1 ids, so the group kernel's LDS hash runs at
    up to two ids per bucket (spilled buckets, continued lookups); at n = 4098
    vertex 0's list of 4097 ids is past the hash and is searched in global
    memory.  Closed forms: C(n,3) triangles, vertex i credited C(n-1-i, 2)."""
    _say("complete", n)
    off = np.arange(n + 1, dtype=np.int64) * (n - 1)
    cols = np.tile(np.arange(n, dtype=np.int64), n).reshape(n, n)
    tails = cols[~np.eye(n, dtype=bool)].reshape(-1)
    tot, pv = ctx.count_csr(off, tails, per_vertex=True)
    i = np.arange(n, dtype=np.int64)
    assert tot == n * (n - 1) * (n - 2) // 6
    assert (pv == (n - 1 - i) * (n - 2 - i) // 2).all()
    assert ctx.timing()["n_heavy"] > 0
    assert ctx.count_csr(off, tails) == tot


def _group_colliding_ids(count, start, bucket):
    """count ids >= start whose home bucket in the group kernel's 2048-bucket
    table is `bucket` (top 11 bits of id * 0x9E3779B1 mod 2^32)."""
    ids = np.arange(start, start + 2_000_000, dtype=np.uint64)
    home = ((ids * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)) >> np.uint64(21)
    sel = ids[home == bucket][:count]
    assert len(sel) == count
    return [int(v) for v in sel]


@pytest.mark.parametrize("group_hash", [1, 2])
@pytest.mark.parametrize("k", [5, 40, 300])
def test_group_table_spills(ctx, k, group_hash):
    """A clique on k+1 vertices whose ids share one home bucket of the group
    kernel's table (under its multiplicative hash, 1, or its low-bits hash, 2:
    ids 2048 apart): the bucket spills, keys sit in its marked slot and in the
    buckets after it (up to 75 of them), and lookups walk the chain.  Closed
    forms: C(k+1, 3) triangles, the j-th smallest id credited C(k - j, 2)."""
    coll = _group_colliding_ids(k, 1000, 1234) if group_hash == 1 else [1000 + 2048 * i for i in range(k)]
    ids = [0] + coll
    n = max(ids) + 1
    edges = [(a, b) for i, a in enumerate(ids) for b in ids[i + 1:]]
    off, tails = csr_from_edges(n, edges)
    exp = (k + 1) * k * (k - 1) // 6
    exp_pv = np.zeros(n, np.uint64)
    for j, v in enumerate(ids):
        exp_pv[v] = (k - j) * (k - j - 1) // 2
    ctx.tune(small_cut=1, wave_owner=-2, group_hash=group_hash)  # every edge heavy, every owner grouped
    try:
        for orient in ("degree", "id"):
            _say("group spill clique", k, group_hash, orient)
            assert ctx.count_csr(off, tails, orient=orient) == exp
            tot, pv = ctx.count_csr(off, tails, orient=orient, per_vertex=True)
            assert tot == exp and (pv == exp_pv).all()
    finally:
        ctx.tune(small_cut=-1, wave_owner=-1, group_hash=-1)


@pytest.mark.parametrize("group_hash", [1, 2])
def test_group_hash_modes_match_oracle(ctx, golden, group_hash):
    """Both table hashes give the golden counts on the R-MAT fixtures (both
    orientations), the oracle's per-vertex counts at s14 and the oracle's
    owned-range sums on a device graph."""
    ctx.tune(group_hash=group_hash)
    try:
        for g in golden["rmat"]:
            off, tails = O.rmat_csr(g["scale"], g["edgefactor"], g["seed"], bool(g["permute"]))
            for orient in ("degree", "id"):
                _say("hash", group_hash, "rmat", g["scale"], g["seed"], orient)
                assert ctx.count_csr(off, tails.astype(np.int64), orient=orient) == g["triangles"]
        off, tails = O.rmat_csr(14, 16, 1, True)
        exp, exp_pv = O.count_degree_ordered(off, tails, per_vertex=True)
        tot, pv = ctx.count_csr(off, tails.astype(np.int64), per_vertex=True)
        assert tot == exp and (pv == exp_pv).all()
        dg = ctx.rmat(13, 16, 4, True)
        o13, t13 = dg.download()
        e13, pv13 = O.count_degree_ordered(o13, t13, per_vertex=True)
        lo, hi = 1000, 3000
        owned, _, tot13 = dg.count_owned(lo, hi)
        assert tot13 == e13 and owned == int(pv13[lo:hi].sum())
        dg.close()
    finally:
        ctx.tune(group_hash=-1)


def test_rmat_unpermuted_labels(ctx):
    """R-MAT s16 without the label permutation (ids mostly zero bits, hubs at
    low ids): counts and per-vertex counts equal the oracle's, in both
    orientations, with and without wave tiles, under both table hashes (the
    low-bits hash spills heavily on these labels)."""
    off, tails = O.rmat_csr(16, 16, 2, False)
    exp, exp_pv = O.count_degree_ordered(off, tails, per_vertex=True)
    t64 = tails.astype(np.int64)
    try:
        for gh in (1, 2):
            for wo in (-2, 256):
                ctx.tune(wave_owner=wo, group_hash=gh)
                for orient in ("degree", "id"):
                    _say("rmat16 unpermuted", gh, wo, orient)
                    tot, pv = ctx.count_csr(off, t64, per_vertex=True, orient=orient)
                    assert tot == exp and (pv == exp_pv).all()
    finally:
        ctx.tune(wave_owner=-1, group_hash=-1)


def test_per_vertex_owner_with_many_tiles(ctx):
    """Id orientation gives vertex 2000, joined to every other vertex, an
    out-list of 4000 ids (inside the LDS hash) and 6000 items: 24 group tiles of
    one owner, more than the 16 after which the per-vertex slot counters
    (16-bit, in LDS) are added to HBM while the owner's table stays.  Random
    edges among all vertices close triangles credited to the owner, to the
    item's partner and to the table id."""
    n, x = 6001, 2000
    rng = np.random.default_rng(11)
    a = rng.integers(0, n, 180000)
    b = rng.integers(0, n, 180000)
    a = np.concatenate([a, np.full(n, x)])
    b = np.concatenate([b, np.arange(n)])
    keep = a != b
    lo, hi = np.minimum(a, b)[keep], np.maximum(a, b)[keep]
    key = np.unique(np.concatenate([lo * n + hi, hi * n + lo]))
    src, dst = key // n, key % n
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off)
    tails = dst.astype(np.int64)
    exp, exp_pv = O.count_reference(off, tails, per_vertex=True)
    _say("owner with many tiles", n, len(tails))
    tot, pv = ctx.count_csr(off, tails, per_vertex=True, orient="id")
    assert ctx.timing()["n_tiles"] >= 24
    assert tot == exp and (pv == exp_pv).all()
    tot, pv = ctx.count_csr(off, tails, per_vertex=True)  # degree orientation too
    assert tot == exp and (pv == exp_pv).all()


def test_invalid_inputs_are_rejected(ctx):
    off, tails = csr_from_edges(4, [(0, 1), (1, 2), (0, 2), (2, 3)])
    bad = tails.copy()
    bad[0], bad[1] = bad[1], bad[0]                   # unsorted row 0
    _say("invalid: unsorted")
    with pytest.raises(xxtric_amd.XtError) as e:
        ctx.count_csr(off, bad, validate=True)
    assert e.value.code == -4
    dup_off = off.copy()
    dup = tails.copy()
    dup[1] = dup[0]                                   # duplicate entry in row 0
    with pytest.raises(xxtric_amd.XtError):
        ctx.count_csr(dup_off, dup, validate=True)
    oob = tails.copy()
    oob[-1] = 99                                      # id out of range: rejected even without validation
    with pytest.raises(xxtric_amd.XtError):
        ctx.count_csr(off, oob)
    asym_off = np.array([0, 2, 3, 4, 4], np.int64)
    asym = np.array([1, 2, 2, 0], np.int64)           # (1,2) without (2,1) ...
    with pytest.raises(xxtric_amd.XtError):
        ctx.count_csr(asym_off, asym, symmetry=True)
    badoff = off.copy()
    badoff[2] = badoff[3] + 1
    with pytest.raises(xxtric_amd.XtError):
        ctx.count_csr(badoff, tails)
    # the context still works afterwards
    assert ctx.count_csr(off, tails) == 1


def test_reference_ffi_shapes(ctx):
    """xt_single_gpu_count / _masked / local_count_gpu_inner on device int64 arrays."""
    off, tails = O.rmat_csr(12, 16, 1, True)
    tails = tails.astype(np.int64)
    exp = O.count_reference(off, tails)
    zero = np.zeros(1, np.int64)
    d_ind = ctx.to_device(off)
    d_el = ctx.to_device(tails)
    d_cnt = ctx.to_device(zero)
    _say("ffi single_gpu_count")
    ctx.single_gpu_count(len(off) - 1, d_ind.ptr, d_el.ptr, d_cnt.ptr)
    ctx.single_gpu_count(len(off) - 1, d_ind.ptr, d_el.ptr, d_cnt.ptr)
    ctx.sync()
    assert int(d_cnt.download(np.int64, 1)[0]) == 2 * exp   # adds into the caller's counter
    # host-built edge mask as xxtric-gpu.hpp:91-107: (e1, m) for i < N[m], m != e1-1
    mask = []
    nv = len(off) - 1
    for i in range(nv):
        e0, e1 = off[i], off[i + 1]
        if e0 + 1 == e1:
            continue
        for m in range(e0, e1):
            if i >= tails[m]:
                continue
            if m != e1 - 1:
                mask += [e1, m]
    d_mask = ctx.to_device(np.array(mask, np.int64))
    d_cnt.upload(zero)
    _say("ffi masked")
    ctx.single_gpu_count_masked(len(mask) // 2, d_mask.ptr, d_ind.ptr, d_el.ptr, d_cnt.ptr)
    ctx.sync()
    assert int(d_cnt.download(np.int64, 1)[0]) == exp
    # one launch per oriented edge, as USE_SPLIT_KERNELS (xxtric-gpu.hpp:204-227), on a small graph
    o2, t2 = random_graph(60, 0.3, 3)
    exp2 = O.count_reference(o2, t2)
    d_el2 = ctx.to_device(np.ascontiguousarray(t2, np.int64))
    d_cnt.upload(zero)
    _say("ffi inner")
    for i in range(len(o2) - 1):
        e0, e1 = o2[i], o2[i + 1]
        for m in range(e0, e1 - 1):
            v = t2[m]
            if i >= v:
                continue
            ctx.local_count_gpu_inner(int(m + 1), int(e1), int(o2[v]), int(o2[v + 1]), d_el2.ptr,
                                      d_cnt.ptr)
    ctx.sync()
    assert int(d_cnt.download(np.int64, 1)[0]) == exp2
    for b in (d_ind, d_el, d_cnt, d_mask, d_el2):
        b.close()


def _hip():
    """The HIP runtime libxtric.so already loaded (same soname, same handle)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.c_void_p, ctypes.c_uint]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def test_reference_ffi_streams(ctx):
    """The launchers on the caller's streams, passed as the reference passes
    them: the inner launcher gets the ADDRESS of one of eight streams
    (xxtric-gpu.hpp:201-227, &streams[id % MAX_CUDA_STREAMS]), the masked and
    whole-graph launchers a stream value."""
    import ctypes
    hip = _hip()
    streams = (ctypes.c_void_p * 8)()
    for t in range(8):
        assert hip.hipStreamCreateWithFlags(ctypes.byref(streams, t * ctypes.sizeof(ctypes.c_void_p)), 1) == 0
    o2, t2 = random_graph(80, 0.3, 5)
    t2 = np.ascontiguousarray(t2, np.int64)
    exp2 = O.count_reference(o2, t2)
    d_ind = ctx.to_device(np.ascontiguousarray(o2, np.int64))
    d_el = ctx.to_device(t2)
    zero = np.zeros(1, np.int64)
    d_cnt = ctx.to_device(zero)
    _say("ffi inner on stream addresses")
    nst = 0
    for i in range(len(o2) - 1):
        e0, e1 = o2[i], o2[i + 1]
        for m in range(e0, e1 - 1):
            v = t2[m]
            if i >= v:
                continue
            addr = ctypes.addressof(streams) + (nst % 8) * ctypes.sizeof(ctypes.c_void_p)
            nst += 1
            ctx.local_count_gpu_inner(int(m + 1), int(e1), int(o2[v]), int(o2[v + 1]), d_el.ptr,
                                      d_cnt.ptr, addr)
    for t in range(8):
        assert hip.hipStreamSynchronize(streams[t]) == 0
    assert int(d_cnt.download(np.int64, 1)[0]) == exp2
    _say("ffi single_gpu_count on a stream value")
    d_cnt.upload(zero)
    ctx.single_gpu_count(len(o2) - 1, d_ind.ptr, d_el.ptr, d_cnt.ptr, streams[3])
    # the call returns once the count has been added: no synchronisation here
    assert int(d_cnt.download(np.int64, 1)[0]) == exp2
    for t in range(8):
        assert hip.hipStreamDestroy(streams[t]) == 0
    for b in (d_ind, d_el, d_cnt):
        b.close()


def test_single_gpu_count_rejects_wide_vertex_counts(ctx):
    """nv >= 2^31 - 1 is refused before anything is read from the device
    arrays (the device ids are int32)."""
    d_ind = ctx.to_device(np.zeros(4, np.int64))
    d_cnt = ctx.to_device(np.zeros(1, np.int64))
    _say("ffi single_gpu_count nv=2^31")
    with pytest.raises(xxtric_amd.XtError) as e:
        ctx.single_gpu_count(1 << 31, d_ind.ptr, d_ind.ptr, d_cnt.ptr)
    assert e.value.code == -1
    assert int(d_cnt.download(np.int64, 1)[0]) == 0
    d_ind.close()
    d_cnt.close()


def test_tric_driver(golden, tmp_path):
    """The reference-compatible driver: same flags, same output lines."""
    import os
    import re
    import subprocess
    tric = os.path.join(os.path.dirname(xxtric_amd.__file__), "bin", "tric")
    g = [x for x in golden["rgg_lcg"] if x["n"] == 100000][0]
    _say("tric -n 100000 -l")
    out = subprocess.run([tric, "-n", "100000", "-l"], capture_output=True, text=True, check=True,
                         timeout=120).stdout
    assert "Number of edges: %d" % g["ne"] in out
    assert "Number of triangles: %d" % g["triangles"] in out
    assert re.search(r"TEPS: [0-9.e+]+", out)
    # vite input with per-vertex output
    off, tails = O.rmat_csr(12, 16, 1, True)
    p = str(tmp_path / "g.bin")
    O.vite_write(p, off, tails)
    vf = str(tmp_path / "pv.txt")
    _say("tric -f vite -v")
    out = subprocess.run([tric, "-f", p, "-v", vf, "-m", "id"], capture_output=True, text=True,
                         check=True, timeout=120).stdout
    exp, exp_pv = O.count_reference(off, tails, per_vertex=True)
    assert "Number of triangles: %d" % exp in out
    assert np.array_equal(np.loadtxt(vf, dtype=np.uint64), exp_pv)
    # device R-MAT
    r = [x for x in golden["rmat"] if x["scale"] == 16 and x["permute"] == 1][0]
    _say("tric -R 16")
    out = subprocess.run([tric, "-R", "16", "-S", "1", "-j"], capture_output=True, text=True,
                         check=True, timeout=120).stdout
    assert "Number of triangles: %d" % r["triangles"] in out


@pytest.mark.parametrize("n,p", [(100000, 2), (100000, 8), (65536, 4), (4096, 8)])
def test_tric_multi_rank_generator(golden, n, p):
    """`tric -n N -l -N P` builds the graph the reference generates on P MPI
    ranks and counts it on the GPU: edges and triangles equal the reference's."""
    import os
    import subprocess
    tric = os.path.join(os.path.dirname(xxtric_amd.__file__), "bin", "tric")
    g = [x for x in golden["rgg_ranks"] if x["n"] == n and x["nprocs"] == p][0]
    _say("tric -n %d -l -N %d" % (n, p))
    out = subprocess.run([tric, "-n", str(n), "-l", "-N", str(p)], capture_output=True, text=True,
                         check=True, timeout=120).stdout
    assert "Number of edges: %d" % g["ne"] in out
    assert "Number of triangles: %d" % g["triangles"] in out


def test_multi_rank_rgg_per_vertex(ctx, golden):
    """Per-vertex counts of a P-rank graph equal the oracle's (min-vertex credit)."""
    g = [x for x in golden["rgg_ranks"] if x["n"] == 65536 and x["nprocs"] == 2][0]
    off, tails = xxtric_amd.rgg_generate(65536, nprocs=2)
    exp, exp_pv = O.count_reference(off, tails, per_vertex=True)
    assert exp == g["triangles"]
    _say("rgg ranks 65536/2 per vertex")
    tot, pv = ctx.count_csr(off, tails, per_vertex=True)
    assert tot == exp and np.array_equal(pv, exp_pv)


def test_resident_graph_and_alg_bytes(ctx):
    off, tails = O.rmat_csr(12, 16, 1, True)
    g = ctx.graph(off, tails, validate=True, symmetry=True)
    exp = O.count_degree_ordered(off, tails)
    _say("resident")
    assert g.count() == exp and g.count(orient="id") == exp
    b, sm, nep = g.alg_bytes()
    assert nep == len(tails) // 2
    # independent sum of min(d+(u), d+(v)) under (degree, id) orientation
    n = len(off) - 1
    deg = np.diff(off)
    rows = np.repeat(np.arange(n), deg)
    key_u = deg[rows] * (1 << 32) + rows
    key_v = deg[tails] * (1 << 32) + tails
    keep = key_v > key_u
    dplus = np.bincount(rows[keep], minlength=n)
    s = np.minimum(dplus[rows[keep]], dplus[tails[keep]]).sum()
    assert sm == s and b == 8 * (n + 1) + 4 * nep + 4 * s
    g.close()


@pytest.mark.parametrize("small_cut", [1, 32, 1 << 30])
def test_wide_layout_gives_the_same_counts(ctx, small_cut):
    """The layout of graphs with 2^32 or more entries (64-bit list offsets, saturated
    16-bit out-degrees), forced on graphs small enough for the oracle: global,
    per-vertex, owned-range and batch counts are unchanged."""
    off, tails = O.rmat_csr(13, 16, 2, True)
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    exp, exp_pv = O.count_reference(off, tails, per_vertex=True)
    hub_o, hub_t = wheel(70000)  # one vertex of degree 69999: its saturated key ties and breaks by id
    hub_exp = O.count_degree_ordered(hub_o, hub_t)
    hub_exp = hub_exp[0] if isinstance(hub_exp, tuple) else hub_exp
    ctx.tune(small_cut=small_cut, wave_owner=-2)  # the group kernel's wide layout (wave tiles: test_gpu_waves)
    ctx.force_wide(True)
    try:
        _say("wide layout, cut", small_cut)
        for orient in ("degree", "id"):
            tot, pv = ctx.count_csr(off, tails, per_vertex=True, orient=orient)
            assert tot == exp and (pv == exp_pv).all(), orient
            assert ctx.count_csr(off, tails, orient=orient) == exp
        g = ctx.graph(off, tails)
        owned, _, tot = g.count_owned(0, n // 3)
        assert tot == exp and owned == int(exp_pv[: n // 3].sum())
        g.close()
        assert ctx.count_csr(hub_o, hub_t) == hub_exp
        b = xxtric_amd.Batch(ctx, 2_000_000, 40_000_000, 64)
        gs = [xxtric_amd.batch_graph_csr(2026, i) for i in range(40)]
        for o, t in gs:
            b.add(o, t)
        b.run()
        assert list(b.results()) == [O.count_degree_ordered(o, t) for o, t in gs]
        b.close()
    finally:
        ctx.force_wide(False)
        ctx.tune(small_cut=xxtric_amd.DEFAULT_LIGHT_CUT, group_blocks=4, group_round=2, wave_owner=-1)


def test_rgg_connectivity_radius_strict_1e5(ctx):
    """The workload's radius variant: 100,000 points of the reference's LCG stream,
    edge iff distance < sqrt(2 ln n / (pi n)).  The cell-grid CSR equals the
    oracle's all-pairs loop with the strict test; global and per-vertex counts
    equal the oracle's in both orientations; `tric -n 100000 -l -C` prints the
    same edge and triangle counts."""
    import os
    import subprocess
    n = 100000
    r = xxtric_amd.rgg_radius_connect(n)
    off, tails = xxtric_amd.rgg_generate(n, seed=1, strict=True, connect=True)
    oo, ot = O.rgg_lcg(n, seed=1, radius=r, strict=True)
    assert (off == oo).all() and (tails == ot).all()
    exp = _check_csr(ctx, off, tails, "rgg1e5 connect strict")
    assert exp > 0
    tric = os.path.join(os.path.dirname(xxtric_amd.__file__), "bin", "tric")
    _say("tric -n 100000 -l -C")
    out = subprocess.run([tric, "-n", str(n), "-l", "-C"], capture_output=True, text=True,
                         check=True, timeout=120).stdout
    assert "Number of edges: %d" % len(tails) in out
    assert "Number of triangles: %d" % exp in out
