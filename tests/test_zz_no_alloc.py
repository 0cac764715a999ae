"""The allocation-free whole-graph count, xt_single_gpu_count_noalloc, against
xt_single_gpu_count (same device int64 CSR, same result, added to the caller's
counter) and the oracle, over calls in a row that reuse its one scratch array:
a large graph, then a smaller one, then others; plus the size limit."""
import ctypes

import numpy as np
import pytest

import oracle_py as O
from helpers import complete, random_graph
from xxtric_amd import XtError, capi

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[no_alloc] " + msg, flush=True)


def _both(ctx, off, tails, stream=None):
    """(allocation-free count, existing entry point's count) of one device CSR."""
    d_ind = ctx.to_device(np.ascontiguousarray(off, np.int64))
    d_el = ctx.to_device(np.ascontiguousarray(tails, np.int64))
    d_a = ctx.to_device(np.zeros(1, np.int64))
    d_b = ctx.to_device(np.zeros(1, np.int64))
    try:
        nv = len(off) - 1
        ctx.single_gpu_count_noalloc(nv, d_ind.ptr, d_el.ptr, d_a.ptr, stream)
        ctx.single_gpu_count(nv, d_ind.ptr, d_el.ptr, d_b.ptr, stream)
        ctx.sync()
        return int(d_a.download(np.int64, 1)[0]), int(d_b.download(np.int64, 1)[0])
    finally:
        for b in (d_ind, d_el, d_a, d_b):
            b.close()


def test_noalloc_equals_single_gpu_count(ctx):
    big = O.rmat_csr(16, 16, 1, True)          # 65,536 vertices, 1.8 M entries (~41 MB of scratch)
    small = O.rmat_csr(12, 16, 7, True)        # 4,096 vertices
    exp_big = O.count_degree_ordered(*big)
    exp_small = O.count_reference(*small)
    _say("rmat s16 then rmat s12")
    assert _both(ctx, *big) == (exp_big, exp_big)
    assert _both(ctx, *small) == (exp_small, exp_small)
    # more calls over the same array: sizes up and down, other shapes
    for name, g in (("k70", complete(70)), ("rand300", random_graph(300, 0.1, 11)), ("rmat s16 again", big),
                    ("rand40", random_graph(40, 0.5, 3))):
        _say(name)
        exp = O.count_reference(*g)
        assert _both(ctx, *g) == (exp, exp), name


def test_noalloc_adds_and_takes_a_stream(ctx):
    """The count is added to the caller's counter, on the caller's stream (a value)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.c_void_p, ctypes.c_uint]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    st = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(st), 1) == 0
    off, tails = O.rmat_csr(11, 16, 2, True)
    exp = O.count_reference(off, tails)
    d_ind = ctx.to_device(np.ascontiguousarray(off, np.int64))
    d_el = ctx.to_device(np.ascontiguousarray(tails, np.int64))
    d_cnt = ctx.to_device(np.array([5], np.int64))
    _say("rmat s11 twice on a caller stream")
    ctx.single_gpu_count_noalloc(len(off) - 1, d_ind.ptr, d_el.ptr, d_cnt.ptr, st.value)
    ctx.single_gpu_count_noalloc(len(off) - 1, d_ind.ptr, d_el.ptr, d_cnt.ptr, st.value)
    assert hip.hipStreamSynchronize(st) == 0
    assert int(d_cnt.download(np.int64, 1)[0]) == 5 + 2 * exp
    assert hip.hipStreamDestroy(st) == 0
    for b in (d_ind, d_el, d_cnt):
        b.close()


def test_noalloc_rejects_what_does_not_fit(ctx):
    """A graph whose scratch exceeds the static array fails with XT_ERR_CAPACITY
    before any kernel runs (the counter stays untouched); bad input is still
    rejected; the next call that fits works."""
    nv = 1 << 21                                   # 96 MiB of per-vertex scratch alone (48 B each)
    d_ind = ctx.to_device(np.zeros(nv + 1, np.int64))
    d_el = ctx.to_device(np.zeros(1, np.int64))
    d_cnt = ctx.to_device(np.array([7], np.int64))
    _say("nv = 2^21, no entries: over the limit")
    with pytest.raises(XtError) as e:
        ctx.single_gpu_count_noalloc(nv, d_ind.ptr, d_el.ptr, d_cnt.ptr)
    assert e.value.code == capi.XT_ERR_CAPACITY
    assert int(d_cnt.download(np.int64, 1)[0]) == 7
    # 2^20 vertices without entries fit (no group tiles), and count nothing
    _say("nv = 2^20, no entries: fits")
    ctx.single_gpu_count_noalloc(nv // 2, d_ind.ptr, d_el.ptr, d_cnt.ptr)
    ctx.sync()
    assert int(d_cnt.download(np.int64, 1)[0]) == 7
    for b in (d_ind, d_el):
        b.close()
    # an id out of range
    off = np.array([0, 1, 2], np.int64)
    tails = np.array([1, 5], np.int64)
    d_ind = ctx.to_device(off)
    d_el = ctx.to_device(tails)
    _say("id out of range")
    with pytest.raises(XtError) as e:
        ctx.single_gpu_count_noalloc(2, d_ind.ptr, d_el.ptr, d_cnt.ptr)
    assert e.value.code == capi.XT_ERR_INPUT
    for b in (d_ind, d_el):
        b.close()
    _say("k5 after the failures")
    off, tails = complete(5)
    d_ind = ctx.to_device(off)
    d_el = ctx.to_device(tails)
    d_cnt.upload(np.zeros(1, np.int64))
    ctx.single_gpu_count_noalloc(5, d_ind.ptr, d_el.ptr, d_cnt.ptr)
    ctx.sync()
    assert int(d_cnt.download(np.int64, 1)[0]) == 10
    for b in (d_ind, d_el, d_cnt):
        b.close()
