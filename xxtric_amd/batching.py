"""Batched counting of many small graphs and the continuous-batching server.

Thin wrappers over xt_batch_* and xt_serve_* of libxtric.so.
"""
import ctypes

import numpy as np

from . import capi
from .capi import XtError


class Batch:
    """Many graphs counted by one persistent device pipeline (xt_batch)."""

    def __init__(self, ctx, max_vertices, max_edges, max_graphs):
        self.ctx = ctx
        self._L = ctx._L
        h = ctypes.c_void_p()
        ctx._check(self._L.xt_batch_create(ctx.h, max_vertices, max_edges, max_graphs,
                                           ctypes.byref(h)))
        self.h = h
        self.n = 0
        ctx._adopt(self)

    def add(self, off, tails):
        off = np.ascontiguousarray(off, np.int64)
        tails = np.ascontiguousarray(tails, np.int32)
        slot = ctypes.c_int64()
        self.ctx._check(self._L.xt_batch_add(self.h, len(off) - 1, len(tails), capi.ptr(off),
                                             capi.ptr(tails), ctypes.byref(slot)))
        self.n += 1
        return slot.value

    def add_generated(self, base_seed, first, count, nmin=200, nmax=20000):
        self.ctx._check(self._L.xt_batch_add_generated(self.h, base_seed, first, count, nmin, nmax))
        self.n += count

    def run(self, mode=capi.XT_BATCH_PERSISTENT, orient="degree"):
        f = capi.XT_ORIENT_ID if orient == "id" else capi.XT_ORIENT_DEGREE
        self.ctx._check(self._L.xt_batch_run(self.h, mode, f))

    def results(self):
        out = np.zeros(self.n, np.uint64)
        self.ctx._check(self._L.xt_batch_results(self.h, capi.ptr(out), self.n))
        return out

    def timing(self):
        t = capi.XtBatchTiming()
        self.ctx._check(self._L.xt_batch_timing_get(self.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def stats(self, orient="degree"):
        """Algorithmic byte figures of the whole batch (sums over its graphs)."""
        st = capi.XtGraphStats()
        flags = capi.XT_ORIENT_ID if orient == "id" else capi.XT_ORIENT_DEGREE
        self.ctx._check(self._L.xt_batch_stats_get(self.h, flags, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def clear(self):
        self.ctx._check(self._L.xt_batch_clear(self.h))
        self.n = 0

    def close(self):
        if self.h and self.ctx.h:  # the context closes its batches before itself
            self._L.xt_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (xt_serve)."""

    def __init__(self, ctx, hbm_budget_bytes=0, max_batch_graphs=0, max_inflight=0, max_queue=0,
                 max_wait_us=0.0):
        self.ctx = ctx
        self._L = ctx._L
        cfg = capi.XtServeConfig(hbm_budget_bytes, max_batch_graphs, max_inflight, max_queue,
                                 max_wait_us)
        h = ctypes.c_void_p()
        ctx._check(self._L.xt_serve_create(ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._borrowed = {}  # ticket -> arrays the server reads in place
        ctx._adopt(self)

    def submit(self, off, tails, arrival_us=-1.0, borrow=False):
        """Queue a graph.  borrow=True skips the copy: the server reads the arrays
        in place, and they are kept referenced here until the ticket is polled."""
        off = np.ascontiguousarray(off, np.int64)
        tails = np.ascontiguousarray(tails, np.int32)
        t = ctypes.c_uint64()
        fn = self._L.xt_serve_submit_borrowed if borrow else self._L.xt_serve_submit
        rc = fn(self.h, len(off) - 1, len(tails), capi.ptr(off), capi.ptr(tails), arrival_us, ctypes.byref(t))
        if rc != capi.XT_OK:
            raise XtError(rc, self._L.xt_last_error(self.ctx.h).decode())
        if borrow:
            self._borrowed[t.value] = (off, tails)
        return t.value

    def poll(self, ticket):
        total = ctypes.c_uint64()
        lat = ctypes.c_double()
        rc = self._L.xt_serve_poll(self.h, ticket, ctypes.byref(total), ctypes.byref(lat))
        if rc == capi.XT_PENDING:
            return None
        # A borrowed request's arrays are let go only once its ticket is retired
        # (its result or its own failure reported); after a server-wide error the
        # request may still be read.
        if rc == capi.XT_OK or self._L.xt_serve_ticket_state(self.h, ticket) == capi.XT_ERR_NOTFOUND:
            self._borrowed.pop(ticket, None)
        if rc != capi.XT_OK:
            raise XtError(rc, "xt_serve_poll")
        return total.value, lat.value

    def ticket_state(self, ticket):
        """0 queued, 1 running, 2 done (not yet polled); XT_ERR_NOTFOUND when retired."""
        return self._L.xt_serve_ticket_state(self.h, ticket)

    def pause(self, paused=True):
        """Hold admission (queued requests wait; running batches finish)."""
        rc = self._L.xt_serve_pause(self.h, 1 if paused else 0)
        if rc != capi.XT_OK:
            raise XtError(rc, "xt_serve_pause")

    def replay(self, pool, expected, order, arrival_us, timeout_s=300.0):
        """Native load generator (xt_serve_replay): request i is pool graph order[i],
        submitted at arrival_us[i] on the server clock.  pool: [(off int64, tails
        int32)], expected: per-graph counts (or None).  Returns (mismatches,
        latencies in us, -1 for requests rejected at submit)."""
        pool = [(np.ascontiguousarray(o, np.int64), np.ascontiguousarray(t, np.int32)) for o, t in pool]
        n = len(pool)
        nv = np.array([len(o) - 1 for o, _ in pool], np.int64)
        ne = np.array([len(t) for _, t in pool], np.int64)
        offs = (ctypes.c_void_p * n)(*[o.ctypes.data for o, _ in pool])
        tls = (ctypes.c_void_p * n)(*[t.ctypes.data if len(t) else 0 for _, t in pool])
        exp = None if expected is None else np.ascontiguousarray(expected, np.uint64)
        order = np.ascontiguousarray(order, np.int64)
        arrival_us = np.ascontiguousarray(arrival_us, np.float64)
        lat = np.empty(len(order), np.float64)
        bad = ctypes.c_int64()
        rc = self._L.xt_serve_replay(self.h, n, capi.ptr(nv), capi.ptr(ne), ctypes.addressof(offs),
                                     ctypes.addressof(tls), capi.ptr(exp), len(order), capi.ptr(order),
                                     capi.ptr(arrival_us), timeout_s, ctypes.byref(bad), capi.ptr(lat))
        if rc != capi.XT_OK:
            raise XtError(rc, "xt_serve_replay")
        return bad.value, lat

    def drain(self, timeout_s=60.0):
        rc = self._L.xt_serve_drain(self.h, timeout_s)
        if rc != capi.XT_OK:
            raise XtError(rc, "xt_serve_drain")

    def now_us(self):
        return self._L.xt_serve_now_us(self.h)

    def stats(self):
        s = capi.XtServeStats()
        rc = self._L.xt_serve_stats_get(self.h, ctypes.byref(s))
        if rc != capi.XT_OK:
            raise XtError(rc, "xt_serve_stats_get")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if self.h and self.ctx.h:  # the context closes its servers before itself
            self._L.xt_serve_destroy(self.h)
        self.h = None
        self._borrowed.clear()
        self._borrowed.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
