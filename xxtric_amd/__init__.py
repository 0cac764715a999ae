"""xxtric_amd -- exact triangle counting on MI355X (gfx950).

Python face of libxtric.so (include/xtric.h).  The counting entry points
mirror the reference's single-device engine (xxtric-gpu.hpp:22-298,
TriangleCountingSingle: construct from a CSR, count(), clear()) and its lower
FFI (counting.cuh:229-323); the ingest helpers mirror graph.hpp (vite files,
the seeded random-geometric generator).  All counting happens in the HIP
kernels of the library; this module only moves arrays and pointers.
"""
import ctypes
import weakref

import numpy as np

from . import capi
from .capi import (XT_BATCH_HIPGRAPH, XT_BATCH_PER_GRAPH, XT_BATCH_PERSISTENT,  # noqa: F401
                   XT_CHECK_SYMMETRY, XT_ORIENT_DEGREE, XT_ORIENT_ID, XT_PER_VERTEX,
                   XT_VALIDATE, XtError)

DEFAULT_LIGHT_CUT = -1  # tune() value that restores the defaults: light = both out-lists <= 32 (8 in batches of 4M+ entries, else 64)

__all__ = ["Context", "DeviceGraph", "TriangleCountingSingle", "rgg_generate", "rgg_radius", "rgg_radius_connect",
           "vite_read", "vite_write", "edgelist_to_csr", "batch_params", "batch_graph_csr",
           "Batch", "Server", "XtError"]


def _flags(orient="degree", per_vertex=False, validate=False, symmetry=False, clean=False):
    f = XT_ORIENT_ID if orient == "id" else XT_ORIENT_DEGREE
    if orient not in ("id", "degree"):
        raise ValueError("orient must be 'degree' or 'id'")
    if per_vertex:
        f |= XT_PER_VERTEX
    if validate:
        f |= XT_VALIDATE
    if symmetry:
        f |= XT_VALIDATE | XT_CHECK_SYMMETRY
    if clean:
        f |= capi.XT_CLEAN
    return f


class Context:
    """A device context (xt_ctx): scratch memory, one HIP stream, events."""

    def __init__(self, device=0):
        self._L = capi.lib()
        h = ctypes.c_void_p()
        rc = self._L.xt_ctx_create(device, ctypes.byref(h))
        if rc != capi.XT_OK:
            raise XtError(rc, "xt_ctx_create(%d) failed (no HIP device?)" % device)
        self.h = h
        self.device = device
        # library objects made on this context (graphs, buffers, batches,
        # servers): closed before the context itself, since each uses it
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)
        return obj

    def _check(self, rc):
        if rc != capi.XT_OK:
            raise XtError(rc, self._L.xt_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            for obj in list(self._children):
                try:
                    obj.close()
                except Exception:
                    pass
            self._L.xt_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, small_cut=0, group_blocks=0, group_round=0, items_per_tile=0, tile_run=0, wave_owner=0,
             group_hash=0):
        """Set the light-edge cut, the group kernel's workgroups per CU, its chunks in
        flight per thread, the items per group tile, the tiles a workgroup draws at
        once, the longest owner list counted in wave tiles and/or the group
        kernel's table hash (1 multiplicative, 2 low bits) (0 = keep; -1 restores
        the default of small_cut, items_per_tile, tile_run, wave_owner, group_hash;
        wave_owner -2 turns wave tiles off).  No setting changes a result."""
        self._check(self._L.xt_ctx_tune(self.h, small_cut, group_blocks, group_round))
        if items_per_tile or tile_run:
            self._check(self._L.xt_ctx_tune_tiles(self.h, items_per_tile, tile_run))
        if wave_owner:
            self._check(self._L.xt_ctx_tune_waves(self.h, wave_owner))
        if group_hash:
            self._check(self._L.xt_ctx_tune_hash(self.h, group_hash))

    def fail_allocs(self, n):
        """Failure injection: the next n scratch allocations fail with XT_ERR_NOMEM."""
        self._check(self._L.xt_ctx_fail_allocs(self.h, int(n)))

    def force_wide(self, on=True):
        """Count with the layout of graphs of 2^32 or more entries (tests, comparisons)."""
        self._check(self._L.xt_ctx_force_wide(self.h, 1 if on else 0))

    def timing(self):
        t = capi.XtTiming()
        self._check(self._L.xt_last_timing(self.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def count_csr(self, off, tails, per_vertex=False, orient="degree", validate=False,
                  symmetry=False):
        """Count a host CSR (GraphElem = int64 arrays); returns total or (total, pv)."""
        off, tails = capi.csr_arrays(off, tails, np.int64)
        nv = len(off) - 1
        total = ctypes.c_uint64()
        pv = np.zeros(nv, np.uint64) if per_vertex else None
        f = _flags(orient, per_vertex, validate, symmetry)
        self._check(self._L.xt_count_csr(self.h, nv, len(tails), capi.ptr(off), capi.ptr(tails), f,
                                         ctypes.byref(total), capi.ptr(pv)))
        return (total.value, pv) if per_vertex else total.value

    def graph(self, off, tails, validate=False, symmetry=False, clean=False):
        """Upload a CSR; clean=True rebuilds it on the device as a simple undirected
        graph (rows sorted, duplicates merged, self-loops dropped, reverse entries added)."""
        off = np.ascontiguousarray(off, np.int64)
        tb = 4 if np.asarray(tails).dtype == np.int32 else 8
        tails = np.ascontiguousarray(tails, np.int32 if tb == 4 else np.int64)
        h = ctypes.c_void_p()
        self._check(self._L.xt_graph_from_csr(self.h, len(off) - 1, len(tails), capi.ptr(off),
                                              capi.ptr(tails), tb,
                                              _flags(validate=validate, symmetry=symmetry, clean=clean),
                                              ctypes.byref(h)))
        return DeviceGraph(self, h)

    def graph_from_edges(self, nv, src, dst):
        """Symmetric CSR built on the device from an edge list (any order or direction,
        duplicates merged, self-loops dropped)."""
        src = np.ascontiguousarray(src, np.int64)
        dst = np.ascontiguousarray(dst, np.int64)
        if len(src) != len(dst):
            raise ValueError("src and dst differ in length")
        h = ctypes.c_void_p()
        self._check(self._L.xt_graph_from_edges(self.h, int(nv), len(src), capi.ptr(src), capi.ptr(dst),
                                                ctypes.byref(h)))
        return DeviceGraph(self, h)

    def graph_from_vite(self, path, layout32=False, validate=False, symmetry=False, clean=False):
        """A vite binary file streamed into a resident graph (bounded host memory);
        clean=True as for graph()."""
        h = ctypes.c_void_p()
        self._check(self._L.xt_graph_from_vite(self.h, str(path).encode(), 1 if layout32 else 0,
                                               _flags(validate=validate, symmetry=symmetry, clean=clean),
                                               ctypes.byref(h)))
        return DeviceGraph(self, h)

    def count_edges(self, nv, src, dst, per_vertex=False, orient="degree"):
        src = np.ascontiguousarray(src, np.int64)
        dst = np.ascontiguousarray(dst, np.int64)
        total = ctypes.c_uint64()
        pv = np.zeros(int(nv), np.uint64) if per_vertex else None
        self._check(self._L.xt_count_edges(self.h, int(nv), len(src), capi.ptr(src), capi.ptr(dst),
                                           _flags(orient, per_vertex), ctypes.byref(total), capi.ptr(pv)))
        return (total.value, pv) if per_vertex else total.value

    def rmat(self, scale, edgefactor=16, seed=1, permute=True, abc=(0.57, 0.19, 0.19)):
        h = ctypes.c_void_p()
        self._check(self._L.xt_graph_rmat(self.h, scale, edgefactor, abc[0], abc[1], abc[2], seed,
                                          int(permute), ctypes.byref(h)))
        return DeviceGraph(self, h)

    def er(self, n, seed, thr64):
        h = ctypes.c_void_p()
        self._check(self._L.xt_graph_er(self.h, n, seed, thr64, ctypes.byref(h)))
        return DeviceGraph(self, h)

    def to_device(self, arr):
        """Copy a numpy array into a new DeviceBuffer."""
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, arr.nbytes)
        buf.upload(arr)
        return buf

    def sync(self):
        self._check(self._L.xt_dev_sync(self.h))

    # ---- reference lower-FFI shapes: raw device pointers (ints) ----
    def single_gpu_count(self, nv, d_indices, d_edge_list, d_local_count, stream=None):
        self._check(self._L.xt_single_gpu_count(self.h, nv, d_indices, d_edge_list, d_local_count,
                                                stream))

    def single_gpu_count_noalloc(self, nv, d_indices, d_edge_list, d_local_count, stream=None):
        """xt_single_gpu_count without any device allocation in the call (64 MiB static scratch)."""
        self._check(self._L.xt_single_gpu_count_noalloc(self.h, nv, d_indices, d_edge_list,
                                                        d_local_count, stream))

    def single_gpu_count_masked(self, nedges, d_edge_mask, d_indices, d_edge_list, d_local_count,
                                stream=None):
        self._check(self._L.xt_single_gpu_count_masked(self.h, nedges, d_edge_mask, d_indices,
                                                       d_edge_list, d_local_count, stream))

    def local_count_gpu_inner(self, start, end, jstart, jend, d_edge_list, d_local_count,
                              stream=None):
        self._check(self._L.xt_local_count_gpu_inner(self.h, start, end, jstart, jend, d_edge_list,
                                                     d_local_count, stream))


class DeviceBuffer:
    """Device memory owned through the library (xt_dev_alloc / xt_memcpy / xt_dev_free)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx._check(ctx._L.xt_dev_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value
        ctx._adopt(self)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx._check(self.ctx._L.xt_memcpy(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes,
                                              capi.XT_H2D))

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        self.ctx._check(self.ctx._L.xt_memcpy(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes,
                                              capi.XT_D2H))
        return out

    def close(self):
        if self.ptr:
            self.ctx._L.xt_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """A symmetric CSR resident in HBM (int64 offsets, int32 ids)."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        nv = ctypes.c_int64()
        ne = ctypes.c_int64()
        ctx._L.xt_graph_info(h, ctypes.byref(nv), ctypes.byref(ne))
        self.nv = nv.value
        self.ne = ne.value
        ctx._adopt(self)

    def count(self, per_vertex=False, orient="degree"):
        total = ctypes.c_uint64()
        pv = np.zeros(self.nv, np.uint64) if per_vertex else None
        self.ctx._check(self.ctx._L.xt_graph_count(self.ctx.h, self.h, _flags(orient, per_vertex),
                                                   ctypes.byref(total), capi.ptr(pv)))
        return (total.value, pv) if per_vertex else total.value

    def count_owned(self, lo, hi, orient="degree", total=True):
        """Triangles whose minimum vertex lies in [lo, hi), and of those the ones
        whose middle vertex does too (one rank of a 1-D vertex split); also the
        graph's total (a second count; None with total=False)."""
        owned = ctypes.c_uint64()
        local = ctypes.c_uint64()
        tot = ctypes.c_uint64()
        self.ctx._check(self.ctx._L.xt_graph_count_owned(self.ctx.h, self.h, int(lo), int(hi), _flags(orient),
                                                         ctypes.byref(owned), ctypes.byref(local),
                                                         ctypes.byref(tot) if total else None))
        return owned.value, local.value, (tot.value if total else None)

    def slice(self, lo, hi):
        """Rows of [lo, hi) and of their neighbours (ghosts) kept, others emptied."""
        h = ctypes.c_void_p()
        self.ctx._check(self.ctx._L.xt_graph_slice(self.ctx.h, self.h, int(lo), int(hi), ctypes.byref(h)))
        return DeviceGraph(self.ctx, h)

    def download(self):
        off = np.empty(self.nv + 1, np.int64)
        tails = np.empty(self.ne, np.int32)
        self.ctx._check(self.ctx._L.xt_graph_download(self.ctx.h, self.h, capi.ptr(off),
                                                      capi.ptr(tails)))
        return off, tails

    def alg_bytes(self, orient="degree"):
        b = ctypes.c_uint64()
        sm = ctypes.c_uint64()
        nep = ctypes.c_int64()
        self.ctx._check(self.ctx._L.xt_graph_alg_bytes(self.ctx.h, self.h, _flags(orient),
                                                       ctypes.byref(b), ctypes.byref(sm),
                                                       ctypes.byref(nep)))
        return b.value, sm.value, nep.value

    def stats(self, orient="degree"):
        """Algorithmic byte figures (whole count and the group kernel's share)."""
        st = capi.XtGraphStats()
        self.ctx._check(self.ctx._L.xt_graph_stats_get(self.ctx.h, self.h, _flags(orient),
                                                       ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.h:
            self.ctx._L.xt_graph_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TriangleCountingSingle:
    """Mirror of the reference's single-device engine (xxtric-gpu.hpp:22-298).

    The constructor takes a CSR (offsets, tails) as Graph::edge_indices_ and the
    tails of Graph::edge_list_ (graph.hpp:267-269) and uploads it; count()
    returns the global triangle count; clear() releases device memory.
    """

    def __init__(self, off, tails, ctx=None, orient="degree"):
        self.ctx = ctx or Context(0)
        self.orient = orient
        self.g = self.ctx.graph(off, np.asarray(tails))

    def count(self):
        return self.g.count(orient=self.orient)

    def per_vertex(self):
        return self.g.count(per_vertex=True, orient=self.orient)[1]

    def clear(self):
        self.g.close()


# ---- host ingest --------------------------------------------------------------
def rgg_radius(n):
    return capi.lib().xt_rgg_radius(n)


def rgg_radius_connect(n):
    """sqrt(2 ln n / (pi n)): the connectivity radius."""
    return capi.lib().xt_rgg_radius_connect(n)


def rgg_generate(n, seed=1, radius=0.0, nprocs=1, strict=False, connect=False):
    """Reference RGG (graph.hpp:745-1374, LCG mode): (off int64[n+1], tails int64).
    nprocs > 1: the graph a run on that many MPI ranks generates, all rows in one CSR.
    strict: edge iff distance < r (the reference tests <=); connect: r = sqrt(2 ln n /
    (pi n)) when radius <= 0 (the reference's radius is graph.hpp:789-791)."""
    L = capi.lib()
    flags = (capi.XT_RGG_STRICT if strict else 0) | (capi.XT_RGG_CONNECT if connect else 0)
    off = np.empty(n + 1, np.int64)
    ne = ctypes.c_int64()
    rc = L.xt_rgg_generate_opts(n, nprocs, seed, radius, flags, capi.ptr(off), None, 0, ctypes.byref(ne))
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_rgg_generate_opts")
    tails = np.empty(ne.value, np.int64)
    rc = L.xt_rgg_generate_opts(n, nprocs, seed, radius, flags, capi.ptr(off), capi.ptr(tails), ne.value,
                                ctypes.byref(ne))
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_rgg_generate_opts")
    return off, tails


def vite_read(path, layout32=False):
    L = capi.lib()
    nv = ctypes.c_int64()
    ne = ctypes.c_int64()
    rc = L.xt_vite_read(path.encode(), int(layout32), ctypes.byref(nv), ctypes.byref(ne), None, None)
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_vite_read(%s)" % path)
    off = np.empty(nv.value + 1, np.int64)
    tails = np.empty(ne.value, np.int64)
    rc = L.xt_vite_read(path.encode(), int(layout32), ctypes.byref(nv), ctypes.byref(ne),
                        capi.ptr(off), capi.ptr(tails))
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_vite_read(%s)" % path)
    return off, tails


def vite_write(path, off, tails, layout32=False):
    off, tails = capi.csr_arrays(off, tails, np.int64)
    rc = capi.lib().xt_vite_write(path.encode(), int(layout32), len(off) - 1, capi.ptr(off),
                                  capi.ptr(tails))
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_vite_write(%s)" % path)


def edgelist_to_csr(path):
    L = capi.lib()
    nv = ctypes.c_int64()
    ne = ctypes.c_int64()
    rc = L.xt_edgelist_to_csr(path.encode(), ctypes.byref(nv), ctypes.byref(ne), None, None)
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_edgelist_to_csr(%s)" % path)
    off = np.empty(nv.value + 1, np.int64)
    tails = np.empty(ne.value, np.int64)
    rc = L.xt_edgelist_to_csr(path.encode(), ctypes.byref(nv), ctypes.byref(ne), capi.ptr(off),
                              capi.ptr(tails))
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_edgelist_to_csr(%s)" % path)
    return off, tails


PARTITION_METHODS = {"greedy": 0, "linear": 1}


def partition_vertices(off, nparts, method="greedy"):
    """Owned vertex ranges of a split over nparts ranks, as the reference's
    balanced read computes them (graph.hpp:397-640): "greedy" is flag -b,
    "linear" flag -d.  Returns bins int64[nparts + 1]; rank r owns
    [bins[r], bins[r + 1])."""
    off = np.ascontiguousarray(off, np.int64)
    bins = np.zeros(nparts + 1, np.int64)
    rc = capi.lib().xt_partition_vertices(len(off) - 1, capi.ptr(off), nparts,
                                          PARTITION_METHODS[method], capi.ptr(bins))
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_partition_vertices(%s, %d parts)" % (method, nparts))
    return bins


def batch_params(base_seed, g, nmin=200, nmax=20000):
    n = ctypes.c_int64()
    kind = ctypes.c_int()
    gseed = ctypes.c_uint64()
    thr = ctypes.c_uint64()
    rc = capi.lib().xt_batch_params(base_seed, g, nmin, nmax, ctypes.byref(n), ctypes.byref(kind),
                                    ctypes.byref(gseed), ctypes.byref(thr))
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_batch_params")
    return n.value, kind.value, gseed.value, thr.value


def batch_graph_csr(base_seed, g, nmin=200, nmax=20000):
    L = capi.lib()
    nv = ctypes.c_int64()
    ne = ctypes.c_int64()
    rc = L.xt_batch_graph_csr(base_seed, g, nmin, nmax, ctypes.byref(nv), ctypes.byref(ne), None, None)
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_batch_graph_csr")
    off = np.empty(nv.value + 1, np.int64)
    tails = np.empty(ne.value, np.int32)
    rc = L.xt_batch_graph_csr(base_seed, g, nmin, nmax, ctypes.byref(nv), ctypes.byref(ne),
                              capi.ptr(off), capi.ptr(tails))
    if rc != capi.XT_OK:
        raise XtError(rc, "xt_batch_graph_csr")
    return off, tails


from .batching import Batch, Server  # noqa: E402,F401
