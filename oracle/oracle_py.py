"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time the reference algorithm; nothing in
xxtric_amd/ imports it.  Each function mirrors one in oracle.h, which cites
the reference file:line it restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_BIN = os.path.join(_HERE, "_ref", "tric_ref")

_lib = None

_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_dbl = ctypes.c_double
_p = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "or_reseeder": (_u32, [_u32]),
            "or_rgg_radius_ref": (_dbl, [_i64]),
            "or_rgg_lcg_coords": (None, [_i64, _u32, _p, _p]),
            "or_rgg_lcg": (_i64, [_i64, _u32, _dbl, _p, _p, _i64, _int]),
            "or_rgg_lcg_opts": (_i64, [_i64, _u32, _dbl, _int, _p, _p, _i64, _int]),
            "or_rgg_lcg_coords_ranks": (None, [_i64, _int, _u32, _p, _p]),
            "or_rgg_lcg_ranks": (_i64, [_i64, _int, _u32, _dbl, _p, _p, _i64, _int]),
            "or_count_reference": (_u64, [_i64, _p, _p, _p, _int, _i64, _i64]),
            "or_count_reference32": (_u64, [_i64, _p, _p, _p, _int, _i64, _i64]),
            "or_count_degree_ordered": (_u64, [_i64, _p, _p, _p, _int]),
            "or_count_reference32_budget": (_i64, [_i64, _p, _p, _int, _dbl, _p, _p]),
            "or_count_reference32_list": (None, [_i64, _p, _p, _p, _i64, _p, _int]),
            "or_vertex_min_credit32": (None, [_i64, _p, _p, _p, _i64, _p, _int]),
            "or_vite_write": (_int, [ctypes.c_char_p, _i64, _p, _p, _int]),
            "or_vite_read": (_int, [ctypes.c_char_p, _int, _p, _p, _p, _p]),
            "or_mix64": (_u64, [_u64]),
            "or_rmat_thresholds": (None, [_dbl, _dbl, _dbl, _p]),
            "or_rmat_tuples": (None, [_int, _u64, _p, _int, _i64, _i64, _p, _p]),
            "or_rmat_csr": (_i64, [_int, _int, _u64, _p, _int, _p, _p, _i64]),
            "or_er_csr": (_i64, [_i64, _u64, _u64, _p, _p, _i64]),
            "or_batch_params": (None, [_u64, _i64, _i64, _i64, _p, _p, _p, _p]),
            "or_partition_greedy": (_int, [_i64, _p, _i64, _p]),
            "or_partition_linear": (_int, [_i64, _p, _i64, _p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def nthreads_default():
    return max(1, min(16, os.cpu_count() or 1))


def reseeder(seed=1):
    return lib().or_reseeder(seed)


def rgg_radius_ref(nv):
    return lib().or_rgg_radius_ref(nv)


def rgg_lcg_coords(n, seed=1):
    X = np.empty(n, np.float64)
    Y = np.empty(n, np.float64)
    lib().or_rgg_lcg_coords(n, seed, _ptr(X), _ptr(Y))
    return X, Y


def rgg_lcg(n, seed=1, radius=0.0, nthreads=None, strict=False):
    """(off int64[n+1], tails int64[ne]) of the O(n^2) reference RGG; strict
    tests distance < radius instead of the reference's <=."""
    nt = nthreads or nthreads_default()
    off = np.empty(n + 1, np.int64)
    ne = lib().or_rgg_lcg_opts(n, seed, radius, int(strict), _ptr(off), None, 0, nt)
    tails = np.empty(ne, np.int64)
    lib().or_rgg_lcg_opts(n, seed, radius, int(strict), _ptr(off), _ptr(tails), ne, nt)
    return off, tails


def rgg_lcg_coords_ranks(n, nprocs, seed=1):
    X = np.empty(n, np.float64)
    Y = np.empty(n, np.float64)
    lib().or_rgg_lcg_coords_ranks(n, nprocs, seed, _ptr(X), _ptr(Y))
    return X, Y


def rgg_lcg_ranks(n, nprocs, seed=1, radius=0.0, nthreads=None):
    """(off, tails) of the RGG a run on nprocs MPI ranks generates (all ranks' rows)."""
    nt = nthreads or nthreads_default()
    off = np.empty(n + 1, np.int64)
    ne = lib().or_rgg_lcg_ranks(n, nprocs, seed, radius, _ptr(off), None, 0, nt)
    if ne < 0:
        raise ValueError("nprocs must be a power of two dividing n, with radius below 1/nprocs")
    tails = np.empty(ne, np.int64)
    lib().or_rgg_lcg_ranks(n, nprocs, seed, radius, _ptr(off), _ptr(tails), ne, nt)
    return off, tails


def count_reference(off, tails, per_vertex=False, nthreads=None, vstart=0, vstride=1):
    """Reference CPU count (xxtric.hpp:60-208).  Returns total or (total, pv)."""
    nv = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    pv = np.zeros(nv, np.uint64) if per_vertex else None
    nt = nthreads or nthreads_default()
    if tails.dtype == np.int32:
        tails = np.ascontiguousarray(tails)
        tot = lib().or_count_reference32(nv, _ptr(off), _ptr(tails), _ptr(pv), nt, vstart, vstride)
    else:
        tails = np.ascontiguousarray(tails, np.int64)
        tot = lib().or_count_reference(nv, _ptr(off), _ptr(tails), _ptr(pv), nt, vstart, vstride)
    return (int(tot), pv) if per_vertex else int(tot)


def count_reference_list(off, tails, vertices, nthreads=None):
    """Reference per-vertex counts (smallest-id credit) of the listed vertices only."""
    off = np.ascontiguousarray(off, np.int64)
    tails = np.ascontiguousarray(tails, np.int32)
    vertices = np.ascontiguousarray(vertices, np.int64)
    pv = np.zeros(len(vertices), np.uint64)
    nt = nthreads or nthreads_default()
    lib().or_count_reference32_list(len(off) - 1, _ptr(off), _ptr(tails), _ptr(vertices),
                                    len(vertices), _ptr(pv), nt)
    return pv


def vertex_min_credit(off, tails, vertices, nthreads=None):
    """Per-vertex counts (smallest-id credit) of the listed vertices by marking
    N(v) and scanning N(u) above u for each neighbour u > v (rows ascending)."""
    off = np.ascontiguousarray(off, np.int64)
    tails = np.ascontiguousarray(tails, np.int32)
    vertices = np.ascontiguousarray(vertices, np.int64)
    out = np.zeros(len(vertices), np.uint64)
    lib().or_vertex_min_credit32(len(off) - 1, _ptr(off), _ptr(tails), _ptr(vertices), len(vertices),
                                 _ptr(out), nthreads or nthreads_default())
    return out


def count_reference_budget(off, tails, seconds, nthreads=None):
    """Reference count on vertices 0,1,2,... until `seconds` elapse.

    Returns (vertices_done, triangles_of_those, degree_sum_of_those)."""
    off = np.ascontiguousarray(off, np.int64)
    tails = np.ascontiguousarray(tails, np.int32)
    tri = ctypes.c_uint64()
    edges = ctypes.c_int64()
    nt = nthreads or nthreads_default()
    done = lib().or_count_reference32_budget(len(off) - 1, _ptr(off), _ptr(tails), nt, seconds,
                                            ctypes.addressof(tri), ctypes.addressof(edges))
    return int(done), int(tri.value), int(edges.value)


def count_degree_ordered(off, tails, per_vertex=False, nthreads=None):
    nv = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    tails = np.ascontiguousarray(tails, np.int32)
    pv = np.zeros(nv, np.uint64) if per_vertex else None
    nt = nthreads or nthreads_default()
    tot = lib().or_count_degree_ordered(nv, _ptr(off), _ptr(tails), _ptr(pv), nt)
    return (int(tot), pv) if per_vertex else int(tot)


def vite_write(path, off, tails, layout32=False):
    off = np.ascontiguousarray(off, np.int64)
    tails = np.ascontiguousarray(tails, np.int64)
    rc = lib().or_vite_write(path.encode(), len(off) - 1, _ptr(off), _ptr(tails), int(layout32))
    if rc != 0:
        raise IOError("vite write failed: %s" % path)


def vite_read(path, layout32=False):
    nv = ctypes.c_int64()
    ne = ctypes.c_int64()
    rc = lib().or_vite_read(path.encode(), int(layout32), ctypes.addressof(nv), ctypes.addressof(ne), None, None)
    if rc != 0:
        raise IOError("vite header read failed: %s" % path)
    off = np.empty(nv.value + 1, np.int64)
    tails = np.empty(ne.value, np.int64)
    rc = lib().or_vite_read(path.encode(), int(layout32), ctypes.addressof(nv), ctypes.addressof(ne), _ptr(off), _ptr(tails))
    if rc != 0:
        raise IOError("vite read failed: %s" % path)
    return off, tails


def mix64(z):
    return lib().or_mix64(z & 0xFFFFFFFFFFFFFFFF)


def rmat_thresholds(a=0.57, b=0.19, c=0.19):
    t = np.empty(3, np.uint32)
    lib().or_rmat_thresholds(a, b, c, _ptr(t))
    return t


def rmat_tuples(scale, seed, start, count, permute=True, abc=(0.57, 0.19, 0.19)):
    thr = rmat_thresholds(*abc)
    s = np.empty(count, np.uint32)
    d = np.empty(count, np.uint32)
    lib().or_rmat_tuples(scale, seed, _ptr(thr), int(permute), start, count, _ptr(s), _ptr(d))
    return s, d


def rmat_csr(scale, edgefactor=16, seed=1, permute=True, abc=(0.57, 0.19, 0.19)):
    thr = rmat_thresholds(*abc)
    nv = 1 << scale
    off = np.empty(nv + 1, np.int64)
    ne = lib().or_rmat_csr(scale, edgefactor, seed, _ptr(thr), int(permute), _ptr(off), None, 0)
    tails = np.empty(ne, np.int32)
    lib().or_rmat_csr(scale, edgefactor, seed, _ptr(thr), int(permute), _ptr(off), _ptr(tails), ne)
    return off, tails


def er_csr(n, seed, thr64):
    off = np.empty(n + 1, np.int64)
    ne = lib().or_er_csr(n, seed, thr64, _ptr(off), None, 0)
    tails = np.empty(ne, np.int32)
    lib().or_er_csr(n, seed, thr64, _ptr(off), _ptr(tails), ne)
    return off, tails


def er_threshold(p):
    return (1 << 32) if p >= 1.0 else int(np.floor(p * 4294967296.0))


def batch_params(base_seed, g, nmin=200, nmax=20000):
    n = ctypes.c_int64()
    kind = ctypes.c_int()
    gseed = ctypes.c_uint64()
    thr = ctypes.c_uint64()
    lib().or_batch_params(base_seed, g, nmin, nmax, ctypes.addressof(n), ctypes.addressof(kind),
                          ctypes.addressof(gseed), ctypes.addressof(thr))
    return n.value, kind.value, gseed.value, thr.value


REF_PARTITION_BIN = os.path.join(_HERE, "_ref", "ref_partition")
REF_RGG_RANKS_BIN = os.path.join(_HERE, "_ref", "ref_rgg_ranks")
MPIEXEC = "/opt/conda/bin/mpiexec"


def partition(off, nparts, method="greedy"):
    """bins int64[nparts+1] of or_partition_greedy / or_partition_linear, or None
    where the restatement reports the reference's undefined cases."""
    off = np.ascontiguousarray(off, np.int64)
    bins = np.zeros(nparts + 1, np.int64)
    f = lib().or_partition_greedy if method == "greedy" else lib().or_partition_linear
    rc = f(len(off) - 1, _ptr(off), nparts, _ptr(bins))
    return bins if rc == 0 else None
