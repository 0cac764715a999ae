"""Small graph builders and independent checkers shared by the tests."""
import hashlib

import numpy as np


def csr_sha256(off, tails):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, np.int64).tobytes())
    h.update(np.ascontiguousarray(tails, np.int64).tobytes())
    return h.hexdigest()


def csr_from_edges(n, edges):
    """Symmetric sorted unique CSR (int64) from undirected pairs; self-loops kept once."""
    s = []
    for a, b in edges:
        s.append((a, b))
        if a != b:
            s.append((b, a))
    s = sorted(set(s))
    off = np.zeros(n + 1, np.int64)
    for a, _ in s:
        off[a + 1] += 1
    off = np.cumsum(off)
    tails = np.array([b for _, b in s], np.int64)
    return off, tails


def complete(n):
    return csr_from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])


def bipartite(a, b):
    return csr_from_edges(a + b, [(i, a + j) for i in range(a) for j in range(b)])


def grid(r, c):
    e = []
    for i in range(r):
        for j in range(c):
            v = i * c + j
            if j + 1 < c:
                e.append((v, v + 1))
            if i + 1 < r:
                e.append((v, v + c))
    return csr_from_edges(r * c, e)


def wheel(n):
    """Hub 0 joined to a cycle 1..n-1: n-1 triangles, all credited to vertex 0."""
    e = [(0, i) for i in range(1, n)]
    e += [(i, i + 1) for i in range(1, n - 1)] + [(n - 1, 1)]
    return csr_from_edges(n, e)


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    e = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]
    return csr_from_edges(n, e)


def dense_adj(off, tails):
    n = len(off) - 1
    A = np.zeros((n, n), np.int64)
    for u in range(n):
        A[u, tails[off[u]:off[u + 1]]] = 1
    np.fill_diagonal(A, 0)
    return A


def brute_force(off, tails):
    """O(n^3): total and per-vertex counts with the smallest-id credit."""
    A = dense_adj(off, tails)
    n = A.shape[0]
    pv = np.zeros(n, np.uint64)
    for u in range(n):
        for v in range(u + 1, n):
            if not A[u, v]:
                continue
            for w in range(v + 1, n):
                if A[u, w] and A[v, w]:
                    pv[u] += 1
    return int(pv.sum()), pv


def sparse_closed_form(off, tails):
    """Per-vertex = row sums of (U U) .* U, U = strictly upper adjacency (scipy)."""
    import scipy.sparse as sp
    n = len(off) - 1
    rows = np.repeat(np.arange(n), np.diff(off))
    cols = np.asarray(tails, np.int64)
    keep = cols > rows
    U = sp.csr_matrix((np.ones(keep.sum(), np.int64), (rows[keep], cols[keep])), shape=(n, n))
    T = (U @ U).multiply(U)
    pv = np.asarray(T.sum(axis=1)).ravel().astype(np.uint64)
    return int(pv.sum()), pv


def min_mid_pairs(off, tails):
    """For every triangle u < v < w, the pair (u, v) -- its minimum and middle
    vertex -- with multiplicity, by brute-force intersection of the higher
    neighbour lists (small graphs only)."""
    off = np.asarray(off, np.int64)
    tails = np.asarray(tails, np.int64)
    n = len(off) - 1
    up = [tails[off[u]:off[u + 1]][tails[off[u]:off[u + 1]] > u] for u in range(n)]
    us, vs, cs = [], [], []
    for u in range(n):
        for v in up[u]:
            c = len(np.intersect1d(up[u], up[v], assume_unique=True))
            if c:
                us.append(u)
                vs.append(v)
                cs.append(c)
    return np.array(us, np.int64), np.array(vs, np.int64), np.array(cs, np.int64)
