"""Reference-independent answers for the LA3 programs, from textbook or
library routines (scipy.sparse.csgraph), used to pin the oracle.

Every function takes LA3 binary records (u32 pairs or triples, edge a->b)
and applies the reference's ingest rules itself: self-loops dropped,
duplicates resolved with the first record winning.
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse import csgraph

BFS_INF = 255
SSSP_INF = 2**31 - 1


def dedup_first(rec):
    """Drop self-loops, keep the first record of each (a, b) pair."""
    rec = np.asarray(rec)
    keep = rec[:, 0] != rec[:, 1]
    rec = rec[keep]
    key = rec[:, 0].astype(np.uint64) << np.uint64(32) | rec[:, 1].astype(np.uint64)
    _, first = np.unique(key, return_index=True)
    return rec[np.sort(first)]


def adjacency(rec, n, undirected=False, weights=False):
    rec = dedup_first(rec)
    a, b = rec[:, 0].astype(np.int64), rec[:, 1].astype(np.int64)
    if undirected:
        a, b = np.concatenate([a, b]), np.concatenate([b, a])
        key = a * (n + 1) + b
        _, first = np.unique(key, return_index=True)
        a, b = a[first], b[first]
        data = np.ones(a.size)
    else:
        data = rec[:, 2].astype(np.float64) if weights else np.ones(a.size)
    return sp.csr_matrix((data, (a, b)), shape=(n + 1, n + 1))


def bfs_levels(rec, n, root, undirected=False):
    A = adjacency(rec, n, undirected)
    d = csgraph.shortest_path(A, directed=not undirected, unweighted=True, indices=int(root))
    out = np.full(n + 1, BFS_INF, dtype=np.uint8)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint8)
    return out


def dijkstra(rec_w, n, root):
    A = adjacency(rec_w, n, weights=True)
    d = csgraph.dijkstra(A, directed=True, indices=int(root))
    out = np.full(n + 1, SSSP_INF, dtype=np.uint32)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint32)
    return out


def min_id_components(rec, n):
    A = adjacency(rec, n, undirected=True)
    _, lab = csgraph.connected_components(A, directed=False)
    mins = np.full(lab.max() + 1, np.iinfo(np.int64).max)
    np.minimum.at(mins, lab, np.arange(n + 1))
    return mins[lab].astype(np.uint32)


def triangles(rec, n):
    """Total triangle count of the underlying simple undirected graph, and per
    vertex the number of triangles in which it has the middle id (brute force
    over wedges; small graphs only)."""
    A = adjacency(rec, n, undirected=True)
    U = sp.triu(A, k=1).tocsr()   # U[w, h] for w < h
    upper = [set(U.indices[U.indptr[i]:U.indptr[i + 1]].tolist()) for i in range(n + 1)]
    lower = [[] for _ in range(n + 1)]
    for w in range(n + 1):
        for h in upper[w]:
            lower[h].append(w)
    per = np.zeros(n + 1, dtype=np.int64)
    for m in range(n + 1):
        for h in upper[m]:
            per[m] += sum(1 for w in lower[m] if h in upper[w])
    return int(per.sum()), per.astype(np.uint32)


def triangle_total_trace(rec, n):
    """sum(A^2 .* A) / 6 on the simple undirected graph."""
    A = adjacency(rec, n, undirected=True).astype(np.int64)
    return int((A @ A).multiply(A).sum()) // 6


def in_degree(rec, n):
    rec = dedup_first(rec)
    return np.bincount(rec[:, 1], minlength=n + 1).astype(np.uint32)


def out_degree(rec, n):
    rec = dedup_first(rec)
    return np.bincount(rec[:, 0], minlength=n + 1).astype(np.uint32)


def classes(rec, n):
    """Vertex classes of a directed load (row = destination, column = source)."""
    rec = dedup_first(rec)
    has_in = np.zeros(n + 1, bool)
    has_out = np.zeros(n + 1, bool)
    has_in[rec[:, 1]] = True
    has_out[rec[:, 0]] = True
    return dict(regular=has_in & has_out, sink=has_in & ~has_out, source=~has_in & has_out,
                isolated=~has_in & ~has_out)


def pagerank_dense(rec, n, iters, alpha=0.15, seed=None):
    """PR (or PPR with teleport at `seed`) as a dense power iteration with the
    reference's structure: regular vertices iterate, sources and isolated
    vertices keep their initial rank, sinks take one step from the final
    regular ranks."""
    rec = dedup_first(rec)
    cls = classes(rec, n)
    outd = out_degree(rec, n).astype(np.float64)
    A = sp.csr_matrix((np.ones(len(rec)), (rec[:, 1].astype(np.int64), rec[:, 0].astype(np.int64))),
                      shape=(n + 1, n + 1))
    base = np.full(n + 1, alpha) if seed is None else np.where(np.arange(n + 1) == seed, alpha, 0.0)
    r = base.copy()
    reg = cls["regular"]
    for _ in range(iters):
        x = np.divide(r, outd, out=np.zeros_like(r), where=outd > 0)
        y = A @ x
        r = np.where(reg, base + (1 - alpha) * y, r)
    x = np.divide(r, outd, out=np.zeros_like(r), where=outd > 0)
    y = A @ x
    r = np.where(cls["sink"], base + (1 - alpha) * y, r)
    return r
