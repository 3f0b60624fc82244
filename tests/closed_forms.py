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
    regular ranks. iters == 1 is the single-iteration path: one step for
    regular and sink vertices alike, from the initial ranks."""
    rec = dedup_first(rec)
    cls = classes(rec, n)
    outd = out_degree(rec, n).astype(np.float64)
    A = sp.csr_matrix((np.ones(len(rec)), (rec[:, 1].astype(np.int64), rec[:, 0].astype(np.int64))),
                      shape=(n + 1, n + 1))
    base = np.full(n + 1, alpha) if seed is None else np.where(np.arange(n + 1) == seed, alpha, 0.0)
    r = base.copy()
    reg = cls["regular"]
    if iters == 1:
        x = np.divide(r, outd, out=np.zeros_like(r), where=outd > 0)
        return np.where(reg | cls["sink"], base + (1 - alpha) * (A @ x), r)
    for _ in range(iters):
        x = np.divide(r, outd, out=np.zeros_like(r), where=outd > 0)
        y = A @ x
        r = np.where(reg, base + (1 - alpha) * y, r)
    x = np.divide(r, outd, out=np.zeros_like(r), where=outd > 0)
    y = A @ x
    r = np.where(cls["sink"], base + (1 - alpha) * y, r)
    return r


# ------------------------------------------------------------ ingest layout ---
def bucket_hash_table(n, nbuckets=128):
    """Hashed id of every id in [0, n] for the reference's bucket hashing at one
    rank (src/matrix/hashers.h:39-67, 128 parts), built from its bucket
    structure rather than its formula: the first height*128 ids are dealt into
    128 buckets by id mod 128 (bucket b holds b, b+128, b+256, ... in
    increasing order), the buckets are laid end to end, and an id's hashed
    value is its position in that sequence; ids from height*128 on keep their
    value."""
    height = n // nbuckets
    max_range = height * nbuckets
    out = np.arange(n + 1, dtype=np.int64)
    if height == 0:
        return out
    pos = 0
    for b in range(nbuckets):
        members = np.arange(b, max_range, nbuckets)
        out[members] = pos + np.arange(members.size)
        pos += members.size
    assert pos == max_range
    return out


def oriented_triples(rec, mode, weighted):
    """Stored (row, col, weight) triples in insertion order for a load mode
    (src/matrix/graph.hpp:165-191), before hashing and deduplication.
    Modes: 0 directed, 1 reversed, 2 undirected, 3 acyclic, 4 reversed acyclic."""
    rec = np.asarray(rec, dtype=np.int64)
    a, b = rec[:, 0], rec[:, 1]
    w = rec[:, 2] if weighted else np.zeros(a.size, dtype=np.int64)
    keep = a != b
    a, b, w = a[keep], b[keep], w[keep]
    if mode == 2:
        # record k inserts (a, b) and then its mirror (b, a)
        row = np.stack([a, b], 1).reshape(-1)
        col = np.stack([b, a], 1).reshape(-1)
        return row, col, np.repeat(w, 2)
    reverse = mode in (1, 4)
    row, col = (a, b) if reverse else (b, a)
    if mode in (3, 4):
        flip = (col > row) if not reverse else (col < row)
        row, col = np.where(flip, col, row), np.where(flip, row, col)
    return row, col, w


def expected_layout(rec, n, weighted, mode, hashing=True):
    """The tile's layout at one rank recomputed from the records alone:
    slot order [regular | sink | source | isolated], each by increasing hashed
    id (src/utils/locator.h:84-120), and the regular-row and sink-row DCSCs
    (src/utils/csc.h:76-130): columns [regular | source] by hashed id, entries
    of a column by row position, colidxs the hashed column id where the
    column has entries in that CSC (0 elsewhere and in the last slot), the
    weight of the first record of a duplicated (row, col) pair
    (unordered_set::insert keeps the first, src/matrix/csc_matrix2d.hpp:62-72)."""
    H = bucket_hash_table(n) if hashing else np.arange(n + 1, dtype=np.int64)
    row, col, w = oriented_triples(rec, mode, weighted)
    row, col = H[row], H[col]
    key = row * (n + 1) + col
    _, first = np.unique(key, return_index=True)
    first = np.sort(first)
    row, col, w = row[first], col[first], w[first]
    S = n + 1
    has_row = np.zeros(S, bool)
    has_col = np.zeros(S, bool)
    has_row[row] = True
    has_col[col] = True
    reg = np.flatnonzero(has_row & has_col)
    sink = np.flatnonzero(has_row & ~has_col)
    src = np.flatnonzero(~has_row & has_col)
    iso = np.flatnonzero(~has_row & ~has_col)
    hid_of = np.concatenate([reg, sink, src, iso]).astype(np.uint32)
    cols = np.concatenate([reg, src])
    colpos = np.full(S, -1, dtype=np.int64)
    colpos[cols] = np.arange(cols.size)
    regpos = np.full(S, -1, dtype=np.int64)
    regpos[reg] = np.arange(reg.size)
    sinkpos = np.full(S, -1, dtype=np.int64)
    sinkpos[sink] = np.arange(sink.size)
    out = dict(hid_of=hid_of, nreg=reg.size, nsink=sink.size, nsrc=src.size, ncols=cols.size)
    for which, posmap in ((0, regpos), (1, sinkpos)):
        sel = posmap[row] >= 0
        r, c, ww = row[sel], col[sel], w[sel]
        gi = posmap[r]
        cp = colpos[c]
        order = np.lexsort((gi, cp))
        counts = np.bincount(cp, minlength=cols.size)
        colptrs = np.zeros(cols.size + 1, dtype=np.uint32)
        colptrs[1:] = np.cumsum(counts)
        colidxs = np.zeros(cols.size + 1, dtype=np.uint32)
        colidxs[:-1] = np.where(counts > 0, cols, 0)
        out[which] = dict(colptrs=colptrs, colidxs=colidxs, global_idx=gi[order].astype(np.uint32),
                          idx=r[order].astype(np.uint32), w=ww[order].astype(np.uint32))
    return out


# ------------------------------------------------------- bounded executions ---
def hop_bounded_dist(rec_w, n, root, k):
    """Shortest distances from root over paths of at most k edges (k rounds
    of synchronous Bellman-Ford relaxation over the deduplicated edges)."""
    rec = dedup_first(rec_w)
    a, b, w = rec[:, 0].astype(np.int64), rec[:, 1].astype(np.int64), rec[:, 2].astype(np.int64)
    inf = np.iinfo(np.int64).max // 4
    d = np.full(n + 1, inf, dtype=np.int64)
    d[root] = 0
    for _ in range(k):
        nd = d.copy()
        ok = d[a] < inf
        np.minimum.at(nd, b[ok], d[a[ok]] + w[ok])
        d = nd
    out = np.full(n + 1, SSSP_INF, dtype=np.uint32)
    out[d < inf] = d[d < inf].astype(np.uint32)
    return out


def cc_bounded(rec, n, m):
    """Labels after m rounds of min-label propagation on the undirected graph:
    the smallest id within m hops of each vertex."""
    A = adjacency(rec, n, undirected=True)
    lab = np.arange(n + 1, dtype=np.int64)
    for _ in range(m):
        nxt = lab.copy()
        coo = A.tocoo()
        np.minimum.at(nxt, coo.row, lab[coo.col])
        lab = nxt
    return lab.astype(np.uint32)


def cc_rounds(rec, n):
    """1 + the largest hop distance from a vertex to its component's
    smallest id: the iteration count of min-label propagation run until an
    iteration changes no label."""
    A = adjacency(rec, n, undirected=True)
    mins = min_id_components(rec, n)
    roots = np.unique(mins)
    d = csgraph.shortest_path(A, directed=False, unweighted=True, indices=roots)
    rowof = np.searchsorted(roots, mins)
    dist = d[rowof, np.arange(n + 1)]
    return 1 + int(dist.max())
