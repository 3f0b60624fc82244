"""Programs on a graph uploaded from the reference's two CSCs
(la3_graph_from_csc, device ingest of the rebuilt records) give exactly the
results of the same graph loaded from its records: BFS levels and parents,
SSSP distances, PageRank (directed), CC labels (undirected), TC (acyclic)."""
import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu


def _upload(ctx, g, n, weighted, mode):
    return L.Graph.from_csc(ctx, n, (g.csc(0), g.csc(1)), weighted, mode)


def test_uploaded_graph_runs_like_the_original(gpu_ctx):
    n = 1 << 14
    rec = L.rmat(14, 16, seed=141, weighted=True)
    g = L.Graph.from_records(gpu_ctx, rec, n, True, L.DIRECTED)
    u = _upload(gpu_ctx, g, n, True, L.DIRECTED)
    for r in [int(x) for x in np.unique(rec[:, 0])[:4]]:
        a, b = L.bfs(g, r), L.bfs(u, r)
        assert all(np.array_equal(x, y) for x, y in zip(a[:2], b[:2])) and a[2] == b[2]
        assert np.array_equal(L.sssp(g, r)[0], L.sssp(u, r)[0])
    assert np.array_equal(L.pagerank(g, 20)[0], L.pagerank(u, 20)[0])
    pairs = np.ascontiguousarray(rec[:, :2])
    gu = L.Graph.from_records(gpu_ctx, pairs, n, False, L.UNDIRECTED)
    uu = _upload(gpu_ctx, gu, n, False, L.UNDIRECTED)
    la, ia = L.cc(gu)
    lb, ib = L.cc(uu)
    assert np.array_equal(la, lb) and ia == ib
    ga = L.Graph.from_records(gpu_ctx, pairs, n, False, L.ACYCLIC)
    ua = _upload(gpu_ctx, ga, n, False, L.ACYCLIC)
    ta, pa = L.tc(ga)
    tb, pb = L.tc(ua)
    assert ta == tb and np.array_equal(pa, pb)
