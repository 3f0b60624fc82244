"""Device top-k of result rows (la3_result_topk, the reference's
VertexProgram::topk, vertex_program.hpp:607-680): per row the k largest or
smallest values, then smaller id first for equal values. Checked exactly
against numpy's stable sort of the same (device-computed) result rows."""
import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu


def _expect(rows, k, largest):
    out_i, out_v = [], []
    for r in rows:
        ids = np.arange(r.size, dtype=np.int64)
        order = np.lexsort((ids, -r.astype(np.float64) if largest else r.astype(np.float64)))
        sel = order[:k]
        out_i.append(np.concatenate([sel.astype(np.uint32), np.full(k - sel.size, 0xffffffff, np.uint32)]))
        out_v.append(np.concatenate([r[sel], np.zeros(k - sel.size, r.dtype)]))
    return np.array(out_i), np.array(out_v)


@pytest.fixture(scope="module")
def graph(gpu_ctx):
    rec = L.rmat(14, 16, seed=101, weighted=True)
    return L.Graph.from_records(gpu_ctx, rec, 1 << 14, True, L.DIRECTED), rec


def test_topk_pagerank_f64(gpu_ctx, graph):
    g, _ = graph
    rank = L.pagerank(g, 20)[0]
    L.pagerank(g, 20, out=L.DEVICE)
    for k in (1, 10, 1000):
        ids, vals = gpu_ctx.topk(np.float64, 1, g.nslots, k, largest=True)
        ei, ev = _expect([rank], k, True)
        assert np.array_equal(ids, ei) and np.array_equal(vals, ev)


def test_topk_batched_rows_with_ties(gpu_ctx, graph):
    g, rec = graph
    roots = np.unique(rec[:, 0])[:70]
    hops, _ = L.bfs_batch(g, roots)
    L.bfs_batch(g, roots, out=L.DEVICE)
    for k in (5, 300):   # u8 levels: thousands of equal values
        ids, vals = gpu_ctx.topk(np.uint8, len(roots), g.nslots, k, largest=False)
        ei, ev = _expect(hops, k, False)
        assert np.array_equal(ids, ei) and np.array_equal(vals, ev)
    dist, _ = L.sssp_batch(g, roots[:9])
    L.sssp_batch(g, roots[:9], out=L.DEVICE)
    ids, vals = gpu_ctx.topk(np.uint32, 9, g.nslots, 64, largest=False)
    ei, ev = _expect(dist, 64, False)
    assert np.array_equal(ids, ei) and np.array_equal(vals, ev)
    ids, vals = gpu_ctx.topk(np.uint32, 9, g.nslots, 64, largest=True)   # the unreached (INF) first
    ei, ev = _expect(dist, 64, True)
    assert np.array_equal(ids, ei) and np.array_equal(vals, ev)


def test_topk_ppr_f32_and_padding(gpu_ctx, graph, g1):
    g, rec = graph
    seeds = np.unique(rec[:, 0])[:5]
    rank, _, _ = L.ppr_batch(g, seeds, 10)
    L.ppr_batch(g, seeds, 10, out=L.DEVICE)
    ids, vals = gpu_ctx.topk(np.float32, 5, g.nslots, 50, largest=True)
    ei, ev = _expect(rank, 50, True)
    assert np.array_equal(ids, ei) and np.array_equal(vals, ev)
    # k above the number of values: every value, then padding
    s = L.Graph.from_records(gpu_ctx, g1["records"], 8, False, L.DIRECTED)
    deg = L.degree(s)
    L.degree(s, out=L.DEVICE)
    ids, vals = gpu_ctx.topk(np.uint32, 1, s.nslots, 12, largest=True)
    ei, ev = _expect([deg], 12, True)
    assert np.array_equal(ids, ei) and np.array_equal(vals, ev)
    with pytest.raises(L.LA3Error):
        gpu_ctx.topk(np.uint32, 2, s.nslots, 3)   # larger than the last result
