"""Device reductions of result rows (la3_result_reduce): the checksums the
reference's drivers print with VertexProgram::reduce (bfs.cpp:28-31,
sssp.cpp:28-36, cc.cpp:27-30, pr.cpp:51-54), computed without copying the
result to the host. Integer sums and counts exactly equal numpy's over the
same (device-computed) rows; floating-point sums within 1e-12 relative of
numpy's FP64 sum. Rows longer than one reduction block (65,536 values) are
covered by a scale-17 graph."""
import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph(gpu_ctx):
    rec = L.rmat(17, 16, seed=111, weighted=True)
    return L.Graph.from_records(gpu_ctx, rec, 1 << 17, True, L.DIRECTED), rec


def test_reduce_pagerank_checksum(gpu_ctx, graph):
    g, _ = graph
    rank = L.pagerank(g, 20)[0]
    L.pagerank(g, 20, out=L.DEVICE)
    s = gpu_ctx.reduce(np.float64, 1, g.nslots, "sum")
    assert abs(s[0] - rank.sum()) <= 1e-12 * abs(rank.sum())
    n = gpu_ctx.reduce(np.float64, 1, g.nslots, "count_ne", 0.15)  # sources and isolated keep 0.15
    assert n[0] == np.count_nonzero(rank != 0.15)


def test_reduce_batched_bfs_sssp_rows(gpu_ctx, graph):
    g, rec = graph
    roots = np.unique(rec[:, 0])[:37]
    hops, _ = L.bfs_batch(g, roots)
    L.bfs_batch(g, roots, out=L.DEVICE)
    reach = gpu_ctx.reduce(np.uint8, len(roots), g.nslots, "count_ne", L.BFS_INF)
    assert np.array_equal(reach, (hops != L.BFS_INF).sum(axis=1).astype(np.uint64))
    tot = gpu_ctx.reduce(np.uint8, len(roots), g.nslots, "sum")
    assert np.array_equal(tot, hops.astype(np.uint64).sum(axis=1))
    dist, _ = L.sssp_batch(g, roots[:9])
    L.sssp_batch(g, roots[:9], out=L.DEVICE)
    reach = gpu_ctx.reduce(np.uint32, 9, g.nslots, "count_ne", L.SSSP_INF)
    assert np.array_equal(reach, (dist != L.SSSP_INF).sum(axis=1).astype(np.uint64))
    tot = gpu_ctx.reduce(np.uint32, 9, g.nslots, "sum")  # unreached count as 2^31 - 1, as in sssp.cpp:33-35
    assert np.array_equal(tot, dist.astype(np.uint64).sum(axis=1))


def test_reduce_cc_and_ppr(gpu_ctx, graph):
    g, rec = graph
    gu = L.Graph.from_records(gpu_ctx, np.ascontiguousarray(rec[:, :2]), 1 << 17, False, L.UNDIRECTED)
    lab, _ = L.cc(gu)
    L.cc(gu, out=L.DEVICE)
    assert gpu_ctx.reduce(np.uint32, 1, gu.nslots, "sum")[0] == lab.astype(np.uint64).sum()
    seeds = np.unique(rec[:, 0])[:5]
    rank, _, _ = L.ppr_batch(g, seeds, 10)
    L.ppr_batch(g, seeds, 10, out=L.DEVICE)
    s = gpu_ctx.reduce(np.float32, len(seeds), g.nslots, "sum")
    exp = rank.astype(np.float64).sum(axis=1)
    assert np.all(np.abs(s - exp) <= 1e-12 * np.abs(exp))


def test_reduce_errors(gpu_ctx, graph):
    g, _ = graph
    L.pagerank(g, 2, out=L.DEVICE)
    with pytest.raises(L.LA3Error):   # more rows than the result holds
        gpu_ctx.reduce(np.float64, 2, g.nslots, "sum")
    gpu_ctx.release_scratch()
    with pytest.raises(L.LA3Error):   # the result is gone
        gpu_ctx.reduce(np.float64, 1, g.nslots, "sum")


def test_fetch_in_slot_order(gpu_ctx, oracle):
    """la3_result_fetch_slots: the last device-resident result reordered to
    the master-slot order equals the id-order result indexed by the slot
    order's ids (which the oracle's own layout gives), for u8 BFS levels, u32
    batched SSSP rows, f64 PageRank and f32 PPR rows; errors for a result too
    small."""
    rec = L.rmat(12, 16, seed=91, weighted=True)
    n = 1 << 12
    g = L.Graph.from_records(gpu_ctx, rec, n, True, L.DIRECTED)
    o = oracle.OracleGraph(rec, n, True, oracle.DIRECTED)
    slot_ids = np.array([o.unhash(h) for h in o.hid_of()], dtype=np.int64)
    assert (slot_ids == g.array("id_of")).all()
    roots = [int(x) for x in np.unique(rec[:, 0])[:5]]
    hops = L.bfs(g, roots[0], parents=False)[0]
    L.bfs(g, roots[0], parents=False, out=L.DEVICE)
    assert (gpu_ctx.fetch_slots(g, np.uint8)[0] == hops[slot_ids]).all()
    dist = L.sssp_batch(g, roots)[0]
    L.sssp_batch(g, roots, out=L.DEVICE)
    got = gpu_ctx.fetch_slots(g, np.uint32, rows=len(roots))
    assert (got == dist[:, slot_ids]).all()
    rank = L.pagerank(g, 10)[0]
    L.pagerank(g, 10, out=L.DEVICE)
    assert (gpu_ctx.fetch_slots(g, np.float64)[0] == rank[slot_ids]).all()
    pr = L.ppr_batch(g, roots[:3], 5)[0]
    L.ppr_batch(g, roots[:3], 5, out=L.DEVICE)
    assert (gpu_ctx.fetch_slots(g, np.float32, rows=3) == pr[:, slot_ids]).all()
    with pytest.raises(L.LA3Error):
        gpu_ctx.fetch_slots(g, np.float32, rows=4)
