"""execute(max_iters) on the device against the oracle's execute(max_iters)
(vertex_program_execute.hpp:11-52): bounded BFS, SSSP and CC; the
single-iteration path (sinks from the initial messages) of PR and PPR on the
single, batched and fused paths; CC on the vertex-program engine, including
graphs where vertex 0 has edges (cc.h:31's 'label 0 is empty' rule) and
directed loads; and the device-resident result after its buffer is freed."""
import os

import numpy as np
import pytest

import la3_amd as L
from test_oracle import random_records

pytestmark = pytest.mark.gpu


def _both(ctx, oracle, rec, n, weighted, mode):
    return (L.Graph.from_records(ctx, rec, n, weighted, mode),
            oracle.OracleGraph(rec, n, weighted, mode))


def test_g1_execute_two(gpu_ctx, oracle, g1):
    """The reference's g1 fixture with execute(2) for BFS and SSSP."""
    g, o = _both(gpu_ctx, oracle, g1["records"], 8, False, L.UNDIRECTED)
    for m in (1, 2, 3):
        hops, par, it = L.bfs(g, 5, max_iters=m)
        eh, ep, eit = oracle.bfs(o, 5, m)
        assert (hops == eh).all() and (par == ep).all() and it == eit == m
    gw, ow = _both(gpu_ctx, oracle, g1["records_w"], 8, True, L.DIRECTED)
    for m in (1, 2, 3):
        d, it = L.sssp(gw, 5, max_iters=m)
        e, eit = oracle.sssp(ow, 5, m)
        assert (d == e).all() and it == eit == m


@pytest.mark.parametrize("seed", [31, 32])
def test_bfs_bounded(gpu_ctx, oracle, seed):
    rec = random_records(seed, 3000, 12000)
    for mode in (L.DIRECTED, L.UNDIRECTED):
        g, o = _both(gpu_ctx, oracle, rec, 3000, False, mode)
        for root in (1, 2, 3, 777):
            for m in (1, 2, 3, 6):
                hops, par, it = L.bfs(g, root, max_iters=m)
                eh, ep, eit = oracle.bfs(o, root, m)
                assert (hops == eh).all() and (par == ep).all() and it == eit, (mode, root, m)


def test_bfs_bounded_rmat(gpu_ctx, oracle):
    rec = L.rmat(14, 16, seed=33)
    g, o = _both(gpu_ctx, oracle, rec, 1 << 14, False, L.DIRECTED)
    roots = [int(x) for x in np.unique(rec[:, 0])[:6]]
    for root in roots:
        for m in (1, 2, 3):
            hops, par, it = L.bfs(g, root, max_iters=m)
            eh, ep, eit = oracle.bfs(o, root, m)
            assert (hops == eh).all() and (par == ep).all() and it == eit
    batch, it = L.bfs_batch(g, roots, max_iters=2)
    assert it == 2
    for q, r in enumerate(roots):
        assert (batch[q] == oracle.bfs(o, r, 2)[0]).all()


@pytest.mark.parametrize("seed", [34, 35])
def test_sssp_bounded(gpu_ctx, oracle, seed):
    rec = random_records(seed, 3000, 16000, weighted=True)
    g, o = _both(gpu_ctx, oracle, rec, 3000, True, L.DIRECTED)
    for root in (1, 2, 5, 1234):
        for m in (1, 2, 3, 5, 9):
            d, it = L.sssp(g, root, max_iters=m)
            e, eit = oracle.sssp(o, root, m)
            assert (d == e).all() and it == eit, (root, m)


def test_sssp_bounded_rmat_and_batch(gpu_ctx, oracle):
    rec = L.rmat(14, 16, seed=36, weighted=True)
    g, o = _both(gpu_ctx, oracle, rec, 1 << 14, True, L.DIRECTED)
    roots = [int(x) for x in np.unique(rec[:, 0])[:5]]
    for m in (1, 2, 4):
        batch, it = L.sssp_batch(g, roots, max_iters=m)
        assert it == m
        for q, r in enumerate(roots):
            e, _ = oracle.sssp(o, r, m)
            assert (batch[q] == e).all(), (r, m)
            assert (L.sssp(g, r, max_iters=m)[0] == e).all()


@pytest.mark.parametrize("seed", [37, 38])
def test_cc_engine_labels_and_iterations(gpu_ctx, oracle, seed):
    rec = random_records(seed, 4000, 3500)
    g, o = _both(gpu_ctx, oracle, rec, 4000, False, L.UNDIRECTED)
    lab, it = L.cc(g)
    elab, eit = oracle.cc(o)
    assert (lab == elab).all() and it == eit
    for m in (1, 2, 3):
        lm, itm = L.cc(g, max_iters=m)
        el, eitm = oracle.cc(o, m)
        assert (lm == el).all() and itm == eitm == m


def test_cc_vertex_zero_has_edges(gpu_ctx, oracle):
    """Ids from 0: the reference's combine treats an accumulator of 0 as empty
    (cc.h:31), so label 0 survives only where it is a row's last message. The
    engine combines in the reference's order and so reproduces it."""
    for seed in (39, 40):
        rec = random_records(seed, 2000, 3000, zero_id=True)
        rec = np.concatenate([rec, np.array([[0, 1], [0, 2], [3, 0]], dtype=np.uint32)])
        g, o = _both(gpu_ctx, oracle, rec, 2000, False, L.UNDIRECTED)
        lab, it = L.cc(g)
        elab, eit = oracle.cc(o)
        assert (lab == elab).all() and it == eit


def test_cc_directed_load(gpu_ctx, oracle):
    """CcVertex on a directed load runs the optimized execution (sinks after
    the loop) in the reference; the engine follows it."""
    rec = random_records(41, 3000, 6000)
    g, o = _both(gpu_ctx, oracle, rec, 3000, False, L.DIRECTED)
    for m in (0, 1, 2, 4):
        lab, it = L.cc(g, max_iters=m)
        elab, eit = oracle.cc(o, m)
        assert (lab == elab).all() and it == eit, m


def test_cc_rmat_scale16(gpu_ctx, oracle):
    rec = L.rmat(16, 16, seed=42)
    g, o = _both(gpu_ctx, oracle, rec, 1 << 16, False, L.UNDIRECTED)
    lab, it = L.cc(g)
    elab, eit = oracle.cc(o)
    assert (lab == elab).all() and it == eit


def test_pagerank_single_iteration(gpu_ctx, oracle):
    """execute(1) is execute_single: sinks from the initial ranks."""
    rec = L.rmat(12, 16, seed=43)
    g = L.Graph.from_records(gpu_ctx, rec, 1 << 12, False, L.DIRECTED)
    o = oracle.OracleGraph(rec, 1 << 12, False, oracle.DIRECTED)
    orr = oracle.OracleGraph(rec, 1 << 12, False, oracle.REVERSED)
    for iters in (1, 2, 3):
        rank, _, it = L.pagerank(g, iters)
        erank, _, eit = oracle.pagerank(o, orr, iters)
        assert np.abs(rank - erank).sum() <= 1e-9 and it == eit == iters


def test_ppr_single_iteration_all_paths(gpu_ctx, oracle):
    rec = L.rmat(12, 16, seed=44)
    n = 1 << 12
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    o = oracle.OracleGraph(rec, n, False, oracle.DIRECTED)
    orr = oracle.OracleGraph(rec, n, False, oracle.REVERSED)
    seeds = [int(x) for x in np.unique(rec[:, 0])[:4]]
    for iters in (1, 2):
        exp = [oracle.ppr(o, orr, s, iters)[0] for s in seeds]
        for q, s in enumerate(seeds):
            assert np.abs(L.ppr(g, s, iters) - exp[q]).sum() <= 1e-9
        batch, _, _ = L.ppr_batch(g, seeds, iters)
        fused = L.batch_run(gpu_ctx, [(g, L.Q_PPR, seeds, iters)])[0]["out"]
        assert np.array_equal(batch, fused)
        for q in range(len(seeds)):
            assert np.abs(batch[q].astype(np.float64) - exp[q]).sum() <= 1e-7, (iters, q)


def test_result_fetch_after_slot_release(gpu_ctx):
    """A session batch finished with a NULL output leaves its rows as the
    last result; when its slot's buffers are freed (the keep limit set to 0),
    fetching it fails cleanly instead of reading freed device memory. With
    the default limit the rows come back."""
    rec = L.rmat(11, 16, seed=45, weighted=True)
    g = L.Graph.from_records(gpu_ctx, rec, 1 << 11, True, L.DIRECTED)
    roots = [int(x) for x in np.unique(rec[:, 0])[:8]]
    exp, _ = L.sssp_batch(g, roots)
    for keep, ok in (("0", False), (None, True)):
        old = os.environ.pop("LA3_SESSION_KEEP_BYTES", None)
        if keep is not None:
            os.environ["LA3_SESSION_KEEP_BYTES"] = keep
        try:
            s = L.Session(gpu_ctx)
            s.add(g, L.Q_SSSP, roots, out=L.DEVICE)
            while s.step():
                pass
            assert [st for _, _, st in s.poll()] == [0]
            s.close()
        finally:
            os.environ.pop("LA3_SESSION_KEEP_BYTES", None)
            if old is not None:
                os.environ["LA3_SESSION_KEEP_BYTES"] = old
        got = np.zeros((len(roots), g.nslots), dtype=np.uint32)
        if ok:
            gpu_ctx.fetch(got)
            assert np.array_equal(got, exp)
        else:
            with pytest.raises(L.LA3Error) as e:
                gpu_ctx.fetch(got)
            assert e.value.status == L.EINVAL
            with pytest.raises(L.LA3Error):
                gpu_ctx.topk(np.uint32, len(roots), g.nslots, 4, largest=False)
    gpu_ctx.release_scratch()
