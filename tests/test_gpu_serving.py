"""Serving loop: resident graphs of different sizes, a mixed BFS / SSSP / PPR
stream submitted while batches run, and every served result equal to the
same query run alone (bit-exact; PPR bit-identical to a batch of one and
within L1 1e-7 of the FP64 single-query program). Both admission modes:
per-(program, graph) batches, and mixed batches across programs and graphs
run as one fused batch run."""
import time

import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs(gpu_ctx):
    out = []
    for scale, seed in ((10, 61), (12, 62), (13, 63)):
        rec = L.rmat(scale, 16, seed=seed, weighted=True)
        g = L.Graph.from_records(gpu_ctx, rec, 1 << scale, True, L.DIRECTED)
        out.append((g, np.unique(rec[:, 0])))
    return out


@pytest.mark.parametrize("fused,max_wait_s", [(False, 0.0), (False, 0.05), (True, 0.05)])
def test_mixed_stream_equals_alone(gpu_ctx, graphs, fused, max_wait_s):
    srv = L.Server(0, nworkers=3, max_batch=16, state_budget=1 << 30, keep_results=True, fused=fused,
                   max_wait_s=max_wait_s)
    gids = [srv.add_graph(g) for g, _ in graphs]
    srv.warm()  # every worker runs each program on each graph (a no-op when fused)
    assert srv.completed() == []  # warm-up runs are not queries
    rng = np.random.default_rng(7)
    subs = []
    for i in range(120):
        gi = int(rng.integers(0, len(graphs)))
        kind = [L.Q_BFS, L.Q_SSSP, L.Q_PPR][int(rng.choice(3, p=[0.5, 0.3, 0.2]))]
        root = int(rng.choice(graphs[gi][1]))
        subs.append((srv.submit(kind, gids[gi], root, 30 if kind == L.Q_PPR else 0), kind, gi, root))
    srv.drain()
    info = {d["qid"]: d for d in srv.completed()}
    assert len(info) == len(subs)
    assert all(d["status"] == 0 for d in info.values())
    assert max(d["batch_size"] for d in info.values()) > 1, "no batching happened"
    for qid, kind, gi, root in subs:
        g = graphs[gi][0]
        got = srv.result(qid, kind, g.nslots)
        d = info[qid]
        assert d["submit_s"] <= d["start_s"] <= d["done_s"]
        if kind == L.Q_BFS:
            assert (got == L.bfs(g, root, parents=False)[0]).all()
        elif kind == L.Q_SSSP:
            assert (got == L.sssp(g, root)[0]).all()
        else:
            alone, _, _ = L.ppr_batch(g, [root], 30)
            assert (got == alone[0]).all()
            assert float(np.abs(got - L.ppr(g, root, 30)).sum()) <= 1e-7
    srv.close()


@pytest.mark.parametrize("fused", [False, True])
def test_admission_budget_limits_batch(gpu_ctx, graphs, fused):
    g, roots = graphs[2]
    srv = L.Server(0, nworkers=1, max_batch=64, state_budget=1, keep_results=False, fused=fused)
    gid = srv.add_graph(g)
    one = srv.query_bytes(L.Q_SSSP, gid)
    assert one > 0
    for r in roots[:20]:
        srv.submit(L.Q_SSSP, gid, int(r))
    srv.drain()
    done = srv.completed()
    # a budget below one query still admits lone queries (nothing else runs), one at a time
    assert len(done) == 20 and max(d["batch_size"] for d in done) == 1
    srv.close()


@pytest.mark.parametrize("fused", [False, True])
def test_graphs_added_while_serving(gpu_ctx, graphs, fused):
    """Graphs registered while workers run batches: the workers copy a
    graph's entry under the server lock, so growing the table is safe."""
    srv = L.Server(0, nworkers=2, max_batch=8, state_budget=1 << 30, keep_results=True, fused=fused)
    g0, roots0 = graphs[0]
    gid0 = srv.add_graph(g0)
    rng = np.random.default_rng(11)
    subs = []
    for i in range(40):
        root = int(rng.choice(roots0))
        subs.append((srv.submit(L.Q_BFS, gid0, root, 0), g0, root))
        if i % 10 == 5:
            for g, roots in graphs[1:]:
                gid = srv.add_graph(g)
                r = int(rng.choice(roots))
                subs.append((srv.submit(L.Q_SSSP, gid, r, 0), g, r))
    srv.drain()
    info = {d["qid"]: d for d in srv.completed()}
    assert len(info) == len(subs) and all(d["status"] == 0 for d in info.values())
    for qid, g, root in subs:
        kind = L.Q_BFS if g is g0 else L.Q_SSSP
        got = srv.result(qid, kind, g.nslots)
        want = L.bfs(g, root, parents=False)[0] if kind == L.Q_BFS else L.sssp(g, root)[0]
        assert (got == want).all()
    srv.close()


@pytest.mark.parametrize("max_wait_s", [0.0, 10.0])
def test_batch_order(gpu_ctx, graphs, max_wait_s):
    """Per-(program, graph) mode, one worker: while a long PPR batch runs, a
    second PPR group on the large graph and a BFS group on the small graph
    queue up. Oldest first (max_wait_s = 0) runs the PPR group next; least
    expected work first (no query has waited max_wait_s) runs the BFS group
    next."""
    (gs, rs), (gl, rl) = graphs[0], graphs[2]
    srv = L.Server(0, nworkers=1, max_batch=8, state_budget=1 << 30, keep_results=False, fused=False,
                   max_wait_s=max_wait_s)
    ids, il = srv.add_graph(gs), srv.add_graph(gl)
    first = srv.submit(L.Q_PPR, il, int(rl[0]), 1000)  # long enough for the rest to queue
    time.sleep(0.005)  # usually the worker has taken it
    ppr = srv.submit(L.Q_PPR, il, int(rl[1]), 29)      # another (kind, graph, iters) group
    bfs = [srv.submit(L.Q_BFS, ids, int(r), 0) for r in rs[:3]]
    srv.drain()
    done = {d["qid"]: d for d in srv.completed()}
    assert len(done) == 5 and all(d["status"] == 0 for d in done.values())
    # (whether or not the worker took `first` before the others arrived, the
    # order between the PPR and BFS groups is fixed by the policy)
    if max_wait_s == 0:
        assert done[ppr]["start_s"] < min(done[q]["start_s"] for q in bfs)
    else:
        assert max(done[q]["start_s"] for q in bfs) < done[ppr]["start_s"]
    srv.close()


def test_one_running_batch_per_group(gpu_ctx, graphs):
    """Per-(program, graph) mode: a group never has two batches on the GPU at
    once; queries arriving while its batch runs form its next batch."""
    g, roots = graphs[2]
    srv = L.Server(0, nworkers=3, max_batch=8, state_budget=1 << 30, keep_results=False, fused=False)
    gid = srv.add_graph(g)
    rng = np.random.default_rng(11)
    for i in range(48):
        kind = (L.Q_BFS, L.Q_PPR)[i % 2]
        srv.submit(kind, gid, int(rng.choice(roots)), 30 if kind == L.Q_PPR else 0)
    srv.drain()
    done = srv.completed()
    assert len(done) == 48 and all(d["status"] == 0 for d in done)
    for kind in (L.Q_BFS, L.Q_PPR):
        spans = sorted({(d["start_s"], d["done_s"]) for d in done if d["kind"] == kind})
        for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
            assert s1 >= e0, f"kind {kind}: batch [{s1}, {e1}] overlaps [{s0}, {e0}]"
    assert max(d["batch_size"] for d in done) > 1
    srv.close()
