"""The benchmark configurations at their own sizes, through the C ABI:
C2 (BFS + SSSP, directed R-MAT scale 22, seed 2), C3 (CC + TC on the
symmetrized scale-20 graph, seed 3), C4 (batched BFS / SSSP / PPR on scale
24, seed 4) and C5 (six resident graphs, a mixed served stream). Exact
programs are compared with the oracle bit for bit on sampled roots; batched,
fused and served results with the single-query programs for every sampled
query; PPR (FP32 + compensation) with the FP64 oracle within L1 1e-7. Graphs
and oracle instances are built once per module; each test stays well under
two minutes on one MI355X and its host."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu


def choose_roots(rec, count, seed):
    """Roots drawn by seed from the vertices with out-degree > 0 (as bench.py)."""
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(np.unique(rec[:, 0]), size=count, replace=False)]


def parallel(fn, jobs):
    with ThreadPoolExecutor(max_workers=8) as ex:  # oracle calls release the GIL
        return list(ex.map(fn, jobs))


# ------------------------------------------------------------------ C2 ---
@pytest.fixture(scope="module")
def c2(gpu_ctx, oracle):
    rec = L.rmat(22, 16, 0.57, 0.19, 0.19, 2, weighted=True)
    n = 1 << 22
    g = L.Graph.from_records(gpu_ctx, rec, n, True, L.DIRECTED)
    o = oracle.OracleGraph(rec, n, True, oracle.DIRECTED)
    return dict(rec=rec, n=n, g=g, o=o, roots=choose_roots(rec, 64, 2))


def test_c2_bfs_sssp_vs_oracle(c2, oracle):
    g, o = c2["g"], c2["o"]
    roots = c2["roots"][:8]
    exp = parallel(lambda r: (oracle.bfs(o, r), oracle.sssp(o, r)), roots)
    for r, ((eh, ep, eit), (ed, _)) in zip(roots, exp):
        hops, par, it = L.bfs(g, r)
        assert (hops == eh).all() and (par == ep).all() and it == eit, r
        assert (L.sssp(g, r)[0] == ed).all(), r


def test_c2_batched_and_fused_equal_single(c2, gpu_ctx):
    g, roots = c2["g"], c2["roots"]
    hb, _ = L.bfs_batch(g, roots)
    db, _ = L.sssp_batch(g, roots)
    for q, r in enumerate(roots):
        assert (hb[q] == L.bfs(g, r, parents=False)[0]).all(), r
        assert (db[q] == L.sssp(g, r)[0]).all(), r
    fr = L.batch_run(gpu_ctx, [(g, L.Q_BFS, roots), (g, L.Q_SSSP, roots)])
    assert (fr[0]["out"] == hb).all() and (fr[1]["out"] == db).all()


def test_c2_release(c2):
    c2.clear()


# ------------------------------------------------------------------ C3 ---
@pytest.fixture(scope="module")
def c3(gpu_ctx, oracle):
    rec = L.rmat(20, 16, 0.57, 0.19, 0.19, 3)
    n = 1 << 20
    return dict(rec=rec, n=n, gu=L.Graph.from_records(gpu_ctx, rec, n, False, L.UNDIRECTED),
                ga=L.Graph.from_records(gpu_ctx, rec, n, False, L.ACYCLIC))


def test_c3_cc_vs_oracle(c3, oracle):
    o = oracle.OracleGraph(c3["rec"], c3["n"], False, oracle.UNDIRECTED)
    elab, eit = oracle.cc(o)
    lab, it = L.cc(c3["gu"])
    assert (lab == elab).all() and it == eit
    assert (L.components(c3["gu"]) == elab).all()


def test_c3_tc_vs_oracle(c3, oracle):
    rec, n = c3["rec"], c3["n"]
    oa, ora = parallel(lambda m: oracle.OracleGraph(rec, n, False, m), [oracle.ACYCLIC, oracle.REVERSED_ACYCLIC])
    etotal, eper = oracle.tc(oa, ora)
    total, per = L.tc(c3["ga"])
    assert total == etotal and (per == eper).all()


def test_c3_release(c3):
    c3.clear()


# ------------------------------------------------------------------ C4 ---
@pytest.fixture(scope="module")
def c4(gpu_ctx):
    rec = L.rmat(24, 16, 0.57, 0.19, 0.19, 4, weighted=True)
    n = 1 << 24
    g = L.Graph.from_records(gpu_ctx, rec, n, True, L.DIRECTED)
    return dict(rec=rec, n=n, g=g, bfs=choose_roots(rec, 256, 41)[:8], sssp=choose_roots(rec, 256, 42)[:8],
                ppr=choose_roots(rec, 128, 43)[:4])


def test_c4_fused_equals_batched_equals_single(c4, gpu_ctx):
    g = c4["g"]
    hb, _ = L.bfs_batch(g, c4["bfs"])
    db, _ = L.sssp_batch(g, c4["sssp"])
    pb, _, _ = L.ppr_batch(g, c4["ppr"], 30)
    for q, r in enumerate(c4["bfs"]):
        assert (hb[q] == L.bfs(g, r, parents=False)[0]).all(), r
    for q, r in enumerate(c4["sssp"]):
        assert (db[q] == L.sssp(g, r)[0]).all(), r
    fr = L.batch_run(gpu_ctx, [(g, L.Q_BFS, c4["bfs"]), (g, L.Q_SSSP, c4["sssp"]), (g, L.Q_PPR, c4["ppr"], 30)])
    assert (fr[0]["out"] == hb).all() and (fr[1]["out"] == db).all() and (fr[2]["out"] == pb).all()
    for q, s in enumerate(c4["ppr"]):
        assert np.abs(pb[q] - L.ppr(g, s, 30)).sum() <= 1e-7, s


def test_c4_bfs_sssp_vs_oracle(c4, oracle):
    o = oracle.OracleGraph(c4["rec"], c4["n"], True, oracle.DIRECTED)
    c4["o"] = o
    g = c4["g"]
    jobs = [("bfs", r) for r in c4["bfs"][:4]] + [("sssp", r) for r in c4["sssp"][:4]]
    exp = parallel(lambda j: oracle.bfs(o, j[1]) if j[0] == "bfs" else oracle.sssp(o, j[1]), jobs)
    for (kind, r), e in zip(jobs, exp):
        if kind == "bfs":
            hops, par, it = L.bfs(g, r)
            assert (hops == e[0]).all() and (par == e[1]).all() and it == e[2], r
        else:
            assert (L.sssp(g, r)[0] == e[0]).all(), r


def test_c4_ppr_vs_fp64_oracle(c4, oracle):
    """FP32 + compensated batched PPR against the FP64 oracle on the full
    graph: 4 seeds at 3 iterations (the oracle's cost grows with the
    iterations); the 30-iteration runs are compared with the FP64 GPU program
    in the test above, itself pinned to the oracle on smaller graphs."""
    o = c4.get("o") or oracle.OracleGraph(c4["rec"], c4["n"], True, oracle.DIRECTED)
    orr = oracle.OracleGraph(c4["rec"], c4["n"], True, oracle.REVERSED)
    seeds = c4["ppr"]
    exp = parallel(lambda s: oracle.ppr(o, orr, s, 3)[0], seeds)
    pb, _, _ = L.ppr_batch(c4["g"], seeds, 3)
    for q in range(len(seeds)):
        assert np.abs(pb[q].astype(np.float64) - exp[q]).sum() <= 1e-7, seeds[q]


def test_c4_release(c4):
    c4.clear()


# ------------------------------------------------------------------ C5 ---
def test_c5_served_mix_equals_alone(gpu_ctx):
    """Six resident graphs (scales 18-22 and 24, seeds 10-15); a seeded mixed
    stream (50% BFS, 30% SSSP, 20% PPR of 30 iterations; graphs Zipf(1.1)
    by size rank) served with continuous batching; each query's result equals
    the same query run alone."""
    scales = [18, 19, 20, 21, 22, 24]
    graphs = []
    for i, sc in enumerate(scales):
        rec = L.rmat(sc, 16, 0.57, 0.19, 0.19, 10 + i, weighted=True)
        graphs.append((L.Graph.from_records(gpu_ctx, rec, 1 << sc, True, L.DIRECTED), np.unique(rec[:, :2])))
        del rec
    rng = np.random.default_rng(2026)
    nq = 120
    zipf = 1.0 / np.arange(1, 7) ** 1.1
    gi = rng.choice(6, size=nq, p=zipf / zipf.sum())
    kinds = rng.choice(3, size=nq, p=[0.5, 0.3, 0.2])
    roots = [int(rng.choice(graphs[i][1])) for i in gi]
    srv = L.Server(0, nworkers=3, max_batch=64, state_budget=64 << 30, keep_results=True)
    gids = [srv.add_graph(g) for g, _ in graphs]
    qids = [srv.submit(int(kinds[j]), gids[gi[j]], roots[j], 30 if kinds[j] == L.Q_PPR else 0) for j in range(nq)]
    srv.drain()
    done = {d["qid"]: d for d in srv.completed()}
    assert len(done) == nq and all(d["status"] == 0 for d in done.values())
    assert max(d["batch_size"] for d in done.values()) > 1  # the stream was batched
    for j, qid in enumerate(qids):
        g = graphs[gi[j]][0]
        got = srv.result(qid, int(kinds[j]), g.nslots)
        if kinds[j] == L.Q_BFS:
            assert (got == L.bfs(g, roots[j], parents=False)[0]).all(), j
        elif kinds[j] == L.Q_SSSP:
            assert (got == L.sssp(g, roots[j])[0]).all(), j
        else:
            assert np.array_equal(got, L.ppr_batch(g, [roots[j]], 30)[0][0]), j
    srv.close()
