"""Pins the CPU oracle (oracle/la3_oracle.cpp) before it is trusted as the
checker: g1 known answers, and textbook / library answers on random graphs
that carry duplicates (with different weights) and self-loops."""
import numpy as np
import pytest

import closed_forms as CF

N_RANDOM = 300


def random_records(seed, n, m, weighted=False, zero_id=False):
    rng = np.random.default_rng(seed)
    lo = 0 if zero_id else 1
    a = rng.integers(lo, n + 1, m)
    # skew: a third of the edges touch a few hubs
    hub = rng.random(m) < 0.3
    a[hub] = rng.integers(lo, lo + 5, hub.sum())
    b = rng.integers(lo, n + 1, m)
    rec = np.stack([a, b], 1)
    dup = rec[rng.integers(0, m, m // 10)]
    loops = np.stack([np.arange(lo, lo + 5)] * 2, 1)
    rec = np.concatenate([rec, dup, loops]).astype(np.uint32)
    rec = rec[rng.permutation(len(rec))]
    if weighted:
        w = rng.integers(1, 129, len(rec)).astype(np.uint32)
        rec = np.concatenate([rec, w[:, None]], 1)
    return np.ascontiguousarray(rec)


# ------------------------------------------------------------ g1 KATs ---
def test_g1_classes(g1, oracle):
    g = oracle.OracleGraph(g1["records"], g1["n"], False, oracle.DIRECTED)
    hid = g.hid_of()
    cls = g1["classes_directed"]
    exp = cls["regular"] + cls["sink"] + cls["source"] + cls["isolated"]
    assert (g.nreg, g.nsink, g.nsrc) == (4, 1, 2)
    # n = 8 < 128 buckets: the hash is the identity, so slots list ids in class order
    assert hid.tolist() == exp


def test_g1_degree(g1, oracle):
    g = oracle.OracleGraph(g1["records"], 8, False, oracle.DIRECTED)
    gr = oracle.OracleGraph(g1["records"], 8, False, oracle.REVERSED)
    ind, outd = oracle.degree(g), oracle.degree(gr)
    for k, v in g1["in_degree"].items():
        assert ind[int(k)] == v
    for k, v in g1["out_degree"].items():
        assert outd[int(k)] == v
    assert outd.sum() == ind.sum() == g1["degree_checksum"]


@pytest.mark.parametrize("iters", [20, 0])
def test_g1_pagerank(g1, oracle, iters):
    g = oracle.OracleGraph(g1["records"], 8, False, oracle.DIRECTED)
    gr = oracle.OracleGraph(g1["records"], 8, False, oracle.REVERSED)
    rank, deg, it = oracle.pagerank(g, gr, iters)
    assert it == 20
    for k, v in g1["pr_20"].items():
        assert abs(rank[int(k)] - v) < 5e-10
    assert f"{rank.sum():f}" == f"{g1['pr_checksum']:f}"


def test_g1_sssp(g1, oracle):
    g = oracle.OracleGraph(g1["records_w"], 8, True, oracle.DIRECTED)
    d, _ = oracle.sssp(g, 5)
    exp = np.full(9, oracle.SSSP_INF, dtype=np.uint32)
    for k, v in g1["sssp_root5"].items():
        exp[int(k)] = v
    assert (d == exp).all()
    assert int(d.astype(np.int64).sum()) == g1["sssp_checksum"]


def test_g1_bfs_undirected(g1, oracle):
    g = oracle.OracleGraph(g1["records"], 8, False, oracle.UNDIRECTED)
    hops, parent, it = oracle.bfs(g, 5)
    for k, v in g1["bfs_undirected_root5_hops"].items():
        assert hops[int(k)] == v
    for k, v in g1["bfs_undirected_root5_parent"].items():
        assert parent[int(k)] == v
    assert int((hops != 255).sum()) == g1["bfs_reachable"]
    assert int((hops.astype(np.int64) * parent).sum()) == g1["bfs_checksum"]
    assert it == 3


def test_g1_cc(g1, oracle):
    g = oracle.OracleGraph(g1["records"], 8, False, oracle.UNDIRECTED)
    lab, _ = oracle.cc(g)
    for k, v in g1["cc_labels"].items():
        assert lab[int(k)] == v
    assert int(lab.sum()) == g1["cc_checksum"]


def test_g1_tc(g1, oracle):
    ga = oracle.OracleGraph(g1["records"], 8, False, oracle.ACYCLIC)
    gra = oracle.OracleGraph(g1["records"], 8, False, oracle.REVERSED_ACYCLIC)
    total, per = oracle.tc(ga, gra)
    assert total == g1["tc_total"]
    _, exp_per = CF.triangles(g1["records"], 8)
    assert (per == exp_per).all()


# --------------------------------------------- random graphs vs closed forms ---
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bfs_matches_textbook(oracle, seed):
    rec = random_records(seed, N_RANDOM, 900)
    for mode, und in ((oracle.DIRECTED, False), (oracle.UNDIRECTED, True)):
        g = oracle.OracleGraph(rec, N_RANDOM, False, mode)
        cls = CF.classes(rec, N_RANDOM)
        regular = cls["regular"] if not und else ~cls["isolated"]
        for root in (1, 2, 17, 150):
            hops, parent, it = oracle.bfs(g, root)
            exp = CF.bfs_levels(rec, N_RANDOM, root, undirected=und)
            assert (hops == exp).all()
            # The loop stops after the first iteration that activates no regular
            # vertex (sinks never count): 1 + the deepest regular level.
            reg_reached = regular & (hops != 255)
            assert it == 1 + (int(hops[reg_reached].max()) if reg_reached.any() else 0)


def _message_order_keys(rec, n):
    """Reference per-row message order of each sender (sources first, then
    regular, each by increasing hashed id), computed from classes and the
    bucket construction of the hash alone (no oracle code)."""
    cls = CF.classes(rec, n)
    hashed = CF.bucket_hash_table(n)
    key = np.full(n + 1, -1, dtype=np.int64)
    src = np.flatnonzero(cls["source"])
    reg = np.flatnonzero(cls["regular"])
    src = src[np.argsort(hashed[src])]
    reg = reg[np.argsort(hashed[reg])]
    key[src] = np.arange(src.size)
    key[reg] = src.size + np.arange(reg.size)
    return key, hashed


@pytest.mark.parametrize("seed", [4, 5])
def test_bfs_parent_is_last_writer(oracle, seed):
    n = 5000
    rec = random_records(seed, n, 30000)
    g = oracle.OracleGraph(rec, n, False, oracle.DIRECTED)
    key, hashed = _message_order_keys(rec, n)
    d = CF.dedup_first(rec)
    for root in (1, 3):
        hops, parent, _ = oracle.bfs(g, root)
        # For each reached vertex v (level L >= 1) the parent is the in-neighbour at level
        # L-1 with the largest message-order key; stored as its hashed id.
        lv = hops.astype(np.int64)
        a, b = d[:, 0].astype(np.int64), d[:, 1].astype(np.int64)
        ok = (lv[a] != 255) & (lv[b] == lv[a] + 1)
        best = np.full(n + 1, -1, dtype=np.int64)
        np.maximum.at(best, b[ok], key[a[ok]])
        inv = np.full(key.max() + 1, -1, dtype=np.int64)
        inv[key[key >= 0]] = np.flatnonzero(key >= 0)
        reached = (hops != 255) & (hops != 0)
        exp = hashed[inv[best[reached]]]
        assert (parent[reached] == exp).all()


@pytest.mark.parametrize("seed", [6, 7, 8])
def test_sssp_matches_dijkstra(oracle, seed):
    rec = random_records(seed, N_RANDOM, 1200, weighted=True)
    g = oracle.OracleGraph(rec, N_RANDOM, True, oracle.DIRECTED)
    for root in (1, 2, 3, 200):
        d, _ = oracle.sssp(g, root)
        assert (d == CF.dijkstra(rec, N_RANDOM, root)).all()


@pytest.mark.parametrize("seed", [9, 10])
def test_cc_matches_union_find(oracle, seed):
    rec = random_records(seed, 2000, 1500)
    g = oracle.OracleGraph(rec, 2000, False, oracle.UNDIRECTED)
    lab, _ = oracle.cc(g)
    assert (lab == CF.min_id_components(rec, 2000)).all()


@pytest.mark.parametrize("seed", [11, 12])
def test_tc_matches_brute_force(oracle, seed):
    rec = random_records(seed, N_RANDOM, 3000)
    ga = oracle.OracleGraph(rec, N_RANDOM, False, oracle.ACYCLIC)
    gra = oracle.OracleGraph(rec, N_RANDOM, False, oracle.REVERSED_ACYCLIC)
    total, per = oracle.tc(ga, gra)
    exp_total, exp_per = CF.triangles(rec, N_RANDOM)
    assert total == exp_total == CF.triangle_total_trace(rec, N_RANDOM)
    assert (per == exp_per).all()


@pytest.mark.parametrize("seed", [13, 14])
def test_degree_counts(oracle, seed):
    rec = random_records(seed, N_RANDOM, 2000)
    g = oracle.OracleGraph(rec, N_RANDOM, False, oracle.DIRECTED)
    gr = oracle.OracleGraph(rec, N_RANDOM, False, oracle.REVERSED)
    assert (oracle.degree(g) == CF.in_degree(rec, N_RANDOM)).all()
    assert (oracle.degree(gr) == CF.out_degree(rec, N_RANDOM)).all()


@pytest.mark.parametrize("seed", [15, 16])
def test_pagerank_matches_dense(oracle, seed):
    rec = random_records(seed, N_RANDOM, 2500)
    g = oracle.OracleGraph(rec, N_RANDOM, False, oracle.DIRECTED)
    gr = oracle.OracleGraph(rec, N_RANDOM, False, oracle.REVERSED)
    rank, deg, _ = oracle.pagerank(g, gr, 20)
    exp = CF.pagerank_dense(rec, N_RANDOM, 20)
    assert np.abs(rank - exp).max() < 1e-12
    assert (deg[CF.classes(rec, N_RANDOM)["regular"]] == CF.out_degree(rec, N_RANDOM)[
        CF.classes(rec, N_RANDOM)["regular"]]).all()
    # the sink step uses the final regular ranks: a variant that updates sinks inside
    # the loop differs (g1 shows the gap); here sinks must equal one step from the end.
    cls = CF.classes(rec, N_RANDOM)
    assert np.allclose(rank[cls["source"] | cls["isolated"]], 0.15)


@pytest.mark.parametrize("seed", [17])
def test_ppr_matches_dense(oracle, seed):
    rec = random_records(seed, N_RANDOM, 2500)
    g = oracle.OracleGraph(rec, N_RANDOM, False, oracle.DIRECTED)
    gr = oracle.OracleGraph(rec, N_RANDOM, False, oracle.REVERSED)
    for s in (1, 2, 50):
        rank, _ = oracle.ppr(g, gr, s, 30)
        exp = CF.pagerank_dense(rec, N_RANDOM, 30, seed=s)
        assert np.abs(rank - exp).max() < 1e-14


def test_oracle_rejects_out_of_range(oracle):
    with pytest.raises(ValueError):
        oracle.OracleGraph(np.array([[1, 9]], dtype=np.uint32), 8, False, oracle.DIRECTED)


def test_empty_graph(oracle):
    g = oracle.OracleGraph(np.zeros((0, 2), dtype=np.uint32), 4, False, oracle.DIRECTED)
    assert (g.nreg, g.nsink, g.nsrc) == (0, 0, 0)
    hops, _, it = oracle.bfs(g, 1)
    assert hops[1] == 0 and (np.delete(hops, 1) == 255).all()
