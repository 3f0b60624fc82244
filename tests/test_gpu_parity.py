"""HIP path vs the oracle on the same seeded inputs, through the C ABI.

Bar: bit-exact for degree, BFS levels and parents, SSSP distances, CC labels
and triangle counts; FP64 PageRank / PPR within L1 <= 1e-6 of the oracle
(the reference's own tolerance for PR) and, tighter, max-abs <= 1e-12, since
only summation grouping on rows longer than 64 entries differs.
"""
import numpy as np
import pytest

import la3_amd as L
from test_oracle import random_records

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-6


def _both(ctx, oracle, rec, n, weighted, mode):
    return (L.Graph.from_records(ctx, rec, n, weighted, mode),
            oracle.OracleGraph(rec, n, weighted, mode))


def _rmat(scale, seed, weighted=False):
    return L.rmat(scale, 16, seed=seed, weighted=weighted)


def _roots(rec, count, seed):
    rng = np.random.default_rng(seed)
    src = np.unique(rec[:, 0])
    return [int(x) for x in rng.choice(src, size=min(count, src.size), replace=False)]


# ------------------------------------------------------------------- g1 ---
def test_g1_all_programs(gpu_ctx, oracle, g1):
    rec, recw, n = g1["records"], g1["records_w"], 8
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    rank, deg, it = L.pagerank(g, 20)
    for k, v in g1["pr_20"].items():
        assert abs(rank[int(k)] - v) < 5e-10
    assert it == 20
    for k, v in g1["out_degree"].items():
        assert deg[int(k)] == v
    gu = L.Graph.from_records(gpu_ctx, rec, n, False, L.UNDIRECTED)
    hops, par, it = L.bfs(gu, 5)
    for k, v in g1["bfs_undirected_root5_hops"].items():
        assert hops[int(k)] == v
    for k, v in g1["bfs_undirected_root5_parent"].items():
        assert par[int(k)] == v
    assert it == 3
    lab, _ = L.cc(gu)
    for k, v in g1["cc_labels"].items():
        assert lab[int(k)] == v
    gw = L.Graph.from_records(gpu_ctx, recw, n, True, L.DIRECTED)
    dist, _ = L.sssp(gw, 5)
    assert int(dist.astype(np.int64).sum()) == g1["sssp_checksum"]
    for k, v in g1["sssp_root5"].items():
        assert dist[int(k)] == v
    ga = L.Graph.from_records(gpu_ctx, rec, n, False, L.ACYCLIC)
    total, per = L.tc(ga)
    assert total == g1["tc_total"]
    oa = oracle.OracleGraph(rec, n, False, oracle.ACYCLIC)
    ora = oracle.OracleGraph(rec, n, False, oracle.REVERSED_ACYCLIC)
    assert (per == oracle.tc(oa, ora)[1]).all()
    gd = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    for k, v in g1["in_degree"].items():
        assert L.degree(gd)[int(k)] == v


# ------------------------------------------------------------ BFS / SSSP ---
@pytest.mark.parametrize("scale,seed", [(10, 1), (14, 2), (16, 3)])
@pytest.mark.parametrize("mode", [L.DIRECTED, L.UNDIRECTED])
def test_bfs_levels_and_parents(gpu_ctx, oracle, scale, seed, mode):
    rec = _rmat(scale, seed)
    n = 1 << scale
    g, o = _both(gpu_ctx, oracle, rec, n, False, mode)
    for root in _roots(rec, 4, seed) + [n, 0]:
        hops, par, it = L.bfs(g, root)
        eh, ep, eit = oracle.bfs(o, root)
        assert (hops == eh).all(), f"levels differ, root {root}"
        assert (par == ep).all(), f"parents differ, root {root}"
        assert it == eit


def test_bfs_random_with_duplicates(gpu_ctx, oracle):
    for seed in (21, 22):
        rec = random_records(seed, 3000, 20000)
        for mode in (L.DIRECTED, L.UNDIRECTED, L.REVERSED):
            g, o = _both(gpu_ctx, oracle, rec, 3000, False, mode)
            for root in (1, 2, 3, 1500, 2999):
                hops, par, it = L.bfs(g, root)
                eh, ep, eit = oracle.bfs(o, root)
                assert (hops == eh).all() and (par == ep).all() and it == eit


@pytest.mark.parametrize("scale,seed", [(10, 4), (14, 5), (16, 6)])
def test_sssp_distances(gpu_ctx, oracle, scale, seed):
    rec = _rmat(scale, seed, weighted=True)
    n = 1 << scale
    g, o = _both(gpu_ctx, oracle, rec, n, True, L.DIRECTED)
    for root in _roots(rec, 4, seed) + [n]:
        d, _ = L.sssp(g, root)
        e, _ = oracle.sssp(o, root)
        assert (d == e).all(), f"distances differ, root {root}"


def test_sssp_random_with_duplicates(gpu_ctx, oracle):
    rec = random_records(23, 3000, 20000, weighted=True)
    g, o = _both(gpu_ctx, oracle, rec, 3000, True, L.DIRECTED)
    for root in (1, 2, 3, 100, 2999):
        assert (L.sssp(g, root)[0] == oracle.sssp(o, root)[0]).all()


# -------------------------------------------------------------- CC / TC ---
@pytest.mark.parametrize("scale,seed", [(10, 7), (16, 8)])
def test_cc_labels(gpu_ctx, oracle, scale, seed):
    rec = _rmat(scale, seed)
    n = 1 << scale
    g, o = _both(gpu_ctx, oracle, rec, n, False, L.UNDIRECTED)
    lab, it = L.cc(g)
    elab, eit = oracle.cc(o)
    assert (lab == elab).all() and it == eit
    assert (L.components(g) == elab).all()


def test_cc_many_components(gpu_ctx, oracle):
    """Sparse graphs with no giant component (many small trees, the sampled
    'largest tree' holds few vertices) and a graph of two equal halves: the
    rows outside the sampled tree carry most of the unions."""
    rng = np.random.default_rng(9)
    n = 40000
    a = rng.integers(1, n + 1, 15000)
    b = rng.integers(1, n + 1, 15000)
    sparse = np.stack([a, b], 1).astype(np.uint32)
    half = n // 2
    c1 = rng.integers(1, half + 1, (200000, 2))
    c2 = rng.integers(half + 1, n + 1, (200000, 2))
    halves = np.concatenate([c1, c2]).astype(np.uint32)
    for rec in (sparse, halves, random_records(24, 3000, 1500)):
        nn = int(rec.max())
        g, o = _both(gpu_ctx, oracle, rec, nn, False, L.UNDIRECTED)
        lab, it = L.cc(g)
        elab, eit = oracle.cc(o)
        assert (lab == elab).all() and it == eit
        assert (L.components(g) == elab).all()


@pytest.mark.parametrize("scale,seed", [(10, 9), (14, 10)])
def test_tc_counts(gpu_ctx, oracle, scale, seed):
    rec = _rmat(scale, seed)
    n = 1 << scale
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.ACYCLIC)
    oa = oracle.OracleGraph(rec, n, False, oracle.ACYCLIC)
    ora = oracle.OracleGraph(rec, n, False, oracle.REVERSED_ACYCLIC)
    total, per = L.tc(g)
    etotal, eper = oracle.tc(oa, ora)
    assert total == etotal
    assert (per == eper).all()


@pytest.mark.parametrize("n", [2200, 4300])
def test_tc_dense_lists_closed_form(gpu_ctx, n):
    """Complete graph on ids 1..n: every pair of long neighbour lists goes
    through the LDS hash passes of both owner bands, and (n = 4300: lists of
    more than 4096 entries) through the wavefront intersections. Closed form: entry (u, v), v < u, adds v - 1 to v, so
    vertex v counts (n - v)(v - 1), and the total is C(n, 3)."""
    ii, jj = np.triu_indices(n, 1)
    rec = np.stack([ii + 1, jj + 1], axis=1).astype(np.uint32)
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.ACYCLIC)
    total, per = L.tc(g)
    assert total == n * (n - 1) * (n - 2) // 6
    v = np.arange(n + 1, dtype=np.int64)
    want = np.where(v >= 1, (n - v) * (v - 1), 0)
    assert (per.astype(np.int64) == want).all()


# ---------------------------------------------------------- PR / degree ---
def test_pagerank_c1_config(gpu_ctx, oracle):
    """PageRank, 20 iterations, on the scale-16 seed-1 R-MAT graph."""
    rec = _rmat(16, 1)
    n = 1 << 16
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    o = oracle.OracleGraph(rec, n, False, oracle.DIRECTED)
    orr = oracle.OracleGraph(rec, n, False, oracle.REVERSED)
    rank, deg, it = L.pagerank(g, 20)
    erank, edeg, eit = oracle.pagerank(o, orr, 20)
    l1 = float(np.abs(rank - erank).sum())
    print(f"PR C1 L1 vs oracle = {l1:.3e}, max abs = {np.abs(rank - erank).max():.3e}")
    assert l1 <= PR_L1_TOL
    assert np.abs(rank - erank).max() <= 1e-12
    assert (deg == edeg).all() and it == eit == 20


@pytest.mark.parametrize("seed", [11, 12])
def test_pagerank_until_convergence(gpu_ctx, oracle, seed):
    rec = _rmat(12, seed)
    n = 1 << 12
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    rank, _, it = L.pagerank(g, 0)
    erank, _, eit = oracle.pagerank(oracle.OracleGraph(rec, n, False, 0),
                                    oracle.OracleGraph(rec, n, False, 1), 0)
    assert it == eit
    assert float(np.abs(rank - erank).sum()) <= PR_L1_TOL


def test_ppr_fp64(gpu_ctx, oracle):
    rec = _rmat(12, 13)
    n = 1 << 12
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    o = oracle.OracleGraph(rec, n, False, 0)
    orr = oracle.OracleGraph(rec, n, False, 1)
    for s in _roots(rec, 3, 13):
        r = L.ppr(g, s, 30)
        e, _ = oracle.ppr(o, orr, s, 30)
        assert float(np.abs(r - e).sum()) <= 1e-12


@pytest.mark.parametrize("mode", [L.DIRECTED, L.REVERSED, L.UNDIRECTED])
def test_degree(gpu_ctx, oracle, mode):
    rec = _rmat(14, 14)
    n = 1 << 14
    g, o = _both(gpu_ctx, oracle, rec, n, False, mode)
    assert (L.degree(g) == oracle.degree(o)).all()


# --------------------------------------------------------------- edges ---
def test_edge_cases(gpu_ctx, oracle):
    # empty graph: every program runs, root is the only reached vertex
    e = np.zeros((0, 2), dtype=np.uint32)
    g = L.Graph.from_records(gpu_ctx, e, 4, False, L.DIRECTED)
    hops, _, it = L.bfs(g, 2)
    assert hops[2] == 0 and (np.delete(hops, 2) == 255).all() and it == 1
    rank, _, _ = L.pagerank(g, 5)
    assert np.allclose(rank, 0.15)
    # root out of range is an error, not a fault
    with pytest.raises(L.LA3Error):
        L.bfs(g, 5)
    # wrong load mode for the program
    with pytest.raises(L.LA3Error):
        L.components(g)
    with pytest.raises(L.LA3Error):
        L.tc(g)
    # sink / source / isolated roots
    rec = np.array([[1, 2], [2, 3], [3, 1], [4, 1], [2, 5]], dtype=np.uint32)
    g, o = _both(gpu_ctx, oracle, rec, 6, False, L.DIRECTED)
    for root in range(7):
        hops, par, it = L.bfs(g, root)
        eh, ep, eit = oracle.bfs(o, root)
        assert (hops == eh).all() and (par == ep).all() and it == eit


# -------------------------------------------------------------- batched ---
@pytest.mark.parametrize("scale,seed,k", [(10, 31, 1), (12, 32, 64), (14, 33, 100), (16, 34, 256)])
def test_bfs_batch_equals_single(gpu_ctx, oracle, scale, seed, k):
    rec = _rmat(scale, seed)
    n = 1 << scale
    g, o = _both(gpu_ctx, oracle, rec, n, False, L.DIRECTED)
    roots = _roots(rec, k, seed)
    if k > 2:
        roots[1] = roots[0]          # repeated root
        roots[2] = n                  # likely isolated / sink
    hops, it = L.bfs_batch(g, roots)
    its = []
    for q, r in enumerate(roots):
        h1, _, it1 = L.bfs(g, r, parents=False)
        assert (hops[q] == h1).all(), f"query {q} (root {r}) differs from single-query"
        its.append(it1)
    assert it == max(its)
    for q in range(0, len(roots), max(1, len(roots) // 4)):
        assert (hops[q] == oracle.bfs(o, roots[q])[0]).all()


def test_bfs_batch_undirected(gpu_ctx, oracle):
    rec = _rmat(14, 35)
    n = 1 << 14
    g, o = _both(gpu_ctx, oracle, rec, n, False, L.UNDIRECTED)
    roots = _roots(rec, 70, 35)
    hops, _ = L.bfs_batch(g, roots)
    for q in (0, 33, 69):
        assert (hops[q] == oracle.bfs(o, roots[q])[0]).all()


@pytest.mark.parametrize("scale,seed,k", [(10, 41, 1), (12, 42, 64), (14, 43, 100), (16, 44, 128),
                                          (13, 45, 256), (12, 46, 1024)])
def test_sssp_batch_equals_single(gpu_ctx, oracle, scale, seed, k):
    rec = _rmat(scale, seed, weighted=True)
    n = 1 << scale
    g, o = _both(gpu_ctx, oracle, rec, n, True, L.DIRECTED)
    roots = _roots(rec, k, seed)
    if k > 2:
        roots[1] = roots[0]
        roots[2] = n
    dist, _ = L.sssp_batch(g, roots)
    for q, r in enumerate(roots):
        d1, _ = L.sssp(g, r)
        assert (dist[q] == d1).all(), f"query {q} (root {r}) differs from single-query"
    for q in range(0, len(roots), max(1, len(roots) // 4)):
        assert (dist[q] == oracle.sssp(o, roots[q])[0]).all()


@pytest.mark.parametrize("delta", [None, 1, 0])
def test_sssp_batch_narrow_rows_equal_wide(gpu_ctx, oracle, delta):
    """Batches of <= 32 queries run on narrow distance rows (1, 2, 4, 8, 16
    or 32 lanes per row, several column entries per wavefront); every query's
    distances equal the same query in a 64-lane batch and the oracle, with
    the near/far split on (delta 1: nearly every pair waits as far), off
    (delta 0) or at the graph's default."""
    rec = _rmat(15, 57, weighted=True)
    n = 1 << 15
    g, o = _both(gpu_ctx, oracle, rec, n, True, L.DIRECTED)
    if delta is not None:
        g.set_tuning(L.TUNE_BATCH_SSSP_DELTA, delta)
    roots = _roots(rec, 80, 57)
    roots[1] = roots[0]
    roots[2] = n
    wide, _ = L.sssp_batch(g, roots)
    for k in (1, 2, 3, 4, 7, 8, 9, 16, 25, 32):
        narrow, _ = L.sssp_batch(g, roots[:k])
        assert (narrow == wide[:k]).all(), f"k={k}: distances differ from the 64-lane batch"
    for q in (0, 3, 31):
        assert (wide[q] == oracle.sssp(o, roots[q])[0]).all()


# ---------------------------------------------------------- batched PPR ---
PPR_L1_TOL = 1e-7


@pytest.mark.parametrize("scale,seed,k", [(12, 51, 5), (13, 55, 20), (14, 52, 64), (16, 53, 128)])
def test_ppr_batch_fp32_vs_fp64_oracle(gpu_ctx, oracle, scale, seed, k):
    """FP32 + compensation within L1 <= 1e-7 of the FP64 oracle, per query."""
    rec = _rmat(scale, seed)
    n = 1 << scale
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    o = oracle.OracleGraph(rec, n, False, oracle.DIRECTED)
    orr = oracle.OracleGraph(rec, n, False, oracle.REVERSED)
    seeds = _roots(rec, k, seed)
    if k > 2:
        seeds[1] = seeds[0]
    rank, comp, resid = L.ppr_batch(g, seeds, 30, comp=True)
    worst = 0.0
    for q in range(0, k, max(1, k // 6)):
        e, _ = oracle.ppr(o, orr, seeds[q], 30)
        l1 = float(np.abs(rank[q].astype(np.float64) - e).sum())
        l1c = float(np.abs(rank[q].astype(np.float64) + comp[q] - e).sum())
        nz = e > 0
        mrel = float((np.abs(rank[q] - e)[nz] / e[nz]).max())
        worst = max(worst, l1)
        assert l1 <= PPR_L1_TOL, f"query {q}: L1 {l1:.3e}"
        assert l1c <= l1 + 1e-12
        assert mrel < 1e-5
        assert np.isfinite(resid[q]) and resid[q] >= 0
    print(f"PPR batch k={k} scale {scale}: worst L1 vs FP64 {worst:.3e}")
    # a query in a batch equals the same query alone
    r1, _, _ = L.ppr_batch(g, [seeds[3]], 30)
    assert (r1[0] == rank[3]).all()


def test_ppr_batch_width_does_not_change_ranks(gpu_ctx):
    """Batches of <= 32 queries run on narrow rows (1, 2, 4, 8, 16 or 32
    lanes per row, several rows per wavefront); every query's ranks and compensation
    terms are bit-identical to the same query in a 64-lane batch."""
    rec = _rmat(15, 56)
    n = 1 << 15
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    seeds = _roots(rec, 80, 56)
    wide, wcomp, _ = L.ppr_batch(g, seeds, 12, comp=True)
    for k in (1, 2, 3, 4, 7, 8, 9, 16, 25, 32):
        narrow, ncomp, resid = L.ppr_batch(g, seeds[:k], 12, comp=True)
        assert (narrow == wide[:k]).all(), f"k={k}: ranks differ from the 64-lane batch"
        assert (ncomp == wcomp[:k]).all(), f"k={k}: compensation differs"
        assert np.isfinite(resid).all() and (resid >= 0).all()


def test_ppr_batch_matches_fp64_single(gpu_ctx):
    rec = _rmat(14, 54)
    n = 1 << 14
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    seeds = _roots(rec, 3, 54)
    rank, _, _ = L.ppr_batch(g, seeds, 20)
    for q, s in enumerate(seeds):
        r64 = L.ppr(g, s, 20)
        assert float(np.abs(rank[q] - r64).sum()) <= PPR_L1_TOL


@pytest.mark.parametrize("delta", [0, 1, 32, 128, 1000])
def test_sssp_near_far_exact_for_any_delta(gpu_ctx, oracle, delta):
    rec = _rmat(14, 71, weighted=True)
    n = 1 << 14
    g, o = _both(gpu_ctx, oracle, rec, n, True, L.DIRECTED)
    g.set_tuning(L.TUNE_SSSP_DELTA, delta)
    g.set_tuning(L.TUNE_BATCH_SSSP_DELTA, delta)
    assert g.get_tuning(L.TUNE_SSSP_DELTA) == delta
    roots = _roots(rec, 70, 71)
    db, _ = L.sssp_batch(g, roots)
    for q in (0, 1, 35, 69):
        e, _ = oracle.sssp(o, roots[q])
        assert (L.sssp(g, roots[q])[0] == e).all()
        assert (db[q] == e).all()


def test_autotune_keeps_results_exact(gpu_ctx, oracle):
    rec = _rmat(13, 72, weighted=True)
    n = 1 << 13
    g, o = _both(gpu_ctx, oracle, rec, n, True, L.DIRECTED)
    roots = _roots(rec, 64, 72)
    rep = g.autotune(roots)
    assert rep["sssp_delta"] in rep["candidates"] and rep["batch_sssp_delta"] in rep["candidates"]
    assert g.get_tuning(L.TUNE_SSSP_DELTA) == rep["sssp_delta"]
    db, _ = L.sssp_batch(g, roots)
    for q in (0, 63):
        assert (db[q] == oracle.sssp(o, roots[q])[0]).all()
