"""Pins the oracle's engine on programs written only against the
vertex-program interface (oracle/la3_oracle.cpp: WDeg, MaxLabel, Level,
PrAlpha, CountGreater, CountNonzero) with closed forms computed from the records, before
those programs check the device VertexProgram<P> template."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

import closed_forms as CF
from test_oracle import random_records, _message_order_keys

N = 300


def test_wdeg_is_weighted_in_degree(oracle):
    rec = random_records(90, N, 2000, weighted=True)
    g = oracle.OracleGraph(rec, N, True, oracle.DIRECTED)
    d = CF.dedup_first(rec)
    exp = np.zeros(N + 1, dtype=np.int64)
    np.add.at(exp, d[:, 1].astype(np.int64), d[:, 2].astype(np.int64))
    for m in (1, 2, 3):  # stationary: every bound gives the same sums
        w, it = oracle.vp_wdeg(g, m)
        assert (w == exp).all() and it == m


def _max_ancestor(rec, n, undirected):
    A = CF.adjacency(rec, n, undirected=undirected)
    reach = np.isfinite(csgraph.shortest_path(A, directed=not undirected, unweighted=True))
    ids = np.arange(n + 1)
    # label(v) = largest id among v and the vertices that reach v
    return np.where(reach, ids[:, None], -1).max(axis=0).astype(np.uint32)


@pytest.mark.parametrize("seed", [91, 92])
def test_maxlabel_is_max_ancestor(oracle, seed):
    rec = random_records(seed, N, 700)
    for mode, und in ((oracle.UNDIRECTED, True), (oracle.DIRECTED, False)):
        g = oracle.OracleGraph(rec, N, False, mode)
        lab, _ = oracle.vp_maxlabel(g)
        assert (lab == _max_ancestor(rec, N, und)).all(), mode
    # bounded, undirected: the largest id within m hops
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED)
    A = CF.adjacency(rec, N, undirected=True)
    d = csgraph.shortest_path(A, directed=False, unweighted=True)
    for m in (1, 2, 3):
        exp = np.where(d <= m, np.arange(N + 1)[:, None], -1).max(axis=0)
        assert (oracle.vp_maxlabel(g, m)[0] == exp).all()


@pytest.mark.parametrize("seed", [93, 94])
def test_level_is_bfs_with_last_writer(oracle, seed):
    n = 5000
    rec = random_records(seed, n, 30000)
    g = oracle.OracleGraph(rec, n, False, oracle.DIRECTED)
    key, hashed = _message_order_keys(rec, n)
    d = CF.dedup_first(rec)
    for root in (1, 7):
        lev, via, _ = oracle.vp_level(g, root)
        hops = CF.bfs_levels(rec, n, root)
        exp = np.where(hops == 255, 0xFFFFFFFF, hops.astype(np.int64)).astype(np.uint32)
        assert (lev == exp).all()
        lv = hops.astype(np.int64)
        a, b = d[:, 0].astype(np.int64), d[:, 1].astype(np.int64)
        ok = (lv[a] != 255) & (lv[b] == lv[a] + 1)
        best = np.full(n + 1, -1, dtype=np.int64)
        np.maximum.at(best, b[ok], key[a[ok]])
        inv = np.full(key.max() + 1, -1, dtype=np.int64)
        inv[key[key >= 0]] = np.flatnonzero(key >= 0)
        reached = (hops != 255) & (hops != 0)
        assert (via[reached] == hashed[inv[best[reached]]]).all()


def test_pralpha_is_dense_pagerank(oracle):
    rec = random_records(95, N, 2500)
    g = oracle.OracleGraph(rec, N, False, oracle.DIRECTED)
    gr = oracle.OracleGraph(rec, N, False, oracle.REVERSED)
    for iters in (1, 2, 5, 20):
        rank, it = oracle.vp_pralpha(g, gr, iters)
        assert np.abs(rank - CF.pagerank_dense(rec, N, iters, alpha=0.2)).max() < 1e-14 and it == iters


@pytest.mark.parametrize("mode", ["directed", "undirected"])
def test_countgreater_reads_the_init_mirror(oracle, mode):
    rec = random_records(96, N, 1500)
    und = mode == "undirected"
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED if und else oracle.DIRECTED)
    d = CF.dedup_first(rec)
    a, b = d[:, 0].astype(np.int64), d[:, 1].astype(np.int64)
    if und:
        a, b = np.concatenate([a, b]), np.concatenate([b, a])
        key = np.unique(a * (N + 1) + b)
        a, b = key // (N + 1), key % (N + 1)
    init = np.arange(N + 1) % 7
    has_in = np.bincount(b, minlength=N + 1) > 0
    has_out = np.bincount(a, minlength=N + 1) > 0
    source = has_out & ~has_in
    y0 = np.zeros(N + 1, dtype=np.int64)
    np.add.at(y0, b, (init[a] > init[b]).astype(np.int64))
    val, cnt, it = oracle.vp_countgreater(g, 1)
    assert it == 1 and (cnt == y0).all() and (val == np.where(has_in, y0 % 7, init)).all()
    # second iteration: sources resend their initial value, the others the
    # value apply left; each row compares with its mirror, the initial value
    msg1 = np.where(source, init, np.where(has_in, y0 % 7, init))
    y1 = np.zeros(N + 1, dtype=np.int64)
    np.add.at(y1, b, (msg1[a] > init[b]).astype(np.int64))
    val, cnt, it = oracle.vp_countgreater(g, 2)
    assert it == 2 and (cnt == y0 + y1).all() and (val == np.where(has_in, y1 % 7, init)).all()


@pytest.mark.parametrize("mode", ["directed", "undirected"])
def test_countnonzero_reads_default_mirrors(oracle, mode):
    """A program that is not stationary has its mirrors left at S(): val 0,
    so each row counts the messages above 0 (vertex_program.hpp:159-173).
    The gather compares with the mirror; a mirror holding the row's own
    initial value (vid % 7) would give other counts."""
    rec = random_records(97, N, 1500)
    und = mode == "undirected"
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED if und else oracle.DIRECTED)
    d = CF.dedup_first(rec)
    a, b = d[:, 0].astype(np.int64), d[:, 1].astype(np.int64)
    if und:
        a, b = np.concatenate([a, b]), np.concatenate([b, a])
        key = np.unique(a * (N + 1) + b)
        a, b = key // (N + 1), key % (N + 1)
    ids = np.arange(N + 1)
    init = ids % 7
    has_in = np.bincount(b, minlength=N + 1) > 0
    has_out = np.bincount(a, minlength=N + 1) > 0
    source = has_out & ~has_in
    sink = has_in & ~has_out
    act0 = ids % 3 == 0
    e0 = act0[a]
    y0 = np.zeros(N + 1, dtype=np.int64)
    np.add.at(y0, b[e0], (init[a[e0]] > 0).astype(np.int64))
    cnt, it = oracle.vp_countnonzero(g, 1)
    assert it == 1 and (cnt == y0).all()
    # iteration 1 of execute_non_opt: regular senders whose apply returned
    # true (y0 > 0); source messages reach sink rows only
    activated = y0 > 0
    e1 = (activated[a] & has_in[a] & has_out[a]) | (act0[a] & source[a] & sink[b])
    y1 = np.zeros(N + 1, dtype=np.int64)
    np.add.at(y1, b[e1], (init[a[e1]] > 0).astype(np.int64))
    cnt, it = oracle.vp_countnonzero(g, 2)
    assert it == 2 and (cnt == y0 + y1).all()


def test_maxlabel_second_execute_continues(oracle):
    """execute(m1) then execute(m2) without reset continues from the
    messages the first run left (the reference initializes once,
    vertex_program_execute.hpp:15-16): on an undirected load, the largest id
    within m1 + m2 hops."""
    rec = random_records(98, N, 700)
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED)
    A = CF.adjacency(rec, N, undirected=True)
    d = csgraph.shortest_path(A, directed=False, unweighted=True)
    for m1, m2 in ((2, 1), (2, 2), (3, 1)):
        lab, it = oracle.vp_maxlabel_seq(g, m1, m2)
        exp = np.where(d <= m1 + m2, np.arange(N + 1)[:, None], -1).max(axis=0)
        assert (lab == exp).all() and it == m2, (m1, m2)
    # execute(1) (execute_single) applies without scattering, and SpMV reads
    # messages without consuming them (streaming_array.h:104-108): the next
    # run re-sends the initial labels, which change nothing
    for m2 in (1, 2):
        lab, it = oracle.vp_maxlabel_seq(g, 1, m2)
        exp = np.where(d <= 1, np.arange(N + 1)[:, None], -1).max(axis=0)
        assert (lab == exp).all() and it == m2, m2


def test_maxlabel_reset_then_initialize_vids(oracle):
    """reset() keeps the states and initialize(vids) runs init on the listed
    vertices only (vertex_program.hpp:51-63, 177-259): after convergence
    (component maxima), re-initialized vertices hold their own ids, nobody
    else sends, so a listed vertex ends with the largest id of its component
    among the listed vertices it reaches through listed vertices, and every
    other vertex keeps its component maximum."""
    rec = random_records(99, N, 600)
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED)
    A = CF.adjacency(rec, N, undirected=True)
    ids = np.arange(N + 1)
    comp = _max_ancestor(rec, N, True).astype(np.int64)
    vids = ids[(ids % 3 == 0) | (ids % 7 == 1)]
    listed = np.zeros(N + 1, bool)
    listed[vids] = True
    sub = A[listed][:, listed]
    ncomp, lab_sub = csgraph.connected_components(sub, directed=False)
    best = np.zeros(ncomp, dtype=np.int64)
    np.maximum.at(best, lab_sub, vids)
    exp = comp.copy()
    exp[vids] = best[lab_sub]
    lab, _ = oracle.vp_maxlabel_seq(g, 0, 0, reset=True, vids=vids)
    assert (lab == exp).all()
    # initialize() of every vertex after reset: back to the fixed point
    lab, _ = oracle.vp_maxlabel_seq(g, 0, 0, reset=True)
    assert (lab == comp).all()


def test_accgreater_mirrors_keep_the_first_initialize(oracle):
    """A stationary state-reading program whose init adds to the state it
    finds, run as initialize; execute(1); reset; initialize; execute(1): the
    second run compares with the mirrors of the first initialize
    (vertex_program.hpp:159-173 mirror once), not with the new states."""
    rec = random_records(100, N, 1800)
    g = oracle.OracleGraph(rec, N, False, oracle.DIRECTED)
    dd = CF.dedup_first(rec)
    a, b = dd[:, 0].astype(np.int64), dd[:, 1].astype(np.int64)
    ids = np.arange(N + 1)
    init = ids % 5 + 1
    has_in = np.bincount(b, minlength=N + 1) > 0
    y1 = np.zeros(N + 1, dtype=np.int64)
    np.add.at(y1, b, (init[a] > init[b]).astype(np.int64))
    val1 = np.where(has_in, y1 % 7, init)
    val2 = val1 + init
    y2 = np.zeros(N + 1, dtype=np.int64)
    np.add.at(y2, b, (val2[a] > init[b]).astype(np.int64))
    val, cnt = oracle.vp_accgreater_twice(g)
    assert (cnt == np.where(has_in, y1 + y2, 0)).all()
    assert (val == np.where(has_in, y2 % 7, val2)).all()
