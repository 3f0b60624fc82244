"""Pins the oracle's execute(max_iters) (vertex_program_execute.hpp:11-52):
max_iters == 1 is execute_single (one step for regular and sink rows, no
re-scatter), m > 1 runs at most m loop iterations (plus, for optimizable
programs, the post-loop sink step) whether or not it converged earlier
(the loop condition is iter < max_iters), 0 runs until an iteration activates
no vertex. Answers come from hop-bounded closed forms (closed_forms.py)."""
import numpy as np
import pytest

import closed_forms as CF
from test_oracle import random_records

N = 400


@pytest.mark.parametrize("seed", [21, 22])
def test_bfs_bounded(oracle, seed):
    rec = random_records(seed, N, 1200)
    for mode, und in ((oracle.DIRECTED, False), (oracle.UNDIRECTED, True)):
        g = oracle.OracleGraph(rec, N, False, mode)
        for root in (1, 2, 33):
            full, fpar, fit = oracle.bfs(g, root)
            exp_full = CF.bfs_levels(rec, N, root, undirected=und)
            assert (full == exp_full).all()
            for m in (1, 2, 3, 5):
                hops, par, it = oracle.bfs(g, root, m)
                exp = np.where(exp_full <= m, exp_full, 255).astype(np.uint8)
                assert (hops == exp).all(), (mode, root, m)
                # parents are set when a vertex is reached, so they match the full run
                assert (par[hops != 255] == fpar[hops != 255]).all()
                # a bounded loop runs exactly m iterations, converged or not
                assert it == m


@pytest.mark.parametrize("seed", [23, 24])
def test_sssp_bounded(oracle, seed):
    rec = random_records(seed, N, 1600, weighted=True)
    g = oracle.OracleGraph(rec, N, True, oracle.DIRECTED)
    cls = CF.classes(rec, N)
    for root in (1, 2, 5, 77):
        # execute(1) = execute_single: every row one relaxation from the root
        d1, it1 = oracle.sssp(g, root, 1)
        assert (d1 == CF.hop_bounded_dist(rec, N, root, 1)).all()
        assert it1 == 1
        for m in (2, 3, 4, 7):
            d, it = oracle.sssp(g, root, m)
            # regular rows: paths of at most m edges; sinks, after the loop,
            # one more step from the regular rows' final distances
            dm, dm1 = CF.hop_bounded_dist(rec, N, root, m), CF.hop_bounded_dist(rec, N, root, m + 1)
            exp = np.where(cls["sink"], dm1, dm)
            exp[root] = 0
            assert (d == exp).all(), (root, m)
            assert it == m


@pytest.mark.parametrize("seed", [25, 26])
def test_cc_bounded_and_rounds(oracle, seed):
    rec = random_records(seed, 2000, 1800)
    g = oracle.OracleGraph(rec, 2000, False, oracle.UNDIRECTED)
    lab, it = oracle.cc(g)
    assert (lab == CF.min_id_components(rec, 2000)).all()
    assert it == CF.cc_rounds(rec, 2000)
    for m in (1, 2, 3):
        lm, itm = oracle.cc(g, m)
        assert (lm == CF.cc_bounded(rec, 2000, m)).all(), m
        assert itm == m


def test_pagerank_single_and_two(oracle):
    rec = random_records(27, N, 2500)
    g = oracle.OracleGraph(rec, N, False, oracle.DIRECTED)
    gr = oracle.OracleGraph(rec, N, False, oracle.REVERSED)
    for iters in (1, 2, 3):
        rank, _, it = oracle.pagerank(g, gr, iters)
        assert np.abs(rank - CF.pagerank_dense(rec, N, iters)).max() < 1e-14
        assert it == iters
