"""Pins the oracle's vertex activity (the state vector's activity set,
src/structures/random_access_array.h:14) with closed forms: apply pushes a
regular vertex that was activated or is stationary and a sink that was
activated (src/vprogram/vertex_program_execute.hpp:693-713); nothing in
execute clears it; the first mirroring of a state-reading program activates
every vertex when stationary, broadcasts the active states and clears the
activity (src/vprogram/vertex_program.hpp:159-173); reset_activity() and
activate_all() clear and set it (:549-555)."""
import numpy as np
import pytest
from scipy.sparse import csgraph

import closed_forms as CF
from test_oracle import random_records
from test_oracle_programs import _max_ancestor

N = 300


def _edges(rec, undirected):
    d = CF.dedup_first(rec)
    a, b = d[:, 0].astype(np.int64), d[:, 1].astype(np.int64)
    if undirected:
        a, b = np.concatenate([a, b]), np.concatenate([b, a])
        key = np.unique(a * (N + 1) + b)
        a, b = key // (N + 1), key % (N + 1)
    return a, b


def test_stationary_program_activates_every_row(oracle):
    """WDeg: stationary, apply always activates: every vertex with an
    in-entry, regular or sink, whatever the bound."""
    rec = random_records(101, N, 1500, weighted=True)
    g = oracle.OracleGraph(rec, N, True, oracle.DIRECTED)
    a, b = _edges(rec, False)
    has_in = np.bincount(b, minlength=N + 1) > 0
    for m in (1, 2, 3):
        _, act = oracle.vp_activity(g, oracle.VP_WDEG, m)
        assert (act.astype(bool) == has_in).all(), m


@pytest.mark.parametrize("mode", ["directed", "undirected"])
def test_activity_is_changed_labels(oracle, mode):
    """MaxLabel is not stationary and activates a vertex when its label
    grows: the activity is the set of vertices whose label is not their id,
    for execute_single, bounded and unbounded runs (sinks by the post-loop
    step on a directed load)."""
    und = mode == "undirected"
    rec = random_records(102, N, 700)
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED if und else oracle.DIRECTED)
    ids = np.arange(N + 1)
    for m in (0, 1, 3):
        lab, act = oracle.vp_activity(g, oracle.VP_MAXLABEL, m)
        assert (act.astype(bool) == (lab != ids)).all(), m
    lab, act = oracle.vp_activity(g, oracle.VP_MAXLABEL, 0)
    assert (lab == _max_ancestor(rec, N, und)).all()
    assert act.sum() > 0 and (act.astype(bool) == (_max_ancestor(rec, N, und) != ids)).all()


def test_level_activity_is_reached_but_root(oracle):
    n = 2000
    rec = random_records(103, n, 9000)
    g = oracle.OracleGraph(rec, n, False, oracle.DIRECTED)
    for root in (1, 9):
        hops = CF.bfs_levels(rec, n, root)
        for m in (0, 2):
            lev, act = oracle.vp_activity(g, oracle.VP_LEVEL, m, root=root)
            reached = (hops != 255) & (hops != 0)
            if m:
                reached &= hops <= m
            assert (act.astype(bool) == reached).all(), (root, m)


@pytest.mark.parametrize("mode", ["directed", "undirected"])
def test_stationary_state_reader_activity_is_regular_rows(oracle, mode):
    """CountGreater is stationary with a state-reading gather and an apply
    that never activates: the first initialize clears the activity after
    mirroring, then every regular row that receives is pushed (stationary)
    and no sink is (apply returned false)."""
    und = mode == "undirected"
    rec = random_records(104, N, 900)
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED if und else oracle.DIRECTED)
    a, b = _edges(rec, und)
    has_in = np.bincount(b, minlength=N + 1) > 0
    has_out = np.bincount(a, minlength=N + 1) > 0
    _, act = oracle.vp_activity(g, oracle.VP_COUNTGREATER, 1)
    assert (act.astype(bool) == (has_in & has_out)).all()
    assert und or (has_in & ~has_out).any()  # the directed case has sinks left out


@pytest.mark.parametrize("mode", ["directed", "undirected"])
def test_activate_all_before_initialize_fills_the_mirrors(oracle, mode):
    """CountNonzero (not stationary, state-reading gather): mirrors take the
    active states at the first initialize. With nothing active they keep
    S() (val 0); after activate_all() every row's mirror holds its initial
    value vid % 7, so a row counts the messages above its own value. Either
    way the activity is cleared by the mirroring, and after execute(1) it is
    the rows that counted something (apply activates when y > 0)."""
    und = mode == "undirected"
    rec = random_records(105, N, 1500)
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED if und else oracle.DIRECTED)
    a, b = _edges(rec, und)
    ids = np.arange(N + 1)
    init = ids % 7
    sender = (ids % 3 == 0)[a]
    for flags, mirror in ((0, np.zeros(N + 1, dtype=np.int64)), (oracle.ACT_ALL_BEFORE, init)):
        y = np.zeros(N + 1, dtype=np.int64)
        np.add.at(y, b[sender], (init[a[sender]] > mirror[b[sender]]).astype(np.int64))
        cnt, act = oracle.vp_activity(g, oracle.VP_COUNTNONZERO, 1, flags=flags)
        assert (cnt == y).all(), flags
        assert (act.astype(bool) == (y > 0)).all(), flags
    y_def = np.zeros(N + 1, dtype=np.int64)
    np.add.at(y_def, b[sender], (init[a[sender]] > 0).astype(np.int64))
    y_all, _ = oracle.vp_activity(g, oracle.VP_COUNTNONZERO, 1, flags=oracle.ACT_ALL_BEFORE)
    assert (y_all != y_def).any()  # the two mirror rules are told apart


def test_reset_activity_and_activate_all_after_execute(oracle):
    rec = random_records(106, N, 700)
    g = oracle.OracleGraph(rec, N, False, oracle.UNDIRECTED)
    _, act = oracle.vp_activity(g, oracle.VP_MAXLABEL, 0, flags=oracle.ACT_RESET_AFTER)
    assert not act.any()
    _, act = oracle.vp_activity(g, oracle.VP_MAXLABEL, 0, flags=oracle.ACT_ALL_AFTER)
    assert act.all()
