"""Programs written only against include/la3_vertex_program.hpp
(tests/vp_programs.hip, compiled outside the library) against the oracle's
Engine<P> with the same operations (oracle/la3_oracle.cpp), on the
reference's g1 fixture and on R-MAT graphs. Integer programs bit-exact
(including an order-dependent combine); the floating-point PageRank within
1e-12 (device code may contract a multiply-add)."""
import ctypes
import os

import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_vp = None


def vp():
    global _vp
    if _vp is None:
        L.lib()
        lib = ctypes.CDLL(os.path.join(ROOT, "tests", "libvp_programs.so"))
        P, u32 = ctypes.c_void_p, ctypes.c_uint32
        lib.vp_error.restype = ctypes.c_char_p
        lib.vp_wdeg.argtypes = [P, P, u32, P, P]
        lib.vp_maxlabel.argtypes = [P, P, u32, P, P]
        lib.vp_level.argtypes = [P, P, u32, u32, P, P, P]
        lib.vp_pralpha.argtypes = [P, P, P, u32, P, P]
        lib.vp_countgreater.argtypes = [P, P, u32, P, P, P]
        lib.vp_countnonzero.argtypes = [P, P, u32, P, P]
        lib.vp_maxlabel_seq.argtypes = [P, P, u32, ctypes.c_int, P, u32, u32, P, P]
        lib.vp_accgreater_twice.argtypes = [P, P, P, P]
        lib.vp_activity.argtypes = [P, P, ctypes.c_int, u32, u32, u32, P, P]
        lib.vp_extract.argtypes = [P, P, u32, ctypes.c_int, P, P, P, P, P, u32, ctypes.c_int, P,
                                   ctypes.c_uint64]
        lib.vp_borrow.argtypes = [P, P, P, P, P, P]
        _vp = lib
    return _vp


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _call(name, *args):
    st = getattr(vp(), name)(*args)
    assert st == 0, vp().vp_error().decode()


def wdeg(g, m):
    out, it = np.zeros(g.nslots, np.uint32), np.zeros(1, np.uint32)
    _call("vp_wdeg", g.ctx.h, g.h, m, _p(out), _p(it))
    return out, int(it[0])


def maxlabel(g, m):
    out, it = np.zeros(g.nslots, np.uint32), np.zeros(1, np.uint32)
    _call("vp_maxlabel", g.ctx.h, g.h, m, _p(out), _p(it))
    return out, int(it[0])


def level(g, root, m):
    lev, via, it = np.zeros(g.nslots, np.uint32), np.zeros(g.nslots, np.uint32), np.zeros(1, np.uint32)
    _call("vp_level", g.ctx.h, g.h, root, m, _p(lev), _p(via), _p(it))
    return lev, via, int(it[0])


def pralpha(g, gr, m):
    out, it = np.zeros(g.nslots, np.float64), np.zeros(1, np.uint32)
    _call("vp_pralpha", g.ctx.h, g.h, gr.h, m, _p(out), _p(it))
    return out, int(it[0])


def countgreater(g, m):
    val, cnt, it = np.zeros(g.nslots, np.uint32), np.zeros(g.nslots, np.uint32), np.zeros(1, np.uint32)
    _call("vp_countgreater", g.ctx.h, g.h, m, _p(val), _p(cnt), _p(it))
    return val, cnt, int(it[0])


def countnonzero(g, m):
    cnt, it = np.zeros(g.nslots, np.uint32), np.zeros(1, np.uint32)
    _call("vp_countnonzero", g.ctx.h, g.h, m, _p(cnt), _p(it))
    return cnt, int(it[0])


def _graphs(ctx, oracle, rec, n, weighted, mode):
    return L.Graph.from_records(ctx, rec, n, weighted, mode), oracle.OracleGraph(rec, n, weighted, mode)


def _cases(g1):
    return [("g1", g1["records_w"], 8, [5, 1, 4]),
            ("rmat14", L.rmat(14, 16, seed=61, weighted=True), 1 << 14, None)]


@pytest.mark.parametrize("case", [0, 1])
def test_interface_programs_match_oracle(gpu_ctx, oracle, g1, case):
    name, recw, n, roots = _cases(g1)[case]
    rec = np.ascontiguousarray(recw[:, :2])
    if roots is None:
        roots = [int(x) for x in np.unique(rec[:, 0])[:3]]
    gw, ow = _graphs(gpu_ctx, oracle, recw, n, True, L.DIRECTED)
    for m in (1, 2):
        w, it = wdeg(gw, m)
        ew, eit = oracle.vp_wdeg(ow, m)
        assert (w == ew).all() and it == eit, (name, m)
    for mode in (L.DIRECTED, L.UNDIRECTED):
        g, o = _graphs(gpu_ctx, oracle, rec, n, False, mode)
        for m in (0, 1, 3):
            lab, it = maxlabel(g, m)
            elab, eit = oracle.vp_maxlabel(o, m)
            assert (lab == elab).all() and it == eit, (name, mode, m)
        for root in roots:
            for m in (0, 2):
                lev, via, it = level(g, root, m)
                elev, evia, eit = oracle.vp_level(o, root, m)
                assert (lev == elev).all() and (via == evia).all() and it == eit, (name, mode, root, m)
        for m in (1, 2, 3):
            val, cnt, it = countgreater(g, m)
            ev, ec, eit = oracle.vp_countgreater(o, m)
            assert (val == ev).all() and (cnt == ec).all() and it == eit, (name, mode, m)
            cnt, it = countnonzero(g, m)
            ec, eit = oracle.vp_countnonzero(o, m)
            assert (cnt == ec).all() and it == eit, (name, mode, m)
    g, o = _graphs(gpu_ctx, oracle, rec, n, False, L.DIRECTED)
    gr, orr = _graphs(gpu_ctx, oracle, rec, n, False, L.REVERSED)
    for m in (1, 2, 20, 0):
        r, it = pralpha(g, gr, m)
        er, eit = oracle.vp_pralpha(o, orr, m)
        assert np.abs(r - er).max() <= 1e-12 and it == eit, (name, m)


def maxlabel_seq(g, m1, m2, reset=False, vids=None):
    lab, it = np.zeros(g.nslots, np.uint32), np.zeros(1, np.uint32)
    v = None if vids is None else np.ascontiguousarray(vids, dtype=np.uint32)
    _call("vp_maxlabel_seq", g.ctx.h, g.h, m1, 1 if reset else 0, None if v is None else _p(v),
          0 if v is None else v.size, m2, _p(lab), _p(it))
    return lab, int(it[0])


@pytest.mark.parametrize("case", [0, 1])
def test_reset_reinitialize_and_continue_match_oracle(gpu_ctx, oracle, g1, case):
    """Second execute without reset (continues from the messages left),
    reset + initialize(vids) (the other vertices keep their states), reset +
    initialize, and mirrors kept from the first initialize (AccGreater)."""
    name, recw, n, _ = _cases(g1)[case]
    rec = np.ascontiguousarray(recw[:, :2])
    g, o = _graphs(gpu_ctx, oracle, rec, n, False, L.UNDIRECTED)
    ids = np.arange(n + 1)
    vids = ids[(ids % 3 == 0) | (ids % 7 == 1)]
    for m1, m2, reset, vv in ((2, 1, False, None), (1, 2, False, None), (3, 2, False, None),
                              (0, 0, True, vids), (0, 0, True, None), (2, 0, True, vids)):
        got = maxlabel_seq(g, m1, m2, reset, vv)
        exp = oracle.vp_maxlabel_seq(o, m1, m2, reset, vv)
        assert (got[0] == exp[0]).all() and got[1] == exp[1], (name, m1, m2, reset)
    gd, od = _graphs(gpu_ctx, oracle, rec, n, False, L.DIRECTED)
    val, cnt = np.zeros(gd.nslots, np.uint32), np.zeros(gd.nslots, np.uint32)
    _call("vp_accgreater_twice", gd.ctx.h, gd.h, _p(val), _p(cnt))
    ev, ec = oracle.vp_accgreater_twice(od)
    assert (val == ev).all() and (cnt == ec).all(), name


def activity(g, prog, m, flags=0, root=0):
    out, act = np.zeros(g.nslots, np.uint32), np.zeros(g.nslots, np.uint8)
    _call("vp_activity", g.ctx.h, g.h, prog, root, m, flags, _p(out), _p(act))
    return out, act


@pytest.mark.parametrize("case", [0, 1])
def test_vertex_activity_matches_oracle(gpu_ctx, oracle, g1, case):
    """The state vector's activity (vertex_program_execute.hpp:693-713,
    vertex_program.hpp:159-173, 549-555) after runs of the interface
    programs, with activate_all() before initialize (mirrors of a
    non-stationary state reader) and reset_activity() / activate_all()
    after execute, bit-exact against the oracle's engine."""
    O = oracle
    name, recw, n, roots = _cases(g1)[case]
    rec = np.ascontiguousarray(recw[:, :2])
    if roots is None:
        roots = [int(x) for x in np.unique(rec[:, 0])[:2]]
    gw, ow = _graphs(gpu_ctx, oracle, recw, n, True, L.DIRECTED)
    for m in (1, 2):
        got, exp = activity(gw, O.VP_WDEG, m), O.vp_activity(ow, O.VP_WDEG, m)
        assert (got[0] == exp[0]).all() and (got[1] == exp[1]).all(), (name, "wdeg", m)
    for mode in (L.DIRECTED, L.UNDIRECTED):
        g, o = _graphs(gpu_ctx, oracle, rec, n, False, mode)
        runs = [(O.VP_MAXLABEL, m, f, 0) for m in (0, 1, 3) for f in (0, O.ACT_RESET_AFTER, O.ACT_ALL_AFTER)]
        runs += [(O.VP_LEVEL, m, 0, r) for m in (0, 2) for r in roots]
        runs += [(O.VP_COUNTNONZERO, m, f, 0) for m in (1, 2) for f in (0, O.ACT_ALL_BEFORE)]
        runs += [(O.VP_COUNTGREATER, m, 0, 0) for m in (1, 2)]
        for prog, m, f, r in runs:
            got = activity(g, prog, m, f, r)
            exp = O.vp_activity(o, prog, m, f, r)
            assert (got[0] == exp[0]).all() and (got[1] == exp[1]).all(), (name, mode, prog, m, f, r)


def extract(g, k, active_only, nvertices=30, values_only=False):
    ids, vals = np.zeros(k, np.uint32), np.zeros(k, np.uint32)
    bids, bvals = np.zeros(2 * k, np.uint32), np.zeros(2 * k, np.uint32)
    sums = np.zeros(2, np.uint64)
    text = ctypes.create_string_buffer(1 << 22)
    _call("vp_extract", g.ctx.h, g.h, k, int(active_only), _p(ids), _p(vals), _p(bids), _p(bvals), _p(sums),
          nvertices, int(values_only), text, len(text))
    return ids, vals, bids.reshape(2, k), bvals.reshape(2, k), [int(x) for x in sums], text.value.decode()


@pytest.mark.parametrize("case", [0, 1])
def test_topk_btopk_reduce_display(gpu_ctx, oracle, g1, case):
    """topk / btopk (vertex_program.hpp:607-767: partial_sort of the pairs in
    slot order, resized to k with I(), V() pairs when fewer), reduce with and
    without active_only (:558-604) and display (:511-546) on max-label
    results, against the oracle's labels and activity and the slot order."""
    name, recw, n, _ = _cases(g1)[case]
    rec = np.ascontiguousarray(recw[:, :2])
    g, o = _graphs(gpu_ctx, oracle, rec, n, False, L.UNDIRECTED)
    lab, act = oracle.vp_activity(o, oracle.VP_MAXLABEL, 0)
    lab = lab.astype(np.int64)
    ids_all = np.arange(n + 1)
    slot_ids = g.array("id_of")
    assert lab[0] == 0 and not act[0]  # vertex 0 has no edge: the pad pair (0, 0) is its own pair
    for active_only in (False, True):
        pool = ids_all[act.astype(bool)] if active_only else ids_all
        for k in (1, 5, len(pool) + 7):
            ids, vals, bids, bvals, sums, _ = extract(g, k, active_only)
            npad = max(0, k - len(pool))
            exp_vals = np.sort(np.concatenate([lab[pool], np.zeros(npad, np.int64)]))[::-1][:k]
            assert (vals.astype(np.int64) == exp_vals).all(), (name, active_only, k)
            assert (lab[ids] == vals).all() and (np.diff(vals.astype(np.int64)) <= 0).all()
            assert len(np.unique(ids[vals > 0])) == int((vals > 0).sum())
            assert np.isin(ids[vals > 0], pool).all()
            # batch 0 by label (as topk), batch 1 by id: distinct values, so
            # exactly the k largest ids of the pool, then pad pairs
            assert (bvals[0].astype(np.int64) == exp_vals).all() and (lab[bids[0]] == bvals[0]).all()
            exp_ids = np.concatenate([np.sort(pool)[::-1], np.zeros(npad, np.int64)])[:k]
            assert (bids[1] == exp_ids).all() and (bvals[1] == exp_ids).all(), (name, active_only, k)
        _, _, _, _, sums, _ = extract(g, 1, active_only)
        assert sums == [int(lab.sum()), int(lab[act.astype(bool)].sum())]
    for nv, vo in ((30, False), (4, True), (n + 50, False)):
        text = extract(g, 1, False, nv, vo)[5]
        nmax = min(nv, n)
        shown = [int(i) for i in slot_ids if i <= nmax]
        lines = [(f"{{label: {lab[i]}}}" if vo else f"{i}: {{label: {lab[i]}}}") for i in shown]
        assert text == "\n".join(lines) + "\n" + ("\n" if vo else ""), (name, nv, vo)


def test_borrowed_states_are_shared(gpu_ctx, oracle, g1):
    """A program constructed from another (vertex_program.h:41,
    vertex_program.hpp:22-30) shares its vertex states: the borrower's
    initialize changes what the owner reads; freeing the borrower keeps the
    owner's states. After both are freed, the borrower starts again from
    value-initialized states."""
    for recw, n in ((g1["records_w"], 8), (L.rmat(13, 16, seed=62, weighted=True), 1 << 13)):
        rec = np.ascontiguousarray(recw[:, :2])
        g, o = _graphs(gpu_ctx, oracle, rec, n, False, L.DIRECTED)
        out = [np.zeros(g.nslots, np.uint32) for _ in range(4)]
        _call("vp_borrow", g.ctx.h, g.h, *[_p(x) for x in out])
        exp, _ = oracle.vp_maxlabel(o, 0)
        assert (out[0] == exp).all()
        assert (out[1] == exp + 1000).all() and (out[2] == exp + 1000).all()
        assert (out[3] == 1000).all()
