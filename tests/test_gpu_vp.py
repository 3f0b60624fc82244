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
