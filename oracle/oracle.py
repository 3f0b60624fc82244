"""ctypes wrapper over oracle/liboracle.so, the CPU restatement of LA3's
vertex-program path at np=1.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker. The product
(la3_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

# Load modes, as in ora_graph_new (graph.h:38-47 load_directed/undirected flags).
DIRECTED, REVERSED, UNDIRECTED, ACYCLIC, REVERSED_ACYCLIC = 0, 1, 2, 3, 4
BFS_INF = 255
SSSP_INF = 2**31 - 1


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.ora_graph_new.restype = P
        L.ora_graph_new.argtypes = [P, u64, u32, i32, i32, i32]
        L.ora_graph_free.argtypes = [P]
        L.ora_graph_info.argtypes = [P, P]
        L.ora_graph_layout.argtypes = [P, P]
        L.ora_graph_csc.argtypes = [P, i32, P, P, P, P, P]
        L.ora_hash.argtypes = [P, u32]
        L.ora_hash.restype = u32
        L.ora_unhash.argtypes = [P, u32]
        L.ora_unhash.restype = u32
        L.ora_degree.argtypes = [P, P]
        L.ora_pr.argtypes = [P, P, u32, P, P, P]
        L.ora_ppr.argtypes = [P, P, u32, u32, ctypes.c_double, P, P]
        L.ora_bfs.argtypes = [P, u32, u32, P, P, P]
        L.ora_sssp.argtypes = [P, u32, u32, P, P]
        L.ora_cc.argtypes = [P, u32, P, P]
        L.ora_tc.argtypes = [P, P, P, P]
        L.ora_vp_wdeg.argtypes = [P, u32, P, P]
        L.ora_vp_maxlabel.argtypes = [P, u32, P, P]
        L.ora_vp_level.argtypes = [P, u32, u32, P, P, P]
        L.ora_vp_pralpha.argtypes = [P, P, u32, P, P]
        L.ora_vp_countgreater.argtypes = [P, u32, P, P, P]
        L.ora_vp_countnonzero.argtypes = [P, u32, P, P]
        L.ora_vp_maxlabel_seq.argtypes = [P, u32, ctypes.c_int, P, u32, u32, P, P]
        L.ora_vp_accgreater_twice.argtypes = [P, P, P]
        L.ora_vp_activity.argtypes = [P, ctypes.c_int, u32, u32, u32, P, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleGraph:
    """One LA3 graph load (Graph<W>::load_directed / load_undirected at np=1)."""

    def __init__(self, records, n, weighted, mode, hashing=1):
        rec = np.ascontiguousarray(records, dtype=np.uint32)
        width = 3 if weighted else 2
        rec = rec.reshape(-1, width)
        self.n = int(n)
        self._rec = rec
        self.h = lib().ora_graph_new(_ptr(rec), rec.shape[0], self.n, int(weighted), int(mode),
                                     int(hashing))
        if not self.h:
            raise ValueError("vertex id out of range [0, n]")
        info = np.zeros(8, dtype=np.uint64)
        lib().ora_graph_info(self.h, _ptr(info))
        (self.nslots, self.nreg, self.nsink, self.nsrc, self.ncols, self.nnz_reg, self.nnz_sink,
         _) = (int(x) for x in info)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_graph_free(self.h)
            self.h = None

    def hash(self, v):
        return int(lib().ora_hash(self.h, int(v)))

    def unhash(self, v):
        return int(lib().ora_unhash(self.h, int(v)))

    def hid_of(self):
        out = np.zeros(self.nslots, dtype=np.uint32)
        lib().ora_graph_layout(self.h, _ptr(out))
        return out

    def csc(self, which):
        nnz = self.nnz_sink if which else self.nnz_reg
        colptrs = np.zeros(self.ncols + 1, dtype=np.uint32)
        colidxs = np.zeros(self.ncols + 1, dtype=np.uint32)
        gidx = np.zeros(max(nnz, 1), dtype=np.uint32)
        idx = np.zeros(max(nnz, 1), dtype=np.uint32)
        w = np.zeros(max(nnz, 1), dtype=np.uint32)
        lib().ora_graph_csc(self.h, int(which), _ptr(colptrs), _ptr(colidxs), _ptr(gidx),
                            _ptr(idx), _ptr(w))
        return dict(colptrs=colptrs, colidxs=colidxs, global_idx=gidx[:nnz], idx=idx[:nnz],
                    w=w[:nnz])


def degree(g):
    out = np.zeros(g.nslots, dtype=np.uint32)
    lib().ora_degree(g.h, _ptr(out))
    return out


def pagerank(g, gr, iters):
    rank = np.zeros(g.nslots, dtype=np.float64)
    deg = np.zeros(g.nslots, dtype=np.uint32)
    it = np.zeros(1, dtype=np.uint32)
    lib().ora_pr(g.h, gr.h, int(iters), _ptr(rank), _ptr(deg), _ptr(it))
    return rank, deg, int(it[0])


def ppr(g, gr, seed, iters, alpha=0.15):
    rank = np.zeros(g.nslots, dtype=np.float64)
    it = np.zeros(1, dtype=np.uint32)
    lib().ora_ppr(g.h, gr.h, int(seed), int(iters), float(alpha), _ptr(rank), _ptr(it))
    return rank, int(it[0])


def bfs(g, root, max_iters=0):
    """BfsVertex execute(max_iters): 0 runs until convergence, 1 is
    execute_single, m > 1 stops after m iterations."""
    hops = np.zeros(g.nslots, dtype=np.uint8)
    parent = np.zeros(g.nslots, dtype=np.uint32)
    it = np.zeros(1, dtype=np.uint32)
    lib().ora_bfs(g.h, int(root), int(max_iters), _ptr(hops), _ptr(parent), _ptr(it))
    return hops, parent, int(it[0])


def sssp(g, root, max_iters=0):
    dist = np.zeros(g.nslots, dtype=np.uint32)
    it = np.zeros(1, dtype=np.uint32)
    lib().ora_sssp(g.h, int(root), int(max_iters), _ptr(dist), _ptr(it))
    return dist, int(it[0])


def cc(g, max_iters=0):
    label = np.zeros(g.nslots, dtype=np.uint32)
    it = np.zeros(1, dtype=np.uint32)
    lib().ora_cc(g.h, int(max_iters), _ptr(label), _ptr(it))
    return label, int(it[0])


def tc(g_acyclic, gr_acyclic):
    total = np.zeros(1, dtype=np.uint64)
    per = np.zeros(g_acyclic.nslots, dtype=np.uint32)
    lib().ora_tc(g_acyclic.h, gr_acyclic.h, _ptr(total), _ptr(per))
    return int(total[0]), per


# --------------------------------------- programs of the interface alone ---
def _u32(n):
    return np.zeros(n, dtype=np.uint32)


def vp_wdeg(g, max_iters=1):
    """Weighted in-degree (stationary; W = u32 weight)."""
    out, it = _u32(g.nslots), _u32(1)
    lib().ora_vp_wdeg(g.h, int(max_iters), _ptr(out), _ptr(it))
    return out, int(it[0])


def vp_maxlabel(g, max_iters=0):
    """Largest id reached by max-label propagation."""
    out, it = _u32(g.nslots), _u32(1)
    lib().ora_vp_maxlabel(g.h, int(max_iters), _ptr(out), _ptr(it))
    return out, int(it[0])


def vp_level(g, root, max_iters=0):
    """Level (iteration of first reach, ~0 unreached) and the last-combined sender."""
    lev, via, it = _u32(g.nslots), _u32(g.nslots), _u32(1)
    lib().ora_vp_level(g.h, int(root), int(max_iters), _ptr(lev), _ptr(via), _ptr(it))
    return lev, via, int(it[0])


def vp_pralpha(g, gr, max_iters):
    """PageRank with alpha 0.2 and tolerance 1e-9, degrees via initialize(other)."""
    rank, it = np.zeros(g.nslots, dtype=np.float64), _u32(1)
    lib().ora_vp_pralpha(g.h, gr.h, int(max_iters), _ptr(rank), _ptr(it))
    return rank, int(it[0])


def vp_countgreater(g, max_iters):
    """State-reading gather: counts messages larger than the mirrored value."""
    val, cnt, it = _u32(g.nslots), _u32(g.nslots), _u32(1)
    lib().ora_vp_countgreater(g.h, int(max_iters), _ptr(val), _ptr(cnt), _ptr(it))
    return val, cnt, int(it[0])


def vp_countnonzero(g, max_iters):
    """Non-stationary state-reading gather: the mirrors keep S(), so each row
    counts the messages above 0 (senders: vid % 3 == 0 at first)."""
    cnt, it = _u32(g.nslots), _u32(1)
    lib().ora_vp_countnonzero(g.h, int(max_iters), _ptr(cnt), _ptr(it))
    return cnt, int(it[0])


def vp_maxlabel_seq(g, m1, m2, reset=False, vids=None):
    """MaxLabel: initialize; execute(m1); then execute(m2) continuing, or
    reset; initialize(vids) (all when None); execute(m2)."""
    label, it = _u32(g.nslots), _u32(1)
    v = None if vids is None else np.ascontiguousarray(vids, dtype=np.uint32)
    lib().ora_vp_maxlabel_seq(g.h, int(m1), 1 if reset else 0, None if v is None else _ptr(v),
                              0 if v is None else int(v.size), int(m2), _ptr(label), _ptr(it))
    return label, int(it[0])


VP_WDEG, VP_MAXLABEL, VP_LEVEL, VP_COUNTNONZERO, VP_COUNTGREATER = range(5)
ACT_ALL_BEFORE, ACT_RESET_AFTER, ACT_ALL_AFTER = 1, 2, 4


def vp_activity(g, prog, max_iters, flags=0, root=0):
    """An interface program's first state field and its vertex activity (by
    original id) after initialize; execute(max_iters), with activate_all()
    before initialize and reset_activity() / activate_all() after execute as
    flags ask."""
    out, act = _u32(g.nslots), np.zeros(g.nslots, dtype=np.uint8)
    st = lib().ora_vp_activity(g.h, int(prog), int(root), int(max_iters), int(flags), _ptr(out), _ptr(act))
    assert st == 0
    return out, act


def vp_accgreater_twice(g):
    """AccGreater: initialize; execute(1); reset; initialize; execute(1)."""
    val, cnt = _u32(g.nslots), _u32(g.nslots)
    lib().ora_vp_accgreater_twice(g.h, _ptr(val), _ptr(cnt))
    return val, cnt
