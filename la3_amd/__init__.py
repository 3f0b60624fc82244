"""Python mirror of LA3's Graph / VertexProgram entry points over the C ABI
of la3_amd/libla3.so (include/la3.h).

The reference's host interface is C++ (src/matrix/graph.h:38-47,
src/vprogram/vertex_program.h:33-220); the C++ drop-in lives in
include/la3.hpp. This module exposes the same loads and programs to Python
for tests, benchmarks and the serving loop. It loads the HIP library and
raises if it is missing: there is no CPU fallback.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libla3.so")

OK, EINVAL, ENOMEM, EIO, EDEVICE, ENOTSUP, EAGAIN = range(7)
DIRECTED, REVERSED, UNDIRECTED, ACYCLIC, REVERSED_ACYCLIC = range(5)
HASH_NONE, HASH_BUCKET = 0, 1
BFS_INF = 255
SSSP_INF = 2**31 - 1

# Every symbol include/la3.h declares (checked by tests/test_capi.py).
EXPORTS = [
    "la3_ctx_create", "la3_ctx_destroy", "la3_last_error", "la3_ctx_synchronize",
    "la3_graph_load_file", "la3_graph_load_records", "la3_graph_free", "la3_graph_info_get",
    "la3_graph_slots", "la3_graph_csc", "la3_graph_hash", "la3_graph_unhash",
    "la3_degree", "la3_pagerank", "la3_bfs", "la3_sssp", "la3_cc", "la3_tc", "la3_ppr",
    "la3_bfs_batch", "la3_sssp_batch", "la3_ppr_batch",
    "la3_last_timing", "la3_set_profiling", "la3_rmat_generate",
    "la3_kernel_stats", "la3_kernel_stats_reset", "la3_host_alloc", "la3_host_free",
    "la3_copy_bandwidth", "la3_result_fetch", "la3_result_fetch_slots",
    "la3_bfs_batch_ctx", "la3_sssp_batch_ctx", "la3_ppr_batch_ctx", "la3_bfs_ctx", "la3_sssp_ctx",
    "la3_server_create", "la3_server_add_graph", "la3_server_query_bytes", "la3_server_submit",
    "la3_server_drain", "la3_server_completed", "la3_server_result", "la3_server_clock",
    "la3_server_destroy", "la3_server_warm", "la3_result_reduce", "la3_graph_from_csc", "la3_graph_set_tuning", "la3_graph_get_tuning", "la3_autotune",
    "la3_batch_run", "la3_session_create", "la3_session_add", "la3_session_step",
    "la3_session_poll", "la3_session_free_slots", "la3_session_destroy", "la3_ctx_release_scratch",
    "la3_result_topk", "la3_ctx_set_ingest", "la3_graph_array", "la3_components",
    "la3_graph_device_view", "la3_ctx_stream", "la3_bfs_noalloc",
]
(TUNE_SSSP_DELTA, TUNE_BATCH_SSSP_DELTA, TUNE_BFS_TASK_K, TUNE_SSSP_TASK_K, TUNE_PPR_TASK_K,
 TUNE_FUSED_BLOCKS) = range(6)
Q_BFS, Q_SSSP, Q_PPR = 0, 1, 2


class LA3Error(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"la3 status {status}: {msg}")
        self.status = status


class GraphInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "n", "nslots", "nregular", "nsink", "nsource", "nisolated", "ncols", "directed",
        "weighted", "mode")] + [(n, ctypes.c_uint64) for n in (
            "nnz", "nnz_regular", "nnz_sink", "nrecords", "device_bytes")]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint64),
                ("ms", ctypes.c_double), ("alg_bytes", ctypes.c_double),
                ("entries", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("nworkers", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("state_budget", ctypes.c_uint64), ("keep_results", ctypes.c_int),
                ("ppr_alpha", ctypes.c_double), ("fused", ctypes.c_int),
                ("max_wait_s", ctypes.c_double)]


class CscArrays(ctypes.Structure):
    _fields_ = [("nvertices", ctypes.c_uint32), ("weighted", ctypes.c_int), ("mode", ctypes.c_int),
                ("hashing", ctypes.c_int), ("ncols", ctypes.c_uint32),
                ("colptrs", ctypes.c_void_p * 2), ("colidxs", ctypes.c_void_p * 2),
                ("global_idx", ctypes.c_void_p * 2), ("idx", ctypes.c_void_p * 2),
                ("weight", ctypes.c_void_p * 2)]


class QueryInfo(ctypes.Structure):
    _fields_ = [("qid", ctypes.c_uint64), ("kind", ctypes.c_uint32), ("graph", ctypes.c_uint32),
                ("root", ctypes.c_uint32), ("batch_size", ctypes.c_uint32),
                ("status", ctypes.c_int32), ("submit_s", ctypes.c_double),
                ("start_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class AutotuneReport(ctypes.Structure):
    _fields_ = [("ncandidates", ctypes.c_uint32), ("candidates", ctypes.c_uint32 * 8),
                ("single_ms", ctypes.c_double * 8), ("batch_ms", ctypes.c_double * 8),
                ("sssp_delta", ctypes.c_uint32), ("batch_sssp_delta", ctypes.c_uint32),
                ("batch_k", ctypes.c_uint32), ("task_k", ctypes.c_uint32 * 3),
                ("task_ms", (ctypes.c_double * 3) * 3), ("fused_blocks", ctypes.c_uint32),
                ("fused_ms", ctypes.c_double * 3)]


class Batch(ctypes.Structure):
    _fields_ = [("graph", ctypes.c_void_p), ("kind", ctypes.c_int), ("k", ctypes.c_uint32),
                ("roots", ctypes.c_void_p), ("iters", ctypes.c_uint32), ("out", ctypes.c_void_p),
                ("out_comp", ctypes.c_void_p), ("residual", ctypes.c_void_p),
                ("iters_run", ctypes.c_uint32)]


class Timing(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("dominant_ms", ctypes.c_double),
                ("dominant_launches", ctypes.c_uint64), ("edges_scanned", ctypes.c_uint64),
                ("dominant_kernel", ctypes.c_char * 64)]


_lib = None


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              "(the HIP library is required; there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        P, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        sig = {
            "la3_ctx_create": (i32, [i32, ctypes.POINTER(P)]),
            "la3_ctx_destroy": (None, [P]),
            "la3_last_error": (ctypes.c_char_p, [P]),
            "la3_ctx_synchronize": (i32, [P]),
            "la3_graph_load_file": (i32, [P, ctypes.c_char_p, u32, i32, i32, i32, ctypes.POINTER(P)]),
            "la3_graph_load_records": (i32, [P, P, u64, u32, i32, i32, i32, ctypes.POINTER(P)]),
            "la3_graph_free": (None, [P]),
            "la3_graph_info_get": (i32, [P, ctypes.POINTER(GraphInfo)]),
            "la3_graph_slots": (i32, [P, P]),
            "la3_graph_csc": (i32, [P, i32, P, P, P, P, P]),
            "la3_graph_hash": (u32, [P, u32]),
            "la3_graph_unhash": (u32, [P, u32]),
            "la3_degree": (i32, [P, P]),
            "la3_pagerank": (i32, [P, u32, P, P, P]),
            "la3_bfs": (i32, [P, u32, u32, P, P, P]),
            "la3_sssp": (i32, [P, u32, u32, P, P]),
            "la3_cc": (i32, [P, u32, P, P]),
            "la3_components": (i32, [P, P]),
            "la3_graph_device_view": (i32, [P, P]),
            "la3_ctx_stream": (P, [P]),
            "la3_bfs_noalloc": (i32, [P, P, u32, u32, P, P, P]),
            "la3_tc": (i32, [P, P, P]),
            "la3_ppr": (i32, [P, u32, u32, ctypes.c_double, P]),
            "la3_bfs_batch": (i32, [P, P, u32, u32, P, P]),
            "la3_sssp_batch": (i32, [P, P, u32, u32, P, P]),
            "la3_ppr_batch": (i32, [P, P, u32, u32, ctypes.c_double, P, P, P]),
            "la3_last_timing": (i32, [P, ctypes.POINTER(Timing)]),
            "la3_set_profiling": (i32, [P, i32]),
            "la3_kernel_stats": (i32, [P, P, u32, ctypes.POINTER(u32)]),
            "la3_kernel_stats_reset": (i32, [P]),
            "la3_host_alloc": (i32, [u64, ctypes.POINTER(P)]),
            "la3_host_free": (None, [P]),
            "la3_copy_bandwidth": (i32, [P, u64, u32, ctypes.POINTER(ctypes.c_double)]),
            "la3_result_fetch": (i32, [P, P, u64]),
            "la3_result_fetch_slots": (i32, [P, P, u32, u32, P]),
            "la3_bfs_ctx": (i32, [P, P, u32, u32, P, P, P]),
            "la3_sssp_ctx": (i32, [P, P, u32, u32, P, P]),
            "la3_bfs_batch_ctx": (i32, [P, P, P, u32, u32, P, P]),
            "la3_sssp_batch_ctx": (i32, [P, P, P, u32, u32, P, P]),
            "la3_ppr_batch_ctx": (i32, [P, P, P, u32, u32, ctypes.c_double, P, P, P]),
            "la3_server_create": (i32, [i32, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]),
            "la3_server_add_graph": (i32, [P, P, ctypes.POINTER(u32)]),
            "la3_server_query_bytes": (u64, [P, i32, u32]),
            "la3_server_submit": (i32, [P, i32, u32, u32, u32, ctypes.POINTER(u64)]),
            "la3_server_drain": (i32, [P]),
            "la3_server_completed": (i32, [P, P, u32, ctypes.POINTER(u32)]),
            "la3_server_result": (i32, [P, u64, P, u64]),
            "la3_server_clock": (ctypes.c_double, [P]),
            "la3_server_destroy": (None, [P]),
            "la3_server_warm": (i32, [P]),
            "la3_result_reduce": (i32, [P, i32, u32, u64, i32, ctypes.c_double, P]),
            "la3_graph_from_csc": (i32, [P, P, ctypes.POINTER(P)]),
            "la3_graph_set_tuning": (i32, [P, i32, u32]),
            "la3_graph_get_tuning": (i32, [P, i32, ctypes.POINTER(u32)]),
            "la3_autotune": (i32, [P, P, u32, ctypes.POINTER(AutotuneReport)]),
            "la3_batch_run": (i32, [P, ctypes.POINTER(Batch), u32, ctypes.c_double]),
            "la3_session_create": (i32, [P, ctypes.c_double, ctypes.POINTER(P)]),
            "la3_session_add": (i32, [P, ctypes.POINTER(Batch), ctypes.POINTER(u64)]),
            "la3_session_step": (i32, [P, ctypes.POINTER(u32)]),
            "la3_session_poll": (i32, [P, P, P, P, u32, ctypes.POINTER(u32)]),
            "la3_session_free_slots": (u32, [P]),
            "la3_session_destroy": (None, [P]),
            "la3_ctx_release_scratch": (i32, [P]),
            "la3_result_topk": (i32, [P, i32, u32, u64, u32, i32, P, P]),
            "la3_ctx_set_ingest": (i32, [P, i32]),
            "la3_graph_array": (i32, [P, i32, P, u64, ctypes.POINTER(u64)]),
            "la3_rmat_generate": (i32, [u32, u32, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, u64, i32, P, ctypes.POINTER(u64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def pinned(shape, dtype):
    """A numpy array in page-locked host memory (freed with the array)."""
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    p = ctypes.c_void_p()
    st = lib().la3_host_alloc(max(nbytes, 1), ctypes.byref(p))
    if st:
        raise LA3Error(st, f"host_alloc({nbytes})")
    buf = (ctypes.c_char * max(nbytes, 1)).from_address(p.value)
    weakref.finalize(buf, lib().la3_host_free, p.value)
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


DEVICE = "device"  # pass as out= to leave a program's result in device memory


def _out(out, shape, dtype):
    if out is DEVICE:
        return None
    if out is None:
        return np.zeros(shape, dtype=dtype)
    assert out.shape == tuple(shape) and out.dtype == np.dtype(dtype) and out.flags.c_contiguous
    return out


def rmat(scale, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=1, weighted=False):
    """Seeded R-MAT records (ids 1..2^scale, no self-loops or duplicates)."""
    n = ctypes.c_uint64(0)
    st = lib().la3_rmat_generate(scale, edgefactor, a, b, c, seed, int(weighted), None,
                                 ctypes.byref(n))
    if st:
        raise LA3Error(st, "rmat_generate")
    out = np.empty(n.value * (3 if weighted else 2), dtype=np.uint32)
    st = lib().la3_rmat_generate(scale, edgefactor, a, b, c, seed, int(weighted), _ptr(out),
                                 ctypes.byref(n))
    if st:
        raise LA3Error(st, "rmat_generate")
    return out.reshape(-1, 3 if weighted else 2)


class Context:
    """One device (or, with device=-1, a host-only context for layout work)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        # The destructors release through this reference: module globals
        # may already be gone when they run at interpreter exit.
        self._L = lib()
        st = self._L.la3_ctx_create(device, ctypes.byref(h))
        if st:
            raise LA3Error(st, f"la3_ctx_create(device={device}) failed")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self._L.la3_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, st, what):
        if st:
            raise LA3Error(st, f"{what}: {lib().la3_last_error(self.h).decode()}")

    def set_profiling(self, on=True):
        self.check(lib().la3_set_profiling(self.h, int(on)), "set_profiling")

    def last_timing(self):
        t = Timing()
        self.check(lib().la3_last_timing(self.h, ctypes.byref(t)), "last_timing")
        return dict(total_ms=t.total_ms, dominant_ms=t.dominant_ms,
                    dominant_launches=t.dominant_launches, edges_scanned=t.edges_scanned,
                    dominant_kernel=t.dominant_kernel.decode())

    def kernel_stats(self):
        """Cumulative per-kernel launches / ms / algorithmic bytes (profiling on)."""
        n = ctypes.c_uint32(0)
        self.check(lib().la3_kernel_stats(self.h, None, 0, ctypes.byref(n)), "kernel_stats")
        arr = (KernelStat * max(n.value, 1))()
        self.check(lib().la3_kernel_stats(self.h, arr, n.value, ctypes.byref(n)), "kernel_stats")
        return {a.name.decode(): dict(launches=a.launches, ms=a.ms, alg_bytes=a.alg_bytes,
                                      entries=a.entries) for a in arr[:n.value]}

    def reset_kernel_stats(self):
        self.check(lib().la3_kernel_stats_reset(self.h), "kernel_stats_reset")

    def synchronize(self):
        self.check(lib().la3_ctx_synchronize(self.h), "synchronize")

    def set_ingest(self, on_device):
        """Build loaded graphs on the device (True, default) or the host."""
        self.check(lib().la3_ctx_set_ingest(self.h, int(bool(on_device))), "set_ingest")

    def release_scratch(self):
        """Frees the device scratch cached between calls."""
        self.check(lib().la3_ctx_release_scratch(self.h), "release_scratch")

    def fetch(self, out):
        """Copies the last program's device-resident main output into `out`."""
        self.check(lib().la3_result_fetch(self.h, _ptr(out), out.nbytes), "result_fetch")
        return out

    def fetch_slots(self, g, dtype, rows=1):
        """The last device-resident result's first `rows` rows (g.nslots values
        of numpy `dtype` by id) in the reference's master-slot order."""
        dt = np.dtype(dtype)
        out = np.empty((rows, g.nslots), dtype=dt)
        self.check(lib().la3_result_fetch_slots(self.h, g.h, dt.itemsize, rows, _ptr(out)), "result_fetch_slots")
        return out

    def topk(self, dtype, rows, n, k, largest=True):
        """Top-k of each row of the last device-resident result (rows x n
        values of numpy `dtype` by id): (ids[rows, k], values[rows, k]),
        ordered by value, then smaller id; padded with id 0xffffffff."""
        dt = np.dtype(dtype)
        code = {np.dtype(np.uint8): 0, np.dtype(np.uint32): 1, np.dtype(np.float32): 2,
                np.dtype(np.float64): 3}[dt]
        ids = np.zeros((rows, k), dtype=np.uint32)
        vals = np.zeros((rows, k), dtype=dt)
        self.check(lib().la3_result_topk(self.h, code, int(rows), int(n), int(k), int(bool(largest)),
                                         _ptr(ids), _ptr(vals)), "result_topk")
        return ids, vals

    def reduce(self, dtype, rows, n, op="sum", value=0):
        """Per-row sum ("sum") or count of values != value ("count_ne") of
        the last device-resident result (la3_result_reduce): uint64 sums of
        integer results and counts, float64 sums of floating-point ones."""
        dt = np.dtype(dtype)
        code = {np.dtype(np.uint8): 0, np.dtype(np.uint32): 1, np.dtype(np.float32): 2,
                np.dtype(np.float64): 3}[dt]
        opc = {"sum": 0, "count_ne": 1}[op]
        out = np.zeros(rows, dtype=np.float64 if (opc == 0 and dt.kind == "f") else np.uint64)
        self.check(lib().la3_result_reduce(self.h, code, int(rows), int(n), opc, float(value), _ptr(out)),
                   "result_reduce")
        return out

    def copy_bandwidth(self, nbytes=1 << 30, reps=20):
        """Measured device copy GB/s (read + write) of a plain copy kernel."""
        v = ctypes.c_double(0)
        self.check(lib().la3_copy_bandwidth(self.h, int(nbytes), int(reps), ctypes.byref(v)),
                   "copy_bandwidth")
        return v.value


class Graph:
    """Graph<W> (src/matrix/graph.h:22-111) at one rank, resident on the device."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self._L = lib()
        self.h = handle
        info = GraphInfo()
        ctx.check(lib().la3_graph_info_get(self.h, ctypes.byref(info)), "graph_info")
        self.info = {f: getattr(info, f) for f, _ in GraphInfo._fields_}
        self.n = info.n
        self.nslots = info.nslots
        self.weighted = bool(info.weighted)
        self.mode = info.mode

    @classmethod
    def from_records(cls, ctx, records, n, weighted=False, mode=DIRECTED, hashing=HASH_BUCKET):
        rec = np.ascontiguousarray(records, dtype=np.uint32)
        nrec = rec.size // (3 if weighted else 2)
        h = ctypes.c_void_p()
        ctx.check(lib().la3_graph_load_records(ctx.h, _ptr(rec), nrec, int(n), int(weighted),
                                               int(mode), int(hashing), ctypes.byref(h)),
                  "graph_load_records")
        return cls(ctx, h)

    @classmethod
    def from_csc(cls, ctx, n, csc, weighted=False, mode=DIRECTED, hashing=HASH_BUCKET):
        """A graph from the reference's two CSCs (la3_graph_from_csc): csc =
        (regular, sink), each a dict with colptrs, colidxs, global_idx, idx, w
        as Graph.csc returns them."""
        arr = CscArrays()
        arr.nvertices, arr.weighted, arr.mode, arr.hashing = int(n), int(weighted), int(mode), int(hashing)
        keep = []
        arr.ncols = int(len(csc[0]["colptrs"]) - 1)
        for w, c in enumerate(csc):
            for field, key in (("colptrs", "colptrs"), ("colidxs", "colidxs"), ("global_idx", "global_idx"),
                               ("idx", "idx"), ("weight", "w")):
                a = c.get(key)
                if a is None or (key == "w" and not weighted):
                    continue
                a = np.ascontiguousarray(a, dtype=np.uint32)
                keep.append(a)
                getattr(arr, field)[w] = a.ctypes.data
        h = ctypes.c_void_p()
        ctx.check(lib().la3_graph_from_csc(ctx.h, ctypes.byref(arr), ctypes.byref(h)), "graph_from_csc")
        return cls(ctx, h)

    @classmethod
    def load_file(cls, ctx, path, n, weighted=False, mode=DIRECTED, hashing=HASH_BUCKET):
        h = ctypes.c_void_p()
        ctx.check(lib().la3_graph_load_file(ctx.h, path.encode(), int(n), int(weighted),
                                            int(mode), int(hashing), ctypes.byref(h)),
                  "graph_load_file")
        return cls(ctx, h)

    # graph.h:38-47 spellings
    @classmethod
    def load_directed(cls, ctx, path, n, reverse_edges=False, remove_cycles=False,
                      weighted=False, hashing=HASH_BUCKET):
        mode = {(False, False): DIRECTED, (True, False): REVERSED,
                (False, True): ACYCLIC, (True, True): REVERSED_ACYCLIC}[
                    (bool(reverse_edges), bool(remove_cycles))]
        return cls.load_file(ctx, path, n, weighted, mode, hashing)

    @classmethod
    def load_undirected(cls, ctx, path, n, weighted=False, hashing=HASH_BUCKET):
        return cls.load_file(ctx, path, n, weighted, UNDIRECTED, hashing)

    def free(self):
        if getattr(self, "h", None):
            self._L.la3_graph_free(self.h)
            self.h = None

    def __del__(self):
        self.free()

    def hash(self, v):
        return int(lib().la3_graph_hash(self.h, int(v)))

    def unhash(self, v):
        return int(lib().la3_graph_unhash(self.h, int(v)))

    def slots(self):
        out = np.zeros(self.nslots, dtype=np.uint32)
        self.ctx.check(lib().la3_graph_slots(self.h, _ptr(out)), "graph_slots")
        return out

    def set_tuning(self, what, value):
        self.ctx.check(lib().la3_graph_set_tuning(self.h, int(what), int(value)), "set_tuning")

    def get_tuning(self, what):
        v = ctypes.c_uint32()
        self.ctx.check(lib().la3_graph_get_tuning(self.h, int(what), ctypes.byref(v)), "get_tuning")
        return v.value

    def autotune(self, roots):
        """Picks by timing the SSSP near/far delta (single and batched) and,
        with >= 64 roots, the fused runs' task widths and blocks per CU."""
        r = np.ascontiguousarray(roots, dtype=np.uint32)
        rep = AutotuneReport()
        self.ctx.check(lib().la3_autotune(self.h, _ptr(r), r.size, ctypes.byref(rep)), "autotune")
        n = rep.ncandidates
        return dict(candidates=list(rep.candidates[:n]), single_ms=list(rep.single_ms[:n]),
                    batch_ms=list(rep.batch_ms[:n]), sssp_delta=rep.sssp_delta,
                    batch_sssp_delta=rep.batch_sssp_delta, batch_k=rep.batch_k,
                    task_k=list(rep.task_k), task_ms=[list(r) for r in rep.task_ms],
                    fused_blocks=rep.fused_blocks, fused_ms=list(rep.fused_ms))

    ARRAYS = ("col_ptr", "col_split", "col_rows", "col_w", "row_ptr", "row_split", "row_cols",
              "row_w", "id_of", "slot_of_id", "long_rows", "plong", "pseg_first", "pseg_row",
              "pseg_beg")

    def array(self, name):
        """One array of the stored layout (la3_graph_array)."""
        which = self.ARRAYS.index(name)
        n = ctypes.c_uint64()
        self.ctx.check(lib().la3_graph_array(self.h, which, None, 0, ctypes.byref(n)), "graph_array")
        out = np.zeros(n.value, dtype=np.uint32)
        self.ctx.check(lib().la3_graph_array(self.h, which, _ptr(out), n.value, ctypes.byref(n)),
                       "graph_array")
        return out

    def csc(self, which):
        nnz = self.info["nnz_sink"] if which else self.info["nnz_regular"]
        nc = self.info["ncols"]
        colptrs = np.zeros(nc + 1, dtype=np.uint32)
        colidxs = np.zeros(nc + 1, dtype=np.uint32)
        g = np.zeros(max(nnz, 1), dtype=np.uint32)
        i = np.zeros(max(nnz, 1), dtype=np.uint32)
        w = np.zeros(max(nnz, 1), dtype=np.uint32)
        self.ctx.check(lib().la3_graph_csc(self.h, int(which), _ptr(colptrs), _ptr(colidxs),
                                           _ptr(g), _ptr(i), _ptr(w)), "graph_csc")
        return dict(colptrs=colptrs, colidxs=colidxs, global_idx=g[:nnz], idx=i[:nnz], w=w[:nnz])


# ------------------------------------------------------------- programs ---
def degree(g, out=None):
    out = _out(out, (g.nslots,), np.uint32)
    g.ctx.check(lib().la3_degree(g.h, _ptr(out)), "degree")
    return out


def pagerank(g, iters, out=None):
    rank = _out(out, (g.nslots,), np.float64)
    deg = np.zeros(g.nslots, dtype=np.uint32) if out is not DEVICE else None
    it = np.zeros(1, dtype=np.uint32)
    g.ctx.check(lib().la3_pagerank(g.h, int(iters), _ptr(rank), _ptr(deg), _ptr(it)), "pagerank")
    return rank, deg, int(it[0])


def ppr(g, seed, iters, alpha=0.15, out=None):
    rank = _out(out, (g.nslots,), np.float64)
    g.ctx.check(lib().la3_ppr(g.h, int(seed), int(iters), float(alpha), _ptr(rank)), "ppr")
    return rank


def bfs(g, root, parents=True, out=None, max_iters=0):
    """BfsVertex execute(max_iters): (hops, parents or None, iterations)."""
    hops = _out(out, (g.nslots,), np.uint8)
    par = np.zeros(g.nslots, dtype=np.uint32) if parents else None
    it = np.zeros(1, dtype=np.uint32)
    g.ctx.check(lib().la3_bfs(g.h, int(root), int(max_iters), _ptr(hops), _ptr(par), _ptr(it)), "bfs")
    return hops, par, int(it[0])


def sssp(g, root, out=None, max_iters=0):
    """SpVertex execute(max_iters): (distances, iterations)."""
    dist = _out(out, (g.nslots,), np.uint32)
    it = np.zeros(1, dtype=np.uint32)
    g.ctx.check(lib().la3_sssp(g.h, int(root), int(max_iters), _ptr(dist), _ptr(it)), "sssp")
    return dist, int(it[0])


def cc(g, out=None, max_iters=0):
    """CcVertex execute(max_iters) on the vertex-program engine: (labels, iterations)."""
    label = _out(out, (g.nslots,), np.uint32)
    it = np.zeros(1, dtype=np.uint32)
    g.ctx.check(lib().la3_cc(g.h, int(max_iters), _ptr(label), _ptr(it)), "cc")
    return label, int(it[0])


def components(g, out=None):
    """Union-find components (min original id per component), undirected loads."""
    label = _out(out, (g.nslots,), np.uint32)
    g.ctx.check(lib().la3_components(g.h, _ptr(label)), "components")
    return label


def tc(g, per_vertex=True):
    total = np.zeros(1, dtype=np.uint64)
    per = np.zeros(g.nslots, dtype=np.uint32) if per_vertex else None
    g.ctx.check(lib().la3_tc(g.h, _ptr(total), _ptr(per)), "tc")
    return int(total[0]), per


def bfs_batch(g, roots, out=None, max_iters=0):
    roots = np.ascontiguousarray(roots, dtype=np.uint32)
    hops = _out(out, (roots.size, g.nslots), np.uint8)
    it = np.zeros(1, dtype=np.uint32)
    g.ctx.check(lib().la3_bfs_batch(g.h, _ptr(roots), roots.size, int(max_iters), _ptr(hops), _ptr(it)),
                "bfs_batch")
    return hops, int(it[0])


def sssp_batch(g, roots, out=None, max_iters=0):
    roots = np.ascontiguousarray(roots, dtype=np.uint32)
    dist = _out(out, (roots.size, g.nslots), np.uint32)
    it = np.zeros(1, dtype=np.uint32)
    g.ctx.check(lib().la3_sssp_batch(g.h, _ptr(roots), roots.size, int(max_iters), _ptr(dist), _ptr(it)),
                "sssp_batch")
    return dist, int(it[0])


def ppr_batch(g, seeds, iters, alpha=0.15, out=None, comp=False):
    """k PPR queries in FP32 with compensation. Returns (rank, comp or None,
    residual): rank[q, id] f32, comp[q, id] its compensation term, residual[q]
    the FP64 L1 change of the last iteration."""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    rank = _out(out, (seeds.size, g.nslots), np.float32)
    cmp = np.zeros((seeds.size, g.nslots), dtype=np.float32) if comp else None
    res = np.zeros(seeds.size, dtype=np.float64)
    g.ctx.check(lib().la3_ppr_batch(g.h, _ptr(seeds), seeds.size, int(iters), float(alpha),
                                    _ptr(rank), _ptr(cmp), _ptr(res)), "ppr_batch")
    return rank, cmp, res


def batch_run(ctx, batches, alpha=0.15, out=None):
    """Runs several batches as one fused, graph-captured traversal
    (la3_batch_run). `batches` is a list of (graph, kind, roots[, iters]) with
    kind Q_BFS / Q_SSSP / Q_PPR. Returns one dict per batch with the result
    ("out": k x nslots by id; None with out=DEVICE), for PPR also "comp" and
    "residual", and "iters". `out` may also be a list with one output array
    (or DEVICE) per batch."""
    arr = (Batch * len(batches))()
    keep = []
    res = []
    for i, b in enumerate(batches):
        g, kind, roots = b[0], b[1], np.ascontiguousarray(b[2], dtype=np.uint32)
        iters = int(b[3]) if len(b) > 3 else 0
        dt = {Q_BFS: np.uint8, Q_SSSP: np.uint32, Q_PPR: np.float32}[kind]
        o = _out(out[i] if isinstance(out, (list, tuple)) else out, (roots.size, g.nslots), dt)
        r = {"out": o}
        arr[i].graph = g.h
        arr[i].kind = kind
        arr[i].k = roots.size
        arr[i].roots = roots.ctypes.data
        arr[i].iters = iters
        arr[i].out = o.ctypes.data if o is not None else None
        if kind == Q_PPR:
            r["comp"] = np.zeros((roots.size, g.nslots), dtype=np.float32) if o is not None else None
            r["residual"] = np.zeros(roots.size, dtype=np.float64)
            arr[i].out_comp = r["comp"].ctypes.data if r["comp"] is not None else None
            arr[i].residual = r["residual"].ctypes.data
        keep.append(roots)
        res.append(r)
    ctx.check(lib().la3_batch_run(ctx.h, arr, len(batches), float(alpha)), "batch_run")
    for i, r in enumerate(res):
        r["iters"] = int(arr[i].iters_run)
    return res


class Session:
    """Continuous batching at step granularity (la3_session_*): batches join
    a running fused run with add(), advance with step(), and finished ones
    come back from poll()."""

    def __init__(self, ctx, alpha=0.15):
        self.ctx = ctx
        self._L = lib()
        h = ctypes.c_void_p()
        ctx.check(lib().la3_session_create(ctx.h, float(alpha), ctypes.byref(h)), "session_create")
        self.h = h
        self._keep = {}

    def add(self, g, kind, roots, iters=0, out=None):
        """Adds a batch; returns (id, result array or None). The array is
        filled when the batch finishes. Raises LA3Error(EAGAIN) if it does
        not fit the free task slots."""
        roots = np.ascontiguousarray(roots, dtype=np.uint32)
        dt = {Q_BFS: np.uint8, Q_SSSP: np.uint32, Q_PPR: np.float32}[kind]
        o = _out(out, (roots.size, g.nslots), dt)
        b = Batch()
        b.graph, b.kind, b.k, b.roots, b.iters = g.h, kind, roots.size, roots.ctypes.data, int(iters)
        b.out = o.ctypes.data if o is not None else None
        bid = ctypes.c_uint64()
        self.ctx.check(lib().la3_session_add(self.h, ctypes.byref(b), ctypes.byref(bid)), "session_add")
        self._keep[bid.value] = (roots, o)
        return bid.value, o

    def step(self):
        """Runs one replay of steps; returns the number of running batches."""
        n = ctypes.c_uint32()
        self.ctx.check(lib().la3_session_step(self.h, ctypes.byref(n)), "session_step")
        return n.value

    def poll(self, cap=256):
        """Finished batches: list of (id, iters_run, status)."""
        ids = np.zeros(cap, dtype=np.uint64)
        its = np.zeros(cap, dtype=np.uint32)
        sts = np.zeros(cap, dtype=np.int32)
        n = ctypes.c_uint32()
        self.ctx.check(lib().la3_session_poll(self.h, _ptr(ids), _ptr(its), _ptr(sts), cap, ctypes.byref(n)),
                       "session_poll")
        out = [(int(ids[i]), int(its[i]), int(sts[i])) for i in range(n.value)]
        for i, _, _ in out:
            self._keep.pop(i, None)
        return out

    def free_slots(self):
        return int(lib().la3_session_free_slots(self.h))

    def close(self):
        if getattr(self, "h", None):
            self._L.la3_session_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Server:
    """Serving loop (include/la3.h, "serving"): resident graphs, continuous
    batching by worker threads, admission by per-query state bytes."""

    def __init__(self, device=0, nworkers=2, max_batch=64, state_budget=32 << 30,
                 keep_results=False, ppr_alpha=0.15, fused=False, max_wait_s=0.05):
        """max_wait_s: batch order of the grouped mode (0: oldest query's group
        first; > 0: least expected work first, unless the oldest query has
        waited longer than max_wait_s seconds)."""
        cfg = ServerConfig(nworkers, max_batch, state_budget, int(keep_results), ppr_alpha,
                           int(fused), float(max_wait_s))
        h = ctypes.c_void_p()
        self._L = lib()
        st = lib().la3_server_create(device, ctypes.byref(cfg), ctypes.byref(h))
        if st:
            raise LA3Error(st, "la3_server_create")
        self.h = h
        self.graphs = []

    def add_graph(self, g):
        gid = ctypes.c_uint32()
        st = lib().la3_server_add_graph(self.h, g.h, ctypes.byref(gid))
        if st:
            raise LA3Error(st, "la3_server_add_graph")
        self.graphs.append(g)  # keeps the graph alive while served
        return gid.value

    def warm(self):
        """Every worker runs each program on each graph once, single and
        max_batch wide (la3_server_warm): scratch at working size before a
        stream."""
        st = lib().la3_server_warm(self.h)
        if st:
            raise LA3Error(st, "la3_server_warm")

    def query_bytes(self, kind, gid):
        return int(lib().la3_server_query_bytes(self.h, kind, gid))

    def submit(self, kind, gid, root, iters=0):
        q = ctypes.c_uint64()
        st = lib().la3_server_submit(self.h, kind, gid, int(root), int(iters), ctypes.byref(q))
        if st:
            raise LA3Error(st, "la3_server_submit")
        return q.value

    def clock(self):
        return lib().la3_server_clock(self.h)

    def drain(self):
        lib().la3_server_drain(self.h)

    def completed(self, cap=4096):
        out, arr, n = [], (QueryInfo * cap)(), ctypes.c_uint32()
        while True:
            lib().la3_server_completed(self.h, arr, cap, ctypes.byref(n))
            out += [{f: getattr(a, f) for f, _ in QueryInfo._fields_} for a in arr[:n.value]]
            if n.value < cap:
                return out

    def result(self, qid, kind, nslots):
        dt = {Q_BFS: np.uint8, Q_SSSP: np.uint32, Q_PPR: np.float32}[kind]
        out = np.zeros(nslots, dtype=dt)
        st = lib().la3_server_result(self.h, qid, _ptr(out), out.nbytes)
        if st:
            raise LA3Error(st, f"la3_server_result({qid})")
        return out

    def close(self):
        if getattr(self, "h", None):
            self._L.la3_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
