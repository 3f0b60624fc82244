"""la3_bfs_noalloc (the allocation-free variant of la3_bfs_ctx) through the C
ABI against la3_bfs on the same graphs and roots: two calls in a row, the
second on a smaller graph (the shared work array keeps the first call's
contents), a bounded call, and the size limit. Output buffers are device
memory the caller allocated (hipMalloc through ctypes here)."""
import ctypes

import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu


_hip = None


def hip():
    """The HIP runtime the library links, for the caller-side device buffers."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return _hip


class DevBuf:
    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(nbytes, 1)) == 0
        assert hip().hipMemset(self.p, 7, max(nbytes, 1)) == 0
        self.n = nbytes

    def to_host(self, dtype):
        out = np.zeros(self.n // np.dtype(dtype).itemsize, dtype=dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, self.n, 2) == 0  # device to host
        return out

    def __del__(self):
        hip().hipFree(self.p)


def _noalloc(ctx, g, root, max_iters=0, parents=True):
    S = g.nslots
    hops = DevBuf(S)
    par = DevBuf(4 * S) if parents else None
    it = ctypes.c_uint32(0)
    st = L.lib().la3_bfs_noalloc(ctx.h, g.h, int(root), int(max_iters), hops.p, par.p if parents else None,
                                 ctypes.byref(it))
    if st:
        raise L.LA3Error(st, L.lib().la3_last_error(ctx.h).decode())
    return hops.to_host(np.uint8), (par.to_host(np.uint32) if parents else None), it.value


def test_noalloc_equals_bfs_two_calls_then_smaller(gpu_ctx):
    big = L.rmat(15, 16, seed=51)
    small = L.rmat(11, 16, seed=52)
    for rec, n in ((big, 1 << 15), (small, 1 << 11)):
        g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
        for root in [int(x) for x in np.unique(rec[:, 0])[:3]]:
            eh, ep, eit = L.bfs(g, root)
            h, p, it = _noalloc(gpu_ctx, g, root)
            assert np.array_equal(h, eh) and np.array_equal(p, ep) and it == eit, (n, root)
            eh2, _, eit2 = L.bfs(g, root, max_iters=2)
            h2, _, it2 = _noalloc(gpu_ctx, g, root, max_iters=2, parents=False)
            assert np.array_equal(h2, eh2) and it2 == eit2 == 2


def test_noalloc_undirected_and_limit(gpu_ctx):
    rec = L.rmat(12, 16, seed=53)
    g = L.Graph.from_records(gpu_ctx, rec, 1 << 12, False, L.UNDIRECTED)
    root = int(rec[0, 0])
    eh, ep, eit = L.bfs(g, root)
    h, p, it = _noalloc(gpu_ctx, g, root)
    assert np.array_equal(h, eh) and np.array_equal(p, ep) and it == eit
    # 4M vertex slots need about 84 MB of work buffers: over the 64 MiB limit
    big = L.Graph.from_records(gpu_ctx, np.zeros((0, 2), dtype=np.uint32), 4_000_000, False, L.DIRECTED)
    with pytest.raises(L.LA3Error) as e:
        _noalloc(gpu_ctx, big, 1)
    assert e.value.status == L.ENOTSUP
