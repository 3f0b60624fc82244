"""Device ingest (la3_graph_load_* on a device context, the default) against
the host builder (the same call on a host-only context, itself checked
byte-for-byte against the oracle's layout in test_capi.py): every array of
the stored layout and every count must be identical, for all five load
modes, weighted and unweighted records with duplicates (the first record
wins) and self-loops, bucket and identity hashing."""
import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def host_ctx():
    c = L.Context(-1)
    yield c
    c.close()


def _records(scale, seed, weighted):
    rng = np.random.default_rng(seed)
    rec = L.rmat(scale, 16, seed=seed, weighted=weighted).copy()
    # duplicates with other weights, and self-loops
    dup = rec[rng.choice(rec.shape[0], size=rec.shape[0] // 10)].copy()
    if weighted:
        dup[:, 2] = rng.integers(1, 129, size=dup.shape[0])
    loops = np.repeat(rng.integers(0, 1 << scale, size=(50, 1)), 2, axis=1).astype(np.uint32)
    if weighted:
        loops = np.hstack([loops, np.ones((50, 1), dtype=np.uint32)])
    allr = np.vstack([rec, dup, loops])
    return np.ascontiguousarray(allr[rng.permutation(allr.shape[0])])


@pytest.mark.parametrize("mode", [L.DIRECTED, L.REVERSED, L.UNDIRECTED, L.ACYCLIC, L.REVERSED_ACYCLIC])
@pytest.mark.parametrize("weighted,hashing,scale", [(True, L.HASH_BUCKET, 12), (False, L.HASH_NONE, 11),
                                                   (False, L.HASH_BUCKET, 14)])
def test_device_ingest_equals_host(gpu_ctx, host_ctx, mode, weighted, hashing, scale):
    rec = _records(scale, 7 + scale + mode, weighted)
    n = (1 << scale) + 3
    gd = L.Graph.from_records(gpu_ctx, rec, n, weighted, mode, hashing)
    gh = L.Graph.from_records(host_ctx, rec, n, weighted, mode, hashing)
    assert gd.info == {**gh.info, "device_bytes": gd.info["device_bytes"]}
    assert np.array_equal(gd.slots(), gh.slots())
    for name in L.Graph.ARRAYS:
        a, b = gd.array(name), gh.array(name)
        assert np.array_equal(a, b), f"{name} differs ({a.size} vs {b.size} entries)"


def test_host_ingest_switch(gpu_ctx, host_ctx):
    rec = _records(10, 3, True)
    gpu_ctx.set_ingest(False)
    try:
        g = L.Graph.from_records(gpu_ctx, rec, 1 << 10, True, L.DIRECTED)
    finally:
        gpu_ctx.set_ingest(True)
    gd = L.Graph.from_records(gpu_ctx, rec, 1 << 10, True, L.DIRECTED)
    for name in L.Graph.ARRAYS:
        assert np.array_equal(g.array(name), gd.array(name)), name
    root = int(rec[0, 0])
    assert np.array_equal(L.sssp(g, root)[0], L.sssp(gd, root)[0])


def test_device_ingest_edge_cases(gpu_ctx, host_ctx):
    empty = np.zeros((0, 2), dtype=np.uint32)
    gd = L.Graph.from_records(gpu_ctx, empty, 10, False, L.DIRECTED)
    assert gd.info["nnz"] == 0 and gd.info["nisolated"] == 11
    assert np.array_equal(L.bfs(gd, 3)[0][3:4], [0])
    with pytest.raises(L.LA3Error):
        L.Graph.from_records(gpu_ctx, np.array([[1, 99]], dtype=np.uint32), 10, False, L.DIRECTED)
    tiny = np.array([[1, 2], [2, 1], [2, 3], [3, 3]], dtype=np.uint32)
    gd = L.Graph.from_records(gpu_ctx, tiny, 3, False, L.DIRECTED)
    gh = L.Graph.from_records(host_ctx, tiny, 3, False, L.DIRECTED)
    for name in L.Graph.ARRAYS:
        assert np.array_equal(gd.array(name), gh.array(name)), name
