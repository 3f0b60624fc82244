"""CPU checks of the C ABI: the library loads and exports every entry point
include/la3.h declares; the host-side ingest (host-only context) builds the
same classes, slot order and DCSC as the oracle; the R-MAT tool is
deterministic; programs refuse to run without a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import la3_amd as L
from test_oracle import random_records

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "la3.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(la3_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(L.LIB_PATH)
    declared = header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/la3.h but not exported"
    assert sorted(L.EXPORTS) == declared


@pytest.fixture(scope="module")
def host_ctx():
    ctx = L.Context(-1)
    yield ctx
    ctx.close()


def _compare_layout(host_ctx, oracle, rec, n, weighted, mode):
    g = L.Graph.from_records(host_ctx, rec, n, weighted, mode)
    o = oracle.OracleGraph(rec, n, weighted, mode)
    assert (g.info["nregular"], g.info["nsink"], g.info["nsource"]) == (o.nreg, o.nsink, o.nsrc)
    assert g.info["ncols"] == o.ncols
    assert (g.info["nnz_regular"], g.info["nnz_sink"]) == (o.nnz_reg, o.nnz_sink)
    assert (g.slots() == o.hid_of()).all()
    for which in (0, 1):
        a, b = g.csc(which), o.csc(which)
        for key in ("colptrs", "global_idx", "idx", "w"):
            assert (a[key] == b[key]).all(), (which, key)
        assert (a["colidxs"][:-1] == b["colidxs"][:-1]).all()
    for v in (0, 1, n // 2, n):
        assert g.hash(v) == o.hash(v) and g.unhash(g.hash(v)) == v


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_ingest_layout_matches_oracle_g1(host_ctx, oracle, g1, mode):
    _compare_layout(host_ctx, oracle, g1["records"], 8, False, mode)
    _compare_layout(host_ctx, oracle, g1["records_w"], 8, True, mode)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_ingest_layout_matches_oracle_random(host_ctx, oracle, mode):
    # duplicates carry different weights: the first record must win
    rec = random_records(31 + mode, 5000, 40000, weighted=True)
    _compare_layout(host_ctx, oracle, rec, 5000, True, mode)


def test_ingest_layout_matches_oracle_rmat(host_ctx, oracle):
    rec = L.rmat(14, 16, seed=5, weighted=True)
    _compare_layout(host_ctx, oracle, rec, 1 << 14, True, L.DIRECTED)


def test_ingest_hashing_none(host_ctx, oracle):
    rec = random_records(40, 1000, 5000)
    g = L.Graph.from_records(host_ctx, rec, 1000, False, L.DIRECTED, hashing=L.HASH_NONE)
    o = oracle.OracleGraph(rec, 1000, False, oracle.DIRECTED, hashing=0)
    assert (g.slots() == o.hid_of()).all()


def test_load_file_and_header(host_ctx, tmp_path, oracle, g1):
    p = tmp_path / "g1.w.bin"
    g1["records_w"].tofile(p)
    g = L.Graph.load_directed(host_ctx, str(p), 8, weighted=True)
    assert g.info["nnz"] == 13 and g.info["nrecords"] == 13
    # 16-byte header {n-1, m-1, nnz}: nvertices = 0 reads it
    h = tmp_path / "g1.h.w.bin"
    with open(h, "wb") as f:
        f.write(np.array([8, 8], dtype=np.uint32).tobytes())
        f.write(np.array([13], dtype=np.uint64).tobytes())
        f.write(g1["records_w"].tobytes())
    gh = L.Graph.load_directed(host_ctx, str(h), 0, weighted=True)
    assert gh.n == 9 and gh.info["nnz"] == 13
    with pytest.raises(L.LA3Error) as e:
        L.Graph.load_directed(host_ctx, str(tmp_path / "missing.bin"), 8)
    assert e.value.status == L.EIO


def test_ingest_rejects_out_of_range(host_ctx):
    with pytest.raises(L.LA3Error) as e:
        L.Graph.from_records(host_ctx, np.array([[1, 9]], dtype=np.uint32), 8)
    assert e.value.status == L.EINVAL


def test_programs_need_a_device(host_ctx, g1):
    g = L.Graph.from_records(host_ctx, g1["records"], 8)
    with pytest.raises(L.LA3Error) as e:
        L.bfs(g, 1)
    assert e.value.status == L.EDEVICE


def test_rmat_deterministic_and_clean():
    a = L.rmat(12, 16, seed=3, weighted=True)
    b = L.rmat(12, 16, seed=3, weighted=True)
    c = L.rmat(12, 16, seed=4, weighted=True)
    assert (a == b).all() and a.shape != c.shape or not (a == c).all()
    assert (a[:, 0] != a[:, 1]).all()
    assert a[:, :2].min() >= 1 and a[:, :2].max() <= 1 << 12
    key = a[:, 0].astype(np.int64) << 32 | a[:, 1]
    assert (np.diff(key) > 0).all()  # sorted and unique
    assert a[:, 2].min() >= 1 and a[:, 2].max() <= 128
    # R-MAT skew: vertex 1 (quadrant a at every level) is the largest hub
    deg = np.bincount(a[:, 0])
    assert deg.argmax() == 1
    # edges kept after dedup/self-loop removal: about 0.9 of the samples at scale 12
    assert 0.8 * 16 * 4096 < a.shape[0] < 16 * 4096
