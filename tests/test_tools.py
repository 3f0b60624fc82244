"""Format tools (the reference's src/tools/mtx2bin.cpp and bin2mtx.cpp,
rebuilt as la3_amd/bin/la3_mtx2bin and la3_bin2mtx): text <-> LA3 binary
records, optional 16-byte header, u32 / f64 weights, and the random weights
1 + rand() % 128 after srand(0) checked against the C library's own rand()."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "la3_amd", "bin")


def tool(name, *args):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.skip(f"{p} not built")
    out = subprocess.run([p, *map(str, args)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    return out.stdout


MTX = """%%MatrixMarket matrix coordinate integer general
% a comment
8 8 5
1 2
2 3
3 1
4 8
8 5
"""


def test_mtx2bin_pairs_header_and_rand_weights(tmp_path):
    src = tmp_path / "g.mtx"
    src.write_text(MTX)
    out = tmp_path / "g.bin"
    tool("la3_mtx2bin", src, out, "-hio", "-woi")
    raw = out.read_bytes()
    n, m = np.frombuffer(raw[:8], dtype=np.uint32)
    nnz = np.frombuffer(raw[8:16], dtype=np.uint64)[0]
    assert (n, m, nnz) == (8, 8, 5)
    rec = np.frombuffer(raw[16:], dtype=np.uint32).reshape(-1, 3)
    assert rec[:, :2].tolist() == [[1, 2], [2, 3], [3, 1], [4, 8], [8, 5]]
    libc = ctypes.CDLL(None)
    libc.srand(0)
    assert rec[:, 2].tolist() == [1 + libc.rand() % 128 for _ in range(5)]


def test_mtx2bin_weights_in_and_doubles(tmp_path):
    src = tmp_path / "w.txt"
    src.write_text("1 2 7\n2 3 9.5\n")
    out = tmp_path / "w.bin"
    tool("la3_mtx2bin", src, out, "-wi", "-woi")
    rec = np.fromfile(out, dtype=np.uint32).reshape(-1, 3)
    assert rec.tolist() == [[1, 2, 7], [2, 3, 9]]
    tool("la3_mtx2bin", src, out, "-wi", "-wod")
    raw = out.read_bytes()
    assert len(raw) == 2 * 16
    assert np.frombuffer(raw[8:16], dtype=np.float64)[0] == 7.0
    assert np.frombuffer(raw[24:32], dtype=np.float64)[0] == 9.5


def test_bin2mtx_round_trip(tmp_path):
    rec = np.array([[1, 2, 5], [3, 4, 6], [4, 1, 128]], dtype=np.uint32)
    b = tmp_path / "r.bin"
    hdr = np.array([4, 4], dtype=np.uint32).tobytes() + np.array([3], dtype=np.uint64).tobytes()
    b.write_bytes(hdr + rec.tobytes())
    t = tmp_path / "r.txt"
    tool("la3_bin2mtx", b, t, "-hio", "-wi", "-wo")
    lines = t.read_text().split("\n")
    assert lines[0] == "4 4 3"
    assert lines[1:4] == ["1 2 5", "3 4 6", "4 1 128"]
    b2 = tmp_path / "r2.bin"
    tool("la3_mtx2bin", t, b2, "-hio", "-wi", "-woi")
    assert b2.read_bytes() == b.read_bytes()
    tool("la3_bin2mtx", b, t, "-hi", "-wi")
    assert t.read_text().split("\n")[:3] == ["1 2", "3 4", "4 1"]


@pytest.mark.parametrize("weighted,header", [(False, False), (True, True)])
def test_rmat_tool_writes_the_generator_records(tmp_path, weighted, header):
    """la3_rmat writes la3_rmat_generate's records (the same as la3_amd.rmat)
    as an LA3 binary file, pairs or weighted triples, with the optional
    16-byte header (u32 n, u32 n, u64 nedges), readable by the loader."""
    import la3_amd as L
    out = tmp_path / "g.bin"
    args = [10, 8, 5, out] + (["-w"] if weighted else []) + (["-ho"] if header else [])
    line = tool("la3_rmat", *args).split()
    rec = L.rmat(10, 8, seed=5, weighted=weighted)
    assert line == [str(1 << 10), str(rec.shape[0])]
    raw = out.read_bytes()
    if header:
        hdr = np.frombuffer(raw[:16], dtype=np.uint32)
        assert hdr[0] == 1 << 10 and hdr[1] == 1 << 10
        assert int(np.frombuffer(raw[8:16], dtype=np.uint64)[0]) == rec.shape[0]
        raw = raw[16:]
    got = np.frombuffer(raw, dtype=np.uint32).reshape(-1, 3 if weighted else 2)
    assert (got == rec).all()
    assert got[:, :2].min() >= 1 and got[:, :2].max() <= 1 << 10 and (got[:, 0] != got[:, 1]).all()
    if weighted:
        assert got[:, 2].min() >= 1 and got[:, 2].max() <= 128
    # the loader reads it; on the header path n is the header's count plus
    # one, as the reference's loader takes it (graph.hpp:99-100)
    ctx = L.Context(-1)
    g = L.Graph.load_file(ctx, str(out), 0 if header else 1 << 10, weighted, L.DIRECTED)
    assert g.info["nrecords"] == rec.shape[0] and g.info["n"] == (1 << 10) + (1 if header else 0)
