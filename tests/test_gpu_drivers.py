"""The C++ drop-in (include/la3.hpp) through la3_amd/bin/la3_run, the
counterpart of the reference's app drivers: same arguments, same output
lines, and the g1 checksums the reference prints at np=1."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "la3_amd", "bin", "la3_run")


@pytest.fixture(scope="module")
def g1_files(tmp_path_factory, g1):
    d = tmp_path_factory.mktemp("g1")
    p, pw = d / "g1_8_8_13.bin", d / "g1_8_8_13.w.bin"
    g1["records"].tofile(p)
    g1["records_w"].tofile(pw)
    return str(p), str(pw)


def run(*args):
    out = subprocess.run([RUN, *map(str, args)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    return out.stdout


def test_pr_driver(g1_files, g1):
    out = run("pr", g1_files[0], 8, 20)
    assert f"Pagerank Checksum = {g1['pr_checksum']:f}" in out
    assert f"Degree Checksum = {g1['degree_checksum']}" in out
    assert "Timer Pagerank Execution stats:" in out


def test_sssp_driver(g1_files, g1):
    out = run("sssp", g1_files[1], 8, 5)
    assert f"Reachable Vertices = {g1['sssp_reachable']}" in out
    assert f"Checksum = {g1['sssp_checksum']}" in out


def test_bfs_driver(g1_files, g1):
    out = run("bfs", g1_files[0], 8, 5)
    assert f"Reachable Vertices = {g1['bfs_reachable']}" in out
    assert f"Checksum = {g1['bfs_checksum']}" in out


def test_cc_tc_degree_drivers(g1_files, g1):
    assert f"Checksum = {g1['cc_checksum']}" in run("cc", g1_files[0], 8)
    assert f"Triangles = {g1['tc_total']}" in run("tc", g1_files[0], 8)
    assert "Out-Degree Checksum = 13" in run("degree", g1_files[0], 8, "-o")


def test_degree_driver_top3(g1_files, g1):
    """degree.cpp:34-43: Top-3 by in-degree (std::partial_sort over the
    states in slot order; equal degrees may come in either order)."""
    out = run("degree", g1_files[0], 8)
    top = [(int(a), int(b)) for a, b in re.findall(r"idx (\d+): degree (\d+)", out)]
    assert "Top-3:" in out and len(top) == 3
    rec = g1["records"]
    indeg = np.bincount(rec[:, 1], minlength=9)
    assert [d for _, d in top] == sorted(indeg, reverse=True)[:3]
    assert all(indeg[i] == d for i, d in top)



def test_ppr_batch_btopk_driver(tmp_path, gpu_ctx):
    """la3_run ppr_topk: four PPR queries as one batch through
    la3::PprBatchVertex<4>, per-query top-k by btopk; each printed (id, rank)
    is that query's rank from la3_ppr_batch, the list is in descending rank
    order and holds the k largest ranks."""
    import la3_amd as L
    rec = L.rmat(12, 16, seed=81)
    n = 1 << 12
    path = tmp_path / "g.bin"
    rec.tofile(path)
    src = np.unique(rec[:, 0])
    seeds = [int(x) for x in np.random.default_rng(81).choice(src, size=4, replace=False)]
    k = 10
    out = run("ppr_topk", path, n, 20, k, *seeds)
    g = L.Graph.from_records(gpu_ctx, rec, n, False, L.DIRECTED)
    rank, _, _ = L.ppr_batch(g, seeds, 20)
    blocks = re.split(r"Query \d+ \(seed \d+\) top-\d+:", out)[1:]
    assert len(blocks) == 4
    for q, blk in enumerate(blocks):
        top = [(int(a), float(b)) for a, b in re.findall(r"idx (\d+): rank (\S+)", blk)]
        assert len(top) == k
        vals = [v for _, v in top]
        assert vals == sorted(vals, reverse=True)
        for i, v in top:
            assert np.float32(v) == rank[q][i]
        assert np.float32(vals[-1]) == np.sort(rank[q])[::-1][k - 1]


def _display(out):
    """The display() lines of a driver's output, in order."""
    return [ln for ln in out.splitlines() if re.match(r"^\d+: \{", ln)]


def test_drivers_display_g1(g1_files, g1):
    """Every driver prints display() before its checksums, as the
    reference's (bfs.cpp:25, sssp.cpp:25, cc.cpp:24, pr.cpp:42,
    degree.cpp:28-33): ids 0..min(30, n) in the state vector's slot order
    (regular, sink, source, isolated; hashed order is the id order at n = 8)
    with the states' to_string formats. Expected lines from the g1 known
    answers."""
    cls = g1["classes_directed"]
    directed = cls["regular"] + cls["sink"] + cls["source"] + cls["isolated"]
    undirected = [1, 2, 3, 4, 5, 6, 8, 0, 7]  # every vertex with an edge is regular
    sssp = {int(k): v for k, v in g1["sssp_root5"].items()}
    out = run("sssp", g1_files[1], 8, 5)
    assert _display(out) == [f"{i}: {{distance: {sssp.get(i, 2147483647)}}}" for i in directed]
    hops = {int(k): v for k, v in g1["bfs_undirected_root5_hops"].items()}
    par = {int(k): v for k, v in g1["bfs_undirected_root5_parent"].items()}
    out = run("bfs", g1_files[0], 8, 5)
    assert _display(out) == [f"{i}: {{parent: {par.get(i, 0)}, hops: {hops[i]}}}" for i in undirected]
    lab = {int(k): v for k, v in g1["cc_labels"].items()}
    out = run("cc", g1_files[0], 8)
    assert _display(out) == [f"{i}: {{label: {lab[i]}}}" for i in undirected]
    rank = {int(k): v for k, v in g1["pr_20"].items()}
    deg = {int(k): v for k, v in g1["out_degree"].items()}
    out = run("pr", g1_files[0], 8, 20)
    shown = [re.match(r"^(\d+): \{rank: (\d+\.\d{6}), degree: (\d+)\}$", ln) for ln in _display(out)]
    assert all(shown) and [int(m.group(1)) for m in shown] == directed
    for m in shown:  # "%f" of the FP64 rank: within rounding of the 6th decimal
        i = int(m.group(1))
        assert abs(float(m.group(2)) - rank[i]) <= 5.1e-7 and int(m.group(3)) == deg.get(i, 0)
    indeg = {int(k): v for k, v in g1["in_degree"].items()}
    out = run("degree", g1_files[0], 8)
    assert "In-Degrees:" in out
    assert _display(out) == [f"{i}: {{degree: {indeg.get(i, 0)}}}" for i in directed]
