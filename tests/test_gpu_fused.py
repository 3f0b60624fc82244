"""Fused, graph-captured batch runs (la3_batch_run): batches of BFS, SSSP and
PPR on several graphs advance as one traversal. Every query's result must be
the single-family batch call's result: bit-exact for BFS levels and SSSP
distances, and for PPR ranks and compensation terms (the row sums run in the
same order whatever block a row lands on); the FP64 residual is a sum of
per-block partials whose grouping depends on the block split, so it is
compared to 1e-12 relative. The single-family batch calls themselves are
pinned to the single-query programs and the oracle in test_gpu_parity.py;
here a few roots are also checked against the oracle directly."""
import numpy as np
import pytest

import la3_amd as L

pytestmark = pytest.mark.gpu


def _graph(ctx, scale, seed, weighted=True, mode=L.DIRECTED):
    rec = L.rmat(scale, 16, seed=seed, weighted=weighted)
    g = L.Graph.from_records(ctx, rec, 1 << scale, weighted, mode)
    return g, rec


def _roots(rec, k, seed, extra=()):
    rng = np.random.default_rng(seed)
    src = np.unique(rec[:, 0])
    r = rng.choice(src, size=k - len(extra), replace=True).astype(np.uint32)
    return np.concatenate([np.asarray(extra, dtype=np.uint32), r])


@pytest.fixture(scope="module")
def two_graphs(gpu_ctx):
    return _graph(gpu_ctx, 12, 71), _graph(gpu_ctx, 11, 72)


def _check(batches, res):
    for (g, kind, roots, *rest), r in zip(batches, res):
        if kind == L.Q_BFS:
            ref, it = L.bfs_batch(g, roots)
            assert np.array_equal(r["out"], ref)
            assert r["iters"] == it
        elif kind == L.Q_SSSP:
            # distances exact; the step count of an unbounded run is not
            # compared: a push reads distances that other pushes lowered
            # earlier in the same step (asynchronous relaxation), so how many
            # steps the last improvement takes depends on the launch's
            # block schedule, which differs between the two runs
            ref, it = L.sssp_batch(g, roots)
            assert np.array_equal(r["out"], ref)
            assert r["iters"] >= 1
        else:
            rank, comp, resid = L.ppr_batch(g, roots, rest[0], comp=True)
            assert np.array_equal(r["out"], rank)
            assert np.array_equal(r["comp"], comp)
            np.testing.assert_allclose(r["residual"], resid, rtol=1e-12, atol=0)
            assert r["iters"] == rest[0]


def test_mixed_two_graphs(gpu_ctx, two_graphs):
    (ga, ra), (gb, rb) = two_graphs
    # id 0 is isolated (ids are 1-based), id 2^scale is likely a sink or
    # isolated; repeated roots share a frontier entry
    sink = ga.nslots - 1
    batches = [
        (ga, L.Q_BFS, _roots(ra, 70, 1, extra=(0, 0, sink))),
        (ga, L.Q_SSSP, _roots(ra, 130, 2, extra=(0, sink, sink))),
        (gb, L.Q_PPR, _roots(rb, 5, 3), 10),
        (gb, L.Q_BFS, _roots(rb, 300, 4)),      # two tasks: 256 + 44
        (gb, L.Q_SSSP, _roots(rb, 1, 5)),
    ]
    res = L.batch_run(gpu_ctx, batches)
    _check(batches, res)


def test_repeated_runs_reuse_the_graph(gpu_ctx, two_graphs):
    (ga, ra), (gb, rb) = two_graphs
    first = [(ga, L.Q_SSSP, _roots(ra, 64, 11)), (gb, L.Q_PPR, _roots(rb, 64, 12), 7)]
    second = [(gb, L.Q_BFS, _roots(rb, 9, 13))]
    for batches in (first, second, first):
        _check(batches, L.batch_run(gpu_ctx, batches))


def test_near_far_threshold_in_fused_run(gpu_ctx, two_graphs):
    (ga, ra), _ = two_graphs
    roots = _roots(ra, 100, 21)
    want, _ = L.sssp_batch(ga, roots)
    for delta in (0, 16, 200):
        ga.set_tuning(L.TUNE_BATCH_SSSP_DELTA, delta)
        res = L.batch_run(gpu_ctx, [(ga, L.Q_SSSP, roots)])
        assert np.array_equal(res[0]["out"], want)
    ga.set_tuning(L.TUNE_BATCH_SSSP_DELTA, 0)


def test_fused_against_oracle(gpu_ctx, oracle):
    rec = L.rmat(10, 8, seed=31, weighted=True)
    n = 1 << 10
    g = L.Graph.from_records(gpu_ctx, rec, n, True, L.DIRECTED)
    og = oracle.OracleGraph(rec, n, True, L.DIRECTED)
    roots = _roots(rec, 6, 32)
    res = L.batch_run(gpu_ctx, [(g, L.Q_BFS, roots), (g, L.Q_SSSP, roots)])
    for q, r in enumerate(roots):
        hops = oracle.bfs(og, int(r))[0]
        dist = oracle.sssp(og, int(r))[0]
        assert np.array_equal(res[0]["out"][q], hops)
        assert np.array_equal(res[1]["out"][q], dist)


def test_undirected_bfs_and_errors(gpu_ctx):
    g, rec = _graph(gpu_ctx, 10, 41, weighted=False, mode=L.UNDIRECTED)
    roots = _roots(rec, 20, 42)
    _check([(g, L.Q_BFS, roots)], L.batch_run(gpu_ctx, [(g, L.Q_BFS, roots)]))
    with pytest.raises(L.LA3Error):
        L.batch_run(gpu_ctx, [(g, L.Q_SSSP, roots)])     # unweighted, undirected
    gd, recd = _graph(gpu_ctx, 10, 43)
    with pytest.raises(L.LA3Error):
        L.batch_run(gpu_ctx, [(gd, L.Q_PPR, roots, 0)])  # PPR needs iterations
    with pytest.raises(L.LA3Error):
        L.batch_run(gpu_ctx, [(gd, L.Q_BFS, np.array([(1 << 10) + 5], dtype=np.uint32))])


def test_session_join_while_running(gpu_ctx, two_graphs):
    """Batches added while others run (step-level continuous batching) get
    the same results as alone."""
    (ga, ra), (gb, rb) = two_graphs
    s = L.Session(gpu_ctx)
    r1 = _roots(ra, 64, 51)
    id1, o1 = s.add(ga, L.Q_SSSP, r1)
    s.step()
    r2 = _roots(rb, 40, 52)
    id2, o2 = s.add(gb, L.Q_BFS, r2)
    r3 = _roots(rb, 3, 53)
    id3, o3 = s.add(gb, L.Q_PPR, r3, 6)
    done = {}
    for _ in range(10000):
        for i, it, st in s.poll():
            assert st == 0
            done[i] = it
        if len(done) == 3:
            break
        s.step()
    assert len(done) == 3
    assert np.array_equal(o1, L.sssp_batch(ga, r1)[0])
    assert np.array_equal(o2, L.bfs_batch(gb, r2)[0])
    assert np.array_equal(o3, L.ppr_batch(gb, r3, 6)[0])
    assert done[id3] == 6
    # the context holds one session at a time
    with pytest.raises(L.LA3Error):
        L.batch_run(gpu_ctx, [(gb, L.Q_BFS, r2)])
    s.close()
    L.batch_run(gpu_ctx, [(gb, L.Q_BFS, r2)])


def test_session_sub_steps_on_small_graphs(gpu_ctx):
    """A session running tasks on a graph 32x smaller than the largest
    running one gives them sub-steps (several iterations per step of the
    large tasks): they finish in fewer session steps than the large PPR
    batch, and every result equals the batched call's."""
    gl, rl = _graph(gpu_ctx, 15, 81)
    gs, rs = _graph(gpu_ctx, 10, 82)
    s = L.Session(gpu_ctx)
    batches = [
        (gl, L.Q_PPR, _roots(rl, 3, 83), 12),
        (gl, L.Q_SSSP, _roots(rl, 20, 84)),
        (gs, L.Q_BFS, _roots(rs, 70, 85)),
        (gs, L.Q_SSSP, _roots(rs, 5, 86)),
        (gs, L.Q_PPR, _roots(rs, 2, 87), 25),
        (gs, L.Q_SSSP, _roots(rs, 100, 88)),
    ]
    ids, outs = [], []
    for g, kind, roots, *rest in batches:
        i, o = s.add(g, kind, roots, *rest)
        ids.append(i)
        outs.append(o)
    finished, step = {}, 0
    for _ in range(10000):
        for i, it, st in s.poll():
            assert st == 0
            finished[i] = (step, it)
        if len(finished) == len(batches):
            break
        s.step()
        step += 1
    assert len(finished) == len(batches)
    for (g, kind, roots, *rest), i, o in zip(batches, ids, outs):
        if kind == L.Q_BFS:
            assert np.array_equal(o, L.bfs_batch(g, roots)[0])
        elif kind == L.Q_SSSP:
            assert np.array_equal(o, L.sssp_batch(g, roots)[0])
        else:
            assert np.array_equal(o, L.ppr_batch(g, roots, rest[0])[0])
            assert finished[i][1] == rest[0]
    big_ppr = finished[ids[0]][0]
    assert finished[ids[0]][0] >= 12
    for i in ids[2:]:
        assert finished[i][0] < big_ppr, "small-graph batches should finish first"
    s.close()


def test_session_slots_exhausted(gpu_ctx, two_graphs):
    _, (gb, rb) = two_graphs
    s = L.Session(gpu_ctx)
    added = 0
    with pytest.raises(L.LA3Error):
        for i in range(17):           # 16 batches x 4 tasks fill the 64 slots
            s.add(gb, L.Q_BFS, _roots(rb, 1024, 60 + i))
            added += 1
    assert added == 16 and s.free_slots() == 0
    s.close()


def test_task_width_and_grid_keep_results(gpu_ctx, two_graphs):
    """Task widths 64 / 128 / 256 and 8 / 16 / 32 blocks per CU (the
    autotuner's choices) change the split of the work, never the results."""
    (ga, ra), _ = two_graphs
    roots = _roots(ra, 200, 81)
    want_b, _ = L.bfs_batch(ga, roots)
    want_s, _ = L.sssp_batch(ga, roots)
    want_p, _, _ = L.ppr_batch(ga, roots[:70], 4)
    for k in (64, 128, 256):
        for blocks in (8, 32):
            for t in (L.TUNE_BFS_TASK_K, L.TUNE_SSSP_TASK_K, L.TUNE_PPR_TASK_K):
                ga.set_tuning(t, k)
            ga.set_tuning(L.TUNE_FUSED_BLOCKS, blocks)
            res = L.batch_run(gpu_ctx, [(ga, L.Q_BFS, roots), (ga, L.Q_SSSP, roots), (ga, L.Q_PPR, roots[:70], 4)])
            assert np.array_equal(res[0]["out"], want_b)
            assert np.array_equal(res[1]["out"], want_s)
            assert np.array_equal(res[2]["out"], want_p)
    with pytest.raises(L.LA3Error):
        ga.set_tuning(L.TUNE_SSSP_TASK_K, 100)
    rep = ga.autotune(roots)
    assert rep["task_k"][1] in (64, 128, 256) and rep["fused_blocks"] in (8, 16, 32)
    res = L.batch_run(gpu_ctx, [(ga, L.Q_SSSP, roots)])
    assert np.array_equal(res[0]["out"], want_s)
    for t in (L.TUNE_BFS_TASK_K, L.TUNE_SSSP_TASK_K, L.TUNE_PPR_TASK_K):
        ga.set_tuning(t, 256)
    ga.set_tuning(L.TUNE_FUSED_BLOCKS, 32)
    ga.set_tuning(L.TUNE_SSSP_DELTA, 0)
    ga.set_tuning(L.TUNE_BATCH_SSSP_DELTA, 0)


def test_release_scratch_then_run_again(gpu_ctx, two_graphs):
    (ga, ra), (gb, rb) = two_graphs
    roots = _roots(ra, 50, 91)
    first = L.batch_run(gpu_ctx, [(ga, L.Q_SSSP, roots), (gb, L.Q_PPR, roots[:5] % 2048, 5)])
    pr1 = L.pagerank(ga, 7)[0]
    gpu_ctx.release_scratch()
    again = L.batch_run(gpu_ctx, [(ga, L.Q_SSSP, roots), (gb, L.Q_PPR, roots[:5] % 2048, 5)])
    assert np.array_equal(first[0]["out"], again[0]["out"])
    assert np.array_equal(first[1]["out"], again[1]["out"])
    assert np.array_equal(pr1, L.pagerank(ga, 7)[0])
    s = L.Session(gpu_ctx)
    with pytest.raises(L.LA3Error):
        gpu_ctx.release_scratch()      # not while a session runs
    s.close()


def test_narrow_tasks_every_width(gpu_ctx, two_graphs):
    """SSSP and PPR batches of 1, 2, 3, 7, 16 and 32 queries run as narrow
    tasks (1, 2, 4, 8, 16 and 32 lanes per vertex row) in one fused run beside
    wide tasks of the same programs, on both graphs; each equals its batched
    call (distances and iteration counts, PPR ranks and compensation terms
    bit-exact). Also a session where narrow batches join running wide ones."""
    (ga, ra), (gb, rb) = two_graphs
    batches = []
    for i, k in enumerate((1, 2, 3, 7, 16, 32)):
        g, r = (ga, ra) if i % 2 == 0 else (gb, rb)
        batches.append((g, L.Q_SSSP, _roots(r, k, 100 + i)))
        batches.append((g, L.Q_PPR, _roots(r, k, 200 + i), 6))
    batches.append((ga, L.Q_SSSP, _roots(ra, 70, 300)))
    batches.append((gb, L.Q_PPR, _roots(rb, 40, 301), 6))
    batches.append((gb, L.Q_BFS, _roots(rb, 5, 302)))
    _check(batches, L.batch_run(gpu_ctx, batches))
    s = L.Session(gpu_ctx)
    wide = [(ga, L.Q_SSSP, _roots(ra, 100, 310), 0), (gb, L.Q_PPR, _roots(rb, 64, 311), 9)]
    late = [(ga, L.Q_SSSP, _roots(ra, 3, 312), 0), (gb, L.Q_PPR, _roots(rb, 1, 313), 9),
            (ga, L.Q_PPR, _roots(ra, 20, 314), 9)]
    outs = [s.add(g, kind, roots, it) for g, kind, roots, it in wide]
    s.step()
    outs += [s.add(g, kind, roots, it) for g, kind, roots, it in late]
    done = set()
    for _ in range(10000):
        for i, _, st in s.poll():
            assert st == 0
            done.add(i)
        if len(done) == len(outs):
            break
        s.step()
    assert len(done) == len(outs)
    for (qid, out), (g, kind, roots, it) in zip(outs, wide + late):
        if kind == L.Q_SSSP:
            assert np.array_equal(out, L.sssp_batch(g, roots)[0])
        else:
            assert np.array_equal(out, L.ppr_batch(g, roots, it)[0])
    s.close()
