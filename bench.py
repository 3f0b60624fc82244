#!/usr/bin/env python3
"""Benchmark of the vertex-program hot path on one MI355X.

Default workload (the batched-queries configuration): a directed R-MAT graph
of scale 24 (16,777,216 vertices), edge factor 16, a=0.57 b=0.19 c=0.19,
seed 4, u32 weights in [1, 128]; 256 BFS roots, 256 SSSP roots and 128
personalized-PageRank seeds (30 iterations, FP32 ranks with compensated
sums) drawn by seed from the vertices with out-degree > 0. One step = all 640
queries as ONE fused run (one fused launch per traversal step, steps replayed
from a captured hipGraph; results stay in HBM). Timed in the same run on the
same queries: the per-family batched calls one after the other, and the
plain GPU vertex program one query at a time. Before timing, every BFS/SSSP
query's batched and fused results are checked equal to its single-query
result, and PPR is checked against the FP64 single-query program.

Prints one JSON line:
  value        GTEPS = sum over queries of the entries traversed (out-degrees
               of the vertices a BFS/SSSP query reached; nnz of regular rows
               x iterations for PPR), divided by the timed wall time;
  roofline     the fused launch's algorithmic bytes (the compulsory traffic of
               each vertex-program step: (4 + w) E_scan + (8 + m) C_act +
               (a + s_r + s_w) R_touch + m R_act, per-vertex terms scaled by
               the live queries, R_touch counted on the device) over its summed
               HIP-event time on the library stream, against 8 TB/s; traffic =
               HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs;
  cpu_baseline the oracle (CPU restatement of the reference's algorithm) on a
               sample of the same BFS/SSSP queries on the same graph, one
               query per thread on the host's cores.

Other workloads (--workload): bfs_sssp_s22 (64 BFS + 64 SSSP roots, scale
22), pr_s16 (PageRank), cc_tc_s20 (CC and TC), serve (serving loop).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS & queries/s: served/batched vs single-query at exact results; p50/p99 ms"
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def choose_roots(rec, count, seed):
    rng = np.random.default_rng(seed)
    cand = np.unique(rec[:, 0])  # vertices with out-degree > 0
    return [int(x) for x in rng.choice(cand, size=count, replace=False)]


def traversed_edges(outdeg, reached_mask):
    return int(outdeg[reached_mask].sum())


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` in `workload` measured by rocprofv3 PMC
    runs, recorded in profiles/traffic.json by the profiling recipe (null if
    absent)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    k = t.get(workload, {}).get(kernel)
    return None if k is None else k.get("bytes_per_dispatch")


def cpu_cores():
    """Host cores this process may use, at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(rec, n, roots, outdeg, gpu_hops, gpu_dist, scale):
    """The oracle (CPU restatement of the reference's np=1 path) on `roots`
    (each as a BFS and as an SSSP query) on the graph of `rec`."""
    from oracle import oracle as O
    t = time.time()
    o = O.OracleGraph(rec, n, True, O.DIRECTED)
    log(f"oracle ingest {time.time() - t:.1f}s (not timed)")
    return cpu_baseline_on(o, outdeg, roots, gpu_hops, gpu_dist, scale)


def cpu_baseline_on(o, outdeg, roots, gpu_hops, gpu_dist, scale):
    """The oracle on `roots` (each as a BFS and as an SSSP query), one query
    per thread across the host cores (queries are independent, as the
    reference's np x tpp ranks and threads would split them); also checks
    that the GPU results for those roots equal the oracle's."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    cores = cpu_cores()

    def one(job):
        kind, r = job
        if kind == "bfs":
            h, _, _ = O.bfs(o, r)
            return traversed_edges(outdeg, h != O.BFS_INF), bool((h == gpu_hops[r]).all())
        d, _ = O.sssp(o, r)
        return traversed_edges(outdeg, d != O.SSSP_INF), bool((d == gpu_dist[r]).all())

    jobs = [("bfs", r) for r in roots] + [("sssp", r) for r in roots]
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:   # ctypes calls release the GIL
        out = list(ex.map(one, jobs))
    secs = time.perf_counter() - t
    edges = sum(e for e, _ in out)
    exact = all(x for _, x in out)
    return dict(value=edges / secs / 1e9, unit="GTEPS", cores=cores, kind="port",
                sample=f"{len(roots)} BFS + {len(roots)} SSSP queries (the same roots) on the full scale-{scale} "
                       f"graph, oracle (CPU restatement), one query per thread on {cores} threads, "
                       f"{secs:.1f} s wall",
                exact_match=exact)


def run_single(L, g, roots, hbuf, dbuf, steps):
    """One step = every root as its own BFS query, then its own SSSP query.
    hbuf/dbuf = L.DEVICE keeps results in HBM."""
    lat = []
    for _ in range(steps):
        for r in roots:
            q = time.perf_counter()
            L.bfs(g, r, parents=False, out=hbuf)
            lat.append(time.perf_counter() - q)
        for r in roots:
            q = time.perf_counter()
            L.sssp(g, r, out=dbuf)
            lat.append(time.perf_counter() - q)
    return lat


def run_batched(L, g, roots, hb, db, steps):
    """One step = all roots in one batched BFS, then one batched SSSP.
    hb/db = L.DEVICE keeps results in HBM."""
    lat = []
    for _ in range(steps):
        q = time.perf_counter()
        L.bfs_batch(g, roots, out=hb)
        lat.append(time.perf_counter() - q)
        q = time.perf_counter()
        L.sssp_batch(g, roots, out=db)
        lat.append(time.perf_counter() - q)
    return lat


def run_fused(L, ctx, g, roots, out, steps):
    """One step = the BFS batch and the SSSP batch advancing together in one
    fused, graph-captured run (la3_batch_run)."""
    lat = []
    for _ in range(steps):
        q = time.perf_counter()
        L.batch_run(ctx, [(g, L.Q_BFS, roots), (g, L.Q_SSSP, roots)], out=out)
        lat.append(time.perf_counter() - q)
    return lat


def timed(ctx, fn, steps, profile=True):
    ctx.set_profiling(profile)
    ctx.reset_kernel_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    lat = fn(steps)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    stats = ctx.kernel_stats() if profile else {}
    ctx.set_profiling(False)
    return elapsed, lat, stats


def roofline(stats, workload=None):
    name, dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["alg_bytes"] / (dom["ms"] * 1e-3) / 1e9 if dom["ms"] > 0 else 0.0
    return {
        "bound": "hbm",
        "kernel": name,
        "achieved": round(achieved, 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4),
        "traffic": load_traffic(name, workload) if workload else None,
        "launches": dom["launches"],
        "avg_launch_us": round(dom["ms"] * 1e3 / max(dom["launches"], 1), 2),
        "alg_bytes_per_launch": round(dom["alg_bytes"] / max(dom["launches"], 1)),
    }


def kernel_table(stats):
    return {k: dict(launches=v["launches"], ms=round(v["ms"], 3),
                    gbs=round(v["alg_bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] else None)
            for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])[:6]}


def base_line(args, value, unit, ms_per_step, dtype, workload, **extra):
    out = {"metric": METRIC, "value": round(value, 4), "unit": unit, "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": {"workload": workload}}
    out.update(extra)
    return out


def workload_pr_s16(args, L):
    """PageRank, 20 iterations, FP64, directed R-MAT scale 16 seed 1."""
    from oracle import oracle as O
    scale, n = 16, 1 << 16
    rec = L.rmat(scale, 16, 0.57, 0.19, 0.19, 1)
    ctx = L.Context(0)
    g = L.Graph.from_records(ctx, rec, n, False, L.DIRECTED)
    rank, _, _ = L.pagerank(g, 20)
    o, orr = O.OracleGraph(rec, n, False, O.DIRECTED), O.OracleGraph(rec, n, False, O.REVERSED)
    t = time.perf_counter()
    erank, _, _ = O.pagerank(o, orr, 20)
    cpu_s = time.perf_counter() - t
    l1 = float(np.abs(rank - erank).sum())
    for _ in range(max(args.warmup, 1)):
        L.pagerank(g, 20, out=L.DEVICE)
    # value: the 20 iterations replayed from the captured hipGraph; then the
    # same iterations as separate launches with per-kernel events (roofline)
    el, _, _ = timed(ctx, lambda k: [L.pagerank(g, 20, out=L.DEVICE) for _ in range(k)], args.steps, profile=False)
    el_p, _, st = timed(ctx, lambda k: [L.pagerank(g, 20, out=L.DEVICE) for _ in range(k)], args.steps)
    edges = g.info["nnz"] * 20
    return base_line(args, edges * args.steps / el / 1e9, "GTEPS", el * 1e3 / args.steps, "f64",
                     "PageRank 20 iterations FP64, directed R-MAT scale 16 ef 16 seed 1",
                     l1_vs_oracle=l1, roofline=roofline(st), kernels=kernel_table(st),
                     profiled_ms_per_step=round(el_p * 1e3 / args.steps, 3),
                     cpu_baseline={"value": edges / cpu_s / 1e9, "unit": "GTEPS", "cores": 1,
                                   "kind": "port", "sample": "one full 20-iteration run (oracle)"})


def workload_cc_tc_s20(args, L):
    """CC (undirected load) and TC (acyclic load) on R-MAT scale 20 seed 3, symmetrized."""
    from oracle import oracle as O
    scale, n = 20, 1 << 20
    rec = L.rmat(scale, 16, 0.57, 0.19, 0.19, 3)
    ctx = L.Context(0)
    gu = L.Graph.from_records(ctx, rec, n, False, L.UNDIRECTED)
    ga = L.Graph.from_records(ctx, rec, n, False, L.ACYCLIC)
    lab, _ = L.cc(gu)
    total, _ = L.tc(ga)
    el_cc, _, _ = timed(ctx, lambda k: [L.cc(gu, out=L.DEVICE) for _ in range(k)], args.steps)
    el_tc, _, st_tc = timed(ctx, lambda k: [L.tc(ga, per_vertex=False) for _ in range(k)], args.steps)
    ou = O.OracleGraph(rec, n, False, O.UNDIRECTED)
    t = time.perf_counter()
    elab, _ = O.cc(ou)
    cpu_cc = time.perf_counter() - t
    oa, ora = O.OracleGraph(rec, n, False, O.ACYCLIC), O.OracleGraph(rec, n, False, O.REVERSED_ACYCLIC)
    t = time.perf_counter()
    etotal, _ = O.tc(oa, ora)
    cpu_tc = time.perf_counter() - t
    cc_e = gu.info["nnz"]
    return base_line(args, cc_e * args.steps / el_cc / 1e9, "GTEPS", el_cc * 1e3 / args.steps, "u32",
                     "CC (value) and TC on symmetrized R-MAT scale 20 ef 16 seed 3",
                     cc_exact=bool((lab == elab).all()), tc_total=total, tc_exact=total == etotal,
                     tc_ms=round(el_tc * 1e3 / args.steps, 3), roofline=None,
                     roofline_note="CC runs on the VertexProgram template, whose kernels are launched outside "
                                   "the library's event timing; a run is about 25 launches of 5-230 us (launch- "
                                   "and latency-bound at this size); per-kernel times in "
                                   "profiles/vp_cc_tc_s20_kernel_stats.md",
                     tc_roofline=roofline(st_tc),
                     cpu_baseline={"value": cc_e / cpu_cc / 1e9, "unit": "GTEPS", "cores": 1,
                                   "kind": "port", "sample": "one CC run (oracle label propagation)",
                                   "tc_s": round(cpu_tc, 3)})


def workload_batched_s24(args, L):
    """256 BFS + 256 SSSP + 128 PPR (30 iterations) on R-MAT scale 24 seed 4:
    one fused run (value) vs the per-family batched calls vs one query at a
    time, with the oracle on a sample of the queries as the CPU baseline."""
    from concurrent.futures import ThreadPoolExecutor
    scale, seed, n = 24, 4, 1 << 24
    t = time.time()
    rec = L.rmat(scale, 16, 0.57, 0.19, 0.19, seed, weighted=True)
    log(f"R-MAT scale {scale}: {rec.shape[0]} edges in {time.time() - t:.1f}s")
    outdeg = np.bincount(rec[:, 0], minlength=n + 1)
    # The oracle's graph for the CPU baseline builds on one host core while
    # the GPU legs run (not timed).
    pool = ThreadPoolExecutor(max_workers=1)
    oracle_graph = None
    if not args.no_cpu_baseline:
        def build_oracle():
            from oracle import oracle as O
            t0 = time.time()
            o = O.OracleGraph(rec, n, True, O.DIRECTED)
            log(f"oracle ingest {time.time() - t0:.1f}s (background, not timed)")
            return o
        oracle_graph = pool.submit(build_oracle)
    ctx = L.Context(0)
    t = time.time()
    g = L.Graph.from_records(ctx, rec, n, True, L.DIRECTED)
    log(f"ingest {time.time() - t:.1f}s: {g.info}")
    bfs_r, sssp_r, ppr_s = choose_roots(rec, 256, 41), choose_roots(rec, 256, 42), choose_roots(rec, 128, 43)
    if args.fused_only:
        return fused_only(args, L, ctx, g, [(g, L.Q_BFS, bfs_r), (g, L.Q_SSSP, sssp_r), (g, L.Q_PPR, ppr_s, 30)],
                          "batched_s24")
    t = time.time()
    tune = g.autotune(sssp_r)
    log(f"autotune ({time.time() - t:.1f}s, not timed): {tune}")
    D = L.DEVICE
    t = time.time()
    hb, _ = L.bfs_batch(g, bfs_r)
    db, _ = L.sssp_batch(g, sssp_r)
    pb, _, resid = L.ppr_batch(g, ppr_s, 30)
    edges = 0
    sample = sssp_r[:args.cpu_roots]
    gpu_hops, gpu_dist = {}, {}
    for q, r in enumerate(bfs_r):
        h1 = L.bfs(g, r, parents=False)[0]
        assert (hb[q] == h1).all(), f"batched BFS differs from single-query for root {r}"
        edges += traversed_edges(outdeg, h1 != L.BFS_INF)
    for q, r in enumerate(sssp_r):
        d1 = L.sssp(g, r)[0]
        assert (db[q] == d1).all(), f"batched SSSP differs from single-query for root {r}"
        edges += traversed_edges(outdeg, d1 != L.SSSP_INF)
        if r in sample:
            gpu_dist[r] = d1
            gpu_hops[r] = L.bfs(g, r, parents=False)[0]
    l1s, rels = [], []
    for q in range(0, 128, 16):
        r64 = L.ppr(g, ppr_s[q], 30)
        l1s.append(float(np.abs(pb[q] - r64).sum()))
        nz = r64 > 0
        rels.append(float((np.abs(pb[q] - r64)[nz] / r64[nz]).max()))
    assert max(l1s) <= 1e-7, f"PPR FP32+compensation vs FP64: L1 {max(l1s)} > 1e-7"
    ppr_edges = g.info["nnz_regular"] * 30
    edges += ppr_edges * 128
    mix = [(g, L.Q_BFS, bfs_r), (g, L.Q_SSSP, sssp_r), (g, L.Q_PPR, ppr_s, 30)]
    fr = L.batch_run(ctx, mix)
    assert (fr[0]["out"] == hb).all() and (fr[1]["out"] == db).all() and (fr[2]["out"] == pb).all(), \
        "fused run differs from the per-family batched calls"
    log(f"exact ({time.time() - t:.1f}s): fused run == per-family batched calls == single-query for all "
        f"512 BFS/SSSP queries; PPR L1 vs FP64 {max(l1s):.2e}")
    del hb, db, pb, fr

    def fused(k):
        for _ in range(k):
            L.batch_run(ctx, mix, out=D)

    def batched(k):
        for _ in range(k):
            L.bfs_batch(g, bfs_r, out=D)
            L.sssp_batch(g, sssp_r, out=D)
            L.ppr_batch(g, ppr_s, 30, out=D)

    def single(k):
        for _ in range(k):
            for r in bfs_r:
                L.bfs(g, r, parents=False, out=D)
            for r in sssp_r:
                L.sssp(g, r, out=D)
            for s_ in ppr_s:
                L.ppr(g, s_, 30, out=D)

    for _ in range(max(args.warmup, 1)):
        fused(1)
    el_f, lat_f, _ = timed(ctx, lambda k: [(time.perf_counter(), fused(1), time.perf_counter()) for _ in range(k)],
                           args.steps, profile=False)
    step_ms = [(b - a) * 1e3 for a, _, b in lat_f]
    log(f"fused: {el_f * 1e3 / args.steps:.1f} ms/step, {edges / (el_f / args.steps) / 1e9:.1f} GTEPS")
    # the same steps as separate launches with per-kernel events and the
    # device's touched-row counts (roofline)
    el_fp, _, st_f = timed(ctx, fused, 1)
    log(f"fused, profiled launches: {el_fp * 1e3:.1f} ms/step; kernels {json.dumps(kernel_table(st_f))}")
    el_b, _, st_b = timed(ctx, batched, args.steps, profile=False)
    log(f"batched (sequential): {el_b * 1e3 / args.steps:.1f} ms/step")
    single(0)
    el_s, _, _ = timed(ctx, single, 1, profile=False)
    log(f"single query: {el_s * 1e3:.1f} ms for 640 queries")
    copy_gbs = ctx.copy_bandwidth(1 << 30, 10)
    log(f"device copy bandwidth {copy_gbs:.0f} GB/s (1 GiB buffer, read + write)")
    out = base_line(args, edges * args.steps / el_f / 1e9, "GTEPS", el_f * 1e3 / args.steps, "u32/f32",
                    "batched queries: 256 BFS + 256 SSSP + 128 PPR (30 iterations, FP32 + compensated sums), "
                    "directed R-MAT scale 24 ef 16 (a=.57 b=.19 c=.19) seed 4, u32 weights [1,128]; all 640 "
                    "queries in one fused run per step, one fused launch per traversal step, steps replayed "
                    "from a captured hipGraph, results left in HBM",
                    queries_per_s=round(640 * args.steps / el_f, 2),
                    p50_ms=round(float(np.percentile(step_ms, 50)), 3),
                    p99_ms=round(float(np.percentile(step_ms, 99)), 3),
                    latency_note="a query's latency is its fused run's wall time (all 640 finish together)",
                    batched_sequential={"value": round(edges * args.steps / el_b / 1e9, 4), "unit": "GTEPS",
                                        "ms_per_step": round(el_b * 1e3 / args.steps, 3),
                                        "note": "la3_bfs_batch, la3_sssp_batch, la3_ppr_batch one after another"},
                    single_query={"value": round(edges / el_s / 1e9, 4), "unit": "GTEPS",
                                  "queries_per_s": round(640 / el_s, 2), "ms": round(el_s * 1e3, 1),
                                  "note": "the plain GPU vertex program, one query per call (FP64 PPR)"},
                    speedup_vs_single=round(el_s * args.steps / el_f, 2),
                    ppr_l1_vs_fp64=max(l1s), ppr_max_rel=max(rels),
                    ppr_residual_last_iter=float(np.max(resid)),
                    roofline=roofline(st_f, "batched_s24"), kernels=kernel_table(st_f),
                    device_copy_gbs=round(copy_gbs, 1),
                    autotune={"sssp_delta": tune["sssp_delta"], "batch_sssp_delta": tune["batch_sssp_delta"],
                              "task_k": tune["task_k"], "fused_blocks_per_cu": tune["fused_blocks"],
                              "task_ms_64_128_256": [[round(x, 3) for x in r] for r in tune["task_ms"]],
                              "fused_ms_8_16_32": [round(x, 3) for x in tune["fused_ms"]]})
    out["config"]["graph_edges"] = int(g.info["nnz"])
    out["config"]["queries_per_step"] = 640
    if oracle_graph is not None:
        o = oracle_graph.result()
        out["cpu_baseline"] = cpu_baseline_on(o, outdeg, sample, gpu_hops, gpu_dist, scale)
    pool.shutdown()
    return out


def fused_only(args, L, ctx, g, mix, workload):
    """Profiling runs (rocprofv3 kernel traces and counter passes): the
    tuning the autotuner chose, given by --tune, instead of the autotuner;
    then only fused steps, every one launched with per-kernel events and
    touched-row counting, so a trace's average for the fused launch covers
    exactly the launches behind the roofline; then the copy kernel of known
    size (1 GiB) that calibrates the HBM counters."""
    sd, bd, kb, ks, kp, blocks = (int(x) for x in args.tune.split(","))
    g.set_tuning(L.TUNE_SSSP_DELTA, sd)
    g.set_tuning(L.TUNE_BATCH_SSSP_DELTA, bd)
    g.set_tuning(L.TUNE_BFS_TASK_K, kb)
    g.set_tuning(L.TUNE_SSSP_TASK_K, ks)
    g.set_tuning(L.TUNE_PPR_TASK_K, kp)
    g.set_tuning(L.TUNE_FUSED_BLOCKS, blocks)
    fused = lambda k: [L.batch_run(ctx, mix, out=L.DEVICE) for _ in range(k)]
    timed(ctx, fused, max(args.warmup, 1))
    el, _, st = timed(ctx, fused, args.steps)
    log(f"fused, profiled launches: {el * 1e3 / args.steps:.1f} ms/step; kernels {json.dumps(kernel_table(st))}")
    copy_gbs = ctx.copy_bandwidth(1 << 30, 1)
    return {"profiling_run": workload, "tune": args.tune, "steps": args.steps,
            "ms_per_step_profiled": round(el * 1e3 / args.steps, 3),
            "roofline": roofline(st, workload), "kernels": kernel_table(st), "device_copy_gbs": round(copy_gbs, 1)}


def workload_serve(args, L):
    """Six resident R-MAT graphs (scales 18-22 and 24, seeds 10-15, weights
    [1,128]); a seeded Poisson stream at fixed QPS mixing 50% BFS, 30% SSSP,
    20% PPR (30 iterations, FP32); graphs chosen Zipf(1.1) over size rank
    (smallest most frequent); roots uniform over non-isolated vertices.
    Served with continuous batching vs one query at a time, same stream."""
    scales = [18, 19, 20, 21, 22, 24]
    ctx = L.Context(0)
    graphs = []
    for i, sc in enumerate(scales):
        rec = L.rmat(sc, 16, 0.57, 0.19, 0.19, 10 + i, weighted=True)
        g = L.Graph.from_records(ctx, rec, 1 << sc, True, L.DIRECTED)
        nonisol = np.unique(rec[:, :2])
        outdeg = np.bincount(rec[:, 0], minlength=(1 << sc) + 1)
        graphs.append((g, nonisol, outdeg))
        del rec
    log("six graphs resident: " + ", ".join(f"s{sc}:{g.info['nnz']}" for sc, (g, _, _) in zip(scales, graphs)))
    rng = np.random.default_rng(2026)
    nq, qps = args.serve_queries, args.qps
    zipf = 1.0 / np.arange(1, 7) ** 1.1
    zipf /= zipf.sum()
    gi = rng.choice(6, size=nq, p=zipf)
    kinds = rng.choice(3, size=nq, p=[0.5, 0.3, 0.2])
    roots = [int(rng.choice(graphs[i][1])) for i in gi]
    arrivals = np.cumsum(rng.exponential(1.0 / qps, size=nq))

    def run(max_batch, nworkers, fused=False):
        srv = L.Server(0, nworkers=nworkers, max_batch=max_batch, state_budget=64 << 30, fused=fused)
        gids = [srv.add_graph(g) for g, _, _ in graphs]
        # warm every (graph, program) once
        for i in range(6):
            for kd in (L.Q_BFS, L.Q_SSSP, L.Q_PPR):
                srv.submit(kd, gids[i], int(graphs[i][1][0]), 30 if kd == L.Q_PPR else 0)
        srv.drain()
        srv.completed()
        if not args.serve_cold:
            srv.warm()  # every worker's scratch at working size: no growth frees in the stream
        t0 = srv.clock()
        for j in range(nq):
            while srv.clock() - t0 < arrivals[j]:
                time.sleep(0.0002)
            srv.submit(int(kinds[j]), gids[gi[j]], roots[j], 30 if kinds[j] == L.Q_PPR else 0)
        srv.drain()
        done = srv.completed()
        srv.close()
        lat = np.array([d["done_s"] - d["submit_s"] for d in done]) * 1e3
        span = max(d["done_s"] for d in done) - t0
        bs = np.array([d["batch_size"] for d in done])
        return lat, span, bs

    # traversed edges per query (deterministic, computed from single-query results)
    edges = 0
    for j in range(nq):
        g, _, outdeg = graphs[gi[j]]
        if kinds[j] == L.Q_PPR:
            edges += g.info["nnz_regular"] * 30
        elif kinds[j] == L.Q_BFS:
            edges += traversed_edges(outdeg, L.bfs(g, roots[j], parents=False)[0] != L.BFS_INF)
        else:
            edges += traversed_edges(outdeg, L.sssp(g, roots[j])[0] != L.SSSP_INF)
    lat_f, span_f, bf = run(args.serve_batch, args.serve_workers, fused=True)
    log(f"served, mixed fused batches: {nq / span_f:.1f} q/s, p50 {np.percentile(lat_f, 50):.1f} ms, "
        f"p99 {np.percentile(lat_f, 99):.1f} ms, mean batch {bf.mean():.1f}")
    lat_b, span_b, bs = run(args.serve_batch, 3)
    log(f"served, per-(program, graph) batches: {nq / span_b:.1f} q/s, p50 {np.percentile(lat_b, 50):.1f} ms, "
        f"p99 {np.percentile(lat_b, 99):.1f} ms, mean batch {bs.mean():.1f}")
    lat_s, span_s, _ = run(1, 1)
    return base_line(args, edges / span_b / 1e9, "GTEPS", span_b * 1e3, "u32/f32",
                     f"serving: {nq} queries, Poisson {qps} QPS, 50/30/20 BFS/SSSP/PPR(30 it), "
                     "six resident R-MAT graphs s18-22,24 seeds 10-15, Zipf(1.1) by size rank; "
                     "continuous batching, batches of one (program, graph), 3 workers",
                     queries_per_s=round(nq / span_b, 2), warm_workers=not args.serve_cold,
                     offered_qps=round(nq / float(arrivals[-1]), 2),
                     p50_ms=round(float(np.percentile(lat_b, 50)), 3),
                     p99_ms=round(float(np.percentile(lat_b, 99)), 3),
                     mean_batch=round(float(bs.mean()), 2), max_batch=int(bs.max()),
                     served_fused={"value": round(edges / span_f / 1e9, 4), "unit": "GTEPS",
                                   "queries_per_s": round(nq / span_f, 2),
                                   "p50_ms": round(float(np.percentile(lat_f, 50)), 3),
                                   "p99_ms": round(float(np.percentile(lat_f, 99)), 3),
                                   "mean_batch": round(float(bf.mean()), 2),
                                   "note": f"fused sessions ({args.serve_workers} workers): mixed groups of "
                                           "every program and graph join a running fused run each step"},
                     single_query={"value": round(edges / span_s / 1e9, 4), "unit": "GTEPS",
                                   "queries_per_s": round(nq / span_s, 2),
                                   "p50_ms": round(float(np.percentile(lat_s, 50)), 3),
                                   "p99_ms": round(float(np.percentile(lat_s, 99)), 3)})


def workload_bfs_sssp_s22(args, L):
    """64 BFS + 64 SSSP roots on R-MAT scale 22 seed 2: fused run vs batched vs single-query."""
    scale, seed = args.scale, 2
    n = 1 << scale
    t = time.time()
    rec = L.rmat(scale, 16, 0.57, 0.19, 0.19, seed, weighted=True)
    log(f"R-MAT scale {scale}: {rec.shape[0]} edges in {time.time() - t:.1f}s")
    outdeg = np.bincount(rec[:, 0], minlength=n + 1)
    roots = choose_roots(rec, args.roots, seed)
    k = len(roots)

    ctx = L.Context(0)
    t = time.time()
    g = L.Graph.from_records(ctx, rec, n, True, L.DIRECTED)
    log(f"ingest + upload {time.time() - t:.1f}s: {g.info}")
    t = time.time()
    tune = g.autotune(roots)
    log(f"autotune ({time.time() - t:.1f}s, not timed): {tune}")
    S = g.nslots
    hb = L.pinned((k, S), np.uint8)
    db = L.pinned((k, S), np.uint32)
    hbuf = L.pinned((S,), np.uint8)
    dbuf = L.pinned((S,), np.uint32)

    # Warmup, exactness (batched == single query for every root) and the
    # per-query traversed-edge counts (results are exact and deterministic).
    for _ in range(max(args.warmup, 1)):
        run_batched(L, g, roots, hb, db, 1)
    q_edges = 0
    hops, dist = {}, {}
    for q, r in enumerate(roots):
        h1, _, _ = L.bfs(g, r, parents=False)
        d1, _ = L.sssp(g, r)
        assert (hb[q] == h1).all(), f"batched BFS differs from single-query for root {r}"
        assert (db[q] == d1).all(), f"batched SSSP differs from single-query for root {r}"
        reach = h1 != L.BFS_INF
        assert (reach == (d1 != L.SSSP_INF)).all(), "BFS and SSSP reach different vertex sets"
        q_edges += 2 * traversed_edges(outdeg, reach)
        hops[r], dist[r] = h1, d1
    log(f"exact: batched == single-query for all {k} BFS and {k} SSSP roots")

    # value: the fused run (both batches advance together, steps replayed
    # from one captured hipGraph); results stay in HBM (inputs resident, no
    # PCIe in the timed region). Checked equal to the batched path first.
    D = L.DEVICE
    fr = L.batch_run(ctx, [(g, L.Q_BFS, roots), (g, L.Q_SSSP, roots)])
    assert (fr[0]["out"] == hb).all() and (fr[1]["out"] == db).all(), "fused run differs from batched"
    log(f"exact: fused run == batched for all {k} BFS and {k} SSSP roots")
    el_f, lat_f, _ = timed(ctx, lambda s: run_fused(L, ctx, g, roots, D, s), args.steps, profile=False)
    gteps_f = q_edges * args.steps / el_f / 1e9
    log(f"fused: {el_f * 1e3 / args.steps:.1f} ms/step, {gteps_f:.2f} GTEPS")
    # the same steps launched one by one with per-kernel events (roofline)
    el_fp, _, st_f = timed(ctx, lambda s: run_fused(L, ctx, g, roots, D, s), args.steps)
    log(f"fused, profiled launches: {el_fp * 1e3 / args.steps:.1f} ms/step; kernels {json.dumps(kernel_table(st_f))}")
    # the two batches one after the other (per-family batched calls)
    el_b, lat_b, st_b = timed(ctx, lambda s: run_batched(L, g, roots, D, D, s), args.steps)
    gteps_b = q_edges * args.steps / el_b / 1e9
    log(f"batched (sequential): {el_b * 1e3 / args.steps:.1f} ms/step, {gteps_b:.2f} GTEPS; kernels {json.dumps(kernel_table(st_b))}")
    # the fused run with every result copied to page-locked host memory
    el_p, _, _ = timed(ctx, lambda s: run_fused(L, ctx, g, roots, [hb, db], s), args.steps, profile=False)
    gteps_p = q_edges * args.steps / el_p / 1e9
    log(f"fused incl. PCIe result copies: {el_p * 1e3 / args.steps:.1f} ms/step, {gteps_p:.2f} GTEPS")
    single = None
    if not args.no_single:
        for _ in range(max(args.warmup, 1)):
            run_single(L, g, roots[:4], D, D, 1)
        el_s, lat_s, st_s = timed(ctx, lambda s: run_single(L, g, roots, D, D, s), args.steps)
        gteps_s = q_edges * args.steps / el_s / 1e9
        lat_ms = np.array(lat_s) * 1e3
        single = {
            "value": round(gteps_s, 4), "unit": "GTEPS",
            "ms_per_step": round(el_s * 1e3 / args.steps, 3),
            "queries_per_s": round(2 * k * args.steps / el_s, 2),
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "kernels": kernel_table(st_s),
        }
        log(f"single: {el_s * 1e3 / args.steps:.1f} ms/step, {gteps_s:.2f} GTEPS; kernels {json.dumps(kernel_table(st_s))}")

    copy_gbs = ctx.copy_bandwidth(1 << 30, 10)
    log(f"device copy bandwidth {copy_gbs:.0f} GB/s (1 GiB buffer, read + write)")
    lat_f_ms = np.array(lat_f) * 1e3
    out = {
        "metric": METRIC,
        "value": round(gteps_f, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el_f * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"BFS+SSSP, {k} roots, directed R-MAT scale {scale} edge factor 16 "
                        f"(a=.57 b=.19 c=.19) seed {seed}, u32 weights [1,128]",
            "mode": f"fused batch run: {k} BFS roots as one k-bit traversal and {k} SSSP roots as one "
                    f"k-wide traversal advancing together, one fused launch per step, steps replayed "
                    f"from a captured hipGraph; results equal the single-query path for every root and "
                    f"stay in HBM (PCIe-inclusive rate under pcie_inclusive)",
            "graph_edges": int(g.info["nnz"]),
            "queries_per_step": 2 * k,
        },
        "queries_per_s": round(2 * k * args.steps / el_f, 2),
        "p50_ms": round(float(np.percentile(lat_f_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_f_ms, 99)), 3),
        "latency_note": "a query's latency is its fused run's wall time (all 2k queries finish together)",
        "pcie_inclusive": {"value": round(gteps_p, 4), "unit": "GTEPS",
                           "ms_per_step": round(el_p * 1e3 / args.steps, 3),
                           "note": "fused run with every per-query result copied to page-locked host memory"},
        "batched_sequential": {"value": round(gteps_b, 4), "unit": "GTEPS",
                               "ms_per_step": round(el_b * 1e3 / args.steps, 3),
                               "note": "la3_bfs_batch then la3_sssp_batch (host-driven steps)",
                               "kernels": kernel_table(st_b)},
        "single_query": single,
        "speedup_vs_single": round(gteps_f / single["value"], 2) if single else None,
        "roofline": roofline(st_f, "bfs_sssp_s22"),
        "device_copy_gbs": round(copy_gbs, 1),
        "autotune": {"sssp_delta": tune["sssp_delta"], "batch_sssp_delta": tune["batch_sssp_delta"],
                     "candidates": tune["candidates"],
                     "single_ms": [round(x, 3) for x in tune["single_ms"]],
                     "batch_ms": [round(x, 3) for x in tune["batch_ms"]],
                     "task_k": tune["task_k"], "task_ms_64_128_256": [[round(x, 3) for x in r] for r in tune["task_ms"]],
                     "fused_blocks_per_cu": tune["fused_blocks"],
                     "fused_ms_8_16_32": [round(x, 3) for x in tune["fused_ms"]]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(rec, n, roots[:args.cpu_roots], outdeg, hops, dist, scale)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--cpu-roots", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the single-query baseline")
    ap.add_argument("--fused-only", action="store_true",
                    help="profiling runs: fixed tuning (--tune), only event-profiled fused steps")
    ap.add_argument("--tune", default="512,512,256,128,128,32",
                    help="--fused-only tuning: sssp_delta,batch_sssp_delta,bfs_k,sssp_k,ppr_k,blocks_per_cu")
    ap.add_argument("--workload", default="batched_s24",
                    choices=["bfs_sssp_s22", "pr_s16", "cc_tc_s20", "batched_s24", "serve"])
    ap.add_argument("--qps", type=float, default=400.0)
    ap.add_argument("--serve-queries", type=int, default=2000)
    ap.add_argument("--serve-batch", type=int, default=64)
    ap.add_argument("--serve-cold", action="store_true",
                    help="serve: skip the server warm-up (workers grow scratch during the stream)")
    ap.add_argument("--serve-workers", type=int, default=1,
                    help="workers (sessions) of the fused serving mode")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the whole graph is one rank's tile set on one "
              "MI355X and this build has no multi-GPU path. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import la3_amd as L
    work = {"batched_s24": workload_batched_s24, "bfs_sssp_s22": workload_bfs_sssp_s22,
            "pr_s16": workload_pr_s16, "cc_tc_s20": workload_cc_tc_s20, "serve": workload_serve}
    print(json.dumps(work[args.workload](args, L)), flush=True)



if __name__ == "__main__":
    main()
