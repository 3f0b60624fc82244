#!/usr/bin/env python3
"""Benchmark of the vertex-program hot path on one MI355X.

Default workload (BFS + SSSP on a directed R-MAT graph): scale 22 (4,194,304
vertices), edge factor 16, a=0.57 b=0.19 c=0.19, seed 2, u32 weights in
[1, 128]; 64 roots drawn by seed from the vertices with out-degree > 0. One
step = the 64 BFS queries then the 64 SSSP queries, each query one call of
the plain GPU vertex program (single-query path).

Prints one JSON line:
  value       GTEPS = sum over queries of the out-degrees of the vertices the
              query reached, divided by the timed wall time;
  roofline    the dominant kernel's algorithmic bytes over its summed
              HIP-event time (events recorded on the library's stream),
              against 8 TB/s;
  cpu_baseline the oracle (CPU restatement of the reference's algorithm,
              one thread) on a sample of the same roots of the same graph.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` measured by rocprofv3 PMC runs,
    recorded in profiles/traffic.json by the profiling recipe (null if absent)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    k = t.get(kernel)
    return None if k is None else k.get("bytes_per_dispatch")


def cpu_baseline(rec, n, roots, outdeg, gpu_hops, gpu_dist):
    """The oracle on `roots` (a sample of the step's roots), one thread; also
    checks that the GPU results for those roots equal the oracle's."""
    from oracle import oracle as O
    t = time.time()
    o = O.OracleGraph(rec, n, True, O.DIRECTED)
    log(f"oracle ingest {time.time() - t:.1f}s (not timed)")
    edges, secs, exact = 0, 0.0, True
    for r in roots:
        t = time.perf_counter()
        h, _, _ = O.bfs(o, r)
        secs += time.perf_counter() - t
        edges += traversed_edges(outdeg, h != O.BFS_INF)
        exact &= bool((h == gpu_hops[r]).all())
        t = time.perf_counter()
        d, _ = O.sssp(o, r)
        secs += time.perf_counter() - t
        edges += traversed_edges(outdeg, d != O.SSSP_INF)
        exact &= bool((d == gpu_dist[r]).all())
    return dict(value=edges / secs / 1e9, unit="GTEPS", cores=1, kind="port",
                sample=f"{len(roots)} BFS + {len(roots)} SSSP roots of the step, full scale-22 "
                       f"graph, oracle (single-threaded CPU restatement), {secs:.1f} s timed",
                exact_match=exact)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--cpu-roots", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the whole graph is one rank's tile set on one "
              "MI355X and this build has no multi-GPU path. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import la3_amd as L

    scale, seed = args.scale, 2
    n = 1 << scale
    t = time.time()
    rec = L.rmat(scale, 16, 0.57, 0.19, 0.19, seed, weighted=True)
    log(f"R-MAT scale {scale}: {rec.shape[0]} edges in {time.time() - t:.1f}s")
    outdeg = np.bincount(rec[:, 0], minlength=n + 1)
    roots = choose_roots(rec, args.roots, seed)

    ctx = L.Context(0)
    t = time.time()
    g = L.Graph.from_records(ctx, rec, n, True, L.DIRECTED)
    log(f"ingest + upload {time.time() - t:.1f}s: {g.info}")

    # Warmup; also fixes the per-query traversed-edge counts (results are exact
    # and deterministic, so they are the same in every step).
    hops, dist = {}, {}
    q_edges = 0
    for w in range(max(args.warmup, 1)):
        for r in roots:
            h, _, _ = L.bfs(g, r)
            hops[r] = h
        for r in roots:
            d, _ = L.sssp(g, r)
            dist[r] = d
    for r in roots:
        reach_b = hops[r] != L.BFS_INF
        reach_s = dist[r] != L.SSSP_INF
        assert (reach_b == reach_s).all(), "BFS and SSSP reach different vertex sets"
        q_edges += traversed_edges(outdeg, reach_b) + traversed_edges(outdeg, reach_s)

    ctx.set_profiling(True)
    ctx.reset_kernel_stats()
    lat = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for r in roots:
            q = time.perf_counter()
            L.bfs(g, r)
            lat.append(time.perf_counter() - q)
        for r in roots:
            q = time.perf_counter()
            L.sssp(g, r)
            lat.append(time.perf_counter() - q)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    ctx.set_profiling(False)

    ms_per_step = elapsed * 1e3 / args.steps
    gteps = q_edges * args.steps / elapsed / 1e9
    qps = 2 * len(roots) * args.steps / elapsed
    dom_name, dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["alg_bytes"] / (dom["ms"] * 1e-3) / 1e9 if dom["ms"] > 0 else 0.0
    lat_ms = np.array(lat) * 1e3
    kernels = {k: dict(launches=v["launches"], ms=round(v["ms"], 3),
                       gbs=round(v["alg_bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] else None)
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])[:6]}
    log(f"kernels: {json.dumps(kernels)}")

    out = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"BFS+SSSP, {len(roots)} roots one query at a time, directed R-MAT "
                        f"scale {scale} edge factor 16 (a=.57 b=.19 c=.19) seed {seed}, "
                        f"u32 weights [1,128]",
            "mode": "single-query",
            "graph_edges": int(g.info["nnz"]),
            "queries_per_step": 2 * len(roots),
        },
        "queries_per_s": round(qps, 2),
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": load_traffic(dom_name),
            "launches": dom["launches"],
            "avg_launch_us": round(dom["ms"] * 1e3 / max(dom["launches"], 1), 2),
            "alg_bytes_per_launch": round(dom["alg_bytes"] / max(dom["launches"], 1)),
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(rec, n, roots[:args.cpu_roots], outdeg, hops, dist)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
