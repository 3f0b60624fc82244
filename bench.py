#!/usr/bin/env python3
"""Benchmark of the vertex-program hot path on one MI355X.

Default workload (BFS + SSSP on a directed R-MAT graph): scale 22 (4,194,304
vertices), edge factor 16, a=0.57 b=0.19 c=0.19, seed 2, u32 weights in
[1, 128]; 64 roots drawn by seed from the vertices with out-degree > 0. One
step = the 64 BFS queries then the 64 SSSP queries. Two paths are timed in
the same run on the same roots: batched (all roots in one traversal per
program; this is `value`) and single-query (one call of the plain GPU vertex
program per root; reported under single_query). Before timing, every root's
batched result is checked equal to its single-query result.

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


def timed(ctx, fn, steps):
    ctx.set_profiling(True)
    ctx.reset_kernel_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    lat = fn(steps)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    ctx.set_profiling(False)
    return elapsed, lat, stats


def roofline(stats):
    name, dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["alg_bytes"] / (dom["ms"] * 1e-3) / 1e9 if dom["ms"] > 0 else 0.0
    return {
        "bound": "hbm",
        "kernel": name,
        "achieved": round(achieved, 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4),
        "traffic": load_traffic(name),
        "launches": dom["launches"],
        "avg_launch_us": round(dom["ms"] * 1e3 / max(dom["launches"], 1), 2),
        "alg_bytes_per_launch": round(dom["alg_bytes"] / max(dom["launches"], 1)),
    }


def kernel_table(stats):
    return {k: dict(launches=v["launches"], ms=round(v["ms"], 3),
                    gbs=round(v["alg_bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] else None)
            for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])[:6]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--cpu-roots", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the single-query baseline")
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
    k = len(roots)

    ctx = L.Context(0)
    t = time.time()
    g = L.Graph.from_records(ctx, rec, n, True, L.DIRECTED)
    log(f"ingest + upload {time.time() - t:.1f}s: {g.info}")
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

    # value: results stay in HBM (inputs resident, no PCIe in the timed region)
    D = L.DEVICE
    el_b, lat_b, st_b = timed(ctx, lambda s: run_batched(L, g, roots, D, D, s), args.steps)
    gteps_b = q_edges * args.steps / el_b / 1e9
    log(f"batched: {el_b * 1e3 / args.steps:.1f} ms/step, {gteps_b:.2f} GTEPS; kernels {json.dumps(kernel_table(st_b))}")
    # the same with every result copied to page-locked host memory
    el_p, _, _ = timed(ctx, lambda s: run_batched(L, g, roots, hb, db, s), args.steps)
    gteps_p = q_edges * args.steps / el_p / 1e9
    log(f"batched incl. PCIe result copies: {el_p * 1e3 / args.steps:.1f} ms/step, {gteps_p:.2f} GTEPS")
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
            "roofline": roofline(st_s),
        }
        log(f"single: {el_s * 1e3 / args.steps:.1f} ms/step, {gteps_s:.2f} GTEPS; kernels {json.dumps(kernel_table(st_s))}")

    copy_gbs = ctx.copy_bandwidth(1 << 30, 10)
    log(f"device copy bandwidth {copy_gbs:.0f} GB/s (1 GiB buffer, read + write)")
    lat_b_ms = np.array(lat_b) * 1e3
    out = {
        "metric": METRIC,
        "value": round(gteps_b, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el_b * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"BFS+SSSP, {k} roots, directed R-MAT scale {scale} edge factor 16 "
                        f"(a=.57 b=.19 c=.19) seed {seed}, u32 weights [1,128]",
            "mode": f"batched: {k} BFS roots in one k-bit traversal, then {k} SSSP roots in one "
                    f"k-wide traversal; results equal the single-query path for every root and "
                    f"stay in HBM (PCIe-inclusive rate under pcie_inclusive)",
            "graph_edges": int(g.info["nnz"]),
            "queries_per_step": 2 * k,
        },
        "queries_per_s": round(2 * k * args.steps / el_b, 2),
        "p50_ms": round(float(np.percentile(lat_b_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_b_ms, 99)), 3),
        "pcie_inclusive": {"value": round(gteps_p, 4), "unit": "GTEPS",
                           "ms_per_step": round(el_p * 1e3 / args.steps, 3),
                           "note": "batched path with every per-query result copied to page-locked host memory"},
        "single_query": single,
        "speedup_vs_single": round(gteps_b / single["value"], 2) if single else None,
        "roofline": roofline(st_b),
        "device_copy_gbs": round(copy_gbs, 1),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(rec, n, roots[:args.cpu_roots], outdeg, hops, dist)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
