"""Benchmark: plain Louvain on one MI355X, edges/s summed over passes.

A step is one full Louvain run (all passes, modularity optimisation and
aggregation) over a device-resident synthetic graph; the graph is restored
from its resident copy between steps, outside the timed region.

    python bench.py [--config c2|c1|c3] [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  value = sum over passes of CSR entries / run time.
roofline: the dominant kernel family's algorithmic bytes / its launch time
(hipEvents on the engine's stream, over the timed steps), against 8 TB/s.
cpu_baseline: the CPU oracle (single core C restatement) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "louvain-community-detection_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MIN_GAIN = 1e-6

CONFIGS = {
    # FEM-band analogue of bcsstk16
    "c1": dict(workload="fem-band-4884", gen="fem", args=(814, 25, 9, 1)),
    # co-purchase-like Chung-Lu over a planted partition
    "c2": dict(workload="copurchase-chung-lu-262111", gen="chung_lu",
               args=(262111, 9.4, 2.5, 420, 2000, 20.0, 2)),
    # R-MAT scale 24, edge factor 16, w ~ U(0.5, 1.5)
    "c3": dict(workload="rmat-24-ef16", gen="rmat", args=(24, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 3)),
    # 4096 stochastic-block-model graphs in one segmented CSR
    "c4": dict(workload="batch-4096-sbm", gen="batch", args=(4096, 100, 50000, 8, 64, 8.0, 32.0,
                                                            10.0, 4)),
}
# the same recipe at a size the single-core oracle finishes in ~10-30 s
CPU_SAMPLES = {
    "c1": None,
    "c2": None,
    "c3": dict(workload="rmat-18-ef16 (same recipe, 1/64 of the vertices)", gen="rmat",
               args=(18, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 3)),
}


def batch_specs(G, vmin, vmax, bmin, bmax, dmin, dmax, p_ratio, seed):
    """Per-graph SBM parameters: V log-uniform, blocks and mean degree uniform."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(vmin), np.log(vmax), G)).astype(np.int64)
    blocks = rng.integers(bmin, bmax + 1, G)
    deg = rng.uniform(dmin, dmax, G)
    return [(int(n[i]), int(min(blocks[i], n[i])), float(deg[i]), p_ratio, seed * 100003 + i)
            for i in range(G)]


def make_batch(lv, args):
    from concurrent.futures import ThreadPoolExecutor
    specs = batch_specs(*args)
    with ThreadPoolExecutor(max_workers=16) as ex:  # generators release the GIL
        return list(ex.map(lambda sp: lv.gen_sbm(*sp), specs))


def make_graph(lv, cfg):
    if cfg["gen"] == "batch":
        return make_batch(lv, cfg["args"])
    return getattr(lv, "gen_" + cfg["gen"])(*cfg["args"])


def cpu_baseline(lv, cfg_name, graph):
    import oracle
    samp = CPU_SAMPLES[cfg_name]
    if samp is not None:
        row, col, w = make_graph(lv, samp)
        desc = samp["workload"]
    else:
        row, col, w = graph
        desc = "full workload"
    o = oracle.Oracle(row.astype(np.uint64), col, w)
    t0 = time.perf_counter()
    r = o.run(MIN_GAIN)
    dt = time.perf_counter() - t0
    return {"value": r.sum_edges / dt, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": f"{desc}: {row.size - 1} vertices, {col.size} entries, {r.passes} passes, "
                      f"{r.total_iters} sweeps, {dt:.2f} s single-core C oracle"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU is design-only "
                                   "(graphs shard by HBM footprint, no collectives)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from loader import load_package
    lv = load_package()
    if lv.device_count() < 1:
        raise RuntimeError("no GPU visible")
    cfg = CONFIGS[args.config]
    if cfg["gen"] == "batch":
        return bench_batch(lv, cfg, args)
    t0 = time.perf_counter()
    graph = make_graph(lv, cfg)
    row, col, w = graph
    gen_s = time.perf_counter() - t0
    print(f"[bench] {cfg['workload']}: {row.size - 1} vertices, {col.size} entries "
          f"(generated in {gen_s:.1f} s)", file=sys.stderr, flush=True)

    g = lv.Graph(row, col, w)
    for i in range(args.warmup):
        g.reset()
        r = g.run(MIN_GAIN)
        print(f"[bench] warmup {i}: {r.alg_ms:.2f} ms Q={r.q_final:.6f}", file=sys.stderr,
              flush=True)
    g.profile(True)
    g.profile_reset()
    times, edges, res = [], [], None
    bytes_alg = 0.0
    for i in range(args.steps):
        g.reset()
        t = time.perf_counter()
        res = g.run(MIN_GAIN)
        times.append(time.perf_counter() - t)
        edges.append(res.sum_edges)
        bytes_alg += res.bytes_alg
        print(f"[bench] step {i}: {res.alg_ms:.2f} ms ({times[-1] * 1e3:.2f} ms wall) "
              f"Q={res.q_final:.6f} passes={res.passes} sweeps={res.total_iters} "
              f"communities={res.n_final}", file=sys.stderr, flush=True)
    stats = g.profile_get()
    g.profile(False)
    total_s = sum(times)
    value = sum(edges) / total_s
    # dominant kernel family by device time
    fam = max((k for k in stats if k != "aggregate"), key=lambda k: stats[k]["ms"])
    st = stats[fam]
    avg_ms = st["ms"] / max(st["launches"], 1)
    avg_bytes = st["bytes"] / max(st["launches"], 1)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    line = {
        "metric": "Louvain edges/s summed over passes (plain, 1 GPU)",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_s / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 gains, exact i64 fixed-point sums",
        "data": "synthetic (seeded generator)",
        "config": {"workload": cfg["workload"], "vertices": int(row.size - 1),
                   "entries": int(col.size), "min_gain": MIN_GAIN,
                   "modularity": float(res.q_final), "passes": int(res.passes),
                   "sweeps": int(res.total_iters), "communities": int(res.n_final)},
        "roofline": {"bound": "hbm", "kernel": fam, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None},
        "whole_run_hbm": {"bytes_alg_per_step": bytes_alg / args.steps,
                          "achieved_gbs": bytes_alg / total_s / 1e9,
                          "frac": bytes_alg / total_s / 1e9 / HBM_PEAK_GBS},
        "kernels": {k: {"ms": round(v["ms"], 4), "launches": int(v["launches"]),
                        "gbs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 else 0.0}
                    for k, v in stats.items()},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lv, args.config, graph)
    g.close()
    print(json.dumps(line))


def bench_batch(lv, cfg, args):
    """graphs/s of the batched engine vs one-graph-at-a-time plain runs on the
    same GPU, with identical per-graph results."""
    t0 = time.perf_counter()
    graphs = make_graph(lv, cfg)
    nv = sum(int(r.size) - 1 for r, _, _ in graphs)
    ne = sum(int(c.size) for _, c, _ in graphs)
    print(f"[bench] {cfg['workload']}: {len(graphs)} graphs, {nv} vertices, {ne} entries "
          f"(generated in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    b = lv.Batch(graphs)
    for i in range(args.warmup):
        b.reset()
        per, tot = b.run(MIN_GAIN)
        print(f"[bench] warmup {i}: {tot.alg_ms:.1f} ms", file=sys.stderr, flush=True)
    times = []
    for i in range(args.steps):
        b.reset()
        t = time.perf_counter()
        per, tot = b.run(MIN_GAIN)
        times.append(time.perf_counter() - t)
        print(f"[bench] step {i}: {tot.alg_ms:.1f} ms ({times[-1] * 1e3:.1f} ms wall) "
              f"passes={tot.passes} sweeps={tot.total_iters}", file=sys.stderr, flush=True)
    b.close()
    batch_gps = len(graphs) / (sum(times) / len(times))
    # plain path, one graph at a time, on a strided sample; same results
    sample = list(range(0, len(graphs), max(1, len(graphs) // 256)))
    plain_ms = 0.0
    same = True
    for i in sample:
        r, c, w = graphs[i]
        with lv.Graph(r, c, w) as g:
            res = g.run(MIN_GAIN)
        plain_ms += res.alg_ms
        same &= (np.float32(res.q_final).view(np.uint32) ==
                 np.float32(per[i].q_final).view(np.uint32)) and res.passes == per[i].passes
    plain_gps = len(sample) / (plain_ms * 1e-3)
    print(f"[bench] plain: {len(sample)} graphs in {plain_ms:.1f} ms -> {plain_gps:.1f} graphs/s; "
          f"identical results: {same}", file=sys.stderr, flush=True)
    line = {
        "metric": "Louvain graphs/s, batched vs plain (1 GPU)",
        "value": batch_gps,
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sum(times) / len(times) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 gains, exact i64 fixed-point sums",
        "data": "synthetic (seeded generator)",
        "config": {"workload": cfg["workload"], "graphs": len(graphs), "vertices": nv,
                   "entries": ne, "min_gain": MIN_GAIN},
        "plain": {"graphs_per_s": plain_gps, "sample_graphs": len(sample),
                  "identical_results": bool(same)},
        "batched_over_plain": batch_gps / plain_gps,
        "edges_per_s": sum(p.sum_edges for p in per) / (sum(times) / len(times)),
    }
    if not args.no_cpu:
        import oracle
        cs = sample[:8]
        t = time.perf_counter()
        for i in cs:
            r, c, w = graphs[i]
            oracle.Oracle(r.astype(np.uint64), c, w).run(MIN_GAIN)
        dt = time.perf_counter() - t
        line["cpu_baseline"] = {"value": len(cs) / dt, "unit": "graphs/s", "cores": 1,
                                "kind": "port",
                                "sample": f"{len(cs)} graphs of the batch, single-core C oracle, "
                                          f"{dt:.2f} s"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
