"""Benchmark: Louvain on one MI355X -- edges/s summed over passes on R-MAT
scale 24, plus graphs/s of the 4,096-graph batch against one graph at a time.

A step is one full Louvain run (all passes, modularity optimisation and
aggregation) over a device-resident synthetic graph; the graph is restored
from its resident copy between steps, outside the timed region.

    python bench.py [--config headline|c1|c2|c3|c4|c5] [--steps K] [--warmup W] [--gpus 1]

The default (headline) prints one JSON line for R-MAT-24 (value = sum over
passes of CSR entries / run time) with a "batch" object for the 4,096-graph
SBM batch (graphs/s batched and one at a time on the same GPU, every graph's
result checked against the oracle's committed digests).
roofline: the dominant kernel family's algorithmic bytes / its launch time
(hipEvents on the engine's stream), against 8 TB/s.
cpu_baseline: the CPU oracle on all host cores and on one core, timed on pass
1's initialisation and first sweep of the same R-MAT-24 graph and
extrapolated to the whole run by entry visits.
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
METRIC = "Louvain edges/s summed over passes + graphs/s batched vs plain, 1 GPU"
DTYPE = "f32 gains, exact i64 fixed-point sums"
GOLDEN = os.path.join(ROOT, "tests", "golden")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def qbits(q):
    return int(np.float32(q).view(np.uint32))


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def measured_traffic(workload, fam, launches_per_run):
    """HBM bytes per launch of a kernel family, from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE + WRITE_SIZE per run, scaled
    by a plain copy kernel), or None without one."""
    import re
    path = os.path.join(ROOT, "profiles", f"pmc_{workload}.md")
    if not os.path.exists(path) or not launches_per_run:
        return None
    with open(path) as f:
        m = re.search(r"```json\n(.*?)```", f.read(), re.S)
    if not m:
        return None
    d = json.loads(m.group(1))["per_run_bytes"].get(fam)
    if not d:
        return None
    return (d["fetch"] + d["write"]) / launches_per_run


# Summaries of rocprofv3 result databases (the text committed under
# profiles/).  Kernel name -> the engine's kernel family (lv_profile_get);
# scans, fills and copies shared by several phases stay unattributed.
FAMILY_PATTERNS = [
    ("relabel", r"k_relabel_(map|deg|rows|heavy)"),
    ("aggregate", r"k_agg_|k_aggp_|k_scatter_by|k_relabel_newid"),
    ("move_global", r"k_hp_|k_heavy_(insert|gain|lanes|decide|clean)"),
    ("r1", r"k_r1_groups|k_heavy_stage|k_heavy_r1"),
    ("move_block", r"k_move_block"),
    ("move_wave", r"k_move_wave"),
    ("move_small", r"k_move_small|k_move_narrow"),
    ("apply", r"k_apply"),
    ("sumsq", r"k_sumsq|k_collapse_decide"),
    ("classify", r"k_classify_|k_heavy_caps|k_heavy_chunks|k_owner_fill"),
]


def family_of(name):
    import re
    for fam, pat in FAMILY_PATTERNS:
        if re.search(pat, name):
            return fam
    return None


def _kernel_short(name):
    import re
    return re.sub(r"\(.*", "", name)


def summarize_kernels(db):
    """Markdown: calls, total / average / max duration per kernel of a
    `rocprofv3 --kernel-trace` database, then per family."""
    import sqlite3
    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    ni = cols.index("name") if "name" in cols else cols.index("kernel_name")
    si, ei = cols.index("start"), cols.index("end")
    agg = {}
    for r in con.execute("select * from kernels"):
        a = agg.setdefault(_kernel_short(r[ni]), [0, 0.0, 0.0])
        d = (r[ei] - r[si]) / 1e3
        a[0] += 1
        a[1] += d
        a[2] = max(a[2], d)
    tot = sum(a[1] for a in agg.values())
    print("| kernel | calls | total us | avg us | max us | % |")
    print("|---|---:|---:|---:|---:|---:|")
    for nm, a in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        if 100 * a[1] / tot < 0.05:
            continue
        print(f"| `{nm[:90]}` | {a[0]} | {a[1]:.1f} | {a[1] / a[0]:.2f} | {a[2]:.1f} | "
              f"{100 * a[1] / tot:.2f} |")
    print(f"\nkernel time {tot / 1e3:.2f} ms over {sum(a[0] for a in agg.values())} dispatches "
          "(kernels under 0.05 % not listed)\n")
    fam = {}
    for nm, a in agg.items():
        d = fam.setdefault(family_of(nm) or "(shared scans/fills/copies)", [0, 0.0])
        d[0] += a[0]
        d[1] += a[1]
    print("| family | dispatches | total ms | avg us | % |")
    print("|---|---:|---:|---:|---:|")
    for f, d in sorted(fam.items(), key=lambda kv: -kv[1][1]):
        print(f"| {f} | {d[0]} | {d[1] / 1e3:.3f} | {d[1] / d[0]:.2f} | {100 * d[1] / tot:.2f} |")


def _counter_sums(db, counter):
    import sqlite3
    con = sqlite3.connect(db)
    tot, disp = {}, {}
    for name, cn, val, did in con.execute(
            "select name, counter_name, counter_value, dispatch_id from pmc_events"):
        if cn != counter:
            continue
        nm = _kernel_short(name)
        tot[nm] = tot.get(nm, 0.0) + float(val)
        disp.setdefault(nm, set()).add(did)
    return tot, {k: len(v) for k, v in disp.items()}


def summarize_pmc(fetch_db, write_db, cfetch_db, cwrite_db, copy_bytes):
    """Markdown + JSON: HBM bytes per kernel family from two counter runs of
    the same command (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each
    scaled by what the counter reads for a plain copy kernel that reads and
    writes copy_bytes (its largest dispatch)."""
    fetch, calls = _counter_sums(fetch_db, "FETCH_SIZE")
    write, _ = _counter_sums(write_db, "WRITE_SIZE")
    kf = max(_counter_sums(cfetch_db, "FETCH_SIZE")[0].values()) * 1024.0
    kw = max(_counter_sums(cwrite_db, "WRITE_SIZE")[0].values()) * 1024.0
    sf, sw = copy_bytes / kf, copy_bytes / kw
    fam = {}
    for nm in set(fetch) | set(write):
        f = family_of(nm)
        if f is None:
            continue
        d = fam.setdefault(f, {"fetch": 0.0, "write": 0.0, "dispatches": 0})
        d["fetch"] += fetch.get(nm, 0.0) * 1024.0 * sf
        d["write"] += write.get(nm, 0.0) * 1024.0 * sw
        d["dispatches"] += calls.get(nm, 0)
    print(f"copy calibration: FETCH x{sf:.4f}, WRITE x{sw:.4f} ({copy_bytes / 1e9:.3f} GB copy "
          f"read {kf / 1e9:.3f} GB, wrote {kw / 1e9:.3f} GB)\n")
    print("| family | dispatches | fetch GB | write GB |")
    print("|---|---:|---:|---:|")
    for f, d in sorted(fam.items(), key=lambda kv: -(kv[1]["fetch"] + kv[1]["write"])):
        print(f"| {f} | {d['dispatches']} | {d['fetch'] / 1e9:.3f} | {d['write'] / 1e9:.3f} |")
    print("\n```json")
    print(json.dumps({"copy_scale": {"fetch": sf, "write": sw},
                      "per_run_bytes": {f: {"fetch": d["fetch"], "write": d["write"]}
                                        for f, d in sorted(fam.items())}}, indent=1))
    print("```")


def cpu_sample(o, threads, ne, res):
    """The oracle `o` (holding the workload's graph) on `threads` host threads:
    pass 1's initialisation and first sweep on the full graph, extrapolated to
    the whole run by entry visits (each pass reads its entries once at init
    and once per sweep, for the moves and again for Q); aggregation time is
    left out, which favours the CPU."""
    o.set_threads(threads)
    t0 = time.perf_counter()
    o.pass_init()
    o.sweep()
    dt = time.perf_counter() - t0
    t_full = dt * res.sum_edge_visits / (2 * ne)
    return dt, t_full, res.sum_edges / t_full


def cpu_baseline_c3(graph, res, single):
    import oracle
    model, nproc, avail = cpu_info()
    threads = max(1, min(16, avail))  # the host share of one GPU
    row, col, w = graph
    ne = int(col.size)
    o = oracle.Oracle(row.astype(np.uint64), col, w)
    dt, t_full, value = cpu_sample(o, threads, ne, res)
    log(f"cpu baseline: {threads} threads, sample {dt:.2f} s, extrapolated run {t_full:.1f} s")
    out = {"value": value, "unit": "edges/s", "cores": threads, "kind": "port",
           "sample": f"pass 1 init + first sweep of the full R-MAT-24 graph ({ne} entries) on "
                     f"{threads} threads of the C oracle: {dt:.2f} s; whole run extrapolated "
                     f"by entry visits ({res.sum_edge_visits} / {2 * ne}) = {t_full:.1f} s, "
                     f"aggregation excluded",
           "cpu_model": model, "nproc": nproc, "affinity_cpus": avail}
    if single:
        dt1, t1, v1 = cpu_sample(o, 1, ne, res)  # pass_init restarts from singletons
        log(f"cpu baseline: 1 thread, sample {dt1:.2f} s, extrapolated run {t1:.1f} s")
        out["single_core"] = {"value": v1, "unit": "edges/s", "cores": 1,
                              "sample_s": dt1, "extrapolated_run_s": t1}
    o.close()
    return out


def bench_plain(lv, W, cfg, args):
    """One graph resident in HBM, K timed runs; returns the JSON line."""
    t0 = time.perf_counter()
    graph = W.make_graph(lv, cfg)
    row, col, w = graph
    log(f"{cfg['workload']}: {row.size - 1} vertices, {col.size} entries "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    g = lv.Graph(row, col, w)  # host checks, device buffers, PCIe upload
    create_s = time.perf_counter() - t0
    if args.tuning:
        g.set_tuning(**args.tuning)
    tuning = g.get_tuning()
    if not args.no_tune:  # untimed: candidate launch shapes on this graph, fastest kept
        t0 = time.perf_counter()
        tuning, tms = g.autotune(W.MIN_GAIN)
        log(f"autotune: {tuning} ({tms:.2f} ms/run, search {time.perf_counter() - t0:.1f} s)")
    # warmup runs profile every kernel family to find the dominant one
    g.profile(True)
    fam, wst = None, None
    for i in range(args.warmup):
        g.reset()
        r = g.run(W.MIN_GAIN)
        log(f"warmup {i}: {r.alg_ms:.2f} ms Q={r.q_final:.6f}")
    if args.warmup:
        wst = g.profile_get()
        fam = max(wst, key=lambda k: wst[k]["ms"])
    # timed runs carry no events: an event record costs a few microseconds,
    # a fifth of a run on the co-purchase graph
    g.profile(False)
    g.profile_reset()
    times, edges, res = [], [], None
    bytes_alg = 0.0
    for i in range(args.steps):
        g.reset()
        t = time.perf_counter()
        res = g.run(W.MIN_GAIN)
        times.append(time.perf_counter() - t)
        edges.append(res.sum_edges)
        bytes_alg += res.bytes_alg
        log(f"step {i}: {res.alg_ms:.2f} ms ({times[-1] * 1e3:.2f} ms wall) Q={res.q_final:.6f} "
            f"passes={res.passes} sweeps={res.total_iters} communities={res.n_final}")
    assign = g.assignment()
    total_s = sum(times)
    value = sum(edges) / total_s
    # one more run of the same workload with HIP events on the dominant family
    # only (every family without a warmup): its launch durations give the
    # roofline
    roofline, stats = None, {}
    if not args.no_roofline:
        g.profile(True)
        g.profile_select([fam] if fam else None)
        g.profile_reset()
        g.reset()
        pres = g.run(W.MIN_GAIN)
        stats = g.profile_get()
        g.profile(False)
        log(f"profiled run: {pres.alg_ms:.2f} ms")
        if fam is None:
            fam = max(stats, key=lambda k: stats[k]["ms"])
        st = stats[fam]
        avg_ms = st["ms"] / max(st["launches"], 1)
        avg_bytes = st["bytes"] / max(st["launches"], 1)
        achieved = avg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        roofline = {"bound": "hbm", "kernel": fam, "achieved": achieved,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic(cfg["workload"], fam, st["launches"]),
                    "launches_per_run": int(st["launches"]),
                    "avg_launch_ms": avg_ms, "bytes_per_launch": avg_bytes}
    g.close()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_s / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE,
        "data": "synthetic (seeded generator)",
        "config": {"workload": cfg["workload"], "vertices": int(row.size - 1),
                   "entries": int(col.size), "min_gain": W.MIN_GAIN,
                   "modularity": float(res.q_final), "passes": int(res.passes),
                   "sweeps": int(res.total_iters), "communities": int(res.n_final),
                   "tuning": tuning, "autotuned": not args.no_tune},
        "roofline": roofline,
        # the boundary takes host buffers: one upload plus one run
        "pcie_inclusive": {"create_ms": create_s * 1e3,
                           "edges_per_s": (sum(edges) / args.steps) /
                                          (create_s + total_s / args.steps)},
        "whole_run_hbm": {"bytes_alg_per_step": bytes_alg / args.steps,
                          "achieved_gbs": bytes_alg / total_s / 1e9,
                          "frac": bytes_alg / total_s / 1e9 / HBM_PEAK_GBS},
        # per family over the warmup runs (every family profiled)
        "kernels": {k: {"ms": round(v["ms"], 4), "launches": int(v["launches"]),
                        "gbs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 else 0.0}
                    for k, v in (wst if args.warmup else stats).items()},
    }
    digest = os.path.join(GOLDEN, f"digests_{cfg['workload']}.json")
    if os.path.exists(digest):  # the oracle's result for this exact input
        with open(digest) as f:
            d = json.load(f)
        line["identical_to_oracle_digest"] = bool(
            qbits(res.q_final) == d["q_bits"] and res.passes == d["passes"] and
            res.total_iters == d["sweeps"] and res.n_final == d["n_final"] and
            res.sum_edges == d["sum_edges"] and W.sha256_arrays(assign) == d["assign"])
    if not args.no_cpu:
        if cfg["gen"] == "rmat" and row.size - 1 >= (1 << 22):
            line["cpu_baseline"] = cpu_baseline_c3(graph, res, not args.no_cpu_single)
        else:  # small graphs: the oracle's whole run, one core
            import oracle
            o = oracle.Oracle(row.astype(np.uint64), col, w)
            t = time.perf_counter()
            r = o.run(W.MIN_GAIN)
            dt = time.perf_counter() - t
            model, nproc, avail = cpu_info()
            line["cpu_baseline"] = {"value": r.sum_edges / dt, "unit": "edges/s", "cores": 1,
                                    "kind": "port", "cpu_model": model, "nproc": nproc,
                                    "sample": f"full workload, single-core C oracle, {dt:.2f} s"}
    return line


def bench_batch(lv, W, cfg, args):
    """graphs/s of the batched engine vs one-graph-at-a-time plain runs on the
    same GPU; every graph's result against the oracle's digests."""
    t0 = time.perf_counter()
    graphs = W.make_graph(lv, cfg)
    nv = sum(int(r.size) - 1 for r, _, _ in graphs)
    ne = sum(int(c.size) for _, c, _ in graphs)
    log(f"{cfg['workload']}: {len(graphs)} graphs, {nv} vertices, {ne} entries "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    b = lv.Batch(graphs)
    if args.tuning:
        b.set_tuning(**args.tuning)
    tuning = b.get_tuning()
    if not args.no_tune:
        t0 = time.perf_counter()
        tuning, tms = b.autotune(W.MIN_GAIN)
        log(f"autotune: {tuning} ({tms:.1f} ms/run, search {time.perf_counter() - t0:.1f} s)")
    for i in range(args.warmup):
        b.reset()
        per, tot = b.run(W.MIN_GAIN)
        log(f"batch warmup {i}: {tot.alg_ms:.1f} ms")
    times = []
    for i in range(args.steps):
        b.reset()
        t = time.perf_counter()
        per, tot = b.run(W.MIN_GAIN)
        times.append(time.perf_counter() - t)
        log(f"batch step {i}: {tot.alg_ms:.1f} ms ({times[-1] * 1e3:.1f} ms wall) "
            f"passes={tot.passes} sweeps={tot.total_iters}")
    assigns = b.assignments()
    b.close()
    ms_batch = sum(times) / len(times) * 1e3
    batch_gps = len(graphs) / (ms_batch * 1e-3)
    # every graph against the oracle's committed result for it
    with open(os.path.join(GOLDEN, f"digests_{cfg['workload']}.json")) as f:
        d = json.load(f)

    def matches(i, q, passes, sweeps, n_final, sum_edges, assign):
        return (qbits(q), passes, sweeps, n_final, sum_edges,
                W.sha256_arrays(np.asarray(assign, np.uint32))[:16]) == \
            (d["q_bits"][i], d["passes"][i], d["sweeps"][i], d["n_final"][i],
             d["sum_edges"][i], d["assign"][i])

    batch_ok = sum(matches(i, per[i].q_final, per[i].passes, per[i].total_iters,
                           per[i].n_final, per[i].sum_edges, assigns[i])
                   for i in range(len(graphs)))
    # plain path, one graph at a time, on a strided sample
    sample = list(range(0, len(graphs), max(1, len(graphs) // 256)))
    plain_ms = 0.0
    plain_ok = 0
    for i in sample:
        r, c, w = graphs[i]
        with lv.Graph(r, c, w) as g:
            res = g.run(W.MIN_GAIN)
            a = g.assignment()
        plain_ms += res.alg_ms
        plain_ok += matches(i, res.q_final, res.passes, res.total_iters, res.n_final,
                            res.sum_edges, a)
    plain_gps = len(sample) / (plain_ms * 1e-3)
    log(f"batch {batch_gps:.0f} graphs/s, plain {plain_gps:.1f} graphs/s ({len(sample)} graphs); "
        f"oracle digests matched: batch {batch_ok}/{len(graphs)}, plain {plain_ok}/{len(sample)}")
    out = {"workload": cfg["workload"], "graphs": len(graphs), "vertices": nv, "entries": ne,
           "graphs_per_s": batch_gps, "ms_per_batch": ms_batch, "steps": args.steps,
           "plain_graphs_per_s": plain_gps, "plain_sample_graphs": len(sample),
           "batched_over_plain": batch_gps / plain_gps,
           "edges_per_s": sum(p.sum_edges for p in per) / (ms_batch * 1e-3),
           "identical_to_oracle_digests": bool(batch_ok == len(graphs) and
                                               plain_ok == len(sample)),
           "graphs_checked": len(graphs) + len(sample),
           "tuning": tuning, "autotuned": not args.no_tune}
    if not args.no_cpu:
        import oracle
        cs = sample[:8]
        t = time.perf_counter()
        for i in cs:
            r, c, w = graphs[i]
            oracle.Oracle(r.astype(np.uint64), c, w).run(W.MIN_GAIN)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": len(cs) / dt, "unit": "graphs/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{len(cs)} graphs of the batch, single-core C oracle, "
                                         f"{dt:.2f} s"}
    return out


def bench_serve(lv, W, cfg, args):
    """Poisson request stream through the serving loop: p50/p99 latency and
    graphs/s, every result checked against the oracle's digests."""
    nreq, rate, seed = cfg["args"]
    t0 = time.perf_counter()
    sbm, rmat = W.serve_pool(lv, seed)
    pool = sbm + rmat
    log(f"pool: {len(sbm)} SBM + {len(rmat)} R-MAT graphs "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    with open(os.path.join(GOLDEN, "digests_serve-pool.json")) as f:
        d = json.load(f)["graphs"]
    solo_ms = []
    for r, c, w in pool:  # one at a time, for the comparison rate
        with lv.Graph(r, c, w) as g:
            solo_ms.append(g.run(W.MIN_GAIN).alg_ms)
    arrivals, picks = W.serve_stream(nreq, rate, seed, len(sbm), len(rmat))
    results = []
    with lv.Server(slots=8, hbm_budget=200 << 30, min_gain=W.MIN_GAIN) as s:
        start = time.perf_counter()
        for rid in range(nreq):
            delay = arrivals[rid] - (time.perf_counter() - start)
            if delay > 0:
                time.sleep(delay)
            s.submit(rid, *pool[picks[rid]])
            results += s.poll()
        s.drain()
        results += s.poll(4096)
    assert len(results) == nreq
    lat = np.array([r.t_end_ms - r.t_submit_ms for r in results])
    span_s = (max(r.t_end_ms for r in results) - min(r.t_submit_ms for r in results)) / 1e3
    ok = sum(r.status == 0 and
             (qbits(r.q_final), r.passes, r.total_iters, r.n_final, r.sum_edges) ==
             (d[picks[r.id]]["q_bits"], d[picks[r.id]]["passes"], d[picks[r.id]]["sweeps"],
              d[picks[r.id]]["n_final"], d[picks[r.id]]["sum_edges"]) for r in results)
    solo_gps = nreq / (sum(solo_ms[p] for p in picks) * 1e-3)
    log(f"{nreq} requests over {span_s:.2f} s; p50 {np.percentile(lat, 50):.1f} ms "
        f"p99 {np.percentile(lat, 99):.1f} ms; oracle digests matched {ok}/{nreq}")
    return {
        "metric": "Louvain served graphs/s (Poisson stream, 1 GPU)",
        "value": nreq / span_s, "unit": "graphs/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": span_s * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": DTYPE, "data": "synthetic (seeded generator)",
        "config": {"workload": cfg["workload"], "requests": nreq, "rate_rps": rate,
                   "pool": "12 SBM (1k-200k vertices) + R-MAT scale 14-21, w ~ U(0.5,1.5)",
                   "slots": 8, "hbm_budget_gb": 200},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean())},
        "one_at_a_time_graphs_per_s": solo_gps,
        "identical_to_oracle_digests": bool(ok == nreq),
    }


def _tuning_arg(text):
    """'heavy_min_degree=8192,bucket_entries=512' -> {field: int}"""
    out = {}
    for item in text.split(","):
        k, _, v = item.partition("=")
        if not k or not v:
            raise argparse.ArgumentTypeError(f"expected FIELD=VALUE, got {item!r}")
        out[k.strip()] = int(v)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="headline",
                    choices=["headline", "c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-cpu-single", action="store_true",
                    help="skip the single-core CPU row of the R-MAT-24 baseline")
    ap.add_argument("--no-batch", action="store_true",
                    help="headline without the 4,096-graph batch object")
    ap.add_argument("--no-roofline", action="store_true",
                    help="skip the extra profiled run (exactly warmup + steps runs, e.g. "
                         "under an external profiler); roofline is then null")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (by default it picks the launch shapes on the "
                         "resident graph or batch before the timed runs)")
    ap.add_argument("--tuning", type=_tuning_arg, metavar="FIELD=V,...",
                    help="launch shapes to run with (fields of lv_tuning, e.g. the 'tuning' "
                         "object of an earlier bench line) instead of the autotuner's search")
    ap.add_argument("--summarize-kernels", metavar="DB",
                    help="print the per-kernel and per-family summary of a rocprofv3 "
                         "--kernel-trace result database and exit (no GPU)")
    ap.add_argument("--summarize-pmc", nargs=5,
                    metavar=("FETCH_DB", "WRITE_DB", "COPY_FETCH_DB", "COPY_WRITE_DB", "BYTES"),
                    help="print the per-family HBM traffic of two rocprofv3 --pmc runs "
                         "(FETCH_SIZE, WRITE_SIZE) scaled by the same counters on a copy "
                         "kernel of BYTES read and written, and exit (no GPU)")
    args = ap.parse_args()
    if args.tuning:
        args.no_tune = True
    if args.summarize_kernels:
        summarize_kernels(args.summarize_kernels)
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc[:4], float(args.summarize_pmc[4]))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU is design-only "
                                   "(graphs shard by HBM footprint, no collectives)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from loader import load_package
    lv = load_package()
    from louvain_amd import workloads as W
    if lv.device_count() < 1:
        raise RuntimeError("no GPU visible")
    if args.config == "c4":
        out = bench_batch(lv, W, W.CONFIGS["c4"], args)
        line = {"metric": METRIC, "value": out["graphs_per_s"], "unit": "graphs/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": out["ms_per_batch"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": DTYPE, "data": "synthetic (seeded generator)",
                "config": {"workload": out["workload"]}, "batch": out}
    elif args.config == "c5":
        line = bench_serve(lv, W, W.CONFIGS["c5"], args)
    else:
        name = "c3" if args.config == "headline" else args.config
        line = bench_plain(lv, W, W.CONFIGS[name], args)
        if args.config == "headline" and not args.no_batch:
            line["batch"] = bench_batch(lv, W, W.CONFIGS["c4"], args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
