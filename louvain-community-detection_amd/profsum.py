"""Summaries of rocprofv3 result databases.

    python3 profsum.py <run_results.db>
        per-kernel calls, total / average / max duration (kernel trace)
    python3 profsum.py --pmc <fetch.db> <write.db> <copy_fetch.db> <copy_write.db> COPY_BYTES
        HBM traffic per kernel family from two counter runs of the same command
        (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), scaled by what the same
        counters read for a plain copy kernel that reads and writes COPY_BYTES;
        prints a table and a JSON block (per-run bytes per family)
"""
import json
import re
import sqlite3
import sys

# kernel name -> kernel family of the engine's profiling (lv_kernel_stats).
# Scans and fills shared by several phases are left unattributed.
FAMILY_PATTERNS = [
    ("relabel", r"k_relabel_(map|deg|rows|heavy)"),
    ("aggregate", r"k_agg_|k_aggp_|k_scatter_by|k_relabel_newid"),
    ("move_global", r"k_hp_|k_heavy_(insert|gain|lanes|decide|clean)"),
    ("r1", r"k_r1<|k_r1v<|k_heavy_stage|k_heavy_r1"),
    ("move_block", r"k_move_block"),
    ("move_wave", r"k_move_wave"),
    ("move_small", r"k_move_small"),
    ("apply", r"k_apply"),
    ("sumsq", r"k_sumsq"),
    ("classify", r"k_classify_|k_heavy_caps|k_heavy_chunks|k_owner_fill"),
]


def family_of(name):
    for fam, pat in FAMILY_PATTERNS:
        if re.search(pat, name):
            return fam
    return None


def short(name):
    return re.sub(r"\(.*", "", name)


def kernel_trace(db):
    con = sqlite3.connect(db)
    cur = con.cursor()
    cols = [r[1] for r in cur.execute("pragma table_info(kernels)")]
    rows = cur.execute("select * from kernels").fetchall()
    ni = cols.index("name") if "name" in cols else cols.index("kernel_name")
    si, ei = cols.index("start"), cols.index("end")
    agg = {}
    for r in rows:
        nm = short(r[ni])
        d = (r[ei] - r[si]) / 1e3
        a = agg.setdefault(nm, [0, 0.0, 0.0])
        a[0] += 1
        a[1] += d
        a[2] = max(a[2], d)
    tot = sum(a[1] for a in agg.values())
    print(f"{'kernel':60s} {'calls':>7s} {'total_us':>11s} {'avg_us':>9s} {'max_us':>9s} {'pct':>6s}")
    for nm, a in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"{nm[:60]:60s} {a[0]:7d} {a[1]:11.1f} {a[1]/a[0]:9.2f} {a[2]:9.1f} "
              f"{100*a[1]/tot:6.2f}")
    print(f"total kernel time {tot/1e3:.2f} ms over {sum(a[0] for a in agg.values())} dispatches")
    fam = {}
    for nm, a in agg.items():
        f = family_of(nm) or "(shared scans/fills/copies)"
        d = fam.setdefault(f, [0, 0.0])
        d[0] += a[0]
        d[1] += a[1]
    print()
    print(f"{'family':30s} {'dispatches':>10s} {'total_ms':>10s} {'pct':>6s}")
    for f, d in sorted(fam.items(), key=lambda kv: -kv[1][1]):
        print(f"{f:30s} {d[0]:10d} {d[1] / 1e3:10.3f} {100 * d[1] / tot:6.2f}")


def counter_sums(db, counter):
    """KiB summed per kernel name, and dispatch counts."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, counter_name, counter_value, dispatch_id from pmc_events"
                       ).fetchall()
    tot, disp = {}, {}
    for name, cn, val, did in rows:
        if cn != counter:
            continue
        nm = short(name)
        tot[nm] = tot.get(nm, 0.0) + float(val)
        disp.setdefault(nm, set()).add(did)
    return tot, {k: len(v) for k, v in disp.items()}


def pmc(fetch_db, write_db, cfetch_db, cwrite_db, copy_bytes):
    fetch, calls = counter_sums(fetch_db, "FETCH_SIZE")
    write, _ = counter_sums(write_db, "WRITE_SIZE")
    cf, _ = counter_sums(cfetch_db, "FETCH_SIZE")
    cw, _ = counter_sums(cwrite_db, "WRITE_SIZE")
    # the copy kernel is the largest dispatch of the calibration run
    kf = max(cf.values()) * 1024.0
    kw = max(cw.values()) * 1024.0
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
    print(f"copy calibration: FETCH x{sf:.4f}, WRITE x{sw:.4f} "
          f"({copy_bytes / 1e9:.3f} GB copy read {kf / 1e9:.3f} GB, wrote {kw / 1e9:.3f} GB)")
    print()
    print(f"| {'family':12s} | {'dispatches':>10s} | {'fetch GB':>10s} | {'write GB':>10s} |")
    print(f"|{'-' * 14}|{'-' * 12}|{'-' * 12}|{'-' * 12}|")
    for f, d in sorted(fam.items(), key=lambda kv: -(kv[1]["fetch"] + kv[1]["write"])):
        print(f"| {f:12s} | {d['dispatches']:10d} | {d['fetch'] / 1e9:10.3f} | "
              f"{d['write'] / 1e9:10.3f} |")
    print()
    print("```json")
    print(json.dumps({"copy_scale": {"fetch": sf, "write": sw},
                      "per_run_bytes": {f: {"fetch": d["fetch"], "write": d["write"]}
                                        for f, d in fam.items()}}, indent=1))
    print("```")


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "--pmc":
        pmc(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], float(sys.argv[6]))
    else:
        kernel_trace(sys.argv[1])
