"""Summarise a rocprofv3 results database: per-kernel calls, total / average / max duration.

Usage: python3 profsum.py <run_results.db>
"""
import sqlite3, sys, re
db = sys.argv[1]
con = sqlite3.connect(db)
cur = con.cursor()
cols = [r[1] for r in cur.execute("pragma table_info(kernels)")]
rows = cur.execute("select * from kernels").fetchall()
ni = cols.index('name') if 'name' in cols else cols.index('kernel_name')
si, ei = cols.index('start'), cols.index('end')
agg = {}
for r in rows:
    nm = r[ni]
    nm = re.sub(r'\(.*', '', nm)
    d = (r[ei] - r[si]) / 1e3
    a = agg.setdefault(nm, [0, 0.0, 0.0])
    a[0] += 1; a[1] += d; a[2] = max(a[2], d)
tot = sum(a[1] for a in agg.values())
print(f"{'kernel':60s} {'calls':>7s} {'total_us':>11s} {'avg_us':>9s} {'max_us':>9s} {'pct':>6s}")
for nm, a in sorted(agg.items(), key=lambda kv: -kv[1][1]):
    print(f"{nm[:60]:60s} {a[0]:7d} {a[1]:11.1f} {a[1]/a[0]:9.2f} {a[2]:9.1f} {100*a[1]/tot:6.2f}")
print(f"total kernel time {tot/1e3:.2f} ms over {sum(a[0] for a in agg.values())} dispatches")
