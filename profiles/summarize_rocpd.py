#!/usr/bin/env python3
"""Print a per-kernel summary (calls, total/avg/min/max duration, VGPRs, LDS) from a
rocprofv3 --kernel-trace database (rocpd SQLite).  Usage: summarize_rocpd.py <results.db>"""
import sqlite3
import sys

db = sqlite3.connect(sys.argv[1])
rows = db.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration), "
                  "max(vgpr_count), max(lds_size), max(grid_x), max(workgroup_x) from kernels "
                  "group by name order by sum(duration) desc").fetchall()
tot = sum(r[2] for r in rows)
print(f"{'kernel':60s} {'calls':>5s} {'total_ms':>9s} {'avg_ms':>9s} {'min_ms':>8s} {'max_ms':>8s} {'pct':>6s} {'vgpr':>4s} {'lds':>6s} {'grid':>8s} {'wg':>4s}")
for n, c, s, a, lo, hi, vg, lds, gx, wg in rows:
    print(f"{n[:60]:60s} {c:5d} {s/1e6:9.3f} {a/1e6:9.4f} {lo/1e6:8.4f} {hi/1e6:8.4f} {100*s/tot:6.2f} {vg:4d} {lds:6d} {gx:8d} {wg:4d}")
