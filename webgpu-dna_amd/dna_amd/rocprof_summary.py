#!/usr/bin/env python3
"""Summaries of rocprofv3 result databases (rocpd SQLite), used for the text files under
profiles/.

  rocprof_summary.py trace <results.db>            per-kernel calls, total/avg/min/max
                                                   duration (ms), VGPRs, LDS, grid, block
  rocprof_summary.py pmc <results.db> [substring]  per-kernel counter totals and the value
                                                   per dispatch (FETCH_SIZE / WRITE_SIZE in KiB)
"""
import sqlite3
import sys
from collections import defaultdict


def trace(path):
    db = sqlite3.connect(path)
    rows = db.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration), "
                      "max(vgpr_count), max(lds_size), max(grid_x), max(workgroup_x) from kernels "
                      "group by name order by sum(duration) desc").fetchall()
    tot = sum(r[2] for r in rows)
    print(f"{'kernel':60s} {'calls':>5s} {'total_ms':>9s} {'avg_ms':>9s} {'min_ms':>8s} {'max_ms':>8s} "
          f"{'pct':>6s} {'vgpr':>4s} {'lds':>6s} {'grid':>8s} {'wg':>4s}")
    for n, c, s, a, lo, hi, vg, lds, gx, wg in rows:
        print(f"{n[:60]:60s} {c:5d} {s / 1e6:9.3f} {a / 1e6:9.4f} {lo / 1e6:8.4f} {hi / 1e6:8.4f} "
              f"{100 * s / tot:6.2f} {vg:4d} {lds:6d} {gx:8d} {wg:4d}")


def pmc(path, sub=None):
    db = sqlite3.connect(path)
    rows = db.execute("select name, counter_name, sum(counter_value), count(distinct dispatch_id) "
                      "from pmc_events group by name, counter_name").fetchall()
    d = defaultdict(dict)
    for n, c, v, k in rows:
        d[n][c] = (v, k)
    print(f"{'kernel':60s} {'counter':16s} {'total':>14s} {'dispatches':>10s} {'per_dispatch':>14s}")
    for n in sorted(d):
        if sub and sub not in n:
            continue
        for c, (v, k) in sorted(d[n].items()):
            print(f"{n[:60]:60s} {c:16s} {v:14.6g} {k:10d} {v / max(k, 1):14.6g}")


if __name__ == "__main__":
    if len(sys.argv) < 3 or sys.argv[1] not in ("trace", "pmc"):
        print(__doc__)
        sys.exit(2)
    if sys.argv[1] == "trace":
        trace(sys.argv[2])
    else:
        pmc(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else None)
