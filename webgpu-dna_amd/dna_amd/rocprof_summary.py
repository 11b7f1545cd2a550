#!/usr/bin/env python3
"""Summaries of rocprofv3 result databases (rocpd SQLite), used for the text files under
profiles/.

  rocprof_summary.py trace <results.db>            per-kernel calls, total/avg/min/max
                                                   duration (ms), VGPRs, LDS, grid, block
  rocprof_summary.py pmc <results.db> [substring]  per-kernel counter totals and the value
                                                   per dispatch (FETCH_SIZE / WRITE_SIZE in KiB)
  rocprof_summary.py dispatches <results.db> <substring>
                                                   every dispatch of the matching kernels in
                                                   start order (queue, start, duration)
  rocprof_summary.py traffic <fetch.db> <write.db> [f_fetch f_write]
                                                   per-dispatch FETCH_SIZE / WRITE_SIZE and the
                                                   bytes after copy-calibration factors
  rocprof_summary.py utilization <util.db> [<fetch.db> <write.db> f_fetch f_write]
                                                   per kernel and launch shape: average duration,
                                                   the derived counters of util.db (VALUBusy,
                                                   OccupancyPercent, ...) averaged over dispatches,
                                                   and HBM GB/s from the traffic runs
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


def dispatches(path, sub):
    """Every dispatch of the kernels whose name contains `sub`, in start order: dispatch id,
    queue, start offset (ms from the first one), duration (ms), grid and workgroup size."""
    db = sqlite3.connect(path)
    rows = db.execute("select dispatch_id, queue_id, start, duration, grid_x, workgroup_x, name "
                      "from kernels where name like ? order by start", (f"%{sub}%",)).fetchall()
    if not rows:
        return
    t0 = rows[0][2]
    print(f"{'dispatch':>8s} {'queue':>5s} {'start_ms':>10s} {'dur_ms':>9s} {'grid':>8s} {'wg':>4s}")
    for d, q, st, du, gx, wg, n in rows:
        print(f"{d:8d} {q:5d} {(st - t0) / 1e6:10.3f} {du / 1e6:9.3f} {gx:8d} {wg:4d}")


def traffic(fetch_db, write_db, f_fetch=1.0, f_write=1.0, top=16):
    """Per-dispatch FETCH_SIZE and WRITE_SIZE (KiB) of two single-counter runs of the same
    command, and the bytes they stand for after the copy-kernel calibration factors."""
    def per_dispatch(path, counter):
        db = sqlite3.connect(path)
        rows = db.execute("select dispatch_id, name, sum(counter_value) from pmc_events "
                          "where counter_name = ? group by dispatch_id", (counter,)).fetchall()
        return {(d, n): v for d, n, v in rows}
    fe, wr = per_dispatch(fetch_db, "FETCH_SIZE"), per_dispatch(write_db, "WRITE_SIZE")
    rows = [(n, d, fe[(d, n)], wr.get((d, n), 0.0)) for (d, n) in fe]
    rows.sort(key=lambda r: -(r[2] * f_fetch + r[3] * f_write))
    print(f"{'kernel':34s} {'dispatch':>8s} {'FETCH_KiB':>12s} {'WRITE_KiB':>12s} {'read_MB':>10s} {'write_MB':>10s}")
    for n, d, f, w in rows[:top]:
        short = n.split("(")[0].replace("void ", "")
        print(f"{short[:34]:34s} {d:8d} {f:12.1f} {w:12.1f} {f * 1024 * f_fetch / 1e6:10.1f} {w * 1024 * f_write / 1e6:10.1f}")


def _short(n):
    return n.split("(")[0].replace("void ", "")


def utilization(util_db, fetch_db=None, write_db=None, f_fetch=1.0, f_write=1.0):
    """Per (kernel, grid, workgroup): dispatches, mean duration (ms) and the mean per-dispatch
    value of every counter in util_db; with the traffic runs, mean HBM bytes per dispatch
    (copy-calibrated) over the mean duration of those same dispatches, in GB/s."""
    def per_dispatch(path, counter=None):
        db = sqlite3.connect(path)
        q = ("select p.dispatch_id, p.counter_name, sum(p.counter_value) from pmc_events p "
             + ("where p.counter_name = ? " if counter else "") + "group by p.dispatch_id, p.counter_name")
        vals = defaultdict(dict)
        for d, c, v in db.execute(q, (counter,) if counter else ()):
            vals[d][c] = v
        shape = {d: (_short(n), gx, wg, du) for d, n, gx, wg, du in
                 db.execute("select dispatch_id, name, grid_x, workgroup_x, duration from kernels")}
        return vals, shape
    vals, shape = per_dispatch(util_db)
    counters = sorted({c for v in vals.values() for c in v})
    groups = defaultdict(list)
    for d, (n, gx, wg, du) in shape.items():
        groups[(n, gx, wg)].append((du, vals.get(d, {})))
    hbm = {}
    if fetch_db and write_db:
        fv, fs = per_dispatch(fetch_db, "FETCH_SIZE")
        wv, _ = per_dispatch(write_db, "WRITE_SIZE")
        acc = defaultdict(lambda: [0.0, 0.0])
        for d, (n, gx, wg, du) in fs.items():
            if d in fv and d in wv:
                b = (fv[d]["FETCH_SIZE"] * f_fetch + wv[d]["WRITE_SIZE"] * f_write) * 1024
                acc[(n, gx, wg)][0] += b
                acc[(n, gx, wg)][1] += du
        hbm = {k: (b / t if t else 0.0) for k, (b, t) in acc.items()}
    head = f"{'kernel':26s} {'grid':>8s} {'wg':>4s} {'calls':>5s} {'avg_ms':>9s}"
    head += "".join(f" {c[:16]:>16s}" for c in counters) + (f" {'HBM_GB/s':>9s}" if hbm else "")
    print(head)
    order = sorted(groups, key=lambda k: -sum(du for du, _ in groups[k]))
    for k in order:
        g = groups[k]
        line = f"{k[0][:26]:26s} {k[1]:8d} {k[2]:4d} {len(g):5d} {sum(du for du, _ in g) / len(g) / 1e6:9.4f}"
        for c in counters:
            xs = [v[c] for _, v in g if c in v]
            line += f" {sum(xs) / len(xs):16.3f}" if xs else f" {'-':>16s}"
        if hbm:
            line += f" {hbm[k]:9.1f}" if k in hbm else f" {'-':>9s}"
        print(line)


if __name__ == "__main__":
    if len(sys.argv) < 3 or sys.argv[1] not in ("trace", "pmc", "traffic", "dispatches", "utilization"):
        print(__doc__)
        sys.exit(2)
    if sys.argv[1] == "trace":
        trace(sys.argv[2])
    elif sys.argv[1] == "dispatches":
        dispatches(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else "")
    elif sys.argv[1] == "pmc":
        pmc(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else None)
    elif sys.argv[1] == "utilization":
        a = sys.argv[2:]
        utilization(a[0], *(a[1:3] if len(a) >= 3 else []), *[float(x) for x in a[3:5]])
    else:
        f = [float(x) for x in sys.argv[4:6]] if len(sys.argv) >= 6 else [1.0, 1.0]
        traffic(sys.argv[2], sys.argv[3], f[0], f[1])
