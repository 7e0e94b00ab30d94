"""Summarise rocprofv3 SQLite results (kernel-trace stats and per-dispatch PMC counters).

  python -m sllg.rocprof_summary trace <results.db>            kernel, calls, mean/min/max us
  python -m sllg.rocprof_summary pmc <results.db> [kernel ...]   mean counter value per dispatch
"""
import sqlite3
import sys
from collections import defaultdict


def trace(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), avg(duration), min(duration), max(duration) from kernels "
                       "group by name order by sum(duration) desc").fetchall()
    out = []
    for name, cnt, avg, mn, mx in rows:
        out.append((name, cnt, avg / 1e3, mn / 1e3, mx / 1e3))
    return out


def pmc(db, kernels=None):
    con = sqlite3.connect(db)
    acc = defaultdict(list)
    for name, counter, value in con.execute("select kernel_name, counter_name, value from counters_collection"):
        if kernels and not any(k in name for k in kernels):
            continue
        acc[(name, counter)].append(value)
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def main():
    mode, db = sys.argv[1], sys.argv[2]
    if mode == "trace":
        print(f"{'kernel':60s} {'calls':>6s} {'mean_us':>10s} {'min_us':>10s} {'max_us':>10s}")
        for name, cnt, avg, mn, mx in trace(db):
            print(f"{name[:60]:60s} {cnt:6d} {avg:10.3f} {mn:10.3f} {mx:10.3f}")
    else:
        # full names: template arguments tell a kernel's instances apart
        for (name, counter), (mean, cnt) in sorted(pmc(db, sys.argv[3:]).items()):
            print(f"{name} {counter} dispatches={cnt} mean={mean:.1f}")


if __name__ == "__main__":
    main()
