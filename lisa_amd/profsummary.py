"""Summarise a rocprofv3 results database (rocpd SQLite) as plain text.

    python -m lisa_amd.profsummary RESULTS.db [--kernel SUBSTR] [--json]

Kernel-trace runs: calls, total and average duration per kernel.
Counter runs (--pmc): the per-dispatch average of every collected counter
per kernel (FETCH_SIZE / WRITE_SIZE are in KiB per dispatch).
"""
from __future__ import annotations

import argparse
import json
import sqlite3
from collections import defaultdict


def kernel_stats(db):
    cur = db.execute("select name, duration from kernels")
    acc = defaultdict(list)
    for name, dur in cur.fetchall():
        acc[name].append(float(dur))
    out = []
    for name, d in acc.items():
        out.append({"kernel": name, "calls": len(d), "total_ms": sum(d) / 1e6,
                    "avg_ms": sum(d) / len(d) / 1e6, "min_ms": min(d) / 1e6,
                    "max_ms": max(d) / 1e6})
    out.sort(key=lambda r: -r["total_ms"])
    return out


def counter_stats(db):
    try:
        cur = db.execute("select name, counter_name, counter_value from pmc_events")
    except sqlite3.Error:
        return []
    acc = defaultdict(list)
    for name, cname, val in cur.fetchall():
        acc[(name, cname)].append(float(val))
    out = [{"kernel": k, "counter": c, "dispatches": len(v), "avg_per_dispatch": sum(v) / len(v)}
           for (k, c), v in acc.items()]
    out.sort(key=lambda r: (r["counter"], -r["avg_per_dispatch"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--kernel", default="")
    ap.add_argument("--json", action="store_true")
    a = ap.parse_args()
    db = sqlite3.connect(a.db)
    ks = [r for r in kernel_stats(db) if a.kernel in r["kernel"]]
    cs = [r for r in counter_stats(db) if a.kernel in r["kernel"]]
    if a.json:
        print(json.dumps({"kernels": ks, "counters": cs}, indent=1))
        return
    print("| kernel | calls | total ms | avg ms | min ms | max ms |")
    print("|---|---|---|---|---|---|")
    for r in ks:
        print(f"| `{r['kernel'][:90]}` | {r['calls']} | {r['total_ms']:.3f} | {r['avg_ms']:.3f} "
              f"| {r['min_ms']:.3f} | {r['max_ms']:.3f} |")
    if cs:
        print()
        print("| kernel | counter | dispatches | avg per dispatch |")
        print("|---|---|---|---|")
        for r in cs:
            print(f"| `{r['kernel'][:90]}` | {r['counter']} | {r['dispatches']} "
                  f"| {r['avg_per_dispatch']:.1f} |")


if __name__ == "__main__":
    main()
