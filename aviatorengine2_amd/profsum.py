"""Summaries of rocprofv3 result databases (rocpd SQLite) as text.

    python -m aviatorengine2_amd.profsum stats  RESULTS.db            per-kernel time
    python -m aviatorengine2_amd.profsum pmc    RESULTS.db [KERNEL]   counter per dispatch
"""
from __future__ import annotations

import sqlite3
import sys
from collections import defaultdict


def _short(name: str) -> str:
    for pre in ("void ", "(anonymous namespace)::"):
        name = name.replace(pre, "")
    return name.split("(")[0]


def kernel_stats(db_path: str):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, duration from kernels").fetchall()
    agg = defaultdict(list)
    for name, dur in rows:
        agg[_short(name)].append(dur)
    total = sum(sum(v) for v in agg.values())
    out = []
    for name, durs in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        s = sum(durs)
        out.append({"kernel": name, "calls": len(durs), "total_ms": s / 1e6, "avg_us": s / len(durs) / 1e3,
                    "min_us": min(durs) / 1e3, "max_us": max(durs) / 1e3, "pct": 100.0 * s / total if total else 0})
    return out


def pmc_per_kernel(db_path: str, kernel_filter: str | None = None):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, counter_name, counter_value, duration from pmc_events").fetchall()
    agg = defaultdict(lambda: defaultdict(list))
    for name, cname, val, dur in rows:
        k = _short(name)
        if kernel_filter and kernel_filter not in k:
            continue
        agg[k][cname].append(val)
    out = []
    for k, counters in agg.items():
        for cname, vals in counters.items():
            out.append({"kernel": k, "counter": cname, "dispatches": len(vals), "avg": sum(vals) / len(vals),
                        "total": sum(vals)})
    return sorted(out, key=lambda r: (r["kernel"], r["counter"]))


def main(argv):
    if len(argv) < 3:
        print(__doc__)
        return 2
    if argv[1] == "stats":
        print(f"{'kernel':<40} {'calls':>6} {'total_ms':>10} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'%':>6}")
        for r in kernel_stats(argv[2]):
            print(f"{r['kernel'][:40]:<40} {r['calls']:>6} {r['total_ms']:>10.3f} {r['avg_us']:>10.1f} "
                  f"{r['min_us']:>10.1f} {r['max_us']:>10.1f} {r['pct']:>6.1f}")
    elif argv[1] == "pmc":
        flt = argv[3] if len(argv) > 3 else None
        print(f"{'kernel':<40} {'counter':<28} {'dispatches':>10} {'avg/dispatch':>16} {'total':>16}")
        for r in pmc_per_kernel(argv[2], flt):
            print(f"{r['kernel'][:40]:<40} {r['counter']:<28} {r['dispatches']:>10} {r['avg']:>16.1f} {r['total']:>16.1f}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
