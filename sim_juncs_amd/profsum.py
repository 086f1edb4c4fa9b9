"""Summarise rocprofv3 result databases (rocpd SQLite) as short text tables.

    python -m sim_juncs_amd.profsum trace  <results.db>            # per-kernel duration stats
    python -m sim_juncs_amd.profsum pmc    <results.db> [kernel]   # per-kernel counter means

Counter values such as FETCH_SIZE / WRITE_SIZE are reported in KiB per dispatch, as
rocprofv3 records them.
"""
from __future__ import annotations

import sqlite3
import sys
from collections import defaultdict


def short(name: str) -> str:
    n = name.split("(")[0]
    return n[:110]


def trace(db: str):
    con = sqlite3.connect(db)
    rows = con.execute("select name, duration from kernels").fetchall()
    agg = defaultdict(list)
    for name, dur in rows:
        agg[name].append(dur)
    tot = sum(sum(v) for v in agg.values())
    print(f"| kernel | calls | total ms | mean us | min us | max us | % time |")
    print("|---|---|---|---|---|---|---|")
    for name, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{short(name)}` | {len(v)} | {sum(v) / 1e6:.3f} | {sum(v) / len(v) / 1e3:.2f} | "
              f"{min(v) / 1e3:.2f} | {max(v) / 1e3:.2f} | {100.0 * sum(v) / tot:.1f} |")


def pmc(db: str, kernel: str | None = None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, counter_name, counter_value, duration from pmc_events").fetchall()
    agg = defaultdict(list)
    for name, cn, cv, dur in rows:
        if kernel and kernel not in name:
            continue
        agg[(name, cn)].append(cv)
    print("| kernel | counter | dispatches | mean per dispatch | min | max |")
    print("|---|---|---|---|---|---|")
    for (name, cn), v in sorted(agg.items(), key=lambda kv: kv[0]):
        print(f"| `{short(name)}` | {cn} | {len(v)} | {sum(v) / len(v):.1f} | {min(v):.1f} | {max(v):.1f} |")


if __name__ == "__main__":
    mode, db = sys.argv[1], sys.argv[2]
    if mode == "trace":
        trace(db)
    else:
        pmc(db, sys.argv[3] if len(sys.argv) > 3 else None)
