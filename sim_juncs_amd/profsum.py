"""Summarise rocprofv3 result databases (rocpd SQLite) as short text tables.

    python -m sim_juncs_amd.profsum trace  <results.db>            # per-kernel duration stats
    python -m sim_juncs_amd.profsum pmc    <results.db> [kernel]   # per-kernel counter means
    python -m sim_juncs_amd.profsum traffic <workload> <kernel> <fetch.db> <write.db> \
        <copy_fetch.db> <copy_write.db> [--store]                  # HBM bytes per launch

Counter values such as FETCH_SIZE / WRITE_SIZE are reported in KiB per dispatch, as
rocprofv3 records them.
"""
from __future__ import annotations

import hashlib
import json
import os
import sqlite3
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
TRAFFIC = os.path.join(HERE, "measured_traffic.json")
COPY_BYTES = 1 << 30  # the calibration copy kernel reads and writes 1 GiB per launch


def kernel_source_hash() -> str:
    """Hash of the device code: a stored traffic figure is valid only for the kernels it was
    measured on (their launch shape is checked separately)."""
    h = hashlib.sha256()
    with open(os.path.join(HERE, "csrc", "sjf_kernels.h"), "rb") as fh:
        h.update(fh.read())
    return h.hexdigest()[:16]


def is_kernel(name: str, kernel: str) -> bool:
    """True for a dispatch of `kernel` itself (k_step matches sjf::k_step<...>, not k_step_rows)."""
    base = name.split("(")[0].split("<")[0].strip()
    return base == kernel or base.endswith("::" + kernel)


def counter_mean(db: str, kernel: str, counter: str) -> float:
    con = sqlite3.connect(db)
    v = [cv for name, cn, cv in con.execute("select name, counter_name, counter_value from pmc_events")
         if is_kernel(name, kernel) and cn == counter]
    if not v:
        raise SystemExit(f"no {counter} samples of {kernel} in {db}")
    return sum(v) / len(v)


def launch_shape(db: str, kernel: str):
    """(threads per workgroup, workgroups) of the dispatches of `kernel` (one shape expected)."""
    con = sqlite3.connect(db)
    shapes = {(wx * wy * wz, (gx // wx) * (gy // wy) * (gz // wz))
              for name, gx, gy, gz, wx, wy, wz in con.execute(
                  "select name, grid_x, grid_y, grid_z, workgroup_x, workgroup_y, workgroup_z from kernels")
              if is_kernel(name, kernel)}
    if len(shapes) != 1:
        raise SystemExit(f"{kernel} in {db}: expected one launch shape, found {sorted(shapes)}")
    return shapes.pop()


def traffic(workload, kernel, fetch_db, write_db, copy_fetch_db, copy_write_db, store=False):
    """HBM bytes per launch of `kernel`: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
    scaled by the ratio of known bytes to the same counter on the calibration copy kernel."""
    ff = COPY_BYTES / (1024.0 * counter_mean(copy_fetch_db, "k_copy", "FETCH_SIZE"))
    fw = COPY_BYTES / (1024.0 * counter_mean(copy_write_db, "k_copy", "WRITE_SIZE"))
    fk = counter_mean(fetch_db, kernel, "FETCH_SIZE")
    wk = counter_mean(write_db, kernel, "WRITE_SIZE")
    threads, groups = launch_shape(fetch_db, kernel)
    if launch_shape(write_db, kernel) != (threads, groups):
        raise SystemExit(f"{kernel}: the FETCH_SIZE and WRITE_SIZE runs launched different shapes")
    rec = dict(fetch_kib=fk, write_kib=wk, fetch_factor=ff, write_factor=fw,
               fetched_bytes=fk * 1024.0 * ff, written_bytes=wk * 1024.0 * fw,
               traffic_bytes=(fk * ff + wk * fw) * 1024.0, kernel_source=kernel_source_hash(),
               threads=threads, groups=groups)
    print(json.dumps(dict(workload=workload, kernel=kernel, **rec), indent=1))
    if store:
        c = {}
        if os.path.exists(TRAFFIC):
            with open(TRAFFIC) as f:
                c = json.load(f)
        c.setdefault(workload, {})[kernel] = rec
        with open(TRAFFIC, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)


def stored_traffic(workload: str, kernel: str, threads: int, groups: int):
    """Measured HBM bytes per launch of `kernel` in `workload`, or None when absent, measured on
    other kernel sources or at another launch shape."""
    if not os.path.exists(TRAFFIC):
        return None
    with open(TRAFFIC) as f:
        rec = json.load(f).get(workload, {}).get(kernel)
    if not rec or rec.get("kernel_source") != kernel_source_hash():
        return None
    if (rec.get("threads"), rec.get("groups")) != (threads, groups):
        return None
    return rec


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
    print(f"| kernel | calls | total ms | mean us | median us | min us | max us | % time |")
    print("|---|---|---|---|---|---|---|---|")
    for name, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        med = sorted(v)[len(v) // 2]
        print(f"| `{short(name)}` | {len(v)} | {sum(v) / 1e6:.3f} | {sum(v) / len(v) / 1e3:.2f} | {med / 1e3:.2f} | "
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
    mode = sys.argv[1]
    if mode == "trace":
        trace(sys.argv[2])
    elif mode == "traffic":
        args = [a for a in sys.argv[2:] if a != "--store"]
        traffic(*args, store="--store" in sys.argv)
    else:
        pmc(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else None)
