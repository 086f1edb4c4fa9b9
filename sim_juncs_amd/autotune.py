"""Tile-shape autotuner for the fused step kernel.

The fused kernel's launch shape has three free parameters: its form (register rows: each
thread updates 1/2/4 y rows; thread rows: a workgroup of rows + 1 thread rows exchanges y and z
neighbours through LDS), the y rows, and the x-chunk each workgroup marches over.  `tune()`
times candidates on a live engine with HIP events (the simulation keeps advancing; timing does
not depend on the field values) and returns the fastest.  Results are cached per (shape, precision, material
mix) in a JSON file next to this module so later runs reuse them.

    python -m sim_juncs_amd.autotune --workload bowtie
"""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuned_tiles.json")

REGISTER_ROWS = (1, 2)
XCHUNKS = (4, 6, 8, 10, 12, 16, 20, 24, 32, 48, 64)


def candidates(shape, xchunks=XCHUNKS):
    """(form, rows, xchunk, zseg) tuples worth timing for a grid shape; zseg 0 is the engine's
    default z passes, other values a z pass length of one-row register tiles (lines of 512 cells
    or more: 128- and 256-cell passes and one pass over the line)."""
    nx, ny, nz = shape
    out = []
    for rows in REGISTER_ROWS:
        if rows == 1 or rows <= ny:
            out += [("register_rows", rows, xc, 0) for xc in xchunks if xc <= nx or xc == xchunks[0]]
    if 512 <= nz <= 1024:
        for zseg in (128, 256, -(-nz // 64) * 64):
            out += [("register_rows", 1, xc, zseg) for xc in xchunks if xc <= nx or xc == xchunks[0]]
    max_rows = 1024 // nz - 1
    for rows in sorted({r for r in (1, 2, 3, 4, 5, 7, 11, 15) if r <= max_rows} | ({max_rows} if max_rows >= 1 else set())):
        if rows > ny and rows != 1:
            continue
        out += [("thread_rows", rows, xc, 0) for xc in xchunks if xc <= nx or xc == xchunks[0]]
    return out


def key(shape, precision, tag="") -> str:
    return f"{shape[0]}x{shape[1]}x{shape[2]}:{precision}:{tag}"


def load_cache() -> dict:
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {}


def lookup(shape, precision, tag=""):
    return load_cache().get(key(shape, precision, tag))


def apply_cached(eng, tag="") -> bool:
    t = lookup(eng.shape, eng.precision, tag)
    if t and "form" in t:
        eng.set_tile(t["form"], t["rows"], t["xchunk"])
        if "zseg" in t:
            eng.set_zsegment(t["zseg"])
        return True
    return False


def tune(eng, steps: int = 20, xchunks=XCHUNKS, verbose: bool = True):
    """Time every candidate tile on `eng`; returns (best dict, all results)."""
    results = []
    eng.step(2)
    for form, rows, xc, zseg in candidates(eng.shape, xchunks):
        try:
            eng.set_tile(form, rows, xc)
            eng.set_zsegment(zseg)
            eng.timed_steps(2)
            _, ms = eng.timed_steps(steps)
        except Exception as exc:  # tile rejected (e.g. LDS budget)
            if verbose:
                print(f"[tune] {form} rows={rows} xc={xc} zseg={zseg}: {exc}", file=sys.stderr, flush=True)
            continue
        results.append(dict(form=form, rows=rows, xchunk=xc, zseg=zseg, ms=ms))
        if verbose:
            print(f"[tune] {form} rows={rows} xc={xc} zseg={zseg}: {ms:.4f} ms", file=sys.stderr, flush=True)
    best = min(results, key=lambda r: r["ms"])
    eng.set_tile(best["form"], best["rows"], best["xchunk"])
    eng.set_zsegment(best["zseg"])
    return best, results


def store(shape, precision, best, tag=""):
    c = load_cache()
    c[key(shape, precision, tag)] = dict(form=best["form"], rows=best["rows"], xchunk=best["xchunk"],
                                         ms=best["ms"])
    if best.get("zseg"):
        c[key(shape, precision, tag)]["zseg"] = best["zseg"]
    with open(CACHE, "w") as f:
        json.dump(c, f, indent=1, sort_keys=True)


def main():
    sys.path.insert(0, os.path.dirname(HERE))
    from sim_juncs_amd import scenario
    from sim_juncs_amd.engine import Engine

    p = argparse.ArgumentParser()
    p.add_argument("--workload", default="bowtie", choices=["bowtie", "plain", "adjoint", "large", "variant"])
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--store", action="store_true")
    a = p.parse_args()
    scn = {"bowtie": scenario.bowtie_junction, "plain": scenario.plain_box,
           "adjoint": scenario.adjoint_junction, "large": scenario.large_junction,
           "variant": lambda: scenario.adjoint_variants()[0]}[a.workload]()
    eng = scn.make(Engine)
    best, res = tune(eng, steps=a.steps)
    bytes_step = eng.step_bytes()
    for r in sorted(res, key=lambda r: r["ms"])[:8]:
        print(json.dumps(dict(r, gbs=bytes_step / (r["ms"] * 1e-3) / 1e9)))
    if a.store:
        store(eng.shape, eng.precision, best)
    eng.close()


if __name__ == "__main__":
    main()
