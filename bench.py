"""Throughput of the fused FDTD step on one MI355X.

Workload (default): the single large junction instance -- 512x512x256 cells of 5 nm, 12-cell
PML, Au bowtie (Drude + 2 Lorentz) on a SiO2 Lorentz substrate with a graphene Drude sheet,
FP32 fields, FP64 DFT accumulators at 4 points x 8 frequencies, Gaussian 800 nm pulse.
A "step" is one FDTD time step over the whole grid (fused update kernel + source + DFT +
monitor sampling every 20 steps).  Fields are resident in HBM before timing starts.

Prints one JSON line (metric: Yee cell-updates/s, cells including PML x steps / wall time),
with the dominant kernel's HBM roofline and the FP64 CPU oracle timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=6000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--workload", default="bowtie", choices=["bowtie", "plain"])
    p.add_argument("--cpu-steps", type=int, default=6)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--kernel-steps", type=int, default=200,
                   help="steps timed with per-launch HIP events for the kernel duration")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are out of scope "
                                   "(instances shard across GPUs without exchange, see DESIGN)"}))
        sys.exit(2)
    import numpy as np  # noqa: F401

    from sim_juncs_amd import scenario
    from sim_juncs_amd.engine import Engine

    t0 = time.time()
    if a.workload == "bowtie":
        scn = scenario.bowtie_junction(steps=a.steps)
    else:
        scn = scenario.plain_box(steps=a.steps, precision="fp64")
    print(f"[bench] scenario {scn.name} built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    eng = scn.make(Engine)
    print(f"[bench] engine ready {eng.launch_info()}", file=sys.stderr, flush=True)
    eng.step(a.warmup)
    eng.synchronize()
    t1 = time.perf_counter()
    eng.step(a.steps)
    eng.synchronize()
    wall = time.perf_counter() - t1
    cells = scn.cells
    value = cells * a.steps / wall
    print(f"[bench] timed {a.steps} steps in {wall:.3f}s", file=sys.stderr, flush=True)

    ms_tot, ms_kernel = eng.timed_steps(a.kernel_steps)
    bytes_step = eng.step_bytes()
    achieved = bytes_step / (ms_kernel * 1e-3) / 1e9
    info = eng.launch_info()
    eng.close()

    cpu = None
    if not a.no_cpu:
        from oracle import oracle as orc

        t2 = time.time()
        o = scn.make(orc.Oracle)
        print(f"[bench] oracle set up in {time.time() - t2:.1f}s", file=sys.stderr, flush=True)
        o.step(1)
        t3 = time.perf_counter()
        o.step(a.cpu_steps)
        cpu_s = time.perf_counter() - t3
        o.close()
        cpu = {"value": cells * a.cpu_steps / cpu_s, "unit": "cell-updates/s",
               "cores": orc.num_threads(), "kind": "port",
               "sample": f"FP64 oracle, same {scn.shape} grid and materials, {a.cpu_steps} steps "
                         f"after 1 warm-up step ({cpu_s:.1f}s)"}

    out = {
        "metric": "Yee cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if scn.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": scn.name, "cells": cells, "precision": scn.precision,
                   "dft": "fp64 compensated", "tile": info},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_step", "kernel_ms": ms_kernel,
                     "algorithmic_bytes_per_launch": bytes_step},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
