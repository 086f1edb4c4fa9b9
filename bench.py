"""Throughput of the FDTD time step on one MI355X.

Default workload (`large`): the largest single-GPU configuration, one 1024x768x512 instance of
the bowtie junction -- cells of 5 nm, 12-cell PML, Au bowtie (Drude + 2 Lorentz) on a SiO2
Lorentz substrate with a graphene Drude sheet, FP32 fields, FP64 compensated DFT accumulators
at 4 points x 8 frequencies, 800 nm Gaussian pulse, 10000 steps.  A "step" is one time step
over the whole grid (fused update kernel + source + DFT + monitor sampling every 20 steps).
Fields are resident in HBM before timing.

Other workloads (--workload): bowtie (the same junction at 512x512x256, 6000 steps), plain
(96^3 FP64), sweep (64 uneven instances batched), adjoint (256x256x128 with a 64x64x16 design
region: forward + discrete adjoint), variants (16 adjoint design variants as one batch: one
fused forward and one fused reverse launch per step over all of them), variants-sequential (the
same 16 variants, one adjoint run after another).

Prints one JSON line: metric Yee cell-updates/s (cells including PML x steps / wall time),
the dominant kernel's HBM roofline, and the FP64 CPU oracle timed on this host's cores.
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
    p.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's)")
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--workload", default="large",
                   choices=["bowtie", "plain", "sweep", "adjoint", "large", "variants", "variants-sequential"])
    p.add_argument("--cpu-steps", type=int, default=None,
                   help="oracle steps timed for the CPU baseline (default: about 20 s of work)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--kernel-steps", type=int, default=100,
                   help="steps timed with per-launch HIP events for the kernel duration")
    p.add_argument("--tile", default=None,
                   help="fused-kernel tile as form,rows,xchunk (default: tuned cache or engine default)")
    p.add_argument("--checkpoint-every", type=int, default=0,
                   help="adjoint workload: checkpoint every K steps and recompute (0: record the design "
                        "fields of every step, no recomputation; both give the same bits)")
    p.add_argument("--zseg", type=int, default=None,
                   help="register-rows tile: z cells per workgroup pass (default: tuned cache or one pass)")
    p.add_argument("--adjoint-tile", default=None,
                   help="adjoint workloads: tile of the fused reverse step as form,rows,xchunk "
                        "(default: the engine's)")
    p.add_argument("--batch-tile", default=None,
                   help="tile of a batch's fused forward launch as form,xchunk (sweep, variants)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_budget_steps(cells, budget=1.4e10):
    """Oracle steps of a bounded CPU sample: about 1.4e10 cell-updates (~20 s on 16 cores)."""
    return max(3, int(budget / cells))


def cpu_baseline(scns, steps, of_steps):
    """FP64 oracle on this host's cores over `steps` steps of the given instances (the workload
    runs `of_steps`; the rate is per step, so the sample is extrapolated by steps)."""
    from oracle import oracle as orc

    cells = 0
    secs = 0.0
    for s in scns:
        o = s.make(orc.Oracle)
        o.step(1)
        t = time.perf_counter()
        o.step(steps)
        secs += time.perf_counter() - t
        cells += s.cells
        o.close()
    if steps < 200:
        span = (f"{steps} of its {of_steps} steps each after 1 warm-up step ({secs:.1f} s; the sample "
                f"is bounded to ~20 s of CPU work, fewer than the 200 steps a longer run would time, and "
                f"the rate is extrapolated by steps)")
    else:
        span = (f"{steps} steps each after 1 warm-up step ({secs:.1f} s of CPU work; the workload runs "
                f"{of_steps}, and the rate is per step)")
    return {"value": cells * steps / secs, "unit": "cell-updates/s", "cores": orc.num_threads(),
            "kind": "port",
            "sample": f"FP64 oracle (build's CPU stepper, not meep) on {len(scns)} full-size instance(s) "
                      f"of the workload, {span}"}


def roofline(bytes_step, ms_kernel, kernel, threads, groups, workload):
    """Roofline of the dominant kernel: algorithmic bytes per launch over its mean duration
    (HIP events on the engine's stream).  `traffic` is the HBM bytes per launch measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (stored by `python -m sim_juncs_amd.profsum
    traffic ... --store`), reported only when it was measured on the kernel sources being run at
    the same launch shape."""
    from sim_juncs_amd import profsum
    from sim_juncs_amd.engine import copy_bandwidth

    achieved = bytes_step / (ms_kernel * 1e-3) / 1e9
    rec = profsum.stored_traffic(workload, kernel, threads, groups)
    copy = copy_bandwidth(1 << 30, 20)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": rec["traffic_bytes"] if rec else None,
            "kernel": kernel, "kernel_ms": ms_kernel, "algorithmic_bytes_per_launch": bytes_step,
            "threads": threads, "groups": groups,
            "measured_copy_gbs": copy, "frac_of_copy": achieved / copy}


def set_batch_tile(a, b):
    if a.batch_tile:
        form, xc = a.batch_tile.split(",")
        b.set_tile(form, int(xc))


def step_kernel(info, nz=None):
    """Name of the fused step kernel the launch shape runs (register rows with fewer threads
    than the z line has cells step it in passes: k_step_z)."""
    if info["form"] != "register_rows":
        return "k_step_rows"
    return "k_step_z" if nz is not None and info["threads"] < nz else "k_step"


def set_tile(a, eng):
    """--tile form,rows,xchunk, else the tuned tile of the shape when cached."""
    from sim_juncs_amd import autotune

    if a.tile:
        form, rows, xc = a.tile.split(",")
        eng.set_tile(form, int(rows), int(xc))
    else:
        autotune.apply_cached(eng)
    if a.zseg is not None:
        eng.set_zsegment(a.zseg)


def single(a, scn, tag):
    from sim_juncs_amd.engine import Engine

    steps = a.steps if a.steps is not None else scn.steps
    eng = scn.make(Engine)
    set_tile(a, eng)
    eng.reserve_steps(a.warmup + steps + a.kernel_steps)  # no allocation in the timed region
    log(f"engine ready {eng.launch_info()}")
    eng.step(a.warmup)
    eng.synchronize()
    t = time.perf_counter()
    eng.step(steps)
    eng.synchronize()
    wall = time.perf_counter() - t
    log(f"timed {steps} steps in {wall:.3f}s")
    _, ms_kernel = eng.timed_steps(a.kernel_steps)
    info = eng.launch_info()
    rl = roofline(eng.step_bytes(), ms_kernel, step_kernel(info, scn.shape[2]), info["threads"], info["groups"],
                  scn.name)
    eng.close()
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline([scn], a.cpu_steps or cpu_budget_steps(scn.cells), scn.steps)
    return dict(value=scn.cells * steps / wall, steps=steps, wall=wall,
                config={"workload": scn.name, "cells": scn.cells, "precision": scn.precision,
                        "dft": "fp64 compensated", "tile": info},
                roofline=rl, cpu=cpu, dtype="f32" if scn.precision == "fp32" else "f64")


def sweep(a):
    from sim_juncs_amd import scenario
    from sim_juncs_amd.engine import Batch, Engine

    scns = scenario.batched_sweep()
    steps = a.steps if a.steps is not None else scns[0].steps
    engs = [s.make(Engine) for s in scns]
    b = Batch(engs)
    set_batch_tile(a, b)
    cells, bytes_step = b.info()
    log(f"batch of {len(engs)} engines, {cells} cells")
    b.reserve_steps(a.warmup + steps + a.kernel_steps)
    b.step(a.warmup)
    b.synchronize()
    t = time.perf_counter()
    b.step(steps)
    b.synchronize()
    wall = time.perf_counter() - t
    _, ms_kernel = b.timed_steps(a.kernel_steps)
    threads, groups = b.launch_info()
    # the batch default for z lines under 96 cells is the row-of-threads form
    kernel = "k_step" if a.batch_tile and a.batch_tile.startswith("register_rows") else "k_step_rows"
    b.close()
    for e in engs:
        e.close()
    sample = scns[:4]
    cpu = None if a.no_cpu else cpu_baseline(sample, a.cpu_steps or cpu_budget_steps(sum(s.cells for s in sample)),
                                             scns[0].steps)
    return dict(value=cells * steps / wall, steps=steps, wall=wall,
                config={"workload": f"sweep_{len(scns)}_instances", "cells": cells, "precision": "fp32",
                        "launch": "one fused launch per step over all instances, hipGraph replay"},
                roofline=roofline(bytes_step, ms_kernel, kernel, threads, groups,
                                  f"sweep_{len(scns)}_instances"), cpu=cpu, dtype="f32")


def adjoint_batched(a):
    """The 16 design variants as one batch: per step one fused forward launch and one fused
    reverse launch over all of them (sjf_batch_adjoint_run; record mode unless
    --checkpoint-every K > 0)."""
    from sim_juncs_amd import scenario
    from sim_juncs_amd.engine import Batch, Engine

    scns = scenario.adjoint_variants()
    steps = a.steps if a.steps is not None else scns[0].steps
    engs = [s.make(Engine) for s in scns]
    for e in engs:
        set_tile(a, e)  # each engine's own tile: a --batch-tile x-chunk of 0 takes its x-chunk
    b = Batch(engs)
    set_batch_tile(a, b)
    cells, _ = b.info()
    bytes_rev = sum(e.step_bytes() for e in engs)
    t = time.perf_counter()
    J, grads = b.adjoint_run(steps, [s.meta["gap"] for s in scns], [s.design[2].size for s in scns],
                             a.checkpoint_every)
    wall = time.perf_counter() - t
    ms, threads, groups = b.adjoint_kernel_info()
    log(f"{len(scns)} variants: J {J.min():.4e}..{J.max():.4e}; reverse launch {ms * 1e3:.1f} us over "
        f"{groups} workgroups of {threads} threads")
    fthreads, fgroups = b.launch_info()
    b.close()
    for e in engs:
        e.close()
    s0 = scns[0]
    workload = f"variants_{len(scns)}_batched"
    rl = roofline(bytes_rev, ms, "k_adj_step_batch", threads, groups, workload)
    cpu = None if a.no_cpu else cpu_baseline([s0], a.cpu_steps or cpu_budget_steps(s0.cells), s0.steps)
    if cpu:
        cpu["value"] /= 2.0
        cpu["sample"] += "; divided by 2 for the forward and reverse sweeps"
    return dict(value=cells * steps / wall, steps=steps, wall=wall,
                config={"workload": workload, "cells": cells, "precision": "fp32",
                        "checkpoint_every": a.checkpoint_every,
                        "launch": f"per step one fused forward launch ({fgroups} workgroups of {fthreads}) and "
                                  f"one fused reverse launch over all {len(scns)} variants",
                        "note": "value counts forward steps per second of a full gradient computation"},
                roofline=rl, cpu=cpu, dtype="f32")


def adjoint(a, variants=False):
    from sim_juncs_amd import scenario
    from sim_juncs_amd.engine import Engine

    scns = scenario.adjoint_variants() if variants else [scenario.adjoint_junction()]
    steps = a.steps if a.steps is not None else scns[0].steps
    cells = 0
    wall = 0.0
    rev = None
    for s in scns:
        eng = s.make(Engine)
        set_tile(a, eng)  # the forward tile (--tile or the tuned cache); --adjoint-tile the reverse one
        if a.adjoint_tile:
            f, r, x = a.adjoint_tile.split(",")
            eng.set_adjoint_tile(f, int(r), int(x))
        nd = s.design[2].size
        t = time.perf_counter()
        J, g = eng.adjoint_run(steps, a.checkpoint_every, s.meta["gap"], nd)
        wall += time.perf_counter() - t
        cells += s.cells
        log(f"{s.name}: J={J:.4e} |grad|={float((g * g).sum()) ** 0.5:.4e}")
        if rev is None:
            # the fused reverse step is the dominant kernel; it moves the forward step's bytes
            rev = (eng.step_bytes(),) + eng.adjoint_kernel_info()
            log(f"reverse kernel {rev[1] * 1e3:.1f} us over {rev[3]} workgroups of {rev[2]} threads")
        eng.close()
    s0 = scns[0]
    workload = ("variants_" if variants else "") + s0.name
    rkernel = "k_adj_rows" if a.adjoint_tile and a.adjoint_tile.startswith("thread_rows") else "k_adj_step"
    rl = roofline(rev[0], rev[1], rkernel, rev[2], rev[3], workload)
    cpu = None if a.no_cpu else cpu_baseline([s0], a.cpu_steps or cpu_budget_steps(s0.cells), s0.steps)
    if cpu:
        # the adjoint costs a forward and a reverse sweep per step (the design fields of every
        # step are recorded, as a CPU implementation could record them too)
        cpu["value"] /= 2.0
        cpu["sample"] += "; divided by 2 for the forward and reverse sweeps"
    return dict(value=cells * steps / wall, steps=steps, wall=wall,
                config={"workload": workload, "cells": cells,
                        "precision": "fp32", "checkpoint_every": a.checkpoint_every,
                        "note": "value counts forward steps per second of a full gradient computation"},
                roofline=rl, cpu=cpu, dtype="f32")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are out of scope "
                                   "(instances shard across GPUs without exchange)"}))
        sys.exit(2)
    from sim_juncs_amd import scenario

    t0 = time.time()
    if a.workload == "bowtie":
        r = single(a, scenario.bowtie_junction(), "bowtie")
    elif a.workload == "plain":
        r = single(a, scenario.plain_box(), "plain")
    elif a.workload == "large":
        r = single(a, scenario.large_junction(), "large")
    elif a.workload == "sweep":
        r = sweep(a)
    elif a.workload == "adjoint":
        r = adjoint(a)
    elif a.workload == "variants-sequential":
        r = adjoint(a, variants=True)
    else:
        r = adjoint_batched(a)
    log(f"done in {time.time() - t0:.1f}s")
    out = {
        "metric": "Yee cell-updates/s",
        "value": r["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": a.warmup,
        "ms_per_step": r["wall"] / r["steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": r["dtype"],
        "data": "synthetic",
        "config": r["config"],
        "roofline": r["roofline"],
        "cpu_baseline": r["cpu"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
