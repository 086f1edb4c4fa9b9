"""Write the host-layer golden fixtures from the reference's own code.

Runs oracle/_ref/ref_harness (the reference's meep-free sources compiled in place by
oracle/build_ref.py) on the reference's test inputs and junction files and stores the results
as JSON next to this script:

  scene_<name>.json   parsed settings, resolved CSG trees and metadata of every root, sources,
                      monitor clusters, numeric context entries, and in() of every root at seeded
                      points (src/cgs.cpp:34-118,422-447)
  conf_cases.json     settings after parse_args/parse_conf_file/correct_defaults
                      (src/argparse.h:120-383) for the reference's config test cases
  fft_cases.json      fft() outputs (src/data_utils.cpp:413-427) for seeded series

The junction files are parsed with the --opts of the reference's production sweep
(scripts/run.sh:54-78: width=0.05, thick=0.2, inf_thick=0, wavelen=0.76, n_cycles=0.5 and
--grid-res 12).  Only data is stored; no reference source text.

    python tests/golden/make_golden.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import build_ref  # noqa: E402

RUN_OPTS = "width=0.05;thick=0.2;inf_thick=0;wavelen=0.76;n_cycles=0.5"

SCENES = [
    # name, cwd (relative to the reference), conf, extra args
    ("test", ".", "tests/test.conf", ["--lo", "0.25", "--hi", "0.95", "--npts", "3000"]),
    # the reference's small end-to-end case sets len 2, pml 1, resolution 5 after loading run.conf
    ("run", ".", "tests/run.conf", ["--npts", "1000", "--post", "len=2,pml_thickness=1,resolution=5"]),
    ("span", ".", "tests/span.conf", ["--npts", "500"]),
    ("box", "junctions/Au_SiO2_box", "params.conf", ["--opts", RUN_OPTS, "--grid-res", "12", "--npts", "2500"]),
    ("bowtie", "junctions/Au_SiO2_bowtie", "params.conf", ["--opts", RUN_OPTS, "--grid-res", "12", "--npts", "2500"]),
    ("sidetie", "junctions/Au_SiO2_sidetie", "params.conf", ["--opts", RUN_OPTS, "--grid-res", "12", "--npts", "2500"]),
    ("bowtie_inf", "junctions/Au_SiO2_bowtie", "params.conf",
     ["--opts", RUN_OPTS.replace("inf_thick=0", "inf_thick=1"), "--grid-res", "12", "--npts", "3000"]),
    # a script of this repository (geom/quirks.geom) that exercises the evaluation rules of the
    # scene language; "@" marks a directory next to this file
    ("quirks", "@geom", "quirks.conf", ["--opts", "extra=4;twice=extra*2", "--npts", "800"]),
    ("small_run", "@geom", "small_run.conf", ["--npts", "500"]),
]

CONF_CASES = [
    ("test_conf", ".", "tests/test.conf", []),
    ("test_conf_cli", ".", "tests/test.conf",
     ["--geom-file", "blah.geom", "--out-dir", "/blah", "--grid-res", "3.0", "--length", "9.0", "--eps1", "2.0",
      "--opts", "a = 0.1; b_option = [1,\"blah\"]"]),
    ("run_conf", ".", "tests/run.conf", []),
    ("span_conf", ".", "tests/span.conf", []),
    ("sweep_conf", "junctions/Au_SiO2_bowtie", "params.conf", ["--grid-res", "12", "--save-span", "10"]),
    ("grid_num", "junctions/Au_SiO2_box", "params.conf", ["--grid-num", "64", "--length", "3.5"]),
]


def run(harness, cwd, args):
    r = subprocess.run([harness, *args], cwd=cwd, capture_output=True, text=True)
    # 4: the reference parser reported an error for the file (recorded in parse_error)
    if r.returncode not in (0, 4):
        raise RuntimeError(f"harness failed ({r.returncode}): {args}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--reference", default="/root/reference")
    p.add_argument("--only", default=None, help="write only the scene fixture of this name")
    a = p.parse_args()
    harness = build_ref.build(a.reference)
    if harness is None:
        raise SystemExit("reference tree not found")
    with tempfile.TemporaryDirectory() as tmp:
        for name, cwd, conf, extra in SCENES:
            if a.only and name != a.only:
                continue
            out = os.path.join(tmp, f"{name}.json")
            where = os.path.join(HERE, cwd[1:]) if cwd.startswith("@") else os.path.join(a.reference, cwd)
            run(harness, where, ["scene", conf, out, *extra])
            d = json.load(open(out))
            d["case"] = {"conf": conf, "args": extra, "dir": cwd}
            with open(os.path.join(HERE, f"scene_{name}.json"), "w") as f:
                json.dump(d, f, separators=(",", ":"))
        if a.only:
            print("wrote", a.only)
            return
        confs = {}
        for name, cwd, conf, extra in CONF_CASES:
            out = os.path.join(tmp, f"{name}.json")
            run(harness, os.path.join(a.reference, cwd), ["conf", conf, out, *extra])
            d = json.load(open(out))
            d["case"] = {"conf": conf, "args": extra, "dir": cwd}
            confs[name] = d
        with open(os.path.join(HERE, "conf_cases.json"), "w") as f:
            json.dump(confs, f, indent=1)
        rng = np.random.default_rng(10)
        ffts = []
        series = [np.sin(2 * np.pi * 2 * np.arange(16) / 16) + 0j]
        for n in (8, 16, 17, 31, 64, 100, 349):
            series.append(rng.standard_normal(n) + 1j * rng.standard_normal(n))
        t = np.arange(300) * 0.1
        series.append(np.exp(-((t - 15) ** 2) / 8) * np.exp(-2j * t))
        for s in series:
            inp = os.path.join(tmp, "series.txt")
            np.savetxt(inp, np.column_stack([s.real, s.imag]), fmt="%.17g")
            out = os.path.join(tmp, "fft.json")
            run(harness, tmp, ["fft", inp, out])
            d = json.load(open(out))
            ffts.append({"input": [[float(x.real), float(x.imag)] for x in s], "output": d["out"]})
        with open(os.path.join(HERE, "fft_cases.json"), "w") as f:
            json.dump(ffts, f, separators=(",", ":"))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
