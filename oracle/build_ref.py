"""TEST INFRASTRUCTURE ONLY: build oracle/_ref/ref_harness from the reference's own sources.

Compiles, in place and without copying them, the reference's meep-free translation units
(src/cgs.cpp, src/cgs_read.cpp, src/cgs_data.cpp, src/data_utils.cpp; src/argparse.h is
header-only) together with oracle/ref_harness.cpp.  The binary goes to oracle/_ref/ (git-
ignored).  The reference's disp.cpp, main.cpp and main_test.cpp need meep (and doctest) and
are not buildable here.  Used by tests/golden/make_golden.py and the CPU tests that compare the
host layer with the reference; absent the reference tree (as on a machine that only runs the
GPU tests) nothing here is built.

    python oracle/build_ref.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
OUT = os.path.join(OUT_DIR, "ref_harness")
UNITS = ("cgs.cpp", "cgs_read.cpp", "cgs_data.cpp", "data_utils.cpp")


def build(reference: str = "/root/reference", force: bool = False) -> str | None:
    src = os.path.join(reference, "src")
    if not os.path.isdir(src):
        return None
    inputs = [os.path.join(src, u) for u in UNITS] + [os.path.join(HERE, "ref_harness.cpp")]
    if not force and os.path.exists(OUT) and all(os.path.getmtime(OUT) >= os.path.getmtime(p) for p in inputs):
        return OUT
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["g++", "-std=c++11", "-O2", "-w", "-I", src, "-o", OUT + ".tmp", *inputs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("building the reference harness failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    p = argparse.ArgumentParser()
    p.add_argument("--reference", default="/root/reference")
    a = p.parse_args()
    print(build(a.reference, force=True))
