"""Compile the engine (libsjf.so, hipcc for gfx950) and the host layer (libsjf_host.so and
the sjf_geom driver, g++) in-tree.

    python -m sim_juncs_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "sjf_engine.hip")
OUT = os.path.join(HERE, "libsjf.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code (the compiler still forms them
# from plain float code unless told not to).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def build_engine(verbose: bool = False) -> str:
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libsjf.so")
    os.replace(OUT + ".tmp", OUT)
    return OUT


HOST_SRC = [os.path.join(HERE, "host", f) for f in
            ("csg.cpp", "geom.cpp", "settings.cpp", "materials.cpp", "fft.cpp", "h5out.cpp", "bound_geom.cpp",
             "capi.cpp")]
HOST_OUT = os.path.join(HERE, "libsjf_host.so")
GEOM_OUT = os.path.join(HERE, "sjf_geom")
CXX_FLAGS = ["-O2", "-std=c++17", "-fopenmp", "-fPIC", "-Wall", "-Wno-unused-result"]


def build_host(verbose: bool = False) -> str:
    """libsjf_host.so (links libsjf.so, found next to it at run time) and the sjf_geom binary."""
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host")]
    link = ["-L", HERE, "-lsjf", "-ldl", "-Wl,-rpath,$ORIGIN"]
    steps = [
        ["g++", *CXX_FLAGS, "-shared", *inc, "-o", HOST_OUT + ".tmp", *HOST_SRC, *link],
        ["g++", *CXX_FLAGS, *inc, "-o", GEOM_OUT + ".tmp", os.path.join(HERE, "host", "sjf_geom.cpp"),
         "-L", HERE, "-lsjf_host", "-lsjf", "-ldl", "-Wl,-rpath,$ORIGIN"],
    ]
    for cmd, out in zip(steps, (HOST_OUT, GEOM_OUT)):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed building {os.path.basename(out)}")
        os.replace(out + ".tmp", out)
    return HOST_OUT


if __name__ == "__main__":
    print(build_engine(verbose=True))
    print(build_host(verbose=True))
