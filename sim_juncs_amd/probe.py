"""Small measurement entry points run on the GPU.

    python -m sim_juncs_amd.probe copy [--bytes N] [--iters K]   # achievable HBM bandwidth
"""
from __future__ import annotations

import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    p = argparse.ArgumentParser()
    p.add_argument("what", choices=["copy"])
    p.add_argument("--bytes", type=int, default=1 << 30)
    p.add_argument("--iters", type=int, default=20)
    a = p.parse_args()
    from sim_juncs_amd.engine import copy_bandwidth

    g = copy_bandwidth(a.bytes, a.iters)
    print(json.dumps({"probe": "copy", "bytes": a.bytes, "iters": a.iters, "gbs_read_plus_write": g,
                      "bytes_per_launch": 2 * a.bytes}))


if __name__ == "__main__":
    main()
