"""The host layer and the CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/host/sanitize_check.cpp drives config parsing, the .geom loader (including a script with a
syntax error), voxelisation with and without smoothing, the fft, the field_samples.h5 writer
and a small oracle run with every feature; it is built here with g++/gcc and -fsanitize=
address,undefined (no recovery), so any out-of-bounds access, use after free, leak or undefined
behaviour aborts it.  CPU only: the host layer's engine calls (BoundGeom) are not linked.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HOST = os.path.join(ROOT, "sim_juncs_amd", "host")
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_host_layer_and_oracle_are_clean_under_asan_and_ubsan(tmp_path):
    obj = str(tmp_path / "orc.o")
    exe = str(tmp_path / "sanitize_check")
    subprocess.run(["gcc", "-std=gnu11", *SAN, "-c", os.path.join(ROOT, "oracle", "fdtd_oracle.c"), "-o", obj],
                   check=True, capture_output=True, text=True)
    srcs = [os.path.join(HOST, f) for f in ("csg.cpp", "geom.cpp", "settings.cpp", "materials.cpp", "fft.cpp",
                                            "h5out.cpp")]
    subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"), "-I", HOST,
                    os.path.join(HERE, "host", "sanitize_check.cpp"), *srcs, obj, "-o", exe, "-ldl", "-lm"],
                   check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, os.path.join(HERE, "golden"), str(tmp_path)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "all parts ran" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
