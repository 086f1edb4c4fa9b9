"""The C++ driver (BoundGeom over the engine) end to end on the GPU.

* The reference's small end-to-end case (src/main_test.cpp:1884-2028: tests/run.conf with
  length 2, PML 1, resolution 5.25 here on a 21^3 grid): field_samples.h5 layout, metadata and
  monitor series.  Its Ey sheet leaves the monitored Ex at round-off level, so a variant with an
  Ex sheet and a Drude box checks the series against the FP64 oracle driven with the same
  voxelised materials, source and monitors (real and imaginary parts as two real runs).
* The Au/SiO2 bowtie junction scene through the sjf_geom command-line binary: 40 monitor
  clusters of 40 points, Drude Au and Lorentz SiO2 from the voxeliser, finite series.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle
from sim_juncs_amd import host
from sim_juncs_amd.materials import MaterialSpec, Pole

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def lut_specs(lut):
    return [MaterialSpec(m["eps_inf"], tuple(Pole(m["f0"][q], m["gamma"][q], m["sigma"][q], bool(m["is_drude"][q]))
                                             for q in range(m["npoles"]))) for m in lut]


def test_driver_small_case_layout(tmp_path):
    conf = os.path.join(G, "conf", "run.conf")
    scene = os.path.join(G, "scene_run.json")
    argv = ["--conf-file", conf, "--length", "2", "--grid-num", "21", "--out-dir", str(tmp_path)]
    print("[case] driver run.conf 21^3 fp64 complex", flush=True)
    s = host.run(conf, scene, argv, precision=8, complex_fields=True)
    assert s["grid_num"] == 21 and s["n_monitors"] == 2 and s["n_clusters"] == 1
    h5 = str(tmp_path / "field_samples.h5")
    assert set(host.h5_list(h5, "/")) == {"info", "cluster_0", "cluster_1"}
    tb = host.h5_read(h5, "info/time_bounds")
    assert tb[0] == 0.0 and tb[1] > 0 and 0 < tb[2] < tb[1]
    assert host.h5_read(h5, "info/n_clusters")[0] == 1
    for k, v in (("l_per_um", 2.0), ("sim_length", 2.0), ("pml_thickness", 1.0), ("tot_len", 4.0)):
        assert host.h5_read(h5, f"info/cgs_params/{k}")[0] == v
    locs = host.h5_read(h5, "cluster_0/locations", ["x", "y", "z"])
    assert locs.tolist() == [[1.0, 1.0, 1.0], [2.0, 4.0, 1.1]]
    n_t = s["n_t_pts"]  # save_span 1
    series = [host.h5_read(h5, f"cluster_0/point_{p}/time", ["Re", "Im"]) for p in range(2)]
    assert all(t.shape == (n_t, 2) for t in series)
    eps = host.h5_read(str(tmp_path / "eps-000000.00.h5"), "eps")
    assert eps.size == 21 ** 3 and np.all(eps == 1.0)
    assert all(np.all(np.isfinite(t)) for t in series)


def test_driver_smoothed_materials_match_oracle(tmp_path):
    """smooth_n > 0 with two roots (a Drude box and a dielectric slab): the smoothed, fractional
    permittivity reaches the engines as per-cell maps (a LUT could not hold it), and the driver's
    monitor series match the FP64 oracle run on the same voxelised maps and materials."""
    with open(os.path.join(G, "scene_run.json")) as f:
        d = json.load(f)
    d["sources"][0]["component"] = 0
    ident = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    box = {"kind": "box", "invert": 0, "trans": ident, "center": [0.95, 2.0, 2.0], "offset": [0.95, 2.0, 2.0]}
    slab = {"kind": "box", "invert": 0, "trans": ident, "center": [2.0, 2.0, 2.6], "offset": [2.0, 2.0, 0.45]}
    d["roots"][0]["tree"] = {"kind": "composite", "invert": 0, "trans": ident, "combine": "union",
                             "children": [box, None]}
    d["roots"][0]["metadata"]["susceptibilities"] = [[1.0, 0.48, 68.5971845, "drude"]]
    d["roots"].append({"metadata": {"eps": 2.5, "color": 7},
                       "tree": {"kind": "composite", "invert": 0, "trans": ident, "combine": "union",
                                "children": [slab, None]}})
    scene = str(tmp_path / "scene.json")
    with open(scene, "w") as f:
        json.dump(d, f)
    with open(os.path.join(G, "conf", "run.conf")) as f:
        text = f.read()
    conf = str(tmp_path / "smooth.conf")
    with open(conf, "w") as f:
        f.write(text.replace("ambient_eps = 1.0", "ambient_eps = 1.0\nsmooth_n = 2\nsmooth_rad = 0.15"))
    out = tmp_path / "out"
    out.mkdir()
    argv = ["--conf-file", conf, "--length", "2", "--grid-num", "21", "--out-dir", str(out)]
    print("[case] driver smoothed materials 21^3 fp64 complex", flush=True)
    s = host.run(conf, scene, argv, precision=8, complex_fields=True)
    h5 = str(out / "field_samples.h5")
    n_t = s["n_t_pts"]
    locs = host.h5_read(h5, "cluster_0/locations", ["x", "y", "z"])
    series = [host.h5_read(h5, f"cluster_0/point_{p}/time", ["Re", "Im"]) for p in range(2)]
    L = 4.0
    dx = L / 21
    words, lut, emap = host.voxelize(scene, 21, dx, ambient_eps=1.0, smooth_n=2, smooth_rad=0.15, um_scale=2.0,
                                     resolution=21 / L, with_map=True)
    assert emap is not None and len(np.unique(emap)) > 4
    info, p = host.source_info(scene, 0, um_scale=2.0)
    ref = []
    for part in (0, 1):
        o = Oracle((21, 21, 21), dx, precision="fp64", part=part)
        o.set_materials(words, lut_specs(lut))
        o.set_eps_maps(emap)
        o.set_pml([5] * 6, 1e-15, 2.0)
        o.add_source(info["component"], [0, 0, 1, 4, 4, 1], kind="gaussian", freq=p["freq"], width=p["width"],
                     phase=p["phase"], t_start=p["t_start"], t_end=p["t_end"], amp=info["amplitude"])
        o.add_monitors("Ex", locs)
        o.set_save_span(1)
        o.step(n_t)
        ref.append(o.read_monitors().real[:n_t])
        o.close()
    scale = max(np.abs(ref[0]).max(), np.abs(ref[1]).max())
    assert scale > 1e-3
    for pt in range(2):
        for part in (0, 1):
            assert np.abs(series[pt][:, part] - ref[part][:, pt]).max() <= 1e-10 * scale, (pt, part)
    eps_dump = host.h5_read(str(out / "eps-000000.00.h5"), "eps")
    assert np.allclose(eps_dump.reshape(21, 21, 21), emap.mean(axis=0), rtol=0, atol=1e-12)


def test_driver_matches_oracle_with_materials(tmp_path):
    """The small case with an Ex sheet and a Drude box filling x < 1.9: the driver's monitor
    series (real and imaginary parts from two real runs) match the FP64 oracle driven with the
    same voxelised materials, source and monitors."""
    with open(os.path.join(G, "scene_run.json")) as f:
        d = json.load(f)
    d["sources"][0]["component"] = 0
    ident = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    box = {"kind": "box", "invert": 0, "trans": ident, "center": [0.95, 2.0, 2.0], "offset": [0.95, 2.0, 2.0]}
    d["roots"][0]["tree"] = {"kind": "composite", "invert": 0, "trans": ident, "combine": "union",
                             "children": [box, None]}
    d["roots"][0]["metadata"]["susceptibilities"] = [[1.0, 0.48, 68.5971845, "drude"]]
    scene = str(tmp_path / "scene.json")
    with open(scene, "w") as f:
        json.dump(d, f)
    conf = os.path.join(G, "conf", "run.conf")
    out = tmp_path / "out"
    out.mkdir()
    argv = ["--conf-file", conf, "--length", "2", "--grid-num", "21", "--out-dir", str(out)]
    print("[case] driver run.conf + Drude box 21^3 fp64 complex", flush=True)
    s = host.run(conf, scene, argv, precision=8, complex_fields=True)
    h5 = str(out / "field_samples.h5")
    n_t = s["n_t_pts"]
    locs = host.h5_read(h5, "cluster_0/locations", ["x", "y", "z"])
    series = [host.h5_read(h5, f"cluster_0/point_{p}/time", ["Re", "Im"]) for p in range(2)]

    L = 4.0
    dx = L / 21
    words, lut = host.voxelize(scene, 21, dx, ambient_eps=1.0, um_scale=2.0, resolution=21 / L)
    assert {m["npoles"] for m in lut} == {0, 1} and {m["eps_inf"] for m in lut} == {1.0, 3.5}
    info, p = host.source_info(scene, 0, um_scale=2.0)
    ref = []
    for part in (0, 1):
        o = Oracle((21, 21, 21), dx, precision="fp64", part=part)
        o.set_materials(words, lut_specs(lut))
        o.set_pml([5] * 6, 1e-15, 2.0)
        o.add_source(info["component"], [0, 0, 1, 4, 4, 1], kind="gaussian", freq=p["freq"], width=p["width"],
                     phase=p["phase"], t_start=p["t_start"], t_end=p["t_end"], amp=info["amplitude"])
        o.add_monitors("Ex", locs)
        o.set_save_span(1)
        assert abs(o.dt - s["dt"]) < 1e-15
        o.step(n_t)
        ref.append(o.read_monitors().real[:n_t])
        o.close()
    scale = max(np.abs(ref[0]).max(), np.abs(ref[1]).max())
    assert scale > 1e-3
    for pt in range(2):
        for part in (0, 1):
            assert np.abs(series[pt][:, part] - ref[part][:, pt]).max() <= 1e-10 * scale, (pt, part)


def test_driver_runs_geom_script_like_its_scene_json(tmp_path):
    """The driver given the .geom script itself (run with the settings of the conf file) writes
    the same field_samples.h5 series as given the scene the reference parser made of it."""
    conf = os.path.join(G, "geom", "small_run.conf")
    out = {}
    for kind, scene in (("geom", os.path.join(G, "geom", "small_run.geom")),
                        ("json", os.path.join(G, "scene_small_run.json"))):
        d = tmp_path / kind
        d.mkdir()
        argv = ["--conf-file", conf, "--grid-num", "17", "--out-dir", str(d)]
        print(f"[case] driver small_run.{kind} 17^3 fp32 complex", flush=True)
        s = host.run(conf, scene, argv, precision=4, complex_fields=True)
        assert s["grid_num"] == 17 and s["n_monitors"] == 3 and s["n_clusters"] == 1
        h5 = str(d / "field_samples.h5")
        out[kind] = [host.h5_read(h5, f"cluster_0/point_{p}/time", ["Re", "Im"]) for p in range(3)]
    for a, b in zip(out["geom"], out["json"]):
        assert np.array_equal(a, b)
    assert max(np.abs(a).max() for a in out["geom"]) > 0


def test_geom_binary_bowtie_junction(tmp_path):
    exe = host.GEOM_BIN
    assert os.path.exists(exe), "sjf_geom is not built"
    conf = os.path.join(G, "conf", "bowtie_params.conf")
    scene = os.path.join(G, "scene_bowtie.json")
    print("[case] sjf_geom bowtie 45^3 fp32 real", flush=True)
    r = subprocess.run([exe, "--conf-file", conf, "--geom-file", scene, "--grid-num", "45", "--out-dir",
                        str(tmp_path), "--fp32", "--real"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "simulation completed" in r.stdout
    h5 = str(tmp_path / "field_samples.h5")
    names = host.h5_list(h5, "/")
    assert len(names) == 42 and "cluster_00" in names and "cluster_40" in names
    n_t = int(host.h5_read(h5, "info/n_time_points")[0])
    assert n_t > 20
    t = host.h5_read(h5, "cluster_05/point_0200/time", ["Re", "Im"])
    assert t.shape == (n_t, 2) and np.all(np.isfinite(t)) and np.abs(t[:, 0]).max() > 0
    assert np.all(t[:, 1] == 0)  # real run: no imaginary part
