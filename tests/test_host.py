"""Host layer against the reference's own code and tests (CPU only).

Golden data: tests/golden/*.json, written by tests/golden/make_golden.py from the reference's
meep-free sources compiled in place (oracle/build_ref.py).  Known answers quoted from the
reference's test suite cite src/main_test.cpp.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from sim_juncs_amd import host

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CONF = {"tests/test.conf": "test.conf", "tests/run.conf": "run.conf", "tests/span.conf": "span.conf"}
SCENES = ["test", "run", "span", "box", "bowtie", "bowtie_inf"]


def scene_path(name):
    return os.path.join(G, f"scene_{name}.json")


def load(name):
    with open(scene_path(name)) as f:
        return json.load(f)


def conf_path(case):
    c = case["conf"]
    if c in CONF:
        return os.path.join(G, "conf", CONF[c])
    return os.path.join(G, "conf", "bowtie_params.conf" if "bowtie" in case["dir"] else "box_params.conf")


# ---- settings (src/argparse.h) -------------------------------------------------------------

@pytest.mark.parametrize("name", ["test_conf", "test_conf_cli", "run_conf", "span_conf", "sweep_conf", "grid_num"])
def test_settings_match_reference(name):
    case = json.load(open(os.path.join(G, "conf_cases.json")))[name]
    path = conf_path(case["case"])
    mine = host.settings(path, ["--conf-file", path, *case["case"]["args"]])
    ref = case["settings"]
    for k, v in ref.items():
        if k == "conf_fname":
            continue
        assert mine[k] == v, (name, k, mine[k], v)


def test_settings_known_answers():
    # src/main_test.cpp:1648-1668 and 1672-1715
    t = os.path.join(G, "conf", "test.conf")
    s = host.settings(t, mode=0)
    assert (s["n_dims"], s["pml_thickness"], s["len"], s["um_scale"]) == (3, 2.0, 16.0, 1.0)
    assert s["resolution"] == 1.55 and s["courant"] == 0.3 and s["smooth_n"] == 1 and s["smooth_rad"] == 0.25
    assert (s["post_source_t"], s["save_span"], s["ambient_eps"]) == (1.0, 171, 1.0)
    assert s["geom_fname"] == "tests/test.geom" and s["out_dir"] == "/test_dir"
    s = host.settings(t, ["--conf-file", t, "--geom-file", "blah.geom", "--out-dir", "/blah", "--grid-res", "3.0",
                          "--length", "9.0", "--eps1", "2.0", "--opts", "a = 0.1; b_option = [1,\"blah\"]"],
                      mode=1)
    assert s["geom_fname"] == "blah.geom" and s["out_dir"] == "/blah" and s["len"] == 9.0
    assert abs(s["resolution"] - 3.1538) < 1e-4 and s["ambient_eps"] == 2.0 and s["pml_thickness"] == 2.0
    # odd cell count: grid_num = 2 int(0.5 (1 + L res)) + 1 and resolution = grid_num / L
    assert s["grid_num"] % 2 == 1 and s["resolution"] == s["grid_num"] / (9.0 + 4.0)


# ---- CSG (src/cgs.cpp) ----------------------------------------------------------------------

@pytest.mark.parametrize("name", SCENES)
def test_point_inclusion_matches_reference(name):
    d = load(name)
    pts = np.array(d["points"])
    mine = host.scene_in(scene_path(name), pts)
    assert mine.shape == (len(pts), len(d["roots"]))
    assert np.array_equal(mine, np.array(d["inside"]).reshape(mine.shape))


def test_point_inclusion_known_answers():
    # src/main_test.cpp:1550-1564
    pts = [(.45, .45, .45), (.45, .41, .6), (.65, .45, .41), (.71, .51, .51), (.55, .45, .45), (.55, .41, .85)]
    assert host.scene_in(scene_path("test"), pts)[:, 0].tolist() == [1, 1, 1, 0, 0, 0]


def _builders_binary(tmp_path):
    exe = str(tmp_path / "builders_check")
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    pkg = os.path.join(root, "sim_juncs_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(root, "include"),
                    os.path.join(here, "host", "builders_check.cpp"), "-o", exe, "-L", pkg, "-lsjf_host", "-lsjf",
                    f"-Wl,-rpath,{pkg}"], check=True)
    return exe


def test_builders_assemble_the_reference_trees(tmp_path):
    """Trees built with the list-assembly builders give the reference parser's in() at every
    golden point: Difference becomes Intersect with a Complement whose own inversion cancels,
    three-point planes use the (p2-p1)x(p3-p1) normal, boxes store centre and half-extent."""
    exe = _builders_binary(tmp_path)
    d = load("test")
    pts = "\n".join(f"{x!r} {y!r} {z!r}" for x, y, z in d["points"])
    out = subprocess.run([exe, "test"], input=pts, capture_output=True, text=True, check=True).stdout.split()
    assert np.array_equal(np.array(out, dtype=int), np.array(d["inside"]).reshape(-1))
    for name in ("bowtie", "bowtie_inf"):
        d = load(name)
        ctx = dict(d["context_numbers"])
        args = [repr(ctx[k]) for k in ("left", "rght", "mid", "top", "bot", "length")]
        pts = "\n".join(f"{x!r} {y!r} {z!r}" for x, y, z in d["points"])
        out = subprocess.run([exe, "bowtie", *args], input=pts, capture_output=True, text=True, check=True).stdout
        got = np.array(out.split(), dtype=int).reshape(-1, 2)
        assert np.array_equal(got, np.array(d["inside"])), name


def test_sidetie_is_rejected_by_the_reference_parser():
    # the reference's own junctions/Au_SiO2_sidetie/junc.geom calls an undefined um_l(): no roots
    d = load("sidetie")
    assert d["parse_error"] != 0 and d["roots"] == []


# ---- material function (src/disp.cpp:56-283) ---------------------------------------------

def test_in_bound_known_answers():
    # src/main_test.cpp:1612-1643: cgs_material_function(root) defaults: def_ret 1, smooth_n 1,
    # smooth_rad DEFAULT_SMOOTH_RAD = 0.05 (eight mirrored offsets)
    pts = [(.45, .45, .45), (.45, .41, .6), (.65, .45, .45), (.71, .51, .51), (.55, .45, .45), (.55, .41, .85)]
    v = host.in_bound(scene_path("test"), pts, def_ret=1.0, smooth_n=1, smooth_rad=0.05)
    assert v.tolist() == [3.5, 3.5, 3.5, 1.0, 1.0, 1.0]


def test_eps_sums_over_roots():
    # Q1: every root adds its background term, so two eps=1 roots on ambient 1 give 2 everywhere
    d = load("bowtie")
    v = host.in_bound(scene_path("bowtie"), np.array(d["points"]), def_ret=1.0)
    assert np.all(v == 2.0)
    # the formula at sample points: sum_i [amb + (eps_i - amb) in_i]
    d = load("test")
    pts = np.array(d["points"])
    inside = np.array(d["inside"])[:, 0]
    v = host.in_bound(scene_path("test"), pts, def_ret=1.5)
    assert np.array_equal(v, 1.5 + (3.5 - 1.5) * inside)


def test_smoothing_offsets():
    p = host.smooth_points(7, 0.25)
    assert p.shape == (56, 3)
    for i in range(7):
        q = p[8 * i: 8 * i + 8]
        # the eight sign reflections of one point, x and y both built from cos(phi)
        assert np.allclose(np.abs(q), np.abs(q[0]))
        assert np.allclose(np.abs(q[:, 0]), np.abs(q[:, 1]))
        assert sorted(map(tuple, np.sign(q).astype(int))) == sorted(
            (sx * int(np.sign(q[7, 0]) or 1), sy * int(np.sign(q[7, 1]) or 1), sz * int(np.sign(q[7, 2]) or 1))
            for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1))
    assert np.array_equal(p, host.smooth_points(7, 0.25))  # fixed seed
    assert host.smooth_points(0, 0.25).shape == (0, 3)


def test_smoothed_in_bound_is_the_sample_average():
    d = load("test")
    pts = np.array(d["points"])[:200]
    off = host.smooth_points(2, 0.03)
    v = host.in_bound(scene_path("test"), pts, def_ret=1.0, smooth_n=2, smooth_rad=0.03)
    samples = np.concatenate([pts[:, None, :], pts[:, None, :] + off[None, :, :]], axis=1)
    frac = host.scene_in(scene_path("test"), samples.reshape(-1, 3))[:, 0].reshape(len(pts), -1).mean(1)
    assert np.allclose(v, 1.0 + 2.5 * frac, rtol=0, atol=1e-15)


# ---- susceptibilities and sources (src/disp.cpp:318-376,418-454) -------------------------------

def test_susceptibilities_known_answers():
    # src/main_test.cpp:1757-1770
    poles, er = host.susceptibilities(scene_path("test"), 0)
    assert er == 0 and len(poles) == 2
    assert (poles[0]["omega_0"], poles[0]["gamma"], poles[0]["sigma"], poles[0]["use_denom"]) == (1.0, 0.48, 68.5971845, False)
    assert (poles[1]["omega_0"], poles[1]["gamma"], poles[1]["sigma"], poles[1]["use_denom"]) == (8.0, 0.816, 452848600800781300, False)


def test_au_lorentz_pole_is_dropped():
    # Q2: the Au Lorentz entry reaches the parser as a 2-element list; parsing stops there
    d = load("bowtie")
    au = [i for i, r in enumerate(d["roots"]) if r["metadata"].get("color") == 42][0]
    poles, er = host.susceptibilities(scene_path("bowtie"), au)
    assert er == -2 and len(poles) == 1 and not poles[0]["use_denom"]
    assert poles[0]["sigma"] == 4.0314052191361974e21
    poles, er = host.susceptibilities(scene_path("bowtie"), 1 - au)
    assert er == 0 and len(poles) == 1 and poles[0]["use_denom"] and poles[0]["omega_0"] == 9.67865314895427


def test_source_info_known_answers():
    # src/main_test.cpp:1775-1807 (cutoff forced to 6 widths: end = 5.2 + 12 * 3)
    g, p = host.source_info(scene_path("test"), 0, um_scale=1.0)
    assert g["gaussian"] and g["component"] == 1
    assert (g["wavelen"], g["amplitude"], g["width"], g["phase"]) == (1.333333, 7.0, 3.0, 0.75)
    assert abs(g["start_time"] - 5.2) < 1e-12 and abs(g["end_time"] - 41.2) < 1e-12
    c, q = host.source_info(scene_path("test"), 1, um_scale=1.0)
    assert not c["gaussian"] and c["component"] == 5
    assert (c["wavelen"], c["amplitude"], c["phase"], c["start_time"], c["end_time"], c["width"]) == (1.66, 8.0, 0.0, 0.2, 1.2, 0.1)
    # unit conversion of src/disp.cpp:604-609
    g, p = host.source_info(scene_path("run"), 0, um_scale=2.0)
    ca = 0.299792458 * 2.0
    assert p["kind"] == 0 and p["freq"] == 1 / (1.33 * 2.0)
    assert p["width"] == g["width"] * ca and p["t_start"] == g["start_time"] * ca and p["t_end"] == g["end_time"] * ca
    assert g["end_time"] == g["start_time"] + 12 * g["width"]


# ---- FFT (src/data_utils.cpp:325-427) -----------------------------------------------------------

def test_fft_matches_reference_bitwise():
    for c in json.load(open(os.path.join(G, "fft_cases.json"))):
        x = np.array(c["input"])
        ref = np.array(c["output"])
        got = host.fft(x[:, 0] + 1j * x[:, 1])
        assert np.array_equal(got, ref[:, 0] + 1j * ref[:, 1]), len(x)


def test_fft_closed_form():
    # sine KAT (src/main_test.cpp:17-40): 16 samples of sin(2 pi 2 n / 16) give -8i at bin 2
    n = np.arange(16)
    X = host.fft(np.sin(2 * np.pi * 2 * n / 16))
    assert abs(X[2] - (-8j)) < 1e-12 and abs(X[-2] - 8j) < 1e-12
    # Q10: the first 2^floor(log2 N) samples with phase step 2 pi / N
    rng = np.random.default_rng(5)
    for N in (5, 16, 23, 100):
        x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        T = 1 << int(np.log(N) / np.log(2))
        k = np.arange(-T // 2, T // 2)
        ref = (x[:T, None] * np.exp(-2j * np.pi * np.arange(T)[:, None] * k[None, :] / N)).sum(0)
        got = host.fft(x)
        assert got.size == T
        assert np.allclose(got[k % T], ref, atol=1e-12 * np.abs(x).sum())


# ---- voxelisation (src/disp.cpp:482-550) --------------------------------------------------------

def test_voxelize_with_smoothing_gives_eps_maps():
    """With smooth_n > 0 the permittivity is the fractional sum over roots (src/disp.cpp:56-112,
    264-283): voxelize returns it per cell and E component for sjf_set_eps_maps, equal at every
    Yee point to the smoothed material function, and a LUT of the pole combinations alone."""
    d = load("test")
    n, L = 17, 1.0
    dx = L / n
    words, lut, emap = host.voxelize(scene_path("test"), n, dx, ambient_eps=1.5, smooth_n=2, smooth_rad=0.03,
                                     um_scale=1.0, resolution=n / L, with_map=True)
    assert emap is not None and emap.shape == (3, n, n, n)
    assert all(m["eps_inf"] == 1.0 for m in lut) and len(lut) <= 255
    assert len(np.unique(emap)) > 3  # fractional values at the boundaries
    off = [(0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5)]
    rng = np.random.default_rng(5)
    cells = rng.integers(0, n, size=(300, 3))
    for c in range(3):
        pts = (cells + np.array(off[c])) * dx
        v = host.in_bound(scene_path("test"), pts, def_ret=1.5, smooth_n=2, smooth_rad=0.03)
        assert np.array_equal(emap[c][cells[:, 0], cells[:, 1], cells[:, 2]], v), c
    # without smoothing: no map, eps_inf in the LUT as before
    w0, lut0, m0 = host.voxelize(scene_path("test"), n, dx, ambient_eps=1.5, um_scale=1.0, resolution=n / L,
                                 with_map=True)
    assert m0 is None and any(m["eps_inf"] != 1.0 for m in lut0)


def test_voxelize_bowtie_materials():
    d = load("bowtie")
    s = d["settings"]
    L = s["len"] + 2 * s["pml_thickness"]
    n = 31
    dx = L / n
    words, lut = host.voxelize(scene_path("bowtie"), n, dx, ambient_eps=1.0, um_scale=s["um_scale"], resolution=n / L)
    assert all(m["eps_inf"] == 2.0 for m in lut)  # Q1 background
    kinds = {(m["npoles"], tuple(m["is_drude"])) for m in lut}
    assert kinds <= {(0, ()), (1, (1,)), (1, (0,))}
    us = s["um_scale"]
    for m in lut:
        if m["npoles"] == 1 and m["is_drude"][0]:
            assert m["f0"][0] == 1e-10 / us and m["gamma"][0] == 0.04274738474121455 / us
            assert m["sigma"][0] == 4.0314052191361974e21
        elif m["npoles"] == 1:
            assert m["f0"][0] == 9.67865314895427 / us and m["sigma"][0] == 1.12
    # every Yee point's material follows in() of the roots at that point
    au = [i for i, r in enumerate(d["roots"]) if r["metadata"].get("color") == 42][0]
    off = [(0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5)]
    rng = np.random.default_rng(3)
    cells = rng.integers(0, n, size=(400, 3))
    for c in range(3):
        pts = (cells + np.array(off[c])) * dx
        inside = host.scene_in(scene_path("bowtie"), pts)
        ids = (words[cells[:, 0], cells[:, 1], cells[:, 2]] >> (8 * c)) & 0xFF
        for (ins, mid) in zip(inside, ids):
            m = lut[mid]
            want_au, want_si = bool(ins[au]), bool(ins[1 - au])
            drudes = sum(m["is_drude"])
            assert drudes == int(want_au) and (m["npoles"] - drudes) == int(want_si)


# ---- field_samples.h5 (src/disp.cpp:758-923) ----------------------------------------------------

def test_field_samples_layout(tmp_path):
    path = str(tmp_path / "field_samples.h5")
    rng = np.random.default_rng(2)
    n_t_pts, span = 37, 4
    nsamp = -(-n_t_pts // span)  # samples taken: ceil(n_t_pts / span)
    series = rng.standard_normal((2, nsamp)) + 1j * rng.standard_normal((2, nsamp))
    um_scale, ttot = 2.0, 13.7
    host.write_field_samples(path, scene_path("run"), um_scale, ttot, n_t_pts, span, series)
    top = set(host.h5_list(path, "/"))
    assert top == {"info", "cluster_0", "cluster_1"}  # one cluster plus the trailing one
    tb = host.h5_read(path, "info/time_bounds")
    ttot_fs = ttot / (0.299792458 * um_scale)
    assert tb[0] == 0.0 and tb[1] == ttot_fs and tb[2] == ttot_fs * span / n_t_pts
    assert host.h5_read(path, "info/n_clusters")[0] == 1
    assert host.h5_read(path, "info/n_time_points")[0] == n_t_pts // span
    # cgs_params: the reference's small end-to-end case (src/main_test.cpp:1981-1988)
    params = set(host.h5_list(path, "info/cgs_params"))
    assert {"l_per_um", "sim_length", "pml_thickness", "tot_len"} <= params
    assert host.h5_read(path, "info/cgs_params/l_per_um")[0] == 2.0
    assert host.h5_read(path, "info/cgs_params/sim_length")[0] == 2.0
    assert host.h5_read(path, "info/cgs_params/pml_thickness")[0] == 1.0
    assert host.h5_read(path, "info/cgs_params/tot_len")[0] == 4.0
    src = host.h5_read(path, "info/sources", ["wavelen", "width", "phase", "start_time", "end_time", "amplitude"])
    g, _ = host.source_info(scene_path("run"), 0)
    assert src.tolist() == [[g["wavelen"], g["width"], g["phase"], g["start_time"], g["end_time"], g["amplitude"]]]
    locs = host.h5_read(path, "cluster_0/locations", ["x", "y", "z"])
    assert locs.tolist() == [[1.0, 1.0, 1.0], [2.0, 4.0, 1.1]]  # src/main_test.cpp:1905-1911
    assert host.h5_read(path, "cluster_1/locations", ["x", "y", "z"]).size == 0
    assert set(host.h5_list(path, "cluster_0")) == {"locations", "point_0", "point_1"}
    for p in range(2):
        t = host.h5_read(path, f"cluster_0/point_{p}/time", ["Re", "Im"])
        assert t.shape == (n_t_pts // span, 2)  # floor, while the series holds ceil samples
        assert np.array_equal(t[:, 0] + 1j * t[:, 1], series[p, : n_t_pts // span])
        f = host.h5_read(path, f"cluster_0/point_{p}/frequency", ["Re", "Im"])
        assert np.array_equal(f[:, 0] + 1j * f[:, 1], host.fft(series[p]))


def test_field_samples_zero_padded_names(tmp_path):
    # 40 clusters of 40 points: cluster_00..cluster_40 and point_0000..point_1599
    path = str(tmp_path / "fs.h5")
    series = np.ones((1600, 3), dtype=complex)
    host.write_field_samples(path, scene_path("bowtie"), 24.0, 5.0, 6, 2, series)
    names = host.h5_list(path, "/")
    assert "cluster_00" in names and "cluster_40" in names and len(names) == 42
    assert host.h5_list(path, "cluster_01")[:2] == ["locations", "point_0040"]
