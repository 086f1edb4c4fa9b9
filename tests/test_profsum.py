"""The profile tools behind bench.py's roofline.traffic: kernel-name matching, the copy-kernel
calibration of FETCH_SIZE / WRITE_SIZE, and when a stored record is reported.  The inputs are
small SQLite files with the columns rocprofv3's results database holds."""
import json
import sqlite3

import pytest

from sim_juncs_amd import profsum

COPY = "void sjf::k_copy(float4 const*, float4*, long)"
STEP = "void sjf::k_step<float, 1, 3, 256>(sjf::StepArgs<float> const*, int const*, int)"
ROWS = "void sjf::k_step_rows<float, 3, 1024>(sjf::StepArgs<float> const*, int const*, int)"


def make_db(path, rows, kernels):
    con = sqlite3.connect(path)
    con.execute("create table pmc_events (name text, counter_name text, counter_value real, duration real)")
    con.executemany("insert into pmc_events values (?, ?, ?, 0)", rows)
    con.execute("create table kernels (name text, grid_x int, grid_y int, grid_z int, workgroup_x int, "
                "workgroup_y int, workgroup_z int, duration real)")
    con.executemany("insert into kernels values (?, ?, ?, ?, ?, ?, ?, 0)", kernels)
    con.commit()
    con.close()
    return str(path)


def test_kernel_names_match_exactly():
    assert profsum.is_kernel(STEP, "k_step")
    assert not profsum.is_kernel(ROWS, "k_step")
    assert profsum.is_kernel(ROWS, "k_step_rows")
    assert profsum.is_kernel(COPY, "k_copy")


def test_traffic_is_scaled_by_the_copy_kernel_and_stored_per_shape(tmp_path, monkeypatch):
    monkeypatch.setattr(profsum, "TRAFFIC", str(tmp_path / "traffic.json"))
    gib_kib = (1 << 30) / 1024.0
    # the copy kernel reads and writes 1 GiB: FETCH_SIZE reads half of it, WRITE_SIZE all of it
    cf = make_db(tmp_path / "cf.db", [(COPY, "FETCH_SIZE", gib_kib / 2)], [(COPY, 1024, 1, 1, 256, 1, 1)])
    cw = make_db(tmp_path / "cw.db", [(COPY, "WRITE_SIZE", gib_kib)], [(COPY, 1024, 1, 1, 256, 1, 1)])
    launches = [(STEP, 256 * 100, 1, 1, 256, 1, 1)] * 2 + [(ROWS, 1024 * 7, 1, 1, 1024, 1, 1)]
    kf = make_db(tmp_path / "kf.db", [(STEP, "FETCH_SIZE", 1000.0), (STEP, "FETCH_SIZE", 3000.0),
                                      (ROWS, "FETCH_SIZE", 1e9)], launches)
    kw = make_db(tmp_path / "kw.db", [(STEP, "WRITE_SIZE", 500.0), (STEP, "WRITE_SIZE", 500.0)], launches)
    profsum.traffic("w", "k_step", kf, kw, cf, cw, store=True)
    rec = json.load(open(profsum.TRAFFIC))["w"]["k_step"]
    assert rec["fetch_factor"] == 2.0 and rec["write_factor"] == 1.0
    assert rec["fetch_kib"] == 2000.0 and rec["write_kib"] == 500.0
    assert rec["traffic_bytes"] == (2000.0 * 2.0 + 500.0) * 1024.0
    assert (rec["threads"], rec["groups"]) == (256, 100)
    assert profsum.stored_traffic("w", "k_step", 256, 100)["traffic_bytes"] == rec["traffic_bytes"]
    # another launch shape, another workload or other device code: no figure
    assert profsum.stored_traffic("w", "k_step", 256, 101) is None
    assert profsum.stored_traffic("v", "k_step", 256, 100) is None
    monkeypatch.setattr(profsum, "kernel_source_hash", lambda: "0" * 16)
    assert profsum.stored_traffic("w", "k_step", 256, 100) is None


def test_traffic_needs_one_launch_shape(tmp_path, monkeypatch):
    monkeypatch.setattr(profsum, "TRAFFIC", str(tmp_path / "traffic.json"))
    cf = make_db(tmp_path / "cf.db", [(COPY, "FETCH_SIZE", 1.0)], [(COPY, 1, 1, 1, 1, 1, 1)])
    cw = make_db(tmp_path / "cw.db", [(COPY, "WRITE_SIZE", 1.0)], [(COPY, 1, 1, 1, 1, 1, 1)])
    mixed = [(STEP, 256 * 100, 1, 1, 256, 1, 1), (STEP, 256 * 50, 1, 1, 256, 1, 1)]
    kf = make_db(tmp_path / "kf.db", [(STEP, "FETCH_SIZE", 1.0)], mixed)
    kw = make_db(tmp_path / "kw.db", [(STEP, "WRITE_SIZE", 1.0)], mixed)
    with pytest.raises(SystemExit):
        profsum.traffic("w", "k_step", kf, kw, cf, cw)
