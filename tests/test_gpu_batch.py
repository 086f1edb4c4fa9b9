"""Graph replay and batching on the GPU.

* Steps replayed from a captured hipGraph give bit-identical fields, monitor samples and DFT
  sums to steps launched one by one, replay after replay with direct steps in between.
* A batch of engines of uneven shapes, advanced by one fused launch per step, gives
  bit-identical results to the same engines stepped on their own, and matches the FP64 oracle
  under every tile form and x-chunk of the fused launch.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from sim_juncs_amd import scenario
from sim_juncs_amd._lib import EngineError
from sim_juncs_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu


def early(scns, fwhm=2.0):
    for s in scns:
        for src in s.sources:
            cep = src["phase"]
            src.update(scenario.gaussian_pulse(0.8, fwhm, cep, start_fs=0.0))
    return scns


def same(a, b):
    for c in range(6):
        assert np.array_equal(a.read_field(c), b.read_field(c)), c
    assert np.array_equal(a.read_monitors(), b.read_monitors())
    for i in range(len(a._dft)):
        assert np.array_equal(a.read_dft(i), b.read_dft(i))


def test_graph_replay_matches_direct_launches():
    scn = early([scenario.plain_box(n=24, steps=0, precision="fp32", pml=3)])[0]
    scn.save_span = 5
    print("[case] graph replay plain_box_24", flush=True)
    a = scn.make(Engine)
    a.set_graph(0)
    b = scn.make(Engine)
    b.set_graph(8)  # chunks of 10 steps (a multiple of the save span)
    for _ in range(4):
        a.step(37)
        b.step(37)
        same(a, b)
    assert np.abs(a.read_field("Ex")).max() > 0
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batch_matches_individual_engines(prec):
    scns = early(scenario.batched_sweep(n=5, seed=7, scale=0.3, precision=prec))
    print("[case] batch", prec, [s.shape for s in scns], flush=True)
    single = [s.make(Engine) for s in scns]
    for e in single:
        e.step(90)
    grouped = [s.make(Engine) for s in scns]
    b = Batch(grouped)
    b.step(50)
    b.step(40)
    for e1, e2 in zip(single, grouped):
        same(e1, e2)
        assert np.abs(e1.read_field("Ex")).max() > 0
    b.close()
    # after the batch is gone the engines step on their own again
    for e1, e2 in zip(single, grouped):
        e1.step(7)
        e2.step(7)
        same(e1, e2)
    for e in single + grouped:
        e.close()


def test_nontemporal_stores_change_no_bits():
    """The fused step stores with the non-temporal cache hint when a launch set's fields (E and
    H read and written, 48 bytes per FP32 cell) exceed 128 MB.  Three FP32 boxes below that
    alone (43-54 MB each) and above it together (139 MB) step to the same bits alone and as one
    batch, both in the one-row register tile."""
    scns = early([scenario.plain_box(n=96, steps=0, precision="fp32", seed=1),
                  scenario.plain_box(n=96, steps=0, precision="fp32", seed=2),
                  scenario.plain_box(n=104, steps=0, precision="fp32", seed=3)])
    fields = [48 * s.cells for s in scns]
    assert max(fields) < 128e6 < sum(fields)
    print("[case] non-temporal stores", [s.shape for s in scns], flush=True)
    single = [s.make(Engine) for s in scns]
    for e in single:
        e.set_tile("register_rows", 1, 8)
        e.step(60)
    grouped = [s.make(Engine) for s in scns]
    b = Batch(grouped)
    b.set_tile("register_rows", 8)
    b.step(60)
    assert b.launch_info()[0] == 128  # one row per thread over the 96- and 104-cell lines
    for e1, e2 in zip(single, grouped):
        same(e1, e2)
        assert np.abs(e1.read_field("Ex")).max() > 0
    b.close()
    for e in single + grouped:
        e.close()


@pytest.mark.parametrize("tile", [None, ("register_rows", 0), ("register_rows", 3), ("thread_rows", 2)])
def test_batch_matches_oracle_fp64(tile):
    scns = early(scenario.batched_sweep(n=3, seed=11, scale=0.25, precision="fp64"))
    print("[case] batch vs oracle", tile, [s.shape for s in scns], flush=True)
    eng = [s.make(Engine) for s in scns]
    b = Batch(eng)
    if tile:
        b.set_tile(*tile)
    with pytest.raises(EngineError):
        b.launch_info()  # nothing built before the first step
    b.step(80)
    threads, groups = b.launch_info()
    if tile and tile[0] == "register_rows":
        # one y row per workgroup over x-chunks of the batch's (or each engine's own) length
        xcs = [min(tile[1], s.shape[0]) if tile[1] else e.launch_info()["xchunk"] for s, e in zip(scns, eng)]
        assert groups == sum(s.shape[1] * -(-s.shape[0] // xc) for s, xc in zip(scns, xcs))
        assert threads == max(-(-s.shape[2] // 64) * 64 for s in scns)
    for s, e in zip(scns, eng):
        o = s.make(Oracle)
        o.step(80)
        refs = [o.read_field(c) for c in range(6)]
        for c in range(6):
            # cross-polarised components are small: normalise by the E (or H) family's norm
            fam = max(np.linalg.norm(refs[q]) for q in (range(3) if c < 3 else range(3, 6)))
            err = np.linalg.norm(e.read_field(c) - refs[c]) / fam
            assert err < 1e-10, (s.name, c, err)
        dref = o.read_dft(0)
        assert np.linalg.norm(e.read_dft(0) - dref) <= 1e-10 * np.linalg.norm(dref)
    b.close()
    for e in eng:
        e.close()
