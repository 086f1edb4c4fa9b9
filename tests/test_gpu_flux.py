"""Flux through DFT planes (sjf_dft_flux) against what a plane wave in vacuum fixes.

A current sheet in a box periodic in x and y (CPML on the z faces) radiates a plane pulse both
ways.  Meep units have unit impedance, so the flux of the +z wave through a plane is
dx^2 * sum |E_x|^2 (up to the half-cell Yee offset between E_x and H_y); the wave loses nothing
between two planes on the same side, and the -z wave behind the sheet carries the same flux
the other way.  The device reduction also equals the sum taken from the read-out spectra.
"""
import numpy as np
import pytest

from sim_juncs_amd import materials as M
from sim_juncs_amd import scenario
from sim_juncs_amd._lib import EngineError
from sim_juncs_amd.engine import Engine

pytestmark = pytest.mark.gpu

SHAPE = (8, 8, 120)
DX = 0.02
FREQS = [1.1, 1.25, 1.4]
TANGENTIAL_Z = ["Ex", "Ey", "Hx", "Hy"]


def plane_wave_engine():
    eng = Engine(SHAPE, DX, precision="fp64", periodic=(1, 1, 0))
    eng.set_materials(np.zeros(SHAPE, np.uint32), [M.VACUUM])
    eng.set_pml([0, 0, 0, 0, 12, 12])
    eng.add_source(**scenario.plane_source(SHAPE, DX, 30, scenario.gaussian_pulse(0.8, 1.5, 0.0, start_fs=0.0)))
    return eng


def test_plane_wave_flux():
    eng = plane_wave_engine()
    planes = {k: eng.add_dft((0, 0, k), (SHAPE[0], SHAPE[1], k + 1), TANGENTIAL_Z, FREQS) for k in (18, 50, 90)}
    steps = int((eng.last_source_time() + 3.0) / eng.dt) + 1
    print(f"[case] plane-wave flux, {steps} steps", flush=True)
    eng.step(steps)
    flux = {k: eng.dft_flux(i, 2) for k, i in planes.items()}
    for k, i in planes.items():
        d = eng.read_dft(i)  # [Ex, Ey, Hx, Hy][f][x][y][1]
        ref = DX * DX * np.real(d[0] * np.conj(d[3]) - d[1] * np.conj(d[2])).sum(axis=(1, 2, 3))
        assert np.allclose(flux[k], ref, rtol=1e-12, atol=0), k
        if k == 50:
            ex2 = DX * DX * (np.abs(d[0]) ** 2).sum(axis=(1, 2, 3))
            assert np.all(np.abs(flux[k] / ex2 - 1) < 2e-2), flux[k] / ex2
    assert np.all(flux[50] > 0)
    assert np.all(np.abs(flux[90] / flux[50] - 1) < 1e-4), flux[90] / flux[50]
    assert np.all(flux[18] < 0)
    assert np.all(np.abs(-flux[18] / flux[50] - 1) < 1e-3), flux[18] / flux[50]
    eng.close()


def test_flux_needs_a_plane_with_the_tangential_components():
    eng = plane_wave_engine()
    thick = eng.add_dft((0, 0, 50), (SHAPE[0], SHAPE[1], 52), TANGENTIAL_Z, FREQS)
    partial = eng.add_dft((0, 0, 60), (SHAPE[0], SHAPE[1], 61), ["Ex", "Ey", "Hx"], FREQS)
    ok = eng.add_dft((0, 0, 70), (SHAPE[0], SHAPE[1], 71), TANGENTIAL_Z, FREQS)
    eng.step(5)
    with pytest.raises(EngineError, match="thick"):
        eng.dft_flux(thick, 2)
    with pytest.raises(EngineError, match="tangential"):
        eng.dft_flux(partial, 2)
    with pytest.raises(EngineError, match="thick"):
        eng.dft_flux(ok, 0)  # along x the plane is 8 cells thick
    assert eng.dft_flux(ok, 2).shape == (len(FREQS),)
    eng.close()


def test_flux_through_a_metal_slab_matches_the_oracle():
    """A 2-cell Au slab (Drude + Lorentz) between two planes: the flux spectra equal the
    oracle's (its own D/E-split stepper and DFT, the flux summed from its spectra) to 1e-7, and
    the slab is passive, 0 < transmitted < net incident."""
    from oracle.oracle import Oracle

    mat = np.zeros(SHAPE + (3,), np.uint8)
    scenario.paint(mat, SHAPE, scenario.box_inside((-1, -1, 64.5), (40, 40, 66.5)), 1)
    lut = [M.VACUUM, M.AU.scaled(1.0)]
    src = scenario.plane_source(SHAPE, DX, 30, scenario.gaussian_pulse(0.8, 1.5, 0.0, start_fs=0.0))
    res = []
    for kind in (Engine, Oracle):
        sim = kind(SHAPE, DX, precision="fp64", periodic=(1, 1, 0))
        sim.set_materials(scenario.pack(mat), lut)
        sim.set_pml([0, 0, 0, 0, 12, 12])
        sim.add_source(**src)
        planes = [sim.add_dft((0, 0, k), (SHAPE[0], SHAPE[1], k + 1), TANGENTIAL_Z, FREQS) for k in (50, 90)]
        dt = sim.dt() if callable(sim.dt) else sim.dt
        sim.step(int((sim.last_source_time() + 3.0) / dt) + 1)
        if kind is Engine:
            res.append([sim.dft_flux(i, 2) for i in planes])
        else:
            fl = []
            for i in planes:
                d = sim.read_dft(i)
                fl.append(DX * DX * np.real(d[0] * np.conj(d[3]) - d[1] * np.conj(d[2])).sum(axis=(1, 2, 3)))
            res.append(fl)
        sim.close()
    (e50, e90), (o50, o90) = res
    assert np.allclose(e50, o50, rtol=1e-7, atol=0) and np.allclose(e90, o90, rtol=1e-7, atol=0)
    assert np.all(e90 > 0) and np.all(e90 < e50)
