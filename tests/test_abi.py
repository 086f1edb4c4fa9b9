"""C ABI and host-side checks that need no GPU."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from sim_juncs_amd import _lib
from sim_juncs_amd import materials as M
from sim_juncs_amd import scenario

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "sjf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sjf_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTS) == names


def test_host_library_exports_every_declared_symbol():
    from sim_juncs_amd import host

    src = open(os.path.join(ROOT, "include", "sjf_host.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(sjfh_[a-z_0-9]+)\s*\(", src)))
    assert len(names) >= 12
    lib = host.lib()
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layouts_match_header():
    # sizes implied by include/sjf.h on x86-64 (int = 4, double = 8, natural alignment)
    assert C.sizeof(_lib.Grid) == 3 * 4 + 4 + 8 + 8 + 4 + 3 * 4 + 4 + 4  # 4 bytes pad before dx
    assert C.sizeof(_lib.Material) == 8 + 8 + 3 * 4 * 8 + 4 * 4
    assert C.sizeof(_lib.Pulse) == 8 + 6 * 8


def _gpu_visible():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-device error path")
def test_create_without_device_fails_cleanly():
    lib = _lib.load()
    g = _lib.Grid()
    g.nx = g.ny = g.nz = 8
    g.dx, g.courant, g.precision = 0.1, 0.5, 4
    h = C.c_void_p()
    rc = lib.sjf_create(C.byref(g), C.byref(h))
    assert rc == -2  # SJF_E_HIP
    assert b"device" in lib.sjf_last_error(None)
    assert not h.value


def test_create_rejects_bad_arguments_before_touching_a_device():
    lib = _lib.load()
    g = _lib.Grid()
    g.nx, g.ny, g.nz = 8, 8, 2000
    g.dx, g.courant, g.precision = 0.1, 0.5, 4
    g.periodic[2] = 1
    h = C.c_void_p()
    assert lib.sjf_create(C.byref(g), C.byref(h)) == -4  # nz > 1024 needs z passes: not periodic
    g.periodic[2] = 0
    g.nz, g.precision = 8, 3
    assert lib.sjf_create(C.byref(g), C.byref(h)) == -1


def test_rakic_gold_poles_match_ev_parameters():
    """AU poles are Rakic 1998 (omega_p = 9.03 eV) converted with hc = 1.23984193 eV um:
    f = omega / hc, gamma = Gamma / hc, sigma = f_j omega_p^2 / omega_j^2 (Drude: omega_0 = 1e-10 um^-1
    as junctions/Au_SiO2_box/junc.geom:60 writes it)."""
    hc = 1.23984193
    wp = 9.03
    d, l1, l2 = M.AU.poles
    assert abs(d.gamma - 0.053 / hc) < 1e-12
    assert abs(d.sigma * d.f0 ** 2 - 0.760 * (wp / hc) ** 2) / d.sigma / d.f0 ** 2 < 1e-9
    assert abs(l1.f0 - 0.415 / hc) < 1e-12 and abs(l1.gamma - 0.241 / hc) < 1e-12
    assert abs(l1.sigma - 0.024 * wp ** 2 / 0.415 ** 2) < 1e-8
    assert abs(l2.f0 - 0.830 / hc) < 1e-9 and abs(l2.gamma - 0.345 / hc) < 1e-9
    assert abs(l2.sigma - 0.010 * wp ** 2 / 0.830 ** 2) < 1e-8


def test_gold_permittivity_at_800nm_is_plasmonic():
    eps = complex(M.AU.eps(1 / 0.8))
    assert -30 < eps.real < -20 and 0.5 < eps.imag < 3


def test_scenarios_voxelise_expected_regions():
    s = scenario.plain_box()
    w = s.matw & 0xFF
    assert s.shape == (96, 96, 96)
    assert (w == 2).sum() > 0 and (w == 1).sum() == 96 * 96 * 20
    b = scenario.bowtie_junction(nx=128, ny=128, nz=64, pml=6)
    wb = b.matw & 0xFF
    assert set(np.unique(wb)) == {0, 1, 2, 3}
    # the two triangles are mirror images about the gap centre
    au = (wb == 2)
    assert au.sum() > 0
    assert abs(int(au[:64].sum()) - int(au[64:].sum())) <= au.sum() * 0.05


def test_graphene_sheet_drude_weight():
    g = M.graphene_sheet(0.005).poles[0]
    e, hbar, eps0, c = 1.602176634e-19, 1.054571817e-34, 8.8541878128e-12, 299792458.0
    wp2 = e * e * 0.4 * e / (math.pi * hbar ** 2 * eps0 * 0.005e-6)
    assert abs(g.sigma * g.f0 ** 2 - wp2 / (2 * math.pi * c) ** 2 * 1e-12) < 1e-12
