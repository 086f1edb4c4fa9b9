"""Pins of the FP64 CPU oracle against closed forms (meep itself is unavailable).

Each test checks a property the mathematics fixes, chosen so that a plausible mistake in the
oracle (wrong curl sign or stencil offset, wrong dt/dx scaling, missing pole term, wrong source
normalisation or time alignment, wrong interpolation offsets) fails it:

* Yee plane-wave dispersion in a periodic box: sin^2(w dt/2)/dt^2 = sum sin^2(k dx/2)/dx^2.
* Discrete energy W = sum eps E^n.E^n + H^{n+1/2}.H^{n-1/2} is invariant in a closed cavity.
* Drude plasma oscillation and Lorentz longitudinal oscillation of a uniform field (curl-free):
  w^2 + i w gamma' = w_p^2 and w^2 + i w gamma' = w0^2 (1 + sigma/eps_inf), and the static limit
  E_inf = D / (eps_inf + sigma) of a Lorentz medium.
* 1-D current sheet radiates E = -K(t - |z|)/2 with K = d(dipole)/dt (c = 1, Z0 = 1).
* PML: reflections below 1e-4 of the incident peak (brute force against a doubled domain).
* DFT accumulators equal the sum over the sampled series; trilinear monitors are exact on
  linear fields; gaussian_src_time_phase metadata (reference src/main_test.cpp:1875-1882).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import Oracle
from sim_juncs_amd import materials as M
from sim_juncs_amd import scenario


def pack_uniform(shape, mid=0):
    m = np.full(shape + (3,), mid, np.uint8)
    return scenario.pack(m)


# ----------------------------------------------------------------------------- Yee ---------

@pytest.mark.parametrize("comp,axis_k,m", [
    ("Ex", (0, 0, 1), 3), ("Ey", (1, 0, 0), 2), ("Ez", (0, 1, 0), 5), ("Ez", (1, 2, 0), 1),
])
def test_yee_plane_wave_dispersion(comp, axis_k, m):
    """A discrete plane wave in a periodic box oscillates at the Yee frequency."""
    n = (12, 10, 14)
    dx, S = 0.1, 0.45
    dt = S * dx
    o = Oracle(n, dx, courant=S, periodic=(1, 1, 1))
    kvec = [2 * math.pi * m * axis_k[d] / (n[d] * dx) for d in range(3)]
    c = ["Ex", "Ey", "Ez"].index(comp)
    off = scenario.COMP_OFF[c]
    ii, jj, kk = np.meshgrid(*[np.arange(n[d]) + off[d] for d in range(3)], indexing="ij")
    phase = kvec[0] * ii * dx + kvec[1] * jj * dx + kvec[2] * kk * dx
    # polarisation must be transverse to k
    assert abs(axis_k[c]) == 0
    o.write_field(comp, np.cos(phase))
    lam = sum(4 / dx ** 2 * math.sin(kvec[d] * dx / 2) ** 2 for d in range(3))
    theta = 2 * math.asin(math.sqrt(lam) * dt / 2)
    probe = (3, 4, 5)
    e0 = math.cos(phase[probe])
    vals = []
    for _ in range(60):
        o.step(1)
        vals.append(o.read_field(comp)[probe])
    steps = np.arange(1, 61)
    # E^n = e0 (cos n theta - tan(theta/2) sin n theta) for E^0 = e0, H^{-1/2} = 0
    want = e0 * (np.cos(steps * theta) - math.tan(theta / 2) * np.sin(steps * theta))
    assert np.max(np.abs(np.array(vals) - want)) < 1e-11


def test_energy_invariant_closed_cavity():
    rng = np.random.default_rng(0)
    shape = (7, 6, 5)
    o = Oracle(shape, dx=0.1, courant=0.5)
    mats = rng.integers(0, 3, size=shape + (3,))
    o.set_materials(scenario.pack(mats), [M.VACUUM, M.MaterialSpec(2.5), M.MaterialSpec(4.0)])
    eps = np.array([1.0, 2.5, 4.0])[mats]
    for c in range(6):
        o.write_field(c, rng.standard_normal(shape))
    W = []
    for _ in range(200):
        E = [o.read_field(c) for c in range(3)]
        H0 = [o.read_field(c) for c in range(3, 6)]
        o.step(1)
        H1 = [o.read_field(c) for c in range(3, 6)]
        W.append(sum((eps[..., c] * E[c] ** 2).sum() for c in range(3)) + sum((H0[c] * H1[c]).sum() for c in range(3)))
    W = np.array(W)
    assert np.abs(W - W[0]).max() / W[0] < 1e-13


def test_unstable_courant_diverges():
    """Above the 3-D Yee limit S = 1/sqrt(3) the scheme must blow up (negative control for the
    dispersion test: it is the dt/dx scaling, not luck, that keeps the runs above bounded)."""
    o = Oracle((6, 6, 6), 0.1, courant=0.62, periodic=(1, 1, 1))
    rng = np.random.default_rng(1)
    o.write_field("Ex", rng.standard_normal((6, 6, 6)))
    o.step(400)
    assert np.abs(o.read_field("Ex")).max() > 1e6


# ----------------------------------------------------------------------------- poles -------

def uniform_medium_series(spec, steps, dt_scale=0.5, dx=0.01, e0=1.0):
    n = (2, 2, 2)
    o = Oracle(n, dx, courant=dt_scale, periodic=(1, 1, 1))
    o.set_materials(pack_uniform(n), [spec])
    o.write_field("Ex", np.full(n, e0))
    out = []
    for _ in range(steps):
        o.step(1)
        out.append(o.read_field("Ex")[0, 0, 0])
    return np.array(out), o.dt


def fit_damped(x, dt):
    """Fit x_n = A exp(-a n dt) cos(w n dt + phi) + c by linear prediction (Prony, order 2):
    x_{n+1} = p x_n + q x_{n-1} + r; returns (w, a, c) with c = r / (1 - p - q)."""
    X = np.stack([x[1:-1], x[:-2], np.ones(len(x) - 2)], 1)
    p, q, r = np.linalg.lstsq(X, x[2:], rcond=None)[0]
    z = np.roots([1, -p, -q])
    z = z[np.argmax(z.imag)]
    w = abs(np.angle(z)) / dt
    a = -math.log(abs(z)) / dt
    return w, a, r / (1 - p - q)


def test_drude_plasma_oscillation():
    f0, gamma, eps_inf = 1e-3, 0.05, 1.5
    sigma = 2.0 / f0 ** 2  # sigma * (2 pi f0)^2 = 8 pi^2
    spec = M.MaterialSpec(eps_inf, (M.Pole(f0, gamma, sigma, True),))
    x, dt = uniform_medium_series(spec, 600)
    w, a, c = fit_damped(x, dt)
    wp2 = sigma * (2 * math.pi * f0) ** 2 / eps_inf
    g = 2 * math.pi * gamma
    w_cont = math.sqrt(wp2 - g * g / 4)
    assert abs(w - w_cont) / w_cont < 2e-3
    assert abs(a - g / 2) / (g / 2) < 2e-3
    # a Drude medium screens a static field completely: the oscillation's fixed point is 0
    assert abs(c) < 1e-9


def test_lorentz_longitudinal_oscillation_and_static_limit():
    f0, gamma, sigma, eps_inf = 1.2, 0.02, 1.7, 2.1
    spec = M.MaterialSpec(eps_inf, (M.Pole(f0, gamma, sigma, False),))
    x, dt = uniform_medium_series(spec, 4000, dx=0.005)
    w0 = 2 * math.pi * f0
    g = 2 * math.pi * gamma
    wl2 = w0 * w0 * (1 + sigma / eps_inf)
    w_cont = math.sqrt(wl2 - g * g / 4)
    w, a, c = fit_damped(x, dt)
    assert abs(w - w_cont) / w_cont < 2e-3
    assert abs(a - g / 2) / (g / 2) < 5e-3
    # static permittivity eps(0) = eps_inf + sigma: D = eps_inf E0 is conserved (no curl), so
    # the field relaxes to D / eps(0)
    assert abs(c - eps_inf / (eps_inf + sigma)) < 1e-9


# ----------------------------------------------------------------------------- source ------

def sheet_run(nz, src_z, steps, pml=40, fwhm=4.0, dx=0.01, mon_dz=(60,)):
    o = Oracle((1, 1, nz), dx, courant=0.5, periodic=(1, 1, 0))
    o.set_pml([0, 0, 0, 0, pml, pml])
    pulse = scenario.gaussian_pulse(0.8, fwhm, 0.4, start_fs=0.0)
    o.add_source("Ex", [0, 0, src_z * dx, dx, dx, src_z * dx], **pulse)
    o.add_monitors("Ex", np.array([[0.5 * dx, 0.0, (src_z + d) * dx] for d in mon_dz]))
    o.set_save_span(1)
    o.step(steps)
    return o, pulse


def test_current_sheet_radiation_matches_green_function():
    nz, src_z, d = 900, 450, 60
    dx = 0.01
    o, pulse = sheet_run(nz, src_z, 2400, dx=dx, mon_dz=(d, -d))
    s = o.read_monitors().real
    dt = o.dt
    t = np.arange(s.shape[0]) * dt
    # K(t) = d/dt Re dipole(t), dipole of gaussian_src_time_phase (closed form)
    w = 2 * math.pi * pulse["freq"]
    width, phi = pulse["width"], pulse["phase"] + math.pi
    peak = 0.5 * (pulse["t_start"] + pulse["t_end"])

    def redip(tt):
        tau = tt - peak
        return np.exp(-tau ** 2 / (2 * width ** 2)) * np.sin(w * tau + phi) / w

    h = 1e-6
    K = lambda tt: (redip(tt + h) - redip(tt - h)) / (2 * h)  # noqa: E731
    for col in range(2):
        want = -0.5 * K(t - d * dx)
        err = np.linalg.norm(s[:, col] - want) / np.linalg.norm(want)
        assert err < 1e-2, err


def test_pml_reflection_small():
    """Monitor series next to the PML equals that of a doubled domain to 1e-4 of the peak."""
    d = 100
    a, _ = sheet_run(500, 250, 2000, pml=24, mon_dz=(d,))
    b, _ = sheet_run(1500, 750, 2000, pml=24, mon_dz=(d,))
    sa, sb = a.read_monitors().real[:, 0], b.read_monitors().real[:, 0]
    assert np.abs(sb).max() > 0
    assert np.abs(sa - sb).max() < 1e-4 * np.abs(sb).max()


def test_gaussian_src_time_phase_metadata():
    """reference src/main_test.cpp:1875-1882: f = 1 -> omega = 2 pi, width 2, fwidth 0.5, and the
    closed form of dipole at the peak: exp(0) exp(-i pi) / (-i omega) = -i/(2 pi)."""
    v, cut = orc.gaussian_dipole(1.0, 2.0, 0.0, 0.0, 2.0, 1.0)
    assert cut == 1.0
    assert abs(v - (-1j / (2 * math.pi))) < 1e-15
    # cutoff shrinks by 0.9 while exp(-c^2/2w^2) < 1e-100, then rounds to float
    _, cut = orc.gaussian_dipole(1.0, 1.0, 0.0, 0.0, 100.0, 0.0)
    assert cut == float(np.float32(50 * 0.9 ** 9))
    # outside the cutoff the dipole is exactly zero
    v, _ = orc.gaussian_dipole(1.0, 2.0, 0.3, 0.0, 2.0, 2.0000001)
    assert v == 0


def test_gaussian_source_window_kat():
    """Gaussian_source end time = start + 2 * 6 * width once a cutoff is given
    (reference src/disp.cpp:369-373; KAT src/main_test.cpp:1793-1794: 5.2 -> 41.2, width 3)."""
    ca = scenario.LIGHT_SPEED
    p = scenario.gaussian_pulse(1.333333, 3.0 / scenario.FWHM_TO_WIDTH, 0.75, start_fs=5.2)
    assert abs(p["t_start"] / ca - 5.2) < 1e-12
    assert abs(p["t_end"] / ca - 41.2) < 1e-12
    assert abs(p["width"] / ca - 3.0) < 1e-12


# ----------------------------------------------------------------------------- monitors ----

def test_trilinear_monitors_exact_on_linear_fields():
    n = (6, 7, 8)
    dx = 0.05
    o = Oracle(n, dx)
    rng = np.random.default_rng(3)
    coef = rng.standard_normal(4)
    for c, name in enumerate(["Ex", "Ey", "Hz"]):
        comp = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"].index(name)
        off = scenario.COMP_OFF[comp]
        ii, jj, kk = np.meshgrid(*[(np.arange(n[d]) + off[d]) * dx for d in range(3)], indexing="ij")
        o.write_field(name, coef[0] + coef[1] * ii + coef[2] * jj + coef[3] * kk)
    pts = rng.uniform(0.15, 0.25, size=(9, 3))
    for name in ["Ex", "Ey", "Hz"]:
        o.add_monitors(name, pts)
    o.step(1)
    s = o.read_monitors().real[0]
    want = coef[0] + pts @ coef[1:]
    for q in range(3):
        assert np.max(np.abs(s[9 * q:9 * q + 9] - want)) < 1e-12


# ----------------------------------------------------------------------------- DFT ---------

def test_dft_equals_sum_over_sampled_series():
    nz, src_z = 400, 200
    dx = 0.01
    o = Oracle((1, 1, nz), dx, periodic=(1, 1, 0))
    o.set_pml([0, 0, 0, 0, 20, 20])
    pulse = scenario.gaussian_pulse(0.8, 3.0, 0.2, start_fs=0.0)
    o.add_source("Ex", [0, 0, src_z * dx, dx, dx, src_z * dx], **pulse)
    zk = src_z + 30
    o.add_monitors("Ex", np.array([[0.5 * dx, 0, zk * dx]]))
    o.add_monitors("Hy", np.array([[0.5 * dx, 0, (zk + 0.5) * dx]]))
    o.set_save_span(1)
    freqs = np.array([0.9, 1.25, 1.6])
    i = o.add_dft((0, 0, zk), (1, 1, zk + 1), ("Ex", "Hy"), freqs)
    N = 900
    o.step(N)
    s = o.read_monitors().real
    ex = np.concatenate([s[1:, 0], [o.read_field("Ex")[0, 0, zk]]])  # E at t = m dt, m = 1..N
    hy = np.concatenate([s[1:, 1], [o.read_field("Hy")[0, 0, zk]]])  # H at t = (m - 1/2) dt
    dt = o.dt
    m = np.arange(1, N + 1)
    dft = o.read_dft(i)
    for f_i, f in enumerate(freqs):
        want_e = np.sum(ex * np.exp(2j * math.pi * f * m * dt)) * dt
        want_h = np.sum(hy * np.exp(2j * math.pi * f * (m - 0.5) * dt)) * dt
        assert abs(dft[0, f_i, 0, 0, 0] - want_e) < 1e-12 * max(1.0, abs(want_e))
        assert abs(dft[1, f_i, 0, 0, 0] - want_h) < 1e-12 * max(1.0, abs(want_h))
