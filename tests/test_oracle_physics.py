"""Pins of the FP64 CPU oracle against closed forms (meep itself is unavailable).

Each test checks a property the mathematics fixes, chosen so that a plausible mistake in the
oracle (wrong curl sign or stencil offset, wrong dt/dx scaling, missing pole term, wrong source
normalisation or time alignment, wrong interpolation offsets) fails it:

* Yee plane-wave dispersion in a periodic box: sin^2(w dt/2)/dt^2 = sum sin^2(k dx/2)/dx^2,
  and with per-component eps maps the same over eps_c; a design box equals the same values
  given as maps.
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


@pytest.mark.parametrize("comp,axis_k,m,eps", [
    ("Ex", (0, 0, 1), 3, (2.25, 1.0, 1.0)), ("Ey", (1, 0, 0), 2, (1.0, 3.0, 1.0)),
    ("Ez", (0, 1, 0), 5, (1.0, 1.0, 1.7)), ("Ex", (0, 1, 1), 1, (4.0, 2.0, 1.5)),
])
def test_eps_maps_set_each_components_permittivity(comp, axis_k, m, eps):
    """Per-cell eps_inf maps (one array per E component): a plane wave polarised along c in a
    medium whose map gives c the value eps_c oscillates at the Yee frequency of lambda / eps_c,
    sin^2(w dt/2)/dt^2 = sum sin^2(k dx/2)/dx^2 / eps_c.  The other components' values differ,
    so a map applied to the wrong component changes the frequency."""
    n = (12, 10, 14)
    dx, S = 0.1, 0.45
    dt = S * dx
    o = Oracle(n, dx, courant=S, periodic=(1, 1, 1))
    o.set_eps_maps(np.stack([np.full(n, e) for e in eps]))
    kvec = [2 * math.pi * m * axis_k[d] / (n[d] * dx) for d in range(3)]
    c = ["Ex", "Ey", "Ez"].index(comp)
    assert axis_k[c] == 0
    off = scenario.COMP_OFF[c]
    ii, jj, kk = np.meshgrid(*[np.arange(n[d]) + off[d] for d in range(3)], indexing="ij")
    phase = kvec[0] * ii * dx + kvec[1] * jj * dx + kvec[2] * kk * dx
    o.write_field(comp, np.cos(phase))
    lam = sum(4 / dx ** 2 * math.sin(kvec[d] * dx / 2) ** 2 for d in range(3)) / eps[c]
    theta = 2 * math.asin(math.sqrt(lam) * dt / 2)
    probe = (3, 4, 5)
    e0 = math.cos(phase[probe])
    vals = []
    for _ in range(60):
        o.step(1)
        vals.append(o.read_field(comp)[probe])
    steps = np.arange(1, 61)
    want = e0 * (np.cos(steps * theta) - math.tan(theta / 2) * np.sin(steps * theta))
    assert np.max(np.abs(np.array(vals) - want)) < 1e-11


def test_design_box_equals_the_same_permittivity_given_as_maps():
    """A design box (one eps per voxel, z fastest, for all three E components, no poles) and eps
    maps holding the same values in the box and 1 elsewhere give the same fields: the box's
    voxel order and extent are those of the maps' cells."""
    n = (11, 9, 13)
    lo, hi = (2, 3, 4), (9, 7, 10)
    rng = np.random.default_rng(7)
    box = rng.uniform(1.2, 4.0, size=tuple(h - l for l, h in zip(lo, hi)))
    runs = []
    for mode in ("box", "maps"):
        o = Oracle(n, 0.02, courant=0.5)
        o.set_pml([2] * 6)
        if mode == "box":
            o.set_design_eps(lo, hi, box)
        else:
            e = np.ones(n)
            e[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = box
            o.set_eps_maps(np.stack([e, e, e]))
        o.add_source("Ey", [0, 0, 0.1, 0.22, 0.18, 0.1], **scenario.gaussian_pulse(0.8, 1.5, 0.4, start_fs=0.0))
        o.step(80)
        runs.append([o.read_field(c) for c in range(6)])
    assert np.abs(runs[0][1]).max() > 0
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


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


def logistic(x):
    return 1.0 / (1.0 + np.exp(-x))


def test_continuous_src_time_closed_form():
    """meep continuous_src_time (reference src/disp.cpp:618): dipole(t) = env(t) e^{-i w t} / (-i w)
    with env = (1 + tanh(ts)) (1 + tanh(te)) / 4, ts = (t - t0)/width - s, te = (t1 - t)/width - s.
    Pinned here through identities a slip would break: the steady state times -i w is exactly
    e^{-i w t}; (1 + tanh x)/2 is the logistic function of 2x (so env = sigma(2 ts) sigma(2 te));
    env is 1/2 at t0 + s width far from the end; turn-on and turn-off mirror each other; the
    float gate switches on where float(t) >= t0."""
    f, st, et = 1.3, 0.1, 40.0
    w = 2 * math.pi * f
    for t in (0.37, 5.0, 17.123, 39.9):
        v = orc.cw_dipole(f, 0.0, st, et, 3.0, t)
        assert abs(v * (-1j * w) - np.exp(-1j * w * t)) < 1e-14
        assert abs(abs(v) - 1 / w) < 1e-15
    width, slow = 0.7, 3.0
    ts_ = np.linspace(st, et, 997)
    for t in ts_:
        v = orc.cw_dipole(f, width, st, et, slow, t)
        env = logistic(2 * ((t - st) / width - slow)) * logistic(2 * ((et - t) / width - slow))
        assert abs(abs(v) * w - env) < 1e-13
        assert abs(v * (-1j * w) - env * np.exp(-1j * w * t)) < 1e-13
    half = abs(orc.cw_dipole(f, width, st, et, slow, st + slow * width)) * w
    assert abs(half - 0.5) < 1e-12
    for x in (0.05, 0.8, 2.1, 3.3):
        a = abs(orc.cw_dipole(f, width, st, et, slow, st + x)) * w
        b = abs(orc.cw_dipole(f, width, st, et, slow, et - x)) * w
        assert abs(a - b) < 1e-12
    # float gate: float(0.1) > 0.1, so a time just below t0 that rounds up to it is inside
    assert orc.cw_dipole(f, 0.0, st, et, slow, st - 1e-10) != 0
    assert orc.cw_dipole(f, 0.0, st, et, slow, st - 1e-4) == 0
    assert orc.cw_dipole(f, 0.0, st, et, slow, et + 1e-4) == 0


def test_continuous_source_current_is_a_sampled_derivative():
    """The injected current (dipole(t + dt) - dipole(t)) / dt of a steady CW source is
    e^{-i w (t + dt/2)} sin(w dt/2)/(w dt/2): unit amplitude up to the Yee sinc factor, centred
    half a step after t."""
    f, dt = 1.25, 0.005
    w = 2 * math.pi * f
    for t in (1.0, 3.7, 9.99):
        J = (orc.cw_dipole(f, 0.0, 0.0, 100.0, 3.0, t + dt) - orc.cw_dipole(f, 0.0, 0.0, 100.0, 3.0, t)) / dt
        want = np.exp(-1j * w * (t + dt / 2)) * math.sin(w * dt / 2) / (w * dt / 2)
        assert abs(J - want) < 1e-10


def cw_sheet_run(nz, src_z, steps, comp="Ex", mon_comp="Ex", off=0.0, pml=40, dx=0.01, mon_dz=(60, -60)):
    o = Oracle((1, 1, nz), dx, courant=0.5, periodic=(1, 1, 0))
    o.set_pml([0, 0, 0, 0, pml, pml])
    pulse = scenario.cw_pulse(0.8, 2.0, 1.0, 1e6)
    z = (src_z + off) * dx
    o.add_source(comp, [0, 0, z, dx, dx, z], **pulse)
    o.add_monitors(mon_comp, np.array([[0.5 * dx, 0.0, (src_z + d) * dx] for d in mon_dz]))
    o.set_save_span(1)
    o.step(steps)
    return o, pulse


def cw_K(pulse):
    """d/dt Re dipole(t) of the continuous source, from the logistic form of its envelope."""
    w = 2 * math.pi * pulse["freq"]
    st, et, wd, s = pulse["t_start"], pulse["t_end"], pulse["width"], pulse["slowness"]

    def redip(tt):
        env = logistic(2 * ((tt - st) / wd - s)) * logistic(2 * ((et - tt) / wd - s))
        return np.where(tt >= st, env * np.sin(w * tt) / w, 0.0)

    h = 1e-6
    return lambda tt: (redip(tt + h) - redip(tt - h)) / (2 * h)


def test_cw_current_sheet_radiates_green_function():
    """An electric CW sheet radiates E = -K(t - |z|)/2 on both sides (c = 1, Z0 = 1), through
    the tanh turn-on into the steady sinusoid."""
    dx, d = 0.01, 60
    o, pulse = cw_sheet_run(900, 450, 3000, dx=dx)
    s = o.read_monitors().real
    t = np.arange(s.shape[0]) * o.dt
    K = cw_K(pulse)
    want = -0.5 * K(t - d * dx)
    assert np.abs(want).max() > 0.4  # steady state reached (|K| -> 1)
    for col in range(2):
        err = np.linalg.norm(s[:, col] - want) / np.linalg.norm(want)
        assert err < 1e-2, err


def test_magnetic_current_sheet_radiates_dual_field():
    """A magnetic current sheet K (component Hy, on the Hy lattice half a cell above z0) radiates
    Hy = -K(t - |z|)/2 on both sides and Ex = -sign(z) K(t - |z|)/2 (dB/dt = -curl E - K, the
    dual of the electric sheet): Ex flips sign across the sheet, Hy does not."""
    dx, d = 0.01, 60
    o, pulse = cw_sheet_run(900, 450, 3000, comp="Hy", mon_comp="Ex", off=0.5, dx=dx)
    s = o.read_monitors().real
    t = np.arange(s.shape[0]) * o.dt
    K = cw_K(pulse)
    up = -0.5 * K(t - (d - 0.5) * dx)
    down = 0.5 * K(t - (d + 0.5) * dx)
    assert np.linalg.norm(s[:, 0] - up) / np.linalg.norm(up) < 1e-2
    assert np.linalg.norm(s[:, 1] - down) / np.linalg.norm(down) < 1e-2
    h, _ = cw_sheet_run(900, 450, 3000, comp="Hy", mon_comp="Hy", off=0.5, dx=dx, mon_dz=(60.5, -59.5))
    sh = h.read_monitors().real
    # Hy samples are H^{n-1/2}: half a step earlier than the E samples
    th = t - 0.5 * o.dt
    for col, dist in ((0, 60.0), (1, 60.0)):
        want = -0.5 * K(th - dist * dx)
        assert np.linalg.norm(sh[:, col] - want) / np.linalg.norm(want) < 1e-2


def column_cw(nz, src, mon, steps, fold):
    """1-D column (periodic 1x1 cross-section) of SiO2 (one Lorentz pole) filling every cell,
    PML on both z faces, a CW Ex sheet at f = 1.25 (800 nm) with a smooth turn-on."""
    dx = 0.01
    mat = np.ones((1, 1, nz, 3), np.uint8)
    o = Oracle((1, 1, nz), dx, courant=0.5, periodic=(1, 1, 0))
    o.set_materials(scenario.pack(mat), [M.VACUUM, M.SIO2.scaled(1.0)])
    o.set_pml([0, 0, 0, 0, 40, 40])
    pulse = scenario.cw_pulse(0.8, 3.0, 0.0, 1e6)
    if fold:
        o.set_pml_dispersion(pulse["freq"])
    o.add_source("Ex", [0, 0, src * dx, dx, dx, src * dx], **pulse)
    o.add_monitors("Ex", np.array([[0.5 * dx, 0.0, mon * dx]]))
    o.set_save_span(1)
    o.step(steps)
    return o.read_monitors().real[:, 0]


def test_pml_with_lorentz_poles_folded_is_matched_at_f_ref():
    """With the Lorentz poles folded out of the PML at f_ref, the slabs hold a non-dispersive
    medium of eps = eps_inf + Re chi(f_ref), the interior's permittivity at f_ref: a CW wave at
    f_ref meets no interface.  The series 150 cells in front of the PML equals that of a column
    long enough that nothing comes back within the run, to 2e-3 of its amplitude; an unfolded
    eps_inf (1 instead of 2.14) in the slabs would reflect ~19%."""
    steps = 3200
    near = column_cw(600, 300, 410, steps, fold=True)
    far = column_cw(3000, 1500, 1610, steps, fold=True)
    assert np.abs(far).max() > 0.1
    assert np.abs(near - far).max() < 2e-3 * np.abs(far).max()
    kept = column_cw(600, 300, 410, steps, fold=False)
    assert np.abs(kept - far).max() < 2e-3 * np.abs(far).max()


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
