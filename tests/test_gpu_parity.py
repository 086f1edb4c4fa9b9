"""Parity of the HIP engine (through the C ABI) with the FP64 CPU oracle.

FP64 engine runs must match the oracle to rounding (the engine uses the E-direct (P, Q) form
of the oracle's D/E split, so agreement is to ~1e-12 relative, not bitwise).  FP32 runs must
stay within 1e-4 relative L2 (the workload tolerance) on the same inputs.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from sim_juncs_amd import materials as M
from sim_juncs_amd import scenario
from sim_juncs_amd._lib import EngineError
from sim_juncs_amd.engine import Engine

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 1e-4


def early_pulse(scn, fwhm_fs=2.0, cep=0.3):
    """Reduced cases run a few hundred steps: use a short pulse that starts at t = 0."""
    for src in scn.sources:
        src.update(scenario.gaussian_pulse(0.8, fwhm_fs, cep, start_fs=0.0))
    return scn


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def run_pair(scn, precision=None, tile=None, steps=None):
    print(f"[case] {scn.name} shape={scn.shape} prec={precision or scn.precision} tile={tile}", flush=True)
    eng = scn.make(Engine, precision=precision or scn.precision)
    if tile:
        eng.set_tile(*tile)
    orc = scn.make(Oracle)
    n = steps if steps is not None else scn.steps
    eng.step(n)
    orc.step(n)
    return eng, orc


def field_errors(eng, orc):
    return [rel(eng.read_field(c), orc.read_field(c)) for c in range(6)]


def random_scenario(shape, seed, pml=None, periodic=(0, 0, 0), poles=True, nmat=4, steps=40):
    """Random materials (dispersive mix), random initial fields, optional PML and source."""
    rng = np.random.default_rng(seed)
    mats = rng.integers(0, nmat, size=tuple(shape) + (3,))
    lut = [M.VACUUM, M.MaterialSpec(2.13), M.AU.scaled(1.0) if poles else M.MaterialSpec(3.0),
           M.SIO2.scaled(1.0) if poles else M.MaterialSpec(1.7)][:nmat]
    dx = 0.02
    scn = scenario.Scenario(name=f"random{shape}", shape=tuple(shape), dx=dx, precision="fp64",
                            periodic=periodic, matw=scenario.pack(mats), lut=lut, steps=steps)
    if pml is not None:
        scn.pml = (pml, 1e-15, 2.0)
    zc = shape[2] // 2
    scn.sources = [scenario.plane_source(shape, dx, zc, scenario.gaussian_pulse(0.8, 2.0, 0.3, start_fs=0.0))]
    return scn, rng


@pytest.mark.parametrize("n", [24, 33])
def test_plain_box_fp64_matches_oracle(n):
    scn = early_pulse(scenario.plain_box(n=n, steps=120, precision="fp64", pml=4))
    eng, orc = run_pair(scn)
    errs = field_errors(eng, orc)
    assert max(errs) < FP64_TOL, errs
    assert np.abs(eng.read_field("Ex")).max() > 0
    m_e, m_o = eng.read_monitors(), orc.read_monitors()
    assert m_e.shape == m_o.shape == (6, 1)
    assert rel(m_e, m_o) < FP64_TOL
    d_e, d_o = eng.read_dft(0), orc.read_dft(0)
    assert rel(d_e, d_o) < FP64_TOL
    eng.close()


@pytest.mark.parametrize("through_source", [False, True])
def test_dft_region_at_and_away_from_the_source_plane_fp64(through_source):
    """Source injection and DFT accumulation share one launch only when no DFT region holds a
    source cell; a region through the source plane keeps them in two launches (the DFT must
    see E after the injection).  Both match the oracle."""
    scn = early_pulse(scenario.plain_box(n=24, steps=90, precision="fp64", pml=4))
    z = scn.meta["src_z"] if through_source else scn.meta["src_z"] + 3
    scn.dfts = list(scn.dfts) + [((0, 0, z), (24, 24, z + 1), ("Ex", "Ey", "Hx"), np.array([0.8, 1.3]))]
    eng, orc = run_pair(scn)
    assert max(field_errors(eng, orc)) < FP64_TOL
    for i in range(len(scn.dfts)):
        d_e, d_o = eng.read_dft(i), orc.read_dft(i)
        assert np.abs(d_o).max() > 0
        assert rel(d_e, d_o) < FP64_TOL, i
    eng.close()


@pytest.mark.parametrize("comp,width,amp,part", [
    ("Ex", 0.6, 1.0, 0), ("Ey", 0.0, 0.7 - 0.4j, 1), ("Hz", 0.6, 1.0, 0), ("Hy", 0.0, 0.5 + 0.5j, 1),
    ("Hx", 0.6, 0.3 + 0.9j, 0),
])
def test_continuous_and_magnetic_sources_fp64(comp, width, amp, part):
    """meep::continuous_src_time sources (reference src/disp.cpp:616-619, tests/test.geom:8 and
    tests/context_test.geom:3 declare CW sources, the latter on Hz) and magnetic-current sources
    on any H component, as a plane sheet and as a point source, in random dispersive media with
    PML: fields and Ex/Hz monitors equal the oracle's to FP64 rounding, for the real and the
    imaginary part of a complex amplitude."""
    shape = (14, 12, 20)
    scn, _ = random_scenario(shape, seed=7 + part, pml=[2, 2, 2, 2, 3, 3], steps=160)
    scn.part = part
    dx = scn.dx
    # width > 0: tanh turn-on and turn-off (end 5 fs = 1.5 time units) inside the 1.6-unit run
    pulse = scenario.cw_pulse(0.8, 0.0, 0.0, 1e6) if width == 0.0 else scenario.cw_pulse(0.8, width, 0.0, 5.0)
    scn.sources = [scenario.plane_source(shape, dx, 9, pulse, comp=comp, amp=amp),
                   dict(comp=comp, box=[5.3 * dx, 6.1 * dx, 12.4 * dx] * 2, amp=0.8 * amp, **pulse)]
    pts = np.array([[7.1 * dx, 6.3 * dx, 4.2 * dx], [3.3 * dx, 8.6 * dx, 15.5 * dx]])
    scn.monitors = [("Ex", pts), ("Hz", pts)]
    eng, orc = run_pair(scn)
    assert abs(eng.last_source_time() - orc.last_source_time()) == 0
    errs = field_errors(eng, orc)
    assert max(errs) < FP64_TOL, errs
    assert np.abs(orc.read_field(comp)).max() > 0
    m_e, m_o = eng.read_monitors(), orc.read_monitors()
    assert np.abs(m_o).max() > 0
    assert rel(m_e, m_o) < FP64_TOL
    eng.close()


@pytest.mark.parametrize("tile", [None, ("thread_rows", 2, 3), ("register_rows", 1, 4)])
def test_pml_without_lorentz_poles_fp64(tile):
    """sjf_set_pml_dispersion: random Au (Drude + 2 Lorentz) / SiO2 (Lorentz) / dielectric media
    filling the grid, PML on every face; inside the slabs the engine runs twin materials without
    the Lorentz poles (folded into eps_inf at f_ref, Au keeps eps_inf and its Drude pole) while the
    oracle skips those poles cell by cell.  Fields and monitors agree to FP64 rounding, and differ
    from the run that keeps the poles in the PML."""
    shape = (16, 14, 22)
    scn, _ = random_scenario(shape, seed=3, pml=[3, 2, 2, 3, 3, 4], steps=150)
    scn.pml_fold = 1.25
    pts = np.array([[7.1 * scn.dx, 6.3 * scn.dx, 4.2 * scn.dx], [1.2 * scn.dx, 1.6 * scn.dx, 20.5 * scn.dx]])
    scn.monitors = [("Ex", pts)]
    eng, orc = run_pair(scn, tile=tile)
    assert max(field_errors(eng, orc)) < FP64_TOL
    assert rel(eng.read_monitors(), orc.read_monitors()) < FP64_TOL
    scn.pml_fold = 0.0
    keep = scn.make(Engine)
    keep.step(scn.steps)
    assert rel(keep.read_field("Ex"), orc.read_field("Ex")) > 1e-6
    keep.close()
    eng.close()


@pytest.mark.parametrize("precision,tile,fold", [("fp64", None, 0.0), ("fp64", ("thread_rows", 2, 3), 1.25),
                                                  ("fp64", ("register_rows", 2, 4), 1.25), ("fp32", None, 1.25)])
def test_eps_maps_match_oracle(precision, tile, fold):
    """sjf_set_eps_maps: a per-cell eps_inf for each E component (the smoothed, fractional
    permittivity) over random dispersive materials, with a design box on top and PML (with and
    without the Lorentz fold, which adds to the map's value in the slabs): fields, monitors and a
    CW source's injection (its coefficient takes 1/eps from the map) against the oracle's own per-
    cell eps arrays.  Adjoint runs refuse a map."""
    shape = (14, 12, 20)
    scn, rng = random_scenario(shape, seed=17, pml=[2, 2, 3, 2, 3, 3], steps=120)
    scn.precision = precision
    scn.pml_fold = fold
    scn.eps_maps = rng.uniform(1.0, 6.0, size=(3,) + shape)
    lo, hi = (5, 4, 8), (9, 7, 12)
    scn.design = (lo, hi, rng.uniform(1.0, 4.0, size=(4, 3, 4)))
    pulse = scenario.cw_pulse(0.8, 0.6, 0.0, 5.0)
    scn.sources = [scenario.plane_source(shape, scn.dx, 10, pulse)]
    scn.monitors = [("Ex", np.array([[6.2 * scn.dx, 5.1 * scn.dx, 14.4 * scn.dx]]))]
    eng, orc = run_pair(scn, tile=tile)
    tol = FP64_TOL if precision == "fp64" else FP32_TOL
    assert max(field_errors(eng, orc)) < tol
    assert rel(eng.read_monitors(), orc.read_monitors()) < tol
    eng.close()
    e = scn.make(Engine)
    with pytest.raises(EngineError, match="eps maps"):
        e.adjoint_run(10, 0, (6, 5, 10), 48)
    e.set_eps_maps(None)
    e.close()


def test_plain_box_fp32_within_tolerance():
    scn = early_pulse(scenario.plain_box(n=32, steps=300, precision="fp32", pml=6))
    eng, orc = run_pair(scn)
    errs = field_errors(eng, orc)
    assert max(errs) < FP32_TOL, errs
    assert np.abs(eng.read_field("Ex")).max() > 0
    assert rel(eng.read_dft(0), orc.read_dft(0)) < FP32_TOL
    eng.close()


R, T = "register_rows", "thread_rows"


@pytest.mark.parametrize("shape,tile", [
    ((5, 7, 9), (R, 1, 1)), ((5, 7, 9), (R, 4, 2)), ((9, 10, 70), (R, 2, 3)), ((9, 10, 70), (R, 4, 64)),
    ((12, 3, 65), (R, 4, 5)), ((1, 1, 130), (R, 1, 1)), ((6, 13, 128), (R, 4, 4)),
    ((5, 7, 9), (T, 1, 1)), ((5, 7, 9), (T, 3, 2)), ((9, 10, 70), (T, 2, 3)), ((9, 10, 70), (T, 13, 64)),
    ((12, 3, 65), (T, 4, 5)), ((1, 1, 130), (T, 1, 1)), ((6, 13, 128), (T, 7, 4)), ((7, 9, 100), (T, 9, 3)),
])
def test_random_media_tiles_fp64(shape, tile):
    scn, rng = random_scenario(shape, seed=sum(shape), pml=[1, 2, 1, 0, 2, 3] if min(shape) > 5 else None)
    eng = scn.make(Engine)
    eng.set_tile(*tile)
    orc = scn.make(Oracle)
    for c in range(6):
        f = rng.standard_normal(shape)
        eng.write_field(c, f)
        orc.write_field(c, f)
    eng.step(scn.steps)
    orc.step(scn.steps)
    errs = field_errors(eng, orc)
    assert max(errs) < FP64_TOL, errs
    eng.close()


@pytest.mark.parametrize("tile", [("register_rows", 2, 3), ("thread_rows", 4, 3)])
@pytest.mark.parametrize("periodic", [(1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)])
def test_periodic_axes_fp64(periodic, tile):
    shape = (8, 6, 66)
    pml = [0 if periodic[d] else 2 for d in range(3) for _ in range(2)]
    scn, rng = random_scenario(shape, seed=7, pml=pml, periodic=periodic)
    eng = scn.make(Engine)
    eng.set_tile(*tile)
    orc = scn.make(Oracle)
    for c in range(6):
        f = rng.standard_normal(shape)
        eng.write_field(c, f)
        orc.write_field(c, f)
    eng.step(30)
    orc.step(30)
    errs = field_errors(eng, orc)
    assert max(errs) < FP64_TOL, errs
    eng.close()


@pytest.mark.parametrize("shape,periodic,tile", [
    ((1, 1, 66), (1, 1, 0), None), ((1, 1, 66), (1, 1, 0), ("register_rows", 4, 3)),
    ((1, 1, 66), (1, 1, 0), ("thread_rows", 5, 3)), ((2, 3, 40), (1, 1, 1), ("register_rows", 4, 5)),
    ((2, 3, 40), (1, 1, 1), ("thread_rows", 9, 4)), ((5, 2, 70), (0, 1, 0), ("thread_rows", 13, 2)),
])
def test_periodic_axes_shorter_than_the_tile_fp64(shape, periodic, tile):
    """A tile's rows, and the planes ahead of its march, reach more than one period past the
    edge of a periodic axis shorter than the tile (one-cell axes: 2-D and 1-D runs)."""
    pml = [0 if periodic[d] else 2 for d in range(3) for _ in range(2)]
    scn, rng = random_scenario(shape, seed=3, pml=pml, periodic=periodic)
    eng = scn.make(Engine)
    if tile:
        eng.set_tile(*tile)
    print(f"[case] periodic {shape} {periodic} {tile} {eng.launch_info()}", flush=True)
    orc = scn.make(Oracle)
    for c in range(6):
        f = rng.standard_normal(shape)
        eng.write_field(c, f)
        orc.write_field(c, f)
    eng.step(30)
    orc.step(30)
    errs = field_errors(eng, orc)
    assert max(errs) < FP64_TOL, errs
    eng.close()


def test_bowtie_small_fp32_and_fp64():
    scn = early_pulse(scenario.bowtie_junction(nx=64, ny=64, nz=64, steps=200, pml=6), fwhm_fs=1.5)
    for prec, tol in (("fp64", FP64_TOL), ("fp32", FP32_TOL)):
        eng, orc = run_pair(scn, precision=prec)
        errs = field_errors(eng, orc)
        assert max(errs) < tol, (prec, errs)
        assert np.abs(eng.read_field("Ex")).max() > 0
        for i in range(len(scn.dfts)):
            assert rel(eng.read_dft(i), orc.read_dft(i)) < tol
        eng.close()


def test_energy_invariant_closed_cavity_fp64():
    """Lossless, non-dispersive closed cavity: W = sum eps E^n.E^n + H^{n+1/2}.H^{n-1/2} is
    exactly conserved by the leapfrog Yee scheme."""
    rng = np.random.default_rng(11)
    shape = (10, 9, 70)
    mats = rng.integers(0, 3, size=shape + (3,))
    lut = [M.VACUUM, M.MaterialSpec(2.5), M.MaterialSpec(4.0)]
    eps = np.array([1.0, 2.5, 4.0])[mats]
    eng = Engine(shape, 0.05, precision="fp64")
    eng.set_materials(scenario.pack(mats), lut)
    for c in range(6):
        eng.write_field(c, rng.standard_normal(shape))
    W = []
    for _ in range(50):
        E = [eng.read_field(c) for c in range(3)]
        H0 = [eng.read_field(c) for c in range(3, 6)]
        eng.step(1)
        H1 = [eng.read_field(c) for c in range(3, 6)]
        W.append(sum((eps[..., c] * E[c] ** 2).sum() for c in range(3)) +
                 sum((H0[c] * H1[c]).sum() for c in range(3)))
    W = np.array(W)
    assert np.abs(W - W[0]).max() / W[0] < 1e-12
    eng.close()


def test_step_one_at_a_time_equals_batch():
    """sjf_step(1) n times is the drop-in for n calls of fields.step(): same state as one
    sjf_step(n), including monitor samples taken before each step with save_span."""
    scn = early_pulse(scenario.plain_box(n=20, steps=45, precision="fp32", pml=3))
    scn.save_span = 7
    a = scn.make(Engine)
    b = scn.make(Engine)
    a.step(45)
    for _ in range(45):
        b.step(1)
    assert np.abs(a.read_field("Ex")).max() > 0
    for c in range(6):
        assert np.array_equal(a.read_field(c), b.read_field(c))
    ma, mb = a.read_monitors(), b.read_monitors()
    assert ma.shape[0] == 7  # samples at steps 0,7,...,42
    assert np.array_equal(ma, mb)
    a.close()
    b.close()


@pytest.mark.parametrize("tile", [None, ("thread_rows", 1, 16)])
def test_field_set_beyond_4gib_matches_small_grid(tile):
    """Per-plane descriptors past 4 GiB: a 720x512x512 FP32 grid (4.5 GB per field set; the Au
    box's polarisation takes 9 planes per x plane, 6.8 GB) runs the same per-cell arithmetic as a
    60^3 grid.  The same point source inside the same Au box, near the far corner of the large
    grid (x plane 690: field and polarisation offsets above 4 GiB) and at the centre of the small
    one, gives bit-identical monitor series and fields around the source for 20 steps.  The
    disturbance travels at most one cell per step, so neither grid's boundary is reached.  The
    cell size is a power of two so that source and monitor positions convert to exactly the same
    fractional cell coordinates in both grids."""
    dx, steps = 1.0 / 64, 20
    lut = [M.VACUUM, M.AU.scaled(1.0)]
    pulse = scenario.gaussian_pulse(0.8, 0.2, 0.3, start_fs=0.0)  # peaks near step 15
    off = np.array([[a, b, c] for a in (-7, 0, 5) for b in (-6, 2) for c in (-4, 0, 3)], dtype=float)
    out = []
    for shape, c in (((60, 60, 60), (30, 30, 30)), ((720, 512, 512), (690, 470, 470))):
        print(f"[case] translation {shape} source {c} tile {tile}", flush=True)
        eng = Engine(shape, dx, precision="fp32")
        if tile:
            eng.set_tile(*tile)
        words = np.zeros(shape, np.uint32)
        words[c[0] - 5:c[0] + 6, c[1] - 5:c[1] + 6, c[2] - 5:c[2] + 6] = 0x00010101  # Au on all three E points
        eng.set_materials(words, lut)
        del words
        at = [c[0] * dx, c[1] * dx, (c[2] + 0.5) * dx]  # an Ez lattice point
        eng.add_source("Ez", at + at, **pulse)
        for comp, o in (("Ex", (0.5, 0, 0)), ("Ez", (0, 0, 0.5)), ("Hy", (0.5, 0, 0.5))):
            eng.add_monitors(comp, (np.array(c) + off + np.array(o)) * dx)
        eng.set_save_span(1)
        eng.step(steps)
        mon = eng.read_monitors()
        box = tuple(slice(v - 22, v + 23) for v in c)
        fields = [eng.read_field(comp)[box] for comp in ("Ex", "Ez", "Hy")]
        eng.close()
        out.append((mon, fields))
    (m0, f0), (m1, f1) = out
    assert np.abs(f0[1]).max() > 0 and np.abs(m0).max() > 0
    assert np.array_equal(m0, m1)
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_read_field_box_is_the_slice_of_the_whole_component(precision):
    """sjf_read_field_box (output_hdf5 over a sub-volume) returns exactly the box of the
    whole-component read-out, for interior, edge and whole-grid boxes; an empty box copies
    nothing and a box outside the grid is an argument error."""
    scn = early_pulse(scenario.plain_box(n=21, steps=30, precision=precision, pml=3))
    eng = scn.make(Engine)
    eng.step(30)
    for c in range(6):
        whole = eng.read_field(c)
        assert np.abs(whole).max() > 0 or c >= 3
        for lo, hi in [((0, 0, 0), (21, 21, 21)), ((3, 5, 7), (10, 6, 20)), ((20, 0, 0), (21, 21, 1)),
                       ((0, 17, 4), (21, 21, 5))]:
            box = eng.read_field(c, lo, hi)
            assert box.shape == tuple(h - l for l, h in zip(lo, hi))
            assert np.array_equal(box, whole[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]])
    assert eng.read_field("Ex", (4, 4, 4), (4, 9, 9)).size == 0
    for lo, hi in [((0, 0, 0), (22, 21, 21)), ((-1, 0, 0), (3, 3, 3)), ((5, 5, 5), (4, 6, 6))]:
        with pytest.raises(EngineError):
            eng.read_field("Ey", lo, hi)
    eng.close()


@pytest.mark.parametrize("precision,zseg,shape", [
    ("fp64", 64, (12, 10, 150)),   # three passes, the last 22 cells
    ("fp64", 128, (10, 9, 200)),
    ("fp32", 256, (8, 8, 512)),
    ("fp64", 0, (8, 7, 1100)),     # longer than a workgroup: default 128-cell passes
])
def test_z_passes_match_one_pass_and_oracle(precision, zseg, shape):
    """A workgroup that covers its z line in passes (sjf_set_zsegment; the last lane of a pass
    hands its H to the first lane of the next) steps exactly as one pass over the line, and
    z lines longer than 1024 cells match the oracle."""
    scn, _ = random_scenario(shape, seed=21, pml=[3] * 6, steps=40)
    print(f"[case] z passes {shape} zseg={zseg} {precision}", flush=True)
    eng = scn.make(Engine, precision=precision)
    eng.set_tile("register_rows", 1, 4)
    eng.set_zsegment(zseg)
    eng.step(40)
    info = eng.launch_info()
    assert info["threads"] == (zseg or 128)
    if shape[2] <= 1024:
        one = scn.make(Engine, precision=precision)
        one.set_tile("register_rows", 1, 4)
        one.set_zsegment(-(-shape[2] // 64) * 64)  # one pass over the line
        assert one.launch_info()["threads"] == -(-shape[2] // 64) * 64
        one.step(40)
        for c in range(6):
            assert np.array_equal(eng.read_field(c), one.read_field(c)), c
        one.close()
    if precision == "fp64":
        orc = scn.make(Oracle)
        orc.step(40)
        errs = field_errors(eng, orc)
        print("  ", ["%.2e" % e for e in errs], flush=True)
        assert max(errs) < FP64_TOL
    eng.close()


def test_z_pass_arguments_are_checked():
    scn, _ = random_scenario((6, 5, 130), seed=3, pml=[2] * 6, steps=4)
    eng = scn.make(Engine, precision="fp64")
    for bad in (32, 100, 2048, -64):
        with pytest.raises(EngineError):
            eng.set_zsegment(bad)
    eng.set_tile("register_rows", 2, 4)
    eng.set_zsegment(64)
    with pytest.raises(EngineError):  # passes run one row per thread
        eng.step(1)
    eng.close()
    # the default: 128-cell passes for a 512-cell line with one row per thread, one pass with two
    scn, _ = random_scenario((6, 5, 512), seed=3, pml=[2] * 6, steps=4)
    eng = scn.make(Engine, precision="fp32")
    eng.set_tile("register_rows", 1, 4)
    assert eng.launch_info()["threads"] == 128
    eng.set_tile("register_rows", 2, 4)
    assert eng.launch_info()["threads"] == 512
    eng.step(2)
    eng.close()
    scn, _ = random_scenario((6, 5, 130), seed=3, pml=None, periodic=(0, 0, 1), steps=4)
    eng = scn.make(Engine, precision="fp64")
    with pytest.raises(EngineError):  # a periodic z axis needs one pass
        eng.set_zsegment(64)
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_z_passes_with_design_region_and_eps_maps(precision):
    """Design voxels and per-cell eps maps that straddle pass boundaries (z 200..300 across the
    128-cell passes of a 512-cell line, and the 64-cell ones) step exactly as in one pass."""
    shape = (8, 8, 512)
    scn, rng = random_scenario(shape, seed=5, pml=[3] * 6, steps=30)
    lo, hi = (2, 2, 200), (6, 7, 300)
    scn.design = (lo, hi, rng.uniform(1.5, 4.0, size=tuple(h - l for l, h in zip(lo, hi))))
    scn.eps_maps = rng.uniform(1.0, 3.0, size=(3,) + shape)
    runs = []
    for zseg in (0, 512, 64):
        print(f"[case] z passes design+maps {precision} zseg={zseg}", flush=True)
        eng = scn.make(Engine, precision=precision)
        eng.set_tile("register_rows", 1, 4)
        eng.set_zsegment(zseg)
        assert eng.launch_info()["threads"] == (zseg or 128)
        eng.step(30)
        runs.append([eng.read_field(c) for c in range(6)])
        eng.close()
    assert np.abs(runs[0][0]).max() > 0
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert np.array_equal(a, b)


def test_batch_of_pass_and_one_pass_lines_equals_each_engine():
    """One fused launch over a 512-cell line stepped in passes, a 130-cell line and a periodic
    96-cell line (both one pass) gives each engine the bits of its own run."""
    from sim_juncs_amd.engine import Batch

    specs = [((6, 5, 512), (0, 0, 0)), ((7, 6, 130), (0, 0, 0)), ((6, 6, 96), (0, 0, 1))]
    scns = []
    for i, (shape, per) in enumerate(specs):
        scn, _ = random_scenario(shape, seed=30 + i, pml=None if per[2] else [2] * 6, periodic=per, steps=40)
        scns.append(scn)
    print("[case] batch of pass and one-pass lines", [s.shape for s in scns], flush=True)
    own = []
    for scn in scns:
        eng = scn.make(Engine, precision="fp64")
        eng.set_tile("register_rows", 1, 4)
        eng.step(40)
        own.append([eng.read_field(c) for c in range(6)])
        eng.close()
    engs = [scn.make(Engine, precision="fp64") for scn in scns]
    b = Batch(engs)
    b.set_tile("register_rows", 4)
    b.step(40)
    threads, _ = b.launch_info()
    assert threads == 192  # the 130-cell line in one pass; the 512-cell one in 128-cell passes
    for i, eng in enumerate(engs):
        for c in range(6):
            assert np.array_equal(eng.read_field(c), own[i][c]), (i, c)
    b.close()
    for eng in engs:
        eng.close()
