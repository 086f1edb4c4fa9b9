"""Adjoint gradients (sjf_adjoint_run) against central finite differences of the FP64 oracle.

J(eps) = dt * sum_{n=1..N} |E^n(gap)|^2; the engine returns dJ/d eps for every design voxel from
the discrete adjoint.  The cases separate the parts of the transposed step (Yee curls only, with
CPML, with Drude/Lorentz media, periodic axes) and the tiling of the fused reverse kernel
(several x planes per workgroup, a z line longer than one wavefront).  The oracle computes J independently (D/E split, its
own PML), so agreement pins both the adjoint algebra and the forward model it transposes.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from sim_juncs_amd import materials as M
from sim_juncs_amd import scenario
from sim_juncs_amd._lib import EngineError
from sim_juncs_amd.engine import Engine

pytestmark = pytest.mark.gpu


# kind -> (shape, periodic axes, PML cells, media, adjoint x-chunk override)
CASES = {
    "vacuum": ((14, 12, 20), (0, 0, 0), None, False, None),
    "pml": ((14, 12, 20), (0, 0, 0), [3] * 6, False, None),
    "poles": ((14, 12, 20), (0, 0, 0), [3] * 6, True, None),
    # periodic x and y, PML on the z faces only
    "periodic": ((14, 12, 20), (1, 1, 0), [0, 0, 0, 0, 3, 3], True, None),
    # several x planes per workgroup, ragged last chunk (14 = 4 + 4 + 4 + 2)
    "poles_xchunk4": ((14, 12, 20), (0, 0, 0), [3] * 6, True, 4),
    # z line of two wavefronts (70 cells in a 128-thread workgroup), chunks of 3 planes
    "poles_long_z": ((10, 8, 70), (0, 0, 0), [3] * 6, True, 3),
}


def design_case(kind, precision="fp64", steps=160, seed=4):
    rng = np.random.default_rng(seed)
    shape, periodic, cells, media, _ = CASES[kind]
    dx = 0.02
    mat = np.zeros(shape + (3,), np.uint8)
    lut = [M.VACUUM]
    pml = (cells, 1e-15, 2.0) if cells else None
    if media:
        lut = [M.VACUUM, M.AU.scaled(1.0), M.SIO2.scaled(1.0)]
        scenario.paint(mat, shape, scenario.box_inside((-1, -1, 12.5), (5.5, 13, 15.5)), 1)
        scenario.paint(mat, shape, scenario.box_inside((8.5, -1, 12.5), (15, 13, 15.5)), 1)
        scenario.paint(mat, shape, scenario.box_inside((-1, -1, 15.5), (15, 13, 21)), 2)
    lo, hi = (5, 4, 9), (9, 8, 12)
    eps = rng.uniform(1, 4, size=tuple(h - l for l, h in zip(lo, hi)))
    scn = scenario.Scenario(name=f"adj_{kind}", shape=shape, dx=dx, precision=precision,
                            periodic=periodic, matw=scenario.pack(mat), lut=lut, pml=pml, steps=steps,
                            design=(lo, hi, eps))
    scn.sources = [scenario.plane_source(shape, dx, 4, scenario.gaussian_pulse(0.8, 1.5, 0.4, start_fs=0.0))]
    gap = (7, 6, 13)
    return scn, gap, eps


def oracle_J(scn, gap, eps):
    s = scenario.Scenario(**{**scn.__dict__})
    s.design = (scn.design[0], scn.design[1], eps)
    dx = scn.dx
    o = s.make(Oracle)
    off = scenario.COMP_OFF
    pts = [[(gap[d] + off[c][d]) * dx for d in range(3)] for c in range(3)]
    for c in range(3):
        o.add_monitors(c, np.array([pts[c]]))
    o.set_save_span(1)
    o.step(scn.steps)
    smp = o.read_monitors().real  # E^0 .. E^{N-1}
    last = np.array([o.read_field(c)[gap] for c in range(3)])
    e = np.vstack([smp[1:], last[None, :]])
    return o.dt * float((e ** 2).sum())


@pytest.mark.parametrize("kind", list(CASES))
def test_adjoint_gradient_matches_oracle_finite_differences(kind):
    scn, gap, eps = design_case(kind)
    print(f"[case] adjoint {kind}", flush=True)
    eng = scn.make(Engine)
    if CASES[kind][4]:
        eng.set_tile("register_rows", 1, CASES[kind][4])
    J, g = eng.adjoint_run(scn.steps, 37, gap, eps.size)
    eng.close()
    J_o = oracle_J(scn, gap, eps)
    assert J > 0
    assert abs(J - J_o) / J_o < 1e-10
    rng = np.random.default_rng(1)
    flat = eps.reshape(-1)
    idx = rng.choice(flat.size, size=8, replace=False)
    for v in idx:
        h = 1e-4 * flat[v]
        ep, em = flat.copy(), flat.copy()
        ep[v] += h
        em[v] -= h
        fd = (oracle_J(scn, gap, ep.reshape(eps.shape)) - oracle_J(scn, gap, em.reshape(eps.shape))) / (2 * h)
        assert abs(g[v] - fd) <= 1e-6 * np.abs(g).max() + 1e-9 * abs(fd), (kind, v, g[v], fd)


# kind, owned rows of the row-of-threads reverse tile, x-chunk (0: default): rows that split
# the y extent unevenly (12 = 5 + 5 + 2, 11 + 1), wrap periodic y rows into the last tile,
# put two or three thread rows into one wavefront (nz = 20) or a z line across two (nz = 70)
ROW_TILES = [("vacuum", 3, 0), ("pml", 5, 4), ("poles", 3, 0), ("poles", 11, 5), ("periodic", 5, 3),
             ("periodic", 11, 0), ("poles_xchunk4", 2, 4), ("poles_long_z", 6, 3)]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kind,rows,xc", ROW_TILES)
def test_adjoint_row_of_threads_tile_matches_register_tile(kind, rows, xc):
    """The row-of-threads reverse step (y and z neighbours through LDS) performs the arithmetic
    of the one-row-per-workgroup tile: the same J and the same gradient to rounding, in FP64
    and FP32, and the launch shape it reports."""
    for precision, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        scn, gap, eps = design_case(kind, precision=precision)
        print(f"[case] adjoint rows {kind} rows={rows} xc={xc} {precision}", flush=True)
        out = []
        for form, r in (("register_rows", 1), ("thread_rows", rows)):
            eng = scn.make(Engine)
            eng.set_adjoint_tile(form, r, xc)
            out.append(eng.adjoint_run(scn.steps, 37, gap, eps.size))
            _, threads, groups = eng.adjoint_kernel_info()
            eng.close()
        nx, ny, nz = scn.shape
        nxc = -(-nx // xc) if xc else None
        assert threads == -(-((rows + 1) * nz) // 64) * 64
        if nxc:
            assert groups == -(-ny // rows) * nxc
        (J0, g0), (J1, g1) = out
        assert J1 == J0  # the forward pass is the same
        assert rel(g1, g0) < tol, (kind, rows, precision, rel(g1, g0))


def test_adjoint_row_of_threads_tile_matches_oracle_finite_differences():
    scn, gap, eps = design_case("poles")
    eng = scn.make(Engine)
    eng.set_adjoint_tile("thread_rows", 3, 4)
    J, g = eng.adjoint_run(scn.steps, 37, gap, eps.size)
    eng.close()
    rng = np.random.default_rng(2)
    flat = eps.reshape(-1)
    for v in rng.choice(flat.size, size=6, replace=False):
        h = 1e-4 * flat[v]
        ep, em = flat.copy(), flat.copy()
        ep[v] += h
        em[v] -= h
        fd = (oracle_J(scn, gap, ep.reshape(eps.shape)) - oracle_J(scn, gap, em.reshape(eps.shape))) / (2 * h)
        assert abs(g[v] - fd) <= 1e-6 * np.abs(g).max() + 1e-9 * abs(fd), (v, g[v], fd)


def test_adjoint_tile_arguments_are_checked():
    scn, gap, eps = design_case("poles_long_z", precision="fp64", steps=10)
    eng = scn.make(Engine)
    with pytest.raises(EngineError):
        eng.set_adjoint_tile("register_rows", 2, 0)
    with pytest.raises(EngineError):
        eng.set_adjoint_tile("thread_rows", 14, 0)  # 15 x 70 threads
    with pytest.raises(EngineError):
        eng.set_adjoint_tile("thread_rows", 7, 0)  # FP64 over 512 threads
    eng.set_adjoint_tile("thread_rows", 6, 0)
    eng.close()


def test_adjoint_independent_of_checkpoint_interval_and_leaves_final_state():
    """Checkpoint intervals 1, 13, 90 and the record mode (0) give the same bits."""
    scn, gap, eps = design_case("poles", steps=90)
    res = []
    for K in (1, 13, 90, 0):
        eng = scn.make(Engine)
        J, g = eng.adjoint_run(scn.steps, K, gap, eps.size)
        assert eng.step_count() == scn.steps
        res.append((J, g, eng.read_field("Ex")))
        eng.close()
    ref = scn.make(Engine)
    ref.step(scn.steps)
    for J, g, ex in res:
        assert J == res[0][0]
        assert np.array_equal(g, res[0][1])
        assert np.array_equal(ex, ref.read_field("Ex"))
    ref.close()


def test_adjoint_kernel_info_reports_the_reverse_launch():
    """The reverse kernel's timing and launch shape: absent before an adjoint run, then one y row
    per workgroup over the x-chunks sjf_set_adjoint_tile gives, else its own default (about
    16384 workgroups), whatever the forward tile."""
    scn, gap, eps = design_case("poles", precision="fp32", steps=30)
    eng = scn.make(Engine)
    eng.set_adjoint_tile("register_rows", 1, 4)
    with pytest.raises(EngineError):
        eng.adjoint_kernel_info()
    eng.adjoint_run(scn.steps, 0, gap, eps.size)
    ms, threads, groups = eng.adjoint_kernel_info()
    nx, ny, nz = scn.shape
    assert 0 < ms < 100
    assert threads == -(-nz // 64) * 64
    assert groups == ny * -(-nx // 4)
    eng.close()
    nxc = min(-(-16384 // ny), nx)
    for tile in (("register_rows", 1, 4), ("thread_rows", 1, 4)):
        eng = scn.make(Engine)
        eng.set_tile(*tile)
        eng.adjoint_run(scn.steps, 0, gap, eps.size)
        _, _, groups = eng.adjoint_kernel_info()
        assert groups == ny * -(-nx // -(-nx // nxc)), tile
        eng.close()


def test_adjoint_fp32_close_to_fp64():
    scn, gap, eps = design_case("poles")
    e64 = scn.make(Engine, precision="fp64")
    J64, g64 = e64.adjoint_run(scn.steps, 40, gap, eps.size)
    e32 = scn.make(Engine, precision="fp32")
    J32, g32 = e32.adjoint_run(scn.steps, 40, gap, eps.size)
    e64.close()
    e32.close()
    assert abs(J32 - J64) / J64 < 1e-4
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-4


def test_torch_autograd_binding():
    import torch

    from sim_juncs_amd.adjoint import GapFieldObjective, run_adjoint

    scn, gap, eps = design_case("poles", precision="fp32", steps=80)
    t = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    J = GapFieldObjective.apply(t, scn, scn.steps, 20, gap)
    J.backward()
    J_ref, g_ref = run_adjoint(scn, eps, scn.steps, 20, gap)
    assert abs(J.item() - J_ref) <= 1e-12 * abs(J_ref)
    assert np.allclose(t.grad.numpy(), g_ref, rtol=1e-12, atol=0)


def test_batched_torch_autograd_binding():
    """BatchedGapFieldObjective: J of every variant and eps.grad under a weighted sum equal the
    single-variant binding's, bit for bit."""
    import torch

    from sim_juncs_amd.adjoint import BatchedGapFieldObjective, run_adjoint

    members = [((14, 12, 20), (0, 0, 0), (5, 4, 9), (9, 8, 12), (7, 6, 13)),
               ((18, 10, 20), (0, 0, 0), (6, 3, 8), (10, 7, 11), (9, 5, 12)),
               ((13, 16, 20), (1, 1, 0), (3, 5, 9), (7, 9, 12), (6, 9, 13))]
    cases = [batch_member(sh, per, lo, hi, gap, i, "fp32", 50) for i, (sh, per, lo, hi, gap) in enumerate(members)]
    eps = torch.tensor(np.stack([e for _, _, e in cases]), dtype=torch.float64, requires_grad=True)
    J = BatchedGapFieldObjective.apply(eps, [c[0] for c in cases], 50, [c[1] for c in cases])
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    (w * J).sum().backward()
    for i, (scn, gap, e) in enumerate(cases):
        J_ref, g_ref = run_adjoint(scn, e, 50, 0, gap)
        assert J[i].item() == J_ref and J_ref > 0
        assert np.array_equal(eps.grad[i].numpy(), w[i].item() * g_ref)


def test_adjoint_beyond_4gib_matches_small_grid():
    """The reverse step's per-plane descriptors past 4 GiB: an adjoint run on a 720x512x512 FP32
    grid (adjoint field sets 4.5 GB each, pole bars 9 planes per x plane) with the design box,
    Au box, source and objective cell near the far corner gives the same J and gradient, bit for
    bit, as the same arrangement at the centre of a 60^3 grid (20 steps: nothing reaches a
    boundary).  The cell size is a power of two so that positions convert exactly."""
    dx, steps = 1.0 / 64, 20
    lut = [M.VACUUM, M.AU.scaled(1.0)]
    pulse = scenario.gaussian_pulse(0.8, 0.2, 0.3, start_fs=0.0)
    rng = np.random.default_rng(3)
    eps = rng.uniform(1, 4, size=(4, 4, 3))
    res = []
    for shape, c in (((60, 60, 60), (30, 30, 30)), ((720, 512, 512), (690, 470, 470))):
        print(f"[case] adjoint translation {shape}", flush=True)
        eng = Engine(shape, dx, precision="fp32")
        words = np.zeros(shape, np.uint32)
        words[c[0] - 5:c[0] + 6, c[1] - 5:c[1] - 1, c[2] - 5:c[2] + 6] = 0x00010101
        eng.set_materials(words, lut)
        del words
        lo = (c[0] - 2, c[1] + 1, c[2] - 1)
        eng.set_design_eps(lo, tuple(l + n for l, n in zip(lo, eps.shape)), eps)
        at = [c[0] * dx, c[1] * dx, (c[2] + 0.5) * dx]
        eng.add_source("Ez", at + at, **pulse)
        J, g = eng.adjoint_run(steps, 7, (c[0], c[1] + 2, c[2]), eps.size)
        eng.close()
        res.append((J, g))
    assert res[0][0] > 0 and np.abs(res[0][1]).max() > 0
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])


def batch_member(shape, periodic, lo, hi, gap, seed, precision, steps):
    """An adjoint case of its own shape, design box and objective cell (Au electrodes and a SiO2
    slab as in design_case, so every member runs the same pole-slot variant)."""
    rng = np.random.default_rng(seed)
    dx = 0.02
    mat = np.zeros(tuple(shape) + (3,), np.uint8)
    lut = [M.VACUUM, M.AU.scaled(1.0), M.SIO2.scaled(1.0)]
    scenario.paint(mat, shape, scenario.box_inside((-1, -1, 12.5), (5.5, 40, 15.5)), 1)
    scenario.paint(mat, shape, scenario.box_inside((shape[0] - 5.5, -1, 12.5), (40, 40, 15.5)), 1)
    scenario.paint(mat, shape, scenario.box_inside((-1, -1, 15.5), (40, 40, 21)), 2)
    pml_cells = [0, 0, 0, 0, 3, 3] if any(periodic) else [3] * 6
    eps = rng.uniform(1, 4, size=tuple(h - l for l, h in zip(lo, hi)))
    scn = scenario.Scenario(name=f"batch_adj_{seed}", shape=tuple(shape), dx=dx, precision=precision,
                            periodic=periodic, matw=scenario.pack(mat), lut=lut, pml=(pml_cells, 1e-15, 2.0),
                            steps=steps, design=(lo, hi, eps))
    scn.sources = [scenario.plane_source(tuple(shape), dx, 4,
                                         scenario.gaussian_pulse(0.8, 1.5, 0.4 + 0.1 * seed, start_fs=0.0))]
    return scn, gap, eps


BATCH_MEMBERS = [
    # shape, periodic, design lo, hi, objective cell, x-chunk of the register-rows mode
    ((14, 12, 20), (0, 0, 0), (5, 4, 9), (9, 8, 12), (7, 6, 13), 5),
    ((18, 10, 20), (0, 0, 0), (6, 3, 8), (11, 6, 13), (9, 5, 12), 4),
    ((13, 16, 20), (1, 1, 0), (3, 5, 10), (7, 12, 12), (6, 9, 13), 3),
]


@pytest.mark.parametrize("mode", ["register_rows", "default", "checkpoints", "reverse_rows"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_adjoint_equals_each_engines_own_run(precision, mode):
    """sjf_batch_adjoint_run over engines of different shapes, design boxes, objective cells and
    periodic axes gives every engine the J and gradient of its own sjf_adjoint_run, bit for bit
    (the per-engine run is pinned to finite differences of the oracle above), and leaves each
    engine in its final forward state.  The batch's forward launch runs one tile form for all
    engines, and the forms round differently in the last bit, so each mode runs the same form
    alone and in the batch: register rows with each engine's x-chunk, or the default
    row-of-threads form, also with the batch checkpointing every 13 steps and recomputing (the
    engines alone record).  The reverse launch takes x-chunks common to all engines, or, in the
    last mode, the row-of-threads reverse tile of 4 rows and 6-plane chunks set on every engine."""
    from sim_juncs_amd.engine import Batch

    steps = 60
    cases = [batch_member(sh, per, lo, hi, gap, i, precision, steps)
             for i, (sh, per, lo, hi, gap, _) in enumerate(BATCH_MEMBERS)]

    def make(scn, m):
        eng = scn.make(Engine)
        if mode == "register_rows":
            eng.set_tile("register_rows", 1, m[5])
        if mode == "reverse_rows":
            eng.set_adjoint_tile("thread_rows", 4, 6)
        return eng

    own = []
    for (scn, gap, eps), m in zip(cases, BATCH_MEMBERS):
        print(f"[case] own adjoint {precision} {mode} {scn.shape}", flush=True)
        eng = make(scn, m)
        J, g = eng.adjoint_run(steps, 0, gap, eps.size)
        own.append((J, g, eng.read_field("Ey")))
        eng.close()
    print(f"[case] batched adjoint {precision} {mode}", flush=True)
    engs = [make(scn, m) for (scn, _, _), m in zip(cases, BATCH_MEMBERS)]
    b = Batch(engs)
    if mode == "register_rows":
        b.set_tile("register_rows", 0)
    with pytest.raises(EngineError):
        b.adjoint_kernel_info()
    J, grads = b.adjoint_run(steps, [gap for _, gap, _ in cases], [eps.size for _, _, eps in cases],
                             13 if mode == "checkpoints" else 0)
    ms, threads, groups = b.adjoint_kernel_info()
    b.close()
    assert 0 < ms < 100 and threads == (128 if mode == "reverse_rows" else 64)
    tiles = 0
    rows = sum(scn.shape[1] for scn, _, _ in cases)
    for (scn, _, _), m in zip(cases, BATCH_MEMBERS):
        nx, ny, _ = scn.shape
        if mode == "reverse_rows":  # 4-row tiles of (4 + 1) x 20 threads
            tiles += -(-ny // 4) * -(-nx // 6)
            continue
        # x-chunks common to all engines, for about 32768 workgroups in all
        nxc = min(-(-32768 // rows), nx)
        xc = -(-nx // nxc)
        tiles += ny * -(-nx // xc)
    assert groups == tiles
    for i, eng in enumerate(engs):
        assert J[i] > 0 and J[i] == own[i][0], i
        assert np.array_equal(grads[i], own[i][1]), i
        assert eng.step_count() == steps
        assert np.array_equal(eng.read_field("Ey"), own[i][2]), i
        eng.close()


def test_batched_adjoint_rejects_members_it_cannot_run_together():
    from sim_juncs_amd.engine import Batch

    a, gap_a, eps_a = batch_member((14, 12, 20), (0, 0, 0), (5, 4, 9), (9, 8, 12), (7, 6, 13), 0, "fp32", 10)
    c, gap_c, eps_c = batch_member((14, 12, 24), (0, 0, 0), (5, 4, 9), (9, 8, 12), (7, 6, 13), 1, "fp32", 10)
    engs = [a.make(Engine), c.make(Engine)]
    b = Batch(engs)
    with pytest.raises(EngineError, match="nz"):
        b.adjoint_run(10, [gap_a, gap_c], [eps_a.size, eps_c.size])
    b.close()
    for e in engs:
        e.close()
