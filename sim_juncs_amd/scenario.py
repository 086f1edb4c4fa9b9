"""Workload descriptions: one `Scenario` drives the HIP engine or the CPU oracle alike.

A scenario is what bound_geom assembles before its time loop (reference src/disp.cpp:482-645):
grid, per-Yee-point material map, PML, sources, monitors (plus the DFT regions this engine
adds).  Builders below produce the sim_juncs junction workloads used for parity and
measurement, with seeded synthetic geometry.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import materials as M

LIGHT_SPEED = 0.299792458  # um per fs; t_meep = t_fs * LIGHT_SPEED * um_scale (src/disp.hpp:157)
FWHM_TO_WIDTH = 1.0 / (2.0 * math.sqrt(math.log(2.0)))  # intensity FWHM -> field sigma

# Yee offsets (in cells) of Ex, Ey, Ez, Hx, Hy, Hz
COMP_OFF = np.array([[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])


@dataclass
class Scenario:
    name: str
    shape: tuple
    dx: float
    courant: float = 0.5
    precision: str = "fp32"
    periodic: tuple = (0, 0, 0)
    part: int = 0
    matw: np.ndarray | None = None
    lut: list = field(default_factory=lambda: [M.VACUUM])
    eps_maps: np.ndarray | None = None  # per-cell eps_inf [3, nx, ny, nz] (sjf_set_eps_maps)
    design: tuple | None = None          # (lo, hi, eps array)
    pml: tuple | None = None             # (cells[6], R, power)
    pml_fold: float = 0.0                # f_ref of sjf_set_pml_dispersion (0: Lorentz poles kept in the PML)
    sources: list = field(default_factory=list)
    monitors: list = field(default_factory=list)  # (comp, xyz[n,3])
    save_span: int = 1
    dfts: list = field(default_factory=list)      # (lo, hi, comps, freqs)
    steps: int = 100
    meta: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))

    def make(self, cls, **kw):
        """Instantiate `cls` (Engine or Oracle) and apply the whole scenario to it."""
        sim = cls(self.shape, self.dx, courant=self.courant, precision=kw.pop("precision", self.precision),
                  periodic=self.periodic, part=self.part, **kw)
        self.apply(sim)
        return sim

    def apply(self, sim):
        if self.matw is not None:
            sim.set_materials(self.matw, self.lut)
        if self.eps_maps is not None:
            sim.set_eps_maps(self.eps_maps)
        if self.design is not None:
            sim.set_design_eps(*self.design)
        if self.pml is not None:
            sim.set_pml(*self.pml)
            if self.pml_fold > 0:
                sim.set_pml_dispersion(self.pml_fold)
        for s in self.sources:
            sim.add_source(**s)
        for comp, xyz in self.monitors:
            sim.add_monitors(comp, xyz)
        sim.set_save_span(self.save_span)
        ids = [sim.add_dft(*d) for d in self.dfts]
        return ids


# ----------------------------------------------------------------------- voxelization ----

def yee_points(shape, comp):
    """Cell-unit coordinates of component `comp`'s Yee points as broadcastable axes
    (nx,1,1), (1,ny,1), (1,1,nz)."""
    nx, ny, nz = shape
    off = COMP_OFF[comp]
    return ((np.arange(nx) + off[0])[:, None, None], (np.arange(ny) + off[1])[None, :, None],
            (np.arange(nz) + off[2])[None, None, :])


def paint(mat: np.ndarray, shape, inside, mid: int):
    """Set material id `mid` on every E Yee point for which inside(x, y, z) (cell units) holds;
    mat has shape (nx, ny, nz, 3)."""
    for c in range(3):
        x, y, z = yee_points(shape, c)
        m = np.broadcast_to(inside(x, y, z), shape)
        mat[..., c][m] = mid


def box_inside(lo, hi):
    """Inclusive box test, as box::in (reference src/cgs.cpp:52-67)."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    return lambda x, y, z: ((x >= lo[0]) & (x <= hi[0]) & (y >= lo[1]) & (y <= hi[1]) &
                            (z >= lo[2]) & (z <= hi[2]))


def triangle_prism_inside(tip, direction, side, z0, z1):
    """Equilateral triangle in x-y with its tip at `tip`, opening along `direction` (+1/-1 in
    x), extruded over z in [z0, z1]."""
    tx, ty = tip
    h = side * math.sqrt(3) / 2

    def f(x, y, z):
        u = (x - tx) * direction  # distance from the tip along the axis
        return (u >= 0) & (u <= h) & (np.abs(y - ty) <= u / math.sqrt(3)) & (z >= z0) & (z <= z1)

    return f


def pack(mat: np.ndarray) -> np.ndarray:
    m = mat.astype(np.uint32)
    return m[..., 0] | (m[..., 1] << 8) | (m[..., 2] << 16)


# ----------------------------------------------------------------------- sources ---------

def gaussian_pulse(wavelength_um, fwhm_fs, cep, um_scale=1.0, start_fs=5.0):
    """Gaussian_source semantics (reference src/cgs_data.cpp:22-53, src/disp.cpp:362-374,
    604-613): width = field sigma in fs, start_time defaults to 5 fs, cutoff forced to 6 so
    end = start + 12 width; converted to meep units with c_a = 0.299792458 um_scale."""
    ca = LIGHT_SPEED * um_scale
    w_fs = fwhm_fs * FWHM_TO_WIDTH
    end_fs = start_fs + 2 * 6 * w_fs
    return dict(kind="gaussian", freq=1.0 / (wavelength_um * um_scale), width=w_fs * ca,
                phase=float(cep), t_start=start_fs * ca, t_end=end_fs * ca)


def cw_pulse(wavelength_um, width_fs, start_fs, end_fs, um_scale=1.0, slowness=3.0):
    """CW_source semantics (reference src/cgs_data.cpp:55-83, src/disp.cpp:318-376,604-619):
    meep::continuous_src_time(frequency, width, start_time, end_time) with meep's default
    slowness 3; times converted to meep units with c_a = 0.299792458 um_scale."""
    ca = LIGHT_SPEED * um_scale
    return dict(kind="continuous", freq=1.0 / (wavelength_um * um_scale), width=width_fs * ca,
                t_start=start_fs * ca, t_end=end_fs * ca, slowness=float(slowness))


def plane_source(shape, dx, z_cell, pulse, comp="Ex", amp=1.0):
    nx, ny, _ = shape
    box = [0.0, 0.0, z_cell * dx, nx * dx, ny * dx, z_cell * dx]
    return dict(comp=comp, box=box, amp=amp, **pulse)


# ----------------------------------------------------------------------- workloads -------

def plain_box(n=96, steps=2000, precision="fp64", seed=1, pml=8):
    """Plain 3-D run: n^3 cells (96^3 full size), `pml`-cell PML on all faces, Courant 0.5,
    Au box (Drude + 2 Lorentz) 24x24x8 cells on a non-dispersive SiO2 slab (eps 2.13) spanning
    x-y and 20 cells of z, x-polarised plane-wave Gaussian pulse at 800 nm, FWHM 10 fs, CEP 0,
    one DFT plane at 5 frequencies.  Box origin and edges jitter by +-1 cell (seeded).
    Sizes scale with n/96 for reduced parity cases."""
    rng = np.random.default_rng(seed)
    s = n / 96.0
    shape = (n, n, n)
    dx = 0.02  # um_scale = 1: 20 nm cells, lambda0 / dx = 40
    mat = np.zeros(shape + (3,), np.uint8)
    slab_lo, slab_hi = pml, pml + max(2, round(20 * s))
    paint(mat, shape, box_inside((-1, -1, slab_lo), (n + 1, n + 1, slab_hi - 0.5)), 1)
    bw, bt = max(2, round(24 * s)), max(2, round(8 * s))
    j = rng.integers(-1, 2, size=6)
    cx = n // 2
    x0 = cx - bw // 2 + j[0]
    x1 = cx + bw // 2 + j[1]
    y0 = cx - bw // 2 + j[2]
    y1 = cx + bw // 2 + j[3]
    z0 = slab_hi
    z1 = slab_hi + bt + j[4]
    paint(mat, shape, box_inside((x0, y0, z0 - 0.5), (x1, y1, z1 - 0.5)), 2)
    lut = [M.VACUUM, M.SIO2_CONST, M.AU.scaled(1.0)]
    src_z = min(z1 + max(2, round(10 * s)), n - pml - 2)
    pulse = gaussian_pulse(0.8, 10.0, 0.0)
    f0 = pulse["freq"]
    dft_z = min(z1 + 2, n - pml - 1)
    freqs = [f0 * r for r in (0.8, 0.9, 1.0, 1.1, 1.2)]
    return Scenario(
        name=f"plain_box_{n}", shape=shape, dx=dx, precision=precision, matw=pack(mat), lut=lut,
        pml=([pml] * 6, 1e-15, 2.0), sources=[plane_source(shape, dx, src_z, pulse)],
        dfts=[((0, 0, dft_z), (n, n, dft_z + 1), ("Ex", "Ey", "Hx", "Hy"), freqs)],
        monitors=[("Ex", np.array([[(cx + 0.5) * dx, cx * dx, (z1 + 1) * dx]]))], save_span=20,
        steps=steps, meta=dict(src_z=src_z, dft_z=dft_z, au=(x0, x1, y0, y1, z0, z1)))


def bowtie_junction(nx=512, ny=512, nz=256, steps=6000, precision="fp32", pml=12):
    """Single large instance: nx x ny x nz cells (512x512x256 full size) of 5 nm, `pml`-cell
    PML, SiO2 substrate (1-pole Lorentz) over z in [pml, pml+64), Au bowtie (Drude + 2
    Lorentz: two equilateral triangles, 160-cell side, 16 cells thick, tips 8 cells apart) on
    the substrate, a 1-cell graphene Drude sheet under the gap, Gaussian pulse 800 nm FWHM 6
    fs, four point DFT monitors x 8 frequencies.  Lateral sizes scale with nx/512."""
    s = nx / 512.0
    shape = (nx, ny, nz)
    dx = 0.005
    mat = np.zeros(shape + (3,), np.uint8)
    sub_hi = pml + max(4, round(64 * s))
    paint(mat, shape, box_inside((-1, -1, pml - 0.5), (nx + 1, ny + 1, sub_hi - 0.5)), 1)
    side = max(8, round(160 * s))
    gap = max(2, round(8 * s))
    thick = max(2, round(16 * s))
    cx, cy = nx / 2, ny / 2
    zt0, zt1 = sub_hi - 0.5, sub_hi + thick - 0.5
    gr_half = max(2, round(40 * s))
    paint(mat, shape, box_inside((cx - gr_half, cy - gr_half, sub_hi - 0.5), (cx + gr_half, cy + gr_half, sub_hi + 0.5)), 3)
    paint(mat, shape, triangle_prism_inside((cx + gap / 2, cy), +1, side, zt0, zt1), 2)
    paint(mat, shape, triangle_prism_inside((cx - gap / 2, cy), -1, side, zt0, zt1), 2)
    lut = [M.VACUUM, M.SIO2.scaled(1.0), M.AU.scaled(1.0), M.graphene_sheet(dx).scaled(1.0)]
    src_z = min(sub_hi + thick + max(4, round(40 * s)), nz - pml - 2)
    pulse = gaussian_pulse(0.8, 6.0, 0.0)
    f0 = pulse["freq"]
    freqs = list(np.linspace(0.8, 1.2, 8) * f0)
    gz = sub_hi + thick // 2
    d20 = max(1, round(20 * s))
    pts = [(int(cx), int(cy), gz), (int(cx) + d20, int(cy), gz), (int(cx) - d20, int(cy), gz),
           (int(cx), int(cy), min(gz + max(1, round(10 * s)), nz - pml - 1))]
    dfts = [((p[0], p[1], p[2]), (p[0] + 1, p[1] + 1, p[2] + 1), ("Ex", "Ey", "Ez"), freqs) for p in pts]
    return Scenario(
        name=f"bowtie_{nx}x{ny}x{nz}", shape=shape, dx=dx, precision=precision, matw=pack(mat),
        lut=lut, pml=([pml] * 6, 1e-15, 2.0), pml_fold=f0, sources=[plane_source(shape, dx, src_z, pulse)],
        dfts=dfts, monitors=[("Ex", np.array([[(cx + 0.5) * dx, cy * dx, gz * dx]]))],
        save_span=20, steps=steps, meta=dict(src_z=src_z, points=pts))


def gap_junction(shape, gap, kind="box", dx=0.005, pml=12, fwhm_fs=10.0, cep=0.0,
                 steps=3000, precision="fp32", thick=None, nfreq=8, name=None):
    """One junction instance of the batched sweep.  kind "box": two Au (Drude + 2 Lorentz)
    electrodes of the junctions/Au_SiO2_box layout -- slabs filling x outside a gap of `gap`
    cells, SiO2 (1-pole Lorentz) filling the gap and the substrate below; kind "sidetie": the
    junctions/Au_SiO2_sidetie layout -- 45-degree Au wedges in the x-z plane pointing at each
    other across the gap, on a SiO2 substrate.  x-polarised Gaussian plane wave from the top,
    one DFT point at the gap centre."""
    nx, ny, nz = shape
    mat = np.zeros(tuple(shape) + (3,), np.uint8)
    cx, cy = nx / 2, ny / 2
    zc = nz / 2
    h = thick if thick is not None else max(4, nz // 8)
    zt, zb = zc - h / 2, zc + h / 2
    xl, xr = cx - gap / 2, cx + gap / 2
    # substrate below the junction layer (to the far edge, as the reference's inf_thick)
    paint(mat, shape, box_inside((-1, -1, zb), (nx + 1, ny + 1, nz + 1)), 1)
    if kind == "box":
        paint(mat, shape, box_inside((xl, -1, zt), (xr, ny + 1, zb)), 1)
        paint(mat, shape, box_inside((-1, -1, zt), (xl, ny + 1, zb)), 2)
        paint(mat, shape, box_inside((xr, -1, zt), (nx + 1, ny + 1, zb)), 2)
    else:
        def wedge(x, y, z):
            u = np.maximum(xl - x, x - xr)  # distance behind the tip plane of either wedge
            return (u >= 0) & (np.abs(z - zc) <= np.minimum(u, h / 2 + 0 * u)) & (y > -1)
        paint(mat, shape, wedge, 2)
    lut = [M.VACUUM, M.SIO2.scaled(1.0), M.AU.scaled(1.0)]
    pulse = gaussian_pulse(0.8, fwhm_fs, cep)
    src_z = pml + 2
    f0 = pulse["freq"]
    p = (int(cx), int(cy), int(zc))
    freqs = list(np.linspace(0.8, 1.2, nfreq) * f0)
    return Scenario(
        name=name or f"{kind}_{nx}x{ny}x{nz}_g{gap}", shape=tuple(shape), dx=dx,
        precision=precision, matw=pack(mat), lut=lut, pml=([pml] * 6, 1e-15, 2.0), pml_fold=f0,
        sources=[plane_source(shape, dx, src_z, pulse)],
        dfts=[(p, (p[0] + 1, p[1] + 1, p[2] + 1), ("Ex", "Ey", "Ez"), freqs)],
        monitors=[("Ex", np.array([[(p[0] + 0.5) * dx, p[1] * dx, p[2] * dx]]))], save_span=20,
        steps=steps, meta=dict(kind=kind, gap=gap, cep=cep, fwhm=fwhm_fs))


def batched_sweep(n=64, seed=3, steps=3000, precision="fp32", scale=1.0):
    """Batched sweep: `n` instances with grid sizes drawn from the seed between 64x64x64 and
    192x192x128 (times `scale` for reduced cases), box or side-tie Au/SiO2 junctions with gap
    widths uniform in 10-40 cells, CEP uniform in [0, 2 pi), pulse FWHM uniform in 5-15 fs,
    12-cell PML (8 below 96 cells), per-instance DFT point at the gap centre."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx, ny = (int(v) for v in rng.integers(64, 193, size=2))
        nz = int(rng.integers(64, 129))
        kind = "box" if rng.random() < 0.5 else "sidetie"
        gap = int(rng.integers(10, 41))
        cep = float(rng.uniform(0, 2 * math.pi))
        fwhm = float(rng.uniform(5, 15))
        if scale != 1.0:
            nx, ny, nz = (max(16, int(round(v * scale))) for v in (nx, ny, nz))
            gap = max(2, int(round(gap * scale)))
        pml = 12 if min(nx, ny, nz) >= 96 else 8
        if scale != 1.0:
            pml = max(3, int(round(pml * scale)))
        gap = min(gap, nx - 2 * pml - 4)
        out.append(gap_junction((nx, ny, nz), gap, kind, pml=pml, fwhm_fs=fwhm, cep=cep,
                                steps=steps, precision=precision, name=f"sweep{i:02d}_{kind}"))
    return out


def adjoint_junction(shape=(256, 256, 128), design=(64, 64, 16), steps=4000, seed=4,
                     precision="fp32", pml=12, dx=0.005, fwhm_fs=10.0, name=None):
    """Adjoint workload: a design box of per-voxel permittivity, uniform in [1, 4] (seeded),
    between two Au (Drude + 2 Lorentz) electrodes of the same thickness on a SiO2 substrate;
    objective J = time-integrated |E|^2 at the centre of the design box.  Full size:
    256x256x128 cells, 64x64x16 design voxels, 4000 steps."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    mat = np.zeros(tuple(shape) + (3,), np.uint8)
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    lo = (cx - design[0] // 2, cy - design[1] // 2, cz - design[2] // 2)
    hi = (lo[0] + design[0], lo[1] + design[1], lo[2] + design[2])
    paint(mat, shape, box_inside((-1, -1, hi[2] - 0.5), (nx + 1, ny + 1, nz + 1)), 1)
    paint(mat, shape, box_inside((-1, -1, lo[2] - 0.5), (lo[0] - 0.5, ny + 1, hi[2] - 0.5)), 2)
    paint(mat, shape, box_inside((hi[0] - 0.5, -1, lo[2] - 0.5), (nx + 1, ny + 1, hi[2] - 0.5)), 2)
    eps = rng.uniform(1.0, 4.0, size=design)
    lut = [M.VACUUM, M.SIO2.scaled(1.0), M.AU.scaled(1.0)]
    pulse = gaussian_pulse(0.8, fwhm_fs, 0.0)
    gap = (cx, cy, cz)
    return Scenario(
        name=name or f"adjoint_{nx}x{ny}x{nz}_d{design[0]}x{design[1]}x{design[2]}", shape=tuple(shape),
        dx=dx, precision=precision, matw=pack(mat), lut=lut, pml=([pml] * 6, 1e-15, 2.0),
        pml_fold=pulse["freq"], sources=[plane_source(shape, dx, pml + 2, pulse)], design=(lo, hi, eps), steps=steps,
        meta=dict(gap=gap))


def adjoint_variants(n=16, seed=5, shape=(192, 192, 96), design=(48, 48, 12), steps=3000,
                     precision="fp32"):
    """Batched adjoint workload: `n` design variants (eps uniform in [1, 4] on the design box,
    CEP uniform in [0, 2 pi)) of the adjoint junction, each with its gradient."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        s = adjoint_junction(shape=shape, design=design, steps=steps, seed=int(rng.integers(1 << 30)),
                             precision=precision, name=f"variant{i:02d}")
        cep = float(rng.uniform(0, 2 * math.pi))
        s.sources[0]["phase"] = cep
        s.meta["cep"] = cep
        out.append(s)
    return out


def large_junction(steps=10000, precision="fp32"):
    """The large plain run: 1024x768x512 cells of the bowtie junction (Au/SiO2/graphene), sized
    to the MI355X's 288 GB of HBM."""
    s = bowtie_junction(nx=1024, ny=768, nz=512, steps=steps, precision=precision, pml=12)
    s.name = "bowtie_1024x768x512"
    return s
