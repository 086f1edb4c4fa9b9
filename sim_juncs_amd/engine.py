"""Python handle on the HIP engine (libsjf.so) through its C ABI.

`Engine` mirrors the meep calls the reference driver makes (src/disp.cpp:482-749): build a
grid and structure, add sources and monitors, `step()`, read monitor series and fields.
Every method calls straight into libsjf; nothing here computes physics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .materials import MaterialSpec

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
TILE_FORMS = {1: "register_rows", 2: "thread_rows"}


def comp_index(c) -> int:
    if isinstance(c, str):
        return COMPONENTS.index(c)
    return int(c)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    arr = (C.c_int * len(a))(*[int(x) for x in a])
    return arr


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 20, device: int = 0) -> float:
    """Achievable HBM bandwidth (GB/s, read + write) of a plain device copy of `nbytes`."""
    lib = _lib.load()
    g = C.c_double()
    rc = lib.sjf_copy_probe(int(device), int(nbytes), int(iters), C.byref(g))
    if rc != 0:
        raise _lib.EngineError(f"sjf_copy_probe failed ({_lib.STATUS.get(rc, rc)})")
    return g.value


class Engine:
    """One simulation on one GPU.  precision: "fp32" or "fp64"."""

    def __init__(self, shape, dx: float, courant: float = 0.5, precision: str = "fp32",
                 periodic=(0, 0, 0), part: int = 0, device: int = 0, lib_path: str | None = None):
        self.lib = _lib.load(lib_path)
        self.shape = tuple(int(s) for s in shape)
        self.dx = float(dx)
        self.precision = precision
        g = _lib.Grid()
        g.nx, g.ny, g.nz = self.shape
        g.dx = self.dx
        g.courant = float(courant)
        g.precision = 4 if precision == "fp32" else 8
        for d in range(3):
            g.periodic[d] = int(periodic[d])
        g.part = int(part)
        g.device = int(device)
        h = C.c_void_p()
        rc = self.lib.sjf_create(C.byref(g), C.byref(h))
        if rc != 0:
            raise _lib.EngineError(
                f"sjf_create failed ({_lib.STATUS.get(rc, rc)}): "
                f"{self.lib.sjf_last_error(None).decode()}")
        self.h = h

    # ------------------------------------------------------------------ plumbing ----------
    def _chk(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.sjf_last_error(self.h).decode()
            raise _lib.EngineError(f"{what} failed ({_lib.STATUS.get(rc, rc)}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            if self.lib.sjf_destroy(self.h) == 0:
                self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    # ------------------------------------------------------------------ structure ---------
    def set_materials(self, matw: np.ndarray, lut: list[MaterialSpec]):
        matw = np.ascontiguousarray(matw, dtype=np.uint32).reshape(-1)
        if matw.size != self.n:
            raise ValueError("material word array must have nx*ny*nz entries")
        arr = (_lib.Material * len(lut))()
        for m, spec in enumerate(lut):
            arr[m].eps_inf = spec.eps_inf
            arr[m].npoles = len(spec.poles)
            for s, p in enumerate(spec.poles):
                arr[m].f0[s] = p.f0
                arr[m].gamma[s] = p.gamma
                arr[m].sigma[s] = p.sigma
                arr[m].is_drude[s] = 1 if p.drude else 0
        rc = self.lib.sjf_set_materials(self.h, matw.ctypes.data_as(C.POINTER(C.c_uint32)), arr,
                                        len(lut))
        self._chk(rc, "sjf_set_materials")

    def set_design_eps(self, lo, hi, eps: np.ndarray):
        eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1)
        self._chk(self.lib.sjf_set_design_eps(self.h, _ip(lo), _ip(hi), _dp(eps)),
                  "sjf_set_design_eps")

    def set_pml(self, cells, R: float = 1e-15, power: float = 2.0):
        self._chk(self.lib.sjf_set_pml(self.h, _ip(cells), float(R), float(power)), "sjf_set_pml")

    def set_eps_maps(self, eps):
        """Per-cell eps_inf of Ex, Ey, Ez (array [3, nx, ny, nz]); None removes the map."""
        if eps is None:
            self._chk(self.lib.sjf_set_eps_maps(self.h, None), "sjf_set_eps_maps")
            return
        v = np.ascontiguousarray(eps, dtype=np.float64).reshape((3,) + tuple(self.shape))
        self._chk(self.lib.sjf_set_eps_maps(self.h, _dp(v)), "sjf_set_eps_maps")

    def set_pml_dispersion(self, f_ref: float):
        """Run the PML slabs without Lorentz poles, folded into eps_inf at f_ref (0: keep them)."""
        self._chk(self.lib.sjf_set_pml_dispersion(self.h, float(f_ref)), "sjf_set_pml_dispersion")

    # ------------------------------------------------------------------ sources -----------
    def add_source(self, comp, box, kind: str = "gaussian", freq: float = 1.0,
                   width: float = 1.0, phase: float = 0.0, t_start: float = 0.0,
                   t_end: float = 0.0, slowness: float = 3.0, amp: complex = 1.0):
        p = _lib.Pulse()
        p.kind = 0 if kind == "gaussian" else 1
        p.freq, p.width, p.phase = float(freq), float(width), float(phase)
        p.t_start, p.t_end, p.slowness = float(t_start), float(t_end), float(slowness)
        b = np.ascontiguousarray(box, dtype=np.float64)
        amp = complex(amp)
        self._chk(self.lib.sjf_add_source(self.h, comp_index(comp), _dp(b), C.byref(p), amp.real,
                                          amp.imag), "sjf_add_source")

    def last_source_time(self) -> float:
        t = C.c_double()
        self._chk(self.lib.sjf_last_source_time(self.h, C.byref(t)), "sjf_last_source_time")
        return t.value

    @property
    def dt(self) -> float:
        t = C.c_double()
        self._chk(self.lib.sjf_dt(self.h, C.byref(t)), "sjf_dt")
        return t.value

    def time(self) -> float:
        t = C.c_double()
        self._chk(self.lib.sjf_time(self.h, C.byref(t)), "sjf_time")
        return t.value

    def step_count(self) -> int:
        n = C.c_long()
        self._chk(self.lib.sjf_step_count(self.h, C.byref(n)), "sjf_step_count")
        return n.value

    # ------------------------------------------------------------------ monitors ----------
    def add_monitors(self, comp, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self._chk(self.lib.sjf_add_monitors(self.h, comp_index(comp), _dp(xyz), xyz.shape[0]),
                  "sjf_add_monitors")
        self._nmon = getattr(self, "_nmon", 0) + xyz.shape[0]

    def set_save_span(self, span: int):
        self._chk(self.lib.sjf_set_save_span(self.h, int(span)), "sjf_set_save_span")

    def read_monitors(self) -> np.ndarray:
        """complex array [n_samples, n_points]"""
        ns = C.c_long()
        self._chk(self.lib.sjf_read_monitors(self.h, None, 0, C.byref(ns)), "sjf_read_monitors")
        npt = getattr(self, "_nmon", 0)
        out = np.zeros((ns.value, npt, 2))
        if ns.value and npt:
            self._chk(self.lib.sjf_read_monitors(self.h, _dp(out), out.size, C.byref(ns)),
                      "sjf_read_monitors")
        return out[..., 0] + 1j * out[..., 1]

    def get_field(self, comp, xyz) -> complex:
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        out = np.zeros(2)
        self._chk(self.lib.sjf_get_field(self.h, comp_index(comp), _dp(x), _dp(out)),
                  "sjf_get_field")
        return complex(out[0], out[1])

    # ------------------------------------------------------------------ DFT ---------------
    def add_dft(self, lo, hi, comps, freqs) -> int:
        mask = 0
        for c in comps:
            mask |= 1 << comp_index(c)
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        i = C.c_int()
        self._chk(self.lib.sjf_add_dft(self.h, _ip(lo), _ip(hi), mask, _dp(f), f.size, C.byref(i)),
                  "sjf_add_dft")
        self._dft = getattr(self, "_dft", {})
        ncomp = bin(mask).count("1")
        ext = tuple(int(h) - int(l) for l, h in zip(lo, hi))
        self._dft[i.value] = (ncomp, f.size, ext)
        return i.value

    def read_dft(self, i: int) -> np.ndarray:
        """complex array [n_comp (ascending component order), n_freq, *box]"""
        ncomp, nf, ext = self._dft[i]
        out = np.zeros((ncomp, nf) + ext + (2,))
        self._chk(self.lib.sjf_read_dft(self.h, int(i), _dp(out), out.size), "sjf_read_dft")
        return out[..., 0] + 1j * out[..., 1]

    def dft_flux(self, i: int, axis: int) -> np.ndarray:
        """Flux along +axis through DFT plane i, one value per frequency (sjf_dft_flux)."""
        nf = self._dft[i][1]
        out = np.zeros(nf)
        self._chk(self.lib.sjf_dft_flux(self.h, int(i), int(axis), _dp(out), out.size), "sjf_dft_flux")
        return out

    def dft_flux_noalloc(self, i: int, axis: int) -> np.ndarray:
        """dft_flux through the allocation-free entry point (sjf_dft_flux_noalloc)."""
        nf = self._dft[i][1]
        out = np.zeros(nf)
        self._chk(self.lib.sjf_dft_flux_noalloc(self.h, int(i), int(axis), _dp(out), out.size),
                  "sjf_dft_flux_noalloc")
        return out

    # ------------------------------------------------------------------ stepping ----------
    def step(self, n: int = 1):
        self._chk(self.lib.sjf_step(self.h, int(n)), "sjf_step")

    def synchronize(self):
        self._chk(self.lib.sjf_synchronize(self.h), "sjf_synchronize")

    def reserve_steps(self, n: int):
        """Size the sample buffer and build the launch tables for the next n steps now, so that
        stepping them allocates nothing (sjf_reserve_steps)."""
        self._chk(self.lib.sjf_reserve_steps(self.h, int(n)), "sjf_reserve_steps")

    def timed_steps(self, n: int):
        """(ms_total, ms_kernel_mean) from HIP events on the engine's stream."""
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.sjf_timed_steps(self.h, int(n), C.byref(a), C.byref(b)),
                  "sjf_timed_steps")
        return a.value, b.value

    # ------------------------------------------------------------------ fields ------------
    def read_field(self, comp, lo=None, hi=None) -> np.ndarray:
        """A whole component, or with lo/hi the cell box [lo, hi) of it (FP64, shape of the box)."""
        if lo is None and hi is None:
            out = np.zeros(self.shape)
            self._chk(self.lib.sjf_read_field(self.h, comp_index(comp), _dp(out)), "sjf_read_field")
            return out
        lo = np.ascontiguousarray(lo if lo is not None else (0, 0, 0), dtype=np.int32)
        hi = np.ascontiguousarray(hi if hi is not None else self.shape, dtype=np.int32)
        out = np.zeros(tuple(max(int(h) - int(l), 0) for l, h in zip(lo, hi)))
        self._chk(self.lib.sjf_read_field_box(self.h, comp_index(comp), _ip(lo), _ip(hi), _dp(out)),
                  "sjf_read_field_box")
        return out

    def write_field(self, comp, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.shape)
        self._chk(self.lib.sjf_write_field(self.h, comp_index(comp), _dp(v)), "sjf_write_field")

    # ------------------------------------------------------------------ accounting --------
    def step_bytes(self) -> float:
        b = C.c_double()
        self._chk(self.lib.sjf_step_bytes(self.h, C.byref(b)), "sjf_step_bytes")
        return b.value

    def launch_info(self) -> dict:
        """Fused-kernel launch shape: form ("register_rows" | "thread_rows"), y rows per thread
        or per workgroup, x-chunk length, threads per workgroup, workgroups."""
        v = [C.c_int() for _ in range(5)]
        self._chk(self.lib.sjf_launch_info(self.h, *[C.byref(x) for x in v]), "sjf_launch_info")
        return dict(form=TILE_FORMS[v[0].value], rows=v[1].value, xchunk=v[2].value,
                    threads=v[3].value, groups=v[4].value)

    def set_tile(self, form: str, rows: int, xchunk: int):
        """Override the default tile: form "register_rows" (rows per thread in {1, 2, 4}) or
        "thread_rows" (rows per workgroup, (rows + 1) * nz <= 1024)."""
        code = {v: k for k, v in TILE_FORMS.items()}[form]
        self._chk(self.lib.sjf_set_tile(self.h, code, int(rows), int(xchunk)), "sjf_set_tile")

    def adjoint_run(self, nsteps: int, checkpoint_every: int, gap, n_design: int):
        """Forward run of `nsteps` steps plus the discrete adjoint: returns (J, dJ/d eps) with
        J = dt * sum_n |E^n(gap)|^2 and the gradient over the design voxels (z fastest).
        checkpoint_every = K > 0 checkpoints every K steps and recomputes; 0 records the design
        fields of every step when they fit (else checkpoints every ~sqrt(nsteps))."""
        J = C.c_double()
        grad = np.zeros(int(n_design))
        self._chk(self.lib.sjf_adjoint_run(self.h, int(nsteps), int(checkpoint_every), _ip(gap),
                                           C.byref(J), _dp(grad), grad.size), "sjf_adjoint_run")
        return J.value, grad

    def set_zsegment(self, cells: int):
        """Register-rows tile: z cells per workgroup pass (a multiple of 64; 0: one pass)."""
        self._chk(self.lib.sjf_set_zsegment(self.h, int(cells)), "sjf_set_zsegment")

    def set_adjoint_tile(self, form: str, rows: int = 1, xchunk: int = 0):
        """Tile of the fused reverse step: "register_rows" (rows = 1) or "thread_rows" (owned
        rows per workgroup, (rows + 1) * nz <= 1024); xchunk 0 picks the default."""
        code = {v: k for k, v in TILE_FORMS.items()}[form]
        self._chk(self.lib.sjf_set_adjoint_tile(self.h, code, int(rows), int(xchunk)), "sjf_set_adjoint_tile")

    def adjoint_kernel_info(self):
        """(mean ms, threads, workgroups) of the fused reverse-step kernel of the last adjoint run."""
        ms, th, gr = C.c_double(), C.c_int(), C.c_int()
        self._chk(self.lib.sjf_adjoint_kernel_info(self.h, C.byref(ms), C.byref(th), C.byref(gr)),
                  "sjf_adjoint_kernel_info")
        return ms.value, th.value, gr.value

    def set_graph(self, chunk: int):
        """Steps per replayed hipGraph (0: launch every step directly)."""
        self._chk(self.lib.sjf_set_graph(self.h, int(chunk)), "sjf_set_graph")


class Batch:
    """Engines of uneven grid shapes advanced together: one fused launch per time step over all
    of them (sjf_batch_* in include/sjf.h)."""

    def __init__(self, engines):
        self.lib = _lib.load()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        rc = self.lib.sjf_batch_create(arr, len(self.engines), C.byref(h))
        if rc != 0:
            raise _lib.EngineError(f"sjf_batch_create failed ({_lib.STATUS.get(rc, rc)}): "
                                   f"{self.lib.sjf_batch_last_error(None).decode()}")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise _lib.EngineError(f"{what} failed ({_lib.STATUS.get(rc, rc)}): "
                                   f"{self.lib.sjf_batch_last_error(self.h).decode()}")

    def set_tile(self, form: str, xchunk: int = 0):
        """Tile form of the fused launch ("register_rows" or "thread_rows") and the x-chunk of
        every instance (0: each engine's own)."""
        code = {v: k for k, v in TILE_FORMS.items()}[form]
        self._chk(self.lib.sjf_batch_set_tile(self.h, code, int(xchunk)), "sjf_batch_set_tile")

    def step(self, n: int = 1):
        self._chk(self.lib.sjf_batch_step(self.h, int(n)), "sjf_batch_step")

    def synchronize(self):
        self._chk(self.lib.sjf_batch_synchronize(self.h), "sjf_batch_synchronize")

    def reserve_steps(self, n: int):
        self._chk(self.lib.sjf_batch_reserve_steps(self.h, int(n)), "sjf_batch_reserve_steps")

    def timed_steps(self, n: int):
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.sjf_batch_timed_steps(self.h, int(n), C.byref(a), C.byref(b)),
                  "sjf_batch_timed_steps")
        return a.value, b.value

    def info(self):
        c, b = C.c_long(), C.c_double()
        self._chk(self.lib.sjf_batch_info(self.h, C.byref(c), C.byref(b)), "sjf_batch_info")
        return c.value, b.value

    def launch_info(self):
        """(threads per workgroup, workgroups) of the fused launch as last built."""
        th, gr = C.c_int(), C.c_int()
        self._chk(self.lib.sjf_batch_launch_info(self.h, C.byref(th), C.byref(gr)), "sjf_batch_launch_info")
        return th.value, gr.value

    def adjoint_run(self, nsteps: int, gaps, n_design, checkpoint_every: int = 0):
        """Engine.adjoint_run for every engine in one pass: one fused reverse launch per step over
        all engines.  gaps: one (i, j, k) per engine; n_design: design voxels per engine;
        checkpoint_every as in Engine.adjoint_run.  Returns (J per engine, list of gradients)."""
        n = len(self.engines)
        if len(gaps) != n or len(n_design) != n:
            raise ValueError("one gap cell and one design size per engine")
        g = np.ascontiguousarray(np.asarray(gaps, dtype=np.int32).reshape(n, 3))
        J = np.zeros(n)
        grads = [np.zeros(int(m)) for m in n_design]
        gp = (C.POINTER(C.c_double) * n)(*[_dp(x) for x in grads])
        caps = (C.c_size_t * n)(*[x.size for x in grads])
        self._chk(self.lib.sjf_batch_adjoint_run(self.h, int(nsteps), int(checkpoint_every), _ip(g.ravel()), _dp(J),
                                                 gp, caps),
                  "sjf_batch_adjoint_run")
        return J, grads

    def adjoint_kernel_info(self):
        """(mean ms, threads, workgroups) of the fused reverse launch of the last batched adjoint run."""
        ms, th, gr = C.c_double(), C.c_int(), C.c_int()
        self._chk(self.lib.sjf_batch_adjoint_kernel_info(self.h, C.byref(ms), C.byref(th), C.byref(gr)),
                  "sjf_batch_adjoint_kernel_info")
        return ms.value, th.value, gr.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.sjf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
