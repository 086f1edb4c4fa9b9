"""ctypes handle on the FP64 CPU oracle (oracle/fdtd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's cpu_baseline
leg.  Same method names and argument meaning as sim_juncs_amd.engine.Engine so one scenario
description can drive both.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fdtd_oracle.c")
LIB = os.path.join(HERE, "liborc.so")
MAXP = 4
COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (OpenMP, no host-specific ISA flags)."""
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        subprocess.check_call(["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-o", LIB, SRC, "-lm"])
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    proto = {
        "orc_create": (vp, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, ip, C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_dt": (C.c_double, [vp]),
        "orc_step_count": (C.c_long, [vp]),
        "orc_set_materials": (C.c_int, [vp, C.POINTER(C.c_uint32), C.c_int, dp, ip, dp]),
        "orc_set_eps_box": (C.c_int, [vp, ip, ip, dp]),
        "orc_set_pml": (C.c_int, [vp, ip, C.c_double, C.c_double]),
        "orc_set_pml_dispersion": (C.c_int, [vp, C.c_double]),
        "orc_set_eps_maps": (C.c_int, [vp, dp]),
        "orc_add_source": (C.c_int, [vp, C.c_int, dp, C.c_int, dp, C.c_double, C.c_double]),
        "orc_last_source_time": (C.c_double, [vp]),
        "orc_add_monitors": (C.c_int, [vp, C.c_int, dp, C.c_int]),
        "orc_set_save_span": (None, [vp, C.c_int]),
        "orc_num_samples": (C.c_long, [vp]),
        "orc_read_monitors": (None, [vp, dp]),
        "orc_add_dft": (C.c_int, [vp, ip, ip, C.c_int, dp, C.c_int]),
        "orc_read_dft": (C.c_int, [vp, C.c_int, dp]),
        "orc_get_field": (None, [vp, C.c_int, dp]),
        "orc_set_field": (None, [vp, C.c_int, dp]),
        "orc_step": (None, [vp, C.c_long]),
        "orc_gaussian_dipole": (None, [C.c_double] * 6 + [dp, dp, dp]),
        "orc_cw_dipole": (None, [C.c_double] * 6 + [dp, dp]),
        "orc_num_threads": (C.c_int, []),
    }
    for name, (res, args) in proto.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return (C.c_int * len(a))(*[int(x) for x in a])


def comp_index(c) -> int:
    return COMPONENTS.index(c) if isinstance(c, str) else int(c)


def gaussian_dipole(f, w, phase, st, et, t):
    """dipole(t) of gaussian_src_time_phase as the oracle evaluates it; returns (value, cutoff)."""
    lib = load()
    re, im, cut = C.c_double(), C.c_double(), C.c_double()
    lib.orc_gaussian_dipole(f, w, phase, st, et, t, C.byref(re), C.byref(im), C.byref(cut))
    return complex(re.value, im.value), cut.value


def cw_dipole(f, w, st, et, slowness, t):
    """dipole(t) of a continuous source (meep continuous_src_time) as the oracle evaluates it."""
    lib = load()
    re, im = C.c_double(), C.c_double()
    lib.orc_cw_dipole(f, w, st, et, slowness, t, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def num_threads() -> int:
    return load().orc_num_threads()


class Oracle:
    def __init__(self, shape, dx: float, courant: float = 0.5, precision: str = "fp64",
                 periodic=(0, 0, 0), part: int = 0, device: int = 0):
        self.lib = load()
        self.shape = tuple(int(s) for s in shape)
        self.dx = float(dx)
        self.h = self.lib.orc_create(*self.shape, self.dx, float(courant), _ip(periodic), int(part))
        self._nmon = 0
        self._dft = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    @property
    def dt(self):
        return self.lib.orc_dt(self.h)

    def time(self):
        return self.lib.orc_step_count(self.h) * self.dt

    def step_count(self):
        return self.lib.orc_step_count(self.h)

    def set_materials(self, matw, lut):
        matw = np.ascontiguousarray(matw, dtype=np.uint32).reshape(-1)
        nm = len(lut)
        eps = np.array([m.eps_inf for m in lut], dtype=np.float64)
        npole = (C.c_int * nm)(*[len(m.poles) for m in lut])
        poles = np.zeros((nm, MAXP, 4))
        for i, m in enumerate(lut):
            for s, p in enumerate(m.poles):
                poles[i, s] = (p.f0, p.gamma, p.sigma, 1.0 if p.drude else 0.0)
        rc = self.lib.orc_set_materials(self.h, matw.ctypes.data_as(C.POINTER(C.c_uint32)), nm,
                                        _dp(eps), npole, _dp(poles))
        if rc:
            raise ValueError(f"orc_set_materials failed ({rc})")

    def set_eps_maps(self, eps):
        """Per-cell eps_inf of Ex, Ey, Ez (array [3, nx, ny, nz])."""
        v = np.ascontiguousarray(eps, dtype=np.float64).reshape((3,) + self.shape)
        self.lib.orc_set_eps_maps(self.h, _dp(v))

    def set_design_eps(self, lo, hi, eps):
        eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1)
        self.lib.orc_set_eps_box(self.h, _ip(lo), _ip(hi), _dp(eps))

    def set_pml(self, cells, R=1e-15, power=2.0):
        self.lib.orc_set_pml(self.h, _ip(cells), float(R), float(power))

    def set_pml_dispersion(self, f_ref):
        """Fold Lorentz poles out of the PML slabs at frequency f_ref (0: keep them)."""
        self.lib.orc_set_pml_dispersion(self.h, float(f_ref))

    def add_source(self, comp, box, kind="gaussian", freq=1.0, width=1.0, phase=0.0,
                   t_start=0.0, t_end=0.0, slowness=3.0, amp=1.0):
        b = np.ascontiguousarray(box, dtype=np.float64)
        if kind == "gaussian":
            prm = np.array([freq, width, phase, t_start, t_end], dtype=np.float64)
            k = 0
        else:
            prm = np.array([freq, width, t_start, t_end, slowness], dtype=np.float64)
            k = 1
        amp = complex(amp)
        self.lib.orc_add_source(self.h, comp_index(comp), _dp(b), k, _dp(prm), amp.real, amp.imag)

    def last_source_time(self):
        return self.lib.orc_last_source_time(self.h)

    def add_monitors(self, comp, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.lib.orc_add_monitors(self.h, comp_index(comp), _dp(xyz), xyz.shape[0])
        self._nmon += xyz.shape[0]

    def set_save_span(self, span):
        self.lib.orc_set_save_span(self.h, int(span))

    def read_monitors(self):
        ns = self.lib.orc_num_samples(self.h)
        out = np.zeros((ns, self._nmon))
        if ns and self._nmon:
            self.lib.orc_read_monitors(self.h, _dp(out))
        return out.astype(complex)

    def add_dft(self, lo, hi, comps, freqs):
        mask = 0
        for c in comps:
            mask |= 1 << comp_index(c)
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        i = self.lib.orc_add_dft(self.h, _ip(lo), _ip(hi), mask, _dp(f), f.size)
        self._dft[i] = (bin(mask).count("1"), f.size, tuple(int(h) - int(l) for l, h in zip(lo, hi)))
        return i

    def read_dft(self, i):
        ncomp, nf, ext = self._dft[i]
        out = np.zeros((ncomp, nf) + ext + (2,))
        self.lib.orc_read_dft(self.h, int(i), _dp(out))
        return out[..., 0] + 1j * out[..., 1]

    def step(self, n=1):
        self.lib.orc_step(self.h, int(n))

    def synchronize(self):
        pass

    def read_field(self, comp):
        out = np.zeros(self.shape)
        self.lib.orc_get_field(self.h, comp_index(comp), _dp(out))
        return out

    def write_field(self, comp, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.shape)
        self.lib.orc_set_field(self.h, comp_index(comp), _dp(v))
