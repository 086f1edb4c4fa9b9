"""ctypes binding of libsjf.so (the C ABI declared in include/sjf.h).

The library is built in-tree by ``sim_juncs_amd.build`` (hipcc for gfx950).  Loading fails
loudly when it is missing: there is no CPU fallback for the engine.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsjf.so")

MAX_POLES = 4
MAX_DFT_FREQS = 64

SJF_OK = 0
STATUS = {
    0: "SJF_OK",
    -1: "SJF_E_ARG",
    -2: "SJF_E_HIP",
    -3: "SJF_E_STATE",
    -4: "SJF_E_UNSUPPORTED",
    -5: "SJF_E_NOMEM",
}

# Every symbol include/sjf.h declares; tests check the library exports all of them.
EXPORTS = [
    "sjf_create", "sjf_destroy", "sjf_last_error", "sjf_set_materials", "sjf_set_design_eps",
    "sjf_set_pml", "sjf_add_source", "sjf_last_source_time", "sjf_dt", "sjf_time",
    "sjf_step_count", "sjf_add_monitors", "sjf_set_save_span", "sjf_read_monitors",
    "sjf_get_field", "sjf_add_dft", "sjf_read_dft", "sjf_step", "sjf_synchronize",
    "sjf_timed_steps", "sjf_set_zsegment", "sjf_read_field", "sjf_read_field_box", "sjf_write_field", "sjf_step_bytes",
    "sjf_launch_info", "sjf_set_tile", "sjf_copy_probe", "sjf_set_graph",
    "sjf_batch_create", "sjf_batch_set_tile", "sjf_batch_step", "sjf_batch_synchronize", "sjf_batch_timed_steps",
    "sjf_batch_info", "sjf_batch_last_error", "sjf_batch_destroy", "sjf_adjoint_run",
    "sjf_adjoint_kernel_info", "sjf_set_adjoint_tile", "sjf_batch_launch_info", "sjf_batch_adjoint_run", "sjf_batch_adjoint_kernel_info",
    "sjf_dft_flux", "sjf_reserve_steps", "sjf_batch_reserve_steps", "sjf_dft_flux_noalloc",
    "sjf_set_pml_dispersion", "sjf_set_eps_maps",
]


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("courant", C.c_double),
        ("precision", C.c_int), ("periodic", C.c_int * 3),
        ("part", C.c_int), ("device", C.c_int),
    ]


class Material(C.Structure):
    _fields_ = [
        ("eps_inf", C.c_double), ("npoles", C.c_int),
        ("f0", C.c_double * MAX_POLES), ("gamma", C.c_double * MAX_POLES),
        ("sigma", C.c_double * MAX_POLES), ("is_drude", C.c_int * MAX_POLES),
    ]


class Pulse(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("freq", C.c_double), ("width", C.c_double),
        ("phase", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
        ("slowness", C.c_double),
    ]


_lib = None


class EngineError(RuntimeError):
    pass


def load(path: str | None = None) -> C.CDLL:
    """Load libsjf.so and declare prototypes.  Raises when the library is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise EngineError(
            f"{p} not found: the HIP engine is not built (run __graft_entry__.build() or "
            "python -m sim_juncs_amd.build); there is no CPU fallback")
    lib = C.CDLL(p)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    lp = C.POINTER(C.c_long)
    proto = {
        "sjf_create": (C.c_int, [C.POINTER(Grid), C.POINTER(vp)]),
        "sjf_destroy": (C.c_int, [vp]),
        "sjf_last_error": (C.c_char_p, [vp]),
        "sjf_set_materials": (C.c_int, [vp, C.POINTER(C.c_uint32), C.POINTER(Material), C.c_int]),
        "sjf_set_design_eps": (C.c_int, [vp, ip, ip, dp]),
        "sjf_set_pml": (C.c_int, [vp, ip, C.c_double, C.c_double]),
        "sjf_add_source": (C.c_int, [vp, C.c_int, dp, C.POINTER(Pulse), C.c_double, C.c_double]),
        "sjf_last_source_time": (C.c_int, [vp, dp]),
        "sjf_dt": (C.c_int, [vp, dp]),
        "sjf_time": (C.c_int, [vp, dp]),
        "sjf_step_count": (C.c_int, [vp, lp]),
        "sjf_add_monitors": (C.c_int, [vp, C.c_int, dp, C.c_int]),
        "sjf_set_save_span": (C.c_int, [vp, C.c_int]),
        "sjf_read_monitors": (C.c_int, [vp, dp, C.c_size_t, lp]),
        "sjf_get_field": (C.c_int, [vp, C.c_int, dp, dp]),
        "sjf_add_dft": (C.c_int, [vp, ip, ip, C.c_int, dp, C.c_int, ip]),
        "sjf_read_dft": (C.c_int, [vp, C.c_int, dp, C.c_size_t]),
        "sjf_step": (C.c_int, [vp, C.c_long]),
        "sjf_synchronize": (C.c_int, [vp]),
        "sjf_reserve_steps": (C.c_int, [vp, C.c_long]),
        "sjf_set_pml_dispersion": (C.c_int, [vp, C.c_double]),
        "sjf_set_eps_maps": (C.c_int, [vp, dp]),
        "sjf_batch_reserve_steps": (C.c_int, [vp, C.c_long]),
        "sjf_timed_steps": (C.c_int, [vp, C.c_long, dp, dp]),
        "sjf_read_field": (C.c_int, [vp, C.c_int, dp]),
        "sjf_read_field_box": (C.c_int, [vp, C.c_int, ip, ip, dp]),
        "sjf_write_field": (C.c_int, [vp, C.c_int, dp]),
        "sjf_step_bytes": (C.c_int, [vp, dp]),
        "sjf_launch_info": (C.c_int, [vp, ip, ip, ip, ip, ip]),
        "sjf_set_tile": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "sjf_copy_probe": (C.c_int, [C.c_int, C.c_size_t, C.c_int, dp]),
        "sjf_set_graph": (C.c_int, [vp, C.c_int]),
        "sjf_batch_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
        "sjf_batch_set_tile": (C.c_int, [vp, C.c_int, C.c_int]),
        "sjf_batch_step": (C.c_int, [vp, C.c_long]),
        "sjf_batch_synchronize": (C.c_int, [vp]),
        "sjf_batch_timed_steps": (C.c_int, [vp, C.c_long, dp, dp]),
        "sjf_batch_info": (C.c_int, [vp, lp, dp]),
        "sjf_batch_last_error": (C.c_char_p, [vp]),
        "sjf_batch_destroy": (C.c_int, [vp]),
        "sjf_adjoint_run": (C.c_int, [vp, C.c_long, C.c_int, ip, dp, dp, C.c_size_t]),
        "sjf_adjoint_kernel_info": (C.c_int, [vp, dp, ip, ip]),
        "sjf_set_adjoint_tile": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "sjf_set_zsegment": (C.c_int, [vp, C.c_int]),
        "sjf_batch_launch_info": (C.c_int, [vp, ip, ip]),
        "sjf_batch_adjoint_run": (C.c_int, [vp, C.c_long, C.c_int, ip, dp, C.POINTER(dp), C.POINTER(C.c_size_t)]),
        "sjf_batch_adjoint_kernel_info": (C.c_int, [vp, dp, ip, ip]),
        "sjf_dft_flux": (C.c_int, [vp, C.c_int, C.c_int, dp, C.c_size_t]),
        "sjf_dft_flux_noalloc": (C.c_int, [vp, C.c_int, C.c_int, dp, C.c_size_t]),
    }
    for name, (res, args) in proto.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib
