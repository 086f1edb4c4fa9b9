"""Python binding of the host layer (libsjf_host.so, C entry points of sjf_host.h).

The host layer restates the reference driver around the engine: configuration parsing
(src/argparse.h), the CSG scene and its material maps (src/cgs.cpp, src/disp.cpp), source
descriptions, the monitor-series transform (src/data_utils.cpp) and field_samples.h5
(src/disp.cpp:758-923).  Scenes are JSON files (see sjfh::Scene::from_json_file) or .geom
scripts in the reference's scene language (geom_scene; the driver takes either).
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsjf_host.so")
GEOM_BIN = os.path.join(HERE, "sjf_geom")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m sim_juncs_amd.build`")
        L = C.CDLL(LIB_PATH)
        dp, ip, cp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_char_p
        up, szp, lp = C.POINTER(C.c_uint), C.POINTER(C.c_size_t), C.POINTER(C.c_long)
        cpp = C.POINTER(C.c_char_p)
        protos = {
            "sjfh_last_error": (C.c_char_p, []),
            "sjfh_settings_json": (C.c_int, [cp, C.c_int, cpp, C.c_int, C.c_char_p, C.c_size_t]),
            "sjfh_scene_in": (C.c_int, [cp, dp, C.c_long, ip, ip]),
            "sjfh_in_bound": (C.c_int, [cp, C.c_double, C.c_uint, C.c_double, dp, C.c_long, dp]),
            "sjfh_smooth_points": (C.c_int, [C.c_uint, C.c_double, dp, C.c_size_t, up]),
            "sjfh_susceptibilities": (C.c_int, [cp, C.c_int, dp, C.c_int, ip, ip]),
            "sjfh_source_info": (C.c_int, [cp, C.c_int, C.c_double, dp, dp]),
            "sjfh_fft": (C.c_int, [dp, C.c_long, dp, C.c_long, lp]),
            "sjfh_voxelize": (C.c_int, [cp, C.c_double, C.c_uint, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_double, up, dp, C.c_int, ip, dp, ip]),
            "sjfh_run": (C.c_int, [cp, C.c_int, cpp, cp, C.c_int, C.c_int, C.c_int, dp]),
            "sjfh_write_field_samples": (C.c_int, [cp, cp, C.c_double, C.c_double, C.c_uint, C.c_uint, dp, C.c_long]),
            "sjfh_h5_read": (C.c_int, [cp, cp, cpp, C.c_int, dp, C.c_size_t, szp]),
            "sjfh_h5_list": (C.c_int, [cp, cp, C.c_char_p, C.c_size_t]),
            "sjfh_geom_scene_json": (C.c_int, [cp, C.c_double, C.c_double, C.c_double, cp, cp, C.c_char_p,
                                               C.c_size_t, ip]),
        }
        for name, (res, args) in protos.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class HostError(RuntimeError):
    pass


def _chk(rc, what):
    if rc != 0:
        raise HostError(f"{what} failed ({rc}): {lib().sjfh_last_error().decode()}")


def _b(s):
    return s.encode() if isinstance(s, str) else s


def _argv(args):
    arr = (C.c_char_p * max(1, len(args)))(*[_b(a) for a in args])
    return arr


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def settings(conf_path, argv=(), mode=2) -> dict:
    """Settings as the reference reads them: mode 0 parse_conf_file only, 1 parse_args(argv) then
    parse_conf_file, 2 as 1 plus correct_defaults once more (the sequence of main())."""
    args = ["sim_geom", *argv]
    buf = C.create_string_buffer(1 << 16)
    _chk(lib().sjfh_settings_json(_b(conf_path), len(args), _argv(args), int(mode), buf, len(buf)),
         "sjfh_settings_json")
    return json.loads(buf.value.decode())


def geom_scene(geom_path, pml_thickness, length, um_scale=1.0, out_dir="", user_opts=""):
    """Run a .geom scene script with these settings values; returns (scene dict in the JSON
    form: roots, sources, monitor_clusters, context_numbers; parse_error code, 0 = success)."""
    cap = 1 << 24
    buf = C.create_string_buffer(cap)
    er = C.c_int()
    _chk(lib().sjfh_geom_scene_json(_b(geom_path), float(pml_thickness), float(length), float(um_scale), _b(out_dir),
                                    _b(user_opts), buf, cap, C.byref(er)), "sjfh_geom_scene_json")
    return json.loads(buf.value.decode()), er.value


def scene_in(scene_json, pts) -> np.ndarray:
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    nr = C.c_int()
    # the root count is only known after loading: size the output for a generous maximum
    out = np.zeros(pts.shape[0] * 64, dtype=np.int32)
    _chk(lib().sjfh_scene_in(_b(scene_json), _dp(pts), pts.shape[0], out.ctypes.data_as(C.POINTER(C.c_int)),
                             C.byref(nr)), "sjfh_scene_in")
    return out[: pts.shape[0] * nr.value].reshape(pts.shape[0], nr.value)


def in_bound(scene_json, pts, def_ret=1.0, smooth_n=0, smooth_rad=0.05) -> np.ndarray:
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(pts.shape[0])
    _chk(lib().sjfh_in_bound(_b(scene_json), def_ret, smooth_n, smooth_rad, _dp(pts), pts.shape[0], _dp(out)),
         "sjfh_in_bound")
    return out


def smooth_points(smooth_n, smooth_rad) -> np.ndarray:
    out = np.zeros(24 * max(1, smooth_n))
    n = C.c_uint()
    _chk(lib().sjfh_smooth_points(smooth_n, smooth_rad, _dp(out), out.size, C.byref(n)), "sjfh_smooth_points")
    return out[: 3 * n.value].reshape(-1, 3)


def susceptibilities(scene_json, root):
    out = np.zeros(4 * 64)
    n, er = C.c_int(), C.c_int()
    _chk(lib().sjfh_susceptibilities(_b(scene_json), root, _dp(out), out.size, C.byref(n), C.byref(er)),
         "sjfh_susceptibilities")
    poles = [dict(omega_0=out[4 * q], gamma=out[4 * q + 1], sigma=out[4 * q + 2], use_denom=bool(out[4 * q + 3]))
             for q in range(n.value)]
    return poles, er.value


def source_info(scene_json, s, um_scale=1.0):
    out, pulse = np.zeros(8), np.zeros(7)
    _chk(lib().sjfh_source_info(_b(scene_json), s, um_scale, _dp(out), _dp(pulse)), "sjfh_source_info")
    info = dict(gaussian=bool(out[0]), component=int(out[1]), wavelen=out[2], width=out[3], phase=out[4],
                start_time=out[5], end_time=out[6], amplitude=out[7])
    p = dict(kind=int(pulse[0]), freq=pulse[1], width=pulse[2], phase=pulse[3], t_start=pulse[4], t_end=pulse[5],
             slowness=pulse[6])
    return info, p


def fft(x) -> np.ndarray:
    x = np.asarray(x, dtype=np.complex128)
    inp = np.ascontiguousarray(np.column_stack([x.real, x.imag]).reshape(-1))
    out = np.zeros(2 * max(1, x.size))
    n = C.c_long()
    _chk(lib().sjfh_fft(_dp(inp), x.size, _dp(out), out.size, C.byref(n)), "sjfh_fft")
    o = out[: 2 * n.value].reshape(-1, 2)
    return o[:, 0] + 1j * o[:, 1]


def voxelize(scene_json, n, dx, ambient_eps=1.0, smooth_n=0, smooth_rad=0.05, um_scale=1.0, resolution=None,
             with_map=False):
    """Material words (n^3, z fastest) and the material table as dicts; with_map=True also returns
    the per-cell eps_inf map [3, n, n, n] (None without smoothing)."""
    words = np.zeros(n ** 3, dtype=np.uint32)
    mats = np.zeros(26 * 256)
    emap = np.zeros(3 * n ** 3)
    nm = C.c_int()
    has = C.c_int()
    res = resolution if resolution is not None else 1.0 / dx
    _chk(lib().sjfh_voxelize(_b(scene_json), ambient_eps, smooth_n, smooth_rad, um_scale, res, n, dx,
                             words.ctypes.data_as(C.POINTER(C.c_uint)), _dp(mats), 256, C.byref(nm), _dp(emap),
                             C.byref(has)), "sjfh_voxelize")
    lut = []
    for m in range(nm.value):
        d = mats[26 * m: 26 * (m + 1)]
        np_ = int(d[1])
        lut.append(dict(eps_inf=d[0], npoles=np_, f0=list(d[2:2 + np_]), gamma=list(d[6:6 + np_]),
                        sigma=list(d[10:10 + np_]), is_drude=[int(v) for v in d[14:14 + np_]]))
    if with_map:
        return words.reshape(n, n, n), lut, (emap.reshape(3, n, n, n) if has.value else None)
    return words.reshape(n, n, n), lut


def run(conf_path, scene_json, argv=(), precision=8, complex_fields=True, device=0) -> dict:
    """Whole driver run: settings, BoundGeom, run(), save_field_times() into the out_dir."""
    args = ["sim_geom", *argv]
    summ = np.zeros(6)
    _chk(lib().sjfh_run(_b(conf_path), len(args), _argv(args), _b(scene_json), precision, int(complex_fields),
                        device, _dp(summ)), "sjfh_run")
    return dict(grid_num=int(summ[0]), dt=summ[1], ttot=summ[2], n_t_pts=int(summ[3]), n_monitors=int(summ[4]),
                n_clusters=int(summ[5]))


def write_field_samples(path, scene_json, um_scale, ttot, n_t_pts, save_span, series):
    series = np.asarray(series, dtype=np.complex128)
    flat = np.ascontiguousarray(np.stack([series.real, series.imag], axis=-1).reshape(-1))
    _chk(lib().sjfh_write_field_samples(_b(path), _b(scene_json), um_scale, ttot, n_t_pts, save_span, _dp(flat),
                                        series.shape[1] if series.ndim == 2 else 0), "sjfh_write_field_samples")


def h5_read(path, dset, members=()):
    count = C.c_size_t()
    mem = (C.c_char_p * max(1, len(members)))(*[_b(m) for m in members])
    rc = lib().sjfh_h5_read(_b(path), _b(dset), mem, len(members), None, 0, C.byref(count))
    if rc not in (0, -3):
        raise HostError(f"cannot read {dset} from {path} ({rc})")
    per = max(1, len(members))
    out = np.zeros(count.value * per)
    if count.value:
        _chk(lib().sjfh_h5_read(_b(path), _b(dset), mem, len(members), _dp(out), out.size, C.byref(count)),
             "sjfh_h5_read")
    return out.reshape(-1, per) if members else out


def h5_list(path, group="/"):
    buf = C.create_string_buffer(1 << 20)
    _chk(lib().sjfh_h5_list(_b(path), _b(group), buf, len(buf)), "sjfh_h5_list")
    return [s for s in buf.value.decode().split("\n") if s]
