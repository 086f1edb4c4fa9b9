// capi.cpp -- C entry points of the host layer (sjf_host.h).
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "sjf_host.hpp"

extern "C" {
#include "sjf_host.h"
}

using namespace sjfh;

namespace {

thread_local std::string g_err;

int fail(const std::exception& e, int code = -1) {
    g_err = e.what();
    return code;
}

std::string jstr(const std::string& s, bool present) {
    if (!present) return "\"\"";
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}

std::string jnum(double x) {
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", x);
    return b;
}

int copy_out(const std::string& s, char* out, size_t cap) {
    if (s.size() + 1 > cap) {
        g_err = "output buffer too small";
        return -3;
    }
    std::memcpy(out, s.c_str(), s.size() + 1);
    return 0;
}

// mode 0: parse_conf_file only; 1: parse_args then parse_conf_file; 2: as main() (1 plus a
// second correct_defaults)
Settings load(const char* conf_path, int argc, const char* const* argv, int mode, int* rc) {
    Settings s;
    *rc = 0;
    if (mode >= 1) {
        std::vector<char*> av;
        for (int i = 0; i < argc; ++i) av.push_back(const_cast<char*>(argv[i]));
        int n = argc;
        *rc = parse_args(&s, &n, av.data());
        if (*rc) return s;
    }
    const std::string conf = s.has_conf_fname ? s.conf_fname : std::string(conf_path ? conf_path : "params.conf");
    *rc = parse_conf_file(&s, conf);
    if (*rc) return s;
    if (mode >= 2) correct_defaults(&s);
    return s;
}

}  // namespace

extern "C" {

const char* sjfh_last_error(void) { return g_err.c_str(); }

int sjfh_settings_json(const char* conf_path, int argc, const char* const* argv, int mode, char* out, size_t cap) {
    try {
        int rc = 0;
        Settings s = load(conf_path, argc, argv, mode, &rc);
        std::ostringstream o;
        o << "{\"rc\":" << rc << ",\"n_dims\":" << s.n_dims << ",\"pml_thickness\":" << jnum(s.pml_thickness)
          << ",\"len\":" << jnum(s.len) << ",\"um_scale\":" << jnum(s.um_scale) << ",\"grid_num\":" << s.grid_num
          << ",\"resolution\":" << jnum(s.resolution) << ",\"courant\":" << jnum(s.courant)
          << ",\"ambient_eps\":" << jnum(s.ambient_eps) << ",\"smooth_n\":" << s.smooth_n
          << ",\"smooth_rad\":" << jnum(s.smooth_rad) << ",\"post_source_t\":" << jnum(s.post_source_t)
          << ",\"save_span\":" << s.save_span << ",\"dump_raw\":" << (int)s.dump_raw << ",\"verbosity\":" << s.verbosity
          << ",\"geom_fname\":" << jstr(s.geom_fname, s.has_geom_fname) << ",\"out_dir\":" << jstr(s.out_dir, s.has_out_dir)
          << ",\"conf_fname\":" << jstr(s.conf_fname, s.has_conf_fname) << ",\"user_opts\":" << jstr(s.user_opts, s.has_user_opts)
          << "}";
        return copy_out(o.str(), out, cap);
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int sjfh_scene_in(const char* scene_json, const double* xyz, long n, int* out, int* n_roots) {
    try {
        const Scene sc = Scene::from_json_file(scene_json);
        const size_t R = sc.roots.size();
        if (n_roots) *n_roots = (int)R;
        for (long p = 0; p < n; ++p)
            for (size_t r = 0; r < R; ++r) out[p * R + r] = in(*sc.roots[r].tree, Vec3{xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2]});
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int sjfh_in_bound(const char* scene_json, double def_ret, unsigned smooth_n, double smooth_rad, const double* xyz,
                  long n, double* out) {
    try {
        const Scene sc = Scene::from_json_file(scene_json);
        MaterialFunction f(def_ret, smooth_n, smooth_rad);
        for (const Root& r : sc.roots) {
            double scale = def_ret;
            auto it = r.metadata.find("eps");
            if (it != r.metadata.end() && it->second.type == Value::Num) scale = it->second.x;
            f.add_region(r.tree.get(), scale);
        }
        for (long p = 0; p < n; ++p) out[p] = f.in_bound(Vec3{xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2]});
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int sjfh_smooth_points(unsigned smooth_n, double smooth_rad, double* out, size_t cap, unsigned* n_points) {
    MaterialFunction f(1.0, smooth_n, smooth_rad);
    const std::vector<double>& p = f.smooth_points();
    if (n_points) *n_points = (unsigned)(p.size() / 3);
    if (p.size() > cap) {
        g_err = "output buffer too small";
        return -3;
    }
    std::memcpy(out, p.data(), sizeof(double) * p.size());
    return 0;
}

int sjfh_susceptibilities(const char* scene_json, int root, double* out, int cap, int* n_poles, int* er) {
    try {
        const Scene sc = Scene::from_json_file(scene_json);
        if (root < 0 || root >= (int)sc.roots.size()) {
            g_err = "root index out of range";
            return -2;
        }
        Value v;
        auto it = sc.roots[root].metadata.find("susceptibilities");
        if (it != sc.roots[root].metadata.end()) v = it->second;
        int e = 0;
        const std::vector<Suscept> p = parse_susceptibilities(v, &e);
        if (er) *er = e;
        if (n_poles) *n_poles = (int)p.size();
        if ((int)p.size() * 4 > cap) {
            g_err = "output buffer too small";
            return -3;
        }
        for (size_t q = 0; q < p.size(); ++q) {
            out[4 * q] = p[q].omega_0;
            out[4 * q + 1] = p[q].gamma;
            out[4 * q + 2] = p[q].sigma;
            out[4 * q + 3] = p[q].use_denom ? 1.0 : 0.0;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int sjfh_source_info(const char* scene_json, int s, double um_scale, double* out, double* pulse) {
    try {
        const Scene sc = Scene::from_json_file(scene_json);
        if (s < 0 || s >= (int)sc.sources.size()) {
            g_err = "source index out of range";
            return -2;
        }
        const SourceInfo si = source_info(sc.sources[s]);
        const double v[8] = {si.gaussian ? 1.0 : 0.0, (double)si.component, si.wavelen, si.width,
                             si.phase, si.start_time, si.end_time, si.amplitude};
        std::memcpy(out, v, sizeof v);
        if (pulse) {
            const sjf_pulse p = pulse_of(si, um_scale);
            const double w[7] = {(double)p.kind, p.freq, p.width, p.phase, p.t_start, p.t_end, p.slowness};
            std::memcpy(pulse, w, sizeof w);
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int sjfh_fft(const double* in, long n, double* out, long cap, long* n_out) {
    std::vector<std::complex<double>> x((size_t)n);
    for (long i = 0; i < n; ++i) x[i] = {in[2 * i], in[2 * i + 1]};
    const std::vector<std::complex<double>> y = fft(x);
    if (n_out) *n_out = (long)y.size();
    if ((long)y.size() * 2 > cap) {
        g_err = "output buffer too small";
        return -3;
    }
    for (size_t k = 0; k < y.size(); ++k) {
        out[2 * k] = y[k].real();
        out[2 * k + 1] = y[k].imag();
    }
    return 0;
}

int sjfh_voxelize(const char* scene_json, double ambient_eps, unsigned smooth_n, double smooth_rad, double um_scale,
                  double resolution, int n, double dx, unsigned* words, double* mats, int cap_mat, int* n_mat,
                  double* eps_map, int* has_map) {
    try {
        const Scene sc = Scene::from_json_file(scene_json);
        Settings s;
        s.ambient_eps = ambient_eps;
        s.smooth_n = smooth_n;
        s.smooth_rad = smooth_rad;
        s.um_scale = um_scale;
        s.resolution = resolution;
        const int n3[3] = {n, n, n};
        const VoxelMaps vm = voxelize(sc, s, n3, dx);
        if (n_mat) *n_mat = (int)vm.lut.size();
        if ((int)vm.lut.size() > cap_mat) {
            g_err = "material buffer too small";
            return -3;
        }
        std::memcpy(words, vm.words.data(), sizeof(unsigned) * vm.words.size());
        if (has_map) *has_map = vm.eps_map.empty() ? 0 : 1;
        if (eps_map && !vm.eps_map.empty()) std::memcpy(eps_map, vm.eps_map.data(), sizeof(double) * vm.eps_map.size());
        for (size_t m = 0; m < vm.lut.size(); ++m) {
            double* d = mats + 26 * m;
            const sjf_material& M = vm.lut[m];
            d[0] = M.eps_inf;
            d[1] = M.npoles;
            for (int q = 0; q < 4; ++q) {
                d[2 + q] = M.f0[q];
                d[6 + q] = M.gamma[q];
                d[10 + q] = M.sigma[q];
                d[14 + q] = M.is_drude[q];
            }
            for (int q = 18; q < 26; ++q) d[q] = 0;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int sjfh_run(const char* conf_path, int argc, const char* const* argv, const char* scene_json, int precision,
             int complex_fields, int device, double* summary) {
    try {
        int rc = 0;
        const Settings s = load(conf_path, argc, argv, 2, &rc);
        if (rc) {
            g_err = "settings could not be read";
            return -2;
        }
        const Scene sc = load_scene(scene_json, s);
        RunOptions opt;
        opt.precision = precision;
        opt.complex_fields = complex_fields != 0;
        opt.device = device;
        BoundGeom g(s, sc, opt);
        g.run(s.out_dir);
        g.save_field_times(s.out_dir);
        if (summary) {
            summary[0] = (double)g.grid_n();
            summary[1] = g.dt();
            summary[2] = g.ttot();
            summary[3] = (double)g.n_t_pts();
            summary[4] = (double)g.get_monitor_locs().size();
            summary[5] = (double)g.get_n_monitor_clusters();
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int sjfh_geom_scene_json(const char* geom_path, double pml_thickness, double len, double um_scale, const char* out_dir,
                         const char* user_opts, char* out, size_t cap, int* parse_error) {
    try {
        Settings s;
        s.pml_thickness = pml_thickness;
        s.len = len;
        s.um_scale = um_scale;
        s.out_dir = out_dir ? out_dir : "";
        s.user_opts = user_opts ? user_opts : "";
        int er = 0;
        const Scene sc = Scene::from_geom_file(geom_path, s, &er);
        if (parse_error) *parse_error = er;
        return copy_out(sc.to_json(), out, cap);
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int sjfh_write_field_samples(const char* path, const char* scene_json, double um_scale, double ttot,
                             unsigned n_t_pts, unsigned save_span, const double* series, long n_samples) {
    try {
        const Scene sc = Scene::from_json_file(scene_json);
        std::vector<SourceInfo> sources;
        for (const SourceSpec& sp : sc.sources)
            if (sp.has_box) sources.push_back(source_info(sp));
        std::vector<Vec3> locs;
        std::vector<size_t> clusters;
        for (const auto& cl : sc.monitor_clusters) {
            for (const Vec3& v : cl) locs.push_back(v);
            clusters.push_back(locs.size());
        }
        std::vector<std::vector<std::complex<double>>> ser(locs.size());
        for (size_t j = 0; j < locs.size(); ++j)
            for (long k = 0; k < n_samples; ++k)
                ser[j].emplace_back(series[(j * n_samples + k) * 2], series[(j * n_samples + k) * 2 + 1]);
        write_field_samples(path, ttot / (0.299792458 * um_scale), n_t_pts, save_span ? save_span : 1, sources,
                            sc.context_numbers, locs, clusters, ser);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

}  // extern "C"
