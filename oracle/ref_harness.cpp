// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// Links the reference's meep-free translation units (cgs.cpp, cgs_read.cpp, cgs_data.cpp,
// data_utils.cpp, header-only argparse.h), compiled in place from the reference tree by
// oracle/build_ref.py, and prints what they compute as JSON so that tests can pin the host
// layer's restatement against the reference itself:
//
//   ref_harness scene <conf> <out.json> [--opts "a=1;b=2"] [--grid-res R] [--npts N] [--seed S]
//                     [--lo A --hi B]   (points uniform in [A,B]^3; default the domain +- 5%)
//                     [--post len=2,pml_thickness=1,resolution=5]   (settings changed after loading)
//       parse_args/parse_conf_file/correct_defaults (src/main.cpp:13-36), the scene built by
//       the reference parser with the settings context (src/disp.cpp:28-51, restated below
//       because disp.cpp needs meep), and for every root the resolved CSG tree (node kinds,
//       inversion flags, transforms, primitive parameters), its metadata, the sources and
//       monitor clusters walked in the order of src/disp.cpp:586-639, the numeric context
//       entries in the order of src/disp.cpp:863-872, and in() of every root at N seeded points.
//   ref_harness conf <conf> <out.json> [cli args ...]    parsed settings
//   ref_harness fft <in.txt> <out.json>                  fft() (src/data_utils.cpp:413-427)
//
// The reference's scene classes keep the tree state protected; the harness opens them up
// with access macros so it can read them without modifying the reference.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>
#include <limits>
#include <iostream>
#include <fstream>
#include <sstream>
#include <algorithm>
#include <complex>
#include <chrono>
#include <functional>
#include <memory>
#include <map>

#define private public
#define protected public
#include "cgs.hpp"
#include "data_utils.hpp"
#include "argparse.h"
#undef private
#undef protected

namespace {

// ---- restated from src/disp.cpp:5-51 (meep-dependent translation unit) ------------------
value h_um_to_l(context& c, cgs_func f, parse_ercode& er) {
    value ret;
    if (f.n_args < 1) { er = E_LACK_TOKENS; return ret; }
    if (f.args[0].type != VAL_NUM) { er = E_BAD_TOKEN; return ret; }
    value l_per_um = c.lookup("l_per_um");
    if (l_per_um.type != VAL_NUM) { er = E_NOT_DEFINED; return ret; }
    return make_val_num(l_per_um.val.x * f.args[0].val.x);
}
value h_fs_to_t(context& c, cgs_func f, parse_ercode& er) {
    value ret;
    if (f.n_args < 1) { er = E_LACK_TOKENS; return ret; }
    if (f.args[0].type != VAL_NUM) { er = E_BAD_TOKEN; return ret; }
    value l_per_um = c.lookup("l_per_um");
    if (l_per_um.type != VAL_NUM) { er = E_NOT_DEFINED; return ret; }
    return make_val_num(0.299792458 * l_per_um.val.x * f.args[0].val.x);
}
context settings_context(const parse_settings& a) {
    context con;
    con.emplace("pi", make_val_num(M_PI));
    con.emplace("pml_thickness", make_val_num(a.pml_thickness));
    con.emplace("sim_length", make_val_num(a.len));
    con.emplace("length", make_val_num(2 * a.pml_thickness + a.len));
    con.emplace("l_per_um", make_val_num(a.um_scale));
    value o = make_val_str(a.out_dir);
    con.emplace("out_dir", o);
    cleanup_val(&o);
    value f = make_val_func("um_to_l", 1, &h_um_to_l);
    con.emplace("um_to_l", f);
    cleanup_val(&f);
    f = make_val_func("fs_to_t", 1, &h_fs_to_t);
    con.emplace("fs_to_t", f);
    cleanup_val(&f);
    if (a.user_opts) {
        line_buffer lb(a.user_opts, ';');
        con.read_from_lines(lb);
    }
    return con;
}

// ---- JSON output --------------------------------------------------------------------------
std::string num(double x) {
    if (std::isinf(x)) return x > 0 ? "1e999" : "-1e999";
    if (std::isnan(x)) return "null";
    char b[64];
    snprintf(b, sizeof b, "%.17g", x);
    return b;
}
std::string str(const char* s) {
    std::string o = "\"";
    for (const char* p = s ? s : ""; *p; ++p) {
        if (*p == '"' || *p == '\\') o += '\\';
        if ((unsigned char)*p < 0x20) { o += ' '; continue; }
        o += *p;
    }
    return o + "\"";
}
std::string vec(vec3 v) { return "[" + num(v.x()) + "," + num(v.y()) + "," + num(v.z()) + "]"; }
std::string mat(const mat3x3& m) {
    std::string o = "[";
    for (int i = 0; i < 9; ++i) o += (i ? "," : "") + num(m.el[i]);
    return o + "]";
}

std::string dump_value(const value& v) {
    switch (v.type) {
        case VAL_NUM: return num(v.val.x);
        case VAL_STR: return str(v.val.s);
        case VAL_3VEC: return vec(*v.val.v);
        case VAL_LIST: {
            std::string o = "[";
            for (size_t i = 0; i < v.n_els; ++i) o += (i ? "," : "") + dump_value(v.val.l[i]);
            return o + "]";
        }
        default: return "null";
    }
}

const char* kind_name(object_type t) {
    switch (t) {
        case CGS_COMPOSITE: return "composite";
        case CGS_SPHERE: return "sphere";
        case CGS_BOX: return "box";
        case CGS_PLANE: return "plane";
        case CGS_CYLINDER: return "cylinder";
        default: return "undef";
    }
}

std::string dump_node(object* o, object_type t) {
    if (!o) return "null";
    std::string s = "{\"kind\":\"" + std::string(kind_name(t)) + "\",\"invert\":" + std::to_string(o->invert) +
                    ",\"trans\":" + mat(o->trans_mat);
    switch (t) {
        case CGS_COMPOSITE: {
            composite_object* c = (composite_object*)o;
            const char* cm = c->cmb == CGS_UNION ? "union" : c->cmb == CGS_INTERSECT ? "intersect"
                           : c->cmb == CGS_DIFFERENCE ? "difference" : "noop";
            s += ",\"combine\":\"" + std::string(cm) + "\",\"children\":[" + dump_node(c->children[0], c->child_types[0]) +
                 "," + dump_node(c->children[1], c->child_types[1]) + "]";
            break;
        }
        case CGS_SPHERE: {
            sphere* p = (sphere*)o;
            s += ",\"center\":" + vec(p->center) + ",\"radius\":" + num(p->rad);
            break;
        }
        case CGS_BOX: {
            box* p = (box*)o;
            s += ",\"center\":" + vec(p->center) + ",\"offset\":" + vec(p->offset);
            break;
        }
        case CGS_PLANE: {
            plane* p = (plane*)o;
            s += ",\"normal\":" + vec(p->normal) + ",\"offset\":" + num(p->offset);
            break;
        }
        case CGS_CYLINDER: {
            cylinder* p = (cylinder*)o;
            s += ",\"center\":" + vec(p->center) + ",\"height\":" + num(p->height) + ",\"r1_sq\":" + num(p->r1_sq) +
                 ",\"r1_sq_x_h\":" + num(p->r1_sq_x_h) + ",\"r2_sq\":" + num(p->r2_sq);
            break;
        }
        default: break;
    }
    return s + "}";
}

std::string dump_settings(const parse_settings& a) {
    std::ostringstream o;
    o << "{\"n_dims\":" << a.n_dims << ",\"pml_thickness\":" << num(a.pml_thickness) << ",\"len\":" << num(a.len)
      << ",\"um_scale\":" << num(a.um_scale) << ",\"grid_num\":" << a.grid_num << ",\"resolution\":" << num(a.resolution)
      << ",\"courant\":" << num(a.courant) << ",\"ambient_eps\":" << num(a.ambient_eps) << ",\"smooth_n\":" << a.smooth_n
      << ",\"smooth_rad\":" << num(a.smooth_rad) << ",\"post_source_t\":" << num(a.post_source_t)
      << ",\"save_span\":" << a.save_span << ",\"dump_raw\":" << (int)a.dump_raw << ",\"verbosity\":" << a.verbosity
      << ",\"geom_fname\":" << str(a.geom_fname) << ",\"out_dir\":" << str(a.out_dir)
      << ",\"conf_fname\":" << str(a.conf_fname) << ",\"user_opts\":" << str(a.user_opts) << "}";
    return o.str();
}

// parse_args + parse_conf_file + correct_defaults as src/main.cpp:16-36 does
int load_settings(parse_settings* a, const char* conf, std::vector<std::string> cli, int* left) {
    std::vector<std::string> argv_s = {"sim_geom", "--conf-file", conf};
    for (auto& x : cli) argv_s.push_back(x);
    std::vector<char*> argv;
    for (auto& x : argv_s) argv.push_back(strdup(x.c_str()));
    int argc = (int)argv.size();
    int ret = parse_args(a, &argc, argv.data());
    if (left) *left = argc;
    if (ret) return ret;
    ret = parse_conf_file(a, a->conf_fname);
    if (ret) return ret;
    correct_defaults(a);
    return 0;
}

int cmd_conf(int argc, char** argv) {
    if (argc < 4) return 2;
    std::vector<std::string> cli(argv + 4, argv + argc);
    parse_settings a;
    int left = 0;
    int rc = load_settings(&a, argv[2], cli, &left);
    std::ofstream f(argv[3]);
    f << "{\"rc\":" << rc << ",\"argc_left\":" << left << ",\"settings\":" << dump_settings(a) << "}\n";
    return 0;
}

int cmd_fft(int argc, char** argv) {
    if (argc < 4) return 2;
    std::ifstream in(argv[2]);
    std::vector<complex> pts;
    double re, im;
    while (in >> re >> im) pts.push_back(complex(re, im));
    data_arr d(pts);
    data_arr r = fft(d);
    std::ofstream f(argv[3]);
    f << "{\"n\":" << pts.size() << ",\"out\":[";
    for (size_t i = 0; i < r.size; ++i) f << (i ? "," : "") << "[" << num(r.buf[i].re) << "," << num(r.buf[i].im) << "]";
    f << "]}\n";
    return 0;
}

int cmd_scene(int argc, char** argv) {
    if (argc < 4) return 2;
    std::vector<std::string> cli;
    long npts = 2000;
    unsigned seed = 1;
    double lo = NAN, hi = NAN;
    std::string post;  // "len=2,pml_thickness=1,resolution=5": set after loading, as a test does
    for (int i = 4; i < argc; ++i) {
        std::string k = argv[i];
        if (k == "--npts" && i + 1 < argc) npts = atol(argv[++i]);
        else if (k == "--lo" && i + 1 < argc) lo = atof(argv[++i]);
        else if (k == "--post" && i + 1 < argc) post = argv[++i];
        else if (k == "--hi" && i + 1 < argc) hi = atof(argv[++i]);
        else if (k == "--seed" && i + 1 < argc) seed = (unsigned)atol(argv[++i]);
        else cli.push_back(k);
    }
    parse_settings a;
    int rc = load_settings(&a, argv[2], cli, nullptr);
    if (rc) return 3;
    {
        std::stringstream ps(post);
        std::string kv;
        while (std::getline(ps, kv, ',')) {
            const size_t eq = kv.find('=');
            if (eq == std::string::npos) continue;
            const std::string k = kv.substr(0, eq);
            const double v = atof(kv.c_str() + eq + 1);
            if (k == "len") a.len = v;
            else if (k == "pml_thickness") a.pml_thickness = v;
            else if (k == "resolution") a.resolution = v;
        }
    }
    parse_ercode er = E_SUCCESS;
    context con = settings_context(a);
    scene s(a.geom_fname, con, &er);
    std::ostringstream o;
    o << "{\"parse_error\":" << (int)er << ",\"settings\":" << dump_settings(a) << ",\"roots\":[";
    std::vector<composite_object*> roots = s.get_roots();
    for (size_t r = 0; r < roots.size(); ++r) {
        composite_object* c = roots[r];
        o << (r ? "," : "") << "{\"metadata\":{";
        bool first = true;
        for (auto& kv : c->metadata) {
            o << (first ? "" : ",") << str(kv.first.c_str()) << ":" << dump_value(kv.second);
            first = false;
        }
        o << "},\"tree\":" << dump_node(c, CGS_COMPOSITE) << "}";
    }
    o << "],\"sources\":[";
    context& c = s.get_context();
    bool firsts = true;
    std::vector<std::string> clusters;
    for (size_t i = c.size(); i > 0; --i) {
        value inst = c.peek_val(i);
        if (inst.type != VAL_INST) continue;
        value ty = inst.val.c->lookup("__type__");
        if (ty.type != VAL_STR) continue;
        if (!strcmp(ty.val.s, "Gaussian_source") || !strcmp(ty.val.s, "CW_source")) {
            o << (firsts ? "" : ",") << "{\"type\":" << str(ty.val.s);
            firsts = false;
            for (const char* key : {"component", "wavelength", "amplitude", "width", "phase", "cutoff", "start_time",
                                    "end_time", "slowness"}) {
                value v = inst.val.c->lookup(key);
                if (v.type == VAL_NUM) o << "," << str(key) << ":" << num(v.val.x);
            }
            value reg = inst.val.c->lookup("region");
            if (is_type(reg, "Box")) {
                parse_ercode e1 = E_SUCCESS, e2 = E_SUCCESS;
                value p1 = reg.val.c->lookup("pt_1").cast_to(VAL_3VEC, e1);
                value p2 = reg.val.c->lookup("pt_2").cast_to(VAL_3VEC, e2);
                if (e1 == E_SUCCESS && e2 == E_SUCCESS) o << ",\"box\":[" << vec(*p1.val.v) << "," << vec(*p2.val.v) << "]";
                cleanup_val(&p1);
                cleanup_val(&p2);
            }
            o << "}";
        } else if (!strcmp(ty.val.s, "monitor")) {
            value vl = inst.val.c->lookup("locations");
            std::string cl = "[";
            if (vl.type == VAL_LIST) {
                for (size_t k = 0; k < vl.n_els; ++k) {
                    parse_ercode e = E_SUCCESS;
                    value v = vl.val.l[k].cast_to(VAL_3VEC, e);
                    if (e == E_SUCCESS) cl += (cl.size() > 1 ? "," : "") + vec(*v.val.v);
                    cleanup_val(&v);
                }
            }
            clusters.push_back(cl + "]");
        }
    }
    o << "],\"monitor_clusters\":[";
    for (size_t k = 0; k < clusters.size(); ++k) o << (k ? "," : "") << clusters[k];
    o << "],\"context_numbers\":[";
    bool firstn = true;
    for (size_t i = c.size(); i > 0; --i) {
        name_val_pair nv = c.peek(i);
        if (nv.get_val().type == VAL_NUM) {
            o << (firstn ? "" : ",") << "[" << str(nv.get_name()) << "," << num(nv.get_val().val.x) << "]";
            firstn = false;
        }
    }
    // in() of every root at seeded points spanning the domain plus a margin
    const double L = 2 * a.pml_thickness + a.len;
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(std::isnan(lo) ? -0.05 * L : lo, std::isnan(hi) ? 1.05 * L : hi);
    o << "],\"points\":[";
    std::string ins = "[";
    for (long p = 0; p < npts; ++p) {
        const double x = u(gen), y = u(gen), z = u(gen);
        o << (p ? "," : "") << "[" << num(x) << "," << num(y) << "," << num(z) << "]";
        ins += p ? ",[" : "[";
        for (size_t r = 0; r < roots.size(); ++r) ins += (r ? "," : "") + std::to_string(roots[r]->in(vec3(x, y, z)));
        ins += "]";
    }
    o << "],\"inside\":" << ins << "]}";
    std::ofstream f(argv[3]);
    f << o.str() << "\n";
    return er == E_SUCCESS ? 0 : 4;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::string cmd = argv[1];
    if (cmd == "scene") return cmd_scene(argc, argv);
    if (cmd == "conf") return cmd_conf(argc, argv);
    if (cmd == "fft") return cmd_fft(argc, argv);
    return 2;
}
