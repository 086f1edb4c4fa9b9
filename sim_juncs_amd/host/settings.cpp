// settings.cpp -- configuration parsing of the reference driver (src/argparse.h:120-383).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sjf_host.hpp"

namespace sjfh {

namespace {

// trim_whitespace (src/argparse.h:102-118): drop leading blanks, cut after the last non-blank.
std::string trim(const std::string& v) {
    size_t b = 0;
    while (b < v.size() && (v[b] == ' ' || v[b] == '\t' || v[b] == '\n')) ++b;
    size_t e = v.size();
    while (e > b && (v[e - 1] == ' ' || v[e - 1] == '\t' || v[e - 1] == '\n')) --e;
    return v.substr(b, e - b);
}

// handle_pair (src/argparse.h:123-169).  Keys whose value is "not given" while negative (or
// unset for strings) are only taken from the file when the command line did not set them.
// The reference trims trailing blanks of the token only: its value trim starts at the
// terminator it has just written, so a value keeps its trailing blanks.
void handle_pair(Settings* s, std::string tok, const std::string& val) {
    while (!tok.empty() && (tok.back() == ' ' || tok.back() == '\t')) tok.pop_back();
    const char* v = val.c_str();
    if (tok == "pml_thickness" && s->pml_thickness < 0) {
        s->pml_thickness = std::strtod(v, nullptr);
    } else if (tok == "resolution" && s->resolution < 0) {
        s->resolution = std::strtod(v, nullptr);
    } else if (tok == "dimensions" && s->n_dims < 0) {
        s->n_dims = (int)std::strtol(v, nullptr, 10);
    } else if (tok == "um_scale") {
        s->um_scale = std::strtod(v, nullptr);
    } else if (tok == "length" && s->len < 0) {
        s->len = std::strtod(v, nullptr);
    } else if (tok == "smooth_rad") {
        s->smooth_rad = std::strtod(v, nullptr);
    } else if (tok == "smooth_n") {
        s->smooth_n = (unsigned)std::strtol(v, nullptr, 10);
    } else if (tok == "courant") {
        s->courant = std::strtod(v, nullptr);
    } else if (tok == "ambient_eps" && s->ambient_eps < 0) {
        s->ambient_eps = std::strtod(v, nullptr);
    } else if (tok == "geom_fname" && !s->has_geom_fname) {
        s->geom_fname = trim(val);
        s->has_geom_fname = true;
    } else if (tok == "out_dir" && !s->has_out_dir) {
        s->out_dir = trim(val);
        s->has_out_dir = true;
    } else if (tok == "post_source_t") {
        s->post_source_t = std::strtod(v, nullptr);
    } else if (tok == "save_span") {
        s->save_span = (unsigned)std::strtol(v, nullptr, 10);
    } else if (tok == "dump_raw") {
        if (val == "true") s->dump_raw = 1;
        else if (val == "false") s->dump_raw = 0;
        else s->dump_raw = (unsigned char)std::strtol(v, nullptr, 10);
    }
}

}  // namespace

void correct_defaults(Settings* s) {
    const double total_len = s->len + 2 * s->pml_thickness;
    if (s->grid_num < 0) s->grid_num = 2 * (long)(int)(0.5 * (1 + total_len * s->resolution)) + 1;
    s->n_dims = std::abs(s->n_dims);
    s->pml_thickness = std::fabs(s->pml_thickness);
    s->len = std::fabs(s->len);
    s->resolution = std::fabs(s->resolution);
    s->ambient_eps = std::fabs(s->ambient_eps);
    s->resolution = (double)s->grid_num / total_len;
    if (!s->has_out_dir) {
        s->out_dir = "/tmp";
        s->has_out_dir = true;
    }
}

// parse_conf_file (src/argparse.h:196-241): ini-style lines; a leading [section] is skipped;
// "key = value" pairs end at a newline or ';' (several per line; a key after ';' keeps its
// leading blanks); '#' ends the scan of a line, and a pair without its terminator is dropped.
int parse_conf_file(Settings* s, const std::string& fname) {
    FILE* fp = std::fopen(fname.c_str(), "r");
    if (!fp) return -1;
    char buf[1024];
    while (std::fgets(buf, sizeof buf, fp)) {
        size_t i = 0;
        while (i < sizeof buf && (buf[i] == ' ' || buf[i] == '\t')) ++i;
        if (buf[i] == '[') {
            while (i < sizeof buf && buf[i] != ']' && buf[i]) ++i;
            if (buf[i]) ++i;
        }
        size_t tok = i;
        long eq = -1;
        for (; i < sizeof buf - 1 && buf[i] && buf[i] != '#'; ++i) {
            if (eq < 0) {
                if (buf[i] == '=') eq = (long)i;
            } else if (buf[i] == '\n' || buf[i] == ';') {
                handle_pair(s, std::string(buf + tok, (size_t)eq - tok), std::string(buf + eq + 1, i - (size_t)eq - 1));
                tok = i + 1;
                eq = -1;
            }
        }
    }
    std::fclose(fp);
    correct_defaults(s);
    return 0;
}

// parse_args (src/argparse.h:244-383): flags are matched by prefix and removed with their
// value; a flag given as the last argument prints its usage and stops the scan.
int parse_args(Settings* a, int* argc, char** argv) {
    unsigned n = (unsigned)*argc;
    auto starts = [](const char* s, const char* p) { return std::strncmp(s, p, std::strlen(p)) == 0; };
    for (unsigned i = 0; i < n; ++i) {
        unsigned to_rm = 0;
        const char* arg = argv[i];
        const bool last = (i == n - 1);
        auto number = [&](double* dst, bool integer) -> int {
            errno = 0;
            if (integer) *dst = (double)std::strtol(argv[i + 1], nullptr, 10);
            else *dst = std::strtod(argv[i + 1], nullptr);
            return errno;
        };
        if (starts(arg, "--conf-file")) {
            if (last) { std::printf("Usage: sim_geom --conf-file file name"); return 0; }
            a->conf_fname = argv[i + 1];
            a->has_conf_fname = true;
            to_rm = 2;
        } else if (starts(arg, "--geom-file")) {
            if (last) { std::printf("Usage: sim_geom --out-dir <prefix to output hdf5 files to>"); return 0; }
            a->geom_fname = trim(argv[i + 1]);
            a->has_geom_fname = true;
            to_rm = 2;
        } else if (starts(arg, "--out-dir")) {
            if (last) { std::printf("Usage: sim_geom --out-dir <prefix to output hdf5 files to>"); return 0; }
            a->out_dir = trim(argv[i + 1]);
            a->has_out_dir = true;
            to_rm = 2;
        } else if (starts(arg, "--grid-res")) {
            if (last) { std::printf("Usage: sim_geom --grid-res <grid points per unit length>"); return 0; }
            if (int e = number(&a->resolution, false)) { std::printf("Invalid floating point supplied to --grid-res"); return e; }
            to_rm = 2;
        } else if (starts(arg, "--grid-num")) {
            if (last) { std::printf("Usage: sim_geom --grid-num total number of grid points to use"); return 0; }
            double v = 0;
            if (int e = number(&v, true)) { std::printf("Invalid floating point supplied to --grid-num"); return e; }
            a->grid_num = (long)v;
            to_rm = 2;
        } else if (starts(arg, "--length")) {
            if (last) { std::printf("Usage: sim_geom --length <length of each axis of the simulation volume in arbitrary units>"); return 0; }
            if (int e = number(&a->len, false)) { std::printf("Invalid floating point supplied to --length"); return e; }
            to_rm = 2;
        } else if (starts(arg, "--eps1")) {
            if (last) { std::printf("Usage: sim_geom --eps1 <epsilon1>"); return 0; }
            if (int e = number(&a->ambient_eps, false)) { std::printf("Invalid floating point supplied to --eps1"); return e; }
            to_rm = 2;
        } else if (starts(arg, "--save-span")) {
            if (last) { std::printf("Usage: sim_geom --save-span <epsilon1>"); return 0; }
            double v = 0;
            if (int e = number(&v, true)) { std::printf("Invalid floating point supplied to --eps1"); return e; }
            a->save_span = (unsigned)(long)v;
            to_rm = 2;
        } else if (starts(arg, "-v")) {
            if (last) { std::printf("Usage: sim_geom -v <verbosity (integer)>"); return 0; }
            double v = 0;
            if (int e = number(&v, true)) { std::printf("Invalid value supplied to -v, should be integer"); return e; }
            a->verbosity = (unsigned)(long)v;
            to_rm = 2;
        } else if (starts(arg, "--opts")) {
            if (last) { std::printf("Usage: sim_geom --opts \"<name1=value1;name2=value2>\""); return 0; }
            a->user_opts = argv[i + 1];
            a->has_user_opts = true;
            to_rm = 2;
        }
        if (to_rm) {
            unsigned j = i;
            for (; j + to_rm < n; ++j) argv[j] = argv[j + to_rm];
            for (; j < n; ++j) argv[j] = nullptr;
            n -= to_rm;
            --i;
        }
    }
    *argc = (int)n;
    if (!a->has_out_dir) {
        a->out_dir = "/tmp";
        a->has_out_dir = true;
    }
    return 0;
}

}  // namespace sjfh
