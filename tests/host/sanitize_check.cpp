// sanitize_check.cpp -- test program built with AddressSanitizer and UndefinedBehaviorSanitizer
// over the host layer (config parsing, the .geom loader, voxelisation with and without
// smoothing, the fft, the field_samples.h5 writer) and the CPU oracle (a small run with PML and
// its Lorentz fold, Drude/Lorentz media, a per-cell eps map, a design box, Gaussian, CW and
// magnetic sources, monitors and DFT regions).  Any sanitizer report aborts the program; it
// prints one line per part and "all parts ran" at the end.
//
//   sanitize_check <golden dir> <scratch dir>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

#include "sjf_host.hpp"

extern "C" {
typedef struct orc orc;
orc* orc_create(int nx, int ny, int nz, double dx, double courant, const int* bc, int part);
void orc_destroy(orc* o);
int orc_set_materials(orc* o, const uint32_t* matw, int nmat, const double* eps_inf, const int* npole,
                      const double* poles);
int orc_set_eps_maps(orc* o, const double* eps);
int orc_set_eps_box(orc* o, const int* lo, const int* hi, const double* eps);
int orc_set_pml(orc* o, const int* cells, double R, double power);
int orc_set_pml_dispersion(orc* o, double f);
int orc_add_source(orc* o, int comp, const double* box, int kind, const double* prm, double amp_re, double amp_im);
int orc_add_monitors(orc* o, int comp, const double* xyz, int npt);
void orc_set_save_span(orc* o, int span);
long orc_num_samples(const orc* o);
void orc_read_monitors(const orc* o, double* out);
int orc_add_dft(orc* o, const int* lo, const int* hi, int mask, const double* freqs, int nf);
int orc_read_dft(const orc* o, int id, double* out);
void orc_get_field(const orc* o, int comp, double* out);
void orc_step(orc* o, long nsteps);
}

using namespace sjfh;

namespace {

int fails = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAILED %s\n", what);
        ++fails;
    }
}

void settings_part(const std::string& g) {
    for (const char* c : {"run.conf", "test.conf", "span.conf", "box_params.conf", "bowtie_params.conf"}) {
        Settings s;
        check(parse_conf_file(&s, g + "/conf/" + c) == 0, c);
        std::vector<std::string> args = {"sim_geom", "--grid-num", "33", "--length", "3.5", "--opts", "a=1;b=2"};
        std::vector<char*> argv;
        for (auto& a : args) argv.push_back(&a[0]);
        int argc = (int)argv.size();
        check(parse_args(&s, &argc, argv.data()) == 0, "parse_args");
        correct_defaults(&s);
    }
    std::printf("settings ok\n");
}

void geom_part(const std::string& g) {
    Settings s;
    parse_conf_file(&s, g + "/geom/small_run.conf");
    for (const char* f : {"small_run.geom", "quirks.geom"}) {
        int er = 0;
        const Scene sc = Scene::from_geom_file(g + "/geom/" + f, s, &er);
        const std::string js = sc.to_json();
        const Scene back = Scene::from_json_text(js);
        check(back.roots.size() == sc.roots.size(), f);
    }
    int er = 0;
    Scene::from_geom_text("x = [1, 2, 3\nBox(", s, &er);  // a script error must not read out of bounds
    std::printf("geom ok\n");
}

void voxel_part(const std::string& g) {
    for (const char* f : {"scene_test.json", "scene_bowtie.json", "scene_box.json", "scene_sidetie.json"}) {
        const Scene sc = Scene::from_json_file(g + "/" + f);
        Settings s;
        s.ambient_eps = 1.0;
        s.um_scale = 1.0;
        s.resolution = 4.0;
        const int n[3] = {13, 11, 9};
        for (unsigned sm : {0u, 2u}) {
            s.smooth_n = sm;
            s.smooth_rad = 0.1;
            const VoxelMaps vm = voxelize(sc, s, n, 0.25);
            check(vm.words.size() == (size_t)13 * 11 * 9 && !vm.lut.empty(), f);
            check(sm == 0 ? vm.eps_map.empty() : vm.eps_map.size() == 3 * vm.words.size(), "eps map size");
        }
    }
    std::printf("voxelize ok\n");
}

void fft_part() {
    for (size_t n : {1, 2, 3, 7, 8, 100, 349, 1024}) {
        std::vector<std::complex<double>> x(n);
        for (size_t i = 0; i < n; ++i) x[i] = {std::sin(0.3 * i), std::cos(0.7 * i)};
        const auto y = fft(x);
        check(!y.empty() && (y.size() & (y.size() - 1)) == 0, "fft size");
    }
    std::printf("fft ok\n");
}

void h5_part(const std::string& scratch) {
    std::vector<SourceInfo> src(1);
    std::vector<std::pair<std::string, double>> params = {{"l_per_um", 2.0}, {"sim_length", 2.0}};
    std::vector<Vec3> locs = {{1, 1, 1}, {2, 3, 1.5}, {0.5, 0.5, 0.5}};
    std::vector<size_t> clusters = {2, 3};
    std::vector<std::vector<std::complex<double>>> series(3, std::vector<std::complex<double>>(37));
    for (auto& s : series)
        for (size_t i = 0; i < s.size(); ++i) s[i] = {std::cos(0.1 * i), std::sin(0.2 * i)};
    try {
        write_field_samples(scratch + "/field_samples.h5", 12.5, 37, 1, src, params, locs, clusters, series);
        std::printf("h5 ok\n");
    } catch (const std::exception& e) {  // no HDF5 runtime on this host: the writer reports it
        std::printf("h5 skipped (%s)\n", e.what());
    }
}

void oracle_part() {
    const int nx = 10, ny = 9, nz = 16;
    const long N = (long)nx * ny * nz;
    const int bc[3] = {0, 0, 0};
    orc* o = orc_create(nx, ny, nz, 0.02, 0.5, bc, 1);
    std::vector<uint32_t> w(N);
    for (long q = 0; q < N; ++q) w[q] = (uint32_t)((q % 3) | (((q / 3) % 3) << 8) | (((q / 7) % 3) << 16));
    const double eps[3] = {1.0, 2.13, 1.0};
    const int np[3] = {0, 1, 3};
    std::vector<double> poles(3 * 4 * 4, 0.0);
    const double sio2[4] = {9.68, 0.08, 1.12, 0};
    const double au[3][4] = {{1e-10, 0.0427, 4.03e21, 1}, {0.335, 0.194, 11.36, 0}, {0.669, 0.278, 1.18, 0}};
    for (int q = 0; q < 4; ++q) poles[(1 * 4 + 0) * 4 + q] = sio2[q];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 4; ++q) poles[(2 * 4 + p) * 4 + q] = au[p][q];
    check(orc_set_materials(o, w.data(), 3, eps, np, poles.data()) == 0, "orc_set_materials");
    std::vector<double> emap(3 * N);
    for (long q = 0; q < 3 * N; ++q) emap[q] = 1.0 + 0.5 * ((q * 7) % 5);
    orc_set_eps_maps(o, emap.data());
    const int lo[3] = {3, 3, 5}, hi[3] = {6, 5, 9};
    std::vector<double> deps(3 * 2 * 4, 2.5);
    orc_set_eps_box(o, lo, hi, deps.data());
    const int cells[6] = {2, 2, 1, 2, 3, 3};
    orc_set_pml(o, cells, 1e-15, 2.0);
    orc_set_pml_dispersion(o, 1.25);
    const double box[6] = {0, 0, 0.2, 0.2, 0.18, 0.2};
    const double gauss[5] = {1.25, 0.2, 0.3, 0.0, 2.4};
    const double cw[5] = {1.25, 0.3, 0.0, 1.5, 3.0};
    orc_add_source(o, 0, box, 0, gauss, 1.0, 0.5);
    orc_add_source(o, 4, box, 1, cw, 0.7, -0.2);
    const double pt[6] = {0.11, 0.07, 0.15, 0.05, 0.05, 0.05};
    orc_add_monitors(o, 0, pt, 2);
    orc_add_monitors(o, 5, pt, 2);
    orc_set_save_span(o, 3);
    const int dlo[3] = {0, 0, 8}, dhi[3] = {nx, ny, 9};
    const double freqs[3] = {1.0, 1.25, 1.5};
    const int id = orc_add_dft(o, dlo, dhi, 0x3f, freqs, 3);
    orc_step(o, 40);
    std::vector<double> mon(orc_num_samples(o) * 4), dft(6 * 3 * nx * ny * 2), f(N);
    orc_read_monitors(o, mon.data());
    orc_read_dft(o, id, dft.data());
    orc_get_field(o, 1, f.data());
    double s = 0;
    for (double v : f) s += v * v;
    check(std::isfinite(s) && s > 0, "oracle fields");
    orc_destroy(o);
    std::printf("oracle ok\n");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string g = argv[1], scratch = argv[2];
    settings_part(g);
    geom_part(g);
    voxel_part(g);
    fft_part();
    h5_part(scratch);
    oracle_part();
    std::printf(fails ? "FAILED\n" : "all parts ran\n");
    return fails ? 1 : 0;
}
