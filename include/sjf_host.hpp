/* sjf_host.hpp -- host layer of the MI355X FDTD engine (C++17).
 *
 * Restates the reference driver's side of the hot path above the C ABI of sjf.h, so a
 * caller of the reference's bound_geom (src/disp.hpp:136-202) finds the same pieces:
 *
 *   Settings            parse_settings, parse_args, parse_conf_file, correct_defaults
 *                       (src/argparse.h:24-241,244-383)
 *   Node / in()         the CSG object tree and its point-inclusion test (src/cgs.hpp:36-166,
 *                       src/cgs.cpp:34-118,422-447), plus builders that assemble a tree from a
 *                       list the way the .geom constructors do (src/cgs.cpp:200-270,
 *                       src/cgs_data.cpp:87-300)
 *   Scene               roots with metadata, sources, monitor clusters and numeric context
 *                       entries: what the reference's scene parser produces (src/cgs.cpp:
 *                       510-575); loaded from JSON (the .geom text interpreter itself is not
 *                       part of this layer)
 *   MaterialFunction    cgs_material_function::in_bound incl. stochastic smoothing
 *                       (src/disp.cpp:56-283)
 *   parse_susceptibilities, source_info (src/disp.cpp:318-376,418-454)
 *   voxelize            per-cell material words + LUT for sjf_set_materials, sampled at the
 *                       E-component Yee points (src/disp.cpp:482-550)
 *   BoundGeom           bound_geom: engine set-up (src/disp.cpp:557-645), run()
 *                       (src/disp.cpp:690-749) and save_field_times() (src/disp.cpp:758-923)
 *   fft                 the reference's fft() with its truncation (src/data_utils.cpp:413-427)
 *
 * Every type here is plain host data; all device work goes through sjf.h.
 */
#ifndef SJF_HOST_HPP
#define SJF_HOST_HPP

#include <complex>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sjf.h"

namespace sjfh {

// ---- settings (src/argparse.h) ---------------------------------------------------------------

/* Field-for-field restatement of parse_settings (src/argparse.h:24-48), defaults included:
 * negative values mark "not given" so the config file fills them and command-line values win. */
struct Settings {
    std::string out_dir, geom_fname, conf_fname, user_opts;
    bool has_out_dir = false, has_geom_fname = false, has_conf_fname = false, has_user_opts = false;
    int n_dims = -3;
    double pml_thickness = -0.3;
    double len = -8.0;
    double um_scale = 1.0;
    long grid_num = -1;
    double resolution = -2.0;
    double courant = 0.5;
    double ambient_eps = -1.0;
    unsigned smooth_n = 0;
    double smooth_rad = 0.05;
    double post_source_t = 10.0;
    unsigned save_span = 20;
    unsigned char dump_raw = 0;
    unsigned verbosity = 1;
};

/* src/argparse.h:244-383: consumes known flags from argv (the rest stay, in order).
 * Returns 0, or errno when a number fails to parse. */
int parse_args(Settings* s, int* argc, char** argv);
/* src/argparse.h:196-241 (calls correct_defaults at the end).  Returns 0, -1 if the file
 * cannot be opened. */
int parse_conf_file(Settings* s, const std::string& fname);
/* src/argparse.h:171-191: odd grid count, absolute values, resolution = grid_num / L_tot. */
void correct_defaults(Settings* s);

// ---- CSG (src/cgs.hpp, src/cgs.cpp) ----------------------------------------------------------

struct Vec3 {
    double x = 0, y = 0, z = 0;
};
struct Mat3 {
    double el[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
};
Mat3 rotation(double theta, Vec3 axis);  // make_rotation (src/geometry.hpp:215-225)

enum class Kind { Composite, Sphere, Box, Plane, Cylinder };
enum class Combine { Union, Intersect, Difference, Noop };

/* A resolved tree node: what composite_object / box / sphere / plane / cylinder hold after the
 * scene is built.  Primitive parameters use the reference's stored forms (box: centre and
 * half-extent; cylinder: r1^2, r1^2 h, r2^2; plane: unit normal and signed offset). */
struct Node {
    Kind kind = Kind::Composite;
    int invert = 0;
    Mat3 trans;
    Vec3 center, offset, normal;
    double radius = 0, plane_offset = 0, height = 0, r1_sq = 0, r1_sq_x_h = 0, r2_sq = 0;
    Combine combine = Combine::Union;
    std::unique_ptr<Node> child[2];
};
/* Point inclusion, 0 or 1, exactly as the reference's in() methods. */
int in(const Node& n, const Vec3& r);

/* Unresolved shape descriptions with the argument conventions of the .geom constructors
 * (src/cgs_data.cpp:87-300).  assemble() turns a list of them into a tree the way
 * composite_object::init_from_list does (src/cgs.cpp:200-270), including how a Complement
 * entry flips the inversion of the entries after it. */
struct Shape {
    std::string type;  // Box Sphere Cylinder Plane Union Intersect Complement Rotate
    Vec3 a, b;         // Box corners; Sphere/Cylinder centre in a; Plane normal in a
    double r = 0, h = 0, r1 = 0, r2 = 0, offset = 0, theta = 0;
    std::vector<Shape> list;
};
Shape Box(Vec3 c1, Vec3 c2);
Shape Sphere(Vec3 center, double radius);
Shape Cylinder(Vec3 center, double h, double r1, double r2);
Shape Cylinder(Vec3 center, double h, double r1);
Shape Plane(Vec3 normal, double offset);
Shape Plane(Vec3 p1, Vec3 p2, Vec3 p3);
Shape Union(std::vector<Shape> l);
Shape Intersect(std::vector<Shape> l);
Shape Complement(std::vector<Shape> l);
Shape Difference(std::vector<Shape> l);  // Intersect([first, Complement(rest)])
Shape Rotate(double theta, Vec3 axis, std::vector<Shape> l);
/* A Composite root: union (or intersection) of the list (src/cgs.cpp:525-535). */
std::unique_ptr<Node> assemble(Combine cmb, const std::vector<Shape>& list);

// ---- scene -----------------------------------------------------------------------------------

/* Metadata value (number, string or list), as the parser's value type holds it. */
struct Value {
    enum Type { Undef, Num, Str, List } type = Undef;
    double x = 0;
    std::string s;
    std::vector<Value> l;
};

struct Root {
    std::map<std::string, Value> metadata;
    std::unique_ptr<Node> tree;
};

/* A Gaussian_source / CW_source instance (src/cgs_data.cpp:22-86): raw fields as given. */
struct SourceSpec {
    bool gaussian = true;
    bool has[9] = {};  // component wavelength amplitude width phase cutoff start_time end_time slowness
    double component = 0, wavelength = 0, amplitude = 0, width = 0, phase = 0, cutoff = 0, start_time = 0,
           end_time = 0, slowness = 0;
    bool has_box = false;
    Vec3 box[2];
};

struct Scene {
    std::vector<Root> roots;                      // in the reference's root order
    std::vector<SourceSpec> sources;              // in bound_geom's walk order
    std::vector<std::vector<Vec3>> monitor_clusters;
    std::vector<std::pair<std::string, double>> context_numbers;  // numeric context entries
    /* Reads a scene in this layer's JSON form (keys roots/tree/metadata, sources,
     * monitor_clusters, context_numbers). */
    static Scene from_json_file(const std::string& path);
    static Scene from_json_text(const std::string& text);
    /* Runs a .geom scene script (src/cgs_read.cpp, src/cgs_data.cpp; scene::read_file,
     * src/cgs.cpp:510-570) in a context holding the settings' values (pi, pml_thickness,
     * sim_length, length, l_per_um, out_dir, um_to_l, fs_to_t: src/disp.cpp:28-50) and the
     * statements of s.user_opts.  parse_error receives the reference's parse_ercode value
     * (0 = success); on an error the statements before it still count (sources, monitors,
     * numbers) but no roots are built, as in the reference. */
    static Scene from_geom_file(const std::string& path, const Settings& s, int* parse_error = nullptr);
    static Scene from_geom_text(const std::string& text, const Settings& s, int* parse_error = nullptr);
    /* This layer's JSON form (the keys from_json_text reads). */
    std::string to_json() const;
};

/* A scene file: path ending in ".json" -> Scene::from_json_file, otherwise a .geom script run
 * with s (Scene::from_geom_file).  A script error throws, as bound_geom stops on one
 * ("Scene parsing failed", src/disp.cpp:562-565). */
Scene load_scene(const std::string& path, const Settings& s);

// ---- materials (src/disp.cpp) ----------------------------------------------------------------

struct Suscept {
    double omega_0 = 0, gamma = 0, sigma = 0;
    bool use_denom = true;  // false: Drude
};
/* src/disp.cpp:418-454: entries [f0, gamma, sigma(, "drude"|"lorentz"|number)]; stops at the
 * first malformed entry (er = -1 not a list / too short, -2 non-numeric), keeping the poles
 * parsed before it. */
std::vector<Suscept> parse_susceptibilities(const Value& v, int* er);

/* cgs_material_function (src/disp.cpp:56-283). */
class MaterialFunction {
public:
    MaterialFunction(double def_ret, unsigned smooth_n, double smooth_rad);
    void add_region(const Node* region, double scale);
    double in_bound(const Vec3& r) const;
    unsigned smooth_n() const { return smooth_n_; }
    const std::vector<double>& smooth_points() const { return pts_; }

private:
    double def_ret_;
    unsigned smooth_n_;
    std::vector<double> pts_;
    std::vector<std::pair<const Node*, double>> regions_;
};

/* Per-cell material words + LUT for sjf_set_materials on an n^3 grid of cell size dx (Yee
 * points: Ex at ((i+1/2)dx, j dx, k dx) etc.).  eps_inf sums over roots (src/disp.cpp:272-
 * 282); each root's poles apply with strength sigma * in_root (src/disp.cpp:529-548). */
struct VoxelMaps {
    int n[3] = {0, 0, 0};
    std::vector<uint32_t> words;
    std::vector<sjf_material> lut;
    // with smoothing (smooth_n > 0) the permittivity is fractional: per-cell eps_inf of Ex, Ey, Ez
    // ([c][cell], for sjf_set_eps_maps) and a LUT of the pole combinations only (eps_inf 1)
    std::vector<double> eps_map;
};
VoxelMaps voxelize(const Scene& scene, const Settings& s, const int n[3], double dx);

// ---- sources (src/disp.cpp:318-376, 597-626) -----------------------------------------------

struct SourceInfo {
    bool gaussian = true;
    int component = 0;  // 0..5 = Ex Ey Ez Hx Hy Hz
    double wavelen = 0.7, width = 1, phase = 0, start_time = 0, end_time = 0, amplitude = 1;
};
SourceInfo source_info(const SourceSpec& sp);
/* Unit conversion of src/disp.cpp:604-609 into an sjf_pulse (meep units). */
sjf_pulse pulse_of(const SourceInfo& si, double um_scale);

// ---- FFT (src/data_utils.cpp:325-427) ----------------------------------------------------------

/* The reference's fft(): the first T = 2^floor(log2 N) samples, transformed with phase
 * 2 pi / N (N the untruncated length), negative frequencies wrapped. */
std::vector<std::complex<double>> fft(const std::vector<std::complex<double>>& x);

// ---- driver (src/disp.hpp:136-202) ------------------------------------------------------------

struct RunOptions {
    int precision = SJF_FP64;   // field storage of the engine(s)
    int device = 0;
    bool complex_fields = true; // two engines (real / imaginary part), as meep's complex run
    int graph_chunk = 32;
    // fold Lorentz poles out of the PML slabs at the first source's frequency
    // (sjf_set_pml_dispersion): a Lorentz substrate reaching the PML otherwise grows without bound
    // in long runs; false keeps the poles there, as meep does
    bool pml_fold = true;
};

class BoundGeom {
public:
    /* bound_geom(parse_settings, ...) (src/disp.cpp:557-645): voxelises the scene, creates the
     * engine(s) on an odd cubic grid of grid_num cells (dx = 1/resolution, Courant 0.5),
     * adds the PML, every source and every monitor cluster. */
    BoundGeom(const Settings& s, const Scene& scene, const RunOptions& opt = RunOptions());
    ~BoundGeom();
    BoundGeom(const BoundGeom&) = delete;
    BoundGeom& operator=(const BoundGeom&) = delete;

    /* run(fname_prefix) (src/disp.cpp:690-749): n_t_pts = (ttot + dt/2)/dt steps, monitors
     * sampled (Ex, before the step) whenever the step index is a multiple of save_span. */
    void run(const std::string& out_dir);
    /* save_field_times (src/disp.cpp:758-923): out_dir/field_samples.h5. */
    void save_field_times(const std::string& out_dir);

    std::vector<Vec3> get_monitor_locs() const { return monitor_locs_; }
    std::vector<std::vector<std::complex<double>>> get_field_times() const { return field_times_; }
    std::vector<SourceInfo> get_sources() const { return sources_; }
    size_t get_n_monitor_clusters() const { return monitor_clusters_.size(); }
    double ttot() const { return ttot_; }
    double dt() const { return dt_; }
    unsigned n_t_pts() const { return n_t_pts_; }
    int grid_n() const { return n_; }
    sjf_engine* engine(int part = 0) const { return eng_[part]; }
    double fs_to_meep_time(double t) const { return t * 0.299792458 * um_scale_; }
    double meep_time_to_fs(double t) const { return t / (0.299792458 * um_scale_); }

private:
    void check(int rc, const char* what, int part = 0) const;
    const Scene* scene_;
    Settings s_;
    RunOptions opt_;
    sjf_engine* eng_[2] = {nullptr, nullptr};
    int nparts_ = 1, n_ = 0;
    double um_scale_ = 1, dt_ = 0, ttot_ = 0, post_source_t_ = 0;
    unsigned n_t_pts_ = 0, save_span_ = 20;
    std::vector<SourceInfo> sources_;
    std::vector<double> eps_cell_;  // mean eps_inf of each cell's three Yee points (dielectric dump)
    std::vector<Vec3> monitor_locs_;
    std::vector<size_t> monitor_clusters_;
    std::vector<std::vector<std::complex<double>>> field_times_;
};

/* field_samples.h5 writer (layout of src/disp.cpp:758-923); used by BoundGeom. */
void write_field_samples(const std::string& path, double ttot_fs, unsigned n_t_pts, unsigned save_span,
                         const std::vector<SourceInfo>& sources,
                         const std::vector<std::pair<std::string, double>>& cgs_params,
                         const std::vector<Vec3>& locs, const std::vector<size_t>& clusters,
                         const std::vector<std::vector<std::complex<double>>>& series);

}  // namespace sjfh

#endif /* SJF_HOST_HPP */
