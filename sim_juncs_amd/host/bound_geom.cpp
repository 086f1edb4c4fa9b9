// bound_geom.cpp -- the reference driver's simulation object on top of the engine
// (bound_geom, src/disp.hpp:136-202, src/disp.cpp:557-923).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "sjf_host.hpp"

namespace sjfh {

void write_grid_h5(const std::string& path, const char* name, const std::vector<double>& data, const int n[3]);

namespace {

// make_dec_str (src/argparse.h:394-430): integer part zero padded to a digits, '.', then b
// truncated fractional digits.
std::string dec_str(double t, unsigned a, unsigned b) {
    long bef = (long)t;
    t -= (double)bef;
    std::string s(a, '0');
    for (unsigned i = 0; i < a && (bef || i == 0); ++i) {
        s[a - i - 1] = (char)('0' + bef % 10);
        bef /= 10;
    }
    s += '.';
    for (unsigned i = 0; i < b; ++i) {
        t *= 10;
        s += (char)('0' + (int)t);
        t -= std::floor(t);
    }
    return s;
}

}  // namespace

void BoundGeom::check(int rc, const char* what, int part) const {
    if (rc) throw std::runtime_error(std::string(what) + ": " + sjf_last_error(eng_[part]));
}

BoundGeom::BoundGeom(const Settings& s, const Scene& scene, const RunOptions& opt)
    : scene_(&scene), s_(s), opt_(opt) {
    if (s.grid_num <= 0 || s.resolution <= 0)
        throw std::invalid_argument("grid_num and resolution must be positive (call correct_defaults)");
    // the reference builds vol1d / vol2d for dimensions 1 / 2 (src/disp.cpp:504-510); this driver
    // runs the 3-D volume only and refuses the others rather than run different physics
    if (s.n_dims != 3)
        throw std::invalid_argument("dimensions = " + std::to_string(s.n_dims) + " is not supported (3-D only)");
    n_ = (int)s.grid_num;
    um_scale_ = s.um_scale;
    post_source_t_ = s.post_source_t;
    save_span_ = s.save_span ? s.save_span : 1;
    nparts_ = opt.complex_fields ? 2 : 1;
    const double dx = 1.0 / s.resolution;

    const int n3[3] = {n_, n_, n_};
    VoxelMaps vm = voxelize(scene, s, n3, dx);
    eps_cell_.resize(vm.words.size());
    const size_t N = vm.words.size();
    for (size_t q = 0; q < N; ++q) {
        double e = 0;
        for (int c = 0; c < 3; ++c)
            e += vm.eps_map.empty() ? vm.lut[(vm.words[q] >> (8 * c)) & 0xff].eps_inf : vm.eps_map[c * N + q];
        eps_cell_[q] = e / 3;
    }

    // meep::pml(thickness) on all faces, in whole cells
    const int pc = (int)std::lround(s.pml_thickness * s.resolution);
    const int cells[6] = {pc, pc, pc, pc, pc, pc};
    for (int p = 0; p < nparts_; ++p) {
        sjf_grid g;
        std::memset(&g, 0, sizeof g);
        g.nx = g.ny = g.nz = n_;
        g.dx = dx;
        g.courant = 0.5;  // the reference never passes its courant key to meep
        g.precision = opt.precision;
        g.part = p;
        g.device = opt.device;
        if (sjf_create(&g, &eng_[p]))
            throw std::runtime_error(std::string("sjf_create: ") + sjf_last_error(nullptr));
        check(sjf_set_materials(eng_[p], vm.words.data(), vm.lut.data(), (int)vm.lut.size()), "sjf_set_materials", p);
        if (!vm.eps_map.empty()) check(sjf_set_eps_maps(eng_[p], vm.eps_map.data()), "sjf_set_eps_maps", p);
        if (pc > 0) check(sjf_set_pml(eng_[p], cells, 1e-15, 2.0), "sjf_set_pml", p);
        check(sjf_set_graph(eng_[p], opt.graph_chunk), "sjf_set_graph", p);
    }
    check(sjf_dt(eng_[0], &dt_), "sjf_dt");

    // sources and monitors in the order of the reference's walk (src/disp.cpp:586-639)
    for (const SourceSpec& sp : scene.sources) {
        if (!sp.has_box) continue;  // only Box regions are sources (src/disp.cpp:597)
        const SourceInfo si = source_info(sp);
        sources_.push_back(si);
        // any component: Ex..Ez are electric currents, Hx..Hz magnetic currents (src/disp.cpp:614,619)
        const sjf_pulse pl = pulse_of(si, um_scale_);
        const double box[6] = {sp.box[0].x, sp.box[0].y, sp.box[0].z, sp.box[1].x, sp.box[1].y, sp.box[1].z};
        for (int p = 0; p < nparts_; ++p)
            check(sjf_add_source(eng_[p], si.component, box, &pl, si.amplitude, 0.0), "sjf_add_source", p);
        double last = 0;
        check(sjf_last_source_time(eng_[0], &last), "sjf_last_source_time");
        ttot_ = last + post_source_t_ * 0.299792458 * um_scale_;
    }
    // Lorentz poles out of the PML slabs, matched at the first source's frequency
    if (opt.pml_fold && pc > 0 && !sources_.empty()) {
        const double f = pulse_of(sources_.front(), um_scale_).freq;
        for (int p = 0; p < nparts_; ++p) check(sjf_set_pml_dispersion(eng_[p], f), "sjf_set_pml_dispersion", p);
    }
    for (const auto& cl : scene.monitor_clusters) {
        for (const Vec3& v : cl) monitor_locs_.push_back(v);
        monitor_clusters_.push_back(monitor_locs_.size());
    }
    if (!monitor_locs_.empty()) {
        std::vector<double> xyz;
        for (const Vec3& v : monitor_locs_) {
            xyz.push_back(v.x);
            xyz.push_back(v.y);
            xyz.push_back(v.z);
        }
        for (int p = 0; p < nparts_; ++p)
            check(sjf_add_monitors(eng_[p], 0 /* Ex */, xyz.data(), (int)monitor_locs_.size()), "sjf_add_monitors", p);
    }
    for (int p = 0; p < nparts_; ++p) check(sjf_set_save_span(eng_[p], (int)save_span_), "sjf_set_save_span", p);
}

BoundGeom::~BoundGeom() {
    for (int p = 0; p < 2; ++p)
        if (eng_[p]) sjf_destroy(eng_[p]);
}

// run (src/disp.cpp:690-749): the dielectric is written once, then n_t_pts steps run with Ex
// sampled at every monitor before each step whose index is a multiple of save_span (the
// engine samples on the device); with dump_raw the Ex field is written at those steps too.
void BoundGeom::run(const std::string& out_dir) {
    const int n3[3] = {n_, n_, n_};
    write_grid_h5(out_dir + "/eps-000000.00.h5", "eps", eps_cell_, n3);
    if (!(ttot_ > 0))
        throw std::runtime_error("no source sets the run time (total time " + std::to_string(ttot_) + ")");
    n_t_pts_ = (unsigned)((ttot_ + dt_ / 2) / dt_);
    const size_t n_locs = monitor_locs_.size();
    field_times_.assign(n_locs, {});

    sjf_batch* batch = nullptr;
    if (nparts_ == 2 && sjf_batch_create(eng_, 2, &batch)) batch = nullptr;
    auto advance = [&](long k) {
        if (k <= 0) return;
        if (batch) {
            if (sjf_batch_step(batch, k)) throw std::runtime_error(std::string("sjf_batch_step: ") + sjf_batch_last_error(batch));
        } else {
            for (int p = 0; p < nparts_; ++p) check(sjf_step(eng_[p], k), "sjf_step", p);
        }
    };
    try {
        if (!s_.dump_raw) {
            advance((long)n_t_pts_);
        } else {
            const unsigned na = (unsigned)std::ceil(std::log(ttot_) / std::log(10));
            const unsigned nb = (unsigned)std::ceil(-std::log(dt_) / std::log(10.0)) + 1;
            std::vector<double> fr((size_t)n_ * n_ * n_), fi;
            for (unsigned i = 0; i < n_t_pts_; i += save_span_) {
                if (batch) sjf_batch_synchronize(batch);
                check(sjf_read_field(eng_[0], 0, fr.data()), "sjf_read_field");
                double t = 0;
                check(sjf_time(eng_[0], &t), "sjf_time");
                const std::string path = out_dir + "/ex-" + dec_str(t, na, nb) + ".h5";
                write_grid_h5(path, "ex.r", fr, n3);
                if (nparts_ == 2) {
                    fi.resize(fr.size());
                    check(sjf_read_field(eng_[1], 0, fi.data()), "sjf_read_field", 1);
                    write_grid_h5(out_dir + "/ex-" + dec_str(t, na, nb) + "-imag.h5", "ex.i", fi, n3);
                }
                advance((long)std::min<unsigned>(save_span_, n_t_pts_ - i));
            }
        }
    } catch (...) {
        if (batch) sjf_batch_destroy(batch);
        throw;
    }
    if (batch) {
        sjf_batch_synchronize(batch);
        sjf_batch_destroy(batch);
    }
    if (n_locs) {
        std::vector<std::vector<double>> raw(nparts_);
        long ns = 0;
        for (int p = 0; p < nparts_; ++p) {
            long have = 0;
            check(sjf_read_monitors(eng_[p], nullptr, 0, &have), "sjf_read_monitors", p);
            raw[p].resize((size_t)have * n_locs * 2);
            check(sjf_read_monitors(eng_[p], raw[p].data(), raw[p].size(), &ns), "sjf_read_monitors", p);
        }
        for (long smp = 0; smp < ns; ++smp)
            for (size_t j = 0; j < n_locs; ++j) {
                const double re = raw[0][((size_t)smp * n_locs + j) * 2];
                const double im = nparts_ == 2 ? raw[1][((size_t)smp * n_locs + j) * 2] : 0.0;
                field_times_[j].emplace_back(re, im);
                if (re > 1000)
                    std::printf("divergence in run at (i,j)=(%ld,%zu) (%f)\n", smp * (long)save_span_, j, re);
            }
    }
}

void BoundGeom::save_field_times(const std::string& out_dir) {
    write_field_samples(out_dir + "/field_samples.h5", meep_time_to_fs(ttot_), n_t_pts_, save_span_, sources_,
                        scene_->context_numbers, monitor_locs_, monitor_clusters_, field_times_);
}

}  // namespace sjfh
