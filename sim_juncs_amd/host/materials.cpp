// materials.cpp -- material maps and source descriptions of the reference driver
// (src/disp.cpp:56-283, 318-376, 418-454, 482-550).
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "sjf_host.hpp"

namespace sjfh {

// ---- susceptibilities (src/disp.cpp:418-454) -------------------------------------------------

std::vector<Suscept> parse_susceptibilities(const Value& v, int* er) {
    std::vector<Suscept> out;
    if (v.type == Value::List) {
        for (const Value& cur : v.l) {
            if (cur.type != Value::List || cur.l.size() < 2) {
                if (er) *er = -1;
                return out;
            }
            // an entry with two elements has no sigma: the type check fails
            if (cur.l[0].type != Value::Num || cur.l[1].type != Value::Num || cur.l.size() < 3 ||
                cur.l[2].type != Value::Num) {
                if (er) *er = -2;
                return out;
            }
            Suscept s;
            s.omega_0 = cur.l[0].x;
            s.gamma = cur.l[1].x;
            s.sigma = cur.l[2].x;
            s.use_denom = true;
            if (cur.l.size() > 3) {
                const Value& k = cur.l[3];
                if (k.type == Value::Num) {
                    s.use_denom = (k.x != 0.0);
                } else if (k.type == Value::Str) {
                    if (k.s == "drude" || k.s == "true") s.use_denom = false;
                    if (k.s == "lorentz" || k.s == "false") s.use_denom = true;
                } else {
                    if (er) *er = -2;
                    return out;
                }
            }
            out.push_back(s);
        }
    }
    if (er) *er = 0;
    return out;
}

// ---- material function (src/disp.cpp:56-283) ---------------------------------------------------

// Smoothing offsets of generate_smooth_pts (src/disp.cpp:56-112): a Mersenne twister seeded
// from the low 16 bits of the seed (the high word is shifted out), a random direction
// (cos theta uniform, phi uniform) and a Gaussian radius per point, x and y both taking cos phi,
// and all eight sign reflections of each point.  Seed: the reference's DEF_SEED 0xd9a28bf3.
MaterialFunction::MaterialFunction(double def_ret, unsigned smooth_n, double smooth_rad)
    : def_ret_(def_ret), smooth_n_(smooth_n) {
    if (smooth_n_ == 0) return;
    const uint64_t key = (uint64_t)(double)0xd9a28bf3;
    std::seed_seq seq{(uint32_t)(key & 0x0000ffff), (uint32_t)((key & 0xffff0000) >> 32)};
    std::mt19937 rng(seq);
    std::normal_distribution<double> radius(0, smooth_rad);
    std::uniform_real_distribution<double> uniform01(0.0, 1.0);
    pts_.resize(24 * (size_t)smooth_n_);
    double* out = pts_.data();
    for (unsigned p = 0; p < smooth_n_; ++p) {
        // draw order: polar parameter, azimuth, radius
        const double u = uniform01(rng);
        const double mu = 1 - 2 * u;                       // cosine of the polar angle
        const double rho = 2 * std::sqrt(u * (1 - u));     // its sine
        const double az = 2 * M_PI * uniform01(rng);
        const double rad = radius(rng);
        const double base[3] = {rad * rho * std::cos(az), rad * rho * std::cos(az), rad * mu};
        // reflections in the order x sign, then y, then z (bit set: +, clear: -)
        for (int refl = 0; refl < 8; ++refl)
            for (int d = 0; d < 3; ++d) *out++ = ((refl >> (2 - d)) & 1) ? base[d] : -base[d];
    }
    smooth_n_ *= 8;
}

void MaterialFunction::add_region(const Node* region, double scale) { regions_.emplace_back(region, scale); }

// in_bound (src/disp.cpp:264-283): every region adds def_ret + (scale - def_ret) * the fraction
// of the sample points (the point and its smoothing offsets) inside it, so overlapping or
// disjoint regions sum their background terms.
double MaterialFunction::in_bound(const Vec3& r) const {
    if (regions_.empty()) return def_ret_;
    double ret = 0;
    for (const auto& reg : regions_) {
        double this_ret = in(*reg.first, r);
        for (unsigned j = 0; j < (pts_.empty() ? 0u : smooth_n_); ++j)
            this_ret += in(*reg.first, Vec3{r.x + pts_[3 * j], r.y + pts_[3 * j + 1], r.z + pts_[3 * j + 2]});
        ret += def_ret_ + (reg.second - def_ret_) * this_ret / (smooth_n_ + 1);
    }
    return ret;
}

// ---- sources (src/disp.cpp:318-376, 597-626) -----------------------------------------------------

SourceInfo source_info(const SourceSpec& sp) {
    SourceInfo si;
    si.gaussian = sp.gaussian;
    if (sp.has[0]) {
        const int c = (int)sp.component;
        si.component = (c >= 1 && c <= 5) ? c : 0;
    }
    if (sp.has[1]) si.wavelen = sp.wavelength;
    if (sp.has[2]) si.amplitude = sp.amplitude;
    if (sp.has[6]) si.start_time = sp.start_time;
    if (sp.has[7]) si.end_time = sp.end_time;
    if (sp.has[8]) si.width = sp.slowness;
    if (sp.gaussian) {
        if (sp.has[3]) si.width = sp.width;
        if (sp.has[4]) si.phase = sp.phase;
        // any given cutoff becomes 6 widths; the end time follows from it
        double cutoff = 5;
        if (sp.has[5]) cutoff = 6;
        si.end_time = si.start_time + 2 * cutoff * si.width;
    }
    return si;
}

sjf_pulse pulse_of(const SourceInfo& si, double um_scale) {
    const double c_by_a = 0.299792458 * um_scale;
    sjf_pulse p;
    std::memset(&p, 0, sizeof p);
    p.kind = si.gaussian ? SJF_SRC_GAUSSIAN : SJF_SRC_CONTINUOUS;
    p.freq = 1 / (si.wavelen * um_scale);
    p.width = si.width * c_by_a;
    p.phase = si.phase;
    p.t_start = si.start_time * c_by_a;
    p.t_end = si.end_time * c_by_a;
    p.slowness = 3.0;  // meep::continuous_src_time's default
    return p;
}

// ---- voxelisation (src/disp.cpp:482-550) -----------------------------------------------------

namespace {

struct Key {
    double eps;
    int np;
    double f0[SJF_MAX_POLES], g[SJF_MAX_POLES], s[SJF_MAX_POLES];
    int drude[SJF_MAX_POLES];
    bool operator<(const Key& o) const {
        if (eps != o.eps) return eps < o.eps;
        if (np != o.np) return np < o.np;
        for (int q = 0; q < np; ++q) {
            if (f0[q] != o.f0[q]) return f0[q] < o.f0[q];
            if (g[q] != o.g[q]) return g[q] < o.g[q];
            if (s[q] != o.s[q]) return s[q] < o.s[q];
            if (drude[q] != o.drude[q]) return drude[q] < o.drude[q];
        }
        return false;
    }
};

struct PoleSpec {
    double f0, g, sigma;
    int drude;
    size_t root;
};

}  // namespace

VoxelMaps voxelize(const Scene& scene, const Settings& s, const int n[3], double dx) {
    VoxelMaps out;
    for (int d = 0; d < 3; ++d) out.n[d] = n[d];
    const size_t N = (size_t)n[0] * n[1] * n[2];
    out.words.assign(N, 0u);

    // make_2d roots are squashed along z and their sigma divided by the sheet thickness
    std::vector<std::unique_ptr<Node>> scaled(scene.roots.size());
    std::vector<const Node*> roots(scene.roots.size());
    std::vector<double> thick(scene.roots.size(), 1.0);
    for (size_t r = 0; r < scene.roots.size(); ++r) {
        const Root& R = scene.roots[r];
        roots[r] = R.tree.get();
        auto it = R.metadata.find("make_2d");
        if (it != R.metadata.end() && it->second.type == Value::Num && it->second.x != 0) {
            thick[r] = 1.0 / s.resolution;
            // rescale (src/cgs.cpp:19-21) on a copy of the root: trans = trans * diag(1, 1, t)
            struct Ref {
                static std::unique_ptr<Node> clone(const Node& a) {
                    auto b = std::make_unique<Node>();
                    b->kind = a.kind; b->invert = a.invert; b->trans = a.trans; b->center = a.center;
                    b->offset = a.offset; b->normal = a.normal; b->radius = a.radius; b->plane_offset = a.plane_offset;
                    b->height = a.height; b->r1_sq = a.r1_sq; b->r1_sq_x_h = a.r1_sq_x_h; b->r2_sq = a.r2_sq;
                    b->combine = a.combine;
                    for (int q = 0; q < 2; ++q) if (a.child[q]) b->child[q] = clone(*a.child[q]);
                    return b;
                }
            };
            auto inner = Ref::clone(*R.tree);
            for (int row = 0; row < 3; ++row) inner->trans.el[3 * row + 2] *= thick[r];
            scaled[r] = std::move(inner);
            roots[r] = scaled[r].get();
        }
    }

    // eps_inf: one region per root with its "eps" metadata (default: the ambient value)
    MaterialFunction eps_fn(s.ambient_eps, s.smooth_n, s.smooth_rad);
    for (size_t r = 0; r < roots.size(); ++r) {
        double scale = s.ambient_eps;
        auto it = scene.roots[r].metadata.find("eps");
        if (it != scene.roots[r].metadata.end() && it->second.type == Value::Num) scale = it->second.x;
        eps_fn.add_region(roots[r], scale);
    }
    // poles: every pole of every root, strength sigma * in_root (default 0 outside)
    std::vector<PoleSpec> poles;
    std::vector<MaterialFunction> pole_fn;
    for (size_t r = 0; r < roots.size(); ++r) {
        auto it = scene.roots[r].metadata.find("susceptibilities");
        if (it == scene.roots[r].metadata.end()) continue;
        int er = 0;
        for (const Suscept& sp : parse_susceptibilities(it->second, &er)) {
            PoleSpec p{sp.omega_0 / s.um_scale, sp.gamma / s.um_scale, sp.sigma / thick[r], sp.use_denom ? 0 : 1, r};
            poles.push_back(p);
            MaterialFunction f(0.0, s.smooth_n, s.smooth_rad);
            f.add_region(roots[r], p.sigma);
            pole_fn.push_back(std::move(f));
        }
    }

    static const double off[3][3] = {{0.5, 0, 0}, {0, 0.5, 0}, {0, 0, 0.5}};
    const int nx = n[0], ny = n[1], nz = n[2];
    // smoothed inclusion fractions make eps_inf a continuum: it goes to a per-cell map and the LUT
    // keys only the poles
    const bool map_eps = s.smooth_n > 0;
    if (map_eps) out.eps_map.assign(3 * N, 1.0);
    // each x plane collects its materials in first-seen order; planes merge in order afterwards
    std::vector<std::vector<Key>> plane_keys(nx);
    std::string error;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nx; ++i) {
        std::map<Key, uint32_t> local;
        std::vector<Key>& order = plane_keys[i];
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                uint32_t word = 0;
                for (int c = 0; c < 3; ++c) {
                    const Vec3 p{(i + off[c][0]) * dx, (j + off[c][1]) * dx, (k + off[c][2]) * dx};
                    Key key;
                    std::memset(&key, 0, sizeof key);
                    key.eps = eps_fn.in_bound(p);
                    if (map_eps) {
                        out.eps_map[(size_t)c * N + ((size_t)i * ny + j) * nz + k] = key.eps;
                        key.eps = 1.0;
                    }
                    for (size_t q = 0; q < poles.size(); ++q) {
                        const double sg = pole_fn[q].in_bound(p);
                        if (sg == 0) continue;
                        if (key.np == SJF_MAX_POLES) {
#pragma omp critical
                            error = "more than 4 poles overlap at one Yee point";
                            continue;
                        }
                        key.f0[key.np] = poles[q].f0;
                        key.g[key.np] = poles[q].g;
                        key.s[key.np] = sg;
                        key.drude[key.np] = poles[q].drude;
                        ++key.np;
                    }
                    auto f = local.find(key);
                    uint32_t id;
                    if (f == local.end()) {
                        id = (uint32_t)order.size();
                        local.emplace(key, id);
                        order.push_back(key);
                    } else {
                        id = f->second;
                    }
                    word |= (id & 0xffu) << (8 * c);
                    if (id > 0xff) {
#pragma omp critical
                        error = "more than 255 distinct materials in one plane";
                    }
                }
                out.words[((size_t)i * ny + j) * nz + k] = word;
            }
    }
    if (!error.empty()) throw std::runtime_error("voxelize: " + error);
    // merge: global ids in plane order, then remap every plane's local ids
    std::map<Key, uint32_t> global;
    std::vector<Key> all;
    std::vector<std::vector<uint32_t>> remap(nx);
    for (int i = 0; i < nx; ++i) {
        for (const Key& k : plane_keys[i]) {
            auto f = global.find(k);
            uint32_t id;
            if (f == global.end()) {
                id = (uint32_t)all.size();
                global.emplace(k, id);
                all.push_back(k);
            } else {
                id = f->second;
            }
            remap[i].push_back(id);
        }
    }
    if (all.size() > 255) throw std::runtime_error("voxelize: more than 255 distinct materials");
#pragma omp parallel for
    for (int i = 0; i < nx; ++i) {
        uint32_t* w = out.words.data() + (size_t)i * ny * nz;
        for (size_t q = 0; q < (size_t)ny * nz; ++q) {
            const uint32_t v = w[q];
            w[q] = remap[i][v & 0xff] | (remap[i][(v >> 8) & 0xff] << 8) | (remap[i][(v >> 16) & 0xff] << 16);
        }
    }
    for (const Key& k : all) {
        sjf_material m;
        std::memset(&m, 0, sizeof m);
        m.eps_inf = k.eps;
        m.npoles = k.np;
        for (int q = 0; q < k.np; ++q) {
            m.f0[q] = k.f0[q];
            m.gamma[q] = k.g[q];
            m.sigma[q] = k.s[q];
            m.is_drude[q] = k.drude[q];
        }
        out.lut.push_back(m);
    }
    return out;
}

}  // namespace sjfh
