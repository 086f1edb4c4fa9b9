// h5out.cpp -- HDF5 output of the driver (field_samples.h5, src/disp.cpp:758-923; grid dumps
// in place of meep's output_hdf5, src/disp.cpp:696,732-737) and a small reader for tests.
//
// libhdf5 (the 1.10 C library of the image's /opt/conda) is loaded with dlopen on first use,
// so the host library has no link-time dependency on it (and does not pull that prefix's
// older C++ runtime into the process).  Only the few entry points used here are resolved.
#include <dlfcn.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "sjf_host.hpp"

extern "C" {
#include "sjf_host.h"
}

namespace sjfh {

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

struct H5 {
    void* lib = nullptr;
    herr_t (*open)();
    herr_t (*eset_auto)(hid_t, void*, void*);
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char*, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t);
    hid_t (*gopen)(hid_t, const char*, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    herr_t (*sclose)(hid_t);
    long long (*sget_npoints)(hid_t);
    hid_t (*tcreate)(int, size_t);
    herr_t (*tinsert)(hid_t, const char*, size_t, hid_t);
    herr_t (*tclose)(hid_t);
    hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen)(hid_t, const char*, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    hid_t (*dget_space)(hid_t);
    herr_t (*dclose)(hid_t);
    herr_t (*gget_info)(hid_t, void*);
    long (*lget_name_by_idx)(hid_t, const char*, int, int, hsize_t, char*, size_t, hid_t);
    hid_t* native_double;
    hid_t* native_hsize;
};

constexpr unsigned kAccTrunc = 0x0002u, kAccRdonly = 0x0000u;
constexpr hid_t kDefault = 0, kAll = 0;
constexpr int kCompound = 6;

H5& h5() {
    static H5 h;
    static std::once_flag once;
    static std::string err;
    std::call_once(once, [] {
        const char* env = std::getenv("SJF_HDF5_LIB");
        const char* cands[] = {env, "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so", "libhdf5.so.103",
                               "libhdf5.so"};
        for (const char* c : cands) {
            if (!c) continue;
            h.lib = dlopen(c, RTLD_NOW | RTLD_LOCAL);
            if (h.lib) break;
        }
        if (!h.lib) {
            err = "libhdf5 (1.10) not found; set SJF_HDF5_LIB";
            return;
        }
        auto sym = [&](const char* n) {
            void* p = dlsym(h.lib, n);
            if (!p) err = std::string("libhdf5 lacks ") + n;
            return p;
        };
        h.open = (herr_t(*)())sym("H5open");
        h.eset_auto = (herr_t(*)(hid_t, void*, void*))sym("H5Eset_auto2");
        h.fcreate = (hid_t(*)(const char*, unsigned, hid_t, hid_t))sym("H5Fcreate");
        h.fopen = (hid_t(*)(const char*, unsigned, hid_t))sym("H5Fopen");
        h.fclose = (herr_t(*)(hid_t))sym("H5Fclose");
        h.gcreate = (hid_t(*)(hid_t, const char*, hid_t, hid_t, hid_t))sym("H5Gcreate2");
        h.gopen = (hid_t(*)(hid_t, const char*, hid_t))sym("H5Gopen2");
        h.gclose = (herr_t(*)(hid_t))sym("H5Gclose");
        h.screate_simple = (hid_t(*)(int, const hsize_t*, const hsize_t*))sym("H5Screate_simple");
        h.sclose = (herr_t(*)(hid_t))sym("H5Sclose");
        h.sget_npoints = (long long (*)(hid_t))sym("H5Sget_simple_extent_npoints");
        h.tcreate = (hid_t(*)(int, size_t))sym("H5Tcreate");
        h.tinsert = (herr_t(*)(hid_t, const char*, size_t, hid_t))sym("H5Tinsert");
        h.tclose = (herr_t(*)(hid_t))sym("H5Tclose");
        h.dcreate = (hid_t(*)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t))sym("H5Dcreate2");
        h.dopen = (hid_t(*)(hid_t, const char*, hid_t))sym("H5Dopen2");
        h.dwrite = (herr_t(*)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*))sym("H5Dwrite");
        h.dread = (herr_t(*)(hid_t, hid_t, hid_t, hid_t, hid_t, void*))sym("H5Dread");
        h.dget_space = (hid_t(*)(hid_t))sym("H5Dget_space");
        h.dclose = (herr_t(*)(hid_t))sym("H5Dclose");
        h.gget_info = (herr_t(*)(hid_t, void*))sym("H5Gget_info");
        h.lget_name_by_idx = (long (*)(hid_t, const char*, int, int, hsize_t, char*, size_t, hid_t))sym("H5Lget_name_by_idx");
        h.native_double = (hid_t*)sym("H5T_NATIVE_DOUBLE_g");
        h.native_hsize = (hid_t*)sym("H5T_NATIVE_HSIZE_g");
        if (err.empty()) {
            h.open();
            h.eset_auto(0, nullptr, nullptr);
        }
    });
    if (!err.empty()) throw std::runtime_error(err);
    return h;
}

void ok(long long rc, const std::string& what) {
    if (rc < 0) throw std::runtime_error("hdf5: " + what + " failed");
}

struct SrcRec {
    double wavelen, width, phase, start_time, end_time, amplitude;
};

// zero-padded decimal of n digits (write_number, src/argparse.h:58-83)
std::string padded(size_t num, size_t digits) {
    std::string s = std::to_string(num);
    if (s.size() < digits) s = std::string(digits - s.size(), '0') + s;
    return s;
}

size_t digits_of(size_t count) {
    if (count == 0) return 1;
    return (size_t)(std::log((double)count) / std::log(10)) + 1;
}

hid_t dataset(hid_t loc, const char* name, hid_t type, const std::vector<hsize_t>& dims) {
    H5& h = h5();
    const hid_t sp = h.screate_simple((int)dims.size(), dims.data(), nullptr);
    ok(sp, "dataspace");
    const hid_t d = h.dcreate(loc, name, type, sp, kDefault, kDefault, kDefault);
    h.sclose(sp);
    ok(d, std::string("create dataset ") + name);
    return d;
}

}  // namespace

// Layout of save_field_times (src/disp.cpp:758-923):
//   /info/time_bounds[3] = {0, ttot in fs, ttot * save_span / n_t_pts}
//   /info/n_clusters, /info/n_time_points (= floor(n_t_pts / save_span)), /info/sources (compound
//   wavelen, width, phase, start_time, end_time, amplitude), /info/cgs_params/<name> scalars
//   /cluster_<j>/locations (compound x, y, z) for every cluster plus one trailing cluster with
//   the points after the last boundary, /cluster_<j>/point_<i>/{time, frequency} (compound Re,
//   Im), point indices global and zero padded; time holds the first floor(n_t_pts/save_span)
//   samples, frequency the reference's fft of the whole series.
void write_field_samples(const std::string& path, double ttot_fs, unsigned n_t_pts, unsigned save_span,
                         const std::vector<SourceInfo>& sources,
                         const std::vector<std::pair<std::string, double>>& cgs_params,
                         const std::vector<Vec3>& locs, const std::vector<size_t>& clusters,
                         const std::vector<std::vector<std::complex<double>>>& series) {
    H5& h = h5();
    const hid_t dbl = *h.native_double;
    const hid_t field = h.tcreate(kCompound, 16);
    ok(h.tinsert(field, "Re", 0, dbl), "insert Re");
    ok(h.tinsert(field, "Im", 8, dbl), "insert Im");
    const hid_t loct = h.tcreate(kCompound, 24);
    ok(h.tinsert(loct, "x", 0, dbl), "insert x");
    ok(h.tinsert(loct, "y", 8, dbl), "insert y");
    ok(h.tinsert(loct, "z", 16, dbl), "insert z");
    const hid_t srct = h.tcreate(kCompound, sizeof(SrcRec));
    ok(h.tinsert(srct, "wavelen", offsetof(SrcRec, wavelen), dbl), "insert wavelen");
    ok(h.tinsert(srct, "width", offsetof(SrcRec, width), dbl), "insert width");
    ok(h.tinsert(srct, "phase", offsetof(SrcRec, phase), dbl), "insert phase");
    ok(h.tinsert(srct, "start_time", offsetof(SrcRec, start_time), dbl), "insert start_time");
    ok(h.tinsert(srct, "end_time", offsetof(SrcRec, end_time), dbl), "insert end_time");
    ok(h.tinsert(srct, "amplitude", offsetof(SrcRec, amplitude), dbl), "insert amplitude");

    const size_t n_locs = locs.size();
    const hsize_t n_t = save_span ? n_t_pts / save_span : n_t_pts;
    std::vector<std::vector<std::complex<double>>> fdom(n_locs);
    hsize_t n_f = 0;
    for (size_t j = 0; j < n_locs; ++j) {
        fdom[j] = fft(series[j]);
        n_f = fdom[j].size();
    }

    const hid_t file = h.fcreate(path.c_str(), kAccTrunc, kDefault, kDefault);
    ok(file, "create " + path);
    const hid_t info = h.gcreate(file, "info", kDefault, kDefault, kDefault);
    ok(info, "group info");
    {
        const double tb[3] = {0.0, ttot_fs, n_t_pts ? (ttot_fs - 0.0) * save_span / n_t_pts : 0.0};
        hid_t d = dataset(info, "time_bounds", dbl, {3});
        ok(h.dwrite(d, dbl, kAll, kAll, kDefault, tb), "write time_bounds");
        h.dclose(d);
        hsize_t v = clusters.size();
        d = dataset(info, "n_clusters", *h.native_hsize, {1});
        ok(h.dwrite(d, *h.native_hsize, kAll, kAll, kDefault, &v), "write n_clusters");
        h.dclose(d);
        v = n_t;
        d = dataset(info, "n_time_points", *h.native_hsize, {1});
        ok(h.dwrite(d, *h.native_hsize, kAll, kAll, kDefault, &v), "write n_time_points");
        h.dclose(d);
        std::vector<SrcRec> recs;
        for (const SourceInfo& s : sources)
            recs.push_back({s.wavelen, s.width, s.phase, s.start_time, s.end_time, s.amplitude});
        d = dataset(info, "sources", srct, {(hsize_t)recs.size()});
        if (!recs.empty()) ok(h.dwrite(d, srct, kAll, kAll, kDefault, recs.data()), "write sources");
        h.dclose(d);
        const hid_t cg = h.gcreate(info, "cgs_params", kDefault, kDefault, kDefault);
        ok(cg, "group cgs_params");
        std::map<std::string, bool> seen;
        for (const auto& kv : cgs_params) {
            if (seen[kv.first]) continue;  // a name defined twice keeps its first (oldest) value
            seen[kv.first] = true;
            d = dataset(cg, kv.first.c_str(), dbl, {1});
            ok(h.dwrite(d, dbl, kAll, kAll, kDefault, &kv.second), "write " + kv.first);
            h.dclose(d);
        }
        h.gclose(cg);
    }
    h.gclose(info);

    const size_t gdig = digits_of(clusters.size()), pdig = digits_of(n_locs);
    size_t i = 0, first = 0;
    for (size_t j = 0; j < clusters.size() + 1; ++j) {
        const size_t max_i = j >= clusters.size() ? n_locs : clusters[j];
        const std::string gname = "cluster_" + padded(j, gdig);
        const hid_t g = h.gcreate(file, gname.c_str(), kDefault, kDefault, kDefault);
        ok(g, "group " + gname);
        std::vector<double> xyz;
        for (size_t q = first; q < max_i; ++q) {
            xyz.push_back(locs[q].x);
            xyz.push_back(locs[q].y);
            xyz.push_back(locs[q].z);
        }
        hid_t d = dataset(g, "locations", loct, {(hsize_t)(max_i - first)});
        if (!xyz.empty()) ok(h.dwrite(d, loct, kAll, kAll, kDefault, xyz.data()), "write locations");
        h.dclose(d);
        first = max_i;
        for (; i < max_i; ++i) {
            if (series[i].size() < 2) {
                std::printf("Error: monitor location %zu has insufficient points\n", i);
                break;
            }
            const std::string pname = "point_" + padded(i, pdig);
            const hid_t pg = h.gcreate(g, pname.c_str(), kDefault, kDefault, kDefault);
            ok(pg, "group " + pname);
            std::vector<double> t(2 * n_t), f(2 * n_f);
            for (hsize_t k = 0; k < n_t && k < series[i].size(); ++k) {
                t[2 * k] = series[i][k].real();
                t[2 * k + 1] = series[i][k].imag();
            }
            for (hsize_t k = 0; k < n_f && k < fdom[i].size(); ++k) {
                f[2 * k] = fdom[i][k].real();
                f[2 * k + 1] = fdom[i][k].imag();
            }
            d = dataset(pg, "time", field, {n_t});
            if (n_t) ok(h.dwrite(d, field, kAll, kAll, kDefault, t.data()), "write time");
            h.dclose(d);
            d = dataset(pg, "frequency", field, {n_f});
            if (n_f) ok(h.dwrite(d, field, kAll, kAll, kDefault, f.data()), "write frequency");
            h.dclose(d);
            h.gclose(pg);
        }
        h.gclose(g);
    }
    h.fclose(file);
    h.tclose(field);
    h.tclose(loct);
    h.tclose(srct);
}

// One 3-D double dataset in its own file (the engine's stand-in for meep's output_hdf5).
void write_grid_h5(const std::string& path, const char* name, const std::vector<double>& data, const int n[3]) {
    H5& h = h5();
    const hid_t file = h.fcreate(path.c_str(), kAccTrunc, kDefault, kDefault);
    ok(file, "create " + path);
    const hid_t d = dataset(file, name, *h.native_double, {(hsize_t)n[0], (hsize_t)n[1], (hsize_t)n[2]});
    ok(h.dwrite(d, *h.native_double, kAll, kAll, kDefault, data.data()), std::string("write ") + name);
    h.dclose(d);
    h.fclose(file);
}

}  // namespace sjfh

// ---- reader entry points (C ABI, for tests and tools) --------------------------------------

extern "C" int sjfh_h5_read(const char* path, const char* dset, const char* const* members, int nmembers,
                            double* out, size_t cap, size_t* count) {
    using namespace sjfh;
    try {
        H5& h = h5();
        const hid_t f = h.fopen(path, kAccRdonly, kDefault);
        if (f < 0) return -1;
        const hid_t d = h.dopen(f, dset, kDefault);
        if (d < 0) {
            h.fclose(f);
            return -2;
        }
        const hid_t sp = h.dget_space(d);
        const long long npts = h.sget_npoints(sp);
        h.sclose(sp);
        const size_t per = nmembers > 0 ? (size_t)nmembers : 1;
        if (count) *count = (size_t)npts;
        int rc = 0;
        if ((size_t)npts * per > cap) {
            rc = -3;
        } else if (npts > 0) {
            hid_t mt = *h.native_double;
            if (nmembers > 0) {
                mt = h.tcreate(kCompound, 8 * per);
                for (int m = 0; m < nmembers; ++m) h.tinsert(mt, members[m], 8 * m, *h.native_double);
            }
            if (h.dread(d, mt, kAll, kAll, kDefault, out) < 0) rc = -4;
            if (nmembers > 0) h.tclose(mt);
        }
        h.dclose(d);
        h.fclose(f);
        return rc;
    } catch (const std::exception&) {
        return -5;
    }
}

extern "C" int sjfh_h5_list(const char* path, const char* group, char* out, size_t cap) {
    using namespace sjfh;
    try {
        H5& h = h5();
        const hid_t f = h.fopen(path, kAccRdonly, kDefault);
        if (f < 0) return -1;
        const hid_t g = h.gopen(f, group, kDefault);
        if (g < 0) {
            h.fclose(f);
            return -2;
        }
        struct {
            int storage_type;
            hsize_t nlinks;
            int64_t max_corder;
            int mounted;
        } info;
        h.gget_info(g, &info);
        std::string names;
        char buf[512];
        for (hsize_t q = 0; q < info.nlinks; ++q) {
            const long len = h.lget_name_by_idx(g, ".", 0, 0, q, buf, sizeof buf, kDefault);
            if (len < 0) continue;
            names += buf;
            names += '\n';
        }
        h.gclose(g);
        h.fclose(f);
        if (names.size() + 1 > cap) return -3;
        std::memcpy(out, names.c_str(), names.size() + 1);
        return 0;
    } catch (const std::exception&) {
        return -5;
    }
}
