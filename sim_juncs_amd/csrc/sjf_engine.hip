// sjf_engine.hip -- host side of the MI355X FDTD engine behind include/sjf.h.
//
// Owns the device state of one simulation (fields, materials, PML auxiliaries, sources,
// monitors, DFT accumulators) and enqueues, per time step, on the engine's own stream:
//   [monitor gather when step % save_span == 0] -> k_step -> k_source (per source) ->
//   k_dft (per DFT region), then swaps the double-buffered E/H/psi_H arrays.
// This is the sequence bound_geom::run performs around fields.step()
// (reference src/disp.cpp:719-741).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sjf.h"
#include "sjf_kernels.h"

using sjf::kMaxPoles;

namespace {

thread_local std::string g_create_error;

struct Source {
    int comp;
    sjf_pulse p;
    double omega = 0, phi = 0, peak = 0, cutoff = 0;  // gaussian, derived
    double amp_re = 1, amp_im = 0;
    std::vector<long> idx;
    std::vector<double> wt;
    long* d_idx = nullptr;
    double* d_coef = nullptr;
};

struct Monitor {
    int comp;
    long idx[8];
    double w[8];
};

struct Dft {
    int lo[3], hi[3];
    int mask;
    std::vector<double> freqs;
    long ncell;
    int ncomp;
    int comps[6];
    double2* acc = nullptr;
    double2* cmp = nullptr;
};

}  // namespace

struct sjf_engine {
    sjf_grid g{};
    int prec = 4;
    long N = 0;
    double dt = 0;
    hipStream_t stream = nullptr;
    std::string err;

    void* Fbuf[2] = {};  // field sets [Ex Ey Ez Hx Hy Hz][N], read/write alternate every step
    void* E[2][3] = {};  // views into Fbuf
    void* H[2][3] = {};
    int ping = 0;

    // materials
    bool have_mat = false;
    std::vector<uint32_t> matw_h;
    std::vector<sjf_material> lut_h;
    uint32_t* matw = nullptr;
    void* lut = nullptr;
    int nmat = 1;
    int maxp = 0;
    bool lorentz_slot[kMaxPoles] = {};
    void* Qbuf = nullptr;  // [slot][comp][N]
    void* Pbuf = nullptr;  // [slot][comp][N], present when a slot holds a Lorentz pole
    bool uniform = true;
    // design region
    int dlo[3] = {0, 0, 0}, dhi[3] = {0, 0, 0};
    std::vector<double> deps_h;
    void* dinv = nullptr;

    // PML
    int pml[6] = {0, 0, 0, 0, 0, 0};
    bool pml_on[3] = {false, false, false};
    void* coefbuf = nullptr;  // per axis: bE[n], aE[n], bH[n], aH[n]
    int coff[3] = {0, 0, 0};
    long slab[3] = {0, 0, 0};
    void* psibuf[3] = {};     // per axis [6][slab]: E aux x2, H aux pair A x2, pair B x2

    std::vector<Source> src;
    bool src_dirty = true;

    std::vector<Monitor> mon;
    int save_span = 1;
    long nsamp = 0, samp_cap = 0;
    long* d_mon_idx = nullptr;
    double* d_mon_w = nullptr;
    int* d_mon_comp = nullptr;
    double* d_samples = nullptr;
    bool mon_dirty = false;

    std::vector<Dft> dft;

    long step = 0;
    int ty = 4, xc = 32;
    bool tile_set = false;
};

namespace {

#define HIPCHK(e, call)                                                                  \
    do {                                                                                 \
        hipError_t _st = (call);                                                         \
        if (_st != hipSuccess) {                                                         \
            (e)->err = std::string(#call) + ": " + hipGetErrorString(_st);               \
            return SJF_E_HIP;                                                            \
        }                                                                                \
    } while (0)

int fail(sjf_engine* e, int code, const std::string& m) {
    e->err = m;
    return code;
}

int dev_alloc(sjf_engine* e, void** p, size_t bytes) {
    if (*p) return SJF_OK;
    hipError_t st = hipMalloc(p, bytes ? bytes : 16);
    if (st != hipSuccess) {
        *p = nullptr;
        e->err = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(st);
        return SJF_E_NOMEM;
    }
    st = hipMemsetAsync(*p, 0, bytes ? bytes : 16, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) return fail(e, SJF_E_HIP, hipGetErrorString(st));
    return SJF_OK;
}

void dev_free(void*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

// All copies are ordered on the engine's (non-blocking) stream and completed before return.
int h2d(sjf_engine* e, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SJF_OK;
    HIPCHK(e, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
}

int d2h(sjf_engine* e, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SJF_OK;
    HIPCHK(e, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
}

const double kCompOff[6][3] = {{0.5, 0.0, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5},
                               {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}};

long cell_index(const sjf_engine* e, int i, int j, int k) {
    return ((long)i * e->g.ny + j) * e->g.nz + k;
}

int axis_n(const sjf_engine* e, int d) { return d == 0 ? e->g.nx : (d == 1 ? e->g.ny : e->g.nz); }

// Weights of an extent [lo, hi] on the lattice (i + off) dx: delta function (linear / dx) for
// zero extent, cell overlap / dx otherwise.
void axis_weights(int n, double dx, double off, int periodic, double lo, double hi,
                  std::vector<int>& ids, std::vector<double>& w) {
    ids.clear();
    w.clear();
    if (hi < lo) std::swap(lo, hi);
    if (hi - lo < 1e-9 * dx) {
        const double s = lo / dx - off;
        const double fl = std::floor(s);
        const int i0 = (int)fl;
        const double f = s - fl;
        for (int q = 0; q < 2; ++q) {
            int i = i0 + q;
            const double ww = (q == 0 ? 1.0 - f : f) / dx;
            if (ww == 0.0) continue;
            if (periodic) i = ((i % n) + n) % n;
            else if (i < 0 || i >= n) continue;
            ids.push_back(i);
            w.push_back(ww);
        }
        return;
    }
    // on a periodic axis the cell of lattice point p also has images p +- n dx
    for (int i = 0; i < n; ++i) {
        double ww = 0;
        for (int m = periodic ? -1 : 0; m <= (periodic ? 1 : 0); ++m) {
            const double p = (i + off) * dx + m * n * dx;
            const double a = std::max(lo, p - 0.5 * dx), b = std::min(hi, p + 0.5 * dx);
            if (b > a) ww += (b - a) / dx;
        }
        if (ww > 0) {
            ids.push_back(i);
            w.push_back(ww);
        }
    }
}

// gaussian_src_time_phase::dipole / continuous_src_time::dipole (see sjf.h)
void dipole(const Source& s, double t, double* re, double* im) {
    if (s.p.kind == SJF_SRC_GAUSSIAN) {
        const double tt = t - s.peak;
        if ((double)(float)std::fabs(tt) > s.cutoff) {
            *re = *im = 0.0;
            return;
        }
        const double g = std::exp(-tt * tt / (2 * s.p.width * s.p.width));
        const double th = -s.omega * tt - s.phi;
        *re = -g * std::sin(th) / s.omega;
        *im = g * std::cos(th) / s.omega;
    } else {
        const float rt = (float)t;
        if (rt < s.p.t_start || rt > s.p.t_end) {
            *re = *im = 0.0;
            return;
        }
        const double w = 2 * M_PI * s.p.freq;
        double env = 1.0;
        if (s.p.width != 0.0) {
            const double ts = (t - s.p.t_start) / s.p.width - s.p.slowness;
            const double te = (s.p.t_end - t) / s.p.width - s.p.slowness;
            env = (1.0 + std::tanh(ts)) * (1.0 + std::tanh(te)) * 0.25;
        }
        const double th = -w * t;
        *re = -env * std::sin(th) / w;
        *im = env * std::cos(th) / w;
    }
}

double inv_eps_at(const sjf_engine* e, long q, int c) {
    if (!e->have_mat) return 1.0;
    const int m = (e->matw_h[q] >> (8 * c)) & 0xff;
    if (m == sjf::kDesignMat) {
        const int k = (int)(q % e->g.nz);
        const long r = q / e->g.nz;
        const int j = (int)(r % e->g.ny), i = (int)(r / e->g.ny);
        const long dq = ((long)(i - e->dlo[0]) * (e->dhi[1] - e->dlo[1]) + (j - e->dlo[1])) *
                            (e->dhi[2] - e->dlo[2]) +
                        (k - e->dlo[2]);
        return 1.0 / e->deps_h[dq];
    }
    return 1.0 / e->lut_h[m].eps_inf;
}

int upload_sources(sjf_engine* e) {
    if (!e->src_dirty) return SJF_OK;
    for (auto& s : e->src) {
        const long n = (long)s.idx.size();
        std::vector<double> coef(n);
        for (long q = 0; q < n; ++q) coef[q] = e->dt * s.wt[q] * inv_eps_at(e, s.idx[q], s.comp);
        if (!s.d_idx) {
            int rc = dev_alloc(e, (void**)&s.d_idx, sizeof(long) * std::max(n, 1L));
            if (rc) return rc;
            rc = dev_alloc(e, (void**)&s.d_coef, sizeof(double) * std::max(n, 1L));
            if (rc) return rc;
            HIPCHK(e, hipMemcpyAsync(s.d_idx, s.idx.data(), sizeof(long) * n,
                                     hipMemcpyHostToDevice, e->stream));
        }
        HIPCHK(e, hipMemcpyAsync(s.d_coef, coef.data(), sizeof(double) * n,
                                 hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    e->src_dirty = false;
    return SJF_OK;
}

int upload_monitors(sjf_engine* e) {
    if (!e->mon_dirty) return SJF_OK;
    const int n = (int)e->mon.size();
    std::vector<long> idx(8L * n);
    std::vector<double> w(8L * n);
    std::vector<int> comp(n);
    for (int p = 0; p < n; ++p) {
        comp[p] = e->mon[p].comp;
        for (int q = 0; q < 8; ++q) {
            idx[8 * p + q] = e->mon[p].idx[q];
            w[8 * p + q] = e->mon[p].w[q];
        }
    }
    dev_free((void*&)e->d_mon_idx);
    dev_free((void*&)e->d_mon_w);
    dev_free((void*&)e->d_mon_comp);
    int rc = dev_alloc(e, (void**)&e->d_mon_idx, sizeof(long) * 8 * n);
    if (!rc) rc = dev_alloc(e, (void**)&e->d_mon_w, sizeof(double) * 8 * n);
    if (!rc) rc = dev_alloc(e, (void**)&e->d_mon_comp, sizeof(int) * n);
    if (rc) return rc;
    HIPCHK(e, hipMemcpyAsync(e->d_mon_idx, idx.data(), sizeof(long) * 8 * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(e->d_mon_w, w.data(), sizeof(double) * 8 * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(e->d_mon_comp, comp.data(), sizeof(int) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->mon_dirty = false;
    return SJF_OK;
}

int ensure_sample_capacity(sjf_engine* e, long need) {
    const int npt = (int)e->mon.size();
    if (npt == 0 || need <= e->samp_cap) return SJF_OK;
    long cap = std::max(need, 2 * e->samp_cap);
    double* p = nullptr;
    hipError_t st = hipMalloc((void**)&p, sizeof(double) * cap * npt);
    if (st != hipSuccess) return fail(e, SJF_E_NOMEM, "sample buffer allocation failed");
    if (e->d_samples && e->nsamp > 0)
        HIPCHK(e, hipMemcpyAsync(p, e->d_samples, sizeof(double) * e->nsamp * npt,
                                 hipMemcpyDeviceToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    dev_free((void*&)e->d_samples);
    e->d_samples = p;
    e->samp_cap = cap;
    return SJF_OK;
}

template <typename T>
sjf::FieldPtrs<T> field_ptrs(const sjf_engine* e, int buf) {
    sjf::FieldPtrs<T> f;
    for (int c = 0; c < 3; ++c) {
        f.f[c] = (const T*)e->E[buf][c];
        f.f[3 + c] = (const T*)e->H[buf][c];
    }
    return f;
}

template <typename T>
sjf::StepArgs<T> step_args(const sjf_engine* e) {
    sjf::StepArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.nx = e->g.nx;
    a.ny = e->g.ny;
    a.nz = e->g.nz;
    a.px = e->g.periodic[0];
    a.py = e->g.periodic[1];
    a.pz = e->g.periodic[2];
    a.nty = (e->g.ny + e->ty - 1) / e->ty;
    a.xc = e->xc;
    a.N = e->N;
    a.cdt = (T)(e->dt / e->g.dx);
    a.F0 = (const T*)e->Fbuf[e->ping];
    a.F1 = (T*)e->Fbuf[e->ping ^ 1];
    a.uniform = e->uniform ? 1 : 0;
    a.nmat = e->nmat;
    a.matw = e->matw;
    a.lut = (const sjf::MatLut<T>*)e->lut;
    a.Q = (T*)e->Qbuf;
    a.P = (T*)e->Pbuf;
    for (int d = 0; d < 3; ++d) {
        a.pml_on[d] = e->pml_on[d] ? 1 : 0;
        a.plo[d] = e->pml[2 * d];
        a.phi[d] = e->pml[2 * d + 1];
        a.S[d] = e->slab[d];
        a.psi[d] = (T*)e->psibuf[d];
        a.coff[d] = e->coff[d];
        a.dlo[d] = e->dlo[d];
        a.dhi[d] = e->dhi[d];
    }
    // the H auxiliaries alternate between segment pairs 2-3 and 4-5 with the field sets
    a.hr = e->ping == 0 ? 2 : 4;
    a.hw = e->ping == 0 ? 4 : 2;
    a.coef = (const T*)e->coefbuf;
    a.dinv = (const T*)e->dinv;
    return a;
}

int block_threads(const sjf_engine* e) { return ((e->g.nz + 63) / 64) * 64; }

int n_groups(const sjf_engine* e) {
    const int nty = (e->g.ny + e->ty - 1) / e->ty;
    const int nxc = (e->g.nx + e->xc - 1) / e->xc;
    return nty * nxc;
}

template <typename T>
size_t step_lds(const sjf_engine* e) {
    return sizeof(T) * 4 * e->ty * block_threads(e) + sizeof(sjf::MatLut<T>) * e->nmat;
}

template <typename T, int NP, int MAXT>
int launch_step_np(sjf_engine* e, const sjf::StepArgs<T>& a, dim3 grid, dim3 block, size_t lds) {
    switch (e->ty) {
        case 1: hipLaunchKernelGGL((sjf::k_step<T, 1, NP, MAXT>), grid, block, lds, e->stream, a); break;
        case 2: hipLaunchKernelGGL((sjf::k_step<T, 2, NP, MAXT>), grid, block, lds, e->stream, a); break;
        case 4: hipLaunchKernelGGL((sjf::k_step<T, 4, NP, MAXT>), grid, block, lds, e->stream, a); break;
        default: return fail(e, SJF_E_ARG, "tile rows must be 1, 2 or 4");
    }
    HIPCHK(e, hipGetLastError());
    return SJF_OK;
}

// Pole slots compiled into the kernel: 0, 1, 3 or 4 (a material with 2 poles runs the
// 3-slot variant; unused slots are predicated off).
int np_variant(const sjf_engine* e) {
    if (e->uniform || e->maxp == 0) return 0;
    if (e->maxp == 1) return 1;
    if (e->maxp <= 3) return 3;
    return 4;
}

template <typename T, int MAXT>
int launch_step_b(sjf_engine* e, const sjf::StepArgs<T>& a, dim3 grid, dim3 block, size_t lds) {
    switch (np_variant(e)) {
        case 0: return launch_step_np<T, 0, MAXT>(e, a, grid, block, lds);
        case 1: return launch_step_np<T, 1, MAXT>(e, a, grid, block, lds);
        case 3: return launch_step_np<T, 3, MAXT>(e, a, grid, block, lds);
        default: return launch_step_np<T, 4, MAXT>(e, a, grid, block, lds);
    }
}

template <typename T>
int launch_step_t(sjf_engine* e) {
    const sjf::StepArgs<T> a = step_args<T>(e);
    const int bt = block_threads(e);
    const dim3 grid(n_groups(e)), block(bt);
    const size_t lds = step_lds<T>(e);
    if (lds > 65536) return fail(e, SJF_E_UNSUPPORTED, "tile needs more than 64 KiB of LDS; lower the rows per thread");
    if (bt <= 256) return launch_step_b<T, 256>(e, a, grid, block, lds);
    if (bt <= 512) return launch_step_b<T, 512>(e, a, grid, block, lds);
    return launch_step_b<T, 1024>(e, a, grid, block, lds);
}

int launch_step(sjf_engine* e) { return e->prec == 4 ? launch_step_t<float>(e) : launch_step_t<double>(e); }

template <typename T>
int launch_monitors_t(sjf_engine* e) {
    const int npt = (int)e->mon.size();
    hipLaunchKernelGGL((sjf::k_monitor<T>), dim3((npt + 255) / 256), dim3(256), 0, e->stream,
                       field_ptrs<T>(e, e->ping), e->d_mon_idx, e->d_mon_w, e->d_mon_comp, npt,
                       e->d_samples + e->nsamp * npt);
    HIPCHK(e, hipGetLastError());
    e->nsamp++;
    return SJF_OK;
}

template <typename T>
int launch_sources_t(sjf_engine* e) {
    // after the step: E lives in buffer ping^1; current J^{n+1/2}
    const double t0 = e->step * e->dt, t1 = (e->step + 1) * e->dt;
    for (auto& s : e->src) {
        if (s.idx.empty()) continue;
        double r0, i0, r1, i1;
        dipole(s, t0, &r0, &i0);
        dipole(s, t1, &r1, &i1);
        const double jr = (r1 - r0) / e->dt, ji = (i1 - i0) / e->dt;
        const double ar = s.amp_re * jr - s.amp_im * ji;
        const double ai = s.amp_re * ji + s.amp_im * jr;
        const double J = e->g.part ? ai : ar;
        if (J == 0.0) continue;
        const long n = (long)s.idx.size();
        hipLaunchKernelGGL((sjf::k_source<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           e->stream, (T*)e->E[e->ping ^ 1][s.comp], s.d_idx, s.d_coef, n, J);
        HIPCHK(e, hipGetLastError());
    }
    return SJF_OK;
}

template <typename T>
int launch_dft_t(sjf_engine* e) {
    for (auto& d : e->dft) {
        sjf::DftArgs<T> a;
        std::memset(&a, 0, sizeof(a));
        a.F = field_ptrs<T>(e, e->ping ^ 1);
        for (int q = 0; q < 3; ++q) {
            a.lo[q] = d.lo[q];
            a.ext[q] = d.hi[q] - d.lo[q];
        }
        a.ncell = d.ncell;
        a.ncomp = d.ncomp;
        for (int s = 0; s < 6; ++s) a.comps[s] = d.comps[s];
        a.nf = (int)d.freqs.size();
        a.acc = d.acc;
        a.cmp = d.cmp;
        const double tE = (e->step + 1) * e->dt, tH = (e->step + 0.5) * e->dt;
        for (int f = 0; f < a.nf; ++f) {
            const double w = 2 * M_PI * d.freqs[f];
            a.ph[0][f] = make_double2(std::cos(w * tE) * e->dt, std::sin(w * tE) * e->dt);
            a.ph[1][f] = make_double2(std::cos(w * tH) * e->dt, std::sin(w * tH) * e->dt);
        }
        hipLaunchKernelGGL((sjf::k_dft<T>), dim3((unsigned)((d.ncell + 255) / 256)), dim3(256), 0,
                           e->stream, a, (long)e->g.ny * e->g.nz, (long)e->g.nz);
        HIPCHK(e, hipGetLastError());
    }
    return SJF_OK;
}

template <typename T>
int one_step_t(sjf_engine* e, hipEvent_t ev_a, hipEvent_t ev_b) {
    int rc;
    if (!e->mon.empty() && e->step % e->save_span == 0) {
        rc = launch_monitors_t<T>(e);
        if (rc) return rc;
    }
    if (ev_a) HIPCHK(e, hipEventRecord(ev_a, e->stream));
    rc = launch_step_t<T>(e);
    if (rc) return rc;
    if (ev_b) HIPCHK(e, hipEventRecord(ev_b, e->stream));
    rc = launch_sources_t<T>(e);
    if (rc) return rc;
    rc = launch_dft_t<T>(e);
    if (rc) return rc;
    e->ping ^= 1;
    e->step++;
    return SJF_OK;
}

int prepare(sjf_engine* e, long n) {
    int rc = upload_sources(e);
    if (rc) return rc;
    rc = upload_monitors(e);
    if (rc) return rc;
    if (!e->mon.empty()) {
        // samples taken by steps [step, step+n): multiples of save_span in that range
        const long a = e->step, b = e->step + n;
        const long first = (a + e->save_span - 1) / e->save_span;
        const long last = (b - 1) / e->save_span;
        const long cnt = n > 0 ? std::max(0L, last - first + 1) : 0;
        rc = ensure_sample_capacity(e, e->nsamp + cnt);
        if (rc) return rc;
    }
    return SJF_OK;
}

int run_steps(sjf_engine* e, long n, std::vector<hipEvent_t>* evs) {
    int rc = prepare(e, n);
    if (rc) return rc;
    for (long s = 0; s < n; ++s) {
        hipEvent_t a = evs ? (*evs)[2 * s] : nullptr;
        hipEvent_t b = evs ? (*evs)[2 * s + 1] : nullptr;
        rc = e->prec == 4 ? one_step_t<float>(e, a, b) : one_step_t<double>(e, a, b);
        if (rc) return rc;
    }
    return SJF_OK;
}

void choose_tile(sjf_engine* e) {
    if (e->tile_set) return;
    // Defaults measured on MI355X: one or two rows per thread and short x-chunks (many
    // workgroups) beat deep register blocking; the autotuner refines per workload.
    e->ty = 1;
    const int nty = (e->g.ny + e->ty - 1) / e->ty;
    const long target_groups = 16384;
    long nxc = std::max(1L, (target_groups + nty - 1) / nty);
    nxc = std::min<long>(nxc, e->g.nx);
    int xc = (int)((e->g.nx + nxc - 1) / nxc);
    e->xc = std::max(std::min(4, e->g.nx), std::min(xc, e->g.nx));
}

int free_all(sjf_engine* e) {
    for (int b = 0; b < 2; ++b) dev_free(e->Fbuf[b]);
    dev_free((void*&)e->matw);
    dev_free(e->lut);
    dev_free(e->dinv);
    dev_free(e->Qbuf);
    dev_free(e->Pbuf);
    dev_free(e->coefbuf);
    for (int d = 0; d < 3; ++d) dev_free(e->psibuf[d]);
    for (auto& s : e->src) {
        dev_free((void*&)s.d_idx);
        dev_free((void*&)s.d_coef);
    }
    dev_free((void*&)e->d_mon_idx);
    dev_free((void*&)e->d_mon_w);
    dev_free((void*&)e->d_mon_comp);
    dev_free((void*&)e->d_samples);
    for (auto& d : e->dft) {
        dev_free((void*&)d.acc);
        dev_free((void*&)d.cmp);
    }
    return SJF_OK;
}

template <typename T>
int upload_lut_t(sjf_engine* e) {
    std::vector<sjf::MatLut<T>> l(e->nmat);
    const double dt = e->dt;
    for (int m = 0; m < e->nmat; ++m) {
        const sjf_material& M = e->lut_h[m];
        sjf::MatLut<T>& L = l[m];
        std::memset(&L, 0, sizeof(L));
        L.inv_eps = (T)(1.0 / M.eps_inf);
        L.np = M.npoles;
        for (int s = 0; s < M.npoles; ++s) {
            // coefficients formed in FP64 (sigma * w0^2 spans ~1e21 * 1e-22 for Au Drude)
            const double w0 = 2 * M_PI * M.f0[s];
            const double gm = 2 * M_PI * M.gamma[s];
            const double g = gm * dt / 2;
            const double Lf = M.is_drude[s] ? 0.0 : 1.0;
            L.alpha[s] = (T)((1 - g) / (1 + g));
            L.beta[s] = (T)(-Lf * w0 * w0 * dt * dt / (1 + g));
            L.c3[s] = (T)(M.sigma[s] * (w0 * w0) * dt * dt / (1 + g));
            if (!M.is_drude[s]) L.lorentz_mask |= 1 << s;
        }
    }
    dev_free(e->lut);
    int rc = dev_alloc(e, &e->lut, sizeof(sjf::MatLut<T>) * e->nmat);
    if (rc) return rc;
    HIPCHK(e, hipMemcpyAsync(e->lut, l.data(), sizeof(sjf::MatLut<T>) * e->nmat,
                             hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
}

int upload_lut(sjf_engine* e) { return e->prec == 4 ? upload_lut_t<float>(e) : upload_lut_t<double>(e); }

template <typename T>
int convert_to_device(sjf_engine* e, void* dst, const double* src, long n) {
    if (sizeof(T) == sizeof(double)) {
        return h2d(e, dst, src, sizeof(double) * n);
    } else {
        std::vector<T> tmp(n);
        for (long q = 0; q < n; ++q) tmp[q] = (T)src[q];
        return h2d(e, dst, tmp.data(), sizeof(T) * n);
    }
    return SJF_OK;
}

template <typename T>
int convert_from_device(sjf_engine* e, double* dst, const void* src, long n) {
    if (sizeof(T) == sizeof(double)) {
        return d2h(e, dst, src, sizeof(double) * n);
    } else {
        std::vector<T> tmp(n);
        int rc = d2h(e, tmp.data(), src, sizeof(T) * n);
        if (rc) return rc;
        for (long q = 0; q < n; ++q) dst[q] = (double)tmp[q];
    }
    return SJF_OK;
}

void trilinear(const sjf_engine* e, int comp, const double* xyz, long* idx, double* w) {
    int i0[3];
    double f[3];
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    for (int d = 0; d < 3; ++d) {
        const double s = xyz[d] / e->g.dx - kCompOff[comp][d];
        const double fl = std::floor(s);
        i0[d] = (int)fl;
        f[d] = s - fl;
    }
    int q = 0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c, ++q) {
                int ii[3] = {i0[0] + a, i0[1] + b, i0[2] + c};
                const double ww = (a ? f[0] : 1 - f[0]) * (b ? f[1] : 1 - f[1]) * (c ? f[2] : 1 - f[2]);
                bool ok = true;
                for (int d = 0; d < 3; ++d) {
                    if (e->g.periodic[d]) ii[d] = ((ii[d] % n[d]) + n[d]) % n[d];
                    else if (ii[d] < 0 || ii[d] >= n[d]) ok = false;
                }
                idx[q] = ok ? cell_index(e, ii[0], ii[1], ii[2]) : -1;
                w[q] = ok ? ww : 0.0;
            }
}

void* field_buf(sjf_engine* e, int comp) {
    return comp < 3 ? e->E[e->ping][comp] : e->H[e->ping][comp - 3];
}

}  // namespace

// =========================================================================== C ABI ====

extern "C" {

int sjf_create(const sjf_grid* g, sjf_engine** out) {
    if (!g || !out) {
        g_create_error = "null argument";
        return SJF_E_ARG;
    }
    *out = nullptr;
    if (g->nx < 1 || g->ny < 1 || g->nz < 1 || !(g->dx > 0) || !(g->courant > 0)) {
        g_create_error = "grid dimensions, dx and courant must be positive";
        return SJF_E_ARG;
    }
    if (g->nz > 1024) {
        g_create_error = "nz > 1024 is not supported (one workgroup spans the z line)";
        return SJF_E_UNSUPPORTED;
    }
    if (g->precision != SJF_FP32 && g->precision != SJF_FP64) {
        g_create_error = "precision must be SJF_FP32 (4) or SJF_FP64 (8)";
        return SJF_E_ARG;
    }
    int ndev = 0;
    hipError_t st = hipGetDeviceCount(&ndev);
    if (st != hipSuccess || ndev == 0) {
        g_create_error = std::string("no HIP device available: ") + hipGetErrorString(st);
        return SJF_E_HIP;
    }
    if (g->device < 0 || g->device >= ndev) {
        g_create_error = "device ordinal out of range";
        return SJF_E_ARG;
    }
    st = hipSetDevice(g->device);
    if (st != hipSuccess) {
        g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(st);
        return SJF_E_HIP;
    }
    sjf_engine* e = new sjf_engine();
    e->g = *g;
    e->prec = g->precision;
    e->N = (long)g->nx * g->ny * g->nz;
    e->dt = g->courant * g->dx;
    st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (st != hipSuccess) {
        g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(st);
        delete e;
        return SJF_E_HIP;
    }
    for (int b = 0; b < 2; ++b) {
        if (dev_alloc(e, &e->Fbuf[b], (size_t)e->prec * 6 * e->N)) {
            g_create_error = e->err;
            free_all(e);
            hipStreamDestroy(e->stream);
            delete e;
            return SJF_E_NOMEM;
        }
        for (int c = 0; c < 3; ++c) {
            e->E[b][c] = (char*)e->Fbuf[b] + (size_t)e->prec * c * e->N;
            e->H[b][c] = (char*)e->Fbuf[b] + (size_t)e->prec * (3 + c) * e->N;
        }
    }
    // PML tables for all axes (b = 1, a = 0 until sjf_set_pml fills an axis)
    {
        const int n3[3] = {g->nx, g->ny, g->nz};
        int off = 0;
        for (int d = 0; d < 3; ++d) {
            e->coff[d] = off;
            off += 4 * n3[d];
        }
        std::vector<double> tab(off, 0.0);
        for (int d = 0; d < 3; ++d)
            for (int i = 0; i < n3[d]; ++i) {
                tab[e->coff[d] + i] = 1.0;
                tab[e->coff[d] + 2 * n3[d] + i] = 1.0;
            }
        int rc = dev_alloc(e, &e->coefbuf, (size_t)e->prec * off);
        if (!rc) rc = e->prec == 4 ? convert_to_device<float>(e, e->coefbuf, tab.data(), off)
                                   : convert_to_device<double>(e, e->coefbuf, tab.data(), off);
        if (rc) {
            g_create_error = e->err;
            free_all(e);
            hipStreamDestroy(e->stream);
            delete e;
            return rc;
        }
    }
    // default material: vacuum
    e->lut_h.assign(1, sjf_material{});
    e->lut_h[0].eps_inf = 1.0;
    e->nmat = 1;
    if (upload_lut(e)) {
        g_create_error = e->err;
        free_all(e);
        hipStreamDestroy(e->stream);
        delete e;
        return SJF_E_HIP;
    }
    choose_tile(e);
    hipStreamSynchronize(e->stream);
    *out = e;
    return SJF_OK;
}

int sjf_destroy(sjf_engine* e) {
    if (!e) return SJF_E_ARG;
    hipSetDevice(e->g.device);
    hipStreamSynchronize(e->stream);
    free_all(e);
    hipStreamDestroy(e->stream);
    delete e;
    return SJF_OK;
}

const char* sjf_last_error(const sjf_engine* e) {
    return e ? e->err.c_str() : g_create_error.c_str();
}

int sjf_set_materials(sjf_engine* e, const uint32_t* mat_word, const sjf_material* lut, int n_mat) {
    if (!e || !mat_word || !lut) return SJF_E_ARG;
    if (n_mat < 1 || n_mat > SJF_MAX_MATERIALS) return fail(e, SJF_E_ARG, "n_mat must be in [1, 255]");
    hipSetDevice(e->g.device);
    int maxp = 0;
    bool lor[kMaxPoles] = {};
    for (int m = 0; m < n_mat; ++m) {
        if (lut[m].npoles < 0 || lut[m].npoles > SJF_MAX_POLES)
            return fail(e, SJF_E_ARG, "npoles must be in [0, 4]");
        if (!(lut[m].eps_inf > 0)) return fail(e, SJF_E_ARG, "eps_inf must be positive");
        maxp = std::max(maxp, lut[m].npoles);
        for (int s = 0; s < lut[m].npoles; ++s)
            if (!lut[m].is_drude[s]) lor[s] = true;
    }
    for (long q = 0; q < e->N; ++q)
        for (int c = 0; c < 3; ++c) {
            const int m = (mat_word[q] >> (8 * c)) & 0xff;
            if (m >= n_mat && m != sjf::kDesignMat)
                return fail(e, SJF_E_ARG, "material word references an id >= n_mat");
            if (m == sjf::kDesignMat && e->deps_h.empty())
                return fail(e, SJF_E_STATE, "design id 255 used before sjf_set_design_eps");
        }
    e->lut_h.assign(lut, lut + n_mat);
    e->nmat = n_mat;
    e->matw_h.assign(mat_word, mat_word + e->N);
    e->have_mat = true;
    e->maxp = maxp;
    e->uniform = (n_mat == 1 && maxp == 0 && e->deps_h.empty());
    int rc = dev_alloc(e, (void**)&e->matw, sizeof(uint32_t) * e->N);
    if (rc) return rc;
    rc = h2d(e, e->matw, mat_word, sizeof(uint32_t) * e->N);
    if (rc) return rc;
    // polarisation state: dense per-slot arrays, zeroed
    dev_free(e->Qbuf);
    dev_free(e->Pbuf);
    if (maxp > 0) {
        bool anylor = false;
        for (int s = 0; s < maxp; ++s) anylor |= lor[s];
        rc = dev_alloc(e, &e->Qbuf, (size_t)e->prec * 3 * maxp * e->N);
        if (!rc && anylor) rc = dev_alloc(e, &e->Pbuf, (size_t)e->prec * 3 * maxp * e->N);
        if (rc) return rc;
    }
    for (int s = 0; s < kMaxPoles; ++s) e->lorentz_slot[s] = lor[s];
    rc = upload_lut(e);
    if (rc) return rc;
    e->src_dirty = true;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
}

int sjf_set_design_eps(sjf_engine* e, const int lo[3], const int hi[3], const double* eps) {
    if (!e || !lo || !hi || !eps) return SJF_E_ARG;
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    long cnt = 1;
    for (int d = 0; d < 3; ++d) {
        if (lo[d] < 0 || hi[d] > n[d] || hi[d] <= lo[d]) return fail(e, SJF_E_ARG, "design box outside grid");
        cnt *= hi[d] - lo[d];
    }
    if (!e->deps_h.empty()) {
        for (int d = 0; d < 3; ++d)
            if (lo[d] != e->dlo[d] || hi[d] != e->dhi[d])
                return fail(e, SJF_E_STATE, "design box cannot change once set");
    }
    for (long q = 0; q < cnt; ++q)
        if (!(eps[q] > 0)) return fail(e, SJF_E_ARG, "design eps must be positive");
    hipSetDevice(e->g.device);
    for (int d = 0; d < 3; ++d) {
        e->dlo[d] = lo[d];
        e->dhi[d] = hi[d];
    }
    e->deps_h.assign(eps, eps + cnt);
    std::vector<double> inv(cnt);
    for (long q = 0; q < cnt; ++q) inv[q] = 1.0 / eps[q];
    int rc = dev_alloc(e, &e->dinv, (size_t)e->prec * cnt);
    if (rc) return rc;
    rc = e->prec == 4 ? convert_to_device<float>(e, e->dinv, inv.data(), cnt)
                      : convert_to_device<double>(e, e->dinv, inv.data(), cnt);
    if (rc) return rc;
    // mark design voxels in the material word (all three E components)
    if (!e->have_mat) {
        e->matw_h.assign(e->N, 0u);
        e->have_mat = true;
        rc = dev_alloc(e, (void**)&e->matw, sizeof(uint32_t) * e->N);
        if (rc) return rc;
    }
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k) e->matw_h[cell_index(e, i, j, k)] = 0x00ffffffu;
    rc = h2d(e, e->matw, e->matw_h.data(), sizeof(uint32_t) * e->N);
    if (rc) return rc;
    e->uniform = false;
    e->src_dirty = true;
    return SJF_OK;
}

int sjf_set_pml(sjf_engine* e, const int cells[6], double R, double power) {
    if (!e || !cells) return SJF_E_ARG;
    if (!(R > 0 && R < 1) || !(power >= 0)) return fail(e, SJF_E_ARG, "R must be in (0,1), power >= 0");
    if (e->step != 0) return fail(e, SJF_E_STATE, "PML must be set before the first step");
    hipSetDevice(e->g.device);
    for (int d = 0; d < 3; ++d) {
        const int n = axis_n(e, d);
        const int lo = cells[2 * d], hi = cells[2 * d + 1];
        if (lo < 0 || hi < 0 || lo + hi > n) return fail(e, SJF_E_ARG, "PML thicker than the grid");
        if ((lo > 0 || hi > 0) && e->g.periodic[d])
            return fail(e, SJF_E_ARG, "PML on a periodic axis");
    }
    std::vector<double> tab;
    int off = 0;
    for (int d = 0; d < 3; ++d) off += 4 * axis_n(e, d);
    tab.assign(off, 0.0);
    for (int d = 0; d < 3; ++d) {
        const int n = axis_n(e, d);
        const int lo = cells[2 * d], hi = cells[2 * d + 1];
        e->pml[2 * d] = lo;
        e->pml[2 * d + 1] = hi;
        e->pml_on[d] = (lo > 0 || hi > 0);
        dev_free(e->psibuf[d]);
        e->slab[d] = 0;
        const double dx = e->g.dx;
        auto sigma = [&](double x) {
            double s = 0;
            if (lo > 0) {
                const double L = lo * dx, u = L - x;
                if (u > 0) s += -(power + 1) * std::log(R) / (2 * L) * std::pow(u / L, power);
            }
            if (hi > 0) {
                const double L = hi * dx, u = x - (n - hi) * dx;
                if (u > 0) s += -(power + 1) * std::log(R) / (2 * L) * std::pow(u / L, power);
            }
            return s;
        };
        for (int i = 0; i < n; ++i) {
            const double bE = std::exp(-sigma(i * dx) * e->dt);
            const double bH = std::exp(-sigma((i + 0.5) * dx) * e->dt);
            tab[e->coff[d] + i] = bE;
            tab[e->coff[d] + n + i] = bE - 1.0;
            tab[e->coff[d] + 2 * n + i] = bH;
            tab[e->coff[d] + 3 * n + i] = bH - 1.0;
        }
        if (!e->pml_on[d]) continue;
        e->slab[d] = (long)(lo + hi) * (e->N / n);
        int rc = dev_alloc(e, &e->psibuf[d], (size_t)e->prec * 6 * e->slab[d]);
        if (rc) return rc;
    }
    {
        int rc = e->prec == 4 ? convert_to_device<float>(e, e->coefbuf, tab.data(), off)
                              : convert_to_device<double>(e, e->coefbuf, tab.data(), off);
        if (rc) return rc;
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
}

int sjf_add_source(sjf_engine* e, int comp, const double box[6], const sjf_pulse* p, double amp_re,
                   double amp_im) {
    if (!e || !box || !p) return SJF_E_ARG;
    if (comp < 0 || comp > 2) return fail(e, SJF_E_UNSUPPORTED, "only electric sources are supported");
    if (p->kind != SJF_SRC_GAUSSIAN && p->kind != SJF_SRC_CONTINUOUS) return fail(e, SJF_E_ARG, "bad pulse kind");
    if (!(p->freq > 0)) return fail(e, SJF_E_ARG, "pulse frequency must be positive");
    if (p->kind == SJF_SRC_GAUSSIAN && !(p->width > 0)) return fail(e, SJF_E_ARG, "pulse width must be positive");
    Source s;
    s.comp = comp;
    s.p = *p;
    s.amp_re = amp_re;
    s.amp_im = amp_im;
    if (p->kind == SJF_SRC_GAUSSIAN) {
        s.omega = 2 * M_PI * p->freq;
        s.phi = p->phase + M_PI;
        s.peak = 0.5 * (p->t_start + p->t_end);
        s.cutoff = (p->t_end - p->t_start) * 0.5;
        while (std::exp(-s.cutoff * s.cutoff / (2 * p->width * p->width)) < 1e-100) s.cutoff *= 0.9;
        s.cutoff = (double)(float)s.cutoff;
    }
    std::vector<int> ids[3];
    std::vector<double> ws[3];
    for (int d = 0; d < 3; ++d)
        axis_weights(axis_n(e, d), e->g.dx, kCompOff[comp][d], e->g.periodic[d], box[d], box[3 + d],
                     ids[d], ws[d]);
    for (size_t a = 0; a < ids[0].size(); ++a)
        for (size_t b = 0; b < ids[1].size(); ++b)
            for (size_t c = 0; c < ids[2].size(); ++c) {
                s.idx.push_back(cell_index(e, ids[0][a], ids[1][b], ids[2][c]));
                s.wt.push_back(ws[0][a] * ws[1][b] * ws[2][c]);
            }
    e->src.push_back(std::move(s));
    e->src_dirty = true;
    return SJF_OK;
}

int sjf_last_source_time(const sjf_engine* e, double* t) {
    if (!e || !t) return SJF_E_ARG;
    double m = 0;
    for (const auto& s : e->src) {
        const double lt = s.p.kind == SJF_SRC_GAUSSIAN ? (double)(float)(s.peak + s.cutoff) : s.p.t_end;
        m = std::max(m, lt);
    }
    *t = m;
    return SJF_OK;
}

int sjf_dt(const sjf_engine* e, double* dt) {
    if (!e || !dt) return SJF_E_ARG;
    *dt = e->dt;
    return SJF_OK;
}

int sjf_time(const sjf_engine* e, double* t) {
    if (!e || !t) return SJF_E_ARG;
    *t = e->step * e->dt;
    return SJF_OK;
}

int sjf_step_count(const sjf_engine* e, long* n) {
    if (!e || !n) return SJF_E_ARG;
    *n = e->step;
    return SJF_OK;
}

int sjf_add_monitors(sjf_engine* e, int comp, const double* xyz, int n) {
    if (!e || (!xyz && n > 0) || n < 0) return SJF_E_ARG;
    if (comp < 0 || comp > 5) return fail(e, SJF_E_ARG, "bad component");
    if (e->nsamp > 0) return fail(e, SJF_E_STATE, "monitors must be added before sampling starts");
    for (int p = 0; p < n; ++p) {
        Monitor m;
        m.comp = comp;
        trilinear(e, comp, xyz + 3 * p, m.idx, m.w);
        e->mon.push_back(m);
    }
    e->mon_dirty = true;
    return SJF_OK;
}

int sjf_set_save_span(sjf_engine* e, int span) {
    if (!e) return SJF_E_ARG;
    e->save_span = span > 0 ? span : 1;
    return SJF_OK;
}

int sjf_read_monitors(sjf_engine* e, double* out, size_t cap, long* n_samples) {
    if (!e || !n_samples) return SJF_E_ARG;
    *n_samples = e->nsamp;
    const long npt = (long)e->mon.size();
    const long tot = e->nsamp * npt;
    if (!out || tot == 0) return SJF_OK;
    if ((size_t)(2 * tot) > cap) return fail(e, SJF_E_ARG, "output buffer too small");
    hipSetDevice(e->g.device);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    std::vector<double> tmp(tot);
    int rc = d2h(e, tmp.data(), e->d_samples, sizeof(double) * tot);
    if (rc) return rc;
    for (long q = 0; q < tot; ++q) {
        out[2 * q] = tmp[q];
        out[2 * q + 1] = 0.0;
    }
    return SJF_OK;
}

int sjf_get_field(sjf_engine* e, int comp, const double xyz[3], double out[2]) {
    if (!e || !xyz || !out || comp < 0 || comp > 5) return SJF_E_ARG;
    hipSetDevice(e->g.device);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    long idx[8];
    double w[8];
    trilinear(e, comp, xyz, idx, w);
    double v = 0;
    for (int q = 0; q < 8; ++q) {
        if (idx[q] < 0) continue;
        const char* base = (const char*)field_buf(e, comp) + (size_t)idx[q] * e->prec;
        if (e->prec == 4) {
            float f;
            int rc = d2h(e, &f, base, 4);
            if (rc) return rc;
            v += w[q] * f;
        } else {
            double f;
            int rc = d2h(e, &f, base, 8);
            if (rc) return rc;
            v += w[q] * f;
        }
    }
    out[0] = v;
    out[1] = 0.0;
    return SJF_OK;
}

int sjf_add_dft(sjf_engine* e, const int lo[3], const int hi[3], int comp_mask, const double* freqs,
                int nf, int* id) {
    if (!e || !lo || !hi || !freqs || !id) return SJF_E_ARG;
    if (nf < 1 || nf > SJF_MAX_DFT_FREQS) return fail(e, SJF_E_ARG, "nf must be in [1, 64]");
    if (comp_mask <= 0 || comp_mask >= 64) return fail(e, SJF_E_ARG, "comp_mask must select components 0..5");
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    Dft d;
    d.ncell = 1;
    for (int q = 0; q < 3; ++q) {
        if (lo[q] < 0 || hi[q] > n[q] || hi[q] <= lo[q]) return fail(e, SJF_E_ARG, "DFT box outside grid");
        d.lo[q] = lo[q];
        d.hi[q] = hi[q];
        d.ncell *= hi[q] - lo[q];
    }
    d.mask = comp_mask;
    d.freqs.assign(freqs, freqs + nf);
    d.ncomp = 0;
    for (int c = 0; c < 6; ++c)
        if (comp_mask >> c & 1) d.comps[d.ncomp++] = c;
    hipSetDevice(e->g.device);
    const size_t bytes = sizeof(double2) * d.ncomp * nf * d.ncell;
    int rc = dev_alloc(e, (void**)&d.acc, bytes);
    if (!rc) rc = dev_alloc(e, (void**)&d.cmp, bytes);
    if (rc) return rc;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->dft.push_back(d);
    *id = (int)e->dft.size() - 1;
    return SJF_OK;
}

int sjf_read_dft(sjf_engine* e, int id, double* out, size_t cap) {
    if (!e || !out || id < 0 || id >= (int)e->dft.size()) return SJF_E_ARG;
    const Dft& d = e->dft[id];
    const long n = (long)d.ncomp * (long)d.freqs.size() * d.ncell;
    if ((size_t)(2 * n) > cap) return fail(e, SJF_E_ARG, "output buffer too small");
    hipSetDevice(e->g.device);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    std::vector<double2> a(n), c(n);
    int rc = d2h(e, a.data(), d.acc, sizeof(double2) * n);
    if (!rc) rc = d2h(e, c.data(), d.cmp, sizeof(double2) * n);
    if (rc) return rc;
    for (long q = 0; q < n; ++q) {
        out[2 * q] = a[q].x + c[q].x;
        out[2 * q + 1] = a[q].y + c[q].y;
    }
    return SJF_OK;
}

int sjf_step(sjf_engine* e, long n) {
    if (!e || n < 0) return SJF_E_ARG;
    hipSetDevice(e->g.device);
    return run_steps(e, n, nullptr);
}

int sjf_synchronize(sjf_engine* e) {
    if (!e) return SJF_E_ARG;
    hipSetDevice(e->g.device);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
}

int sjf_timed_steps(sjf_engine* e, long n, double* ms_total, double* ms_kernel_mean) {
    if (!e || n < 1 || !ms_total || !ms_kernel_mean) return SJF_E_ARG;
    hipSetDevice(e->g.device);
    int rc = prepare(e, n);
    if (rc) return rc;
    std::vector<hipEvent_t> evs(2 * n);
    for (auto& ev : evs) HIPCHK(e, hipEventCreate(&ev));
    hipEvent_t t0, t1;
    HIPCHK(e, hipEventCreate(&t0));
    HIPCHK(e, hipEventCreate(&t1));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipEventRecord(t0, e->stream));
    rc = run_steps(e, n, &evs);
    if (rc) return rc;
    HIPCHK(e, hipEventRecord(t1, e->stream));
    HIPCHK(e, hipEventSynchronize(t1));
    float tot = 0;
    HIPCHK(e, hipEventElapsedTime(&tot, t0, t1));
    double ksum = 0;
    for (long s = 0; s < n; ++s) {
        float ms = 0;
        HIPCHK(e, hipEventElapsedTime(&ms, evs[2 * s], evs[2 * s + 1]));
        ksum += ms;
    }
    for (auto& ev : evs) hipEventDestroy(ev);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    *ms_total = tot;
    *ms_kernel_mean = ksum / n;
    return SJF_OK;
}

int sjf_read_field(sjf_engine* e, int comp, double* out) {
    if (!e || !out || comp < 0 || comp > 5) return SJF_E_ARG;
    hipSetDevice(e->g.device);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    void* src = field_buf(e, comp);
    return e->prec == 4 ? convert_from_device<float>(e, out, src, e->N)
                        : convert_from_device<double>(e, out, src, e->N);
}

int sjf_write_field(sjf_engine* e, int comp, const double* in) {
    if (!e || !in || comp < 0 || comp > 5) return SJF_E_ARG;
    hipSetDevice(e->g.device);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    void* dst = field_buf(e, comp);
    return e->prec == 4 ? convert_to_device<float>(e, dst, in, e->N)
                        : convert_to_device<double>(e, dst, in, e->N);
}

int sjf_step_bytes(const sjf_engine* e, double* bytes) {
    if (!e || !bytes) return SJF_E_ARG;
    const double p = e->prec;
    double b = e->N * (12 * p + (e->uniform ? 0 : 4));
    // polarisation: Drude pole = Q read+write, Lorentz pole = P and Q read+write
    if (e->have_mat && e->maxp > 0) {
        for (long q = 0; q < e->N; ++q)
            for (int c = 0; c < 3; ++c) {
                const int m = (e->matw_h[q] >> (8 * c)) & 0xff;
                if (m == sjf::kDesignMat) continue;
                const sjf_material& M = e->lut_h[m];
                for (int s = 0; s < M.npoles; ++s) b += (M.is_drude[s] ? 2 : 4) * p;
            }
    }
    if (!e->deps_h.empty()) b += p * (double)e->deps_h.size();
    // PML: per cell and PML axis, 2 E + 2 H auxiliaries, each read + written
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    for (int d = 0; d < 3; ++d)
        if (e->pml_on[d]) b += 8 * p * (double)(e->pml[2 * d] + e->pml[2 * d + 1]) * (e->N / n[d]);
    *bytes = b;
    return SJF_OK;
}

int sjf_launch_info(const sjf_engine* e, int* ty, int* xchunk, int* threads, int* groups) {
    if (!e) return SJF_E_ARG;
    if (ty) *ty = e->ty;
    if (xchunk) *xchunk = e->xc;
    if (threads) *threads = block_threads(e);
    if (groups) *groups = n_groups(e);
    return SJF_OK;
}

int sjf_set_tile(sjf_engine* e, int ty, int xchunk) {
    if (!e) return SJF_E_ARG;
    if (ty != 1 && ty != 2 && ty != 4) return fail(e, SJF_E_ARG, "ty must be 1, 2 or 4");
    if (xchunk < 1) return fail(e, SJF_E_ARG, "xchunk must be >= 1");
    e->ty = ty;
    e->xc = std::min(xchunk, e->g.nx);
    e->tile_set = true;
    return SJF_OK;
}

int sjf_copy_probe(int device, size_t bytes, int iters, double* gbs) {
    if (!gbs || bytes < 16 || iters < 1) return SJF_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return SJF_E_HIP;
    const long n = (long)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n * 16) != hipSuccess) return SJF_E_NOMEM;
    if (hipMalloc((void**)&b, n * 16) != hipSuccess) {
        hipFree(a);
        return SJF_E_NOMEM;
    }
    hipMemset(a, 0, n * 16);
    hipMemset(b, 0, n * 16);
    hipEvent_t t0, t1;
    hipEventCreate(&t0);
    hipEventCreate(&t1);
    const int blocks = 256 * 8 * 4;
    hipLaunchKernelGGL(sjf::k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
    hipEventRecord(t0, 0);
    for (int it = 0; it < iters; ++it) hipLaunchKernelGGL(sjf::k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
    hipEventRecord(t1, 0);
    hipError_t st = hipEventSynchronize(t1);
    float ms = 0;
    hipEventElapsedTime(&ms, t0, t1);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    hipFree(a);
    hipFree(b);
    if (st != hipSuccess) return SJF_E_HIP;
    *gbs = 2.0 * n * 16 * iters / (ms * 1e-3) / 1e9;
    return SJF_OK;
}

}  // extern "C"
