// sjf_engine.hip -- host side of the MI355X FDTD engine behind include/sjf.h.
//
// An engine owns the device state of one simulation (fields, materials, PML auxiliaries,
// polarisation state, sources, monitors, DFT accumulators).  Stepping goes through a launch
// set: one engine, or a batch of engines, compiled into small device tables so that each
// time step is a fixed sequence of launches on one stream,
//   [k_monitor when step % save_span == 0] -> k_step -> [k_source] -> [k_dft] (or k_tick),
// the last of which advances the step counters, whatever the number of instances; source and
// DFT share one launch (k_source_dft) when no DFT region holds a source cell.  The per-step
// scalars (step index, pulse value, DFT phase) are computed on the device from each engine's
// step counter, so a run of steps is
// captured once in a hipGraph and replayed.  This is the sequence bound_geom::run performs
// around fields.step() (reference src/disp.cpp:719-741).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
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
    sjf::PulseDev dev;
    std::vector<long> idx;
    std::vector<double> wt;
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
    double* d_freq = nullptr;
};

struct LaunchSet;

}  // namespace

struct sjf_engine {
    sjf_grid g{};
    int prec = 4;
    long N = 0;
    double dt = 0;
    hipStream_t stream = nullptr;
    std::string err;
    long version = 0;  // bumped by every change that invalidates compiled launch tables

    void* Fbuf[2] = {};  // field sets [x][Ex Ey Ez Hx Hy Hz][y][z]; step n reads set n&1
    long* d_step = nullptr;
    long step = 0;

    // materials: given_* as the caller set them; matw_h / lut_h / nmat / maxp are what the kernels
    // run (the given map, with Lorentz poles folded out of the PML slabs when pml_fref > 0)
    bool have_mat = false;
    std::vector<uint32_t> given_w;
    std::vector<sjf_material> given_lut;
    double pml_fref = 0;
    std::vector<uint32_t> matw_h;
    std::vector<sjf_material> lut_h;
    uint32_t* matw = nullptr;
    void* lut = nullptr;
    int nmat = 1;
    int maxp = 0;
    void* Qbuf = nullptr;  // [x][slot * 3 + comp][y][z]
    void* Pbuf = nullptr;  // the same, present when a slot holds a Lorentz pole
    bool uniform = true;
    int dlo[3] = {0, 0, 0}, dhi[3] = {0, 0, 0};
    std::vector<double> deps_h;
    void* dinv = nullptr;
    std::vector<double> emap_h;  // per-cell eps_inf [Ex Ey Ez][x][y][z] as given (sjf_set_eps_maps)
    std::vector<double> emap_eff;  // as run (Lorentz poles folded in inside the PML slabs)
    void* emap = nullptr;          // 1/eps_inf on the device, [x][Ex Ey Ez][y][z]

    // PML
    int pml[6] = {0, 0, 0, 0, 0, 0};
    bool pml_on[3] = {false, false, false};
    void* coefbuf = nullptr;  // per axis: bE[n], aE[n], bH[n], aH[n]
    int coff[3] = {0, 0, 0};
    long slab[3] = {0, 0, 0};
    void* psibuf[3] = {};  // per axis [6][slab]: E aux x2, H aux pair A x2, pair B x2

    std::vector<Source> src;
    std::vector<Monitor> mon;
    int save_span = 1;
    long nsamp = 0, samp_cap = 0;
    double* d_samples = nullptr;
    std::vector<Dft> dft;

    int form = SJF_TILE_REGISTER_ROWS;  // fused-kernel form
    int ty = 1, xc = 16;                // y rows per thread (register form) or per workgroup
    int zseg = 0;                       // register form: z cells per workgroup pass (0: default)
    bool tile_set = false;
    int graph_chunk = 32;  // steps per captured graph (0: no graphs)

    // fused reverse-step kernel of the last adjoint run: mean duration and launch shape
    double adj_ms = 0;
    int adj_threads = 0, adj_groups = 0;
    // its tile (sjf_set_adjoint_tile): form, owned rows of a row-of-threads tile, x-chunk (0: default)
    int adj_form = SJF_TILE_REGISTER_ROWS;
    int adj_rows = 0, adj_xc = 0;
    bool adj_set = false;  // set by sjf_set_adjoint_tile (else the default of adj_tile_of)

    LaunchSet* own = nullptr;
    sjf_batch* batch = nullptr;
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

template <typename P>
void dev_free(P*& p) {
    if (p) (void)hipFree((void*)p);
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

// the field set read by the next step
void* field_set(sjf_engine* e) { return e->Fbuf[e->step & 1]; }

// byte offset of component comp of cell q in a field set
size_t field_offset(const sjf_engine* e, int comp, long q) {
    return (size_t)e->prec * sjf::field_index(q, comp, (long)e->g.ny * e->g.nz);
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
    if (!e->emap_eff.empty()) return 1.0 / e->emap_eff[(size_t)c * e->N + q];
    return 1.0 / e->lut_h[m].eps_inf;
}

// Row-of-threads form: (rows + 1) thread rows of nz threads per workgroup, 32-bit offsets.
bool thread_rows_fit(const sjf_engine* e, int rows) { return rows >= 1 && (long)(rows + 1) * e->g.nz <= 1024; }

// Field bytes one step of these engines moves (E and H read and written) above which the one-row
// step kernels store with the non-temporal hint.  Measured on MI355X (FP32): 1024x768x512 7.17 ->
// 6.86 ms per step, 512x512x256 1.230 -> 1.205 ms, the 256x256x128 reverse step 0.282 -> 0.267
// ms, one 192x192x96 design variant (170 MB of fields) 12.46 -> 12.75 G cell-updates/s per
// gradient; the 96^3 FP64 box (85 MB, which the 256 MB Infinity Cache holds from one step to the
// next) 0.046 -> 0.059 ms with the hint, so it keeps plain stores.
constexpr double kStreamBytes = 128e6;
double field_bytes(const sjf_engine* e) { return 12.0 * e->prec * (double)e->N; }

// z cells per pass of a register-rows workgroup with `rows` rows per thread: the set value, else
// 128-cell passes for one row per thread on a non-periodic z line that is a multiple of 256
// cells above 256 or longer than 1024, else the whole line.  Measured on MI355X, 1024x768x512
// FP32: one pass of 512 threads leaves two 8-wave workgroups per CU (4 waves per SIMD of the 5
// the registers allow) and runs 8.07 ms per step; two passes of 256 threads (5 workgroups per
// CU) run 7.31 ms.  With the non-temporal stores, two boxes: 256-cell passes 6.45 and 6.87 ms,
// 128-cell passes (ten 2-wave workgroups per CU) 6.20 and 6.69 ms, 64-cell passes 6.80 ms.
// On 512x512x256 one pass wins: 128-cell passes 1.295 against 1.183 ms.
int zseg_of(const sjf_engine* e, int rows) {
    const int nz = e->g.nz;
    if (e->zseg > 0) return std::min(e->zseg, nz);
    if (rows == 1 && !e->g.periodic[2] && nz > 256 && (nz % 256 == 0 || nz > 1024)) return 128;
    return nz;
}

// FP64 tiles stop at 512 threads: at 1024 the register budget of 128 per thread spills
int default_thread_rows(const sjf_engine* e) { return std::max(1, (e->prec == 8 ? 512 : 1024) / e->g.nz - 1); }

int block_threads_for(const sjf_engine* e, int form, int ty) {
    const int t = form == SJF_TILE_THREAD_ROWS ? (ty + 1) * e->g.nz : zseg_of(e, ty);
    return ((t + 63) / 64) * 64;
}

int block_threads(const sjf_engine* e) { return block_threads_for(e, e->form, e->ty); }

int tiles_of(const sjf_engine* e, int ty, int xc) {
    const int nty = (e->g.ny + ty - 1) / ty;
    const int nxc = (e->g.nx + xc - 1) / xc;
    return nty * nxc;
}

// Pole slots compiled into the step kernel: 0, 1, 3 or 4 (a material with 2 poles runs the
// 3-slot variant; unused slots are predicated off).
int np_variant_of(int maxp, bool uniform) {
    if (uniform || maxp == 0) return 0;
    if (maxp == 1) return 1;
    if (maxp <= 3) return 3;
    return 4;
}

template <typename T>
sjf::StepArgs<T> step_args(const sjf_engine* e, int ty, int xc) {
    sjf::StepArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.rt = ty + 1;
    a.zseg = zseg_of(e, ty);
    a.nx = e->g.nx;
    a.ny = e->g.ny;
    a.nz = e->g.nz;
    a.px = e->g.periodic[0];
    a.py = e->g.periodic[1];
    a.pz = e->g.periodic[2];
    a.nty = (e->g.ny + ty - 1) / ty;
    a.xc = xc;
    a.N = e->N;
    a.cdt = (T)(e->dt / e->g.dx);
    a.F[0] = (T*)e->Fbuf[0];
    a.F[1] = (T*)e->Fbuf[1];
    a.d_step = e->d_step;
    a.uniform = e->uniform ? 1 : 0;
    a.nmat = e->nmat;
    a.matw = e->matw;
    a.lut = (const sjf::MatLut<T>*)e->lut;
    a.Q = (T*)e->Qbuf;
    a.P = (T*)e->Pbuf;
    a.nqp = e->Qbuf ? 3 * e->maxp : 0;
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
    a.coef = (const T*)e->coefbuf;
    a.dinv = (const T*)e->dinv;
    a.emap = (const T*)e->emap;
    return a;
}

// ======================================================================== launch sets =====

struct LaunchSet {
    std::vector<sjf_engine*> eng;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int prec = 4, ty = 1, npv = 0, block = 64;
    int form = SJF_TILE_REGISTER_ROWS;
    int xc = 0;  // x-chunk of every instance of a batch (0: each engine's own)
    bool zpass = false;  // register form with some z line stepped in passes (k_step_z)
    bool nts = false;    // one-row tiles store with the non-temporal hint (field_bytes > kStreamBytes)
    size_t lds = 0;
    std::vector<long> versions;
    bool built = false;
    // step kernel tables
    void* d_args = nullptr;
    int* d_tile_first = nullptr;
    int ntiles = 0;
    // sources
    void* d_src = nullptr;
    int* d_blk_src = nullptr;
    long* d_blk_first = nullptr;
    long* d_blk_end = nullptr;
    long* d_src_idx = nullptr;
    double* d_src_coef = nullptr;
    int nsrcblk = 0;
    // magnetic sources (injected into H before the fused update), same table layout
    void* d_hsrc = nullptr;
    int* d_hblk_src = nullptr;
    long* d_hblk_first = nullptr;
    long* d_hblk_end = nullptr;
    long* d_hsrc_idx = nullptr;
    double* d_hsrc_coef = nullptr;
    int nhsrcblk = 0;
    // monitors
    void* d_mon = nullptr;
    int nmon = 0;
    // DFT
    void* d_dft = nullptr;
    int ndft = 0;
    long dft_total = 0;
    int dft_maxnf = 0;  // most frequencies of any region (the launch's second dimension)
    bool src_dft_one_launch = false;  // no DFT region holds a source cell: one launch for both
    // step counters, and the sign-off word of the last kernel of a step that advances them
    long** d_counters = nullptr;
    unsigned* d_done = nullptr;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    long gK = 0;
};

void set_free_tables(LaunchSet* s) {
    dev_free(s->d_args);
    dev_free(s->d_tile_first);
    dev_free(s->d_src);
    dev_free(s->d_blk_src);
    dev_free(s->d_blk_first);
    dev_free(s->d_blk_end);
    dev_free(s->d_src_idx);
    dev_free(s->d_src_coef);
    dev_free(s->d_hsrc);
    dev_free(s->d_hblk_src);
    dev_free(s->d_hblk_first);
    dev_free(s->d_hblk_end);
    dev_free(s->d_hsrc_idx);
    dev_free(s->d_hsrc_coef);
    s->nsrcblk = s->nhsrcblk = 0;
    dev_free(s->d_mon);
    dev_free(s->d_dft);
    dev_free(s->d_counters);
    dev_free(s->d_done);
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->gexec = nullptr;
    s->graph = nullptr;
    s->gK = 0;
    s->built = false;
}

template <typename V>
int upload_vec(sjf_engine* e, hipStream_t st, void** dst, const std::vector<V>& v) {
    const size_t bytes = sizeof(V) * std::max<size_t>(v.size(), 1);
    hipError_t rc = hipMalloc(dst, bytes);
    if (rc != hipSuccess) return fail(e, SJF_E_NOMEM, "launch table allocation failed");
    if (!v.empty()) {
        rc = hipMemcpyAsync(*dst, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice, st);
        if (rc == hipSuccess) rc = hipStreamSynchronize(st);
        if (rc != hipSuccess) return fail(e, SJF_E_HIP, hipGetErrorString(rc));
    }
    return SJF_OK;
}

template <typename T>
int build_set_t(LaunchSet* s) {
    sjf_engine* e0 = s->eng[0];
    set_free_tables(s);
    // step kernel
    std::vector<sjf::StepArgs<T>> args;
    std::vector<int> first;
    int maxp = 0, block = 64;
    bool all_uniform = true;
    int t = 0;
    size_t zh = 0;  // register form with z passes: values of the pass halo in LDS
    for (sjf_engine* e : s->eng) {
        // the engine's own set runs its tuned rows; a batch picks rows per engine for its form
        const int rows = (s == e->own) ? e->ty
                                       : (s->form == SJF_TILE_THREAD_ROWS ? default_thread_rows(e) : s->ty);
        const int xc = (s == e->own || s->xc <= 0) ? e->xc : std::min(s->xc, e->g.nx);
        args.push_back(step_args<T>(e, rows, xc));
        first.push_back(t);
        t += tiles_of(e, rows, xc);
        maxp = std::max(maxp, e->maxp);
        all_uniform = all_uniform && e->uniform;
        block = std::max(block, block_threads_for(e, s->form, rows));
        if (s->form == SJF_TILE_REGISTER_ROWS && zseg_of(e, rows) < e->g.nz) {
            if (rows != 1) return fail(e0, SJF_E_UNSUPPORTED, "z passes run one row per thread");
            zh = std::max<size_t>(zh, 4 * (size_t)xc);
        }
        if (block_threads_for(e, s->form, rows) > 1024)
            return fail(e0, SJF_E_UNSUPPORTED, "a z line longer than 1024 cells needs one row per thread (z passes)");
    }
    s->zpass = zh > 0;
    double fb = 0;
    for (sjf_engine* e : s->eng) fb += field_bytes(e);
    s->nts = fb > kStreamBytes;
    first.push_back(t);
    s->ntiles = t;
    s->block = block;
    s->npv = np_variant_of(maxp, all_uniform);
    int nmat = 1;
    for (sjf_engine* e : s->eng) nmat = std::max(nmat, e->nmat);
    s->lds = (s->form == SJF_TILE_THREAD_ROWS ? sizeof(T) * 6 * block : sizeof(T) * 4 * s->ty * block) +
             sizeof(sjf::MatLut<T>) * nmat + sizeof(T) * zh;
    if (s->lds > 65536)
        return fail(e0, SJF_E_UNSUPPORTED, "tile needs more than 64 KiB of LDS; use fewer rows");
    int rc = upload_vec(e0, s->stream, &s->d_args, args);
    if (!rc) rc = upload_vec(e0, s->stream, (void**)&s->d_tile_first, first);
    if (rc) return rc;

    // sources: entries of every source of every engine, in blocks of 256 of one source; electric
    // sources (coefficient dt * weight / eps) and magnetic ones (dt * weight, mu = 1) in separate
    // tables
    for (int mag = 0; mag < 2; ++mag) {
        std::vector<sjf::SrcDev<T>> srcs;
        std::vector<int> blk_src;
        std::vector<long> blk_first, blk_end, idx;
        std::vector<double> coef;
        for (sjf_engine* e : s->eng) {
            for (const Source& src : e->src) {
                if (src.idx.empty() || (src.comp >= 3) != (mag == 1)) continue;
                sjf::SrcDev<T> d;
                std::memset(&d, 0, sizeof(d));
                d.p = src.dev;
                d.p.part = e->g.part;
                d.p.dt = e->dt;
                for (int b = 0; b < 2; ++b) d.F[b] = (T*)e->Fbuf[b];
                d.d_step = e->d_step;
                const int sid = (int)srcs.size();
                srcs.push_back(d);
                const long base = (long)idx.size();
                for (size_t q = 0; q < src.idx.size(); ++q) {
                    idx.push_back(sjf::field_index(src.idx[q], src.comp, (long)e->g.ny * e->g.nz));
                    coef.push_back(e->dt * src.wt[q] * (mag ? 1.0 : inv_eps_at(e, src.idx[q], src.comp)));
                }
                const long n = (long)src.idx.size();
                for (long b = 0; b < n; b += 256) {
                    blk_src.push_back(sid);
                    blk_first.push_back(base + b);
                    blk_end.push_back(base + std::min(n, b + 256));
                }
            }
        }
        const int nblk = (int)blk_src.size();
        if (!nblk) continue;
        void** tsrc = mag ? &s->d_hsrc : &s->d_src;
        int** tbs = mag ? &s->d_hblk_src : &s->d_blk_src;
        long** tbf = mag ? &s->d_hblk_first : &s->d_blk_first;
        long** tbe = mag ? &s->d_hblk_end : &s->d_blk_end;
        long** tidx = mag ? &s->d_hsrc_idx : &s->d_src_idx;
        double** tcoef = mag ? &s->d_hsrc_coef : &s->d_src_coef;
        (mag ? s->nhsrcblk : s->nsrcblk) = nblk;
        rc = upload_vec(e0, s->stream, tsrc, srcs);
        if (!rc) rc = upload_vec(e0, s->stream, (void**)tbs, blk_src);
        if (!rc) rc = upload_vec(e0, s->stream, (void**)tbf, blk_first);
        if (!rc) rc = upload_vec(e0, s->stream, (void**)tbe, blk_end);
        if (!rc) rc = upload_vec(e0, s->stream, (void**)tidx, idx);
        if (!rc) rc = upload_vec(e0, s->stream, (void**)tcoef, coef);
        if (rc) return rc;
    }

    // monitors
    std::vector<sjf::MonDev<T>> mons;
    for (sjf_engine* e : s->eng) {
        for (size_t p = 0; p < e->mon.size(); ++p) {
            sjf::MonDev<T> m;
            std::memset(&m, 0, sizeof(m));
            const int c = e->mon[p].comp;
            for (int b = 0; b < 2; ++b) m.F[b] = (const T*)e->Fbuf[b];
            m.d_step = e->d_step;
            m.out = e->d_samples;
            m.npt = (int)e->mon.size();
            m.p = (int)p;
            m.span = e->save_span;
            for (int q = 0; q < 8; ++q) {
                const long cq = e->mon[p].idx[q];
                m.idx[q] = cq >= 0 ? sjf::field_index(cq, c, (long)e->g.ny * e->g.nz) : -1;
                m.w[q] = e->mon[p].w[q];
            }
            mons.push_back(m);
        }
    }
    s->nmon = (int)mons.size();
    if (s->nmon) {
        rc = upload_vec(e0, s->stream, &s->d_mon, mons);
        if (rc) return rc;
    }

    // DFT regions
    std::vector<sjf::DftDev<T>> regs;
    long tot = 0;
    for (sjf_engine* e : s->eng) {
        for (const Dft& d : e->dft) {
            sjf::DftDev<T> r;
            std::memset(&r, 0, sizeof(r));
            for (int b = 0; b < 2; ++b) r.F[b] = (const T*)e->Fbuf[b];
            r.d_step = e->d_step;
            r.N = e->N;
            r.nyz = (long)e->g.ny * e->g.nz;
            r.nz = e->g.nz;
            for (int q = 0; q < 3; ++q) {
                r.lo[q] = d.lo[q];
                r.ext[q] = d.hi[q] - d.lo[q];
            }
            r.ncell = d.ncell;
            r.first = tot;
            for (int q = 0; q < 6; ++q) r.comps[q] = d.comps[q];
            r.ncomp = d.ncomp;
            r.nf = (int)d.freqs.size();
            r.freq = d.d_freq;
            r.dt = e->dt;
            r.acc = d.acc;
            r.cmp = d.cmp;
            regs.push_back(r);
            tot += d.ncell * d.ncomp;  // one thread per (component, cell) of each frequency
        }
    }
    s->ndft = (int)regs.size();
    s->dft_total = tot;
    // sources and DFT can share a launch when no DFT region of an engine holds one of its
    // source cells
    s->src_dft_one_launch = s->nsrcblk > 0 && s->ndft > 0;
    for (sjf_engine* e : s->eng) {
        const long nyz = (long)e->g.ny * e->g.nz;
        for (const Source& src : e->src)
            for (const long q : (src.comp < 3 ? src.idx : std::vector<long>())) {
                const int i = (int)(q / nyz), j = (int)((q % nyz) / e->g.nz), k = (int)(q % e->g.nz);
                for (const Dft& d : e->dft)
                    if (i >= d.lo[0] && i < d.hi[0] && j >= d.lo[1] && j < d.hi[1] && k >= d.lo[2] && k < d.hi[2])
                        s->src_dft_one_launch = false;
            }
    }
    s->dft_maxnf = 0;
    for (const auto& r : regs) s->dft_maxnf = std::max(s->dft_maxnf, r.nf);
    if (s->ndft) {
        rc = upload_vec(e0, s->stream, &s->d_dft, regs);
        if (rc) return rc;
    }

    std::vector<long*> ctr;
    for (sjf_engine* e : s->eng) ctr.push_back(e->d_step);
    rc = upload_vec(e0, s->stream, (void**)&s->d_counters, ctr);
    if (rc) return rc;
    rc = upload_vec(e0, s->stream, (void**)&s->d_done, std::vector<unsigned>(sjf::kDoneWords, 0u));  // armed at zero
    if (rc) return rc;

    s->versions.clear();
    for (sjf_engine* e : s->eng) s->versions.push_back(e->version);
    s->built = true;
    return SJF_OK;
}

int build_set(LaunchSet* s) { return s->prec == 4 ? build_set_t<float>(s) : build_set_t<double>(s); }

bool set_stale(const LaunchSet* s) {
    if (!s->built) return true;
    for (size_t i = 0; i < s->eng.size(); ++i)
        if (s->eng[i]->version != s->versions[i]) return true;
    return false;
}

#define SETCHK(s, call)                                                                  \
    do {                                                                                 \
        hipError_t _st = (call);                                                         \
        if (_st != hipSuccess) {                                                         \
            (s)->eng[0]->err = std::string(#call) + ": " + hipGetErrorString(_st);       \
            return SJF_E_HIP;                                                            \
        }                                                                                \
    } while (0)

template <typename T, int NP, int MAXT>
int launch_step_np(LaunchSet* s) {
    const dim3 grid(s->ntiles), block(s->block);
    const auto* args = (const sjf::StepArgs<T>*)s->d_args;
    const int n = (int)s->eng.size();
    if (s->form == SJF_TILE_THREAD_ROWS) {
        hipLaunchKernelGGL((sjf::k_step_rows<T, NP, MAXT>), grid, block, s->lds, s->stream, args, s->d_tile_first, n);
        SETCHK(s, hipGetLastError());
        return SJF_OK;
    }
    if (s->zpass) {
        if (s->nts)
            hipLaunchKernelGGL((sjf::k_step_z<T, NP, MAXT, true>), grid, block, s->lds, s->stream, args, s->d_tile_first, n);
        else
            hipLaunchKernelGGL((sjf::k_step_z<T, NP, MAXT>), grid, block, s->lds, s->stream, args, s->d_tile_first, n);
        SETCHK(s, hipGetLastError());
        return SJF_OK;
    }
    switch (s->ty) {
        case 1:
            if (s->nts)
                hipLaunchKernelGGL((sjf::k_step<T, 1, NP, MAXT, true>), grid, block, s->lds, s->stream, args, s->d_tile_first, n);
            else
                hipLaunchKernelGGL((sjf::k_step<T, 1, NP, MAXT>), grid, block, s->lds, s->stream, args, s->d_tile_first, n);
            break;
        case 2: hipLaunchKernelGGL((sjf::k_step<T, 2, NP, MAXT>), grid, block, s->lds, s->stream, args, s->d_tile_first, n); break;
        case 4: hipLaunchKernelGGL((sjf::k_step<T, 4, NP, MAXT>), grid, block, s->lds, s->stream, args, s->d_tile_first, n); break;
        default: return fail(s->eng[0], SJF_E_ARG, "tile rows must be 1, 2 or 4");
    }
    SETCHK(s, hipGetLastError());
    return SJF_OK;
}

template <typename T, int MAXT>
int launch_step_b(LaunchSet* s) {
    switch (s->npv) {
        case 0: return launch_step_np<T, 0, MAXT>(s);
        case 1: return launch_step_np<T, 1, MAXT>(s);
        case 3: return launch_step_np<T, 3, MAXT>(s);
        default: return launch_step_np<T, 4, MAXT>(s);
    }
}

template <typename T>
int launch_step_t(LaunchSet* s) {
    if (s->block <= 256) return launch_step_b<T, 256>(s);
    if (s->block <= 512) return launch_step_b<T, 512>(s);
    return launch_step_b<T, 1024>(s);
}

// One time step of every engine in the set.  `sample` launches the monitor gather first.
template <typename T>
int launch_one_step_t(LaunchSet* s, bool sample, hipEvent_t ev_a, hipEvent_t ev_b, bool bare = false) {
    if (sample && s->nmon && !bare) {
        hipLaunchKernelGGL((sjf::k_monitor<T>), dim3((s->nmon + 255) / 256), dim3(256), 0, s->stream,
                           (const sjf::MonDev<T>*)s->d_mon, s->nmon);
        SETCHK(s, hipGetLastError());
    }
    if (s->nhsrcblk) {
        const sjf::SrcTables<T> ht{(const sjf::SrcDev<T>*)s->d_hsrc, s->d_hblk_src, s->d_hblk_first,
                                   s->d_hblk_end, s->d_hsrc_idx, s->d_hsrc_coef};
        hipLaunchKernelGGL((sjf::k_source_h<T>), dim3(s->nhsrcblk), dim3(256), 0, s->stream, ht);
        SETCHK(s, hipGetLastError());
    }
    if (ev_a) SETCHK(s, hipEventRecord(ev_a, s->stream));
    int rc = launch_step_t<T>(s);
    if (rc) return rc;
    if (ev_b) SETCHK(s, hipEventRecord(ev_b, s->stream));
    // the last kernel of the step advances the step counters: the DFT kernel when there is
    // one, else the source kernel, else k_tick
    const int n = (int)s->eng.size();
    const bool dft = s->ndft && !bare;
    const sjf::SrcTables<T> st{(const sjf::SrcDev<T>*)s->d_src, s->d_blk_src, s->d_blk_first, s->d_blk_end,
                               s->d_src_idx, s->d_src_coef};
    const sjf::Tick tick{s->d_counters, n, s->d_done};
    const int dftx = (int)((s->dft_total + 255) / 256);
    if (dft && s->src_dft_one_launch) {
        hipLaunchKernelGGL((sjf::k_source_dft<T>), dim3((unsigned)(s->nsrcblk + dftx * s->dft_maxnf)), dim3(256), 0,
                           s->stream, st, s->nsrcblk, (const sjf::DftDev<T>*)s->d_dft, s->ndft, s->dft_total, dftx,
                           tick);
        SETCHK(s, hipGetLastError());
        return SJF_OK;
    }
    if (s->nsrcblk) {
        const sjf::Tick src_tick{s->d_counters, n, dft ? nullptr : s->d_done};
        hipLaunchKernelGGL((sjf::k_source<T>), dim3(s->nsrcblk), dim3(256), 0, s->stream, st, src_tick);
        SETCHK(s, hipGetLastError());
    }
    if (dft) {
        hipLaunchKernelGGL((sjf::k_dft<T>), dim3((unsigned)dftx, (unsigned)s->dft_maxnf), dim3(256), 0, s->stream,
                           (const sjf::DftDev<T>*)s->d_dft, s->ndft, s->dft_total, tick);
        SETCHK(s, hipGetLastError());
    } else if (!s->nsrcblk) {
        hipLaunchKernelGGL(sjf::k_tick, dim3((n + 63) / 64), dim3(64), 0, s->stream, s->d_counters, n);
        SETCHK(s, hipGetLastError());
    }
    return SJF_OK;
}

int launch_one_step(LaunchSet* s, bool sample, hipEvent_t a = nullptr, hipEvent_t b = nullptr) {
    return s->prec == 4 ? launch_one_step_t<float>(s, sample, a, b) : launch_one_step_t<double>(s, sample, a, b);
}

// step kernel + sources + counter only (forward recomputation inside the adjoint)
int launch_bare_step(LaunchSet* s) {
    return s->prec == 4 ? launch_one_step_t<float>(s, false, nullptr, nullptr, true)
                        : launch_one_step_t<double>(s, false, nullptr, nullptr, true);
}

bool samples_at(const LaunchSet* s, long step) {
    for (const sjf_engine* e : s->eng)
        if (!e->mon.empty() && step % e->save_span == 0) return true;
    return false;
}

int sync_engine(sjf_engine* e);

int ensure_sample_capacity(sjf_engine* e, long need) {
    const int npt = (int)e->mon.size();
    if (npt == 0 || need <= e->samp_cap) return SJF_OK;
    // the old samples may still be in flight on a batch stream
    if (int rc = sync_engine(e)) return rc;
    const long cap = std::max(need, 2 * e->samp_cap);
    double* p = nullptr;
    if (hipMalloc((void**)&p, sizeof(double) * cap * npt) != hipSuccess)
        return fail(e, SJF_E_NOMEM, "sample buffer allocation failed");
    if (e->d_samples && e->nsamp > 0)
        HIPCHK(e, hipMemcpyAsync(p, e->d_samples, sizeof(double) * e->nsamp * npt, hipMemcpyDeviceToDevice,
                                 e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    dev_free(e->d_samples);
    e->d_samples = p;
    e->samp_cap = cap;
    e->version++;
    return SJF_OK;
}

long samples_in(long a, long b, int span) {
    // multiples of span in [a, b)
    if (b <= a) return 0;
    const long first = (a + span - 1) / span, last = (b - 1) / span;
    return std::max(0L, last - first + 1);
}

int prepare_set(LaunchSet* s, long n) {
    for (sjf_engine* e : s->eng) {
        if (e->mon.empty()) continue;
        const long need = (e->step + n + e->save_span - 1) / e->save_span;
        int rc = ensure_sample_capacity(e, need);
        if (rc) return rc;
    }
    if (set_stale(s)) return build_set(s);
    return SJF_OK;
}

void advance_hosts(LaunchSet* s, long n) {
    for (sjf_engine* e : s->eng) {
        if (!e->mon.empty()) e->nsamp += samples_in(e->step, e->step + n, e->save_span);
        e->step += n;
    }
}

long set_span(const LaunchSet* s) {
    long span = 1;
    for (const sjf_engine* e : s->eng)
        if (!e->mon.empty()) span = e->save_span;
    return span;
}

// Steps per captured graph: a multiple of the save span so every replay starts aligned.
long graph_chunk(const LaunchSet* s) {
    const sjf_engine* e0 = s->eng[0];
    if (e0->graph_chunk <= 0) return 0;
    const long span = set_span(s);
    return std::max(1L, (e0->graph_chunk + span - 1) / span) * span;
}

int capture_graph(LaunchSet* s, long K) {
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->gexec = nullptr;
    s->graph = nullptr;
    const long s0 = s->eng[0]->step;
    SETCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = SJF_OK;
    for (long q = 0; q < K && !rc; ++q) rc = launch_one_step(s, samples_at(s, s0 + q));
    hipGraph_t g = nullptr;
    const hipError_t st = hipStreamEndCapture(s->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    SETCHK(s, st);
    s->graph = g;
    SETCHK(s, hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
    s->gK = K;
    return SJF_OK;
}

int run_set(LaunchSet* s, long n) {
    int rc = prepare_set(s, n);
    if (rc) return rc;
    const long K = graph_chunk(s);
    const long span = set_span(s);
    long done = 0;
    auto eager = [&](long m) -> int {
        for (long q = 0; q < m; ++q) {
            const int r = launch_one_step(s, samples_at(s, s->eng[0]->step));
            if (r) return r;
            advance_hosts(s, 1);
        }
        return SJF_OK;
    };
    const long lead = K > 0 ? (span - s->eng[0]->step % span) % span : 0;
    if (K > 0 && n - lead >= K) {
        // align the first replay to the save span, then replay whole chunks
        rc = eager(lead);
        if (rc) return rc;
        done += lead;
        if (!s->gexec || s->gK != K) {
            rc = capture_graph(s, K);
            if (rc) return rc;
        }
        while (n - done >= K) {
            SETCHK(s, hipGraphLaunch(s->gexec, s->stream));
            advance_hosts(s, K);
            done += K;
        }
    }
    return eager(n - done);
}

LaunchSet* own_set(sjf_engine* e) {
    if (!e->own) {
        e->own = new LaunchSet();
        e->own->eng.push_back(e);
        e->own->stream = e->stream;
        e->own->prec = e->prec;
    }
    if (e->own->ty != e->ty || e->own->form != e->form) {
        e->own->ty = e->ty;
        e->own->form = e->form;
        e->own->built = false;
    }
    return e->own;
}

void destroy_set(LaunchSet* s) {
    if (!s) return;
    set_free_tables(s);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int timed_set(LaunchSet* s, long n, double* ms_total, double* ms_kernel_mean) {
    int rc = prepare_set(s, n);
    if (rc) return rc;
    std::vector<hipEvent_t> evs(2 * n);
    for (auto& ev : evs) SETCHK(s, hipEventCreate(&ev));
    hipEvent_t t0, t1;
    SETCHK(s, hipEventCreate(&t0));
    SETCHK(s, hipEventCreate(&t1));
    SETCHK(s, hipStreamSynchronize(s->stream));
    SETCHK(s, hipEventRecord(t0, s->stream));
    for (long q = 0; q < n; ++q) {
        rc = launch_one_step(s, samples_at(s, s->eng[0]->step), evs[2 * q], evs[2 * q + 1]);
        if (rc) return rc;
        advance_hosts(s, 1);
    }
    SETCHK(s, hipEventRecord(t1, s->stream));
    SETCHK(s, hipEventSynchronize(t1));
    float tot = 0;
    SETCHK(s, hipEventElapsedTime(&tot, t0, t1));
    double ksum = 0;
    for (long q = 0; q < n; ++q) {
        float ms = 0;
        SETCHK(s, hipEventElapsedTime(&ms, evs[2 * q], evs[2 * q + 1]));
        ksum += ms;
    }
    for (auto& ev : evs) (void)hipEventDestroy(ev);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    *ms_total = tot;
    *ms_kernel_mean = ksum / n;
    return SJF_OK;
}

// ======================================================================== engine helpers ==

void choose_tile(sjf_engine* e) {
    if (e->tile_set) return;
    // Defaults measured on MI355X: for short z lines (nz < 96) the row-of-threads form wins
    // (up to ~1024 threads per workgroup, few halo rows); from 96 cells on one row per thread
    // keeps more waves resident (192x192x96 FP32: 0.120 ms against 0.126 ms for 4 thread rows;
    // 96^3 FP64: 0.0467 against 0.0485 ms).  The autotuner refines per workload.
    long target_groups = 16384;
    e->form = SJF_TILE_REGISTER_ROWS;
    e->ty = 1;
    const int rows = e->g.nz > 96 ? default_thread_rows(e) : std::min(3, default_thread_rows(e));
    if (e->g.nz < 96 && thread_rows_fit(e, rows)) {
        e->form = SJF_TILE_THREAD_ROWS;
        e->ty = rows;
        target_groups = 2048;
    }
    const int nty = (e->g.ny + e->ty - 1) / e->ty;
    long nxc = std::max(1L, (target_groups + nty - 1) / nty);
    nxc = std::min<long>(nxc, e->g.nx);
    const int xc = (int)((e->g.nx + nxc - 1) / nxc);
    e->xc = std::max(std::min(4, e->g.nx), std::min(xc, e->g.nx));
}

int free_all(sjf_engine* e) {
    destroy_set(e->own);
    e->own = nullptr;
    for (int b = 0; b < 2; ++b) dev_free(e->Fbuf[b]);
    dev_free(e->d_step);
    dev_free(e->matw);
    dev_free(e->lut);
    dev_free(e->dinv);
    dev_free(e->emap);
    dev_free(e->Qbuf);
    dev_free(e->Pbuf);
    dev_free(e->coefbuf);
    for (int d = 0; d < 3; ++d) dev_free(e->psibuf[d]);
    dev_free(e->d_samples);
    for (auto& d : e->dft) {
        dev_free(d.acc);
        dev_free(d.cmp);
        dev_free(d.d_freq);
    }
    return SJF_OK;
}

template <typename T>
int convert_to_device(sjf_engine* e, void* dst, const double* src, long n) {
    if (sizeof(T) == sizeof(double)) return h2d(e, dst, src, sizeof(double) * n);
    std::vector<T> tmp(n);
    for (long q = 0; q < n; ++q) tmp[q] = (T)src[q];
    return h2d(e, dst, tmp.data(), sizeof(T) * n);
}

// One component of the current field set to / from a dense z-fastest host array: the component
// is one plane of ny * nz values in each x plane of six, a pitched copy.
template <typename T>
int write_component(sjf_engine* e, int comp, const double* src) {
    const size_t plane = sizeof(T) * (size_t)e->g.ny * e->g.nz;
    std::vector<T> tmp(e->N);
    for (long q = 0; q < e->N; ++q) tmp[q] = (T)src[q];
    char* dst = (char*)field_set(e) + plane * comp;
    HIPCHK(e, hipMemcpy2DAsync(dst, 6 * plane, tmp.data(), plane, plane, e->g.nx, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
}

template <typename T>
int read_component(sjf_engine* e, int comp, double* dst) {
    const size_t plane = sizeof(T) * (size_t)e->g.ny * e->g.nz;
    std::vector<T> tmp(e->N);
    const char* src = (const char*)field_set(e) + plane * comp;
    HIPCHK(e, hipMemcpy2DAsync(tmp.data(), plane, src, 6 * plane, plane, e->g.nx, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (long q = 0; q < e->N; ++q) dst[q] = (double)tmp[q];
    return SJF_OK;
}

// Copies the cell box [lo, hi) of one component: one strided 2-D copy (y rows of the box's z
// span) per x plane of the box into a pinned-free host staging array, then converts to FP64.
template <typename T>
int read_component_box(sjf_engine* e, int comp, const int* lo, const int* hi, double* dst) {
    const size_t plane = sizeof(T) * (size_t)e->g.ny * e->g.nz;
    const size_t bx = (size_t)(hi[0] - lo[0]), by = (size_t)(hi[1] - lo[1]), bz = (size_t)(hi[2] - lo[2]);
    std::vector<T> tmp(bx * by * bz);
    const char* base = (const char*)field_set(e) + plane * comp;
    for (size_t i = 0; i < bx; ++i) {
        const char* src = base + 6 * plane * (size_t)(lo[0] + i) + sizeof(T) * ((size_t)lo[1] * e->g.nz + lo[2]);
        HIPCHK(e, hipMemcpy2DAsync(tmp.data() + i * by * bz, sizeof(T) * bz, src, sizeof(T) * (size_t)e->g.nz,
                                   sizeof(T) * bz, by, hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (size_t q = 0; q < tmp.size(); ++q) dst[q] = (double)tmp[q];
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
        L.eps_inf = (T)M.eps_inf;
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
    return h2d(e, e->lut, l.data(), sizeof(sjf::MatLut<T>) * e->nmat);
}

int upload_lut(sjf_engine* e) { return e->prec == 4 ? upload_lut_t<float>(e) : upload_lut_t<double>(e); }

// 1/eps_inf of the effective map on the device in the plane layout [x][Ex Ey Ez][y][z]
template <typename T>
int upload_emap_t(sjf_engine* e) {
    const long nyz = (long)e->g.ny * e->g.nz;
    std::vector<T> h((size_t)3 * e->N);
    for (int i = 0; i < e->g.nx; ++i)
        for (int c = 0; c < 3; ++c)
            for (long r = 0; r < nyz; ++r)
                h[((size_t)i * 3 + c) * nyz + r] = (T)(1.0 / e->emap_eff[(size_t)c * e->N + (long)i * nyz + r]);
    int rc = dev_alloc(e, &e->emap, sizeof(T) * 3 * e->N);
    if (rc) return rc;
    return h2d(e, e->emap, h.data(), sizeof(T) * 3 * e->N);
}

int upload_emap(sjf_engine* e) { return e->prec == 4 ? upload_emap_t<float>(e) : upload_emap_t<double>(e); }

// Real part of a Lorentz pole's susceptibility at meep frequency f:
// sigma w0^2 (w0^2 - w^2) / ((w0^2 - w^2)^2 + w^2 gamma'^2), w = 2 pi f, w0 = 2 pi f0, gamma' = 2 pi gamma.
double lorentz_re_chi(const sjf_material& M, int s, double f) {
    const double w = 2 * M_PI * f, w0 = 2 * M_PI * M.f0[s], g = 2 * M_PI * M.gamma[s];
    const double d = w0 * w0 - w * w;
    return M.sigma[s] * w0 * w0 * d / (d * d + w * w * g * g);
}

bool cell_in_pml(const sjf_engine* e, long q) {
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    const long nyz = (long)n[1] * n[2];
    const int ijk[3] = {(int)(q / nyz), (int)((q % nyz) / n[2]), (int)(q % n[2])};
    for (int d = 0; d < 3; ++d)
        if (e->pml_on[d] && (ijk[d] < e->pml[2 * d] || ijk[d] >= n[d] - e->pml[2 * d + 1])) return true;
    return false;
}

// The material map the kernels run: the given map, except that with pml_fref > 0 every cell inside
// a PML slab whose material has Lorentz poles takes a twin of that material with them folded out
// (their real susceptibility at pml_fref added to eps_inf when the sum stays >= 1, else eps_inf kept;
// Drude poles kept).  Uploads the words and the LUT and resets the polarisation state.
int apply_materials(sjf_engine* e) {
    const int ng = (int)e->given_lut.size();
    e->lut_h = e->given_lut;
    e->matw_h = e->given_w;
    bool any_pml = e->pml_on[0] || e->pml_on[1] || e->pml_on[2];
    std::vector<double> fold_add(ng, 0.0);  // sum of the folded Lorentz poles' Re chi per material
    std::vector<int> twin(ng, -1);
    const bool fold = e->pml_fref > 0 && any_pml;
    if (fold) {
        for (int m = 0; m < ng; ++m) {
            const sjf_material& M = e->given_lut[m];
            sjf_material T = M;
            T.npoles = 0;
            double eps = M.eps_inf;
            bool has_lor = false;
            for (int s = 0; s < M.npoles; ++s) {
                if (M.is_drude[s]) {
                    const int t = T.npoles++;
                    T.f0[t] = M.f0[s];
                    T.gamma[t] = M.gamma[s];
                    T.sigma[t] = M.sigma[s];
                    T.is_drude[t] = 1;
                } else {
                    has_lor = true;
                    eps += lorentz_re_chi(M, s, e->pml_fref);
                }
            }
            if (!has_lor) continue;
            fold_add[m] = eps - M.eps_inf;
            T.eps_inf = eps >= 1.0 ? eps : M.eps_inf;
            for (int s = T.npoles; s < SJF_MAX_POLES; ++s) T.f0[s] = T.gamma[s] = T.sigma[s] = 0, T.is_drude[s] = 0;
            if ((int)e->lut_h.size() >= SJF_MAX_MATERIALS)
                return fail(e, SJF_E_UNSUPPORTED, "materials and their PML twins exceed 255");
            twin[m] = (int)e->lut_h.size();
            e->lut_h.push_back(T);
        }
        for (long q = 0; q < e->N; ++q) {
            if (!cell_in_pml(e, q)) continue;
            uint32_t w = e->matw_h[q];
            for (int c = 0; c < 3; ++c) {
                const int m = (w >> (8 * c)) & 0xff;
                if (m < ng && twin[m] >= 0) w = (w & ~(0xffu << (8 * c))) | ((uint32_t)twin[m] << (8 * c));
            }
            e->matw_h[q] = w;
        }
    }
    // per-cell eps_inf map: the given values, with the same fold inside the PML slabs
    if (!e->emap_h.empty()) {
        e->emap_eff = e->emap_h;
        if (fold)
            for (long q = 0; q < e->N; ++q) {
                if (!cell_in_pml(e, q)) continue;
                for (int c = 0; c < 3; ++c) {
                    const int m = (e->given_w[q] >> (8 * c)) & 0xff;
                    if (m >= ng || twin[m] < 0) continue;
                    double& v = e->emap_eff[(size_t)c * e->N + q];
                    if (v + fold_add[m] >= 1.0) v += fold_add[m];
                }
            }
        if (int rc = upload_emap(e)) return rc;
    }
    int maxp = 0;
    bool anylor = false;
    for (const sjf_material& M : e->lut_h) {
        maxp = std::max(maxp, M.npoles);
        for (int s = 0; s < M.npoles; ++s)
            if (!M.is_drude[s]) anylor = true;
    }
    e->nmat = (int)e->lut_h.size();
    e->maxp = maxp;
    e->uniform = (e->nmat == 1 && maxp == 0 && e->deps_h.empty() && e->emap_h.empty());
    int rc = dev_alloc(e, (void**)&e->matw, sizeof(uint32_t) * e->N);
    if (!rc) rc = h2d(e, e->matw, e->matw_h.data(), sizeof(uint32_t) * e->N);
    if (rc) return rc;
    // polarisation state, zeroed
    dev_free(e->Qbuf);
    dev_free(e->Pbuf);
    if (maxp > 0) {
        rc = dev_alloc(e, &e->Qbuf, (size_t)e->prec * 3 * maxp * e->N);
        if (!rc && anylor) rc = dev_alloc(e, &e->Pbuf, (size_t)e->prec * 3 * maxp * e->N);
        if (rc) return rc;
    }
    rc = upload_lut(e);
    if (rc) return rc;
    e->version++;
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



double engine_step_bytes(const sjf_engine* e) {
    const double p = e->prec;
    double b = e->N * (12 * p + (e->uniform ? 0 : 4));
    // polarisation: per pole and dispersive component 3p, the compulsory state of the textbook
    // recurrence (read P^n and P^{n-1}, write P^{n+1})
    if (e->have_mat && e->maxp > 0) {
        long pc = 0;
        for (long q = 0; q < e->N; ++q)
            for (int c = 0; c < 3; ++c) {
                const int m = (e->matw_h[q] >> (8 * c)) & 0xff;
                if (m != sjf::kDesignMat) pc += e->lut_h[m].npoles;
            }
        b += 3 * p * (double)pc;
    }
    if (!e->deps_h.empty()) b += p * (double)e->deps_h.size();
    if (!e->emap_h.empty()) b += 3 * p * (double)e->N;  // per-cell 1/eps_inf of the three E components
    // PML: per cell and PML axis, 2 E + 2 H auxiliaries, each read + written
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    for (int d = 0; d < 3; ++d)
        if (e->pml_on[d]) b += 8 * p * (double)(e->pml[2 * d] + e->pml[2 * d + 1]) * (e->N / n[d]);
    return b;
}

}  // namespace


// ======================================================================== adjoint =========

namespace {

// Everything that evolves with the time step (the current field set, polarisation state and all
// PML auxiliaries), in bytes.
struct StateLayout {
    size_t fields = 0, q = 0, p = 0, psi[3] = {0, 0, 0}, total = 0;
};

StateLayout state_layout(const sjf_engine* e) {
    StateLayout L;
    L.fields = (size_t)e->prec * 6 * e->N;
    L.q = e->Qbuf ? (size_t)e->prec * 3 * e->maxp * e->N : 0;
    L.p = e->Pbuf ? (size_t)e->prec * 3 * e->maxp * e->N : 0;
    L.total = L.fields + L.q + L.p;
    for (int d = 0; d < 3; ++d) {
        L.psi[d] = e->psibuf[d] ? (size_t)e->prec * 6 * e->slab[d] : 0;
        L.total += L.psi[d];
    }
    return L;
}

// (on the engine's stream, or on `st` when given: a batch's stream)
int copy_state(sjf_engine* e, char* ck, const StateLayout& L, bool save, hipStream_t st = nullptr) {
    if (!st) st = e->stream;
    auto cp = [&](void* dev, size_t off, size_t bytes) -> hipError_t {
        if (!bytes) return hipSuccess;
        return save ? hipMemcpyAsync(ck + off, dev, bytes, hipMemcpyDeviceToDevice, st)
                    : hipMemcpyAsync(dev, ck + off, bytes, hipMemcpyDeviceToDevice, st);
    };
    size_t off = 0;
    HIPCHK(e, cp(e->Fbuf[e->step & 1], off, L.fields));
    off += L.fields;
    HIPCHK(e, cp(e->Qbuf, off, L.q));
    off += L.q;
    HIPCHK(e, cp(e->Pbuf, off, L.p));
    off += L.p;
    for (int d = 0; d < 3; ++d) {
        HIPCHK(e, cp(e->psibuf[d], off, L.psi[d]));
        off += L.psi[d];
    }
    return SJF_OK;
}

int set_step(sjf_engine* e, long n) {
    e->step = n;
    HIPCHK(e, hipMemcpyAsync(e->d_step, &e->step, sizeof(long), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
}

// x-chunk of the fused reverse step (one y row per workgroup) unless sjf_set_adjoint_tile sets
// one: enough chunks for about `target` workgroups over `rows` y rows (the engine's own, or those
// of every engine of a batch), independent of the forward tile.  Measured on MI355X: alone,
// 256x256x128 runs 1.019 ms per gradient step at 4 planes (16384 workgroups), 1.041 at 8 and
// 1.13 at 16; sixteen 192x192x96 engines in one batch run the reverse launch in 1.96 ms at 3
// planes (196608 workgroups), 1.82 ms at 6 and 1.79 ms at 12 on one box, and in 1.61 ms at 18
// planes (33792 workgroups) against 1.64-1.68 ms at 12 and 1.69-1.71 ms at 32 on another.
int adj_xchunk(const sjf_engine* e, long target = 16384, long rows = 0) {
    if (rows <= 0) rows = e->g.ny;
    long nxc = std::max(1L, (target + rows - 1) / rows);
    nxc = std::min<long>(nxc, e->g.nx);
    return std::max(1, (int)((e->g.nx + nxc - 1) / nxc));
}

// workgroups the reverse launch of a batch aims at (x-chunks common to all engines)
constexpr long kBatchAdjGroups = 32768;

// Tile of the reverse step: the one sjf_set_adjoint_tile gave, else one row per workgroup (the
// form a batch runs by default, so an engine alone and in a batch give the same bits).  Measured
// on MI355X, one 192x192x96 FP32 engine: 0.139 ms per reverse step with row-of-threads tiles of
// 9 rows and 6 planes, 0.155 ms one row per workgroup; sixteen such engines in one batch: 1.78 ms
// per reverse launch with the 9-row tiles, 1.65 ms one row per workgroup; 256x256x128: 0.268 ms
// for the best row-of-threads tile against 0.266 ms one row per workgroup.
struct AdjTile {
    bool rows;
    int nrows, xc;  // xc 0: the default chunks
};
AdjTile adj_tile_of(const sjf_engine* e) {
    if (e->adj_set && e->adj_form == SJF_TILE_THREAD_ROWS)
        return {true, e->adj_rows, e->adj_xc > 0 ? std::min(e->adj_xc, e->g.nx) : 0};
    return {false, 1, e->adj_set && e->adj_xc > 0 ? std::min(e->adj_xc, e->g.nx) : 0};
}

template <typename T, int NP, bool NTS>
void launch_adj_one_row(sjf_engine* e, const sjf::AdjArgs<T>& g, int block, unsigned tiles, size_t lds) {
    if (block <= 256) hipLaunchKernelGGL((sjf::k_adj_step<T, NP, 256, NTS>), dim3(tiles), dim3(block), lds, e->stream, g);
    else if (block <= 512) hipLaunchKernelGGL((sjf::k_adj_step<T, NP, 512, NTS>), dim3(tiles), dim3(block), lds, e->stream, g);
    else hipLaunchKernelGGL((sjf::k_adj_step<T, NP, 1024, NTS>), dim3(tiles), dim3(block), lds, e->stream, g);
}

template <typename T, int NP>
void launch_adj_np(sjf_engine* e, const sjf::AdjArgs<T>& g, int block, unsigned tiles, size_t lds) {
    // the reverse step streams the adjoint fields as the forward step streams E and H
    if (field_bytes(e) > kStreamBytes) launch_adj_one_row<T, NP, true>(e, g, block, tiles, lds);
    else launch_adj_one_row<T, NP, false>(e, g, block, tiles, lds);
}

template <typename T, int NP>
void launch_adj_rows_np(sjf_engine* e, const sjf::AdjArgs<T>& g, int block, unsigned tiles, size_t lds) {
    if (block <= 256) hipLaunchKernelGGL((sjf::k_adj_rows<T, NP, 256>), dim3(tiles), dim3(block), lds, e->stream, g);
    else if (block <= 512) hipLaunchKernelGGL((sjf::k_adj_rows<T, NP, 512>), dim3(tiles), dim3(block), lds, e->stream, g);
    else hipLaunchKernelGGL((sjf::k_adj_rows<T, NP, 1024>), dim3(tiles), dim3(block), lds, e->stream, g);
}

template <typename T>
int launch_adj_step(sjf_engine* e, const sjf::AdjArgs<T>& g, int npv, int block, unsigned tiles, size_t lds,
                    bool rows) {
    switch (npv) {
        case 0: rows ? launch_adj_rows_np<T, 0>(e, g, block, tiles, lds) : launch_adj_np<T, 0>(e, g, block, tiles, lds); break;
        case 1: rows ? launch_adj_rows_np<T, 1>(e, g, block, tiles, lds) : launch_adj_np<T, 1>(e, g, block, tiles, lds); break;
        case 3: rows ? launch_adj_rows_np<T, 3>(e, g, block, tiles, lds) : launch_adj_np<T, 3>(e, g, block, tiles, lds); break;
        default: rows ? launch_adj_rows_np<T, 4>(e, g, block, tiles, lds) : launch_adj_np<T, 4>(e, g, block, tiles, lds); break;
    }
    HIPCHK(e, hipGetLastError());
    return SJF_OK;
}

// LDS of the reverse step: E-term and H-term exchange planes and the material table
template <typename T>
size_t adj_lds_bytes(const sjf_engine* e, bool rows, int block) {
    const size_t lut = sizeof(sjf::MatLut<T>) * e->nmat;
    return rows ? sizeof(T) * (8 * (size_t)block + 2 * (size_t)e->g.nz) + lut : sizeof(T) * 12 * (size_t)block + lut;
}

template <typename T>
int adjoint_run_t(sjf_engine* e, long nsteps, long K, long gap, double* J_out, double* grad, long nd) {
    LaunchSet* s = own_set(e);
    int rc = prepare_set(s, nsteps);
    if (rc) return rc;
    const StateLayout L = state_layout(e);
    const long slot = 3 * nd + 3;  // E of the design region and of the objective cell
    // Record mode (K == 0 and the record fits a quarter of the free device memory): the forward
    // pass records E of the design region and the objective cell after every step, so the
    // reverse pass needs no recomputation.  Otherwise the forward pass checkpoints the full state
    // every K steps (K == 0: about sqrt(nsteps)) and the reverse pass recomputes each segment.
    bool rec_all = false;
    if (K == 0) {
        size_t fr = 0, tot = 0;
        HIPCHK(e, hipMemGetInfo(&fr, &tot));
        rec_all = (double)sizeof(T) * slot * (nsteps + 1) <= 0.25 * (double)fr;
        if (!rec_all) K = std::max(1L, std::lround(std::sqrt((double)nsteps)));
    }
    K = std::min(K, nsteps);  // a longer interval holds one segment of nsteps all the same
    const long nck = rec_all ? 0 : (nsteps + K - 1) / K;
    const long nslots = rec_all ? nsteps + 1 : K + 1;
    char* cks = nullptr;
    if (!rec_all && hipMalloc((void**)&cks, L.total * (nck + 1)) != hipSuccess)
        return fail(e, SJF_E_NOMEM, "checkpoint storage allocation failed (" +
                                        std::to_string((double)L.total * (nck + 1) / 1e9) + " GB)");
    double* dJ = nullptr;
    double* gacc = nullptr;
    T* ring = nullptr;
    T* A = nullptr;
    T* Aq = nullptr;
    T* Ap = nullptr;
    T* Apsi[3] = {nullptr, nullptr, nullptr};
    auto cleanup = [&]() {
        dev_free(cks);
        dev_free(dJ);
        dev_free(gacc);
        dev_free(ring);
        dev_free(A);
        dev_free(Aq);
        dev_free(Ap);
        for (int d = 0; d < 3; ++d) dev_free(Apsi[d]);
    };
    auto bail = [&](int r) {
        cleanup();
        return r;
    };
    // adjoint fields: two sets of six components (read / written per reverse step)
    if (dev_alloc(e, (void**)&dJ, sizeof(double)) || dev_alloc(e, (void**)&gacc, sizeof(double) * nd) ||
        dev_alloc(e, (void**)&ring, sizeof(T) * slot * nslots) || dev_alloc(e, (void**)&A, sizeof(T) * 12 * e->N) ||
        (L.q && dev_alloc(e, (void**)&Aq, L.q)) || (L.p && dev_alloc(e, (void**)&Ap, L.p)))
        return bail(SJF_E_NOMEM);
    // auxiliary bars: per axis 8 slots (E pair, two H pairs, second E pair)
    for (int d = 0; d < 3; ++d) {
        // the reverse step addresses the eight slots of a slab within one 2 GiB buffer range
        if ((double)L.psi[d] / 6 * 8 >= 2147483648.0)
            return bail(fail(e, SJF_E_UNSUPPORTED, "adjoint PML auxiliaries exceed 2 GiB on one axis"));
        if (L.psi[d] && dev_alloc(e, (void**)&Apsi[d], L.psi[d] / 6 * 8)) return bail(SJF_E_NOMEM);
    }

    const double dt = e->dt;
    const unsigned gdes = (unsigned)((nd + 3 + 255) / 256);
    const int de0 = e->dhi[0] - e->dlo[0], de1 = e->dhi[1] - e->dlo[1], de2 = e->dhi[2] - e->dlo[2];
    const long nyz = (long)e->g.ny * e->g.nz;
    auto record = [&](long sl) -> int {
        const T* E = (const T*)e->Fbuf[e->step & 1];
        hipLaunchKernelGGL((sjf::k_adj_record<T>), dim3(gdes), dim3(256), 0, e->stream, E, nyz, e->g.ny, e->g.nz,
                           e->dlo[0], e->dlo[1], e->dlo[2], de0, de1, de2, gap, ring, sl);
        HIPCHK(e, hipGetLastError());
        return SJF_OK;
    };
    // ---- forward pass: the objective after every step; checkpoints every K steps or the record
    if (rec_all && (rc = record(0))) return bail(rc);
    for (long n = 0; n < nsteps; ++n) {
        if (!rec_all && n % K == 0) {
            if ((rc = copy_state(e, cks + L.total * (n / K), L, true))) return bail(rc);
        }
        if ((rc = launch_one_step(s, samples_at(s, e->step)))) return bail(rc);
        advance_hosts(s, 1);
        const T* E = (const T*)e->Fbuf[e->step & 1];
        hipLaunchKernelGGL((sjf::k_adj_objective<T>), dim3(1), dim3(64), 0, e->stream, E, nyz, gap, dt, dJ);
        HIPCHK(e, hipGetLastError());
        if (rec_all && (rc = record(n + 1))) return bail(rc);
    }
    if (!rec_all && (rc = copy_state(e, cks + L.total * nck, L, true))) return bail(rc);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if ((rc = d2h(e, J_out, dJ, sizeof(double)))) return bail(rc);

    // ---- backward pass
    sjf::AdjArgs<T> g;
    std::memset(&g, 0, sizeof(g));
    // register form: one y row per workgroup; row-of-threads form: arows owned rows
    const AdjTile at = adj_tile_of(e);
    const bool rows = at.rows;
    const int arows = at.nrows;
    int axc = at.xc > 0 ? at.xc : adj_xchunk(e);
    if (rows && at.xc == 0) {  // about 16384 workgroups over the tile rows
        const long trows = (e->g.ny + arows - 1) / arows;
        const long nxc = std::min<long>(std::max(1L, (16384 + trows - 1) / trows), e->g.nx);
        axc = (int)((e->g.nx + nxc - 1) / nxc);
    }
    g.f = step_args<T>(e, arows, axc);
    g.A[0] = A;
    g.A[1] = A + (size_t)6 * e->N;
    g.rd = 0;
    g.Aq = Aq;
    g.Ap = Ap;
    for (int d = 0; d < 3; ++d) g.Apsi[d] = Apsi[d];
    const int block = (((rows ? arows + 1 : 1) * e->g.nz + 63) / 64) * 64;
    const unsigned tiles = (unsigned)((long)g.f.nty * ((e->g.nx + g.f.xc - 1) / g.f.xc));
    const size_t lds = adj_lds_bytes<T>(e, rows, block);
    if (lds > 65536) return bail(fail(e, SJF_E_UNSUPPORTED, "adjoint tile needs more than 64 KiB of LDS"));
    const int npv = np_variant_of(e->maxp, e->uniform);
    auto ebar = [&]() { return g.A[g.rd]; };
    const double ie[3] = {inv_eps_at(e, gap, 0), inv_eps_at(e, gap, 1), inv_eps_at(e, gap, 2)};
    auto inject = [&](long sl) -> int {
        hipLaunchKernelGGL((sjf::k_adj_inject<T>), dim3(1), dim3(64), 0, e->stream, ebar(), nyz, ring, nd, sl, gap, dt,
                           ie[0], ie[1], ie[2]);
        HIPCHK(e, hipGetLastError());
        return SJF_OK;
    };
    // HIP events around the fused reverse kernel of the first kTimed reverse steps (its mean
    // duration is reported by sjf_adjoint_kernel_info)
    constexpr int kTimed = 64;
    std::vector<hipEvent_t> evs;
    long ntimed = 0;
    auto free_events = [&]() {
        for (auto& ev : evs) (void)hipEventDestroy(ev);
        evs.clear();
    };
    for (int q = 0; q < 2 * kTimed; ++q) {
        hipEvent_t ev = nullptr;
        if (hipEventCreate(&ev) != hipSuccess) {
            free_events();
            return bail(fail(e, SJF_E_HIP, "hipEventCreate failed"));
        }
        evs.push_back(ev);
    }
    // reverse steps n1 - 1 .. n0 with the record slots of steps n0 .. n1 at slot (n - base)
    auto reverse = [&](long n0, long n1, long base) -> int {
        for (long n = n1 - 1; n >= n0; --n) {
            hipLaunchKernelGGL((sjf::k_adj_grad<T>), dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, e->stream,
                               g.f, ebar(), ring, n + 1 - base, n - base, gacc);
            HIPCHK(e, hipGetLastError());
            g.hr = (n & 1) ? 2 : 4;
            g.hw = (n & 1) ? 4 : 2;
            g.er = (n & 1) ? 0 : 6;
            g.ew = (n & 1) ? 6 : 0;
            const bool timed = ntimed < kTimed;
            if (timed) HIPCHK(e, hipEventRecord(evs[2 * ntimed], e->stream));
            int r = launch_adj_step<T>(e, g, npv, block, tiles, lds, rows);
            if (r) return r;
            if (timed) HIPCHK(e, hipEventRecord(evs[2 * ntimed++ + 1], e->stream));
            g.rd = 1 - g.rd;
            if (n >= 1 && (r = inject(n - base))) return r;
        }
        return SJF_OK;
    };
    auto bail_ev = [&](int r) {
        free_events();
        return bail(r);
    };
    if (rec_all) {
        if ((rc = inject(nsteps))) return bail_ev(rc);
        if ((rc = reverse(0, nsteps, 0))) return bail_ev(rc);
    } else {
        for (long k = nck - 1; k >= 0; --k) {
            const long n0 = k * K, n1 = std::min(n0 + K, nsteps);
            if ((rc = set_step(e, n0))) return bail_ev(rc);
            if ((rc = copy_state(e, cks + L.total * k, L, false))) return bail_ev(rc);
            if ((rc = record(0))) return bail_ev(rc);
            for (long n = n0; n < n1; ++n) {
                if ((rc = launch_bare_step(s))) return bail_ev(rc);
                e->step++;
                if ((rc = record(n + 1 - n0))) return bail_ev(rc);
            }
            if (k == nck - 1 && (rc = inject(n1 - n0))) return bail_ev(rc);
            if ((rc = reverse(n0, n1, n0))) return bail_ev(rc);
        }
    }
    if (hipStreamSynchronize(e->stream) != hipSuccess) return bail_ev(fail(e, SJF_E_HIP, "hipStreamSynchronize failed"));
    double msum = 0;
    for (long q = 0; q < ntimed; ++q) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, evs[2 * q], evs[2 * q + 1]) != hipSuccess)
            return bail_ev(fail(e, SJF_E_HIP, "hipEventElapsedTime failed"));
        msum += ms;
    }
    free_events();
    e->adj_ms = ntimed ? msum / ntimed : 0;
    e->adj_threads = block;
    e->adj_groups = (int)tiles;
    std::vector<double> gh(nd);
    if ((rc = d2h(e, gh.data(), gacc, sizeof(double) * nd))) return bail(rc);
    // the state is Ubar = Ebar / eps, so the accumulated sum already carries the 1/eps_v factor
    for (long v = 0; v < nd; ++v) grad[v] = -gh[v];
    // leave the engine in its final forward state (record mode never left it)
    if (!rec_all) {
        if ((rc = set_step(e, nsteps))) return bail(rc);
        if ((rc = copy_state(e, cks + L.total * nck, L, false))) return bail(rc);
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    cleanup();
    return SJF_OK;
}

template <typename T, int NP, bool NTS>
void launch_adj_batch_step(LaunchSet* s, const sjf::AdjArgs<T>* args, const int* first, int n, int block,
                           unsigned tiles, size_t lds) {
    if (block <= 256)
        hipLaunchKernelGGL((sjf::k_adj_step_batch<T, NP, 256, NTS>), dim3(tiles), dim3(block), lds, s->stream, args, first, n);
    else if (block <= 512)
        hipLaunchKernelGGL((sjf::k_adj_step_batch<T, NP, 512, NTS>), dim3(tiles), dim3(block), lds, s->stream, args, first, n);
    else
        hipLaunchKernelGGL((sjf::k_adj_step_batch<T, NP, 1024, NTS>), dim3(tiles), dim3(block), lds, s->stream, args,
                           first, n);
}

template <typename T, int NP>
void launch_adj_batch_np(LaunchSet* s, const sjf::AdjArgs<T>* args, const int* first, int n, int block,
                         unsigned tiles, size_t lds, bool rows) {
    if (rows) {
        if (block <= 256)
            hipLaunchKernelGGL((sjf::k_adj_rows_batch<T, NP, 256>), dim3(tiles), dim3(block), lds, s->stream, args, first, n);
        else if (block <= 512)
            hipLaunchKernelGGL((sjf::k_adj_rows_batch<T, NP, 512>), dim3(tiles), dim3(block), lds, s->stream, args, first, n);
        else
            hipLaunchKernelGGL((sjf::k_adj_rows_batch<T, NP, 1024>), dim3(tiles), dim3(block), lds, s->stream, args,
                               first, n);
        return;
    }
    if (s->nts) launch_adj_batch_step<T, NP, true>(s, args, first, n, block, tiles, lds);
    else launch_adj_batch_step<T, NP, false>(s, args, first, n, block, tiles, lds);
}

// Adjoint gradients of every engine of a launch set in one pass: the forward steps are the set's
// own fused launches; after each of them one launch adds the objectives of all engines (and, in
// record mode, records their design fields); each reverse step is one gradient launch, one fused
// reverse launch over all engines' tiles and one injection launch.  With checkpoints (K > 0, or
// when the record does not fit a quarter of the free memory) every engine's state is saved every
// K steps and each segment is recomputed with bare batch steps, as adjoint_run_t does for one
// engine.  The arithmetic of every cell is that of adjoint_run_t, so each engine's J and gradient
// are those of its own run.  gaps[i]: objective cell index of engine i; nds[i]: its design voxels.
template <typename T>
int batch_adjoint_run_t(LaunchSet* s, long nsteps, long K, const long* gaps, const long* nds, double* J_out,
                        double* const* grads, double* ms_out, int* threads_out, int* groups_out) {
    sjf_engine* e0 = s->eng[0];
    const int n = (int)s->eng.size();
    int rc = prepare_set(s, nsteps);
    if (rc) return rc;
    bool rec_all = K == 0;
    if (rec_all) {
        double rec_bytes = 0;
        for (int i = 0; i < n; ++i) rec_bytes += (double)sizeof(T) * (3 * nds[i] + 3) * (nsteps + 1);
        size_t fr = 0, tot = 0;
        HIPCHK(e0, hipMemGetInfo(&fr, &tot));
        rec_all = rec_bytes <= 0.25 * (double)fr;
        if (!rec_all) K = std::max(1L, std::lround(std::sqrt((double)nsteps)));
    }
    K = std::min(K, nsteps);  // a longer interval holds one segment of nsteps all the same
    const long nck = rec_all ? 0 : (nsteps + K - 1) / K;
    // record slots per engine; with checkpoints slot K + 1 takes the forward pass's writes, which
    // nothing reads (that pass only needs the objectives)
    const long nslots = rec_all ? nsteps + 1 : K + 2;
    std::vector<StateLayout> Ls(n);
    std::vector<char*> cks(n, nullptr);
    std::vector<void*> mem;
    auto cleanup = [&]() {
        for (void*& p : mem) dev_free(p);
        mem.clear();
    };
    auto bail = [&](int r) {
        cleanup();
        return r;
    };
    auto alloc = [&](void** p, size_t bytes) -> int {
        *p = nullptr;
        const int r = dev_alloc(e0, p, bytes);
        if (!r) mem.push_back(*p);
        return r;
    };
    std::vector<sjf::AdjBatchDev<T>> tab(n);
    std::vector<sjf::AdjArgs<T>> args(n);
    std::vector<int> first;
    int t = 0, nmat = 1;
    long maxv = 0, rows = 0;
    // one row per workgroup with x-chunks common to all engines, or, when the first engine's tile
    // was set to the row-of-threads form, its rows (at most each engine's ny) and its x-chunk
    const AdjTile bt = adj_tile_of(e0);
    const bool brows = bt.rows;
    const int bxc = bt.xc > 0 ? bt.xc : 6;
    int bblock = 64;
    for (const sjf_engine* e : s->eng) rows += e->g.ny;
    for (int i = 0; i < n; ++i) {
        sjf_engine* e = s->eng[i];
        const StateLayout L = state_layout(e);
        const long nd = nds[i];
        void *A = nullptr, *Aq = nullptr, *Ap = nullptr, *ring = nullptr, *gacc = nullptr, *dJ = nullptr;
        void* Apsi[3] = {nullptr, nullptr, nullptr};
        if (alloc(&A, sizeof(T) * 12 * e->N) || (L.q && alloc(&Aq, L.q)) || (L.p && alloc(&Ap, L.p)) ||
            alloc(&ring, sizeof(T) * (3 * nd + 3) * nslots) || alloc(&gacc, sizeof(double) * nd) ||
            alloc(&dJ, sizeof(double)))
            return bail(SJF_E_NOMEM);
        Ls[i] = L;
        if (!rec_all && alloc((void**)&cks[i], L.total * (nck + 1)))
            return bail(fail(e0, SJF_E_NOMEM, "checkpoint storage allocation failed (" +
                                                  std::to_string((double)L.total * (nck + 1) / 1e9) + " GB)"));
        for (int d = 0; d < 3; ++d) {
            if ((double)L.psi[d] / 6 * 8 >= 2147483648.0)
                return bail(fail(e0, SJF_E_UNSUPPORTED, "adjoint PML auxiliaries exceed 2 GiB on one axis"));
            if (L.psi[d] && alloc(&Apsi[d], L.psi[d] / 6 * 8)) return bail(SJF_E_NOMEM);
        }
        sjf::AdjBatchDev<T>& d = tab[i];
        std::memset(&d, 0, sizeof(d));
        d.F[0] = (const T*)e->Fbuf[0];
        d.F[1] = (const T*)e->Fbuf[1];
        d.A[0] = (T*)A;
        d.A[1] = (T*)A + (size_t)6 * e->N;
        d.ring = (T*)ring;
        d.gacc = (double*)gacc;
        d.J = (double*)dJ;
        d.nyz = (long)e->g.ny * e->g.nz;
        d.ny = e->g.ny;
        d.nz = e->g.nz;
        for (int q = 0; q < 3; ++q) {
            d.dlo[q] = e->dlo[q];
            d.de[q] = e->dhi[q] - e->dlo[q];
            d.ie[q] = inv_eps_at(e, gaps[i], q);
        }
        d.nd = nd;
        d.gap = gaps[i];
        d.dt = e->dt;
        sjf::AdjArgs<T>& g = args[i];
        std::memset(&g, 0, sizeof(g));
        // the first engine's tile form for all (they share nz and the precision)
        if (brows) {
            const int r = std::min(bt.nrows, e->g.ny);
            g.f = step_args<T>(e, r, std::min(bxc, e->g.nx));
            bblock = std::max(bblock, (((r + 1) * e->g.nz + 63) / 64) * 64);
        } else {
            g.f = step_args<T>(e, 1, adj_xchunk(e, kBatchAdjGroups, rows));  // one y row per workgroup
            g.f.nty = e->g.ny;
        }
        g.A[0] = d.A[0];
        g.A[1] = d.A[1];
        g.Aq = (T*)Aq;
        g.Ap = (T*)Ap;
        for (int q = 0; q < 3; ++q) g.Apsi[q] = (T*)Apsi[q];
        first.push_back(t);
        t += (int)((long)g.f.nty * ((e->g.nx + g.f.xc - 1) / g.f.xc));
        nmat = std::max(nmat, e->nmat);
        maxv = std::max(maxv, nd + 3);
    }
    first.push_back(t);
    // the reverse step's buffer roles alternate with the parity p of the step index n: read set
    // rd = (nsteps - 1 - n) & 1, H auxiliary pair read 2 / 4 and E pair read 0 / 6
    std::vector<sjf::AdjArgs<T>> argp[2] = {args, args};
    for (int p = 0; p < 2; ++p)
        for (sjf::AdjArgs<T>& g : argp[p]) {
            g.rd = (int)((nsteps - 1 - p) & 1);
            g.hr = p ? 2 : 4;
            g.hw = p ? 4 : 2;
            g.er = p ? 0 : 6;
            g.ew = p ? 6 : 0;
        }
    void *d_tab = nullptr, *d_args0 = nullptr, *d_args1 = nullptr, *d_first = nullptr;
    auto upload = [&](void** dst, auto const& v) -> int {
        *dst = nullptr;
        const int r = upload_vec(e0, s->stream, dst, v);
        if (*dst) mem.push_back(*dst);  // freed with the rest, also when the copy failed
        return r;
    };
    if ((rc = upload(&d_tab, tab)) || (rc = upload(&d_args0, argp[0])) || (rc = upload(&d_args1, argp[1])) ||
        (rc = upload(&d_first, first)))
        return bail(rc);
    const auto* tabp = (const sjf::AdjBatchDev<T>*)d_tab;
    const sjf::AdjArgs<T>* argd[2] = {(const sjf::AdjArgs<T>*)d_args0, (const sjf::AdjArgs<T>*)d_args1};
    const int block = brows ? bblock : ((e0->g.nz + 63) / 64) * 64;
    const size_t lds = (brows ? sizeof(T) * (8 * (size_t)block + 2 * (size_t)e0->g.nz) : sizeof(T) * 12 * (size_t)block) +
                       sizeof(sjf::MatLut<T>) * nmat;
    if (lds > 65536) return bail(fail(e0, SJF_E_UNSUPPORTED, "adjoint tile needs more than 64 KiB of LDS"));
    const int npv = np_variant_of(e0->maxp, e0->uniform);
    const dim3 grid_v((unsigned)((maxv + 255) / 256), (unsigned)n);
    auto record = [&](long sl, int objective) -> int {
        hipLaunchKernelGGL((sjf::k_adj_record_batch<T>), grid_v, dim3(256), 0, s->stream, tabp,
                           (int)(e0->step & 1), sl, objective);
        SETCHK(s, hipGetLastError());
        return SJF_OK;
    };
    auto save_all = [&](long k, bool save) -> int {
        for (int i = 0; i < n; ++i) {
            const int r = copy_state(s->eng[i], cks[i] + Ls[i].total * k, Ls[i], save, s->stream);
            if (r) return r;
        }
        return SJF_OK;
    };
    // ---- forward pass: the objectives after every step; the record or checkpoints every K steps
    if (rec_all && (rc = record(0, 0))) return bail(rc);
    for (long q = 0; q < nsteps; ++q) {
        if (!rec_all && q % K == 0 && (rc = save_all(q / K, true))) return bail(rc);
        if ((rc = launch_one_step(s, samples_at(s, e0->step)))) return bail(rc);
        advance_hosts(s, 1);
        if ((rc = record(rec_all ? q + 1 : K + 1, 1))) return bail(rc);
    }
    if (!rec_all && (rc = save_all(nck, true))) return bail(rc);
    // ---- reverse pass
    constexpr int kTimed = 64;
    std::vector<hipEvent_t> evs;
    auto free_events = [&]() {
        for (auto& ev : evs) (void)hipEventDestroy(ev);
        evs.clear();
    };
    auto bail_ev = [&](int r) {
        free_events();
        return bail(r);
    };
    for (int q = 0; q < 2 * kTimed; ++q) {
        hipEvent_t ev = nullptr;
        if (hipEventCreate(&ev) != hipSuccess) return bail_ev(fail(e0, SJF_E_HIP, "hipEventCreate failed"));
        evs.push_back(ev);
    }
    int rd = 0;
    auto inject = [&](long sl) -> int {
        hipLaunchKernelGGL((sjf::k_adj_inject_batch<T>), dim3(1, (unsigned)n), dim3(64), 0, s->stream, tabp, rd, sl);
        SETCHK(s, hipGetLastError());
        return SJF_OK;
    };
    const dim3 grid_g((unsigned)((maxv - 3 + 255) / 256), (unsigned)n);
    long ntimed = 0;
    // one reverse step: gradient term, fused reverse launch (timed for the first kTimed steps)
    auto reverse_step = [&](long q, long base) -> int {
        hipLaunchKernelGGL((sjf::k_adj_grad_batch<T>), grid_g, dim3(256), 0, s->stream, tabp, rd, q + 1 - base,
                           q - base);
        SETCHK(s, hipGetLastError());
        const bool timed = ntimed < kTimed;
        if (timed) SETCHK(s, hipEventRecord(evs[2 * ntimed], s->stream));
        const sjf::AdjArgs<T>* ad = argd[q & 1];
        switch (npv) {
            case 0: launch_adj_batch_np<T, 0>(s, ad, (const int*)d_first, n, block, (unsigned)t, lds, brows); break;
            case 1: launch_adj_batch_np<T, 1>(s, ad, (const int*)d_first, n, block, (unsigned)t, lds, brows); break;
            case 3: launch_adj_batch_np<T, 3>(s, ad, (const int*)d_first, n, block, (unsigned)t, lds, brows); break;
            default: launch_adj_batch_np<T, 4>(s, ad, (const int*)d_first, n, block, (unsigned)t, lds, brows); break;
        }
        SETCHK(s, hipGetLastError());
        if (timed) SETCHK(s, hipEventRecord(evs[2 * ntimed++ + 1], s->stream));
        return SJF_OK;
    };
    // reverse steps q1 - 1 .. q0 with the record slots of steps q0 .. q1 at slot (q - base)
    auto reverse = [&](long q0, long q1, long base) -> int {
        for (long q = q1 - 1; q >= q0; --q) {
            int r = reverse_step(q, base);
            if (r) return r;
            rd = 1 - rd;
            if (q >= 1 && (r = inject(q - base))) return r;
        }
        return SJF_OK;
    };
    auto set_steps = [&](long step) -> int {
        if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(e0, SJF_E_HIP, "hipStreamSynchronize failed");
        for (sjf_engine* e : s->eng) {
            const int r = set_step(e, step);
            if (r) return r;
        }
        return SJF_OK;
    };
    if (rec_all) {
        if ((rc = inject(nsteps)) || (rc = reverse(0, nsteps, 0))) return bail_ev(rc);
    } else {
        for (long k = nck - 1; k >= 0; --k) {
            const long q0 = k * K, q1 = std::min(q0 + K, nsteps);
            if ((rc = set_steps(q0)) || (rc = save_all(k, false)) || (rc = record(0, 0))) return bail_ev(rc);
            for (long q = q0; q < q1; ++q) {
                if ((rc = launch_bare_step(s))) return bail_ev(rc);
                for (sjf_engine* e : s->eng) e->step++;
                if ((rc = record(q + 1 - q0, 0))) return bail_ev(rc);
            }
            if (k == nck - 1 && (rc = inject(q1 - q0))) return bail_ev(rc);
            if ((rc = reverse(q0, q1, q0))) return bail_ev(rc);
        }
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess)
        return bail_ev(fail(e0, SJF_E_HIP, "hipStreamSynchronize failed"));
    double msum = 0;
    for (long q = 0; q < ntimed; ++q) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, evs[2 * q], evs[2 * q + 1]) != hipSuccess)
            return bail_ev(fail(e0, SJF_E_HIP, "hipEventElapsedTime failed"));
        msum += ms;
    }
    free_events();
    *ms_out = ntimed ? msum / ntimed : 0;
    *threads_out = block;
    *groups_out = t;
    for (int i = 0; i < n; ++i) {
        std::vector<double> gh(nds[i]);
        if ((rc = d2h(e0, gh.data(), tab[i].gacc, sizeof(double) * nds[i])) ||
            (rc = d2h(e0, &J_out[i], tab[i].J, sizeof(double))))
            return bail(rc);
        // the state is Ubar = Ebar / eps, so the sums already carry the 1/eps_v factor
        for (long v = 0; v < nds[i]; ++v) grads[i][v] = -gh[v];
    }
    // leave every engine in its final forward state (record mode never left it)
    if (!rec_all) {
        if ((rc = set_steps(nsteps)) || (rc = save_all(nck, false))) return bail(rc);
        if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(e0, SJF_E_HIP, "hipStreamSynchronize failed"));
    }
    cleanup();
    return SJF_OK;
}

}  // namespace

extern "C" int sjf_adjoint_run(sjf_engine* e, long nsteps, int checkpoint_every, const int gap[3], double* J,
                               double* grad, size_t cap) {
    if (!e || nsteps < 1 || checkpoint_every < 0 || !gap || !J || !grad) return SJF_E_ARG;
    if (!e->emap_h.empty()) return fail(e, SJF_E_UNSUPPORTED, "adjoint runs take the LUT's eps_inf (no eps maps)");
    if (e->g.nz > 1024) return fail(e, SJF_E_UNSUPPORTED, "the reverse step spans the z line: nz <= 1024");
    if (e->batch) return fail(e, SJF_E_STATE, "engine is stepped by its batch");
    if (e->step != 0) return fail(e, SJF_E_STATE, "the adjoint run starts from step 0");
    if (e->deps_h.empty()) return fail(e, SJF_E_STATE, "no design region (sjf_set_design_eps)");
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    for (int d = 0; d < 3; ++d)
        if (gap[d] < 0 || gap[d] >= n[d]) return fail(e, SJF_E_ARG, "objective cell outside the grid");
    const long nd = (long)e->deps_h.size();
    if (cap < (size_t)nd) return fail(e, SJF_E_ARG, "gradient buffer smaller than the design region");
    HIPCHK(e, hipSetDevice(e->g.device));
    const long q = cell_index(e, gap[0], gap[1], gap[2]);
    return e->prec == 4 ? adjoint_run_t<float>(e, nsteps, checkpoint_every, q, J, grad, nd)
                        : adjoint_run_t<double>(e, nsteps, checkpoint_every, q, J, grad, nd);
}

extern "C" int sjf_set_adjoint_tile(sjf_engine* e, int form, int rows, int xchunk) {
    if (!e || xchunk < 0) return SJF_E_ARG;
    if (form == SJF_TILE_REGISTER_ROWS) {
        if (rows != 1) return fail(e, SJF_E_ARG, "the register form of the reverse step owns one row per workgroup");
    } else if (form == SJF_TILE_THREAD_ROWS) {
        if (!thread_rows_fit(e, rows))
            return fail(e, SJF_E_ARG, "row-of-threads tile needs (rows + 1) * nz <= 1024");
        const int block = (((rows + 1) * e->g.nz + 63) / 64) * 64;
        // FP64 tiles stop at 512 threads: at 1024 the register budget of 128 per thread spills
        if (e->prec == 8 && block > 512)
            return fail(e, SJF_E_ARG, "FP64 row-of-threads reverse tile needs (rows + 1) * nz <= 512");
        const size_t lds = e->prec == 4 ? adj_lds_bytes<float>(e, true, block) : adj_lds_bytes<double>(e, true, block);
        if (lds > 65536) return fail(e, SJF_E_ARG, "adjoint row-of-threads tile needs more than 64 KiB of LDS");
    } else {
        return fail(e, SJF_E_ARG, "unknown tile form");
    }
    e->adj_form = form;
    e->adj_rows = rows;
    e->adj_xc = xchunk;
    e->adj_set = true;
    return SJF_OK;
}

extern "C" int sjf_adjoint_kernel_info(const sjf_engine* e, double* ms_mean, int* threads, int* groups) {
    if (!e || !ms_mean) return SJF_E_ARG;
    if (e->adj_groups == 0) return SJF_E_STATE;
    *ms_mean = e->adj_ms;
    if (threads) *threads = e->adj_threads;
    if (groups) *groups = e->adj_groups;
    return SJF_OK;
}

// ============================================================================ batches ====

namespace {

// ======================================================================== flux read-out ===
// The flux of a DFT plane is reduced on the device in two launches (k_dft_flux_part,
// k_dft_flux_sum) through a buffer of nf * (nb + 1) doubles: nb partial sums per frequency, then
// the nf results that are copied out.  sjf_dft_flux allocates that buffer per call;
// sjf_dft_flux_noalloc uses `scratch`, a static device array of at most 64 frequencies x 65
// doubles (33 KiB) that every device the library runs on holds until the process ends.  Every
// slot a call reads is written earlier in the same call, so nothing needs clearing; calls that
// overlap on one device share the array and race.
constexpr int kFluxMaxBlocks = 64;
constexpr int kFluxScratchDoubles = SJF_MAX_DFT_FREQS * (kFluxMaxBlocks + 1);
static __device__ __align__(256) double scratch[kFluxScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // per device: the address of `scratch` there

// Resolves this device's copy of `scratch` (the device must be current).  The array's name in
// host code is only a stand-in; kernels and copies take the address resolved here.
int scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return SJF_OK;
    void* p = nullptr;
    const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (st != hipSuccess || !p) {
        g_create_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st);
        return SJF_E_HIP;
    }
    g_scratch_ptr[device] = (double*)p;
    return SJF_OK;
}

double* scratch_on(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

struct FluxPlan {
    const Dft* d;
    int slot[4];  // E_a, E_b, H_a, H_b
    int nf, nb;
    long scratch_doubles() const { return (long)nf * (nb + 1); }
};

int flux_plan(sjf_engine* e, int id, int axis, const double* flux, size_t cap, FluxPlan* fp) {
    if (!e || !flux || id < 0 || id >= (int)e->dft.size() || axis < 0 || axis > 2) return SJF_E_ARG;
    const Dft& d = e->dft[id];
    if (d.hi[axis] - d.lo[axis] != 1) return fail(e, SJF_E_ARG, "a flux plane is one cell thick along its normal");
    const int a = (axis + 1) % 3, b = (axis + 2) % 3;
    int slot[6] = {-1, -1, -1, -1, -1, -1};
    for (int q = 0; q < d.ncomp; ++q) slot[d.comps[q]] = q;
    if (slot[a] < 0 || slot[b] < 0 || slot[3 + a] < 0 || slot[3 + b] < 0)
        return fail(e, SJF_E_ARG, "a flux plane needs both tangential E and both tangential H components");
    fp->d = &d;
    fp->slot[0] = slot[a];
    fp->slot[1] = slot[b];
    fp->slot[2] = slot[3 + a];
    fp->slot[3] = slot[3 + b];
    fp->nf = (int)d.freqs.size();
    fp->nb = (int)std::min<long>(kFluxMaxBlocks, std::max<long>(1, (d.ncell + 255) / 256));
    if (cap < (size_t)fp->nf) return fail(e, SJF_E_ARG, "output buffer too small");
    return SJF_OK;
}

// buf: nf * nb partial sums, then the nf results
int flux_run(sjf_engine* e, const FluxPlan& fp, double* buf, double* flux) {
    const Dft& d = *fp.d;
    hipLaunchKernelGGL(sjf::k_dft_flux_part, dim3(fp.nb, fp.nf), dim3(256), 0, e->stream, (const double2*)d.acc,
                       (const double2*)d.cmp, d.ncell, fp.nf, fp.slot[0], fp.slot[1], fp.slot[2], fp.slot[3], fp.nb,
                       buf);
    if (hipGetLastError() != hipSuccess) return fail(e, SJF_E_HIP, "k_dft_flux_part launch failed");
    double* out = buf + (long)fp.nf * fp.nb;
    hipLaunchKernelGGL(sjf::k_dft_flux_sum, dim3(1), dim3(64), 0, e->stream, (const double*)buf, fp.nf, fp.nb,
                       e->g.dx * e->g.dx, out);
    if (hipGetLastError() != hipSuccess) return fail(e, SJF_E_HIP, "k_dft_flux_sum launch failed");
    return d2h(e, flux, out, sizeof(double) * fp.nf);
}

}  // namespace

struct sjf_batch {
    LaunchSet* set = nullptr;
    std::string err;
    double adj_ms = 0;  // fused reverse launch of the last batched adjoint run
    int adj_threads = 0, adj_groups = 0;
};

namespace {
int sync_engine(sjf_engine* e) {
    HIPCHK(e, hipSetDevice(e->g.device));
    if (e->batch) HIPCHK(e, hipStreamSynchronize(e->batch->set->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SJF_OK;
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
    if (g->nz > 1024 && g->periodic[2]) {
        g_create_error = "a z line longer than 1024 cells is stepped in passes, which needs a non-periodic z axis";
        return SJF_E_UNSUPPORTED;
    }
    if (g->precision != SJF_FP32 && g->precision != SJF_FP64) {
        g_create_error = "precision must be SJF_FP32 (4) or SJF_FP64 (8)";
        return SJF_E_ARG;
    }
    if ((double)g->nx * g->ny * g->nz * 6 >= 4294967296.0) {
        g_create_error = "grid too large for one engine (6 * cells must stay below 2^32)";
        return SJF_E_UNSUPPORTED;
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
    if (int rc = scratch_setup(g->device)) return rc;
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
    auto bail = [&](int rc) {
        g_create_error = e->err;
        free_all(e);
        (void)hipStreamDestroy(e->stream);
        delete e;
        return rc;
    };
    for (int b = 0; b < 2; ++b) {
        int rc = dev_alloc(e, &e->Fbuf[b], (size_t)e->prec * 6 * e->N);
        if (rc) return bail(rc);
    }
    if (int rc = dev_alloc(e, (void**)&e->d_step, sizeof(long))) return bail(rc);
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
        if (rc) return bail(rc);
    }
    // default material: vacuum
    e->lut_h.assign(1, sjf_material{});
    e->lut_h[0].eps_inf = 1.0;
    e->nmat = 1;
    if (int rc = upload_lut(e)) return bail(rc);
    choose_tile(e);
    *out = e;
    return SJF_OK;
}

int sjf_destroy(sjf_engine* e) {
    if (!e) return SJF_E_ARG;
    if (e->batch) return fail(e, SJF_E_STATE, "destroy the batch before its engines");
    (void)hipSetDevice(e->g.device);
    (void)hipStreamSynchronize(e->stream);
    free_all(e);
    (void)hipStreamDestroy(e->stream);
    delete e;
    return SJF_OK;
}

const char* sjf_last_error(const sjf_engine* e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int sjf_set_materials(sjf_engine* e, const uint32_t* mat_word, const sjf_material* lut, int n_mat) {
    if (!e || !mat_word || !lut) return SJF_E_ARG;
    if (n_mat < 1 || n_mat > SJF_MAX_MATERIALS) return fail(e, SJF_E_ARG, "n_mat must be in [1, 255]");
    if (int rc = sync_engine(e)) return rc;
    for (int m = 0; m < n_mat; ++m) {
        if (lut[m].npoles < 0 || lut[m].npoles > SJF_MAX_POLES)
            return fail(e, SJF_E_ARG, "npoles must be in [0, 4]");
        if (!(lut[m].eps_inf > 0)) return fail(e, SJF_E_ARG, "eps_inf must be positive");
    }
    for (long q = 0; q < e->N; ++q)
        for (int c = 0; c < 3; ++c) {
            const int m = (mat_word[q] >> (8 * c)) & 0xff;
            if (m >= n_mat && m != sjf::kDesignMat)
                return fail(e, SJF_E_ARG, "material word references an id >= n_mat");
            if (m == sjf::kDesignMat && e->deps_h.empty())
                return fail(e, SJF_E_STATE, "design id 255 used before sjf_set_design_eps");
        }
    e->given_lut.assign(lut, lut + n_mat);
    e->given_w.assign(mat_word, mat_word + e->N);
    e->have_mat = true;
    return apply_materials(e);
}

int sjf_set_eps_maps(sjf_engine* e, const double* eps) {
    if (!e) return SJF_E_ARG;
    if (int rc = sync_engine(e)) return rc;
    if (!eps) {
        e->emap_h.clear();
        e->emap_eff.clear();
        dev_free(e->emap);
    } else {
        for (long q = 0; q < 3 * e->N; ++q)
            if (!(eps[q] > 0) || !std::isfinite(eps[q])) return fail(e, SJF_E_ARG, "map eps must be positive and finite");
        e->emap_h.assign(eps, eps + 3 * e->N);
        if (!e->have_mat) {
            e->given_w.assign(e->N, 0u);
            e->given_lut = e->lut_h;
            e->have_mat = true;
        }
    }
    return apply_materials(e);
}

int sjf_set_pml_dispersion(sjf_engine* e, double f_ref) {
    if (!e) return SJF_E_ARG;
    if (!(f_ref >= 0) || !std::isfinite(f_ref)) return fail(e, SJF_E_ARG, "f_ref must be >= 0");
    if (e->step != 0) return fail(e, SJF_E_STATE, "the PML media must be set before the first step");
    if (int rc = sync_engine(e)) return rc;
    e->pml_fref = f_ref;
    return e->have_mat ? apply_materials(e) : SJF_OK;
}

int sjf_set_design_eps(sjf_engine* e, const int lo[3], const int hi[3], const double* eps) {
    if (!e || !lo || !hi || !eps) return SJF_E_ARG;
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    long cnt = 1;
    for (int d = 0; d < 3; ++d) {
        if (lo[d] < 0 || hi[d] > n[d] || hi[d] <= lo[d]) return fail(e, SJF_E_ARG, "design box outside grid");
        cnt *= hi[d] - lo[d];
    }
    if (!e->deps_h.empty())
        for (int d = 0; d < 3; ++d)
            if (lo[d] != e->dlo[d] || hi[d] != e->dhi[d])
                return fail(e, SJF_E_STATE, "design box cannot change once set");
    for (long q = 0; q < cnt; ++q)
        if (!(eps[q] > 0)) return fail(e, SJF_E_ARG, "design eps must be positive");
    if (int rc = sync_engine(e)) return rc;
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
        e->given_w.assign(e->N, 0u);
        e->given_lut = e->lut_h;
        e->have_mat = true;
        rc = dev_alloc(e, (void**)&e->matw, sizeof(uint32_t) * e->N);
        if (rc) return rc;
    }
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k) {
                const long q = cell_index(e, i, j, k);
                e->matw_h[q] = e->given_w[q] = 0x00ffffffu;
            }
    rc = h2d(e, e->matw, e->matw_h.data(), sizeof(uint32_t) * e->N);
    if (rc) return rc;
    e->uniform = false;
    e->version++;
    return SJF_OK;
}

int sjf_set_pml(sjf_engine* e, const int cells[6], double R, double power) {
    if (!e || !cells) return SJF_E_ARG;
    if (!(R > 0 && R < 1) || !(power >= 0)) return fail(e, SJF_E_ARG, "R must be in (0,1), power >= 0");
    if (e->step != 0) return fail(e, SJF_E_STATE, "PML must be set before the first step");
    for (int d = 0; d < 3; ++d) {
        const int n = axis_n(e, d);
        const int lo = cells[2 * d], hi = cells[2 * d + 1];
        if (lo < 0 || hi < 0 || lo + hi > n) return fail(e, SJF_E_ARG, "PML thicker than the grid");
        if ((lo > 0 || hi > 0) && e->g.periodic[d]) return fail(e, SJF_E_ARG, "PML on a periodic axis");
    }
    if (int rc = sync_engine(e)) return rc;
    int off = 0;
    for (int d = 0; d < 3; ++d) off += 4 * axis_n(e, d);
    std::vector<double> tab(off, 0.0);
    for (int d = 0; d < 3; ++d) {
        const int n = axis_n(e, d);
        const int lo = cells[2 * d], hi = cells[2 * d + 1];
        e->pml[2 * d] = lo;
        e->pml[2 * d + 1] = hi;
        e->pml_on[d] = (lo > 0 || hi > 0);
        dev_free(e->psibuf[d]);
        e->slab[d] = 0;
        const double dx = e->g.dx;
        // sigma(u) = sigma_max (u/d)^p, sigma_max = -(p+1) ln R / (2 d): meep::pml(d) defaults
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
        // the step kernels address the six auxiliary slots of a slab within one 2 GiB buffer range
        if ((double)e->prec * 6 * e->slab[d] >= 2147483648.0)
            return fail(e, SJF_E_UNSUPPORTED, "PML slab auxiliaries exceed 2 GiB on one axis");
        int rc = dev_alloc(e, &e->psibuf[d], (size_t)e->prec * 6 * e->slab[d]);
        if (rc) return rc;
    }
    int rc = e->prec == 4 ? convert_to_device<float>(e, e->coefbuf, tab.data(), off)
                          : convert_to_device<double>(e, e->coefbuf, tab.data(), off);
    if (rc) return rc;
    if (e->have_mat && e->pml_fref > 0) return apply_materials(e);  // the slabs moved
    e->version++;
    return SJF_OK;
}

int sjf_add_source(sjf_engine* e, int comp, const double box[6], const sjf_pulse* p, double amp_re,
                   double amp_im) {
    if (!e || !box || !p) return SJF_E_ARG;
    if (comp < 0 || comp > 5) return fail(e, SJF_E_ARG, "source component must be Ex..Hz");
    if (p->kind != SJF_SRC_GAUSSIAN && p->kind != SJF_SRC_CONTINUOUS) return fail(e, SJF_E_ARG, "bad pulse kind");
    if (!(p->freq > 0)) return fail(e, SJF_E_ARG, "pulse frequency must be positive");
    if (p->kind == SJF_SRC_GAUSSIAN && !(p->width > 0)) return fail(e, SJF_E_ARG, "pulse width must be positive");
    Source s;
    s.comp = comp;
    s.p = *p;
    std::memset(&s.dev, 0, sizeof(s.dev));
    s.dev.kind = p->kind;
    s.dev.amp_re = amp_re;
    s.dev.amp_im = amp_im;
    if (p->kind == SJF_SRC_GAUSSIAN) {
        // gaussian_src_time_phase constructor (reference src/disp.cpp:378-393)
        s.dev.omega = 2 * M_PI * p->freq;
        s.dev.width = p->width;
        s.dev.phi = p->phase + M_PI;
        s.dev.peak = 0.5 * (p->t_start + p->t_end);
        double cut = (p->t_end - p->t_start) * 0.5;
        while (std::exp(-cut * cut / (2 * p->width * p->width)) < 1e-100) cut *= 0.9;
        s.dev.cutoff = (double)(float)cut;
    } else {
        s.dev.freq = p->freq;
        s.dev.cw_width = p->width;
        s.dev.t_start = p->t_start;
        s.dev.t_end = p->t_end;
        s.dev.slow = p->slowness;
    }
    std::vector<int> ids[3];
    std::vector<double> ws[3];
    for (int d = 0; d < 3; ++d)
        axis_weights(axis_n(e, d), e->g.dx, kCompOff[comp][d], e->g.periodic[d], box[d], box[3 + d], ids[d], ws[d]);
    for (size_t a = 0; a < ids[0].size(); ++a)
        for (size_t b = 0; b < ids[1].size(); ++b)
            for (size_t c = 0; c < ids[2].size(); ++c) {
                s.idx.push_back(cell_index(e, ids[0][a], ids[1][b], ids[2][c]));
                s.wt.push_back(ws[0][a] * ws[1][b] * ws[2][c]);
            }
    e->src.push_back(std::move(s));
    e->version++;
    return SJF_OK;
}

int sjf_last_source_time(const sjf_engine* e, double* t) {
    if (!e || !t) return SJF_E_ARG;
    double m = 0;
    for (const auto& s : e->src) {
        // src_time::last_time rounds to float (gaussian_src_time_phase: src/disp.hpp:119)
        const double lt = s.p.kind == SJF_SRC_GAUSSIAN ? (double)(float)(s.dev.peak + s.dev.cutoff)
                                                       : (double)(float)s.p.t_end;
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
    dev_free(e->d_samples);
    e->samp_cap = 0;
    e->version++;
    return SJF_OK;
}

int sjf_set_save_span(sjf_engine* e, int span) {
    if (!e) return SJF_E_ARG;
    const int s = span > 0 ? span : 1;
    if (e->nsamp > 0 && s != e->save_span) return fail(e, SJF_E_STATE, "save span cannot change once sampling started");
    e->save_span = s;
    e->version++;
    return SJF_OK;
}

int sjf_read_monitors(sjf_engine* e, double* out, size_t cap, long* n_samples) {
    if (!e || !n_samples) return SJF_E_ARG;
    *n_samples = e->nsamp;
    const long npt = (long)e->mon.size();
    const long tot = e->nsamp * npt;
    if (!out || tot == 0) return SJF_OK;
    if ((size_t)(2 * tot) > cap) return fail(e, SJF_E_ARG, "output buffer too small");
    if (int rc = sync_engine(e)) return rc;
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
    if (int rc = sync_engine(e)) return rc;
    long idx[8];
    double w[8];
    trilinear(e, comp, xyz, idx, w);
    double v = 0;
    for (int q = 0; q < 8; ++q) {
        if (idx[q] < 0) continue;
        const char* base = (const char*)field_set(e) + field_offset(e, comp, idx[q]);
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

int sjf_add_dft(sjf_engine* e, const int lo[3], const int hi[3], int comp_mask, const double* freqs, int nf,
                int* id) {
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
    if (int rc = sync_engine(e)) return rc;
    const size_t bytes = sizeof(double2) * d.ncomp * nf * d.ncell;
    int rc = dev_alloc(e, (void**)&d.acc, bytes);
    if (!rc) rc = dev_alloc(e, (void**)&d.cmp, bytes);
    if (!rc) rc = dev_alloc(e, (void**)&d.d_freq, sizeof(double) * nf);
    if (!rc) rc = h2d(e, d.d_freq, d.freqs.data(), sizeof(double) * nf);
    if (rc) return rc;
    e->dft.push_back(d);
    *id = (int)e->dft.size() - 1;
    e->version++;
    return SJF_OK;
}

int sjf_read_dft(sjf_engine* e, int id, double* out, size_t cap) {
    if (!e || !out || id < 0 || id >= (int)e->dft.size()) return SJF_E_ARG;
    const Dft& d = e->dft[id];
    const long n = (long)d.ncomp * (long)d.freqs.size() * d.ncell;
    if ((size_t)(2 * n) > cap) return fail(e, SJF_E_ARG, "output buffer too small");
    if (int rc = sync_engine(e)) return rc;
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

int sjf_dft_flux(sjf_engine* e, int id, int axis, double* flux, size_t cap) {
    FluxPlan fp;
    if (int rc = flux_plan(e, id, axis, flux, cap, &fp)) return rc;
    if (int rc = sync_engine(e)) return rc;
    double* buf = nullptr;
    if (int rc = dev_alloc(e, (void**)&buf, sizeof(double) * fp.scratch_doubles())) return rc;
    int rc = flux_run(e, fp, buf, flux);
    dev_free(buf);
    return rc;
}

int sjf_dft_flux_noalloc(sjf_engine* e, int id, int axis, double* flux, size_t cap) {
    FluxPlan fp;
    if (int rc = flux_plan(e, id, axis, flux, cap, &fp)) return rc;
    if (fp.scratch_doubles() > (long)kFluxScratchDoubles)
        return fail(e, SJF_E_ARG, "flux needs " + std::to_string(fp.scratch_doubles()) +
                                      " doubles of scratch; the allocation-free buffer holds " +
                                      std::to_string(kFluxScratchDoubles));
    double* buf = scratch_on(e->g.device);
    if (!buf) return fail(e, SJF_E_STATE, "no scratch address for this device (sjf_create sets it up)");
    if (int rc = sync_engine(e)) return rc;
    return flux_run(e, fp, buf, flux);
}

int sjf_step(sjf_engine* e, long n) {
    if (!e || n < 0) return SJF_E_ARG;
    if (e->batch) return fail(e, SJF_E_STATE, "engine is stepped by its batch");
    HIPCHK(e, hipSetDevice(e->g.device));
    return run_set(own_set(e), n);
}

int sjf_reserve_steps(sjf_engine* e, long n) {
    if (!e || n < 0) return SJF_E_ARG;
    if (e->batch) return fail(e, SJF_E_STATE, "engine is stepped by its batch");
    HIPCHK(e, hipSetDevice(e->g.device));
    return prepare_set(own_set(e), n);
}

int sjf_synchronize(sjf_engine* e) {
    if (!e) return SJF_E_ARG;
    return sync_engine(e);
}

int sjf_set_graph(sjf_engine* e, int chunk) {
    if (!e || chunk < 0) return SJF_E_ARG;
    e->graph_chunk = chunk;
    e->version++;
    return SJF_OK;
}

int sjf_timed_steps(sjf_engine* e, long n, double* ms_total, double* ms_kernel_mean) {
    if (!e || n < 1 || !ms_total || !ms_kernel_mean) return SJF_E_ARG;
    if (e->batch) return fail(e, SJF_E_STATE, "engine is stepped by its batch");
    HIPCHK(e, hipSetDevice(e->g.device));
    return timed_set(own_set(e), n, ms_total, ms_kernel_mean);
}

int sjf_read_field(sjf_engine* e, int comp, double* out) {
    if (!e || !out || comp < 0 || comp > 5) return SJF_E_ARG;
    if (int rc = sync_engine(e)) return rc;
    return e->prec == 4 ? read_component<float>(e, comp, out) : read_component<double>(e, comp, out);
}

int sjf_read_field_box(sjf_engine* e, int comp, const int lo[3], const int hi[3], double* out) {
    if (!e || !lo || !hi || comp < 0 || comp > 5) return SJF_E_ARG;
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    for (int d = 0; d < 3; ++d)
        if (lo[d] < 0 || hi[d] > n[d] || lo[d] > hi[d]) return fail(e, SJF_E_ARG, "field box outside the grid");
    if (lo[0] == hi[0] || lo[1] == hi[1] || lo[2] == hi[2]) return SJF_OK;  // empty box: nothing to copy
    if (!out) return SJF_E_ARG;
    if (int rc = sync_engine(e)) return rc;
    return e->prec == 4 ? read_component_box<float>(e, comp, lo, hi, out)
                        : read_component_box<double>(e, comp, lo, hi, out);
}

int sjf_write_field(sjf_engine* e, int comp, const double* in) {
    if (!e || !in || comp < 0 || comp > 5) return SJF_E_ARG;
    if (int rc = sync_engine(e)) return rc;
    return e->prec == 4 ? write_component<float>(e, comp, in) : write_component<double>(e, comp, in);
}

int sjf_step_bytes(const sjf_engine* e, double* bytes) {
    if (!e || !bytes) return SJF_E_ARG;
    *bytes = engine_step_bytes(e);
    return SJF_OK;
}

int sjf_launch_info(const sjf_engine* e, int* form, int* ty, int* xchunk, int* threads, int* groups) {
    if (!e) return SJF_E_ARG;
    if (form) *form = e->form;
    if (ty) *ty = e->ty;
    if (xchunk) *xchunk = e->xc;
    if (threads) *threads = block_threads(e);
    if (groups) *groups = tiles_of(e, e->ty, e->xc);
    return SJF_OK;
}

int sjf_set_zsegment(sjf_engine* e, int cells) {
    if (!e) return SJF_E_ARG;
    if (cells != 0 && (cells < 64 || cells > 1024 || cells % 64 != 0))
        return fail(e, SJF_E_ARG, "z segment must be 0 or a multiple of 64 from 64 to 1024");
    if (cells > 0 && cells < e->g.nz && e->g.periodic[2])
        return fail(e, SJF_E_ARG, "z passes need a non-periodic z axis");
    e->zseg = cells;
    e->version++;
    return SJF_OK;
}

int sjf_set_tile(sjf_engine* e, int form, int ty, int xchunk) {
    if (!e) return SJF_E_ARG;
    if (form == SJF_TILE_REGISTER_ROWS) {
        if (ty != 1 && ty != 2 && ty != 4) return fail(e, SJF_E_ARG, "rows per thread must be 1, 2 or 4");
    } else if (form == SJF_TILE_THREAD_ROWS) {
        if (!thread_rows_fit(e, ty))
            return fail(e, SJF_E_ARG, "row-of-threads tile needs (rows + 1) * nz <= 1024 and 4-byte-addressable arrays");
    } else {
        return fail(e, SJF_E_ARG, "unknown tile form");
    }
    if (xchunk < 1) return fail(e, SJF_E_ARG, "xchunk must be >= 1");
    e->form = form;
    e->ty = ty;
    e->xc = std::min(xchunk, e->g.nx);
    e->tile_set = true;
    e->version++;
    return SJF_OK;
}

// ---- batches --------------------------------------------------------------------------------

int sjf_batch_create(sjf_engine* const* engines, int n, sjf_batch** out) {
    if (!engines || n < 1 || !out) {
        g_create_error = "null argument";
        return SJF_E_ARG;
    }
    *out = nullptr;
    sjf_engine* e0 = engines[0];
    if (!e0) return SJF_E_ARG;
    long span = -1;
    for (int i = 0; i < n; ++i) {
        sjf_engine* e = engines[i];
        if (!e) return SJF_E_ARG;
        std::string why;
        if (e->batch) why = "engine already belongs to a batch";
        else if (e->g.device != e0->g.device) why = "engines of a batch must share a device";
        else if (e->prec != e0->prec) why = "engines of a batch must share a precision";
        else if (e->step != e0->step) why = "engines of a batch must be at the same step";
        else if (!e->mon.empty()) {
            if (span >= 0 && span != e->save_span) why = "engines of a batch must share a save span";
            span = e->save_span;
        }
        for (int j = 0; j < i && why.empty(); ++j)
            if (engines[j] == e) why = "engine listed twice";
        if (!why.empty()) {
            g_create_error = why;
            return SJF_E_ARG;
        }
    }
    if (hipSetDevice(e0->g.device) != hipSuccess) {
        g_create_error = "hipSetDevice failed";
        return SJF_E_HIP;
    }
    for (int i = 0; i < n; ++i)
        if (hipStreamSynchronize(engines[i]->stream) != hipSuccess) {
            g_create_error = "stream synchronisation failed";
            return SJF_E_HIP;
        }
    sjf_batch* b = new sjf_batch();
    LaunchSet* s = new LaunchSet();
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        delete b;
        g_create_error = "hipStreamCreate failed";
        return SJF_E_HIP;
    }
    s->own_stream = true;
    s->prec = e0->prec;
    s->ty = 1;
    // measured on MI355X: with some z line shorter than 96 cells (the sweep of 64 instances,
    // 64^3 to 192x192x128) the row-of-threads form with 6-plane x-chunks (2.78 ms per step)
    // beats one row per thread (2.89-2.96 ms at 6-32 planes); with every line of 96 cells or more
    // (the 16 design variants, 192x192x96) one row per thread with 16-plane chunks wins (3.38
    // against 3.54 ms per gradient step)
    bool short_z = false;
    for (int i = 0; i < n; ++i) short_z = short_z || engines[i]->g.nz < 96;
    s->form = short_z ? SJF_TILE_THREAD_ROWS : SJF_TILE_REGISTER_ROWS;
    s->xc = short_z ? 6 : 16;
    for (int i = 0; i < n; ++i)
        if (!thread_rows_fit(engines[i], default_thread_rows(engines[i]))) {
            s->form = SJF_TILE_REGISTER_ROWS;
            s->xc = 16;
        }
    for (int i = 0; i < n; ++i) {
        s->eng.push_back(engines[i]);
        engines[i]->batch = b;
    }
    b->set = s;
    *out = b;
    return SJF_OK;
}

int sjf_batch_set_tile(sjf_batch* b, int form, int xchunk) {
    if (!b || xchunk < 0) return SJF_E_ARG;
    LaunchSet* s = b->set;
    if (form == SJF_TILE_THREAD_ROWS) {
        for (sjf_engine* e : s->eng)
            if (!thread_rows_fit(e, default_thread_rows(e))) {
                b->err = "an instance does not fit the row-of-threads tile";
                return SJF_E_ARG;
            }
    } else if (form != SJF_TILE_REGISTER_ROWS) {
        b->err = "unknown tile form";
        return SJF_E_ARG;
    }
    s->form = form;
    s->ty = 1;
    s->xc = xchunk;
    s->built = false;
    return SJF_OK;
}

int sjf_batch_step(sjf_batch* b, long n) {
    if (!b || n < 0) return SJF_E_ARG;
    LaunchSet* s = b->set;
    if (hipSetDevice(s->eng[0]->g.device) != hipSuccess) return SJF_E_HIP;
    const int rc = run_set(s, n);
    if (rc) b->err = s->eng[0]->err;
    return rc;
}

int sjf_batch_reserve_steps(sjf_batch* b, long n) {
    if (!b || n < 0) return SJF_E_ARG;
    LaunchSet* s = b->set;
    if (hipSetDevice(s->eng[0]->g.device) != hipSuccess) return SJF_E_HIP;
    const int rc = prepare_set(s, n);
    if (rc) b->err = s->eng[0]->err;
    return rc;
}

int sjf_batch_timed_steps(sjf_batch* b, long n, double* ms_total, double* ms_kernel_mean) {
    if (!b || n < 1 || !ms_total || !ms_kernel_mean) return SJF_E_ARG;
    LaunchSet* s = b->set;
    if (hipSetDevice(s->eng[0]->g.device) != hipSuccess) return SJF_E_HIP;
    const int rc = timed_set(s, n, ms_total, ms_kernel_mean);
    if (rc) b->err = s->eng[0]->err;
    return rc;
}

int sjf_batch_synchronize(sjf_batch* b) {
    if (!b) return SJF_E_ARG;
    if (hipStreamSynchronize(b->set->stream) != hipSuccess) return SJF_E_HIP;
    return SJF_OK;
}

int sjf_batch_info(const sjf_batch* b, long* cells, double* bytes_per_step) {
    if (!b) return SJF_E_ARG;
    long c = 0;
    double by = 0;
    for (const sjf_engine* e : b->set->eng) {
        c += e->N;
        by += engine_step_bytes(e);
    }
    if (cells) *cells = c;
    if (bytes_per_step) *bytes_per_step = by;
    return SJF_OK;
}

int sjf_batch_launch_info(const sjf_batch* b, int* threads, int* groups) {
    if (!b) return SJF_E_ARG;
    if (!b->set->built) return SJF_E_STATE;
    if (threads) *threads = b->set->block;
    if (groups) *groups = b->set->ntiles;
    return SJF_OK;
}

int sjf_batch_adjoint_run(sjf_batch* b, long nsteps, int checkpoint_every, const int* gaps, double* J,
                          double* const* grads, const size_t* caps) {
    if (!b || nsteps < 1 || checkpoint_every < 0 || !gaps || !J || !grads || !caps) return SJF_E_ARG;
    for (const sjf_engine* e : b->set->eng)
        if (!e->emap_h.empty()) {
            b->err = "adjoint runs take the LUT's eps_inf (no eps maps)";
            return SJF_E_UNSUPPORTED;
        }
    LaunchSet* s = b->set;
    sjf_engine* e0 = s->eng[0];
    const int n = (int)s->eng.size();
    std::vector<long> q(n, 0), nd(n, 0);
    std::string why;
    for (int i = 0; i < n && why.empty(); ++i) {
        const sjf_engine* e = s->eng[i];
        const int ext[3] = {e->g.nx, e->g.ny, e->g.nz};
        if (e->step != 0) why = "the adjoint run starts from step 0";
        else if (e->deps_h.empty()) why = "an engine has no design region (sjf_set_design_eps)";
        else if (e->g.nz != e0->g.nz) why = "engines of a batched adjoint run must share nz";
        else if (e->g.nz > 1024) why = "the reverse step spans the z line: nz <= 1024";
        else if (np_variant_of(e->maxp, e->uniform) != np_variant_of(e0->maxp, e0->uniform))
            why = "engines of a batched adjoint run must share the pole-slot variant";
        else if (!grads[i]) why = "null gradient buffer";
        for (int d = 0; d < 3 && why.empty(); ++d)
            if (gaps[3 * i + d] < 0 || gaps[3 * i + d] >= ext[d]) why = "objective cell outside the grid";
        if (why.empty()) {
            nd[i] = (long)e->deps_h.size();
            if (caps[i] < (size_t)nd[i]) why = "gradient buffer smaller than the design region";
            q[i] = cell_index(e, gaps[3 * i], gaps[3 * i + 1], gaps[3 * i + 2]);
        }
    }
    if (!why.empty()) {
        b->err = why;
        return SJF_E_ARG;
    }
    if (hipSetDevice(e0->g.device) != hipSuccess) return SJF_E_HIP;
    double ms = 0;
    int th = 0, gr = 0;
    const int rc = s->prec == 4
                       ? batch_adjoint_run_t<float>(s, nsteps, checkpoint_every, q.data(), nd.data(), J, grads, &ms,
                                                    &th, &gr)
                       : batch_adjoint_run_t<double>(s, nsteps, checkpoint_every, q.data(), nd.data(), J, grads, &ms,
                                                     &th, &gr);
    if (rc) {
        b->err = e0->err;
        return rc;
    }
    b->adj_ms = ms;
    b->adj_threads = th;
    b->adj_groups = gr;
    return SJF_OK;
}

int sjf_batch_adjoint_kernel_info(const sjf_batch* b, double* ms_mean, int* threads, int* groups) {
    if (!b || !ms_mean) return SJF_E_ARG;
    if (b->adj_groups == 0) return SJF_E_STATE;
    *ms_mean = b->adj_ms;
    if (threads) *threads = b->adj_threads;
    if (groups) *groups = b->adj_groups;
    return SJF_OK;
}

const char* sjf_batch_last_error(const sjf_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int sjf_batch_destroy(sjf_batch* b) {
    if (!b) return SJF_E_ARG;
    LaunchSet* s = b->set;
    (void)hipSetDevice(s->eng[0]->g.device);
    (void)hipStreamSynchronize(s->stream);
    for (sjf_engine* e : s->eng) e->batch = nullptr;
    destroy_set(s);
    delete b;
    return SJF_OK;
}

// ---- measurement probe ------------------------------------------------------------------------

int sjf_copy_probe(int device, size_t bytes, int iters, double* gbs) {
    if (!gbs || bytes < 16 || iters < 1) return SJF_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return SJF_E_HIP;
    const long n = (long)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n * 16) != hipSuccess) return SJF_E_NOMEM;
    if (hipMalloc((void**)&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return SJF_E_NOMEM;
    }
    (void)hipMemset(a, 0, n * 16);
    (void)hipMemset(b, 0, n * 16);
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(sjf::k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(t0, 0);
    for (int it = 0; it < iters; ++it) hipLaunchKernelGGL(sjf::k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(t1, 0);
    const hipError_t st = hipEventSynchronize(t1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (st != hipSuccess) return SJF_E_HIP;
    *gbs = 2.0 * n * 16 * iters / (ms * 1e-3) / 1e9;
    return SJF_OK;
}

}  // extern "C"
