/*
 * fdtd_oracle.c -- CPU FP64 restatement of the FDTD time step that sim_juncs hands to
 * meep (src/disp.cpp:740, `fields.step()`), plus the driver-side pieces that feed it.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed CPU baseline.
 * Nothing in the product (libsjf) links, loads or calls it; only tests/, the smoke check
 * and bench.py's cpu_baseline leg do.
 *
 * Structure restated (meep 1.24 is not vendored in the reference; its step order is
 * restated from meep's published model and pinned here by closed-form physics checks):
 *   1. B/H update      H^{n+1/2} = H^{n-1/2} - dt * curl~ E^n          (mu = 1)
 *   2. D update        D^{n+1}   = D^n + dt * (curl~ H^{n+1/2} - J^{n+1/2})
 *   3. polarization    P_k^{n+1} = [(2 - L w0^2 dt^2) P^n - (1 - g) P^{n-1}
 *                                    + sigma w0^2 dt^2 E^n] / (1 + g),  g = gamma' dt / 2
 *                      (Lorentz L = 1, Drude L = 0: src/disp.cpp:529-548 builds
 *                       lorentzian_susceptibility(f0/um_scale, gamma/um_scale, drude))
 *   4. E update        E^{n+1} = (D^{n+1} - sum_k P_k^{n+1}) / eps_inf
 * curl~ is the Yee curl with stretched-coordinate PML realised as CPML auxiliary psi
 * (kappa = 1, alpha = 0), sigma(u) = sigma_max (u/d)^p, sigma_max = -(p+1) ln R / (2 d)
 * (src/disp.cpp:527 passes meep::pml(d); R = 1e-15, p = 2 are meep's defaults).
 *
 * Source time functions: gaussian_src_time_phase (src/disp.cpp:378-400) restated exactly;
 * continuous_src_time restated from meep 1.24's published definition (tanh turn-on).
 * The injected current is the centred difference of dipole(t) (meep src_time::current).
 * Monitors follow bound_geom::run (src/disp.cpp:719-741): sample Ex before the step when
 * step % save_span == 0.  DFT accumulators are new (no counterpart in the reference).
 *
 * Grid conventions (shared with the HIP engine, documented in include/sjf.h):
 *   idx = (i*ny + j)*nz + k, z fastest.  Ex at ((i+1/2)dx, j dx, k dx), Ey at
 *   (i dx, (j+1/2)dx, k dx), Ez at (i dx, j dx, (k+1/2)dx), Hx at (i, j+1/2, k+1/2),
 *   Hy at (i+1/2, j, k+1/2), Hz at (i+1/2, j+1/2, k).  Fields outside the grid are zero
 *   (metallic outer edge) unless the axis is periodic.
 *
 * What pins each part (tests/test_oracle_physics.py unless named otherwise):
 *   update_H / update_D curls   Yee plane-wave dispersion, discrete energy invariant, Courant
 *                               blow-up
 *   poles (update_PE)           Drude plasma and Lorentz longitudinal oscillations, static limit
 *   CPML                        reflection against a doubled domain
 *   orc_set_pml_dispersion      CW wave at f_ref meets no interface at the folded slabs
 *   orc_set_eps_maps            plane-wave dispersion with a different eps per component
 *   orc_set_eps_box             equals the same values given as maps
 *   gauss_dipole, add_sources   window and phase known answers (reference src/main_test.cpp),
 *                               current-sheet Green's function
 *   cw_dipole                   closed form (logistic envelope, steady state, float gate), sampled
 *                               current, CW sheet Green's function
 *   add_h_sources               magnetic sheet: dual of the electric sheet's field
 *   sample_monitors             trilinear interpolation exact on linear fields
 *   accumulate_dft              equals the sum over the sampled series
 *   flux (engine only)          plane-wave invariants (tests/test_gpu_flux.py)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXP 4
#define ORC_MAXMAT 256

enum { C_EX = 0, C_EY, C_EZ, C_HX, C_HY, C_HZ };

typedef struct {
    int kind; /* 0 gaussian, 1 continuous */
    double omega, width, phi, peak, cutoff;       /* gaussian */
    double freq, cw_width, t_start, t_end, slow;  /* continuous */
    double amp_re, amp_im;
    int comp;
    long n;
    long* idx;
    double* wt;
} orc_src;

typedef struct {
    int comp;
    long c_idx[8];
    double c_wt[8];
} orc_mon;

typedef struct {
    int lo[3], hi[3];
    int mask;
    int nf;
    double* freq;
    double* acc; /* [comp 6][nf][cells][2] (only masked comps used) */
    long ncell;
} orc_dft;

typedef struct orc {
    int n[3];
    long N;
    double dx, dt;
    int bc[3];
    int part; /* 0: real part of the complex source, 1: imaginary part */
    double* F[6];  /* Ex Ey Ez Hx Hy Hz */
    double* D[3];
    double* inv_eps[3];
    int has_pml[3];
    int npml[6];
    double* bE[3];
    double* aE[3];
    double* bH[3];
    double* aH[3];
    double* psiE[3][3]; /* [axis][E comp] */
    double* psiH[3][3]; /* [axis][H comp] */
    uint32_t* matw;
    double fold_fref; /* > 0: Lorentz poles folded out of the PML slabs at this frequency */
    double* emap[3];  /* per-cell eps_inf of Ex, Ey, Ez (orc_set_eps_maps), or NULL */
    int nmat;
    double eps_inf[ORC_MAXMAT];
    int npole[ORC_MAXMAT];
    double pole[ORC_MAXMAT][ORC_MAXP][4];
    int maxp;
    double* P[ORC_MAXP][3];
    double* Pm[ORC_MAXP][3];
    long step;
    int nsrc;
    orc_src* src;
    int nmon;
    orc_mon* mon;
    int save_span;
    long nsamp, samp_cap;
    double* samples; /* [nsamp][nmon] */
    int ndft;
    orc_dft* dft;
} orc;

static long IDX(const orc* o, int i, int j, int k) { return ((long)i * o->n[1] + j) * o->n[2] + k; }

/* ------------------------------------------------------------------ allocation ---- */

static double* zalloc(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

orc* orc_create(int nx, int ny, int nz, double dx, double courant, const int* bc, int part) {
    orc* o = (orc*)calloc(1, sizeof(orc));
    o->n[0] = nx; o->n[1] = ny; o->n[2] = nz;
    o->N = (long)nx * ny * nz;
    o->dx = dx;
    o->dt = courant * dx;
    for (int d = 0; d < 3; d++) o->bc[d] = bc ? bc[d] : 0;
    o->part = part;
    for (int c = 0; c < 6; c++) o->F[c] = zalloc(o->N);
    for (int c = 0; c < 3; c++) {
        o->D[c] = zalloc(o->N);
        o->inv_eps[c] = (double*)malloc(sizeof(double) * o->N);
        for (long q = 0; q < o->N; q++) o->inv_eps[c][q] = 1.0;
    }
    o->save_span = 1;
    return o;
}

void orc_destroy(orc* o) {
    if (!o) return;
    for (int c = 0; c < 6; c++) free(o->F[c]);
    for (int c = 0; c < 3; c++) { free(o->D[c]); free(o->inv_eps[c]); free(o->emap[c]); }
    for (int d = 0; d < 3; d++) {
        free(o->bE[d]); free(o->aE[d]); free(o->bH[d]); free(o->aH[d]);
        for (int c = 0; c < 3; c++) { free(o->psiE[d][c]); free(o->psiH[d][c]); }
    }
    for (int p = 0; p < ORC_MAXP; p++)
        for (int c = 0; c < 3; c++) { free(o->P[p][c]); free(o->Pm[p][c]); }
    free(o->matw);
    for (int s = 0; s < o->nsrc; s++) { free(o->src[s].idx); free(o->src[s].wt); }
    free(o->src);
    free(o->mon);
    free(o->samples);
    for (int q = 0; q < o->ndft; q++) { free(o->dft[q].freq); free(o->dft[q].acc); }
    free(o->dft);
    free(o);
}

double orc_dt(const orc* o) { return o->dt; }
long orc_step_count(const orc* o) { return o->step; }

/* ------------------------------------------------------------------ materials ----- */

/* matw: one word per cell, bits 0-7 / 8-15 / 16-23 = material of Ex / Ey / Ez.
 * poles: [nmat][ORC_MAXP][4] = {f0, gamma, sigma, is_drude} in meep units.
 * Restates structure_from_settings (src/disp.cpp:482-550): eps_inf from chi1p1, one
 * lorentzian_susceptibility per parsed pole with (omega_0, gamma) already / um_scale. */
int orc_set_materials(orc* o, const uint32_t* matw, int nmat, const double* eps_inf,
                      const int* npole, const double* poles) {
    if (nmat < 1 || nmat > ORC_MAXMAT - 1) return -1; /* id 255 is the design voxel */
    o->nmat = nmat;
    o->maxp = 0;
    for (int m = 0; m < nmat; m++) {
        o->eps_inf[m] = eps_inf[m];
        o->npole[m] = npole[m];
        if (npole[m] < 0 || npole[m] > ORC_MAXP) return -1;
        if (npole[m] > o->maxp) o->maxp = npole[m];
        for (int p = 0; p < ORC_MAXP; p++)
            for (int q = 0; q < 4; q++) o->pole[m][p][q] = poles[(m * ORC_MAXP + p) * 4 + q];
    }
    free(o->matw);
    o->matw = (uint32_t*)malloc(sizeof(uint32_t) * o->N);
    memcpy(o->matw, matw, sizeof(uint32_t) * o->N);
    for (long q = 0; q < o->N; q++)
        for (int c = 0; c < 3; c++) {
            int m = (matw[q] >> (8 * c)) & 0xff;
            if (m >= nmat) return -2;
            o->inv_eps[c][q] = 1.0 / o->eps_inf[m];
        }
    for (int p = 0; p < o->maxp; p++)
        for (int c = 0; c < 3; c++) {
            if (!o->P[p][c]) { o->P[p][c] = zalloc(o->N); o->Pm[p][c] = zalloc(o->N); }
        }
    return 0;
}

/* Per-cell eps_inf of the three E components (eps: 3 * N values, [Ex Ey Ez][cell]), in place of
 * the materials' eps_inf: the fractional permittivity of the reference's smoothed material
 * function (chi1p1, src/disp.cpp:264-283 with smooth_n > 0).  Call after orc_set_materials and
 * before orc_set_eps_box and orc_set_pml_dispersion. */
int orc_set_eps_maps(orc* o, const double* eps) {
    for (int c = 0; c < 3; c++) {
        free(o->emap[c]);
        o->emap[c] = (double*)malloc(sizeof(double) * o->N);
        memcpy(o->emap[c], eps + (long)c * o->N, sizeof(double) * o->N);
        for (long q = 0; q < o->N; q++) o->inv_eps[c][q] = 1.0 / eps[(long)c * o->N + q];
    }
    return 0;
}

/* Per-voxel permittivity over a box of cells (all three E components of a voxel take
 * the voxel's eps).  Used for design regions; overrides eps_inf there, no poles. */
int orc_set_eps_box(orc* o, const int* lo, const int* hi, const double* eps) {
    long q = 0;
    for (int i = lo[0]; i < hi[0]; i++)
        for (int j = lo[1]; j < hi[1]; j++)
            for (int k = lo[2]; k < hi[2]; k++, q++) {
                long id = IDX(o, i, j, k);
                for (int c = 0; c < 3; c++) o->inv_eps[c][id] = 1.0 / eps[q];
                if (o->matw) o->matw[id] = 0x00ffffffu; /* id 255: no poles */
            }
    return 0;
}

/* ------------------------------------------------------------------ PML media ----- */

static int in_pml_cell(const orc* o, long q) {
    int ijk[3];
    ijk[2] = (int)(q % o->n[2]);
    ijk[1] = (int)((q / o->n[2]) % o->n[1]);
    ijk[0] = (int)(q / ((long)o->n[1] * o->n[2]));
    for (int d = 0; d < 3; d++)
        if (o->has_pml[d] && (ijk[d] < o->npml[2 * d] || ijk[d] >= o->n[d] - o->npml[2 * d + 1])) return 1;
    return 0;
}

/* eps_inf of a cell of material m (base: the material's eps_inf, or the cell's map value) inside a
 * PML slab when Lorentz poles are folded out at frequency f: base + sum over Lorentz poles of
 * Re chi(f), chi = sigma w0^2 / (w0^2 - w^2 - i w gamma'), kept at base when the sum falls below 1
 * (metals). */
static double folded_eps(const orc* o, int m, double base, double f) {
    double w = 2 * M_PI * f, e = base;
    for (int p = 0; p < o->npole[m]; p++) {
        if (o->pole[m][p][3] != 0.0) continue; /* Drude pole: kept as a pole */
        double w0 = 2 * M_PI * o->pole[m][p][0], g = 2 * M_PI * o->pole[m][p][1], sg = o->pole[m][p][2];
        double complex_re = sg * w0 * w0 * (w0 * w0 - w * w) /
                            ((w0 * w0 - w * w) * (w0 * w0 - w * w) + w * w * g * g);
        e += complex_re;
    }
    return e >= 1.0 ? e : base;
}

static int has_lorentz(const orc* o, int m) {
    for (int p = 0; p < o->npole[m]; p++)
        if (o->pole[m][p][3] == 0.0) return 1;
    return 0;
}

/* Lorentz poles inside a PML admit growing modes; with f > 0 the cells of the PML slabs run
 * their material without its Lorentz poles, whose real susceptibility at f joins eps_inf
 * (update_PE skips those poles there).  Call after orc_set_materials, orc_set_eps_box and
 * orc_set_pml. */
int orc_set_pml_dispersion(orc* o, double f) {
    o->fold_fref = f;
    if (!o->matw || f <= 0) return 0;
    for (long q = 0; q < o->N; q++) {
        if (!in_pml_cell(o, q)) continue;
        for (int c = 0; c < 3; c++) {
            int m = (o->matw[q] >> (8 * c)) & 0xff;
            if (m < o->nmat && has_lorentz(o, m)) {
                double base = o->emap[c] ? o->emap[c][q] : o->eps_inf[m];
                o->inv_eps[c][q] = 1.0 / folded_eps(o, m, base, f);
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ PML ----------- */

/* sigma at coordinate x (meep units) along axis d; the low PML is [0, nlo dx], the high
 * PML is [(n - nhi) dx, n dx]. */
static double sigma_at(int n, double dx, int nlo, int nhi, double R, double p, double x) {
    double s = 0.0;
    if (nlo > 0) {
        double L = nlo * dx, u = L - x;
        if (u > 0) s += -(p + 1) * log(R) / (2 * L) * pow(u / L, p);
    }
    if (nhi > 0) {
        double L = nhi * dx, u = x - (n - nhi) * dx;
        if (u > 0) s += -(p + 1) * log(R) / (2 * L) * pow(u / L, p);
    }
    return s;
}

int orc_set_pml(orc* o, const int* cells, double R, double power) {
    for (int d = 0; d < 3; d++) {
        o->npml[2 * d] = cells[2 * d];
        o->npml[2 * d + 1] = cells[2 * d + 1];
        o->has_pml[d] = (cells[2 * d] > 0 || cells[2 * d + 1] > 0);
        if (!o->has_pml[d]) continue;
        int n = o->n[d];
        o->bE[d] = zalloc(n); o->aE[d] = zalloc(n);
        o->bH[d] = zalloc(n); o->aH[d] = zalloc(n);
        for (int i = 0; i < n; i++) {
            double sE = sigma_at(n, o->dx, cells[2 * d], cells[2 * d + 1], R, power, i * o->dx);
            double sH = sigma_at(n, o->dx, cells[2 * d], cells[2 * d + 1], R, power, (i + 0.5) * o->dx);
            o->bE[d][i] = exp(-sE * o->dt);
            o->aE[d][i] = o->bE[d][i] - 1.0;
            o->bH[d][i] = exp(-sH * o->dt);
            o->aH[d][i] = o->bH[d][i] - 1.0;
        }
        for (int c = 0; c < 3; c++) {
            if (c == d) continue;
            o->psiE[d][c] = zalloc(o->N);
            o->psiH[d][c] = zalloc(o->N);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ sources ------- */

static double complex_re_im(double re, double im, int part) { return part ? im : re; }

/* gaussian_src_time_phase::dipole (src/disp.cpp:395-400) */
static void gauss_dipole(const orc_src* s, double t, double* re, double* im) {
    double tt = t - s->peak;
    if ((double)(float)fabs(tt) > s->cutoff) { *re = 0; *im = 0; return; }
    double g = exp(-tt * tt / (2 * s->width * s->width));
    double th = -s->omega * tt - s->phi;
    /* polar(1, th) * 1/(-i omega) = (cos th + i sin th) * (i / omega) */
    *re = -g * sin(th) / s->omega;
    *im = g * cos(th) / s->omega;
}

/* meep 1.24 continuous_src_time::dipole (restated from meep's published source) */
static void cw_dipole(const orc_src* s, double t, double* re, double* im) {
    float rt = (float)t;
    if (rt < s->t_start || rt > s->t_end) { *re = 0; *im = 0; return; }
    double w = 2 * M_PI * s->freq;
    double env = 1.0;
    if (s->cw_width != 0.0) {
        double ts = (t - s->t_start) / s->cw_width - s->slow;
        double te = (s->t_end - t) / s->cw_width - s->slow;
        env = (1.0 + tanh(ts)) * (1.0 + tanh(te)) * 0.25;
    }
    double th = -w * t;
    *re = -env * sin(th) / w;
    *im = env * cos(th) / w;
}

static void src_dipole(const orc_src* s, double t, double* re, double* im) {
    if (s->kind == 0) gauss_dipole(s, t, re, im);
    else cw_dipole(s, t, re, im);
}

/* Exposed for tests: dipole(t) of a continuous source (f, width, t_start, t_end, slowness). */
void orc_cw_dipole(double f, double w, double st, double et, double slow, double t, double* re,
                   double* im) {
    orc_src s;
    memset(&s, 0, sizeof(s));
    s.kind = 1;
    s.freq = f;
    s.cw_width = w;
    s.t_start = st;
    s.t_end = et;
    s.slow = slow;
    cw_dipole(&s, t, re, im);
}

/* Exposed for tests: dipole(t) of a gaussian source (f, w, phase, st, et). */
void orc_gaussian_dipole(double f, double w, double phase, double st, double et, double t,
                         double* re, double* im, double* cutoff_out) {
    orc_src s;
    memset(&s, 0, sizeof(s));
    s.omega = 2 * M_PI * f;
    s.width = w;
    s.phi = phase + M_PI;
    s.peak = 0.5 * (st + et);
    s.cutoff = (et - st) * 0.5;
    while (exp(-s.cutoff * s.cutoff / (2 * w * w)) < 1e-100) s.cutoff *= 0.9;
    s.cutoff = (double)(float)s.cutoff;
    gauss_dipole(&s, t, re, im);
    if (cutoff_out) *cutoff_out = s.cutoff;
}

/* 1-D weights of a source extent [lo, hi] onto a lattice with positions (i + off) dx.
 * Zero extent: linear interpolation / dx (delta function).  Otherwise: overlap of [lo, hi]
 * with the cell [p - dx/2, p + dx/2] divided by dx. */
static int axis_weights(int n, double dx, double off, int periodic, double lo, double hi,
                        int* ids, double* w) {
    int cnt = 0;
    if (hi < lo) { double t = lo; lo = hi; hi = t; }
    if (hi - lo < 1e-9 * dx) {
        double s = lo / dx - off;
        double fl = floor(s);
        int i0 = (int)fl;
        double f = s - fl;
        for (int q = 0; q < 2; q++) {
            int i = i0 + q;
            double ww = (q == 0 ? 1.0 - f : f) / dx;
            if (ww == 0.0) continue;
            if (periodic) i = ((i % n) + n) % n;
            else if (i < 0 || i >= n) continue;
            ids[cnt] = i; w[cnt] = ww; cnt++;
        }
        return cnt;
    }
    /* on a periodic axis the cell of lattice point p also has images p +- n dx */
    for (int i = 0; i < n; i++) {
        double ww = 0;
        for (int m = periodic ? -1 : 0; m <= (periodic ? 1 : 0); m++) {
            double p = (i + off) * dx + m * n * dx;
            double a = fmax(lo, p - 0.5 * dx), b = fmin(hi, p + 0.5 * dx);
            if (b > a) ww += (b - a) / dx;
        }
        if (ww > 0) { ids[cnt] = i; w[cnt] = ww; cnt++; }
    }
    return cnt;
}

static const double COMP_OFF[6][3] = {
    {0.5, 0.0, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5},
    {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}};

/* kind 0 gaussian: params = {f, w, phase, t_start, t_end}
 * kind 1 continuous: params = {f, width, t_start, t_end, slowness}
 * box = {x0, y0, z0, x1, y1, z1} in meep units.  comp 0..2: electric current J added to the
 * D update; comp 3..5: magnetic current K added to the B (= H) update (the reference passes any
 * component to fields::add_volume_source, src/disp.cpp:614,619). */
int orc_add_source(orc* o, int comp, const double* box, int kind, const double* prm,
                   double amp_re, double amp_im) {
    if (comp < 0 || comp > 5) return -1;
    o->src = (orc_src*)realloc(o->src, sizeof(orc_src) * (o->nsrc + 1));
    orc_src* s = &o->src[o->nsrc++];
    memset(s, 0, sizeof(*s));
    s->kind = kind;
    s->comp = comp;
    s->amp_re = amp_re;
    s->amp_im = amp_im;
    if (kind == 0) {
        s->omega = 2 * M_PI * prm[0];
        s->width = prm[1];
        s->phi = prm[2] + M_PI;
        s->peak = 0.5 * (prm[3] + prm[4]);
        s->cutoff = (prm[4] - prm[3]) * 0.5;
        while (exp(-s->cutoff * s->cutoff / (2 * s->width * s->width)) < 1e-100) s->cutoff *= 0.9;
        s->cutoff = (double)(float)s->cutoff;
    } else {
        s->freq = prm[0];
        s->cw_width = prm[1];
        s->t_start = prm[2];
        s->t_end = prm[3];
        s->slow = prm[4];
    }
    int* ids[3];
    double* ws[3];
    int cnt[3];
    for (int d = 0; d < 3; d++) {
        ids[d] = (int*)malloc(sizeof(int) * (o->n[d] + 2));
        ws[d] = (double*)malloc(sizeof(double) * (o->n[d] + 2));
        cnt[d] = axis_weights(o->n[d], o->dx, COMP_OFF[comp][d], o->bc[d], box[d], box[3 + d],
                              ids[d], ws[d]);
    }
    s->n = (long)cnt[0] * cnt[1] * cnt[2];
    s->idx = (long*)malloc(sizeof(long) * (s->n ? s->n : 1));
    s->wt = (double*)malloc(sizeof(double) * (s->n ? s->n : 1));
    long q = 0;
    for (int a = 0; a < cnt[0]; a++)
        for (int b = 0; b < cnt[1]; b++)
            for (int c = 0; c < cnt[2]; c++, q++) {
                s->idx[q] = IDX(o, ids[0][a], ids[1][b], ids[2][c]);
                s->wt[q] = ws[0][a] * ws[1][b] * ws[2][c];
            }
    for (int d = 0; d < 3; d++) { free(ids[d]); free(ws[d]); }
    return o->nsrc - 1;
}

/* Latest time any source is on: gaussian last_time() = float(peak + cutoff)
 * (src/disp.hpp:119); continuous: end time. */
double orc_last_source_time(const orc* o) {
    double t = 0;
    for (int s = 0; s < o->nsrc; s++) {
        /* src_time::last_time rounds to float (gaussian_src_time_phase: src/disp.hpp:119) */
        double lt = o->src[s].kind == 0 ? (double)(float)(o->src[s].peak + o->src[s].cutoff)
                                          : (double)(float)o->src[s].t_end;
        if (lt > t) t = lt;
    }
    return t;
}

/* ------------------------------------------------------------------ monitors ------ */

int orc_add_monitors(orc* o, int comp, const double* xyz, int npt) {
    o->mon = (orc_mon*)realloc(o->mon, sizeof(orc_mon) * (o->nmon + npt));
    for (int p = 0; p < npt; p++) {
        orc_mon* m = &o->mon[o->nmon + p];
        m->comp = comp;
        int i0[3];
        double f[3];
        for (int d = 0; d < 3; d++) {
            double s = xyz[3 * p + d] / o->dx - COMP_OFF[comp][d];
            double fl = floor(s);
            i0[d] = (int)fl;
            f[d] = s - fl;
        }
        int q = 0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++, q++) {
                    int ii[3] = {i0[0] + a, i0[1] + b, i0[2] + c};
                    double w = (a ? f[0] : 1 - f[0]) * (b ? f[1] : 1 - f[1]) * (c ? f[2] : 1 - f[2]);
                    int ok = 1;
                    for (int d = 0; d < 3; d++) {
                        if (o->bc[d]) ii[d] = ((ii[d] % o->n[d]) + o->n[d]) % o->n[d];
                        else if (ii[d] < 0 || ii[d] >= o->n[d]) ok = 0;
                    }
                    m->c_idx[q] = ok ? IDX(o, ii[0], ii[1], ii[2]) : -1;
                    m->c_wt[q] = ok ? w : 0.0;
                }
    }
    o->nmon += npt;
    return 0;
}

void orc_set_save_span(orc* o, int span) { o->save_span = span > 0 ? span : 1; }

static void sample_monitors(orc* o) {
    if (o->nmon == 0) return;
    if (o->nsamp == o->samp_cap) {
        o->samp_cap = o->samp_cap ? 2 * o->samp_cap : 64;
        o->samples = (double*)realloc(o->samples, sizeof(double) * o->samp_cap * o->nmon);
    }
    for (int p = 0; p < o->nmon; p++) {
        const orc_mon* m = &o->mon[p];
        double v = 0;
        for (int q = 0; q < 8; q++)
            if (m->c_idx[q] >= 0) v += m->c_wt[q] * o->F[m->comp][m->c_idx[q]];
        o->samples[o->nsamp * o->nmon + p] = v;
    }
    o->nsamp++;
}

long orc_num_samples(const orc* o) { return o->nsamp; }
void orc_read_monitors(const orc* o, double* out) {
    memcpy(out, o->samples, sizeof(double) * o->nsamp * o->nmon);
}

/* ------------------------------------------------------------------ DFT ----------- */

int orc_add_dft(orc* o, const int* lo, const int* hi, int mask, const double* freqs, int nf) {
    o->dft = (orc_dft*)realloc(o->dft, sizeof(orc_dft) * (o->ndft + 1));
    orc_dft* d = &o->dft[o->ndft];
    memset(d, 0, sizeof(*d));
    for (int a = 0; a < 3; a++) { d->lo[a] = lo[a]; d->hi[a] = hi[a]; }
    d->mask = mask;
    d->nf = nf;
    d->freq = (double*)malloc(sizeof(double) * nf);
    memcpy(d->freq, freqs, sizeof(double) * nf);
    d->ncell = (long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    d->acc = zalloc(6L * nf * d->ncell * 2);
    return o->ndft++;
}

static void accumulate_dft(orc* o) {
    for (int q = 0; q < o->ndft; q++) {
        orc_dft* d = &o->dft[q];
        for (int c = 0; c < 6; c++) {
            if (!(d->mask >> c & 1)) continue;
            double t = c < 3 ? (o->step + 1) * o->dt : (o->step + 0.5) * o->dt;
            for (int f = 0; f < d->nf; f++) {
                double th = 2 * M_PI * d->freq[f] * t;
                double cr = cos(th) * o->dt, ci = sin(th) * o->dt;
                double* acc = d->acc + ((long)c * d->nf + f) * d->ncell * 2;
                long p = 0;
                for (int i = d->lo[0]; i < d->hi[0]; i++)
                    for (int j = d->lo[1]; j < d->hi[1]; j++)
                        for (int k = d->lo[2]; k < d->hi[2]; k++, p++) {
                            double v = o->F[c][IDX(o, i, j, k)];
                            acc[2 * p] += v * cr;
                            acc[2 * p + 1] += v * ci;
                        }
            }
        }
    }
}

/* out: [comp][nf][ncell][2] for the components in the mask, in component order */
int orc_read_dft(const orc* o, int id, double* out) {
    const orc_dft* d = &o->dft[id];
    long per = (long)d->nf * d->ncell * 2;
    long w = 0;
    for (int c = 0; c < 6; c++) {
        if (!(d->mask >> c & 1)) continue;
        memcpy(out + w, d->acc + c * per, sizeof(double) * per);
        w += per;
    }
    return 0;
}

/* ------------------------------------------------------------------ fields -------- */

void orc_get_field(const orc* o, int comp, double* out) { memcpy(out, o->F[comp], sizeof(double) * o->N); }

/* Sets a field array; for E components D is made consistent (D = eps_inf E + sum P). */
void orc_set_field(orc* o, int comp, const double* in) {
    memcpy(o->F[comp], in, sizeof(double) * o->N);
    if (comp < 3) {
        for (long q = 0; q < o->N; q++) {
            double d = in[q] / o->inv_eps[comp][q];
            for (int p = 0; p < o->maxp; p++) d += o->P[p][comp][q];
            o->D[comp][q] = d;
        }
    }
}

/* ------------------------------------------------------------------ stepping ------ */

/* neighbour index along axis d at offset +1 (dir=+1) or -1; -1 when outside a metallic
 * edge. */
static inline long nb(const orc* o, int i, int j, int k, int d, int dir) {
    int ii[3] = {i, j, k};
    ii[d] += dir;
    if (ii[d] < 0 || ii[d] >= o->n[d]) {
        if (!o->bc[d]) return -1;
        ii[d] = (ii[d] + o->n[d]) % o->n[d];
    }
    return IDX(o, ii[0], ii[1], ii[2]);
}

static inline double at(const double* a, long q) { return q < 0 ? 0.0 : a[q]; }

static inline int in_pml(const orc* o, int d, int i) {
    return o->has_pml[d] && (i < o->npml[2 * d] || i >= o->n[d] - o->npml[2 * d + 1]);
}

/* derivative of field `f` along axis d at cell (i,j,k), forward (+1) for H updates and
 * backward (-1) for E updates, with the CPML correction applied in place to psi. */
static inline double deriv(orc* o, const double* f, long id, int i, int j, int k, int d,
                           int fwd, double* psi) {
    double v;
    if (fwd) v = (at(f, nb(o, i, j, k, d, +1)) - f[id]) / o->dx;
    else v = (f[id] - at(f, nb(o, i, j, k, d, -1))) / o->dx;
    int ii = d == 0 ? i : (d == 1 ? j : k);
    if (psi && in_pml(o, d, ii)) {
        double b = fwd ? o->bH[d][ii] : o->bE[d][ii];
        double a = fwd ? o->aH[d][ii] : o->aE[d][ii];
        psi[id] = b * psi[id] + a * v;
        v += psi[id];
    }
    return v;
}

static void update_H(orc* o) {
    double *Ex = o->F[0], *Ey = o->F[1], *Ez = o->F[2];
    double *Hx = o->F[3], *Hy = o->F[4], *Hz = o->F[5];
    double dt = o->dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->n[0]; i++)
        for (int j = 0; j < o->n[1]; j++)
            for (int k = 0; k < o->n[2]; k++) {
                long id = IDX(o, i, j, k);
                /* Hx: -(dEz/dy - dEy/dz) */
                double cx = deriv(o, Ez, id, i, j, k, 1, 1, o->psiH[1][0]) -
                            deriv(o, Ey, id, i, j, k, 2, 1, o->psiH[2][0]);
                double cy = deriv(o, Ex, id, i, j, k, 2, 1, o->psiH[2][1]) -
                            deriv(o, Ez, id, i, j, k, 0, 1, o->psiH[0][1]);
                double cz = deriv(o, Ey, id, i, j, k, 0, 1, o->psiH[0][2]) -
                            deriv(o, Ex, id, i, j, k, 1, 1, o->psiH[1][2]);
                Hx[id] -= dt * cx;
                Hy[id] -= dt * cy;
                Hz[id] -= dt * cz;
            }
}

static void update_D(orc* o) {
    double *Hx = o->F[3], *Hy = o->F[4], *Hz = o->F[5];
    double dt = o->dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->n[0]; i++)
        for (int j = 0; j < o->n[1]; j++)
            for (int k = 0; k < o->n[2]; k++) {
                long id = IDX(o, i, j, k);
                double cx = deriv(o, Hz, id, i, j, k, 1, 0, o->psiE[1][0]) -
                            deriv(o, Hy, id, i, j, k, 2, 0, o->psiE[2][0]);
                double cy = deriv(o, Hx, id, i, j, k, 2, 0, o->psiE[2][1]) -
                            deriv(o, Hz, id, i, j, k, 0, 0, o->psiE[0][1]);
                double cz = deriv(o, Hy, id, i, j, k, 0, 0, o->psiE[0][2]) -
                            deriv(o, Hx, id, i, j, k, 1, 0, o->psiE[1][2]);
                o->D[0][id] += dt * cx;
                o->D[1][id] += dt * cy;
                o->D[2][id] += dt * cz;
            }
}

/* Electric currents: D^{n+1} -= dt J^{n+1/2}, J^{n+1/2} = (dipole(t_{n+1}) - dipole(t_n)) / dt. */
static void add_sources(orc* o) {
    double t0 = o->step * o->dt, t1 = (o->step + 1) * o->dt;
    for (int s = 0; s < o->nsrc; s++) {
        const orc_src* src = &o->src[s];
        if (src->comp > 2) continue;
        double r0, i0, r1, i1;
        src_dipole(src, t0, &r0, &i0);
        src_dipole(src, t1, &r1, &i1);
        double jr = (r1 - r0) / o->dt, ji = (i1 - i0) / o->dt;
        /* amp * J */
        double ar = src->amp_re * jr - src->amp_im * ji;
        double ai = src->amp_re * ji + src->amp_im * jr;
        double J = complex_re_im(ar, ai, o->part);
        double* D = o->D[src->comp];
        for (long q = 0; q < src->n; q++) D[src->idx[q]] -= o->dt * J * src->wt[q];
    }
}

/* Magnetic currents, after the H update: H^{n+1/2} -= dt K^n, K^n = (dipole(t_n + dt/2) -
 * dipole(t_n - dt/2)) / dt (mu = 1; dB/dt = -curl E - K). */
static void add_h_sources(orc* o) {
    double t0 = (o->step - 0.5) * o->dt, t1 = (o->step + 0.5) * o->dt;
    for (int s = 0; s < o->nsrc; s++) {
        const orc_src* src = &o->src[s];
        if (src->comp < 3) continue;
        double r0, i0, r1, i1;
        src_dipole(src, t0, &r0, &i0);
        src_dipole(src, t1, &r1, &i1);
        double kr = (r1 - r0) / o->dt, ki = (i1 - i0) / o->dt;
        double ar = src->amp_re * kr - src->amp_im * ki;
        double ai = src->amp_re * ki + src->amp_im * kr;
        double K = complex_re_im(ar, ai, o->part);
        double* H = o->F[src->comp];
        for (long q = 0; q < src->n; q++) H[src->idx[q]] -= o->dt * K * src->wt[q];
    }
}

static void update_PE(orc* o) {
    double dt = o->dt;
    /* per material, per pole: coefficients of the textbook ADE */
    double cA[ORC_MAXMAT][ORC_MAXP], cB[ORC_MAXMAT][ORC_MAXP], cC[ORC_MAXMAT][ORC_MAXP];
    for (int m = 0; m < o->nmat; m++)
        for (int p = 0; p < o->npole[m]; p++) { /* npole[255] stays 0: design voxels */
            double w0 = 2 * M_PI * o->pole[m][p][0];
            double gm = 2 * M_PI * o->pole[m][p][1];
            double sg = o->pole[m][p][2];
            double L = o->pole[m][p][3] != 0.0 ? 0.0 : 1.0;
            double g = gm * dt / 2;
            cA[m][p] = (2 - L * w0 * w0 * dt * dt) / (1 + g);
            cB[m][p] = -(1 - g) / (1 + g);
            cC[m][p] = sg * w0 * w0 * dt * dt / (1 + g);
        }
#pragma omp parallel for schedule(static)
    for (long q = 0; q < o->N; q++) {
        int fold = o->fold_fref > 0 && o->maxp && in_pml_cell(o, q);
        for (int c = 0; c < 3; c++) {
            double E = o->F[c][q];
            double sumP = 0;
            if (o->maxp) {
                int m = (o->matw[q] >> (8 * c)) & 0xff;
                for (int p = 0; p < o->npole[m]; p++) {
                    if (fold && o->pole[m][p][3] == 0.0) continue; /* Lorentz pole folded out */
                    double pn = o->P[p][c][q], pm = o->Pm[p][c][q];
                    double pp = cA[m][p] * pn + cB[m][p] * pm + cC[m][p] * E;
                    o->Pm[p][c][q] = pn;
                    o->P[p][c][q] = pp;
                    sumP += pp;
                }
            }
            o->F[c][q] = (o->D[c][q] - sumP) * o->inv_eps[c][q];
        }
    }
}

void orc_step(orc* o, long nsteps) {
    for (long s = 0; s < nsteps; s++) {
        if (o->step % o->save_span == 0) sample_monitors(o);
        update_H(o);
        add_h_sources(o);
        update_D(o);
        add_sources(o);
        update_PE(o);
        accumulate_dft(o);
        o->step++;
    }
}

/* Discrete field energy W = 1/2 sum(eps E^n.E^n + H^{n+1/2}.H^{n-1/2}) requires H at two
 * half steps; tests compute it from field snapshots.  This helper returns sum eps E.E. */
double orc_e_energy(const orc* o) {
    double w = 0;
    for (int c = 0; c < 3; c++)
        for (long q = 0; q < o->N; q++) w += o->F[c][q] * o->F[c][q] / o->inv_eps[c][q];
    return w;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
