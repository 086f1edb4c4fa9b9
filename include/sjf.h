/*
 * sjf.h -- C ABI of the MI355X FDTD engine (libsjf.so).
 *
 * The engine replaces the meep objects that sim_juncs' driver `bound_geom` drives
 * (reference src/disp.hpp:142-200, src/disp.cpp:482-749).  Every entry point below names
 * the reference call it stands in for.  No C++ or HIP types cross this boundary: opaque
 * handle, plain pointers and sizes, `int` status codes (0 = success), never an exception.
 *
 * Units are meep units (c = 1, length unit a = 1/um_scale um; reference Appendix of
 * src/disp.cpp:5-22, src/disp.hpp:157-158).  Grid conventions:
 *   cell (i,j,k), i < nx, j < ny, k < nz; linear index (i*ny + j)*nz + k (z fastest);
 *   Ex at ((i+1/2)dx, j dx, k dx), Ey at (i dx, (j+1/2)dx, k dx), Ez at (i dx, j dx, (k+1/2)dx),
 *   Hx at (i dx, (j+1/2)dx, (k+1/2)dx), Hy at ((i+1/2)dx, j dx, (k+1/2)dx),
 *   Hz at ((i+1/2)dx, (j+1/2)dx, k dx).
 *   After n steps E is at time n*dt and H at (n - 1/2)*dt, dt = courant*dx.
 *   Fields outside the grid are zero (metallic outer edge) unless an axis is periodic.
 *
 * Threading: one engine owns one HIP stream and is not thread-safe; separate engines may
 * be driven from separate threads.  Entry points enqueue work on the engine's stream and
 * return; the read functions synchronise that stream before copying.
 */
#ifndef SJF_H
#define SJF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sjf_engine sjf_engine;
typedef struct sjf_batch sjf_batch;

/* status codes (mirror the reference's parse_ercode style: src/cgs_read.hpp:32) */
enum {
    SJF_OK = 0,
    SJF_E_ARG = -1,         /* invalid argument */
    SJF_E_HIP = -2,         /* HIP runtime error (message in sjf_last_error) */
    SJF_E_STATE = -3,       /* call not valid in the engine's current state */
    SJF_E_UNSUPPORTED = -4, /* valid request outside what this build supports */
    SJF_E_NOMEM = -5        /* device or host allocation failed */
};

/* field components (meep::Ex..meep::Hz order used by the reference, src/disp.cpp:335-341) */
enum { SJF_EX = 0, SJF_EY = 1, SJF_EZ = 2, SJF_HX = 3, SJF_HY = 4, SJF_HZ = 5 };

enum { SJF_FP32 = 4, SJF_FP64 = 8 };
enum { SJF_SRC_GAUSSIAN = 0, SJF_SRC_CONTINUOUS = 1 };

#define SJF_MAX_POLES 4
#define SJF_MAX_MATERIALS 255 /* material id 255 is reserved for design voxels */
#define SJF_MAX_DFT_FREQS 64

typedef struct {
    int nx, ny, nz;  /* cells per axis; any parity (nz <= 1024) */
    double dx;       /* cell size in meep length units (1/resolution) */
    double courant;  /* dt = courant * dx; 0.5 is meep's default and what the reference runs
                        (the reference parses `courant` but never passes it: src/argparse.h:142,
                        src/disp.cpp:527) */
    int precision;   /* SJF_FP32 or SJF_FP64 field storage and arithmetic */
    int periodic[3]; /* per axis: 0 metallic outer edge, 1 periodic */
    int part;        /* 0: evolve the real part of meep's complex fields, 1: the imaginary
                        part (the reference never calls use_real_fields(), so meep runs
                        complex fields; with real materials the parts are independent) */
    int device;      /* HIP device ordinal */
} sjf_grid;

/* One material: eps_inf plus Lorentz/Drude poles, meep units (f0 and gamma already divided
 * by um_scale, sigma already divided by the 2-D thickness: src/disp.cpp:537-546).
 * chi(omega) = sum sigma w0^2 / (L w0^2 - w^2 - i w gamma'), w0 = 2 pi f0, gamma' = 2 pi gamma,
 * L = 0 for Drude ("drude" poles: lorentzian_susceptibility(..., no_omega_0_denominator)). */
typedef struct {
    double eps_inf;
    int npoles;
    double f0[SJF_MAX_POLES];
    double gamma[SJF_MAX_POLES];
    double sigma[SJF_MAX_POLES];
    int is_drude[SJF_MAX_POLES];
} sjf_material;

/* Source time function.  SJF_SRC_GAUSSIAN restates gaussian_src_time_phase
 * (src/disp.cpp:378-400): omega = 2 pi freq, phi' = phase + pi, peak = (t_start+t_end)/2,
 * cutoff = (t_end - t_start)/2 shrunk while exp(-cutoff^2/2w^2) < 1e-100 then rounded to
 * float, dipole(t) = exp(-tt^2/2w^2) exp(-i(omega tt + phi')) / (-i omega) for
 * float(|tt|) <= cutoff.  SJF_SRC_CONTINUOUS restates meep::continuous_src_time
 * (src/disp.cpp:618): exp(-i omega t)/(-i omega) with a tanh turn-on of `width` and
 * `slowness`.  The injected current is (dipole(t_{n+1}) - dipole(t_n)) / dt. */
typedef struct {
    int kind;
    double freq;     /* meep frequency = 1/(wavelength_um * um_scale) (src/disp.cpp:606) */
    double width;    /* meep time units (src/disp.cpp:607) */
    double phase;    /* CEP phase, radians */
    double t_start;  /* meep time units (src/disp.cpp:608) */
    double t_end;    /* meep time units (src/disp.cpp:609) */
    double slowness; /* continuous only; meep's default is 3 */
} sjf_pulse;

/* ---- lifetime ------------------------------------------------------------------------ */

/* Replaces meep::vol3d(...) + new meep::structure(...) + meep::fields(strct)
 * (src/disp.cpp:509,527,560).  Allocates all field arrays on `g->device`. */
int sjf_create(const sjf_grid* g, sjf_engine** out);
/* Replaces ~bound_geom's delete strct (src/disp.cpp:647-649). */
int sjf_destroy(sjf_engine* e);
/* Message of the last failure on `e` (or of the last failed sjf_create when e == NULL). */
const char* sjf_last_error(const sjf_engine* e);

/* ---- structure ----------------------------------------------------------------------- */

/* Replaces the structure's chi1p1 sampling and add_susceptibility calls
 * (src/disp.cpp:515-548, in_bound at src/disp.cpp:264-283).  mat_word holds one uint32 per
 * cell: bits 0-7, 8-15, 16-23 = material id of the cell's Ex, Ey, Ez Yee point (ids index
 * `lut`, n_mat <= 255).  Resets all polarisation state. */
int sjf_set_materials(sjf_engine* e, const uint32_t* mat_word, const sjf_material* lut,
                      int n_mat);
/* Per-cell eps_inf of the three E components, replacing the LUT's eps_inf everywhere outside the
 * design region: eps holds 3 * nx*ny*nz values, [Ex Ey Ez][x][y][z] (z fastest), each > 0; poles
 * stay those of each cell's material.  This is how the fractional permittivity of the reference's
 * smoothed material function reaches the engine: chi1p1 = sum over roots of [ambient + (eps_root -
 * ambient) * in_bound fraction] (src/disp.cpp:56-112, 264-283 with smooth_n > 0), a continuum that a
 * 255-entry LUT cannot hold.  Inside the PML slabs under sjf_set_pml_dispersion the folded Lorentz
 * susceptibility is added to the map's value.  eps = NULL removes the map.  Adjoint runs need the
 * LUT's eps_inf (SJF_E_UNSUPPORTED with a map).  Resets the polarisation state. */
int sjf_set_eps_maps(sjf_engine* e, const double* eps);
/* Per-voxel permittivity (no poles) over the cell box [lo, hi).  New; used for the design
 * region of the adjoint workloads.  eps is (hi-lo) sized, z fastest. */
int sjf_set_design_eps(sjf_engine* e, const int lo[3], const int hi[3], const double* eps);
/* Replaces meep::pml(thickness) (src/disp.cpp:527).  cells = {x_lo, x_hi, y_lo, y_hi, z_lo,
 * z_hi} PML thicknesses in cells; R = asymptotic reflection (meep default 1e-15); power =
 * profile exponent (meep default 2). */
int sjf_set_pml(sjf_engine* e, const int cells[6], double R, double power);
/* Media of the PML slabs.  A Lorentz medium inside the PML admits modes that grow without bound
 * once the pulse has gone (measured on the junctions: a Lorentz SiO2 substrate reaching the PML
 * grows x3-5 per 500 steps from the rounding floor; FP32 runs leave the 1e-4 tolerance within
 * ~6000 steps and blow up past ~10000, FP64 ones later).  With f_ref > 0 every cell inside a PML
 * slab runs its material without Lorentz poles: their real susceptibility at f_ref joins eps_inf
 * (kept at eps_inf when the sum falls below 1, as for metals) and Drude poles stay, which keeps
 * the slabs matched to the medium near f_ref.  f_ref = 0 (the default) runs the given materials
 * in the PML as well.  Callable in any order with sjf_set_materials and sjf_set_pml before the
 * first step; each re-derives the map and resets the polarisation state.  New (meep keeps the
 * poles in its PML, whose formulation differs). */
int sjf_set_pml_dispersion(sjf_engine* e, double f_ref);

/* ---- sources, monitors, DFT ------------------------------------------------------------ */

/* Replaces fields::add_volume_source(c, src, volume, amp) (src/disp.cpp:603,614,619,680) and
 * add_point_source (src/disp.cpp:659, zero-extent box).  comp in Ex..Hz: Ex..Ez are electric
 * currents J added to the E update, E^{n+1} -= dt J^{n+1/2} w / eps, J^{n+1/2} = (dipole(t_{n+1}) -
 * dipole(t_n)) / dt; Hx..Hz are magnetic currents K added to the H update, H^{n+1/2} -= dt K^n w,
 * K^n = (dipole(t_n + dt/2) - dipole(t_n - dt/2)) / dt (dB/dt = -curl E - K, mu = 1).  box = {x0,
 * y0,z0,x1,y1,z1} in meep units; an axis of zero extent is a delta function (linear weights / dx),
 * otherwise cells weigh by their overlap with the box. */
int sjf_add_source(sjf_engine* e, int comp, const double box[6], const sjf_pulse* p,
                   double amp_re, double amp_im);
/* Replaces fields::last_source_time() (src/disp.cpp:625). */
int sjf_last_source_time(const sjf_engine* e, double* t);
/* Replaces fields.dt and fields.time() (src/disp.cpp:704,733). */
int sjf_dt(const sjf_engine* e, double* dt);
int sjf_time(const sjf_engine* e, double* t);
int sjf_step_count(const sjf_engine* e, long* n);

/* Replaces the per-span fields.get_field(c, loc) loop of bound_geom::run
 * (src/disp.cpp:721-730): registers points sampled on the device, by trilinear interpolation
 * on the component's lattice, at the start of every step whose index is a multiple of the
 * save span (sample, then step, as src/disp.cpp:721-740). */
int sjf_add_monitors(sjf_engine* e, int comp, const double* xyz, int n);
/* save_span of bound_geom (src/disp.hpp:192, src/disp.cpp:699: 0 means 1). */
int sjf_set_save_span(sjf_engine* e, int span);
/* out[s][p][2] = (value, 0) for sample s, point p.  *n_samples = samples taken so far; at
 * most cap doubles are written. */
int sjf_read_monitors(sjf_engine* e, double* out, size_t cap, long* n_samples);
/* Replaces fields::get_field(c, vec) (src/disp.cpp:724) for a one-off read. */
int sjf_get_field(sjf_engine* e, int comp, const double xyz[3], double out[2]);

/* DFT accumulation over the cell box [lo, hi) (new: the reference only FFTs monitor series
 * after the run, src/disp.cpp:796-806).  After each step adds F(t) exp(+i 2 pi f t) dt with t
 * = the component's own time, in FP64 with Neumaier-compensated sums.  comp_mask bit c
 * selects component c.  nf <= SJF_MAX_DFT_FREQS. */
int sjf_add_dft(sjf_engine* e, const int lo[3], const int hi[3], int comp_mask,
                const double* freqs, int nf, int* id);
/* out[c][f][cell][2] for the masked components in ascending order. */
int sjf_read_dft(sjf_engine* e, int id, double* out, size_t cap);
/* Flux through DFT region `id`, a plane one cell thick along `axis` that holds both tangential
 * E and both tangential H components: flux[f] = dx^2 * Re sum_cells (E_a conj(H_b) - E_b conj(H_a))
 * with (axis, a, b) in cyclic order, the Poynting flux along +axis of the accumulated spectra
 * (each component at its own Yee point, not interpolated), reduced on the device per frequency
 * in a fixed order; cap >= the region's frequencies.  New (the reference takes no flux: the
 * derived monitor of meep's add_flux / get_fluxes, for the DFT planes of this engine). */
int sjf_dft_flux(sjf_engine* e, int id, int axis, double* flux, size_t cap);
/* sjf_dft_flux without any device allocation: the reduction's partial sums and results go to a
 * static device buffer of 64 x 65 doubles (33 KiB) that each device holds until the process ends,
 * so the region may have at most SJF_MAX_DFT_FREQS frequencies (SJF_E_ARG beyond).  The buffer is
 * shared by every call on a device, so calls that overlap on one device race.  Same results as
 * sjf_dft_flux, bit for bit.  No counterpart in the reference: for callers that cannot allocate
 * device memory (code inside a hipGraph capture, serving loops, harnesses that forbid hipMalloc). */
int sjf_dft_flux_noalloc(sjf_engine* e, int id, int axis, double* flux, size_t cap);

/* ---- time stepping ---------------------------------------------------------------------- */

/* Replaces fields.step() (src/disp.cpp:740).  n = 1 is the exact drop-in; n > 1 runs n steps
 * back to back (with the monitor sampling rule above). */
int sjf_step(sjf_engine* e, long n);
/* Waits for all work enqueued on the engine's stream. */
int sjf_synchronize(sjf_engine* e);
/* Makes the next n steps allocation-free: sizes the monitor sample buffer for them and builds the
 * launch tables now (sjf_step otherwise does both on demand, which allocates and synchronises).
 * New (callers that time or capture a run of steps). */
int sjf_reserve_steps(sjf_engine* e, long n);
/* Steps per captured hipGraph replayed by sjf_step (rounded up to a multiple of the save span);
 * 0 launches every step directly.  Results are identical either way. */
int sjf_set_graph(sjf_engine* e, int chunk);
/* Runs n steps and reports, from HIP events on the engine's stream, the wall time of the
 * whole region and the mean duration of the fused update kernel (ms). */
int sjf_timed_steps(sjf_engine* e, long n, double* ms_total, double* ms_kernel_mean);

/* ---- field I/O --------------------------------------------------------------------------- */

/* Replaces output_hdf5 (src/disp.cpp:696,735) as the data source: copies a whole component
 * (nx*ny*nz values, converted to FP64) to host memory. */
int sjf_read_field(sjf_engine* e, int comp, double* out);
/* The cell box [lo, hi) of one component (output_hdf5 over a sub-volume, src/disp.cpp:696,735
 * passes vol.surroundings(), the whole grid, which sjf_read_field covers), converted to FP64,
 * row-major over (x, y, z) of the box.  An empty box copies nothing; a box outside the grid is
 * SJF_E_ARG. */
int sjf_read_field_box(sjf_engine* e, int comp, const int lo[3], const int hi[3], double* out);
/* Sets a whole component from host FP64 values (initial conditions, tests). */
int sjf_write_field(sjf_engine* e, int comp, const double* in);

/* ---- accounting -------------------------------------------------------------------------- */

/* Algorithmic HBM bytes one launch of the fused update kernel must move, from the actual masks:
 * N (12 p + 4) for E and H read and written and the material word (no word for a uniform
 * medium), 3 p per pole and dispersive E component (P^n, P^{n-1} read, P^{n+1} written), 8 p per
 * PML cell and axis (two E and two H auxiliaries read and written) and p per design voxel (its
 * 1/eps); p = 4 (FP32) or 8 (FP64). */
int sjf_step_bytes(const sjf_engine* e, double* bytes);
/* Forms of the fused update kernel.  REGISTER_ROWS: one workgroup per z line, each thread
 * updates `rows` (1, 2 or 4) consecutive y rows and recomputes H on the row below them.
 * THREAD_ROWS: a workgroup holds rows + 1 rows of nz threads and updates `rows` y rows,
 * exchanging y and z neighbours through LDS; needs (rows + 1) * nz <= 1024 and every array
 * addressable with 32-bit byte offsets (cells * precision < 2^32). */
#define SJF_TILE_REGISTER_ROWS 1
#define SJF_TILE_THREAD_ROWS 2
/* Launch configuration of the fused update kernel: form, y rows (per thread or per
 * workgroup), x-chunk length, threads per workgroup, workgroups. */
int sjf_launch_info(const sjf_engine* e, int* form, int* rows, int* xchunk, int* threads, int* groups);
/* Override the default tile shape (xchunk >= 1: planes each workgroup marches along x) of the
 * forward step.  The reverse step of sjf_adjoint_run has its own (sjf_set_adjoint_tile). */
int sjf_set_tile(sjf_engine* e, int form, int rows, int xchunk);
/* REGISTER_ROWS tile: a workgroup covers its z line in passes of `cells` cells (a multiple of
 * 64, at most 1024), the last lane of a pass handing its H to the first lane of the next through
 * LDS.  0 = the default: 128-cell passes with one row per thread on a non-periodic z line that is
 * a multiple of 256 cells above 256 or longer than 1024, else one pass over the whole line.
 * Passes need one row per thread and a non-periodic z axis.  Results are the same for every
 * segment length.  New (tuning, and z lines longer than one workgroup). */
int sjf_set_zsegment(sjf_engine* e, int cells);

/* ---- adjoint -----------------------------------------------------------------------------
 * New (the reference has no gradients).  Runs `nsteps` steps from step 0 and returns the
 * objective J = dt * sum_{n=1..nsteps} |E^n(gap)|^2 (E of the cell `gap`, all three components)
 * and its exact gradient with respect to every design-region permittivity (sjf_set_design_eps;
 * grad in the same z-fastest order, cap >= number of design voxels).  The gradient is the
 * discrete adjoint of the update the engine runs, applied in reverse (Yee curls, CPML and pole
 * recursions).  The reverse pass needs E of the design region and of `gap` at every step:
 *   checkpoint_every = K > 0: the forward pass stores the full state every K steps and the
 *     reverse pass recomputes each segment from its checkpoint;
 *   checkpoint_every = 0: the forward pass records those values after every step when the record
 *     fits a quarter of the free device memory (no recomputation), else it checkpoints every
 *     ~sqrt(nsteps) steps.
 * Both give identical results.  Monitors and DFT regions see the forward pass once; the engine
 * ends in its final forward state. */
int sjf_adjoint_run(sjf_engine* e, long nsteps, int checkpoint_every, const int gap[3], double* J,
                    double* grad, size_t cap);
/* Tile of the fused reverse step of sjf_adjoint_run: SJF_TILE_REGISTER_ROWS (rows = 1: one y row
 * per workgroup) or SJF_TILE_THREAD_ROWS (`rows` owned y rows of nz threads each plus one halo
 * row per workgroup, the y and z neighbours exchanged through LDS; needs (rows + 1) * nz <= 1024
 * and the exchange planes within 64 KiB of LDS), marching over `xchunk` x planes (0: enough
 * chunks for about 16384 workgroups).  The default is one row per workgroup.  The gradient is
 * the same to rounding either way (the same arithmetic per cell).  A batched adjoint run runs one row per workgroup unless its first engine's tile was set
 * to the row-of-threads form: then every engine runs that many rows (at most its ny) over the
 * first engine's x-chunk (6 planes where it gave none).  New (tuning). */
int sjf_set_adjoint_tile(sjf_engine* e, int form, int rows, int xchunk);
/* The fused reverse-step kernel of the last sjf_adjoint_run: mean duration in ms (HIP events on
 * the engine's stream around the first min(nsteps, 64) reverse launches), threads per workgroup
 * and workgroups.  SJF_E_STATE before any adjoint run.  It moves the same algorithmic bytes per
 * launch as the forward update (sjf_step_bytes).  New (accounting). */
int sjf_adjoint_kernel_info(const sjf_engine* e, double* ms_mean, int* threads, int* groups);

/* ---- batches ----------------------------------------------------------------------------
 * New (the reference sweeps parameters with one meep process per SLURM task,
 * scripts/run.sh:47-78): engines of any grid shapes, same device, precision, step count and
 * save span, advanced together.  Each time step runs as one launch of the fused update over
 * every engine's tiles, one source launch, one DFT launch and (every save span) one monitor
 * launch; runs of steps are replayed from a captured hipGraph.  Engines stay readable with the
 * sjf_read_* calls (which wait for the batch) but cannot be stepped on their own until the
 * batch is destroyed. */
int sjf_batch_create(sjf_engine* const* engines, int n, sjf_batch** out);
/* Tile form of the batch's fused launch (SJF_TILE_REGISTER_ROWS with one row per thread, or
 * SJF_TILE_THREAD_ROWS with each instance's default rows) and the x-chunk of every instance
 * (0: each engine's own).  New (tuning of batched runs). */
int sjf_batch_set_tile(sjf_batch* b, int form, int xchunk);
int sjf_batch_step(sjf_batch* b, long n);
/* sjf_reserve_steps for every engine of the batch and the batch's launch tables. */
int sjf_batch_reserve_steps(sjf_batch* b, long n);
int sjf_batch_synchronize(sjf_batch* b);
int sjf_batch_timed_steps(sjf_batch* b, long n, double* ms_total, double* ms_kernel_mean);
/* Total cells and algorithmic bytes of one fused launch over the batch. */
int sjf_batch_info(const sjf_batch* b, long* cells, double* bytes_per_step);
/* Threads per workgroup and workgroups of the batch's fused launch as last built (SJF_E_STATE
 * before the first step). */
int sjf_batch_launch_info(const sjf_batch* b, int* threads, int* groups);
/* sjf_adjoint_run for every engine of the batch in one pass: the forward steps are the batch's
 * fused launches, and each reverse step is one fused reverse launch over every engine's tiles
 * (plus one gradient and one injection launch for all engines).  checkpoint_every as in
 * sjf_adjoint_run (0: record mode when the record of all engines fits a quarter of the free
 * device memory, else checkpoints every ~sqrt(nsteps) steps).  Engine i gets J[i] and its
 * gradient in grads[i] (caps[i] >= its design voxels), with objective cell gaps[3i..3i+2]; each
 * result equals that engine's own sjf_adjoint_run, and every engine ends in its final forward
 * state.  Every engine needs a design region and step 0, and all share nz and the pole-slot
 * variant (the same materials with poles, or none); SJF_E_ARG otherwise.
 * New (the batched design-variant gradients). */
int sjf_batch_adjoint_run(sjf_batch* b, long nsteps, int checkpoint_every, const int* gaps, double* J,
                          double* const* grads, const size_t* caps);
/* The fused reverse launch of the last sjf_batch_adjoint_run, as sjf_adjoint_kernel_info. */
int sjf_batch_adjoint_kernel_info(const sjf_batch* b, double* ms_mean, int* threads, int* groups);
const char* sjf_batch_last_error(const sjf_batch* b);
int sjf_batch_destroy(sjf_batch* b);

/* Achievable HBM bandwidth: times `iters` launches of a plain copy of `bytes` bytes (one 16-byte
 * load and store per thread) on `device` and reports (read + write bytes) / time in GB/s.  New;
 * used as the measured bandwidth ceiling next to the nominal peak and to calibrate traffic
 * counters. */
int sjf_copy_probe(int device, size_t bytes, int iters, double* gbs);

#ifdef __cplusplus
}
#endif

#endif /* SJF_H */
