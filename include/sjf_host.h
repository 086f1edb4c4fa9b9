/* sjf_host.h -- C entry points of the host layer (libsjf_host.so).
 *
 * The C++ API is sjf_host.hpp; these wrappers expose the same pieces with plain pointers so
 * that other languages (and the test suite, through ctypes) can drive them.  Scene arguments
 * are paths of scene JSON files (the form sjfh::Scene::from_json_file reads).  Every
 * function returns 0 on success and a negative code otherwise; sjfh_last_error() describes
 * the last failure of the calling thread.
 */
#ifndef SJF_HOST_H
#define SJF_HOST_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* sjfh_last_error(void);

/* Settings as the reference driver reads them, written as a JSON object to out.  mode 0:
 * parse_conf_file(conf_path) alone; 1: parse_args over argv (argv[0] is the program name)
 * then parse_conf_file of its --conf-file argument (or conf_path); 2: as 1 plus a second
 * correct_defaults, the sequence of the reference's main() (src/main.cpp:16-36). */
int sjfh_settings_json(const char* conf_path, int argc, const char* const* argv, int mode, char* out, size_t cap);

/* in() of every root of the scene at n points (xyz triples); out[p * n_roots + r]. */
int sjfh_scene_in(const char* scene_json, const double* xyz, long n, int* out, int* n_roots);

/* cgs_material_function::in_bound with every root as a region scaled by its "eps" metadata
 * (default def_ret), smoothing from (smooth_n, smooth_rad) (src/disp.cpp:56-283). */
int sjfh_in_bound(const char* scene_json, double def_ret, unsigned smooth_n, double smooth_rad,
                  const double* xyz, long n, double* out);

/* The smoothing offsets generated for (smooth_n, smooth_rad): 8 * smooth_n points. */
int sjfh_smooth_points(unsigned smooth_n, double smooth_rad, double* out, size_t cap, unsigned* n_points);

/* parse_susceptibilities of root r: out[4q..4q+3] = {omega_0, gamma, sigma, use_denom}. */
int sjfh_susceptibilities(const char* scene_json, int root, double* out, int cap, int* n_poles, int* er);

/* source_info of source s: out = {gaussian, component, wavelen, width, phase, start_time,
 * end_time, amplitude}; pulse (optional, 7 doubles) = the sjf_pulse fields after the unit
 * conversion with um_scale: {kind, freq, width, phase, t_start, t_end, slowness}. */
int sjfh_source_info(const char* scene_json, int s, double um_scale, double* out, double* pulse);

/* The reference's fft of n complex samples (re, im pairs): *n_out = 2^floor(log2 n) bins. */
int sjfh_fft(const double* in, long n, double* out, long cap, long* n_out);

/* voxelize on an n^3 grid of cell size dx: words (n^3) and up to cap_mat materials (each 26
 * doubles: eps_inf, npoles, f0[4], gamma[4], sigma[4], is_drude[4], 4 unused).  With smooth_n > 0
 * *has_map = 1 and eps_map (3 n^3 doubles, [Ex Ey Ez][cell], may be NULL) receives the per-cell
 * eps_inf for sjf_set_eps_maps; the materials then carry the poles only. */
int sjfh_voxelize(const char* scene_json, double ambient_eps, unsigned smooth_n, double smooth_rad,
                  double um_scale, double resolution, int n, double dx, unsigned* words, double* mats,
                  int cap_mat, int* n_mat, double* eps_map, int* has_map);

/* Runs the .geom scene script geom_path (the reference's scene language: scene::read_file,
 * src/cgs.cpp:510-570) in the context context_from_settings builds (src/disp.cpp:28-50) from
 * the given settings values and the user_opts statements ("a=1;b=2"), and writes the scene in
 * this layer's JSON form (roots, sources, monitor_clusters, context_numbers) to out.
 * *parse_error = the script's parse_ercode (0 = success; on an error no roots are built). */
int sjfh_geom_scene_json(const char* geom_path, double pml_thickness, double len, double um_scale, const char* out_dir,
                         const char* user_opts, char* out, size_t cap, int* parse_error);

/* Whole driver run (BoundGeom): settings from conf_path + argv as sjfh_settings_json with
 * mode 2, scene from scene_json (a scene JSON, or a .geom script run with those settings),
 * then run() and save_field_times() into the settings' out_dir.  precision: 4 or 8; complex_fields: two engines (Re, Im).  summary =
 * {grid_num, dt, ttot, n_t_pts, n_monitors, n_clusters} (optional). */
int sjfh_run(const char* conf_path, int argc, const char* const* argv, const char* scene_json, int precision,
             int complex_fields, int device, double* summary);

/* field_samples.h5 for externally produced monitor series (the writer BoundGeom uses): sources,
 * monitor clusters and cgs_params from the scene, ttot in meep time units, series[point]
 * [sample][re, im] with n_samples per point. */
int sjfh_write_field_samples(const char* path, const char* scene_json, double um_scale, double ttot,
                             unsigned n_t_pts, unsigned save_span, const double* series, long n_samples);

/* Reads dataset dset of an HDF5 file as doubles: compound datasets by member names
 * (members[0..nmembers), packed per element), plain numeric ones with nmembers = 0.
 * *count = elements in the dataset; nothing is read when count * max(nmembers, 1) > cap. */
int sjfh_h5_read(const char* path, const char* dset, const char* const* members, int nmembers, double* out,
                 size_t cap, size_t* count);
/* Names of the links in a group, newline separated. */
int sjfh_h5_list(const char* path, const char* group, char* out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* SJF_HOST_H */
