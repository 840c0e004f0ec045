/*
 * ff_plugin -- flat C binding of the HIP plugin systems (the reference's
 * system_t method tables) for FFI callers (ctypes, cgo, JNI ...).
 *
 * The structs mirror the reference's types byte for byte (see
 * include/form_factors_amd/api.hpp for the C++ view and the reference
 * file:line of each); systems are opaque pointers to the reference-shaped
 * system_t. Return codes are the reference's: 0 on success, a negative
 * RAY_CASTER_/EMISSION_/FORM_FACTORS_/THERMAL_* code on failure.
 */
#ifndef FF_PLUGIN_H
#define FF_PLUGIN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } ffp_vec3;                         /* math/types.h:36-41 */
typedef struct { ffp_vec3 points[3]; } ffp_face;                    /* math/types.h:51-54 */
typedef struct { ffp_vec3 origin, direction; } ffp_ray;             /* math/types.h:72-76 */
typedef struct {                                                    /* ray_caster/system.h:37-43 */
  int n_tasks;
  ffp_ray* ray;
  ffp_face** hit_face;
  ffp_vec3* hit_point;
} ffp_rc_task;
typedef struct { int n_faces; ffp_face* faces; } ffp_rc_scene;      /* ray_caster/system.h:46-50 */
typedef struct { int first_idx, n_faces, padding0; } ffp_mesh;      /* form_factors/system.h:33-38 */
typedef struct {                                                    /* form_factors/system.h:41-47 */
  int n_faces;
  ffp_face* faces;
  int n_meshes;
  ffp_mesh* meshes;
} ffp_ff_scene;
typedef struct { int n_rays; float* form_factors; } ffp_ff_task;    /* form_factors/system.h:61-68 */
typedef struct {                                                    /* subject/system.h:31-58 */
  float density, heat_capacity, thermal_conductivity, thickness;
  float front[5]; /* specular, diffuse, absorbance, transmittance, emissivity */
  float rear[5];
  char name[128];
} ffp_material;
typedef struct { int first_idx, n_faces, material_idx; } ffp_subject_mesh;  /* subject/system.h:64-69 */
typedef struct {                                                    /* subject/system.h:77-85 */
  int n_faces;
  ffp_face* faces;
  int n_meshes;
  ffp_subject_mesh* meshes;
  int n_materials;
  ffp_material* materials;
} ffp_subject_scene;
typedef struct { const float* temperatures; float* emission; float* absorption; } ffp_te_task;  /* thermal_equation/system.h:28-33 */
typedef struct { int n_step; float time_delta; float* temperatures; } ffp_ts_task;  /* thermal_solution/system.h:42-47 */

/* ---- ray caster: ray_caster::system_create / system_* (ray_caster/system.h:116-131).
 * type 2 = naive HIP, 4 = grid HIP (the CUDA casters' slots); NULL for others. */
void* ffp_ray_caster_create(int type);
int ffp_ray_caster_set_scene(void* system, ffp_rc_scene* scene);
int ffp_ray_caster_prepare(void* system);
int ffp_ray_caster_cast(void* system, ffp_rc_task* task);
void ffp_ray_caster_free(void* system);

/* ---- emission: emission::system_create / system_* (emission/system.h:102-111).
 * type 3 = Malley HIP, 4 = parallel rays HIP. The caster may be any
 * reference-shaped ray caster (ours or the reference's CPU casters). */
void* ffp_emission_create(int type, void* ray_caster);
int ffp_emission_set_scene(void* system, ffp_rc_scene* scene);
/* malley_emission::task_t{n_rays, rays, total_weight, weights} (emission/malley_emission.h:27-45);
 * *rays_out receives the ray caster task the emitter allocated (free with ffp_task_free). */
int ffp_emission_calculate_malley(void* system, int n_rays, float total_weight, float* weights,
                                  ffp_rc_task** rays_out);
/* parallel_rays_emission_cpu::task_t (emission/parallel_rays_cpu.h:27-34). */
int ffp_emission_calculate_parallel(void* system, int n_rays, const float* direction3,
                                    const float* center3, float distance, float radius,
                                    ffp_rc_task** rays_out);
/* Counter-RNG key and emission frame of a HIP emitter (default 0 and the
 * reference's unnormalised frame). Call k of calculate() uses the key
 * seed + k * 0x9E3779B97F4A7C15. Types 3 and 4 (the frame is ignored for
 * type 4); others: -EMISSION_ERROR. */
int ffp_emission_set_seed(void* system, unsigned long long seed, int frame);
void ffp_emission_free(void* system);
void ffp_task_free(ffp_rc_task* task);

/* ---- form factors: form_factors::system_create / system_* (form_factors/system.h:125-139).
 * type 2 = HIP (fused emit/cast/tally when the emitter is Malley HIP over
 * the grid HIP caster). */
void* ffp_form_factors_create(int type, void* emitter);
int ffp_form_factors_set_scene(void* system, ffp_ff_scene* scene);
int ffp_form_factors_prepare(void* system);
int ffp_form_factors_calculate(void* system, ffp_ff_task* task);
void ffp_form_factors_free(void* system);

/* ---- thermal equations (thermal_equation/system.h:56-92).
 * type 6 = SB_FF HIP: params = {form_factors system, n_rays, f_dtype}
 * type 7 = radiance HIP: params = {emitter, n_rays}
 * type 8 = conductive HIP: no parameters (dep, n_rays, f_dtype unused) */
void* ffp_thermal_equation_create(int type, void* form_factors_or_emitter, int n_rays, int f_dtype);
/* type 9 = heat sources (thermal_equation/heat_source_cpu.h:28-41); the
 * caller keeps `sources` alive, as with the reference's params. */
typedef struct ffp_heat_source {
  int mesh_idx;
  float power; /* negative: a heat sink */
} ffp_heat_source;
void* ffp_heat_source_create(int n_sources, ffp_heat_source* sources);
/* type 10 = distant source (thermal_equation/parallel_rays_cpu.h:26-39) over
 * a parallel-rays emitter (fused on the device for the HIP emitter over the
 * grid caster); the caller keeps `source` alive and may change it between
 * calls (the reference reads it through its source callback each step).
 * source NULL: no distant source, the equation adds nothing. */
typedef struct ffp_source {
  float power; /* W/m^2 */
  ffp_vec3 direction;
} ffp_source;
void* ffp_distant_source_create(void* emitter, int n_rays, ffp_source* source);
int ffp_thermal_equation_set_scene(void* system, ffp_subject_scene* scene);
int ffp_thermal_equation_calculate(void* system, ffp_te_task* task);
void ffp_thermal_equation_free(void* system);

/* ---- thermal solution (thermal_solution/system.h:72-106): type 2 = Adams HIP (FP64). */
void* ffp_thermal_solution_create(int type, int n_equations, void** equations);
int ffp_thermal_solution_set_scene(void* system, ffp_subject_scene* scene, float* temperatures);
int ffp_thermal_solution_calculate(void* system, ffp_ts_task* task);
/* No counterpart in the reference: where every equation of the solution is one
 * of this library's SB_FF, radiance (HIP Malley emitter over the HIP grid
 * caster), conduction, heat-source or distant-source (HIP parallel-rays emitter
 * over the grid caster) systems, a step keeps its emission and absorption in device
 * memory from the equations to the FP64 update and copies only the new
 * temperatures back (on = 1, the default); on = 0 runs the equations on host
 * arrays, as the reference's solutions do. Both give the same temperatures
 * bit for bit. For callers comparing the two or bounding device work. */
int ffp_thermal_solution_set_device_loop(void* system, int on);
void ffp_thermal_solution_free(void* system);

#ifdef __cplusplus
}
#endif
#endif /* FF_PLUGIN_H */
