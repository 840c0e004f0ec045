// The reference's plugin API, restated with identical layouts, for the HIP
// implementations in libff_amd.so.
//
// Every struct below has the layout of the reference type it names, so a
// caller of the reference library can pass its own objects (scene_t,
// task_t, mesh_t ...) unchanged. The method tables have the reference's
// member order and signatures; implementations are malloc-allocated and
// released by shutdown() + free(), as the reference's system_free does
// (ray_caster/system.cpp:97-101).
#pragma once

namespace math {
typedef float point_t;
struct vec3 { point_t x, y, z; };                 // math/types.h:36-41
struct triangle_t { vec3 points[3]; };            // math/types.h:51-54
typedef triangle_t face_t;
struct ray_t { vec3 origin, direction; };         // math/types.h:72-76 (direction = end point)
}  // namespace math

namespace ray_caster {
typedef math::triangle_t face_t;
typedef math::ray_t ray_t;
struct task_t {                                   // ray_caster/system.h:37-43
  int n_tasks;
  math::ray_t* ray;
  face_t** hit_face;
  math::vec3* hit_point;
};
struct scene_t { int n_faces; face_t* faces; };   // ray_caster/system.h:46-50
struct system_t { const struct system_methods_t* methods; };
struct system_methods_t {                         // ray_caster/system.h:85-104
  int (*init)(system_t* system);
  int (*shutdown)(system_t* system);
  int (*set_scene)(system_t* system, scene_t* scene);
  int (*prepare)(system_t* system);
  int (*cast)(system_t* system, task_t* task);
};
}  // namespace ray_caster

#define RAY_CASTER_OK 0
#define RAY_CASTER_ERROR 1
#define RAY_CASTER_NOT_SUPPORTED 2
// Factory slots: the CUDA casters' IDs (ray_caster/system.h:106-109) are
// taken by the HIP casters.
#define RAY_CASTER_NAIVE_CPU 1
#define RAY_CASTER_NAIVE_HIP 2
#define RAY_CASTER_ZGRID_CPU 3
#define RAY_CASTER_GRID_HIP 4

namespace emission {
struct task_t {                                   // emission/system.h:36-56
  task_t(int n) : n_rays(n), rays(0) {}
  int n_rays;
  ray_caster::task_t* rays;
};
struct system_t { const struct system_methods_t* methods; };
struct system_methods_t {                         // emission/system.h:77-90
  int (*init)(system_t* system, ray_caster::system_t* ray_caster);
  int (*shutdown)(system_t* system);
  int (*set_scene)(system_t* system, ray_caster::scene_t* scene);
  int (*calculate)(system_t* system, task_t* task);
};
}  // namespace emission

#define EMISSION_OK 0
#define EMISSION_ERROR 200
#define EMISSION_MALLEY_CPU 1
#define EMISSION_PARALLEL_RAYS_CPU 2
#define EMISSION_MALLEY_HIP 3
#define EMISSION_PARALLEL_RAYS_HIP 4

namespace malley_emission {
struct task_t : emission::task_t {                // emission/malley_emission.h:27-45
  task_t(int n, float total, float* ws) : emission::task_t(n), total_weight(total), weights(ws) {}
  float total_weight;
  float* weights;
};
}  // namespace malley_emission

namespace parallel_rays_emission {
struct task_t : emission::task_t {                // emission/parallel_rays_cpu.h:27-34
  task_t(int n) : emission::task_t(n) {}
  math::vec3 direction;
  math::vec3 center;
  float distance;
  float radius;
};
}  // namespace parallel_rays_emission

namespace form_factors {
typedef math::triangle_t face_t;
struct mesh_t { int first_idx, n_faces, padding0; };                      // form_factors/system.h:33-38
struct scene_t { int n_faces; face_t* faces; int n_meshes; mesh_t* meshes; };  // :41-47
struct task_t { int n_rays; float* form_factors; };                       // :61-68
struct system_t { const struct system_methods_t* methods; };
struct system_methods_t {                                                 // :95-117
  int (*init)(system_t* system, emission::system_t* emitter);
  int (*shutdown)(system_t* system);
  int (*set_scene)(system_t* system, scene_t* scene);
  int (*prepare)(system_t* system);
  int (*calculate)(system_t* system, task_t* task);
};
}  // namespace form_factors

#define FORM_FACTORS_OK 0
#define FORM_FACTORS_ERROR 100
#define FORM_FACTORS_EMPTY_SCENE 101
#define FORM_FACTORS_SCENE_TOO_SMALL 102
#define FORM_FACTORS_CPU 1
#define FORM_FACTORS_HIP 2

namespace subject {
typedef math::triangle_t face_t;
struct shell_properties_t { float density, heat_capacity, thermal_conductivity, thickness; };  // subject/system.h:31-37
struct optical_properties_t {                                             // :41-48
  float specular_reflectance, diffuse_reflectance, absorbance, transmittance, emissivity;
};
struct material_t { shell_properties_t shell; optical_properties_t front, rear; char name[128]; };  // :52-58
struct mesh_t { int first_idx, n_faces, material_idx; };                  // :64-69
struct scene_t {                                                          // :77-85
  int n_faces; face_t* faces; int n_meshes; mesh_t* meshes; int n_materials; material_t* materials;
};
}  // namespace subject

namespace thermal_equation {
struct task_t { const float* temperatures; float* emission; float* absorption; };  // thermal_equation/system.h:28-33
struct system_t { const struct system_methods_t* methods; };
struct system_methods_t {                                                 // :56-70
  int (*init)(system_t* system, void* params);
  int (*shutdown)(system_t* system);
  int (*set_scene)(system_t* system, subject::scene_t* scene);
  int (*calculate)(system_t* system, task_t* task);
};
}  // namespace thermal_equation

#define THERMAL_EQUATION_OK 0
#define THERMAL_EQUATION_ERROR 400
#define THERMAL_EQUATION_SB_FF_CPU 1
#define THERMAL_EQUATION_RADIANCE_CPU 2
#define THERMAL_EQUATION_CONDUCTIVE_CPU 3
#define THERMAL_EQUATION_HEAT_SOURCE_CPU 4
#define THERMAL_EQUATION_DISNTANT_SOURCE_CPU 5
#define THERMAL_EQUATION_SB_FF_HIP 6
#define THERMAL_EQUATION_RADIANCE_HIP 7
#define THERMAL_EQUATION_CONDUCTIVE_HIP 8
#define THERMAL_EQUATION_HEAT_SOURCE_HIP 9
#define THERMAL_EQUATION_DISTANT_SOURCE_HIP 10

namespace heat_source_equation {                                          // thermal_equation/heat_source_cpu.h:28-41
struct heat_source_t { int mesh_idx; float power; };
struct params_t { int n_sources; heat_source_t* sources; };
}
namespace parallel_rays_cpu {                                             // thermal_equation/parallel_rays_cpu.h:26-39
struct source_t { float power; math::vec3 direction; };
struct params_t { int n_rays; emission::system_t* emitter; void* source_param; source_t (*source)(void* param); };
}

namespace form_factors_equation {
struct params_t { form_factors::system_t* form_factors_calculator; int n_rays; };  // thermal_equation/form_factors.h
}
namespace radiance_equation {
struct params_t { emission::system_t* emitter; int n_rays; };             // thermal_equation/radiance_cpu.h
}
// SB_FF_HIP parameters: the reference's, plus the storage type of F on the
// device (FFK_F32 / FFK_BF16 / FFK_F16 of ffk.h).
namespace hip_sb_ff {
struct params_t { form_factors::system_t* form_factors_calculator; int n_rays; int f_dtype; };
}

namespace thermal_solution {
struct params_t { int n_equations; thermal_equation::system_t** equations; };  // thermal_solution/system.h:32-36
struct task_t { int n_step; float time_delta; float* temperatures; };         // :42-47
struct system_t { const struct system_methods_t* methods; };
struct system_methods_t {                                                     // :72-89
  int (*init)(system_t* system, params_t* params);
  int (*shutdown)(system_t* system);
  int (*set_scene)(system_t* system, subject::scene_t* scene, float* temperatures);
  int (*calculate)(system_t* system, task_t* task);
};
}  // namespace thermal_solution

#define THERMAL_SOLUTION_OK 0
#define THERMAL_SOLUTION_ERROR 300
#define THERMAL_SOLUTION_CPU_ADAMS 1
#define THERMAL_SOLUTION_HIP_ADAMS 2

// ---- the HIP implementations (malloc'd, methods set, NOT yet initialised) --
namespace hip_ray_caster {
ray_caster::system_t* system_create_naive();  // slot RAY_CASTER_NAIVE_HIP
ray_caster::system_t* system_create_grid();   // slot RAY_CASTER_GRID_HIP
}
namespace hip_malley_emission { emission::system_t* system_create(); }
namespace hip_parallel_rays_emission { emission::system_t* system_create(); }
namespace hip_form_factors { form_factors::system_t* system_create(); }
namespace hip_sb_ff { thermal_equation::system_t* system_create(); }
namespace hip_radiance { thermal_equation::system_t* system_create(); }
namespace hip_conduction { thermal_equation::system_t* system_create(); }     // params: none
namespace hip_heat_source { thermal_equation::system_t* system_create(); }    // heat_source_equation::params_t
namespace hip_distant_source { thermal_equation::system_t* system_create(); } // parallel_rays_cpu::params_t
namespace hip_adams { thermal_solution::system_t* system_create(); }
