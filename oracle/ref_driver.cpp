/*
 * Driver over the COMPILED REFERENCE -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is this project's own code. It is compiled together with the
 * reference's CPU sources (read in place under the reference tree by
 * oracle/Makefile, target `ref`) into oracle/_ref/libffref.so, which is
 * never committed. It exposes flat extern "C" entry points over the
 * reference's public C++ API so that tests/golden/make_golden.py can write
 * golden vectors and bench.py can time the reference CPU caster. Nothing in
 * the product library links or loads it.
 */
#include "cpuFactorsCalculator/cpu_system.h"
#include "emission/malley_cpu.h"
#include "emission/malley_emission.h"
#include "emission/parallel_rays_cpu.h"
#include "emission/system.h"
#include "import_export/csv_export.h"
#include "import_export/obj_export.h"
#include "import_export/obj_import.h"
#include "form_factors/system.h"
#include "math/grid.h"
#include "math/mat.h"
#include "math/operations.h"
#include "math/ray.h"
#include "math/triangle.h"
#include "ray_caster/naive_cpu.h"
#include "ray_caster/system.h"
#include "ray_caster/zgrid_cpu.h"
#include "subject/generator.h"
#include "subject/system.h"
#include "thermal_equation/conductive_cpu.h"
#include "thermal_equation/form_factors.h"
#include "thermal_equation/heat_source_cpu.h"
#include "thermal_equation/parallel_rays_cpu.h"
#include "thermal_equation/radiance_cpu.h"
#include "thermal_equation/system.h"
#include "thermal_solution/cpu_adams.h"
#include "thermal_solution/system.h"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace {

ray_caster::system_t* make_caster(int kind) {
  // The factory switch (ray_caster/system.cpp:65-90) also names the CUDA
  // casters; the CPU constructors are called directly instead.
  ray_caster::system_t* s = kind == 0 ? raycaster_naive_cpu::system_create()
                                      : ray_caster_zgrid_cpu::system_create();
  ray_caster::system_init(s);
  return s;
}

void free_caster(ray_caster::system_t* s) {
  ray_caster::system_shutdown(s);
  free(s);
}

struct cell_sink {
  std::vector<int>* out;
  int max;
};

bool collect(math::grid_coord_t p, cell_sink* s) {
  if ((int)s->out->size() / 2 < s->max) {
    s->out->push_back(p.x);
    s->out->push_back(p.y);
  }
  return (int)s->out->size() / 2 >= s->max;
}

}  // namespace

EXPORT int ffref_triangle_intersect(const float* ray6, const float* tri9, float* point3) {
  math::ray_t r;
  math::triangle_t t;
  memcpy(&r, ray6, sizeof(r));
  memcpy(&t, tri9, sizeof(t));
  math::vec3 p = {0, 0, 0};
  int code = math::triangle_intersect(r, t, &p);
  if (code == TRIANGLE_INTERSECTION_UNIQUE && point3) memcpy(point3, &p, sizeof(p));
  return code;
}

EXPORT float ffref_triangle_area(const float* tri9) {
  math::triangle_t t;
  memcpy(&t, tri9, sizeof(t));
  return math::triangle_area(t);
}

EXPORT void ffref_rotate_towards(const float* from3, const float* to3, float* m9) {
  math::vec3 a, b;
  memcpy(&a, from3, sizeof(a));
  memcpy(&b, to3, sizeof(b));
  math::mat33 m = math::rotate_towards(a, b);
  memcpy(m9, &m, sizeof(m));
}

EXPORT void ffref_mat_vec(const float* m9, const float* v3, float* out3) {
  math::mat33 m;
  math::vec3 v;
  memcpy(&m, m9, sizeof(m));
  memcpy(&v, v3, sizeof(v));
  math::vec3 r = m * v;
  memcpy(out3, &r, sizeof(r));
}

EXPORT void ffref_axis_rotation(float x, float y, float z, float* m9) {
  math::mat33 m = math::axis_rotation(x, y, z);
  memcpy(m9, &m, sizeof(m));
}

EXPORT int ffref_ray_intersect_segment(const float* p6, const float* q6, float* s) {
  math::ray_t p, q;
  memcpy(&p, p6, sizeof(p));
  memcpy(&q, q6, sizeof(q));
  return math::ray_intersect_segment(p, q, *s) ? 1 : 0;
}

EXPORT int ffref_grid_traverse(const float* base3, const float* size3, int nx, int ny,
                               const float* ray6, int is_segment, int* cells, int max_cells) {
  math::vec3 base, size;
  math::ray_t r;
  memcpy(&base, base3, sizeof(base));
  memcpy(&size, size3, sizeof(size));
  memcpy(&r, ray6, sizeof(r));
  math::grid_2d_t g = math::make_grid(base, size, nx, ny);
  std::vector<int> out;
  cell_sink sink = {&out, max_cells};
  if (is_segment)
    math::grid_put(&g, r, (math::grid_traversal_callback)&collect, &sink);
  else
    math::grid_traverse(&g, r, (math::grid_traversal_callback)&collect, &sink);
  if (!out.empty()) memcpy(cells, out.data(), out.size() * sizeof(int));
  return (int)out.size() / 2;
}

EXPORT int ffref_grid_rasterize(const float* base3, const float* size3, int nx, int ny,
                                const float* tri9, int* cells, int max_cells) {
  math::vec3 base, size;
  math::triangle_t t;
  memcpy(&base, base3, sizeof(base));
  memcpy(&size, size3, sizeof(size));
  memcpy(&t, tri9, sizeof(t));
  math::grid_2d_t g = math::make_grid(base, size, nx, ny);
  std::vector<int> out;
  cell_sink sink = {&out, max_cells};
  math::grid_rasterize(&g, t, (math::grid_traversal_callback)&collect, &sink);
  if (!out.empty()) memcpy(cells, out.data(), out.size() * sizeof(int));
  return (int)out.size() / 2;
}

/* zgrid structure as ray_caster/zgrid_cpu.cpp:57-68 builds it. */
EXPORT int ffref_zgrid_info(const float* faces, int n_faces, int* nx, int* ny, float* base3,
                            float* side3, float* size3, int* counts, int* offsets, int* refs,
                            int max_refs) {
  const math::triangle_t* t = (const math::triangle_t*)faces;
  math::triangles_analysis_t stat = math::triangles_analyze(t, n_faces);
  math::grid_2d_t g = math::grid_deduce_optimal(stat);
  *nx = g.n_x;
  *ny = g.n_y;
  memcpy(base3, &g.base, 12);
  memcpy(side3, &g.side, 12);
  memcpy(size3, &g.size, 12);
  if (!counts) return 0;
  math::grid_2d_index_t* index = math::grid_make_index(&g);
  math::grid_index_triangles(&g, index, t, n_faces);
  for (int i = 0; i < g.n_x * g.n_y; ++i) {
    counts[i] = index->cells[i].count;
    offsets[i] = index->cells[i].offset;
  }
  int n = index->n_triangles;
  if (n <= max_refs && n > 0) memcpy(refs, index->triangles, n * sizeof(int));
  math::grid_free_index(index);
  return n;
}

/* kind 0 = naive_cpu, 1 = zgrid_cpu. Rays are cast on n_threads disjoint
 * slices concurrently through one system, as the CPU casters allow. */
EXPORT int ffref_cast(int kind, const float* faces, int n_faces, const float* rays, long n_rays,
                      int* hit_idx, float* hit_pts, int n_threads) {
  ray_caster::system_t* s = make_caster(kind);
  ray_caster::scene_t scene = {n_faces, (ray_caster::face_t*)faces};
  int r = ray_caster::system_set_scene(s, &scene);
  if (r == 0) r = ray_caster::system_prepare(s);
  if (r < 0) {
    free_caster(s);
    return r;
  }
  std::vector<ray_caster::face_t*> hit(n_rays);
  if (n_threads < 1) n_threads = 1;
  long chunk = (n_rays + n_threads - 1) / n_threads;
  std::vector<std::thread> th;
  std::vector<int> rc(n_threads, 0);
  for (int i = 0; i < n_threads; ++i) {
    long lo = i * chunk;
    long hi = lo + chunk < n_rays ? lo + chunk : n_rays;
    if (lo >= hi) break;
    th.emplace_back([=, &hit, &rc] {
      ray_caster::task_t task = {(int)(hi - lo), (math::ray_t*)(rays + 6 * lo), &hit[lo],
                                 (math::vec3*)(hit_pts + 3 * lo)};
      rc[i] = ray_caster::system_cast(s, &task);
    });
  }
  for (auto& t : th) t.join();
  for (long i = 0; i < n_rays; ++i) hit_idx[i] = hit[i] ? (int)(hit[i] - scene.faces) : -1;
  free_caster(s);
  for (int v : rc)
    if (v < 0) return v;
  return 0;
}

EXPORT void ffref_emitted_counts(const float* weights, int n_faces, int n_rays, float total,
                                 int* front, int* rear) {
  malley_emission::task_t task(n_rays, total, (float*)weights);
  for (int f = 0; f < n_faces; ++f) {
    front[f] = malley_emission::emitted_front(&task, f);
    rear[f] = malley_emission::emitted_rear(&task, f);
  }
}

/* Malley rays exactly as emission/malley_cpu.cpp generates them on a fresh
 * system (mt19937 seeds 1..4), together with the naive caster's hits for
 * them. Returns the number of rays (<= max_rays are copied). */
EXPORT long ffref_malley(const float* faces, int n_faces, const float* weights, float total,
                         int n_rays, float* rays_out, int* hit_idx, long max_rays) {
  ray_caster::system_t* rc = make_caster(0);
  emission::system_t* em = emission::system_create(EMISSION_MALLEY_CPU, rc);
  ray_caster::scene_t scene = {n_faces, (ray_caster::face_t*)faces};
  long n = -1;
  if (emission::system_set_scene(em, &scene) == 0) {
    malley_emission::task_t task(n_rays, total, (float*)weights);
    if (emission::system_calculate(em, &task) == 0 && task.rays) {
      n = task.rays->n_tasks;
      long m = n < max_rays ? n : max_rays;
      memcpy(rays_out, task.rays->ray, m * sizeof(math::ray_t));
      for (long i = 0; i < m; ++i)
        hit_idx[i] = task.rays->hit_face[i] ? (int)(task.rays->hit_face[i] - scene.faces) : -1;
      ray_caster::task_free(task.rays);
    } else {
      n = 0;
    }
  }
  emission::system_free(em);
  free_caster(rc);
  return n;
}

EXPORT void ffref_parallel_rays(const float* dir3, const float* center3, float distance,
                                float radius, int n_rays, const float* faces, int n_faces,
                                float* rays_out) {
  ray_caster::system_t* rc = make_caster(0);
  emission::system_t* em = emission::system_create(EMISSION_PARALLEL_RAYS_CPU, rc);
  ray_caster::scene_t scene = {n_faces, (ray_caster::face_t*)faces};
  emission::system_set_scene(em, &scene);
  // The task type has no default constructor; the reference keeps it as a
  // member of a malloc'd struct (thermal_equation/parallel_rays_cpu.cpp:40)
  // and fills its fields, which is what is done here.
  alignas(parallel_rays_emission_cpu::task_t) char buf[sizeof(parallel_rays_emission_cpu::task_t)];
  parallel_rays_emission_cpu::task_t& task = *reinterpret_cast<parallel_rays_emission_cpu::task_t*>(buf);
  task.n_rays = n_rays;
  task.rays = 0;
  memcpy(&task.direction, dir3, 12);
  memcpy(&task.center, center3, 12);
  task.distance = distance;
  task.radius = radius;
  emission::system_calculate(em, &task);
  memcpy(rays_out, task.rays->ray, (size_t)n_rays * sizeof(math::ray_t));
  ray_caster::task_free(task.rays);
  emission::system_free(em);
  free_caster(rc);
}

/* Whole reference FF pipeline: FORM_FACTORS_CPU over MALLEY_CPU over the
 * chosen CPU caster. meshes are (first_idx, n_faces, pad) triples. */
EXPORT int ffref_ff_calculate(const float* faces, int n_faces, const int* meshes, int n_meshes,
                              int n_rays, int caster, float* F) {
  ray_caster::system_t* rc = make_caster(caster);
  emission::system_t* em = emission::system_create(EMISSION_MALLEY_CPU, rc);
  form_factors::system_t* ff = form_factors::system_create(FORM_FACTORS_CPU, em);
  form_factors::scene_t scene = {n_faces, (form_factors::face_t*)faces, n_meshes,
                                 (form_factors::mesh_t*)meshes};
  int r = form_factors::system_set_scene(ff, &scene);
  if (r == 0) r = form_factors::system_prepare(ff);
  if (r == 0) {
    form_factors::task_t task = {n_rays, F};
    r = form_factors::system_calculate(ff, &task);
  }
  form_factors::system_free(ff);
  emission::system_free(em);
  free_caster(rc);
  return r;
}

namespace {
/* Timing plugins: a caster and an emitter that forward every method to the
 * reference's own system and add the wall time of cast() / calculate(). The
 * form-factor calculator calls the emitter, which calls the caster, so the
 * time of one FORM_FACTORS_CPU calculate() splits into ray generation
 * (emitter minus caster), casting, and the tally and normalisation
 * (calculator minus emitter). */
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
struct timed_caster_t : ray_caster::system_t {
  ray_caster::system_t* inner;
  double seconds;
};
int tc_init(timed_caster_t*) { return 0; }
int tc_shutdown(timed_caster_t*) { return 0; }
int tc_set_scene(timed_caster_t* s, ray_caster::scene_t* sc) { return ray_caster::system_set_scene(s->inner, sc); }
int tc_prepare(timed_caster_t* s) { return ray_caster::system_prepare(s->inner); }
int tc_cast(timed_caster_t* s, ray_caster::task_t* t) {
  const double t0 = now_s();
  int r = ray_caster::system_cast(s->inner, t);
  s->seconds += now_s() - t0;
  return r;
}
const ray_caster::system_methods_t tc_methods = {
    (int (*)(ray_caster::system_t*))&tc_init,
    (int (*)(ray_caster::system_t*))&tc_shutdown,
    (int (*)(ray_caster::system_t*, ray_caster::scene_t*))&tc_set_scene,
    (int (*)(ray_caster::system_t*))&tc_prepare,
    (int (*)(ray_caster::system_t*, ray_caster::task_t*))&tc_cast,
};
struct timed_emitter_t : emission::system_t {
  emission::system_t* inner;
  double seconds;
};
int te_init(timed_emitter_t*, ray_caster::system_t*) { return 0; }
int te_shutdown(timed_emitter_t*) { return 0; }
int te_set_scene(timed_emitter_t* s, ray_caster::scene_t* sc) { return emission::system_set_scene(s->inner, sc); }
int te_calculate(timed_emitter_t* s, emission::task_t* t) {
  const double t0 = now_s();
  int r = emission::system_calculate(s->inner, t);
  s->seconds += now_s() - t0;
  return r;
}
const emission::system_methods_t te_methods = {
    (int (*)(emission::system_t*, ray_caster::system_t*))&te_init,
    (int (*)(emission::system_t*))&te_shutdown,
    (int (*)(emission::system_t*, ray_caster::scene_t*))&te_set_scene,
    (int (*)(emission::system_t*, emission::task_t*))&te_calculate,
};
}  // namespace

/* One FORM_FACTORS_CPU calculate() (cpuFactorsCalculator/cpu_system.cpp:149-201)
 * over MALLEY_CPU over the chosen CPU caster, single-threaded as the
 * reference runs it, with its time split into ray generation, casting and
 * tally + normalisation (seconds3). */
EXPORT int ffref_ff_timed(const float* faces, int n_faces, const int* meshes, int n_meshes, int n_rays,
                          int caster, float* F, double* seconds3) {
  timed_caster_t tc;
  tc.methods = &tc_methods;
  tc.inner = make_caster(caster);
  tc.seconds = 0;
  emission::system_t* em_inner = emission::system_create(EMISSION_MALLEY_CPU, &tc);
  timed_emitter_t te;
  te.methods = &te_methods;
  te.inner = em_inner;
  te.seconds = 0;
  form_factors::system_t* ff = form_factors::system_create(FORM_FACTORS_CPU, &te);
  form_factors::scene_t scene = {n_faces, (form_factors::face_t*)faces, n_meshes,
                                 (form_factors::mesh_t*)meshes};
  int r = form_factors::system_set_scene(ff, &scene);
  if (r == 0) r = form_factors::system_prepare(ff);
  double total = 0;
  if (r == 0) {
    form_factors::task_t task = {n_rays, F};
    const double t0 = now_s();
    r = form_factors::system_calculate(ff, &task);
    total = now_s() - t0;
  }
  seconds3[0] = te.seconds - tc.seconds;
  seconds3[1] = tc.seconds;
  seconds3[2] = total - te.seconds;
  form_factors::system_free(ff);
  emission::system_free(em_inner);
  free_caster(tc.inner);
  return r;
}

/* Error codes of the reference FF calculator for a given scene. */
EXPORT int ffref_ff_prepare_code(const float* faces, int n_faces, const int* meshes,
                                 int n_meshes) {
  ray_caster::system_t* rc = make_caster(0);
  emission::system_t* em = emission::system_create(EMISSION_MALLEY_CPU, rc);
  form_factors::system_t* ff = form_factors::system_create(FORM_FACTORS_CPU, em);
  form_factors::scene_t scene = {n_faces, (form_factors::face_t*)faces, n_meshes,
                                 (form_factors::mesh_t*)meshes};
  int r = form_factors::system_set_scene(ff, &scene);
  int p = form_factors::system_prepare(ff);
  form_factors::system_free(ff);
  emission::system_free(em);
  free_caster(rc);
  return r < 0 ? r : p;
}

namespace {
/* A constant-rate equation: emission e[m], absorption a[m] each step
 * (the same shape as the mock equation of the reference's Adams tests). */
struct const_eq_t : thermal_equation::system_t {
  const float* e;
  const float* a;
  int n;
};
int ce_init(const_eq_t*, void*) { return 0; }
int ce_shutdown(const_eq_t*) { return 0; }
int ce_set_scene(const_eq_t*, subject::scene_t*) { return 0; }
int ce_calculate(const_eq_t* s, thermal_equation::task_t* task) {
  for (int m = 0; m < s->n; ++m) {
    task->emission[m] += s->e[m];
    task->absorption[m] += s->a[m];
  }
  return 0;
}
const thermal_equation::system_methods_t ce_methods = {
    (int (*)(thermal_equation::system_t*, void*))&ce_init,
    (int (*)(thermal_equation::system_t*))&ce_shutdown,
    (int (*)(thermal_equation::system_t*, subject::scene_t*))&ce_set_scene,
    (int (*)(thermal_equation::system_t*, thermal_equation::task_t*))&ce_calculate,
};
}  // namespace

/* cpu_adams trajectory: n_steps steps of dt, temperatures after each step
 * written to T_out[step * n_meshes + m]. Materials: shell (density,
 * heat_capacity, thickness) per mesh; optical properties are unused here. */
EXPORT int ffref_adams(const float* faces, int n_faces, const int* meshes, int n_meshes,
                       const float* shell3, const float* T0, const float* e, const float* a,
                       float dt, int n_steps, float* T_out) {
  std::vector<subject::material_t> mats(n_meshes);
  std::vector<subject::mesh_t> ms(n_meshes);
  for (int m = 0; m < n_meshes; ++m) {
    mats[m] = subject::black_body();
    mats[m].shell.density = shell3[3 * m];
    mats[m].shell.heat_capacity = shell3[3 * m + 1];
    mats[m].shell.thickness = shell3[3 * m + 2];
    ms[m] = subject::make_mesh(meshes[3 * m], meshes[3 * m + 1], m);
  }
  subject::scene_t scene = {n_faces, (subject::face_t*)faces, n_meshes, ms.data(), n_meshes,
                            mats.data()};
  const_eq_t eq;
  eq.methods = &ce_methods;
  eq.e = e;
  eq.a = a;
  eq.n = n_meshes;
  thermal_equation::system_t* eqp = &eq;
  thermal_solution::params_t params = {1, &eqp};
  thermal_solution::system_t* sol = thermal_solution::system_create(THERMAL_SOLUTION_CPU_ADAMS, &params);
  std::vector<float> T(T0, T0 + n_meshes);
  int r = thermal_solution::system_set_scene(sol, &scene, T.data());
  thermal_solution::task_t* task = thermal_solution::task_create(n_meshes);
  task->time_delta = dt;
  for (int s = 0; s < n_steps && r == 0; ++s) {
    r = thermal_solution::system_calculate(sol, task);
    memcpy(T_out + (size_t)s * n_meshes, task->temperatures, n_meshes * sizeof(float));
  }
  thermal_solution::task_free(task);
  thermal_solution::system_free(sol);
  return r;
}

/* conductive_equation (thermal_equation/conductive_cpu.cpp): one calculate()
 * with per-mesh conductivity and thickness, emission/absorption out. */
EXPORT int ffref_conduction(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* k,
                            const float* thickness, const float* T, float* emission, float* absorption) {
  std::vector<subject::material_t> mats(n_meshes);
  std::vector<subject::mesh_t> ms(n_meshes);
  for (int m = 0; m < n_meshes; ++m) {
    mats[m] = subject::black_body();
    mats[m].shell.thermal_conductivity = k[m];
    mats[m].shell.thickness = thickness[m];
    ms[m] = subject::make_mesh(meshes[3 * m], meshes[3 * m + 1], m);
  }
  subject::scene_t scene = {n_faces, (subject::face_t*)faces, n_meshes, ms.data(), n_meshes, mats.data()};
  thermal_equation::system_t* eq = conductive_equation::system_create();
  conductive_equation::params_t params;
  thermal_equation::system_init(eq, &params);
  int r = thermal_equation::system_set_scene(eq, &scene);
  thermal_equation::task_t task = {T, emission, absorption};
  memset(emission, 0, n_meshes * sizeof(float));
  memset(absorption, 0, n_meshes * sizeof(float));
  if (r == 0) r = thermal_equation::system_calculate(eq, &task);
  thermal_equation::system_free(eq);
  return r;
}

/* radiance_equation (thermal_equation/radiance_cpu.cpp) over MALLEY_CPU and
 * the naive caster on a fresh system: one calculate() call. Materials carry
 * front/rear emissivity per mesh. */
EXPORT int ffref_radiance(const float* faces, int n_faces, const int* meshes, int n_meshes,
                          const float* eps2, const float* T, int n_rays, float* emission,
                          float* absorption) {
  std::vector<subject::material_t> mats(n_meshes);
  std::vector<subject::mesh_t> ms(n_meshes);
  for (int m = 0; m < n_meshes; ++m) {
    mats[m] = subject::black_body();
    mats[m].front.emissivity = eps2[2 * m];
    mats[m].rear.emissivity = eps2[2 * m + 1];
    ms[m] = subject::make_mesh(meshes[3 * m], meshes[3 * m + 1], m);
  }
  subject::scene_t scene = {n_faces, (subject::face_t*)faces, n_meshes, ms.data(), n_meshes,
                            mats.data()};
  ray_caster::system_t* rc = make_caster(0);
  emission::system_t* em = emission::system_create(EMISSION_MALLEY_CPU, rc);
  radiance_equation::params_t params = {em, n_rays};
  thermal_equation::system_t* eq = radiance_equation::system_create();
  thermal_equation::system_init(eq, &params);
  int r = thermal_equation::system_set_scene(eq, &scene);
  thermal_equation::task_t task = {T, emission, absorption};
  memset(emission, 0, n_meshes * sizeof(float));
  memset(absorption, 0, n_meshes * sizeof(float));
  if (r == 0) r = thermal_equation::system_calculate(eq, &task);
  thermal_equation::system_free(eq);
  emission::system_free(em);
  free_caster(rc);
  return r;
}

/* form_factors_equation (thermal_equation/form_factors.cpp:77-128) over
 * FORM_FACTORS_CPU, MALLEY_CPU and the naive caster: set_scene computes F
 * with n_rays, then one calculate() adds the per-mesh emission and the
 * absorption q = F^T E. F is computed a second time by a fresh calculator with
 * the same seeds (identical rays, so identical bits) and returned in F_out;
 * areas_out receives subject::build_meshes_areas. Materials carry front/rear
 * emissivity per mesh. */
EXPORT int ffref_sbff(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* eps2,
                      const float* T, int n_rays, float* F_out, float* areas_out, float* emission,
                      float* absorption) {
  std::vector<subject::material_t> mats(n_meshes);
  std::vector<subject::mesh_t> ms(n_meshes);
  for (int m = 0; m < n_meshes; ++m) {
    mats[m] = subject::black_body();
    mats[m].front.emissivity = eps2[2 * m];
    mats[m].rear.emissivity = eps2[2 * m + 1];
    ms[m] = subject::make_mesh(meshes[3 * m], meshes[3 * m + 1], m);
  }
  subject::scene_t scene = {n_faces, (subject::face_t*)faces, n_meshes, ms.data(), n_meshes,
                            mats.data()};
  float* areas = 0;
  subject::build_meshes_areas(&scene, &areas);
  memcpy(areas_out, areas, n_meshes * sizeof(float));
  free(areas);

  ray_caster::system_t* rc = make_caster(0);
  emission::system_t* em = emission::system_create(EMISSION_MALLEY_CPU, rc);
  form_factors::system_t* ff = form_factors::system_create(FORM_FACTORS_CPU, em);
  form_factors_equation::params_t params = {ff, n_rays};
  thermal_equation::system_t* eq = form_factors_equation::system_create();
  thermal_equation::system_init(eq, &params);
  int r = thermal_equation::system_set_scene(eq, &scene);
  thermal_equation::task_t task = {T, emission, absorption};
  memset(emission, 0, n_meshes * sizeof(float));
  memset(absorption, 0, n_meshes * sizeof(float));
  if (r == 0) r = thermal_equation::system_calculate(eq, &task);
  thermal_equation::system_free(eq);
  form_factors::system_free(ff);
  emission::system_free(em);
  free_caster(rc);
  if (r == 0) r = ffref_ff_calculate(faces, n_faces, meshes, n_meshes, n_rays, 0, F_out);
  return r;
}

/* perfTest's workload (perfTest/perfTest.cpp:42-83) from the reference's own
 * spherical generator, in its order: n_faces confetti triangles (three unit
 * sphere points offset by a point of the ball of `radius`), then n_rays
 * collapsing rays (end point on the sphere of radius 2 + radius, origin 10 %
 * further out). */
EXPORT void ffref_perftest_scene(int n_faces, float radius, int n_rays, float* faces_out, float* rays_out) {
  subject::generator_t* g = subject::generator_create_spherical();
  for (int i = 0; i != n_faces; ++i) {
    math::vec3 p[3];
    subject::generator_surface_point(g, 3, p);
    math::vec3 position;
    subject::generator_volume_point(g, 1, &position);
    position *= radius;
    for (int k = 0; k != 3; ++k) {
      p[k] += position;
      memcpy(faces_out + 9 * (size_t)i + 3 * k, &p[k], sizeof(p[k]));
    }
  }
  for (int i = 0; i != n_rays; ++i) {
    math::vec3 direction;
    subject::generator_surface_point(g, 1, &direction);
    direction *= (2 + radius);
    math::vec3 origin = direction * 1.1f;
    memcpy(rays_out + 6 * (size_t)i, &origin, sizeof(origin));
    memcpy(rays_out + 6 * (size_t)i + 3, &direction, sizeof(direction));
  }
  subject::generator_free(g);
}

/* math::triangles_bsphere (math/triangle.cpp:361-383). */
EXPORT void ffref_bsphere(const float* faces, int n_faces, float* center3, float* radius) {
  math::sphere_t sp = math::triangles_bsphere((const math::triangle_t*)faces, n_faces);
  memcpy(center3, &sp.center, 12);
  *radius = sp.radius;
}

namespace {
parallel_rays_cpu::source_t sun_source(void* p) { return *(parallel_rays_cpu::source_t*)p; }
}  // namespace

/* parallel_rays_cpu (thermal_equation/parallel_rays_cpu.cpp) over
 * PARALLEL_RAYS_CPU emission and the naive caster: one calculate() call.
 * absorb2 holds the front/rear absorbance of each mesh. */
EXPORT int ffref_sun(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* absorb2,
                     float power, const float* dir3, int n_rays, float* absorption) {
  std::vector<subject::material_t> mats(n_meshes);
  std::vector<subject::mesh_t> ms(n_meshes);
  for (int m = 0; m < n_meshes; ++m) {
    mats[m] = subject::black_body();
    mats[m].front.absorbance = absorb2[2 * m];
    mats[m].rear.absorbance = absorb2[2 * m + 1];
    ms[m] = subject::make_mesh(meshes[3 * m], meshes[3 * m + 1], m);
  }
  subject::scene_t scene = {n_faces, (subject::face_t*)faces, n_meshes, ms.data(), n_meshes, mats.data()};
  ray_caster::system_t* rc = make_caster(0);
  emission::system_t* em = emission::system_create(EMISSION_PARALLEL_RAYS_CPU, rc);
  parallel_rays_cpu::source_t src;
  src.power = power;
  memcpy(&src.direction, dir3, 12);
  parallel_rays_cpu::params_t params = {n_rays, em, &src, &sun_source};
  thermal_equation::system_t* eq = parallel_rays_cpu::system_create();
  thermal_equation::system_init(eq, &params);
  int r = thermal_equation::system_set_scene(eq, &scene);
  std::vector<float> T(n_meshes, 300.f), emission(n_meshes, 0.f);
  memset(absorption, 0, n_meshes * sizeof(float));
  thermal_equation::task_t task = {T.data(), emission.data(), absorption};
  if (r == 0) r = thermal_equation::system_calculate(eq, &task);
  thermal_equation::system_free(eq);
  emission::system_free(em);
  free_caster(rc);
  return r;
}

/* heat_source_equation (thermal_equation/heat_source_cpu.cpp): one calculate()
 * adding to the given arrays. */
EXPORT int ffref_heat_source(const int* mesh_idx, const float* power, int n_sources, float* emission,
                             float* absorption) {
  std::vector<heat_source_equation::heat_source_t> src(n_sources);
  for (int i = 0; i < n_sources; ++i) {
    src[i].mesh_idx = mesh_idx[i];
    src[i].power = power[i];
  }
  heat_source_equation::params_t params = {n_sources, src.data()};
  thermal_equation::system_t* eq = heat_source_equation::system_create();
  thermal_equation::system_init(eq, &params);
  thermal_equation::task_t task = {nullptr, emission, absorption};
  int r = thermal_equation::system_calculate(eq, &task);
  thermal_equation::system_free(eq);
  return r;
}

// ---- import_export: obj_import::scene / task, obj_export, csv_export ----

// obj_import::scene; on success *handle holds the scene for ffref_obj_copy.
EXPORT int ffref_obj_import(const char* path, int* n_faces, int* n_meshes, int* n_materials, void** handle) {
  subject::scene_t* scene = nullptr;
  const int r = obj_import::scene(path, &scene);
  *handle = scene;
  if (r != OBJ_IMPORT_OK || !scene) return r;
  *n_faces = scene->n_faces;
  *n_meshes = scene->n_meshes;
  *n_materials = scene->n_materials;
  return r;
}

// faces: 9 floats per face; meshes: (first, n_faces, material); materials:
// 14 floats (shell 4, front 5, rear 5) and a 128-byte name each.
EXPORT void ffref_obj_copy(void* handle, float* faces, int* meshes, float* props, char* names) {
  const subject::scene_t* s = static_cast<const subject::scene_t*>(handle);
  std::memcpy(faces, s->faces, sizeof(float) * 9 * (size_t)s->n_faces);
  for (int m = 0; m < s->n_meshes; ++m) {
    meshes[3 * m] = s->meshes[m].first_idx;
    meshes[3 * m + 1] = s->meshes[m].n_faces;
    meshes[3 * m + 2] = s->meshes[m].material_idx;
  }
  for (int k = 0; k < s->n_materials; ++k) {
    const subject::material_t& mt = s->materials[k];
    std::memcpy(props + 14 * k, &mt.shell, 4 * sizeof(float));
    std::memcpy(props + 14 * k + 4, &mt.front, 5 * sizeof(float));
    std::memcpy(props + 14 * k + 9, &mt.rear, 5 * sizeof(float));
    std::memcpy(names + 128 * k, mt.name, 128);
  }
}

EXPORT void ffref_obj_free(void* handle) { subject::scene_free(static_cast<subject::scene_t*>(handle)); }

// obj_import::task called n_calls times on one open file. Per call: the
// return code, the step and the first n_meshes temperatures; after the last
// call the heat sources (up to src_cap) and the distant source.
EXPORT int ffref_task_import(const char* path, int n_meshes, int n_calls, int* codes, float* dts, float* temps,
                             int* n_src, int* src_mesh, float* src_power, int src_cap, float* dst4, int* has_dst) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  thermal_solution::task_t* t = thermal_solution::task_create(0);
  heat_source_equation::params_t hs = {0, 0};
  parallel_rays_cpu::params_t ds = {1, 0, 0, 0};
  for (int c = 0; c < n_calls; ++c) {
    codes[c] = obj_import::task(f, n_meshes, t, &hs, &ds);
    dts[c] = t->time_delta;
    for (int m = 0; m < n_meshes && m < (codes[c] > 0 ? codes[c] : 0); ++m) temps[(size_t)c * n_meshes + m] = t->temperatures[m];
  }
  fclose(f);
  *n_src = hs.n_sources;
  for (int i = 0; i < hs.n_sources && i < src_cap; ++i) {
    src_mesh[i] = hs.sources[i].mesh_idx;
    src_power[i] = hs.sources[i].power;
  }
  *has_dst = ds.source != 0;
  if (ds.source) {
    parallel_rays_cpu::source_t src = ds.source(ds.source_param);
    dst4[0] = src.power;
    dst4[1] = src.direction.x;
    dst4[2] = src.direction.y;
    dst4[3] = src.direction.z;
  }
  free(hs.sources);
  thermal_solution::task_free(t);
  return 0;
}

// obj_import::scene then obj_export::scene into out_path.
EXPORT int ffref_obj_reexport(const char* obj_path, const char* out_path) {
  subject::scene_t* scene = nullptr;
  int r = obj_import::scene(obj_path, &scene);
  if (r != OBJ_IMPORT_OK) return r;
  FILE* f = fopen(out_path, "w");
  if (!f) return -1;
  r = obj_export::scene(f, scene);
  fclose(f);
  subject::scene_free(scene);
  return r;
}

EXPORT int ffref_task_export(const char* out_path, int n_step, float dt, const float* temps, int n_meshes,
                             int binary) {
  FILE* f = fopen(out_path, "w");
  if (!f) return -1;
  thermal_solution::task_t t;
  t.n_step = n_step;
  t.time_delta = dt;
  t.temperatures = const_cast<float*>(temps);
  const int r = binary ? obj_export::task_binary(f, n_meshes, &t) : obj_export::task(f, n_meshes, &t);
  fclose(f);
  return r;
}

EXPORT int ffref_csv_export(const char* path, const float* F, int n_meshes) {
  form_factors::scene_t scene = {0, 0, n_meshes, 0};
  form_factors::task_t task = {0, const_cast<float*>(F)};
  return csv_export::export_csv(path, &scene, &task);
}
