// HIP implementations of the reference's plugin systems (C-style method
// tables) on top of the ffk C ABI, plus the flat C binding of ff_plugin.h.
//
// Each system is malloc-allocated with its C++ state placement-constructed
// in init() and destroyed in shutdown(), because the reference frees
// systems with free() after shutdown() (ray_caster/system.cpp:97-101).
#include <hip/hip_runtime_api.h>

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "ff_plugin.h"
#include "ffk.h"
#include "form_factors_amd/api.hpp"

namespace {

// A stream of default priority owned by each system.
hipStream_t make_stream() {
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  return s;
}
void drop_stream(hipStream_t s) {
  if (s) (void)hipStreamDestroy(s);
}

template <class T>
struct dbuf {
  T* p = nullptr;
  size_t cap = 0;
  bool reserve(size_t n) {
    if (n <= cap) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    cap = n;
    return true;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

ray_caster::task_t* rc_task_create(int n) {  // ray_caster/system.cpp:44-52
  ray_caster::task_t* t = (ray_caster::task_t*)malloc(sizeof(ray_caster::task_t));
  t->n_tasks = n;
  t->ray = (math::ray_t*)malloc((size_t)n * sizeof(math::ray_t));
  t->hit_face = (ray_caster::face_t**)malloc((size_t)n * sizeof(ray_caster::face_t*));
  t->hit_point = (math::vec3*)malloc((size_t)n * sizeof(math::vec3));
  return t;
}
void rc_task_free(ray_caster::task_t* t) {  // ray_caster/system.cpp:54-63
  if (!t) return;
  free(t->ray);
  free(t->hit_face);
  free(t->hit_point);
  free(t);
}

// Two-sided area weights and total, continuing from `total`
// (cpuFactorsCalculator/cpu_system.cpp:120-129).
float area_of(const math::triangle_t& f) {
  const float ux = f.points[1].x - f.points[0].x, uy = f.points[1].y - f.points[0].y,
              uz = f.points[1].z - f.points[0].z;
  const float vx = f.points[2].x - f.points[0].x, vy = f.points[2].y - f.points[0].y,
              vz = f.points[2].z - f.points[0].z;
  const float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
  return sqrtf(nx * nx + ny * ny + nz * nz) / 2;
}

// A step's equation arrays in device memory (the device-resident step of the
// Adams solution): each equation adds to d_em / d_ab what its calculate adds
// to thermal_equation::task_t's emission / absorption, in the same float
// arithmetic, on the solution's stream.
struct dev_task {
  const float* h_T;  // the step's temperatures on the host
  const float* d_T;  // the same on the device
  float* d_em;
  float* d_ab;
  hipStream_t st;
};

}  // namespace

// ============================================================ ray caster ==
namespace hip_ray_caster {

struct state {
  int method = FFK_CAST_GRID;
  ray_caster::scene_t* scene = nullptr;
  ffk_scene* s = nullptr;
  hipStream_t st = nullptr;
  bool prepared = false;
  dbuf<float> rays, pts;
  dbuf<int> idx;
  std::vector<int> h_idx;
  std::vector<float> h_pts;
};

struct system : ray_caster::system_t {
  state* st;
  int method;
};

int init(system* sys) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -RAY_CASTER_ERROR;
  sys->st->method = sys->method;
  sys->st->st = make_stream();
  return sys->st->st ? RAY_CASTER_OK : -RAY_CASTER_ERROR;
}

int shutdown(system* sys) {
  state* s = sys->st;
  if (!s) return RAY_CASTER_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  ffk_scene_destroy(s->s);
  s->rays.release();
  s->pts.release();
  s->idx.release();
  drop_stream(s->st);
  delete s;
  sys->st = nullptr;
  return RAY_CASTER_OK;
}

int set_scene(system* sys, ray_caster::scene_t* scene) {
  sys->st->scene = scene;
  sys->st->prepared = false;
  return RAY_CASTER_OK;
}

int prepare(system* sys) {
  state* s = sys->st;
  if (!s->scene || s->scene->n_faces == 0) return -RAY_CASTER_ERROR;  // ray_caster/naive_cpu.cpp:56-61
  if (!s->s && ffk_scene_create(&s->s) < 0) return -RAY_CASTER_ERROR;
  if (ffk_scene_set_faces(s->s, (const float*)s->scene->faces, s->scene->n_faces, s->st) < 0)
    return -RAY_CASTER_ERROR;
  if (s->method == FFK_CAST_GRID && ffk_scene_build_grid(s->s, 0, 0, 0, s->st) < 0) return -RAY_CASTER_ERROR;
  if (hipStreamSynchronize(s->st) != hipSuccess) return -RAY_CASTER_ERROR;
  s->prepared = true;
  return RAY_CASTER_OK;
}

// Device-resident cast used by the HIP emitter when both sit on this caster.
int cast_device(system* sys, const float* d_rays, long long n, int* d_idx, float* d_pts, hipStream_t st) {
  state* s = sys->st;
  if (!s->prepared) return -RAY_CASTER_ERROR;
  return ffk_cast(s->s, s->method, d_rays, n, d_idx, d_pts, nullptr, st) < 0 ? -RAY_CASTER_ERROR : RAY_CASTER_OK;
}

// Fill the reference's output convention from device results.
void scatter_hits(state* s, ray_caster::task_t* task, long long lo, long long n) {
  for (long long i = 0; i < n; ++i) {
    const int f = s->h_idx[i];
    if (f >= 0) {
      task->hit_face[lo + i] = &s->scene->faces[f];
      task->hit_point[lo + i].x = s->h_pts[3 * i];
      task->hit_point[lo + i].y = s->h_pts[3 * i + 1];
      task->hit_point[lo + i].z = s->h_pts[3 * i + 2];
    } else {
      task->hit_face[lo + i] = 0;
    }
  }
}

int cast(system* sys, ray_caster::task_t* task) {
  state* s = sys->st;
  if (!s->prepared) return -RAY_CASTER_ERROR;
  const long long total = task->n_tasks;
  const long long chunk = 1LL << 24;
  for (long long lo = 0; lo < total; lo += chunk) {
    const long long n = total - lo < chunk ? total - lo : chunk;
    if (!s->rays.reserve(6 * n) || !s->idx.reserve(n) || !s->pts.reserve(3 * n)) return -RAY_CASTER_ERROR;
    if (hipMemcpyAsync(s->rays.p, task->ray + lo, 24 * n, hipMemcpyHostToDevice, s->st) != hipSuccess)
      return -RAY_CASTER_ERROR;
    if (ffk_cast(s->s, s->method, s->rays.p, n, s->idx.p, s->pts.p, nullptr, s->st) < 0) return -RAY_CASTER_ERROR;
    s->h_idx.resize(n);
    s->h_pts.resize(3 * n);
    if (hipMemcpyAsync(s->h_idx.data(), s->idx.p, 4 * n, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
        hipMemcpyAsync(s->h_pts.data(), s->pts.p, 12 * n, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
        hipStreamSynchronize(s->st) != hipSuccess)
      return -RAY_CASTER_ERROR;
    scatter_hits(s, task, lo, n);
  }
  return RAY_CASTER_OK;
}

const ray_caster::system_methods_t methods = {
    (int (*)(ray_caster::system_t*))&init,
    (int (*)(ray_caster::system_t*))&shutdown,
    (int (*)(ray_caster::system_t*, ray_caster::scene_t*))&set_scene,
    (int (*)(ray_caster::system_t*))&prepare,
    (int (*)(ray_caster::system_t*, ray_caster::task_t*))&cast,
};

static ray_caster::system_t* create(int method) {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  s->method = method;
  return s;
}

ray_caster::system_t* system_create_naive() { return create(FFK_CAST_NAIVE); }
ray_caster::system_t* system_create_grid() { return create(FFK_CAST_GRID); }

bool is_hip(const ray_caster::system_t* s) { return s && s->methods == &methods; }
ffk_scene* device_scene(ray_caster::system_t* s) { return ((system*)s)->st->s; }
int method_of(ray_caster::system_t* s) { return ((system*)s)->st->method; }
ray_caster::scene_t* scene_of(ray_caster::system_t* s) { return ((system*)s)->st->scene; }

}  // namespace hip_ray_caster

// ====================================================== Malley emission ==
namespace hip_malley_emission {

struct state {
  ray_caster::system_t* caster = nullptr;
  ray_caster::scene_t* scene = nullptr;
  ffk_scene* s = nullptr;
  hipStream_t st = nullptr;
  unsigned long long seed = 0, calls = 0;
  int frame = FFK_FRAME_REFERENCE;
  std::vector<int> front, rear;
  dbuf<float> rays, pts;
  dbuf<int> idx;
  std::vector<int> h_idx;
  std::vector<float> h_pts;
};

struct system : emission::system_t {
  state* st;
};

int init(system* sys, ray_caster::system_t* caster) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -EMISSION_ERROR;
  sys->st->caster = caster;
  sys->st->st = make_stream();
  return sys->st->st ? EMISSION_OK : -EMISSION_ERROR;
}

int shutdown(system* sys) {
  state* s = sys->st;
  if (!s) return EMISSION_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  ffk_scene_destroy(s->s);
  s->rays.release();
  s->pts.release();
  s->idx.release();
  drop_stream(s->st);
  delete s;
  sys->st = nullptr;
  return EMISSION_OK;
}

// emission/malley_cpu.cpp:77-92
int set_scene(system* sys, ray_caster::scene_t* scene) {
  state* s = sys->st;
  if (scene == 0 || scene->n_faces == 0) return -EMISSION_ERROR;
  s->scene = scene;
  int r;
  if ((r = s->caster->methods->set_scene(s->caster, scene)) < 0) return r;
  if ((r = s->caster->methods->prepare(s->caster)) < 0) return r;
  if (!s->s && ffk_scene_create(&s->s) < 0) return -EMISSION_ERROR;
  if (ffk_scene_set_faces(s->s, (const float*)scene->faces, scene->n_faces, s->st) < 0) return -EMISSION_ERROR;
  int whole[3] = {0, scene->n_faces, 0};  // the emitter itself needs no mesh split
  if (ffk_scene_set_meshes(s->s, whole, 1, s->st) < 0) return -EMISSION_ERROR;
  s->front.clear();
  return EMISSION_OK;
}

// Key of the next calculate(): successive calls draw fresh rays, as the
// reference's generators advance between calls.
unsigned long long next_key(state* s) { return s->seed + 0x9E3779B97F4A7C15ull * (s->calls++); }

// Per-face counts and the device layout for this task's weights.
int layout(state* s, int n_rays, float total, const float* weights, long long* n_real) {
  const int nf = s->scene->n_faces;
  std::vector<int> fr(nf), rr(nf);
  ffk_emitted_counts(weights, nf, n_rays, total, fr.data(), rr.data());
  if (fr != s->front || rr != s->rear) {
    if (ffk_scene_set_emission(s->s, fr.data(), rr.data(), s->frame, n_real, s->st) < 0) return -EMISSION_ERROR;
    s->front.swap(fr);
    s->rear.swap(rr);
  } else {
    ffk_scene_counts(s->s, nullptr, nullptr, n_real);
  }
  return EMISSION_OK;
}

// emission/malley_cpu.cpp:195-211: generate the rays, then cast them.
int calculate(system* sys, emission::task_t* base) {
  state* s = sys->st;
  malley_emission::task_t* task = (malley_emission::task_t*)base;
  if (task->rays) {
    rc_task_free(task->rays);
    task->rays = 0;
  }
  if (!s->scene) return -EMISSION_ERROR;
  long long n = 0;
  int r;
  if ((r = layout(s, task->n_rays, task->total_weight, task->weights, &n)) < 0) return r;
  if (n == 0) return EMISSION_OK;
  if (n > 0x7fffffffLL) return -EMISSION_ERROR;  // task_t counts rays in an int
  if (!s->rays.reserve(6 * n)) return -EMISSION_ERROR;
  if (ffk_emit_malley(s->s, next_key(s), s->rays.p, s->st) < 0) return -EMISSION_ERROR;
  ray_caster::task_t* rt = rc_task_create((int)n);
  task->rays = rt;
  if (hipMemcpyAsync(rt->ray, s->rays.p, 24 * n, hipMemcpyDeviceToHost, s->st) != hipSuccess) return -EMISSION_ERROR;
  if (hip_ray_caster::is_hip(s->caster) && hip_ray_caster::scene_of(s->caster) == s->scene) {
    // Both on the GPU: cast the device-resident rays, copy results back once.
    if (!s->idx.reserve(n) || !s->pts.reserve(3 * n)) return -EMISSION_ERROR;
    if (hip_ray_caster::cast_device((hip_ray_caster::system*)s->caster, s->rays.p, n, s->idx.p, s->pts.p, s->st) < 0)
      return -EMISSION_ERROR;
    s->h_idx.resize(n);
    s->h_pts.resize(3 * n);
    if (hipMemcpyAsync(s->h_idx.data(), s->idx.p, 4 * n, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
        hipMemcpyAsync(s->h_pts.data(), s->pts.p, 12 * n, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
        hipStreamSynchronize(s->st) != hipSuccess)
      return -EMISSION_ERROR;
    for (long long i = 0; i < n; ++i) {
      const int f = s->h_idx[i];
      if (f >= 0) {
        rt->hit_face[i] = &s->scene->faces[f];
        rt->hit_point[i].x = s->h_pts[3 * i];
        rt->hit_point[i].y = s->h_pts[3 * i + 1];
        rt->hit_point[i].z = s->h_pts[3 * i + 2];
      } else {
        rt->hit_face[i] = 0;
      }
    }
    return EMISSION_OK;
  }
  if (hipStreamSynchronize(s->st) != hipSuccess) return -EMISSION_ERROR;
  return s->caster->methods->cast(s->caster, rt);
}

const emission::system_methods_t methods = {
    (int (*)(emission::system_t*, ray_caster::system_t*))&init,
    (int (*)(emission::system_t*))&shutdown,
    (int (*)(emission::system_t*, ray_caster::scene_t*))&set_scene,
    (int (*)(emission::system_t*, emission::task_t*))&calculate,
};

emission::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  return s;
}

bool is_hip(const emission::system_t* s) { return s && s->methods == &methods; }
state* state_of(emission::system_t* s) { return ((system*)s)->st; }

}  // namespace hip_malley_emission

// ================================================ parallel-rays emission ==
namespace hip_parallel_rays_emission {

struct state {
  ray_caster::system_t* caster = nullptr;
  ray_caster::scene_t* scene = nullptr;
  unsigned long long seed = 0, calls = 0;
  dbuf<float> rays;
};
struct system : emission::system_t {
  state* st;
};

int init(system* sys, ray_caster::system_t* caster) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -EMISSION_ERROR;
  sys->st->caster = caster;
  return EMISSION_OK;
}
int shutdown(system* sys) {
  if (sys->st) {
    sys->st->rays.release();
    delete sys->st;
    sys->st = nullptr;
  }
  return EMISSION_OK;
}
// emission/parallel_rays_cpu.cpp:68-84
int set_scene(system* sys, ray_caster::scene_t* scene) {
  if (scene == 0 || scene->n_faces == 0) return -EMISSION_ERROR;
  sys->st->scene = scene;
  int r;
  if ((r = sys->st->caster->methods->set_scene(sys->st->caster, scene)) < 0) return r;
  if ((r = sys->st->caster->methods->prepare(sys->st->caster)) < 0) return r;
  return EMISSION_OK;
}
// emission/parallel_rays_cpu.cpp:108-121
int calculate(system* sys, emission::task_t* base) {
  state* s = sys->st;
  parallel_rays_emission::task_t* task = (parallel_rays_emission::task_t*)base;
  if (task->rays) {
    rc_task_free(task->rays);
    task->rays = 0;
  }
  const long long n = task->n_rays;
  if (n <= 0) return EMISSION_OK;
  if (!s->rays.reserve(6 * n)) return -EMISSION_ERROR;
  const unsigned long long key = s->seed + 0x9E3779B97F4A7C15ull * (s->calls++);  // as next_key
  if (ffk_emit_parallel(&task->direction.x, &task->center.x, task->distance, task->radius, n, key, s->rays.p,
                        nullptr) < 0)
    return -EMISSION_ERROR;
  ray_caster::task_t* rt = rc_task_create((int)n);
  task->rays = rt;
  if (hipMemcpy(rt->ray, s->rays.p, 24 * n, hipMemcpyDeviceToHost) != hipSuccess) return -EMISSION_ERROR;
  return s->caster->methods->cast(s->caster, rt);
}

const emission::system_methods_t methods = {
    (int (*)(emission::system_t*, ray_caster::system_t*))&init,
    (int (*)(emission::system_t*))&shutdown,
    (int (*)(emission::system_t*, ray_caster::scene_t*))&set_scene,
    (int (*)(emission::system_t*, emission::task_t*))&calculate,
};
emission::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  return s;
}

bool is_hip(const emission::system_t* s) { return s && s->methods == &methods; }
state* state_of(emission::system_t* s) { return ((system*)s)->st; }
unsigned long long next_key(state* s) { return s->seed + 0x9E3779B97F4A7C15ull * (s->calls++); }

}  // namespace hip_parallel_rays_emission

// ========================================================= form factors ==
namespace hip_form_factors {

struct state {
  emission::system_t* emitter = nullptr;
  form_factors::scene_t* scene = nullptr;
  ray_caster::scene_t rc_scene = {0, 0};
  float total_area = 0;
  std::vector<float> weights;
  ffk_scene* s = nullptr;
  hipStream_t st = nullptr;
  bool prepared = false;
  bool grid = false;
  std::vector<int> front, rear;
  int frame = -1;
  dbuf<float> F;
  dbuf<unsigned long long> tally;
  dbuf<int> hits;
  std::vector<int> h_hits;
};
struct system : form_factors::system_t {
  state* st;
};

int init(system* sys, emission::system_t* emitter) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -FORM_FACTORS_ERROR;
  sys->st->emitter = emitter;
  sys->st->st = make_stream();
  return sys->st->st ? FORM_FACTORS_OK : -FORM_FACTORS_ERROR;
}

int shutdown(system* sys) {
  state* s = sys->st;
  if (!s) return FORM_FACTORS_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  ffk_scene_destroy(s->s);
  s->F.release();
  s->tally.release();
  s->hits.release();
  drop_stream(s->st);
  delete s;
  sys->st = nullptr;
  return FORM_FACTORS_OK;
}

// cpuFactorsCalculator/cpu_system.cpp:82-92
int set_scene(system* sys, form_factors::scene_t* scene) {
  state* s = sys->st;
  s->scene = scene;
  s->rc_scene.n_faces = scene->n_faces;
  s->rc_scene.faces = scene->faces;
  s->prepared = false;
  return s->emitter->methods->set_scene(s->emitter, &s->rc_scene);
}

// cpuFactorsCalculator/cpu_system.cpp:95-136. The running total is not
// reset between calls, as in the reference.
int prepare(system* sys) {
  state* s = sys->st;
  if (s->scene == 0 || s->scene->n_faces == 0 || s->scene->n_meshes == 0) return -FORM_FACTORS_EMPTY_SCENE;
  const int nf = s->scene->n_faces;
  s->weights.assign(2 * (size_t)nf, 0.f);
  for (int i = 0; i != nf; ++i) {
    const float a = area_of(s->scene->faces[i]);
    s->total_area += 2.f * a;
    s->weights[2 * i] = a;
    s->weights[2 * i + 1] = a;
  }
  if (s->total_area < FLT_EPSILON) return -FORM_FACTORS_SCENE_TOO_SMALL;
  if (!s->s && ffk_scene_create(&s->s) < 0) return -FORM_FACTORS_ERROR;
  if (ffk_scene_set_faces(s->s, (const float*)s->scene->faces, nf, s->st) < 0) return -FORM_FACTORS_ERROR;
  if (ffk_scene_set_meshes(s->s, (const int*)s->scene->meshes, s->scene->n_meshes, s->st) < 0)
    return -FORM_FACTORS_ERROR;
  s->grid = false;
  s->front.clear();
  s->prepared = true;
  return FORM_FACTORS_OK;
}

int ensure_layout(state* s, int n_rays, int frame) {
  const int nf = s->scene->n_faces;
  std::vector<int> fr(nf), rr(nf);
  ffk_emitted_counts(s->weights.data(), nf, n_rays, s->total_area, fr.data(), rr.data());
  if (fr != s->front || rr != s->rear || frame != s->frame) {
    if (ffk_scene_set_emission(s->s, fr.data(), rr.data(), frame, nullptr, s->st) < 0) return -FORM_FACTORS_ERROR;
    s->front.swap(fr);
    s->rear.swap(rr);
    s->frame = frame;
  }
  return FORM_FACTORS_OK;
}

// cpuFactorsCalculator/cpu_system.cpp:149-201 on the device.
int calculate(system* sys, form_factors::task_t* task) {
  state* s = sys->st;
  if (!s->prepared) return -FORM_FACTORS_ERROR;
  const int N = s->scene->n_meshes;
  const size_t NN = (size_t)N * N;
  if (!s->F.reserve(NN)) return -FORM_FACTORS_ERROR;
  int r;
  hip_malley_emission::state* em =
      hip_malley_emission::is_hip(s->emitter) ? hip_malley_emission::state_of(s->emitter) : nullptr;
  if (em && hip_ray_caster::is_hip(em->caster) && hip_ray_caster::method_of(em->caster) == FFK_CAST_GRID) {
    // Fused: emission, grid cast, tally and normalisation in one device pass.
    if ((r = ensure_layout(s, task->n_rays, em->frame)) < 0) return r;
    if (!s->grid) {
      if (ffk_scene_build_grid(s->s, 0, 0, 0, s->st) < 0) return -FORM_FACTORS_ERROR;
      s->grid = true;
    }
    if (!s->tally.reserve(NN)) return -FORM_FACTORS_ERROR;
    const unsigned long long key = hip_malley_emission::next_key(em);
    if (ffk_ff_fused(s->s, FFK_CAST_GRID, key, s->tally.p, s->F.p, nullptr, s->st) < 0) return -FORM_FACTORS_ERROR;
  } else {
    // Any other emitter: it emits and casts; the tally runs on the device.
    malley_emission::task_t et(task->n_rays, s->total_area, s->weights.data());
    if ((r = s->emitter->methods->calculate(s->emitter, &et)) < 0) {
      rc_task_free(et.rays);
      return r;
    }
    if ((r = ensure_layout(s, task->n_rays, s->frame < 0 ? FFK_FRAME_REFERENCE : s->frame)) < 0) {
      rc_task_free(et.rays);
      return r;
    }
    const long long n = et.rays ? et.rays->n_tasks : 0;
    s->h_hits.resize(n > 0 ? n : 1);
    for (long long i = 0; i < n; ++i)
      s->h_hits[i] = et.rays->hit_face[i] ? (int)(et.rays->hit_face[i] - s->scene->faces) : -1;
    rc_task_free(et.rays);
    if (!s->tally.reserve(NN) || !s->hits.reserve(n)) return -FORM_FACTORS_ERROR;
    if (ffk_tally_zero(s->s, s->tally.p, s->st) < 0) return -FORM_FACTORS_ERROR;
    if (n > 0) {
      if (hipMemcpyAsync(s->hits.p, s->h_hits.data(), 4 * n, hipMemcpyHostToDevice, s->st) != hipSuccess)
        return -FORM_FACTORS_ERROR;
      if (ffk_tally(s->s, s->hits.p, n, s->tally.p, s->st) < 0) return -FORM_FACTORS_ERROR;
    }
    if (ffk_normalise(s->s, s->tally.p, s->F.p, s->st) < 0) return -FORM_FACTORS_ERROR;
  }
  if (hipMemcpyAsync(task->form_factors, s->F.p, 4 * NN, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
      hipStreamSynchronize(s->st) != hipSuccess)
    return -FORM_FACTORS_ERROR;
  return FORM_FACTORS_OK;
}

const form_factors::system_methods_t methods = {
    (int (*)(form_factors::system_t*, emission::system_t*))&init,
    (int (*)(form_factors::system_t*))&shutdown,
    (int (*)(form_factors::system_t*, form_factors::scene_t*))&set_scene,
    (int (*)(form_factors::system_t*))&prepare,
    (int (*)(form_factors::system_t*, form_factors::task_t*))&calculate,
};
form_factors::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  return s;
}

}  // namespace hip_form_factors

// ================================= SB_FF equation (radiative exchange) ==
namespace hip_sb_ff {

struct state {
  params_t params = {nullptr, 0, FFK_F32};
  subject::scene_t* scene = nullptr;
  form_factors::scene_t ff_scene = {0, 0, 0, 0};
  std::vector<float> F;
  std::vector<double> T, area, ef, er, em, ab;
  dbuf<float> dF32;
  dbuf<unsigned short> dS;  // packed exchange storage (2 or 4 bytes per cell)
  dbuf<double> dvec;  // T, area, ef, er, em, ab, row scale, E in FP32 (n/2 + 1 doubles)
  hipStream_t st = nullptr;
};
struct system : thermal_equation::system_t {
  state* st;
};

int init(system* sys, params_t* params) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -THERMAL_EQUATION_ERROR;
  sys->st->params = *params;
  sys->st->st = make_stream();
  return sys->st->st ? THERMAL_EQUATION_OK : -THERMAL_EQUATION_ERROR;
}
int shutdown(system* sys) {
  state* s = sys->st;
  if (!s) return THERMAL_EQUATION_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  s->dF32.release();
  s->dS.release();
  s->dvec.release();
  drop_stream(s->st);
  delete s;
  sys->st = nullptr;
  return THERMAL_EQUATION_OK;
}

// thermal_equation/form_factors.cpp:77-104: F is computed once here.
int set_scene(system* sys, subject::scene_t* scene) {
  state* s = sys->st;
  s->scene = scene;
  const int N = scene->n_meshes;
  s->area.assign(N, 0.0);
  s->ef.assign(N, 0.0);
  s->er.assign(N, 0.0);
  for (int m = 0; m < N; ++m) {
    float a = 0;  // subject/system.cpp:48-56, float sum per mesh
    for (int f = 0; f != scene->meshes[m].n_faces; ++f) a += area_of(scene->faces[scene->meshes[m].first_idx + f]);
    s->area[m] = a;
    const subject::material_t& mat = scene->materials[scene->meshes[m].material_idx];
    s->ef[m] = mat.front.emissivity;
    s->er[m] = mat.rear.emissivity;
  }
  s->ff_scene.n_faces = scene->n_faces;
  s->ff_scene.faces = scene->faces;
  s->ff_scene.n_meshes = N;
  s->ff_scene.meshes = (form_factors::mesh_t*)scene->meshes;  // layout-compatible (form_factors.cpp:87)
  form_factors::system_t* ff = s->params.form_factors_calculator;
  int r;
  if ((r = ff->methods->set_scene(ff, &s->ff_scene)) < 0) return r;
  if ((r = ff->methods->prepare(ff)) < 0) return r;
  s->F.assign((size_t)N * N, 0.f);
  form_factors::task_t task = {s->params.n_rays, s->F.data()};
  if ((r = ff->methods->calculate(ff, &task)) < 0) return r;
  // F in the chosen storage type on the device
  const size_t NN = (size_t)N * N;
  if (!s->dF32.reserve(NN) || !s->dvec.reserve(7 * (size_t)N + (size_t)N / 2 + 1)) return -THERMAL_EQUATION_ERROR;
  if (hipMemcpyAsync(s->dF32.p, s->F.data(), 4 * NN, hipMemcpyHostToDevice, s->st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  double* d = s->dvec.p;
  // packed (transposed) storage; count-valued rows are stored as exact counts
  if (!s->dS.reserve(NN * (s->params.f_dtype == FFK_F32 ? 2 : 1))) return -THERMAL_EQUATION_ERROR;
  if (ffk_pack_exchange(s->dF32.p, N, s->params.f_dtype, nullptr, s->dS.p, d + 6 * N, s->st) < 0)
    return -THERMAL_EQUATION_ERROR;
  s->dF32.release();
  if (hipMemcpyAsync(d + N, s->area.data(), 8 * N, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(d + 2 * N, s->ef.data(), 8 * N, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(d + 3 * N, s->er.data(), 8 * N, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipStreamSynchronize(s->st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  return THERMAL_EQUATION_OK;
}

// thermal_equation/form_factors.cpp:109-128; adds to the task's arrays.
int calculate(system* sys, thermal_equation::task_t* task) {
  state* s = sys->st;
  const int N = s->scene->n_meshes;
  s->T.assign(task->temperatures, task->temperatures + N);
  double* d = s->dvec.p;
  const void* F = s->dS.p;
  if (hipMemcpyAsync(d, s->T.data(), 8 * N, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemsetAsync(d + 4 * N, 0, 16 * (size_t)N, s->st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  const double* scale = d + 6 * N;
  float* work = reinterpret_cast<float*>(d + 7 * N);
  if (ffk_sbff(N, F, s->params.f_dtype, scale, d, d + N, d + 2 * N, d + 3 * N, d + 4 * N, d + 5 * N, work, s->st) <
      0)
    return -THERMAL_EQUATION_ERROR;
  s->em.resize(N);
  s->ab.resize(N);
  if (hipMemcpyAsync(s->em.data(), d + 4 * N, 8 * N, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
      hipMemcpyAsync(s->ab.data(), d + 5 * N, 8 * N, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
      hipStreamSynchronize(s->st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  for (int m = 0; m < N; ++m) {
    task->emission[m] += (float)s->em[m];
    task->absorption[m] += (float)s->ab[m];
  }
  return THERMAL_EQUATION_OK;
}

bool can_dev(system* sys) { return sys->st && sys->st->scene && sys->st->scene->n_meshes > 0; }

// calculate with the step's arrays on the device: the same double exchange
// from the float temperatures, its sums rounded and added.
int calculate_dev(system* sys, const dev_task& t) {
  state* s = sys->st;
  const int N = s->scene->n_meshes;
  double* d = s->dvec.p;
  return ffk_sbff_dev(N, s->dS.p, s->params.f_dtype, d + 6 * N, t.d_T, d + N, d + 2 * N, d + 3 * N, d, d + 4 * N,
                      d + 5 * N, reinterpret_cast<float*>(d + 7 * N), t.d_em, t.d_ab, t.st) < 0
             ? -THERMAL_EQUATION_ERROR
             : THERMAL_EQUATION_OK;
}

const thermal_equation::system_methods_t methods = {
    (int (*)(thermal_equation::system_t*, void*))&init,
    (int (*)(thermal_equation::system_t*))&shutdown,
    (int (*)(thermal_equation::system_t*, subject::scene_t*))&set_scene,
    (int (*)(thermal_equation::system_t*, thermal_equation::task_t*))&calculate,
};
thermal_equation::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  return s;
}

}  // namespace hip_sb_ff

// ======================================== radiance equation (Monte Carlo) ==
namespace hip_radiance {

struct state {
  radiance_equation::params_t params = {nullptr, 0};
  subject::scene_t* scene = nullptr;
  ray_caster::scene_t emission_scene = {0, 0};
  ffk_scene* s = nullptr;
  hipStream_t st = nullptr;
  bool grid = false;
  std::vector<int> face2mesh;
  std::vector<float> eps2, weights, power;
  std::vector<double> absorption;
  bool dev_pending = false;  // a device step whose status is still to be read
};
struct system : thermal_equation::system_t {
  state* st;
};

int init(system* sys, radiance_equation::params_t* params) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -THERMAL_EQUATION_ERROR;
  sys->st->params = *params;
  sys->st->st = make_stream();
  return sys->st->st ? THERMAL_EQUATION_OK : -THERMAL_EQUATION_ERROR;
}

int shutdown(system* sys) {
  state* s = sys->st;
  if (!s) return THERMAL_EQUATION_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  ffk_scene_destroy(s->s);
  drop_stream(s->st);
  delete s;
  sys->st = nullptr;
  return THERMAL_EQUATION_OK;
}

// thermal_equation/radiance_cpu.cpp:78-94: the emitter gets the faces of the
// subject scene (a layout-compatible prefix); the device scene is built here.
int set_scene(system* sys, subject::scene_t* scene) {
  state* s = sys->st;
  s->scene = scene;
  s->emission_scene.n_faces = scene->n_faces;
  s->emission_scene.faces = scene->faces;
  int r;
  emission::system_t* em = s->params.emitter;
  if ((r = em->methods->set_scene(em, &s->emission_scene)) < 0) return r;
  const int nf = scene->n_faces, nm = scene->n_meshes;
  s->face2mesh.assign(nf, -1);  // subject/system.cpp:98-112; faces outside every mesh: -1
  s->eps2.assign(2 * (size_t)nm, 0.f);
  for (int m = 0; m < nm; ++m) {
    const subject::mesh_t& mesh = scene->meshes[m];
    if (mesh.first_idx < 0 || mesh.n_faces < 0 || mesh.first_idx + mesh.n_faces > nf) return -THERMAL_EQUATION_ERROR;
    for (int f = 0; f < mesh.n_faces; ++f) s->face2mesh[mesh.first_idx + f] = m;
    const subject::material_t& mat = scene->materials[mesh.material_idx];
    s->eps2[2 * m] = mat.front.emissivity;
    s->eps2[2 * m + 1] = mat.rear.emissivity;
  }
  s->grid = false;
  if (nf == 0 || nm == 0) return THERMAL_EQUATION_OK;
  if (!s->s && ffk_scene_create(&s->s) < 0) return -THERMAL_EQUATION_ERROR;
  if (ffk_scene_set_faces(s->s, (const float*)scene->faces, nf, s->st) < 0) return -THERMAL_EQUATION_ERROR;
  if (ffk_scene_set_meshes(s->s, (const int*)scene->meshes, nm, s->st) < 0) return -THERMAL_EQUATION_ERROR;
  return THERMAL_EQUATION_OK;
}

// thermal_equation/radiance_cpu.cpp:132-190; adds to the task's arrays. The
// emission is added face by face with the reference's float expressions; the
// absorbed power of each mesh is the exact sum of its ray powers, added once.
int calculate(system* sys, thermal_equation::task_t* task) {
  state* s = sys->st;
  const subject::scene_t* scene = s->scene;
  if (!scene) return -THERMAL_EQUATION_ERROR;
  const int nf = scene->n_faces, nm = scene->n_meshes;
  if (nf == 0 || nm == 0) return THERMAL_EQUATION_OK;
  const int n_rays = s->params.n_rays;
  emission::system_t* emitter = s->params.emitter;
  hip_malley_emission::state* em =
      hip_malley_emission::is_hip(emitter) ? hip_malley_emission::state_of(emitter) : nullptr;
  int r;
  s->absorption.assign(nm, 0.0);
  if (em && hip_ray_caster::is_hip(em->caster) && hip_ray_caster::method_of(em->caster) == FFK_CAST_GRID) {
    // Fused on the device: weights, layout and powers on the host with the
    // scene's cached areas, then emission, grid cast and deposit in one pass.
    if (!s->grid) {
      if (ffk_scene_build_grid(s->s, 0, 0, 0, s->st) < 0) return -THERMAL_EQUATION_ERROR;
      s->grid = true;
    }
    const unsigned long long key = hip_malley_emission::next_key(em);
    if (ffk_radiance_step(s->s, s->eps2.data(), task->temperatures, n_rays, em->frame, key, task->emission,
                          s->absorption.data(), nullptr, s->st) < 0)
      return -THERMAL_EQUATION_ERROR;
  } else {
    float total = 0;
    s->weights.assign(2 * (size_t)nf, 0.f);
    if (ffk_radiance_weights((const float*)scene->faces, nf, (const int*)scene->meshes, nm, s->eps2.data(),
                             task->temperatures, s->weights.data(), &total) < 0)
      return -THERMAL_EQUATION_ERROR;
    std::vector<int> fr(nf), rr(nf);
    ffk_emitted_counts(s->weights.data(), nf, n_rays, total, fr.data(), rr.data());
    s->power.assign(2 * (size_t)nf, 0.f);
    if (ffk_radiance_powers((const float*)scene->faces, nf, (const int*)scene->meshes, nm, s->eps2.data(),
                            task->temperatures, fr.data(), rr.data(), s->power.data(), task->emission) < 0)
      return -THERMAL_EQUATION_ERROR;
    // Any other emitter emits and casts; the walk runs here, in double.
    malley_emission::task_t et(n_rays, total, s->weights.data());
    if ((r = emitter->methods->calculate(emitter, &et)) < 0) {
      rc_task_free(et.rays);
      return r;
    }
    const long long n_total = et.rays ? et.rays->n_tasks : 0;
    long long n_ray = 0;
    for (int m = 0; m < nm; ++m) {
      const subject::mesh_t& mesh = scene->meshes[m];
      for (int f = mesh.first_idx; f < mesh.first_idx + mesh.n_faces; ++f) {
        const int face_rays = fr[f] + rr[f];
        for (int j = 0; j < face_rays && n_ray < n_total; ++j, ++n_ray) {
          const math::triangle_t* h = et.rays->hit_face[n_ray];
          if (!h) continue;
          const int hm = s->face2mesh[h - scene->faces];
          if (hm >= 0) s->absorption[hm] += (double)s->power[2 * (size_t)f + (j < fr[f] ? 0 : 1)];
        }
      }
    }
    rc_task_free(et.rays);
  }
  for (int m = 0; m < nm; ++m) task->absorption[m] += (float)s->absorption[m];
  return THERMAL_EQUATION_OK;
}

// The device-resident form runs where calculate fuses the step on the device
// (the HIP Malley emitter over the HIP grid caster), or where there is
// nothing to add.
bool can_dev(system* sys) {
  state* s = sys->st;
  if (!s || !s->scene) return false;
  if (s->scene->n_faces == 0 || s->scene->n_meshes == 0) return true;
  emission::system_t* emitter = s->params.emitter;
  if (!hip_malley_emission::is_hip(emitter)) return false;
  const hip_malley_emission::state* em = hip_malley_emission::state_of(emitter);
  return hip_ray_caster::is_hip(em->caster) && hip_ray_caster::method_of(em->caster) == FFK_CAST_GRID;
}

// calculate's fused branch with the step's arrays on the device.
int calculate_dev(system* sys, const dev_task& t) {
  state* s = sys->st;
  const subject::scene_t* scene = s->scene;
  if (scene->n_faces == 0 || scene->n_meshes == 0) return THERMAL_EQUATION_OK;
  hip_malley_emission::state* em = hip_malley_emission::state_of(s->params.emitter);
  if (!s->grid) {
    if (ffk_scene_build_grid(s->s, 0, 0, 0, t.st) < 0) return -THERMAL_EQUATION_ERROR;
    s->grid = true;
  }
  const unsigned long long key = hip_malley_emission::next_key(em);
  if (ffk_radiance_step_dev(s->s, s->eps2.data(), t.h_T, s->params.n_rays, em->frame, key, t.d_em, t.d_ab, t.st) <
      0)
    return -THERMAL_EQUATION_ERROR;
  s->dev_pending = true;
  return THERMAL_EQUATION_OK;
}

// After the solution's stream has drained: the pass's power check.
int status_dev(system* sys) {
  state* s = sys->st;
  if (!s->dev_pending) return THERMAL_EQUATION_OK;
  s->dev_pending = false;
  return ffk_radiance_status(s->s) < 0 ? -THERMAL_EQUATION_ERROR : THERMAL_EQUATION_OK;
}

const thermal_equation::system_methods_t methods = {
    (int (*)(thermal_equation::system_t*, void*))&init,
    (int (*)(thermal_equation::system_t*))&shutdown,
    (int (*)(thermal_equation::system_t*, subject::scene_t*))&set_scene,
    (int (*)(thermal_equation::system_t*, thermal_equation::task_t*))&calculate,
};
thermal_equation::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  return s;
}

}  // namespace hip_radiance

// ================================================ conduction (device) ==
namespace hip_conduction {

struct state {
  subject::scene_t* scene = nullptr;
  hipStream_t st = nullptr;
  int n = 0;
  dbuf<int> row, col;
  dbuf<double> G, vec;  // vec: T, emission, absorption
  std::vector<double> T, em, ab;
};
struct system : thermal_equation::system_t {
  state* st;
};

int init(system* sys, void*) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -THERMAL_EQUATION_ERROR;
  sys->st->st = make_stream();
  return sys->st->st ? THERMAL_EQUATION_OK : -THERMAL_EQUATION_ERROR;
}
int shutdown(system* sys) {
  state* s = sys->st;
  if (!s) return THERMAL_EQUATION_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  s->row.release();
  s->col.release();
  s->G.release();
  s->vec.release();
  drop_stream(s->st);
  delete s;
  sys->st = nullptr;
  return THERMAL_EQUATION_OK;
}

// thermal_equation/conductive_cpu.cpp:50-64: the face graph, here reduced
// once to the mesh conductance operator (shell conductivity and thickness of
// each mesh's material).
int set_scene(system* sys, subject::scene_t* scene) {
  state* s = sys->st;
  s->scene = scene;
  const int nm = scene->n_meshes;
  s->n = nm;
  if (nm == 0) return THERMAL_EQUATION_OK;
  std::vector<float> k(nm), t(nm);
  std::vector<int> meshes(3 * (size_t)nm);
  for (int m = 0; m < nm; ++m) {
    const subject::material_t& mat = scene->materials[scene->meshes[m].material_idx];
    k[m] = mat.shell.thermal_conductivity;
    t[m] = mat.shell.thickness;
    meshes[3 * m] = scene->meshes[m].first_idx;
    meshes[3 * m + 1] = scene->meshes[m].n_faces;
    meshes[3 * m + 2] = 0;
  }
  int nnz = 0;
  const float* faces = (const float*)scene->faces;
  if (ffk_conduction_graph(faces, scene->n_faces, meshes.data(), nm, k.data(), t.data(), nullptr, nullptr, nullptr,
                           0, &nnz) < 0)
    return -THERMAL_EQUATION_ERROR;
  std::vector<int> row(nm + 1), col(std::max(1, nnz));
  std::vector<double> G(std::max(1, nnz));
  if (ffk_conduction_graph(faces, scene->n_faces, meshes.data(), nm, k.data(), t.data(), row.data(), col.data(),
                           G.data(), (int)G.size(), &nnz) < 0)
    return -THERMAL_EQUATION_ERROR;
  if (!s->row.reserve(nm + 1) || !s->col.reserve(col.size()) || !s->G.reserve(G.size()) ||
      !s->vec.reserve(3 * (size_t)nm))
    return -THERMAL_EQUATION_ERROR;
  if (hipMemcpyAsync(s->row.p, row.data(), 4 * row.size(), hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(s->col.p, col.data(), 4 * col.size(), hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(s->G.p, G.data(), 8 * G.size(), hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipStreamSynchronize(s->st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  return THERMAL_EQUATION_OK;
}

// thermal_equation/conductive_cpu.cpp:83-141; adds to the task's arrays.
int calculate(system* sys, thermal_equation::task_t* task) {
  state* s = sys->st;
  const int n = s->n;
  if (n == 0) return THERMAL_EQUATION_OK;
  s->T.assign(task->temperatures, task->temperatures + n);
  s->em.assign(n, 0.0);
  s->ab.assign(n, 0.0);
  double* d = s->vec.p;
  if (hipMemcpyAsync(d, s->T.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(d + n, s->em.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(d + 2 * n, s->ab.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s->st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  if (ffk_conduction(n, s->row.p, s->col.p, s->G.p, d, d + n, d + 2 * n, s->st) < 0) return -THERMAL_EQUATION_ERROR;
  if (hipMemcpyAsync(s->em.data(), d + n, 8 * (size_t)n, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
      hipMemcpyAsync(s->ab.data(), d + 2 * n, 8 * (size_t)n, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
      hipStreamSynchronize(s->st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  for (int m = 0; m < n; ++m) {
    task->emission[m] += (float)s->em[m];
    task->absorption[m] += (float)s->ab[m];
  }
  return THERMAL_EQUATION_OK;
}

bool can_dev(system* sys) { return sys->st && sys->st->scene; }

// calculate with the step's arrays on the device: one kernel from the float
// temperatures, the same double sums, rounded and added.
int calculate_dev(system* sys, const dev_task& t) {
  state* s = sys->st;
  if (s->n == 0) return THERMAL_EQUATION_OK;
  return ffk_conduction_dev(s->n, s->row.p, s->col.p, s->G.p, t.d_T, t.d_em, t.d_ab, t.st) < 0
             ? -THERMAL_EQUATION_ERROR
             : THERMAL_EQUATION_OK;
}

const thermal_equation::system_methods_t methods = {
    (int (*)(thermal_equation::system_t*, void*))&init,
    (int (*)(thermal_equation::system_t*))&shutdown,
    (int (*)(thermal_equation::system_t*, subject::scene_t*))&set_scene,
    (int (*)(thermal_equation::system_t*, thermal_equation::task_t*))&calculate,
};
thermal_equation::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  return s;
}

}  // namespace hip_conduction

// ======================================================== heat sources ==
// thermal_equation/heat_source_cpu.cpp:48-61: a handful of additions per
// step, kept on the host (the device-resident run applies the same sources
// inside its step kernel).
namespace hip_heat_source {

// Device copies of the sources for the device-resident step: uploaded each
// step through pinned staging (the caller may change its sources between
// steps, as the reference reads them on every calculate).
struct dev_state {
  dbuf<int> mesh;
  dbuf<float> power;
  int* h_mesh = nullptr;
  float* h_power = nullptr;
  int cap = 0;
  hipEvent_t ev = nullptr;  // the last upload from the staging
};
struct system : thermal_equation::system_t {
  heat_source_equation::params_t* params;
  dev_state* dev;
};

int init(system* sys, heat_source_equation::params_t* params) {
  sys->params = new (std::nothrow) heat_source_equation::params_t(*params);
  return sys->params ? THERMAL_EQUATION_OK : -THERMAL_EQUATION_ERROR;
}
int shutdown(system* sys) {
  delete sys->params;
  sys->params = nullptr;
  if (dev_state* d = sys->dev) {
    if (d->ev) {
      (void)hipEventSynchronize(d->ev);
      (void)hipEventDestroy(d->ev);
    }
    if (d->h_mesh) (void)hipHostFree(d->h_mesh);
    if (d->h_power) (void)hipHostFree(d->h_power);
    d->mesh.release();
    d->power.release();
    delete d;
    sys->dev = nullptr;
  }
  return THERMAL_EQUATION_OK;
}
int set_scene(system*, subject::scene_t*) { return THERMAL_EQUATION_OK; }
int calculate(system* sys, thermal_equation::task_t* task) {
  const heat_source_equation::params_t& p = *sys->params;
  for (int i = 0; i != p.n_sources; ++i) {
    const heat_source_equation::heat_source_t& src = p.sources[i];
    if (src.power > 0) task->absorption[src.mesh_idx] += src.power;
    else task->emission[src.mesh_idx] += -src.power;
  }
  return THERMAL_EQUATION_OK;
}

bool can_dev(system* sys) { return sys->params != nullptr; }

// calculate with the step's arrays on the device (the same additions, in
// list order, by one device thread).
int calculate_dev(system* sys, const dev_task& t) {
  const heat_source_equation::params_t& p = *sys->params;
  if (p.n_sources <= 0) return THERMAL_EQUATION_OK;
  if (!sys->dev && !(sys->dev = new (std::nothrow) dev_state())) return -THERMAL_EQUATION_ERROR;
  dev_state& d = *sys->dev;
  if (!d.ev && hipEventCreateWithFlags(&d.ev, hipEventDisableTiming) != hipSuccess) return -THERMAL_EQUATION_ERROR;
  if (hipEventSynchronize(d.ev) != hipSuccess) return -THERMAL_EQUATION_ERROR;
  if (d.cap < p.n_sources) {
    if (d.h_mesh) (void)hipHostFree(d.h_mesh);
    if (d.h_power) (void)hipHostFree(d.h_power);
    d.h_mesh = nullptr;
    d.h_power = nullptr;
    d.cap = 0;
    if (hipHostMalloc((void**)&d.h_mesh, 4 * (size_t)p.n_sources, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&d.h_power, 4 * (size_t)p.n_sources, hipHostMallocDefault) != hipSuccess ||
        !d.mesh.reserve(p.n_sources) || !d.power.reserve(p.n_sources))
      return -THERMAL_EQUATION_ERROR;
    d.cap = p.n_sources;
  }
  for (int i = 0; i != p.n_sources; ++i) {
    d.h_mesh[i] = p.sources[i].mesh_idx;
    d.h_power[i] = p.sources[i].power;
  }
  if (hipMemcpyAsync(d.mesh.p, d.h_mesh, 4 * (size_t)p.n_sources, hipMemcpyHostToDevice, t.st) != hipSuccess ||
      hipMemcpyAsync(d.power.p, d.h_power, 4 * (size_t)p.n_sources, hipMemcpyHostToDevice, t.st) != hipSuccess ||
      hipEventRecord(d.ev, t.st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  return ffk_heat_sources_dev(p.n_sources, d.mesh.p, d.power.p, t.d_em, t.d_ab, t.st) < 0 ? -THERMAL_EQUATION_ERROR
                                                                                        : THERMAL_EQUATION_OK;
}

const thermal_equation::system_methods_t methods = {
    (int (*)(thermal_equation::system_t*, void*))&init,
    (int (*)(thermal_equation::system_t*))&shutdown,
    (int (*)(thermal_equation::system_t*, subject::scene_t*))&set_scene,
    (int (*)(thermal_equation::system_t*, thermal_equation::task_t*))&calculate,
};
thermal_equation::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->params = nullptr;
  s->dev = nullptr;
  return s;
}

}  // namespace hip_heat_source

// ============================================ distant source (parallel rays) ==
namespace hip_distant_source {

struct state {
  parallel_rays_cpu::params_t params = {0, nullptr, nullptr, nullptr};
  subject::scene_t* scene = nullptr;
  ray_caster::scene_t emission_scene = {0, 0};
  ffk_scene* s = nullptr;
  hipStream_t st = nullptr;
  bool grid = false;
  float center[3] = {0, 0, 0}, bs_radius = 0;
  std::vector<int> face2mesh;
  std::vector<float> absorb2;
  std::vector<unsigned long long> counts;
  std::vector<double> absorption;
  dbuf<float> d_absorb2;  // device-resident step: absorbances (uploaded once per scene)
  dbuf<unsigned long long> d_counts;
  bool dev_ready = false;
  // a step's hit counts launched ahead on this system's stream (prelaunch_dev)
  hipEvent_t ev_start = nullptr, ev_counted = nullptr;
  bool counting = false;
  float counted_power = 0.f;
};
struct system : thermal_equation::system_t {
  state* st;
};

int init(system* sys, parallel_rays_cpu::params_t* params) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -THERMAL_EQUATION_ERROR;
  sys->st->params = *params;
  sys->st->st = make_stream();
  return sys->st->st ? THERMAL_EQUATION_OK : -THERMAL_EQUATION_ERROR;
}
int shutdown(system* sys) {
  state* s = sys->st;
  if (!s) return THERMAL_EQUATION_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  ffk_scene_destroy(s->s);
  s->d_absorb2.release();
  s->d_counts.release();
  if (s->ev_start) (void)hipEventDestroy(s->ev_start);
  if (s->ev_counted) (void)hipEventDestroy(s->ev_counted);
  drop_stream(s->st);
  delete s;
  sys->st = nullptr;
  return THERMAL_EQUATION_OK;
}

// thermal_equation/parallel_rays_cpu.cpp:68-83
int set_scene(system* sys, subject::scene_t* scene) {
  state* s = sys->st;
  s->scene = scene;
  s->emission_scene.n_faces = scene->n_faces;
  s->emission_scene.faces = scene->faces;
  int r;
  emission::system_t* em = s->params.emitter;
  if ((r = em->methods->set_scene(em, &s->emission_scene)) < 0) return r;
  const int nf = scene->n_faces, nm = scene->n_meshes;
  s->face2mesh.assign(nf, -1);
  s->absorb2.assign(2 * (size_t)nm, 0.f);
  for (int m = 0; m < nm; ++m) {
    const subject::mesh_t& mesh = scene->meshes[m];
    if (mesh.first_idx < 0 || mesh.n_faces < 0 || mesh.first_idx + mesh.n_faces > nf) return -THERMAL_EQUATION_ERROR;
    for (int f = 0; f < mesh.n_faces; ++f) s->face2mesh[mesh.first_idx + f] = m;
    const subject::material_t& mat = scene->materials[mesh.material_idx];
    s->absorb2[2 * m] = mat.front.absorbance;
    s->absorb2[2 * m + 1] = mat.rear.absorbance;
  }
  if (nf > 0 && ffk_bsphere((const float*)scene->faces, nf, s->center, &s->bs_radius) < 0)
    return -THERMAL_EQUATION_ERROR;
  s->grid = false;
  s->dev_ready = false;
  if (nf == 0 || nm == 0) return THERMAL_EQUATION_OK;
  if (!s->s && ffk_scene_create(&s->s) < 0) return -THERMAL_EQUATION_ERROR;
  if (ffk_scene_set_faces(s->s, (const float*)scene->faces, nf, s->st) < 0) return -THERMAL_EQUATION_ERROR;
  if (ffk_scene_set_meshes(s->s, (const int*)scene->meshes, nm, s->st) < 0) return -THERMAL_EQUATION_ERROR;
  return THERMAL_EQUATION_OK;
}

// thermal_equation/parallel_rays_cpu.cpp:92-131; adds to task->absorption.
int calculate(system* sys, thermal_equation::task_t* task) {
  state* s = sys->st;
  if (s->params.source == 0) return THERMAL_EQUATION_OK;
  const subject::scene_t* scene = s->scene;
  if (!scene || scene->n_faces == 0 || scene->n_meshes == 0) return THERMAL_EQUATION_OK;
  const parallel_rays_cpu::source_t source = s->params.source(s->params.source_param);
  const float radius = s->bs_radius * 1.1f;
  const float distance = 1.1f + radius;
  const int n_rays = s->params.n_rays;
  const float ray_power = source.power * (float(M_PI) * radius * radius) / n_rays;
  const int nm = scene->n_meshes;
  s->absorption.assign(nm, 0.0);
  emission::system_t* emitter = s->params.emitter;
  hip_parallel_rays_emission::state* em =
      hip_parallel_rays_emission::is_hip(emitter) ? hip_parallel_rays_emission::state_of(emitter) : nullptr;
  int r;
  if (em && hip_ray_caster::is_hip(em->caster) && hip_ray_caster::method_of(em->caster) == FFK_CAST_GRID) {
    // Fused on the device: emission, grid cast and per-side hit counts.
    if (!s->grid) {
      if (ffk_scene_build_grid(s->s, 0, 0, 0, s->st) < 0) return -THERMAL_EQUATION_ERROR;
      s->grid = true;
    }
    s->counts.assign(2 * (size_t)nm, 0ull);
    const unsigned long long key = hip_parallel_rays_emission::next_key(em);
    if (ffk_sun(s->s, &source.direction.x, s->center, distance, radius, 0, n_rays, key, s->counts.data(), nullptr,
                s->st) < 0)
      return -THERMAL_EQUATION_ERROR;
    if (ffk_sun_deposit(nm, s->counts.data(), s->absorb2.data(), ray_power, s->absorption.data()) < 0)
      return -THERMAL_EQUATION_ERROR;
  } else {
    parallel_rays_emission::task_t et(n_rays);
    et.rays = nullptr;
    et.center = math::vec3{s->center[0], s->center[1], s->center[2]};
    et.distance = distance;
    et.direction = source.direction;
    et.radius = radius;
    if ((r = emitter->methods->calculate(emitter, &et)) < 0) {
      rc_task_free(et.rays);
      return r;
    }
    const ray_caster::task_t* rt = et.rays;
    for (int i = 0; rt && i != rt->n_tasks; ++i) {
      const math::triangle_t* h = rt->hit_face[i];
      if (!h) continue;
      const int m = s->face2mesh[h - scene->faces];
      if (m < 0) continue;
      const math::vec3 u = {h->points[1].x - h->points[0].x, h->points[1].y - h->points[0].y,
                            h->points[1].z - h->points[0].z};
      const math::vec3 v = {h->points[2].x - h->points[0].x, h->points[2].y - h->points[0].y,
                            h->points[2].z - h->points[0].z};
      const float nx = u.y * v.z - u.z * v.y, ny = u.z * v.x - u.x * v.z, nz = u.x * v.y - u.y * v.x;
      const float side = nx * source.direction.x + ny * source.direction.y + nz * source.direction.z;
      s->absorption[m] += (double)((side < 0 ? s->absorb2[2 * m] : s->absorb2[2 * m + 1]) * ray_power);
    }
    rc_task_free(et.rays);
  }
  for (int m = 0; m < nm; ++m) task->absorption[m] += (float)s->absorption[m];
  return THERMAL_EQUATION_OK;
}

// As hip_radiance::can_dev: where calculate fuses the step on the device, or
// where it adds nothing.
bool can_dev(system* sys) {
  state* s = sys->st;
  if (!s || !s->scene) return false;
  if (s->params.source == 0 || s->scene->n_faces == 0 || s->scene->n_meshes == 0) return true;
  emission::system_t* emitter = s->params.emitter;
  if (!hip_parallel_rays_emission::is_hip(emitter)) return false;
  const hip_parallel_rays_emission::state* em = hip_parallel_rays_emission::state_of(emitter);
  return hip_ray_caster::is_hip(em->caster) && hip_ray_caster::method_of(em->caster) == FFK_CAST_GRID;
}

// The first half of calculate's fused branch, launched when the step starts:
// the parallel rays' hit counts depend on the scene and the ray key only, so
// they are counted on this system's own stream while the solution's stream
// runs the equations before this one.
int prelaunch_dev(system* sys, const dev_task& t) {
  state* s = sys->st;
  s->counting = false;
  if (s->params.source == 0) return THERMAL_EQUATION_OK;
  const subject::scene_t* scene = s->scene;
  if (scene->n_faces == 0 || scene->n_meshes == 0) return THERMAL_EQUATION_OK;
  const parallel_rays_cpu::source_t source = s->params.source(s->params.source_param);
  const float radius = s->bs_radius * 1.1f;
  const float distance = 1.1f + radius;
  const int n_rays = s->params.n_rays;
  const int nm = scene->n_meshes;
  if (!s->grid) {
    if (ffk_scene_build_grid(s->s, 0, 0, 0, t.st) < 0) return -THERMAL_EQUATION_ERROR;
    s->grid = true;
  }
  if (!s->dev_ready) {
    if (!s->d_absorb2.reserve(2 * (size_t)nm) || !s->d_counts.reserve(2 * (size_t)nm) ||
        hipMemcpy(s->d_absorb2.p, s->absorb2.data(), 8 * (size_t)nm, hipMemcpyHostToDevice) != hipSuccess)
      return -THERMAL_EQUATION_ERROR;
    if ((!s->ev_start && hipEventCreateWithFlags(&s->ev_start, hipEventDisableTiming) != hipSuccess) ||
        (!s->ev_counted && hipEventCreateWithFlags(&s->ev_counted, hipEventDisableTiming) != hipSuccess))
      return -THERMAL_EQUATION_ERROR;
    s->dev_ready = true;
  }
  hip_parallel_rays_emission::state* em = hip_parallel_rays_emission::state_of(s->params.emitter);
  const unsigned long long key = hip_parallel_rays_emission::next_key(em);
  // the counts follow what the solution's stream did before (the grid, the
  // previous step's reading of d_counts)
  if (hipEventRecord(s->ev_start, t.st) != hipSuccess || hipStreamWaitEvent(s->st, s->ev_start, 0) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  if (ffk_sun_count_dev(s->s, &source.direction.x, s->center, distance, radius, 0, n_rays, key, s->d_counts.p,
                        s->st) < 0 ||
      hipEventRecord(s->ev_counted, s->st) != hipSuccess)
    return -THERMAL_EQUATION_ERROR;
  s->counted_power = source.power * (float(M_PI) * radius * radius) / n_rays;  // the reference's ray power
  s->counting = true;
  return THERMAL_EQUATION_OK;
}

// The second half, in the equations' order: the counts' powers added to the
// step's absorption on the solution's stream (counted here first when the
// step did not launch them ahead).
int calculate_dev(system* sys, const dev_task& t) {
  state* s = sys->st;
  if (!s->counting) {
    const int r = prelaunch_dev(sys, t);
    if (r < 0 || !s->counting) return r;
  }
  s->counting = false;
  if (hipStreamWaitEvent(t.st, s->ev_counted, 0) != hipSuccess) return -THERMAL_EQUATION_ERROR;
  return ffk_sun_deposit_dev(s->scene->n_meshes, s->d_counts.p, s->d_absorb2.p, s->counted_power, t.d_ab, t.st) < 0
             ? -THERMAL_EQUATION_ERROR
             : THERMAL_EQUATION_OK;
}

const thermal_equation::system_methods_t methods = {
    (int (*)(thermal_equation::system_t*, void*))&init,
    (int (*)(thermal_equation::system_t*))&shutdown,
    (int (*)(thermal_equation::system_t*, subject::scene_t*))&set_scene,
    (int (*)(thermal_equation::system_t*, thermal_equation::task_t*))&calculate,
};
thermal_equation::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  return s;
}

}  // namespace hip_distant_source

// ============================================== device-resident equations ==
// The equations of this library that can add to a step's arrays in device
// memory (dev_task), found by their method tables.
namespace dev_equations {

bool can(thermal_equation::system_t* e) {
  if (!e) return false;
  if (e->methods == &hip_sb_ff::methods) return hip_sb_ff::can_dev((hip_sb_ff::system*)e);
  if (e->methods == &hip_radiance::methods) return hip_radiance::can_dev((hip_radiance::system*)e);
  if (e->methods == &hip_conduction::methods) return hip_conduction::can_dev((hip_conduction::system*)e);
  if (e->methods == &hip_heat_source::methods) return hip_heat_source::can_dev((hip_heat_source::system*)e);
  if (e->methods == &hip_distant_source::methods)
    return hip_distant_source::can_dev((hip_distant_source::system*)e);
  return false;
}

int run(thermal_equation::system_t* e, const dev_task& t) {
  if (e->methods == &hip_sb_ff::methods) return hip_sb_ff::calculate_dev((hip_sb_ff::system*)e, t);
  if (e->methods == &hip_radiance::methods) return hip_radiance::calculate_dev((hip_radiance::system*)e, t);
  if (e->methods == &hip_conduction::methods) return hip_conduction::calculate_dev((hip_conduction::system*)e, t);
  if (e->methods == &hip_heat_source::methods)
    return hip_heat_source::calculate_dev((hip_heat_source::system*)e, t);
  if (e->methods == &hip_distant_source::methods)
    return hip_distant_source::calculate_dev((hip_distant_source::system*)e, t);
  return -THERMAL_EQUATION_ERROR;
}

// Work an equation can launch when the step starts, ahead of the equations
// before it (on its own stream; run() then joins it in order).
int prelaunch(thermal_equation::system_t* e, const dev_task& t) {
  if (e->methods == &hip_distant_source::methods)
    return hip_distant_source::prelaunch_dev((hip_distant_source::system*)e, t);
  return THERMAL_EQUATION_OK;
}

// Checks made once the step's stream has drained.
int status(thermal_equation::system_t* e) {
  if (e->methods == &hip_radiance::methods) return hip_radiance::status_dev((hip_radiance::system*)e);
  return THERMAL_EQUATION_OK;
}

}  // namespace dev_equations

// ==================================================== Adams solver (FP64) ==
namespace hip_adams {

struct state {
  thermal_solution::params_t params = {0, nullptr};
  subject::scene_t* scene = nullptr;
  int n_step = 0;
  std::vector<double> C, T, Tn;
  std::vector<float> Tf, em, ab;
  dbuf<double> d;  // C, em, ab, E ring (5N), T_n, T_next
  dbuf<float> f;   // device-resident step: T, emission, absorption (float, N each)
  bool dev_loop = true;
  hipStream_t st = nullptr;
};
struct system : thermal_solution::system_t {
  state* st;
};

int init(system* sys, thermal_solution::params_t* params) {
  sys->st = new (std::nothrow) state();
  if (!sys->st) return -THERMAL_SOLUTION_ERROR;
  sys->st->params = *params;
  sys->st->st = make_stream();
  return sys->st->st ? THERMAL_SOLUTION_OK : -THERMAL_SOLUTION_ERROR;
}
int shutdown(system* sys) {
  state* s = sys->st;
  if (!s) return THERMAL_SOLUTION_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  s->d.release();
  s->f.release();
  drop_stream(s->st);
  delete s;
  sys->st = nullptr;
  return THERMAL_SOLUTION_OK;
}

// thermal_solution/cpu_adams.cpp:62-85
int set_scene(system* sys, subject::scene_t* scene, float* temperatures) {
  state* s = sys->st;
  s->scene = scene;
  int r;
  for (int i = 0; i != s->params.n_equations; ++i) {
    thermal_equation::system_t* eq = s->params.equations[i];
    if ((r = eq->methods->set_scene(eq, scene)) < 0) return r;
  }
  const int N = scene->n_meshes;
  s->n_step = 0;
  s->C.assign(N, 0.0);
  for (int m = 0; m < N; ++m) {
    float a = 0;
    for (int f = 0; f != scene->meshes[m].n_faces; ++f) a += area_of(scene->faces[scene->meshes[m].first_idx + f]);
    const subject::shell_properties_t& sh = scene->materials[scene->meshes[m].material_idx].shell;
    s->C[m] = (double)a * sh.thickness * sh.density * sh.heat_capacity;  // cpu_adams.cpp:127-129
  }
  s->T.assign(temperatures, temperatures + N);
  // the equations set their scenes up on their own streams; the steps run on
  // this one
  if (hipDeviceSynchronize() != hipSuccess) return -THERMAL_SOLUTION_ERROR;
  if (!s->d.reserve(10 * (size_t)N)) return -THERMAL_SOLUTION_ERROR;
  if (hipMemsetAsync(s->d.p, 0, 80 * (size_t)N, s->st) != hipSuccess ||
      hipMemcpyAsync(s->d.p, s->C.data(), 8 * N, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(s->d.p + 8 * (size_t)N, s->T.data(), 8 * N, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipStreamSynchronize(s->st) != hipSuccess)
    return -THERMAL_SOLUTION_ERROR;
  return THERMAL_SOLUTION_OK;
}

// thermal_solution/cpu_adams.cpp:107-182 with the update in FP64: the
// equations add to host arrays, which go to the device for the update.
int calculate_host(system* sys, thermal_solution::task_t* task) {
  state* s = sys->st;
  const int N = s->scene->n_meshes;
  s->Tf.resize(N);
  for (int m = 0; m < N; ++m) s->Tf[m] = (float)s->T[m];
  s->em.assign(N, 0.f);
  s->ab.assign(N, 0.f);
  thermal_equation::task_t et = {s->Tf.data(), s->em.data(), s->ab.data()};
  int r;
  for (int i = 0; i != s->params.n_equations; ++i) {
    thermal_equation::system_t* eq = s->params.equations[i];
    if ((r = eq->methods->calculate(eq, &et)) < 0) return r;
  }
  std::vector<double> emd(s->em.begin(), s->em.end()), abd(s->ab.begin(), s->ab.end());
  double* d = s->d.p;
  double *dC = d, *dem = d + N, *dab = d + 2 * (size_t)N, *dE = d + 3 * (size_t)N, *dT = d + 8 * (size_t)N,
         *dTn = d + 9 * (size_t)N;
  if (hipMemcpyAsync(dem, emd.data(), 8 * N, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(dab, abd.data(), 8 * N, hipMemcpyHostToDevice, s->st) != hipSuccess)
    return -THERMAL_SOLUTION_ERROR;
  if (ffk_adams_step(N, s->n_step, task->time_delta, dC, dem, dab, dE, dT, dTn, nullptr, s->st) < 0)
    return -THERMAL_SOLUTION_ERROR;
  if (hipMemcpyAsync(dT, dTn, 8 * N, hipMemcpyDeviceToDevice, s->st) != hipSuccess ||
      hipMemcpyAsync(s->T.data(), dTn, 8 * N, hipMemcpyDeviceToHost, s->st) != hipSuccess ||
      hipStreamSynchronize(s->st) != hipSuccess)
    return -THERMAL_SOLUTION_ERROR;
  for (int m = 0; m < N; ++m) task->temperatures[m] = (float)s->T[m];
  task->n_step = ++s->n_step;
  return THERMAL_SOLUTION_OK;
}

// The same step with the equations' arrays in device memory: the step's
// float temperatures, emission and absorption stay on the device from the
// equations to the FP64 update, and one copy brings the new temperatures
// back for the task (radiance keeps its weights and counts on the host,
// from the host temperatures). The arithmetic is calculate_host's, so the
// temperatures are its bits. A failed step leaves the solution's state as
// it was.
int calculate_dev(system* sys, thermal_solution::task_t* task) {
  state* s = sys->st;
  const int N = s->scene->n_meshes;
  s->Tf.resize(N);
  for (int m = 0; m < N; ++m) s->Tf[m] = (float)s->T[m];
  if (!s->f.reserve(3 * (size_t)N)) return -THERMAL_SOLUTION_ERROR;
  double* d = s->d.p;
  double *dC = d, *dE = d + 3 * (size_t)N, *dT = d + 8 * (size_t)N, *dTn = d + 9 * (size_t)N;
  float *dTf = s->f.p, *dem = dTf + N, *dab = dTf + 2 * (size_t)N;
  if (ffk_step_begin_dev(N, dT, dTf, dem, dab, s->st) < 0) return -THERMAL_SOLUTION_ERROR;
  const dev_task t = {s->Tf.data(), dTf, dem, dab, s->st};
  int r = THERMAL_SOLUTION_OK;
  for (int i = 0; i != s->params.n_equations && r >= 0; ++i) r = dev_equations::prelaunch(s->params.equations[i], t);
  for (int i = 0; i != s->params.n_equations && r >= 0; ++i) r = dev_equations::run(s->params.equations[i], t);
  if (r >= 0 && ffk_adams_step_f(N, s->n_step, task->time_delta, dC, dem, dab, dE, dT, dTn, nullptr, s->st) < 0)
    r = -THERMAL_SOLUTION_ERROR;
  s->Tn.resize(N);
  if (r >= 0 && hipMemcpyAsync(s->Tn.data(), dTn, 8 * (size_t)N, hipMemcpyDeviceToHost, s->st) != hipSuccess)
    r = -THERMAL_SOLUTION_ERROR;
  if (hipStreamSynchronize(s->st) != hipSuccess && r >= 0) r = -THERMAL_SOLUTION_ERROR;
  for (int i = 0; i != s->params.n_equations; ++i) {  // every pending check is read
    const int q = dev_equations::status(s->params.equations[i]);
    if (r >= 0 && q < 0) r = q;
  }
  if (r < 0) return r;
  if (hipMemcpyAsync(dT, dTn, 8 * (size_t)N, hipMemcpyDeviceToDevice, s->st) != hipSuccess)
    return -THERMAL_SOLUTION_ERROR;
  s->T.swap(s->Tn);
  for (int m = 0; m < N; ++m) task->temperatures[m] = (float)s->T[m];
  task->n_step = ++s->n_step;
  return THERMAL_SOLUTION_OK;
}

// The device-resident step when every equation has a device form (and the
// solution has not been told otherwise), else the host-array step.
int calculate(system* sys, thermal_solution::task_t* task) {
  state* s = sys->st;
  bool dev = s->dev_loop && s->scene && s->scene->n_meshes > 0;
  for (int i = 0; dev && i != s->params.n_equations; ++i) dev = dev_equations::can(s->params.equations[i]);
  return dev ? calculate_dev(sys, task) : calculate_host(sys, task);
}

const thermal_solution::system_methods_t methods = {
    (int (*)(thermal_solution::system_t*, thermal_solution::params_t*))&init,
    (int (*)(thermal_solution::system_t*))&shutdown,
    (int (*)(thermal_solution::system_t*, subject::scene_t*, float*))&set_scene,
    (int (*)(thermal_solution::system_t*, thermal_solution::task_t*))&calculate,
};
thermal_solution::system_t* system_create() {
  system* s = (system*)malloc(sizeof(system));
  if (!s) return nullptr;
  s->methods = &methods;
  s->st = nullptr;
  return s;
}

}  // namespace hip_adams

// ============================================================ C binding ==
extern "C" {

void* ffp_ray_caster_create(int type) {
  ray_caster::system_t* s = nullptr;
  if (type == RAY_CASTER_NAIVE_HIP) s = hip_ray_caster::system_create_naive();
  else if (type == RAY_CASTER_GRID_HIP) s = hip_ray_caster::system_create_grid();
  if (!s) return nullptr;
  if (s->methods->init(s) < 0) {
    s->methods->shutdown(s);
    free(s);
    return nullptr;
  }
  return s;
}
int ffp_ray_caster_set_scene(void* s, ffp_rc_scene* scene) {
  ray_caster::system_t* r = (ray_caster::system_t*)s;
  return r->methods->set_scene(r, (ray_caster::scene_t*)scene);
}
int ffp_ray_caster_prepare(void* s) {
  ray_caster::system_t* r = (ray_caster::system_t*)s;
  return r->methods->prepare(r);
}
int ffp_ray_caster_cast(void* s, ffp_rc_task* task) {
  ray_caster::system_t* r = (ray_caster::system_t*)s;
  return r->methods->cast(r, (ray_caster::task_t*)task);
}
void ffp_ray_caster_free(void* s) {
  if (!s) return;
  ray_caster::system_t* r = (ray_caster::system_t*)s;
  r->methods->shutdown(r);
  free(r);
}

void* ffp_emission_create(int type, void* caster) {
  emission::system_t* s = nullptr;
  if (type == EMISSION_MALLEY_HIP) s = hip_malley_emission::system_create();
  else if (type == EMISSION_PARALLEL_RAYS_HIP) s = hip_parallel_rays_emission::system_create();
  if (!s) return nullptr;
  if (s->methods->init(s, (ray_caster::system_t*)caster) < 0) {
    s->methods->shutdown(s);
    free(s);
    return nullptr;
  }
  return s;
}
int ffp_emission_set_scene(void* s, ffp_rc_scene* scene) {
  emission::system_t* e = (emission::system_t*)s;
  return e->methods->set_scene(e, (ray_caster::scene_t*)scene);
}
int ffp_emission_calculate_malley(void* s, int n_rays, float total_weight, float* weights, ffp_rc_task** rays_out) {
  emission::system_t* e = (emission::system_t*)s;
  malley_emission::task_t task(n_rays, total_weight, weights);
  int r = e->methods->calculate(e, &task);
  *rays_out = (ffp_rc_task*)task.rays;
  return r;
}
int ffp_emission_calculate_parallel(void* s, int n_rays, const float* dir3, const float* center3, float distance,
                                    float radius, ffp_rc_task** rays_out) {
  emission::system_t* e = (emission::system_t*)s;
  parallel_rays_emission::task_t task(n_rays);
  task.direction = {dir3[0], dir3[1], dir3[2]};
  task.center = {center3[0], center3[1], center3[2]};
  task.distance = distance;
  task.radius = radius;
  int r = e->methods->calculate(e, &task);
  *rays_out = (ffp_rc_task*)task.rays;
  return r;
}
int ffp_emission_set_seed(void* s, unsigned long long seed, int frame) {
  emission::system_t* e = (emission::system_t*)s;
  if (hip_parallel_rays_emission::is_hip(e)) {  // no emission frame
    hip_parallel_rays_emission::state* st = hip_parallel_rays_emission::state_of(e);
    st->seed = seed;
    st->calls = 0;
    return EMISSION_OK;
  }
  if (!hip_malley_emission::is_hip(e)) return -EMISSION_ERROR;
  if (frame != FFK_FRAME_REFERENCE && frame != FFK_FRAME_ORTHONORMAL) return -EMISSION_ERROR;
  hip_malley_emission::state* st = hip_malley_emission::state_of(e);
  st->seed = seed;
  st->calls = 0;
  if (st->frame != frame) {
    st->frame = frame;
    st->front.clear();  // rebuild the layout with the new frame
  }
  return EMISSION_OK;
}
void ffp_emission_free(void* s) {
  if (!s) return;
  emission::system_t* e = (emission::system_t*)s;
  e->methods->shutdown(e);
  free(e);
}
void ffp_task_free(ffp_rc_task* t) { rc_task_free((ray_caster::task_t*)t); }

void* ffp_form_factors_create(int type, void* emitter) {
  if (type != FORM_FACTORS_HIP) return nullptr;
  form_factors::system_t* s = hip_form_factors::system_create();
  if (!s) return nullptr;
  if (s->methods->init(s, (emission::system_t*)emitter) < 0) {
    s->methods->shutdown(s);
    free(s);
    return nullptr;
  }
  return s;
}
int ffp_form_factors_set_scene(void* s, ffp_ff_scene* scene) {
  form_factors::system_t* f = (form_factors::system_t*)s;
  return f->methods->set_scene(f, (form_factors::scene_t*)scene);
}
int ffp_form_factors_prepare(void* s) {
  form_factors::system_t* f = (form_factors::system_t*)s;
  return f->methods->prepare(f);
}
int ffp_form_factors_calculate(void* s, ffp_ff_task* task) {
  form_factors::system_t* f = (form_factors::system_t*)s;
  return f->methods->calculate(f, (form_factors::task_t*)task);
}
void ffp_form_factors_free(void* s) {
  if (!s) return;
  form_factors::system_t* f = (form_factors::system_t*)s;
  f->methods->shutdown(f);
  free(f);
}

void* ffp_thermal_equation_create(int type, void* dep, int n_rays, int f_dtype) {
  thermal_equation::system_t* s = nullptr;
  if (type == THERMAL_EQUATION_SB_FF_HIP) {
    s = hip_sb_ff::system_create();
    if (!s) return nullptr;
    hip_sb_ff::params_t p = {(form_factors::system_t*)dep, n_rays, f_dtype};
    if (s->methods->init(s, &p) < 0) {
      s->methods->shutdown(s);
      free(s);
      return nullptr;
    }
  } else if (type == THERMAL_EQUATION_CONDUCTIVE_HIP) {
    s = hip_conduction::system_create();
    if (!s) return nullptr;
    if (s->methods->init(s, nullptr) < 0) {
      s->methods->shutdown(s);
      free(s);
      return nullptr;
    }
  } else if (type == THERMAL_EQUATION_RADIANCE_HIP) {
    if (!dep) return nullptr;
    s = hip_radiance::system_create();
    if (!s) return nullptr;
    radiance_equation::params_t p = {(emission::system_t*)dep, n_rays};
    if (s->methods->init(s, &p) < 0) {
      s->methods->shutdown(s);
      free(s);
      return nullptr;
    }
  }
  return s;
}
void* ffp_heat_source_create(int n_sources, ffp_heat_source* sources) {
  if (n_sources < 0 || (n_sources > 0 && !sources)) return nullptr;
  thermal_equation::system_t* s = hip_heat_source::system_create();
  if (!s) return nullptr;
  heat_source_equation::params_t p = {n_sources, (heat_source_equation::heat_source_t*)sources};
  if (s->methods->init(s, &p) < 0) {
    s->methods->shutdown(s);
    free(s);
    return nullptr;
  }
  return s;
}

namespace {
parallel_rays_cpu::source_t read_source(void* p) { return *(parallel_rays_cpu::source_t*)p; }
}  // namespace

void* ffp_distant_source_create(void* emitter, int n_rays, ffp_source* source) {
  if (!emitter || n_rays <= 0) return nullptr;
  thermal_equation::system_t* s = hip_distant_source::system_create();
  if (!s) return nullptr;
  // No source: the equation adds nothing (parallel_rays_cpu.cpp:94-95).
  parallel_rays_cpu::params_t p = {n_rays, (emission::system_t*)emitter, source, source ? &read_source : nullptr};
  if (s->methods->init(s, &p) < 0) {
    s->methods->shutdown(s);
    free(s);
    return nullptr;
  }
  return s;
}

int ffp_thermal_equation_set_scene(void* s, ffp_subject_scene* scene) {
  thermal_equation::system_t* e = (thermal_equation::system_t*)s;
  return e->methods->set_scene(e, (subject::scene_t*)scene);
}
int ffp_thermal_equation_calculate(void* s, ffp_te_task* task) {
  thermal_equation::system_t* e = (thermal_equation::system_t*)s;
  return e->methods->calculate(e, (thermal_equation::task_t*)task);
}
void ffp_thermal_equation_free(void* s) {
  if (!s) return;
  thermal_equation::system_t* e = (thermal_equation::system_t*)s;
  e->methods->shutdown(e);
  free(e);
}

void* ffp_thermal_solution_create(int type, int n_equations, void** equations) {
  if (type != THERMAL_SOLUTION_HIP_ADAMS) return nullptr;
  thermal_solution::system_t* s = hip_adams::system_create();
  if (!s) return nullptr;
  thermal_solution::params_t p = {n_equations, (thermal_equation::system_t**)equations};
  if (s->methods->init(s, &p) < 0) {
    s->methods->shutdown(s);
    free(s);
    return nullptr;
  }
  return s;
}
int ffp_thermal_solution_set_scene(void* s, ffp_subject_scene* scene, float* T) {
  thermal_solution::system_t* t = (thermal_solution::system_t*)s;
  return t->methods->set_scene(t, (subject::scene_t*)scene, T);
}
int ffp_thermal_solution_calculate(void* s, ffp_ts_task* task) {
  thermal_solution::system_t* t = (thermal_solution::system_t*)s;
  return t->methods->calculate(t, (thermal_solution::task_t*)task);
}
int ffp_thermal_solution_set_device_loop(void* s, int on) {
  thermal_solution::system_t* t = (thermal_solution::system_t*)s;
  if (!t || t->methods != &hip_adams::methods || !((hip_adams::system*)t)->st) return -THERMAL_SOLUTION_ERROR;
  ((hip_adams::system*)t)->st->dev_loop = on != 0;
  return THERMAL_SOLUTION_OK;
}
void ffp_thermal_solution_free(void* s) {
  if (!s) return;
  thermal_solution::system_t* t = (thermal_solution::system_t*)s;
  t->methods->shutdown(t);
  free(t);
}

}  // extern "C"
