// Launch-shape autotuner for the form-factor path.
//
// The grid resolution decides how many cells a ray steps through against how
// many triangles it tests per cell; the best trade depends on the scene
// (triangle sizes, empty space, ray coherence). The tuner casts a fixed
// sample of the scene's own Malley rays (the emission layout scaled down to
// about `sample_rays`, same faces and directions) over a ladder of
// resolutions, keeps the fastest, then times the fused kernel's block size
// on the full layout when a run is short. The choice is cached per scene
// signature (faces, meshes, ray counts, frame) for the life of the process.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "ffk.h"

namespace {

struct tuned {
  int dims[3];
  int block;
};

std::mutex g_mu;
std::map<unsigned long long, tuned> g_cache;

unsigned long long fnv1a(const void* p, size_t n, unsigned long long h) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

// ms of fn() on stream st, best of `reps`.
template <class F>
int time_ms(hipStream_t st, int reps, F fn, float* best) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return FFK_ERR_HIP;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return FFK_ERR_HIP;
  }
  int r = FFK_OK;
  *best = 3.0e38f;
  for (int k = 0; k < reps && r == FFK_OK; ++k) {
    (void)hipEventRecord(a, st);
    r = fn();
    (void)hipEventRecord(b, st);
    if (r == FFK_OK && hipEventSynchronize(b) == hipSuccess) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, a, b);
      *best = std::min(*best, ms);
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return r;
}

void dims_for(const float lo[3], const float hi[3], double cells_per_face, int n_faces, int dims[3]) {
  double ext[3], vol = 1.0;
  for (int a = 0; a < 3; ++a) {
    ext[a] = std::max(1e-12, (double)hi[a] - (double)lo[a]);
    vol *= ext[a];
  }
  const double cell = std::cbrt(vol / std::max(1.0, cells_per_face * n_faces));
  for (int a = 0; a < 3; ++a) dims[a] = (int)std::min(512.0, std::max(1.0, std::ceil(ext[a] / cell)));
}

}  // namespace

extern "C" {

int ffk_autotune(ffk_scene* s, long long sample_rays, int* out_dims3, int* out_block, void* stream) {
  if (!s || sample_rays <= 0) return FFK_ERR_ARG;
  int nf = 0, nm = 0;
  long long n_rays = 0;
  if (ffk_scene_counts(s, &nf, &nm, &n_rays) < 0 || nf == 0 || nm == 0 || n_rays == 0) return FFK_ERR_ARG;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  std::vector<float> faces(9 * (size_t)nf);
  std::vector<int> meshes(3 * (size_t)nm), front(nf), rear(nf);
  int frame = 0;
  if (ffk_scene_layout(s, faces.data(), meshes.data(), front.data(), rear.data(), &frame) < 0) return FFK_ERR_ARG;
  unsigned long long key = 1469598103934665603ull;
  key = fnv1a(faces.data(), faces.size() * 4, key);
  key = fnv1a(meshes.data(), meshes.size() * 4, key);
  key = fnv1a(front.data(), front.size() * 4, key);
  key = fnv1a(rear.data(), rear.size() * 4, key);
  key = fnv1a(&frame, sizeof(frame), key);
  tuned t;
  bool hit = false;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      t = it->second;
      hit = true;
    }
  }
  int r;
  if (!hit) {
    // 1. a representative sample: the layout scaled to about sample_rays
    const double scale = std::min(1.0, (double)sample_rays / (double)n_rays);
    std::vector<int> sf(nf), sr(nf);
    for (int f = 0; f < nf; ++f) {
      sf[f] = front[f] ? std::max(1, (int)std::lround(front[f] * scale)) : 0;
      sr[f] = rear[f] ? std::max(1, (int)std::lround(rear[f] * scale)) : 0;
    }
    long long n_sample = 0;
    if ((r = ffk_scene_set_emission(s, sf.data(), sr.data(), frame, &n_sample, stream)) < 0) return r;
    float* d_rays = nullptr;
    int* d_idx = nullptr;
    if (hipMalloc(&d_rays, 24 * (size_t)n_sample) != hipSuccess ||
        hipMalloc(&d_idx, 4 * (size_t)n_sample) != hipSuccess) {
      (void)hipFree(d_rays);
      (void)ffk_scene_set_emission(s, front.data(), rear.data(), frame, nullptr, stream);
      return FFK_ERR_NOMEM;
    }
    r = ffk_emit_malley(s, 0x7E57ull, d_rays, stream);
    // 2. grid resolution ladder (cells per face), casting the sample
    int best_dims[3] = {0, 0, 0};
    float best = 3.0e38f;
    if (r == FFK_OK) r = ffk_scene_build_grid(s, 0, 0, 0, stream);
    float lo[3], hi[3];
    int d0[3];
    long long nr = 0;
    if (r == FFK_OK) r = ffk_scene_grid_info(s, d0, lo, hi, &nr);
    static const double ladder[] = {0.25, 0.5, 1.0, 2.0, 4.0, 8.0, 16.0};
    int prev[3] = {-1, -1, -1};
    for (double cpf : ladder) {
      if (r != FFK_OK) break;
      int dims[3];
      dims_for(lo, hi, cpf, nf, dims);
      if (std::equal(dims, dims + 3, prev)) continue;
      std::copy(dims, dims + 3, prev);
      if ((r = ffk_scene_build_grid(s, dims[0], dims[1], dims[2], stream)) < 0) break;
      float ms = 0.f;
      r = time_ms(st, 2, [&] { return ffk_cast(s, FFK_CAST_GRID, d_rays, n_sample, d_idx, nullptr, nullptr, stream); },
                  &ms);
      if (r == FFK_OK && ms < best) {
        best = ms;
        std::copy(dims, dims + 3, best_dims);
      }
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_rays);
    (void)hipFree(d_idx);
    const int r2 = ffk_scene_set_emission(s, front.data(), rear.data(), frame, nullptr, stream);
    if (r == FFK_OK) r = r2;
    if (r < 0) return r;
    std::copy(best_dims, best_dims + 3, t.dims);
    if ((r = ffk_scene_build_grid(s, t.dims[0], t.dims[1], t.dims[2], stream)) < 0) return r;
    // 3. block size of the fused kernel on the full layout, when a run is short
    // (estimated from the sample) and F is small enough to allocate here
    t.block = 0;
    const double est_ms = best * (double)n_rays / (double)std::max(1LL, n_sample);
    const size_t nn = (size_t)nm * nm;
    if (est_ms < 250.0 && nn * 12 < (size_t(1) << 31)) {
      float* d_F = nullptr;
      unsigned long long* d_tally = nullptr;
      if (hipMalloc(&d_F, 4 * nn) == hipSuccess && hipMalloc(&d_tally, 8 * nn) == hipSuccess) {
        float best_b = 3.0e38f;
        for (int block : {256, 512, 1024}) {
          if ((r = ffk_set_tuning(s, block, 0, -1)) < 0) break;
          float ms = 0.f;
          r = time_ms(st, 2, [&] { return ffk_ff_fused(s, FFK_CAST_GRID, 1, d_tally, d_F, nullptr, stream); }, &ms);
          if (r < 0) break;
          if (ms < best_b) {
            best_b = ms;
            t.block = block;
          }
        }
        (void)ffk_set_tuning(s, 0, 0, -1);
      }
      (void)hipFree(d_F);
      (void)hipFree(d_tally);
      if (r < 0) return r;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = t;
  } else {
    if ((r = ffk_scene_build_grid(s, t.dims[0], t.dims[1], t.dims[2], stream)) < 0) return r;
  }
  if ((r = ffk_set_tuning(s, t.block, 0, -1)) < 0) return r;
  if (out_dims3) std::copy(t.dims, t.dims + 3, out_dims3);
  if (out_block) *out_block = t.block;
  return hit ? 1 : FFK_OK;
}

}  // extern "C"
