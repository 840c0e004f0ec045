// Launch-shape autotuner for the form-factor path.
//
// The grid resolution decides how many cells a ray steps through against how
// many triangles it tests per cell; the best trade depends on the scene
// (triangle sizes, empty space, ray coherence). The tuner casts a fixed
// sample of the scene's own Malley rays (the emission layout scaled down to
// about `sample_rays`, same faces and directions) over a ladder of
// resolutions with the scene's ray scheduling. When a full run is short
// (estimated from the sample) and F fits, the three fastest resolutions are
// then timed with the fused kernel itself on the full layout, and the
// fused kernel's block size, walk mode and ray scheduling (refill threshold,
// pooled-test ratio) after that. The choice is cached per scene
// signature (faces, meshes, ray counts, frame) for the life of the process.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
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
  int walk;
  int refill;
  int pool;
};

// (refill threshold, pooled-test ratio) pairs the tuner times: the default
// first, then schedules that start rays sooner and pool tests later, which
// suit scenes where rays meet many occupied cells.
constexpr int kSchedules[][2] = {{32, 2}, {24, 3}, {16, 4}};

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

int ffk_autotune(ffk_scene* s, long long sample_rays, int* out_dims3, int* out_block, int* out_walk, void* stream) {
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
    struct trial {
      float ms;
      double cpf;
      int dims[3];
    };
    std::vector<trial> trials;
    if (r == FFK_OK) r = ffk_scene_build_grid(s, 0, 0, 0, stream);
    float lo[3], hi[3];
    int d0[3];
    long long nr = 0;
    if (r == FFK_OK) r = ffk_scene_grid_info(s, d0, lo, hi, &nr);
    static const double ladder[] = {0.25, 0.5, 1.0, 2.0, 4.0, 8.0, 16.0};
    int prev[3] = {-1, -1, -1};
    for (double cpf : ladder) {
      if (r != FFK_OK) break;
      trial tr;
      tr.cpf = cpf;
      dims_for(lo, hi, cpf, nf, tr.dims);
      if (std::equal(tr.dims, tr.dims + 3, prev)) continue;
      std::copy(tr.dims, tr.dims + 3, prev);
      if ((r = ffk_scene_build_grid(s, tr.dims[0], tr.dims[1], tr.dims[2], stream)) < 0) break;
      r = time_ms(st, 2, [&] { return ffk_cast(s, FFK_CAST_GRID, d_rays, n_sample, d_idx, nullptr, nullptr, stream); },
                  &tr.ms);
      if (r == FFK_OK) trials.push_back(tr);
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_rays);
    (void)hipFree(d_idx);
    const int r2 = ffk_scene_set_emission(s, front.data(), rear.data(), frame, nullptr, stream);
    if (r == FFK_OK) r = r2;
    if (r < 0) return r;
    if (trials.empty()) return FFK_ERR_ARG;
    std::stable_sort(trials.begin(), trials.end(), [](const trial& a, const trial& b) { return a.ms < b.ms; });
    std::copy(trials[0].dims, trials[0].dims + 3, t.dims);
    // 3. the fused kernel on the full layout, when a run is short (estimated
    // from the sample) and F (plus the tally, if the scene needs one) takes
    // under half of the free device memory: the two fastest resolutions and
    // the resolutions a factor sqrt(2) in cells either side of the fastest,
    // then the block size
    t.block = 0;
    t.walk = 0;
    if ((r = ffk_get_refill(s, &t.refill, &t.pool)) < 0) return r;
    const double est_ms = trials[0].ms * (double)n_rays / (double)std::max(1LL, n_sample);
    const size_t nn = (size_t)nm * nm;
    const bool need_tally = ffk_ff_needs_tally(s) != 0;
    const size_t scratch = 4 * nn + (need_tally ? 8 * nn : 0);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    bool fused_timed = false;
    if (est_ms < 250.0 && scratch <= free_b / 2) {
      float* d_F = nullptr;
      unsigned long long* d_tally = nullptr;
      if (hipMalloc(&d_F, 4 * nn) == hipSuccess && (!need_tally || hipMalloc(&d_tally, 8 * nn) == hipSuccess)) {
        fused_timed = true;
        std::vector<std::array<int, 3>> cand;
        auto add = [&](const int* d) {
          std::array<int, 3> a = {d[0], d[1], d[2]};
          if (std::find(cand.begin(), cand.end(), a) == cand.end()) cand.push_back(a);
        };
        add(trials[0].dims);
        if (trials.size() > 1) add(trials[1].dims);
        for (double f : {std::sqrt(2.0), std::sqrt(0.5)}) {
          int d[3];
          dims_for(lo, hi, trials[0].cpf * f, nf, d);
          add(d);
        }
        float best_g = 3.0e38f;
        for (size_t k = 0; k < cand.size() && r == FFK_OK; ++k) {
          const int* dims = cand[k].data();
          if ((r = ffk_scene_build_grid(s, dims[0], dims[1], dims[2], stream)) < 0) break;
          float ms = 0.f;
          r = time_ms(st, 2, [&] { return ffk_ff_fused(s, FFK_CAST_GRID, 1, d_tally, d_F, nullptr, stream); }, &ms);
          if (r == FFK_OK && ms < best_g) {
            best_g = ms;
            std::copy(dims, dims + 3, t.dims);
          }
        }
        if (r == FFK_OK) r = ffk_scene_build_grid(s, t.dims[0], t.dims[1], t.dims[2], stream);
        float best_b = 3.0e38f;
        for (int block : {256, 512, 1024}) {
          if (r != FFK_OK) break;
          if ((r = ffk_set_tuning(s, block, 0, -1)) < 0) break;
          float ms = 0.f;
          r = time_ms(st, 2, [&] { return ffk_ff_fused(s, FFK_CAST_GRID, 1, d_tally, d_F, nullptr, stream); }, &ms);
          if (r < 0) break;
          if (ms < best_b) {
            best_b = ms;
            t.block = block;
          }
        }
        float best_w = 3.0e38f;
        for (int walk : {0, 1}) {
          if (r != FFK_OK) break;
          if ((r = ffk_set_tuning(s, t.block, 0, -1)) < 0 || (r = ffk_set_walk(s, walk)) < 0) break;
          float ms = 0.f;
          r = time_ms(st, 2, [&] { return ffk_ff_fused(s, FFK_CAST_GRID, 1, d_tally, d_F, nullptr, stream); }, &ms);
          if (r == FFK_OK && ms < best_w) {
            best_w = ms;
            t.walk = walk;
          }
        }
        float best_s = 3.0e38f;
        const int refill0 = t.refill, pool0 = t.pool;
        for (const auto& sch : kSchedules) {
          if (r != FFK_OK) break;
          if ((r = ffk_set_walk(s, t.walk)) < 0 || (r = ffk_set_refill(s, sch[0], sch[1])) < 0) break;
          float ms = 0.f;
          r = time_ms(st, 2, [&] { return ffk_ff_fused(s, FFK_CAST_GRID, 1, d_tally, d_F, nullptr, stream); }, &ms);
          if (r == FFK_OK && ms < best_s) {
            best_s = ms;
            t.refill = sch[0];
            t.pool = sch[1];
          }
        }
        (void)ffk_set_tuning(s, 0, 0, -1);
        (void)ffk_set_walk(s, 0);
        (void)ffk_set_refill(s, refill0, pool0);
      }
      (void)hipFree(d_F);
      (void)hipFree(d_tally);
      if (r < 0) return r;
    }
    if (!fused_timed && (r = ffk_scene_build_grid(s, t.dims[0], t.dims[1], t.dims[2], stream)) < 0) return r;
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = t;
  } else {
    if ((r = ffk_scene_build_grid(s, t.dims[0], t.dims[1], t.dims[2], stream)) < 0) return r;
  }
  if ((r = ffk_set_tuning(s, t.block, 0, -1)) < 0) return r;
  if ((r = ffk_set_walk(s, t.walk)) < 0) return r;
  if ((r = ffk_set_refill(s, t.refill, t.pool)) < 0) return r;
  if (out_dims3) std::copy(t.dims, t.dims + 3, out_dims3);
  if (out_block) *out_block = t.block;
  if (out_walk) *out_walk = t.walk;
  return hit ? 1 : FFK_OK;
}

}  // extern "C"
