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
// signature (faces, meshes, ray counts, frame) and device, in memory and in
// a text file (ffk_tune_cache_path), so that later processes skip the
// timing.
//
// The batch tuner (ffk_batch_autotune) times each LDS class of a batch alone
// over a sample of its tile queue, greedily over the walk mode, the ray
// scheduling and the block size, and keeps a change only when it wins by 3 %
// (repeated timings of one setting spread by 1-2 %); its results are cached
// the same way under the batch's signature.
#include <hip/hip_runtime_api.h>

#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
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
constexpr int kSchedules[][2] = {{32, 2}, {24, 3}, {16, 4}, {40, 3}, {40, 4}};

unsigned long long fnv1a(const void* p, size_t n, unsigned long long h) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

// ---- results cache: in memory, backed by a text file ---------------------------
// One line per result, "<kind> <key in hex> <n> <n integers>", appended as
// results are found; when the file is read, the last line of a key wins.
// kind "scene": dims0 dims1 dims2 block walk refill pool; kind "batch": per
// class id block refill pool walk.
std::mutex g_mu;
std::map<std::pair<std::string, unsigned long long>, std::vector<int>> g_results;
bool g_loaded = false;

std::string cache_file() {
  const char* env = std::getenv("FFK_TUNE_CACHE");
  if (env) return (!*env || !std::strcmp(env, "off")) ? std::string() : std::string(env);
  const char* xdg = std::getenv("XDG_CACHE_HOME");
  const char* home = std::getenv("HOME");
  std::string dir = xdg && *xdg ? std::string(xdg) : home && *home ? std::string(home) + "/.cache" : std::string();
  return dir.empty() ? std::string() : dir + "/form_factors_amd/autotune.txt";
}

// mkdir -p of the file's directory (best effort: without it there is no file cache)
void make_dirs(const std::string& file) {
  for (size_t pos = file.find('/', 1); pos != std::string::npos; pos = file.find('/', pos + 1))
    (void)mkdir(file.substr(0, pos).c_str(), 0755);
}

void load_locked() {
  if (g_loaded) return;
  g_loaded = true;
  const std::string path = cache_file();
  if (path.empty()) return;
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return;
  char kind[16];
  unsigned long long key;
  int n;
  while (std::fscanf(f, "%15s %llx %d", kind, &key, &n) == 3) {
    if (n < 0 || n > 4096) break;
    std::vector<int> v(n);
    bool ok = true;
    for (int i = 0; i < n && ok; ++i) ok = std::fscanf(f, "%d", &v[i]) == 1;
    if (!ok) break;
    g_results[{kind, key}] = v;
  }
  std::fclose(f);
}

bool cache_get(const char* kind, unsigned long long key, std::vector<int>& v) {
  std::lock_guard<std::mutex> lk(g_mu);
  load_locked();
  auto it = g_results.find({kind, key});
  if (it == g_results.end()) return false;
  v = it->second;
  return true;
}

void cache_put(const char* kind, unsigned long long key, const std::vector<int>& v) {
  std::lock_guard<std::mutex> lk(g_mu);
  load_locked();
  g_results[{kind, key}] = v;
  const std::string path = cache_file();
  if (path.empty()) return;
  make_dirs(path);
  FILE* f = std::fopen(path.c_str(), "a");
  if (!f) return;
  std::fprintf(f, "%s %016llx %d", kind, key, (int)v.size());
  for (int x : v) std::fprintf(f, " %d", x);
  std::fprintf(f, "\n");
  std::fclose(f);
}

// The current device's identity in cache keys: results depend on it.
unsigned long long device_key(unsigned long long h) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return h;
  h = fnv1a(prop.gcnArchName, std::strlen(prop.gcnArchName), h);
  return fnv1a(&prop.multiProcessorCount, sizeof(prop.multiProcessorCount), h);
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
  key = device_key(key);
  tuned t;
  bool hit = false;
  {
    std::vector<int> v;
    if (cache_get("scene", key, v) && v.size() == 7) {
      t = tuned{{v[0], v[1], v[2]}, v[3], v[4], v[5], v[6]};
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
    t.walk = FFK_WALK_DEFAULT;
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
        auto time_grid = [&](const int* dims, float* ms) {
          int e = ffk_scene_build_grid(s, dims[0], dims[1], dims[2], stream);
          if (e == FFK_OK)
            e = time_ms(st, 2, [&] { return ffk_ff_fused(s, FFK_CAST_GRID, 1, d_tally, d_F, nullptr, stream); }, ms);
          return e;
        };
        for (size_t k = 0; k < cand.size() && r == FFK_OK; ++k) {
          float ms = 0.f;
          r = time_grid(cand[k].data(), &ms);
          if (r == FFK_OK && ms < best_g) {
            best_g = ms;
            std::copy(cand[k].data(), cand[k].data() + 3, t.dims);
          }
        }
        // The fused kernel often gains from finer grids than the sample cast
        // ranks first (empty cells are cheap to cross, smaller cells mean
        // fewer tests), and its timings differ by a few per cent between
        // nearby resolutions: step a factor sqrt(2) finer than the finest
        // candidate, keeping a step only when it wins by 1 %, until a step
        // is 3 % slower than the best, after 4 steps, or past 2^22 cells (an
        // axis capped at 512 cells can repeat a resolution).
        double cpf = trials[0].cpf * std::sqrt(2.0);  // the finest candidate
        if (trials.size() > 1) cpf = std::max(cpf, trials[1].cpf);
        for (int step = 0; step < 4 && r == FFK_OK; ++step) {
          cpf *= std::sqrt(2.0);
          int d[3];
          dims_for(lo, hi, cpf, nf, d);
          if ((long long)d[0] * d[1] * d[2] > (1LL << 22)) break;
          std::array<int, 3> a = {d[0], d[1], d[2]};
          if (std::find(cand.begin(), cand.end(), a) != cand.end()) continue;
          cand.push_back(a);
          float ms = 0.f;
          if ((r = time_grid(d, &ms)) < 0) break;
          if (ms < 0.99f * best_g) {
            best_g = ms;
            std::copy(d, d + 3, t.dims);
          } else if (ms > 1.03f * best_g) {
            break;
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
        // the default walk mode first; another must be faster by 1 %, more
        // than the spread of repeated timings
        float best_w = 3.0e38f;
        for (int walk : {FFK_WALK_DEFAULT, 0, 1, 2, 4, 5, 6, 7}) {
          if (r != FFK_OK) break;
          if ((r = ffk_set_tuning(s, t.block, 0, -1)) < 0 || (r = ffk_set_walk(s, walk)) < 0) break;
          float ms = 0.f;
          r = time_ms(st, 2, [&] { return ffk_ff_fused(s, FFK_CAST_GRID, 1, d_tally, d_F, nullptr, stream); }, &ms);
          if (r == FFK_OK && (best_w > 1.0e38f || ms < 0.99f * best_w)) {
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
        (void)ffk_set_walk(s, FFK_WALK_DEFAULT);
        (void)ffk_set_refill(s, refill0, pool0);
      }
      (void)hipFree(d_F);
      (void)hipFree(d_tally);
      if (r < 0) return r;
    }
    if (!fused_timed && (r = ffk_scene_build_grid(s, t.dims[0], t.dims[1], t.dims[2], stream)) < 0) return r;
    cache_put("scene", key, {t.dims[0], t.dims[1], t.dims[2], t.block, t.walk, t.refill, t.pool});
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

int ffk_tune_cache_path(char* buf, int capacity) {
  const std::string p = cache_file();
  if (buf && capacity > 0) {
    std::strncpy(buf, p.c_str(), (size_t)capacity - 1);
    buf[capacity - 1] = 0;
  }
  return (int)p.size();
}

void ffk_tune_cache_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_results.clear();
  g_loaded = false;
}

int ffk_batch_autotune(ffk_batch* b, const unsigned long long* seeds, unsigned long long* const* d_tallies,
                       float* const* d_Fs, long long sample_rays, void* stream) {
  if (!b || !seeds || !d_Fs) return FFK_ERR_ARG;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int r;
  // the untuned run: binds the buffers (and gives valid results)
  if ((r = ffk_batch_run(b, seeds, d_tallies, d_Fs, 0, nullptr, stream)) < 0) return r;
  const int n_classes = ffk_batch_classes(b, nullptr, nullptr, nullptr, 0);
  if (n_classes < 0) return n_classes;
  std::vector<int> ids(n_classes), tiles(n_classes);
  std::vector<long long> rays(n_classes);
  if ((r = ffk_batch_classes(b, ids.data(), rays.data(), tiles.data(), n_classes)) < 0) return r;
  unsigned long long key = 0;
  if ((r = ffk_batch_signature(b, &key)) < 0) return r;
  key = device_key(fnv1a(&sample_rays, sizeof(sample_rays), key));
  int block0 = 0, rpt0 = 0, nb0 = 0;
  if ((r = ffk_batch_get_tuning(b, &block0, &rpt0, &nb0)) < 0) return r;
  std::vector<int> v;
  // cached: five words per class, then the batch-wide tile size
  if (cache_get("batch", key, v) && v.size() == 5 * (size_t)n_classes + 1) {
    if (v.back() != rpt0 && (r = ffk_batch_set_tuning(b, block0, v.back(), nb0)) < 0) return r;
    for (int c = 0; c < n_classes; ++c)
      if ((r = ffk_batch_set_class_tuning(b, v[5 * c], v[5 * c + 1], v[5 * c + 2], v[5 * c + 3], v[5 * c + 4])) < 0)
        return r;
    return 1;
  }
  v.clear();
  for (int c = 0; c < n_classes; ++c) {
    const int id = ids[c];
    struct cfg {
      int block, refill, pool, walk;
    };
    // the batch-wide defaults, spelled out: automatic block, 32/2, walk 3
    cfg best = {-1, 32, 2, FFK_WALK_DEFAULT};
    float best_ms = 3.0e38f;
    auto trial = [&](const cfg& x) -> int {
      int rr = ffk_batch_set_class_tuning(b, id, x.block, x.refill, x.pool, x.walk);
      if (rr < 0) return rr;
      float ms = 0.f;
      rr = ffk_batch_run_class(b, id, sample_rays, stream);  // warm-up of this configuration
      if (rr == FFK_OK) rr = time_ms(st, 2, [&] { return ffk_batch_run_class(b, id, sample_rays, stream); }, &ms);
      if (rr == FFK_ERR_HIP) {  // a launch shape the hardware refuses (LDS): not a candidate
        (void)hipGetLastError();
        return FFK_OK;
      }
      if (rr < 0) return rr;
      // the first trial (the defaults) sets the bar; a change must win by 3 %
      // on the sample, more than the spread of repeated timings
      if (best_ms > 1.0e38f || ms < 0.97f * best_ms) {
        best_ms = ms;
        best = x;
      }
      return FFK_OK;
    };
    if ((r = trial(best)) < 0) return r;
    const cfg base = best;
    for (int walk : {0, 1, 2, 7})
      if ((r = trial({base.block, base.refill, base.pool, walk})) < 0) return r;
    const cfg after_walk = best;
    for (const auto& sch : kSchedules) {
      if (sch[0] == after_walk.refill && sch[1] == after_walk.pool) continue;
      if ((r = trial({after_walk.block, sch[0], sch[1], after_walk.walk})) < 0) return r;
    }
    const cfg after_sched = best;
    for (int block : {256, 512, 1024})
      if ((r = trial({block, after_sched.refill, after_sched.pool, after_sched.walk})) < 0) return r;
    // The sample is the head of the queue (the class's largest scenes); a
    // change is kept only if it also wins over the whole class.
    const bool changed = best.block != base.block || best.refill != base.refill || best.pool != base.pool ||
                         best.walk != base.walk;
    if (changed && sample_rays > 0 && sample_rays < rays[c]) {
      float full[2] = {0.f, 0.f};
      const cfg pair[2] = {base, best};
      for (int k = 0; k < 2 && r == FFK_OK; ++k) {
        if ((r = ffk_batch_set_class_tuning(b, id, pair[k].block, pair[k].refill, pair[k].pool, pair[k].walk)) < 0)
          return r;
        if ((r = ffk_batch_run_class(b, id, 0, stream)) < 0) return r;
        r = time_ms(st, 1, [&] { return ffk_batch_run_class(b, id, 0, stream); }, &full[k]);
      }
      if (r < 0) return r;
      if (!(full[1] < 0.99f * full[0])) best = base;
    }
    if ((r = ffk_batch_set_class_tuning(b, id, best.block, best.refill, best.pool, best.walk)) < 0) return r;
    v.insert(v.end(), {id, best.block, best.refill, best.pool, best.walk});
  }
  // Batch-wide tile size, on whole runs (the per-class settings stay: a new
  // tile size plans the same classes).
  const int n_sc = ffk_batch_n_scenes(b);
  if (n_sc < 0) return n_sc;
  std::vector<int> flags0(n_sc), flags(n_sc);
  if ((r = ffk_batch_tally_flags(b, flags0.data())) < 0) return r;
  auto whole = [&](float* ms) {
    int rr = ffk_batch_run(b, seeds, d_tallies, d_Fs, 0, nullptr, stream);  // warm-up (re-plans the graph)
    if (rr == FFK_OK) rr = time_ms(st, 2, [&] { return ffk_batch_run(b, seeds, d_tallies, d_Fs, 0, nullptr, stream); }, ms);
    return rr;
  };
  float best_ms = 0.f;
  if ((r = whole(&best_ms)) < 0) return r;
  int best_rpt = rpt0;
  for (int rpt : {32768, 16384, 8192}) {
    if (rpt == rpt0) continue;
    if ((r = ffk_batch_set_tuning(b, block0, rpt, nb0)) < 0) return r;
    if ((r = ffk_batch_tally_flags(b, flags.data())) < 0) return r;
    if (flags != flags0) continue;  // would need tally buffers the caller did not pass
    float ms = 0.f;
    if ((r = whole(&ms)) < 0) return r;
    if (ms < 0.99f * best_ms) {
      best_ms = ms;
      best_rpt = rpt;
    }
  }
  if ((r = ffk_batch_set_tuning(b, block0, best_rpt, nb0)) < 0) return r;
  v.push_back(best_rpt);
  // valid results for the chosen settings
  if ((r = ffk_batch_run(b, seeds, d_tallies, d_Fs, 0, nullptr, stream)) < 0) return r;
  (void)hipStreamSynchronize(st);
  cache_put("batch", key, v);
  return FFK_OK;
}

}  // extern "C"
