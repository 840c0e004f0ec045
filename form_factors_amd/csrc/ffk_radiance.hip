// radiance_equation::calculate on the device (thermal_equation/radiance_cpu.cpp:132-190).
//
// The reference emits Malley rays with weights T * eps * A per face side,
// casts them, and walks the rays mesh by mesh, face by face (front rays, then
// rear rays), depositing front_emission / front_rays (or the rear analogue)
// on the mesh of the hit face, in float. Here one kernel emits, casts and
// deposits each ray; the walk position of ray g (the reference walks rays in
// emission order) is found in a table of walk entries (mesh-order faces with
// their first walk position), and the powers are summed per mesh in 64-bit
// fixed point: integer additions, so the result does not depend on the order
// of the atomics and carries no float accumulation error.
//
// Ray powers can span many decades (they scale with T^3, and a face whose
// share of the rays rounds below one still emits one ray carrying its whole
// power), so one fixed-point scale for all rays would round the smallest
// ones coarsely. Each ray power is therefore assigned to a band of 2^8 in
// magnitude below the largest power; every band has its own scale and its
// own accumulator row, and the band travels in the top 4 bits of the ray's
// fixed-point word.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ffk_trace.hpp"

namespace ffk {

constexpr int kPowerBands = 16;      // accumulator rows (bands of 2^kBandOrders)
constexpr int kBandOrders = 8;       // binary orders of magnitude per band
constexpr int kBandShift = 60;       // band index in bits 60..63 of a q word
constexpr unsigned long long kValueMask = (1ull << kBandShift) - 1;

// Adds the fixed-point power word v to mesh m's row of its band.
__device__ inline void deposit(unsigned long long* __restrict__ acc, int n_meshes, int m, unsigned long long v) {
  if (v) atomicAdd(&acc[(size_t)(v >> kBandShift) * n_meshes + m], v & kValueMask);
}

__global__ void k_fill_acc(unsigned long long* __restrict__ p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

template <bool STATS>
__global__ void __launch_bounds__(256) k_radiance(scene_desc S, const long long* __restrict__ walk_start,
                                                  const int* __restrict__ walk_face, int n_walk,
                                                  const unsigned long long* __restrict__ q, long long n_rays,
                                                  uint32_t k0, uint32_t k1, unsigned long long* __restrict__ acc,
                                                  unsigned long long* __restrict__ stats) {
  uint32_t nt = 0, nc = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  int hint = 0, whint = 0;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n_rays; g += stride) {
    const int f = find_face(S.face_off, S.n_faces, g, hint);
    hint = f;
    v3 o, e;
    malley_ray(S, f, (int)(g - gld(S.face_off + f)), k0, k1, &o, &e);
    hit_rec h = {3.402823466e38f, -1, mk(0.f, 0.f, 0.f)};
    cast_grid<STATS>(S, o, e - o, h, nt, nc);
    if (h.f < 0) continue;
    const int m = gld(S.face2mesh + h.f);
    if (m < 0) continue;
    const int k = find_face(walk_start, n_walk, g, whint);
    whint = k;
    const int wf = walk_face[k];
    deposit(acc, S.n_meshes, m, q[2 * k + ((g - walk_start[k]) < gld(S.front + wf) ? 0 : 1)]);
  }
  if (STATS) {
    unsigned long long t = nt, c = nc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      t += __shfl_xor(t, off, 64);
      c += __shfl_xor(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&stats[0], t);
      atomicAdd(&stats[1], c);
    }
  }
}

// The same deposit with lane refill and pooled triangle tests
// (pooled_walk): block b casts walk positions [b * chunk, (b + 1) * chunk).
template <bool STATS, int WALK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
    k_radiance_pooled(scene_desc S, const long long* __restrict__ walk_start, const int* __restrict__ walk_face,
                      int n_walk, const unsigned long long* __restrict__ q, long long n_rays, long long chunk,
                      uint32_t k0, uint32_t k1, unsigned long long* __restrict__ acc,
                      unsigned long long* __restrict__ stats, int refill, int pool) {
  __shared__ unsigned long long best[4 * kPoolWords];
  __shared__ int s_next;
  // The faces and the walk entries of the block's first and last rays bound
  // their lookups.
  __shared__ int s_faces[2], s_walk[2];
  const long long r0 = (long long)blockIdx.x * chunk;
  const long long cnt = min(chunk, n_rays - r0);
  if (threadIdx.x < 4) {  // the four searches side by side on lanes 0-3
    const bool walk = threadIdx.x >= 2;
    const long long g = (threadIdx.x & 1) ? r0 + cnt - 1 : r0;
    const int v = find_face(walk ? walk_start : S.face_off, walk ? n_walk : S.n_faces, g, 0);
    (walk ? s_walk : s_faces)[threadIdx.x & 1] = v;
    if (threadIdx.x == 0) s_next = 0;
  }
  __syncthreads();
  uint32_t nt = 0, nc = 0;
  int hint = s_faces[0], whint = s_walk[0];
  pooled_walk<STATS, WALK>(
      S, best + kPoolWords * (threadIdx.x >> 6), &s_next, cnt, refill, pool,
      [&](long long i, v3& o, v3& dir) {
        const long long g = r0 + i;
        const int f = find_face_in(S.face_off, s_faces, g, hint);
        hint = f;
        v3 e;
        malley_ray(S, f, (int)(g - gld(S.face_off + f)), k0, k1, &o, &e);
        dir = e - o;
      },
      [&](long long i, v3, v3, unsigned long long key) {
        float d2;
        int hf;
        decode_key(key, d2, hf);
        if (hf < 0) return;
        const int m = gld(S.face2mesh + hf);
        if (m < 0) return;
        const long long g = r0 + i;
        const int k = find_face_in(walk_start, s_walk, g, whint);
        whint = k;
        const int wf = walk_face[k];
        deposit(acc, S.n_meshes, m, q[2 * k + ((g - walk_start[k]) < gld(S.front + wf) ? 0 : 1)]);
      },
      nt, nc);
  if (STATS) {
    unsigned long long t = nt, c = nc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      t += __shfl_xor(t, off, 64);
      c += __shfl_xor(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&stats[0], t);
      atomicAdd(&stats[1], c);
    }
  }
}

}  // namespace ffk

using namespace ffk;

namespace {

// math::triangle_area (math/triangle.cpp:86-92) in float.
float face_area(const float* p) {
  v3 u = mk(p[3], p[4], p[5]) - mk(p[0], p[1], p[2]);
  v3 w = mk(p[6], p[7], p[8]) - mk(p[0], p[1], p[2]);
  v3 n = cross(u, w);
  return sqrtf(dot(n, n)) / 2;
}

bool meshes_valid(const int* h_meshes, int n_meshes, int n_faces) {
  for (int m = 0; m < n_meshes; ++m) {
    const long long f0 = h_meshes[3 * m], n = h_meshes[3 * m + 1];
    if (f0 < 0 || n < 0 || f0 + n > n_faces) return false;
  }
  return true;
}

// Area of face f: from the scene's cache when given (computed there with the
// same expression), else from the face.
inline float area_of(const float* h_area, const float* h_faces, int f) {
  return h_area ? h_area[f] : face_area(h_faces + 9 * (size_t)f);
}

// radiance_equation::calculate_weights (thermal_equation/radiance_cpu.cpp:96-130):
// T * eps * area per face side, and their float total in mesh order.
float weights_impl(const float* h_faces, const float* h_area, int n_faces, const int* h_meshes, int n_meshes,
                   const float* h_eps2, const float* h_T, float* h_weights) {
  for (int i = 0; i < 2 * n_faces; ++i) h_weights[i] = 0.f;  // faces outside every mesh
  float t = 0;
  for (int m = 0; m < n_meshes; ++m) {
    const float density = h_T[m];
    const float front_density = density * h_eps2[2 * m];
    const float rear_density = density * h_eps2[2 * m + 1];
    for (int f = h_meshes[3 * m]; f < h_meshes[3 * m] + h_meshes[3 * m + 1]; ++f) {
      const float a = area_of(h_area, h_faces, f);
      const float front = front_density * a;
      const float rear = rear_density * a;
      t += front + rear;
      h_weights[2 * f] = front;
      h_weights[2 * f + 1] = rear;
    }
  }
  return t;
}

// thermal_equation/radiance_cpu.cpp:148-178: the power of one front / rear
// ray of each face and the emission added to each mesh.
void powers_impl(const float* h_faces, const float* h_area, int n_faces, const int* h_meshes, int n_meshes,
                 const float* h_eps2, const float* h_T, const int* h_front, const int* h_rear, float* h_ray_power,
                 float* h_emission) {
  const float sigma = 5.670400e-8f;  // thermal_equation/radiance_cpu.cpp:29
  for (int i = 0; i < 2 * n_faces; ++i) h_ray_power[i] = 0.f;
  for (int m = 0; m < n_meshes; ++m) {
    const float T = h_T[m];
    const float power_density = sigma * (T * T * T * T);
    const float front_density = power_density * h_eps2[2 * m];
    const float rear_density = power_density * h_eps2[2 * m + 1];
    for (int f = h_meshes[3 * m]; f < h_meshes[3 * m] + h_meshes[3 * m + 1]; ++f) {
      const float a = area_of(h_area, h_faces, f);
      const float front_emission = front_density * a;
      const float rear_emission = rear_density * a;
      h_emission[m] += front_emission + rear_emission;
      h_ray_power[2 * f] = h_front[f] > 0 ? front_emission / h_front[f] : 0.f;
      h_ray_power[2 * f + 1] = h_rear[f] > 0 ? rear_emission / h_rear[f] : 0.f;
    }
  }
}

}  // namespace

extern "C" {

int ffk_radiance_weights(const float* h_faces, int n_faces, const int* h_meshes, int n_meshes,
                         const float* h_eps2, const float* h_T, float* h_weights, float* total) {
  FFK_REQUIRE(h_faces && h_meshes && h_eps2 && h_T && h_weights && total && n_faces >= 0 && n_meshes >= 0,
              FFK_ERR_ARG, "ffk_radiance_weights: bad args");
  FFK_REQUIRE(meshes_valid(h_meshes, n_meshes, n_faces), FFK_ERR_ARG, "ffk_radiance_weights: mesh out of range");
  *total = weights_impl(h_faces, nullptr, n_faces, h_meshes, n_meshes, h_eps2, h_T, h_weights);
  return FFK_OK;
}

int ffk_radiance_powers(const float* h_faces, int n_faces, const int* h_meshes, int n_meshes,
                        const float* h_eps2, const float* h_T, const int* h_front, const int* h_rear,
                        float* h_ray_power, float* h_emission) {
  FFK_REQUIRE(h_faces && h_meshes && h_eps2 && h_T && h_front && h_rear && h_ray_power && h_emission &&
                  n_faces >= 0 && n_meshes >= 0,
              FFK_ERR_ARG, "ffk_radiance_powers: bad args");
  FFK_REQUIRE(meshes_valid(h_meshes, n_meshes, n_faces), FFK_ERR_ARG, "ffk_radiance_powers: mesh out of range");
  powers_impl(h_faces, nullptr, n_faces, h_meshes, n_meshes, h_eps2, h_T, h_front, h_rear, h_ray_power, h_emission);
  return FFK_OK;
}

int ffk_radiance(ffk_scene* s, unsigned long long seed, const float* h_ray_power, double* h_absorption,
                 unsigned long long* d_stats, void* stream) {
  FFK_REQUIRE(s && h_ray_power && h_absorption, FFK_ERR_ARG, "ffk_radiance: bad args");
  FFK_REQUIRE(s->has_emission && s->has_grid, FFK_ERR_ARG, "ffk_radiance: emission layout and grid must be set");
  hipStream_t st = as_stream(stream);
  const int nf = s->n_faces, nm = s->n_meshes;
  const long long total = s->n_rays;
  // Staging for W walk entries and nm meshes, in this order: start[W + 1]
  // (int64), q[2W] (uint64), acc[kPowerBands * nm] (uint64), face[W + 1]
  // (int32). Every walk entry is a face, so W = n_faces bounds them.
  const size_t W = (size_t)nf;
  const size_t off_q = 8 * (W + 1), off_acc = off_q + 16 * W, off_face = off_acc + 8 * (size_t)kPowerBands * nm;
  const size_t bytes = off_face + 4 * (W + 1);
  if ((int)W + 1 > s->rad_cap_walk || nm > s->rad_cap_meshes) {
    FFK_HIP(hipStreamSynchronize(st));
    dfree(s->d_walk_block);
    s->d_walk_block = nullptr;
    if (s->h_walk_block) (void)hipHostFree(s->h_walk_block);
    s->h_walk_block = nullptr;
    s->rad_cap_walk = s->rad_cap_meshes = 0;
    int r;
    if ((r = dmalloc(&s->d_walk_block, bytes)) < 0) return r;
    FFK_HIP(hipHostMalloc(&s->h_walk_block, bytes, hipHostMallocDefault));
    s->rad_cap_walk = (int)W + 1;
    s->rad_cap_meshes = nm;
  }
  char* hb = static_cast<char*>(s->h_walk_block);
  char* db = static_cast<char*>(s->d_walk_block);
  long long* start = reinterpret_cast<long long*>(hb);
  unsigned long long* q = reinterpret_cast<unsigned long long*>(hb + off_q);
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(hb + off_acc);
  int* face = reinterpret_cast<int*>(hb + off_face);
  s->d_walk_start = reinterpret_cast<long long*>(db);
  s->d_walk_q = reinterpret_cast<unsigned long long*>(db + off_q);
  s->d_rad_acc = reinterpret_cast<unsigned long long*>(db + off_acc);
  s->d_walk_face = reinterpret_cast<int*>(db + off_face);
  // Walk entries: faces in mesh order with their first walk position; the
  // walk stops at the last emitted ray (radiance_cpu.cpp:170).
  int n_walk = 0;
  long long pos = 0;
  float p_max = 0.f, p_min = 0.f;  // over the ray powers of sides that emit rays
  double upper_all = 0.0;
  for (int m = 0; m < nm && pos < total; ++m) {
    const int f0 = s->h_meshes[3 * m], n = s->h_meshes[3 * m + 1];
    for (int f = f0; f < f0 + n && pos < total; ++f) {
      const float pf = h_ray_power[2 * f], pr = h_ray_power[2 * f + 1];
      FFK_REQUIRE(pf >= 0.f && pr >= 0.f && std::isfinite(pf) && std::isfinite(pr), FFK_ERR_ARG,
                  "ffk_radiance: ray power of face %d is negative or not finite", f);
      start[n_walk] = pos;
      face[n_walk] = f;
      ++n_walk;
      const int cf = s->h_front[f], cr = s->h_rear[f];
      if (cf > 0 && pf > 0.f) {
        p_max = std::max(p_max, pf);
        p_min = p_min > 0.f ? std::min(p_min, pf) : pf;
      }
      if (cr > 0 && pr > 0.f) {
        p_max = std::max(p_max, pr);
        p_min = p_min > 0.f ? std::min(p_min, pr) : pr;
      }
      upper_all += (double)cf * pf + (double)cr * pr;
      pos += (long long)cf + cr;
    }
  }
  const long long walk_total = std::min(pos, total);
  for (int m = 0; m < nm; ++m) h_absorption[m] = 0.0;
  if (n_walk == 0 || p_max <= 0.f) return FFK_OK;
  start[n_walk] = walk_total;
  // Band of a ray power: its binary order below the largest power, in steps
  // of kBandOrders (the last band takes everything smaller). Per band, the
  // largest possible sum (every ray of the band absorbed by one mesh) and the
  // largest single power fix the scale: sums stay below 2^63 and a single
  // ray's value below 2^60, clear of the band bits. A ray of power p in band
  // b is then rounded by at most 0.5 / scale_b, i.e. relatively by at most
  // 2^-55 times the rays in the band, or 2^-52 (2^kBandOrders / 2^60).
  auto exponent = [](float p) {  // ilogb of a positive float from its bits (subnormals rare)
    uint32_t u;
    std::memcpy(&u, &p, 4);
    const int e = (int)((u >> 23) & 0xff);
    return e ? e - 127 : std::ilogb(p);
  };
  const int e_max = exponent(p_max);
  const bool one_band = (e_max - exponent(p_min)) / kBandOrders == 0;  // the usual case
  double upper[kPowerBands] = {}, pmax_b[kPowerBands] = {};
  int n_bands = 1;
  unsigned char* band = nullptr;
  if (one_band) {
    upper[0] = upper_all;
    pmax_b[0] = p_max;
  } else {
    s->rad_band.resize(2 * (size_t)n_walk);
    band = s->rad_band.data();
    for (int k = 0; k < n_walk; ++k) {
      const int f = face[k];
      const int cnt[2] = {s->h_front[f], s->h_rear[f]};
      for (int side = 0; side < 2; ++side) {
        const float p = h_ray_power[2 * f + side];
        const bool emits = p > 0.f && cnt[side] > 0;
        const int b = emits ? std::min(kPowerBands - 1, (e_max - exponent(p)) / kBandOrders) : 0;
        band[2 * k + side] = (unsigned char)b;
        if (!emits) continue;
        upper[b] += (double)cnt[side] * p;
        pmax_b[b] = std::max(pmax_b[b], (double)p);
        n_bands = std::max(n_bands, b + 1);
      }
    }
  }
  double scale[kPowerBands];
  for (int b = 0; b < kPowerBands; ++b) {
    int e_sum = 0, e_one = 0;
    std::frexp(upper[b], &e_sum);   // upper < 2^e_sum
    std::frexp(pmax_b[b], &e_one);  // pmax < 2^e_one
    scale[b] = upper[b] > 0.0 ? std::ldexp(1.0, std::min(63 - e_sum, kBandShift - e_one)) : 1.0;
  }
  // round half to even of a non-negative x < 2^60 (what llrint does in the
  // default rounding mode; x - trunc(x) is exact, so the comparison is too)
  auto round_even = [](double x) {
    const long long v = (long long)x;
    const double frac = x - (double)v;
    return (unsigned long long)v + (frac > 0.5 || (frac == 0.5 && (v & 1)) ? 1ull : 0ull);
  };
  for (int k = 0; k < n_walk; ++k) {
    const int f = face[k];
    const int cnt[2] = {s->h_front[f], s->h_rear[f]};
    for (int side = 0; side < 2; ++side) {
      const float p = h_ray_power[2 * f + side];
      const int b = band ? band[2 * k + side] : 0;
      // sides without rays are never read; their power may lie outside the bands
      q[2 * k + side] = p > 0.f && cnt[side] > 0
                            ? ((unsigned long long)b << kBandShift) | round_even((double)p * scale[b])
                            : 0ull;
    }
  }
  const size_t n_acc = (size_t)n_bands * nm;
  // start, q and (past acc) face in one copy; acc is zeroed on the device
  FFK_HIP(hipMemcpyAsync(db, hb, off_acc, hipMemcpyHostToDevice, st));
  FFK_HIP(hipMemcpyAsync(db + off_face, hb + off_face, 4 * ((size_t)n_walk + 1), hipMemcpyHostToDevice, st));
  k_fill_acc<<<(unsigned)((n_acc + 255) / 256), 256, 0, st>>>(s->d_rad_acc, (int)n_acc);
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  if (walk_total > 0 && s->tune_refill > 0 && s->tune_pool > 0) {
    // Whole rounds of the blocks the device holds at once (the last round is
    // not left part empty), at most 16384 rays and at least a ray per lane
    // per block.
    auto go = [&](auto kernel) {
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
      const long long resident = (long long)per_cu * device_cu_count();
      const long long rounds = std::max<long long>(1, (walk_total + resident * 16384 - 1) / (resident * 16384));
      const long long chunk = std::max<long long>(256, (walk_total + resident * rounds - 1) / (resident * rounds));
      const unsigned blocks = (unsigned)((walk_total + chunk - 1) / chunk);
      kernel<<<blocks, 256, 0, st>>>(s->desc(), s->d_walk_start, s->d_walk_face, n_walk, s->d_walk_q, walk_total,
                                     chunk, k0, k1, s->d_rad_acc, d_stats, s->tune_refill,
                                     pool_arg(s->tune_pool, s->tune_walk));
    };
    const bool jump = (s->tune_walk & kWalkJump) != 0;
    if (d_stats) jump ? go(k_radiance_pooled<true, 1>) : go(k_radiance_pooled<true, 2>);
    else jump ? go(k_radiance_pooled<false, 1>) : go(k_radiance_pooled<false, 2>);
  } else if (walk_total > 0) {
    const long long blocks_needed = (walk_total + 255) / 256;
    const int blocks = (int)std::max<long long>(1, std::min<long long>(blocks_needed, 64LL * device_cu_count()));
    if (d_stats)
      k_radiance<true><<<blocks, 256, 0, st>>>(s->desc(), s->d_walk_start, s->d_walk_face, n_walk, s->d_walk_q,
                                               walk_total, k0, k1, s->d_rad_acc, d_stats);
    else
      k_radiance<false><<<blocks, 256, 0, st>>>(s->desc(), s->d_walk_start, s->d_walk_face, n_walk, s->d_walk_q,
                                                walk_total, k0, k1, s->d_rad_acc, nullptr);
  }
  FFK_HIP(hipGetLastError());
  FFK_HIP(hipMemcpyAsync(acc, s->d_rad_acc, 8 * n_acc, hipMemcpyDeviceToHost, st));
  FFK_HIP(hipStreamSynchronize(st));
  for (int m = 0; m < nm; ++m) {
    double a = 0.0;
    for (int b = 0; b < n_bands; ++b) a += (double)acc[(size_t)b * nm + m] / scale[b];
    h_absorption[m] = a;
  }
  return FFK_OK;
}

int ffk_radiance_step(ffk_scene* s, const float* h_eps2, const float* h_T, int n_rays, int frame,
                      unsigned long long seed, float* h_emission, double* h_absorption,
                      unsigned long long* d_stats, void* stream) {
  FFK_REQUIRE(s && s->has_faces && s->has_meshes && h_eps2 && h_T && h_emission && h_absorption && n_rays >= 0,
              FFK_ERR_ARG, "ffk_radiance_step: bad args");
  FFK_REQUIRE(s->has_grid, FFK_ERR_ARG, "ffk_radiance_step: the grid must be built");
  const int nf = s->n_faces, nm = s->n_meshes;
  s->rad_w.resize(2 * (size_t)nf);
  const float total =
      weights_impl(s->h_faces.data(), s->h_area.data(), nf, s->h_meshes.data(), nm, h_eps2, h_T, s->rad_w.data());
  s->rad_front.resize(nf);
  s->rad_rear.resize(nf);
  int r;
  if ((r = ffk_emitted_counts(s->rad_w.data(), nf, n_rays, total, s->rad_front.data(), s->rad_rear.data())) < 0)
    return r;
  if (!s->has_emission || frame != s->frame || s->rad_front != s->h_front || s->rad_rear != s->h_rear) {
    if ((r = ffk_scene_set_emission(s, s->rad_front.data(), s->rad_rear.data(), frame, nullptr, stream)) < 0)
      return r;
  }
  s->rad_p.resize(2 * (size_t)nf);
  powers_impl(s->h_faces.data(), s->h_area.data(), nf, s->h_meshes.data(), nm, h_eps2, h_T, s->h_front.data(),
              s->h_rear.data(), s->rad_p.data(), h_emission);
  return ffk_radiance(s, seed, s->rad_p.data(), h_absorption, d_stats, stream);
}

}  // extern "C"
