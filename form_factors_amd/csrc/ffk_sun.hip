// Absorption of a distant source (thermal_equation/parallel_rays_cpu.cpp:92-131)
// on the device: parallel rays from an aperture before the scene, cast on the
// grid, hits counted per (mesh, side). Every ray carries the same power, so
// the absorbed power of a mesh side is its exact hit count times
// absorbance * ray_power; the counts are integers, so the result is
// independent of the order of the atomics.
#include "ffk_trace.hpp"

namespace ffk {

__global__ void k_fill_counts(unsigned long long* __restrict__ p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Ray i: counter (i, 0, 1, i >> 32) as ffk_emit_parallel. Disk aperture
// (emission/parallel_rays_cpu.cpp:88-106): rho = radius * sqrt(u), angle
// 2 pi v. Square aperture: x = radius * (2u - 1), y = radius * (2v - 1).
// The aperture's rotation (rotate_towards from +z to the direction), passed by value.
struct rot33 {
  float m[9];
};

template <bool STATS>
__global__ void __launch_bounds__(256) k_sun(scene_desc S, v3 dir, v3 center, float distance, float radius,
                                             int aperture, rot33 rot, long long n, uint32_t k0,
                                             uint32_t k1, unsigned long long* __restrict__ counts,
                                             unsigned long long* __restrict__ stats) {
  uint32_t nt = 0, nc = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const u4 r = philox(u4{(uint32_t)i, 0u, 1u, (uint32_t)(i >> 32)}, k0, k1);
    float x, y;
    if (aperture == 1) {
      x = radius * (2.f * u01(r.x) - 1.f);
      y = radius * (2.f * u01(r.y) - 1.f);
    } else {
      const float rho = radius * sqrtf(u01(r.x));
      float s, c;
      sincos_turns(u01(r.y), &s, &c);
      x = rho * c;
      y = rho * s;
    }
    const v3 o = mat_vec(rot.m, mk(x, y, -distance)) + center;
    const v3 e = o + dir * 0.1f;
    hit_rec h = {3.402823466e38f, -1, mk(0.f, 0.f, 0.f)};
    cast_grid<STATS>(S, o, e - o, h, nt, nc);
    if (h.f < 0) continue;
    const int m = gld(S.face2mesh + h.f);
    if (m < 0) continue;
    // math::triangle_normal = (p1 - p0) x (p2 - p0), the per-face u x v
    const float4* q = reinterpret_cast<const float4*>(S.tri + h.f);
    const float4 q1 = gld4(q + 1), q2 = gld4(q + 2);
    const v3 nrm = cross(mk(q1.x, q1.y, q1.z), mk(q2.x, q2.y, q2.z));
    const int side = dot(nrm, dir) < 0 ? 0 : 1;
    atomicAdd(&counts[2 * (size_t)m + side], 1ull);
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

}  // namespace ffk

using namespace ffk;

// ffk_sun_deposit on the device: the double sum of the counts times the
// reference's per-ray float values (from zero), rounded to float and added to
// the step's absorption.
__global__ void k_sun_deposit_f(int nm, const unsigned long long* __restrict__ counts,
                                const float* __restrict__ absorb2, float ray_power, float* __restrict__ absorption) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const float vf = absorb2[2 * m] * ray_power, vr = absorb2[2 * m + 1] * ray_power;
  const double a = 0.0 + ((double)counts[2 * (size_t)m] * (double)vf + (double)counts[2 * (size_t)m + 1] * (double)vr);
  absorption[m] += (float)a;
}

extern "C" {

int ffk_bsphere(const float* h_faces, int n_faces, float* center3, float* radius) {
  FFK_REQUIRE(h_faces && center3 && radius && n_faces > 0, FFK_ERR_ARG, "ffk_bsphere: bad args");
  v3 gc = mk(0.f, 0.f, 0.f);
  for (int i = 0; i < n_faces; ++i) {
    const float* t = h_faces + 9 * (size_t)i;
    v3 c = (mk(t[0], t[1], t[2]) + mk(t[3], t[4], t[5])) + mk(t[6], t[7], t[8]);
    gc = gc + mk(c.x / 3.f, c.y / 3.f, c.z / 3.f);
  }
  const float inv = 1.f / (float)n_faces;
  gc = mk(gc.x * inv, gc.y * inv, gc.z * inv);
  float r = 0.f;
  for (int i = 1; i < n_faces; ++i) {  // math/triangle.cpp:372 starts at 1
    const float* t = h_faces + 9 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      const v3 d = gc - mk(t[3 * k], t[3 * k + 1], t[3 * k + 2]);
      const float len = sqrtf(dot(d, d));
      if (len > r) r = len;
    }
  }
  center3[0] = gc.x;
  center3[1] = gc.y;
  center3[2] = gc.z;
  *radius = r;
  return FFK_OK;
}

int ffk_sun(ffk_scene* s, const float* h_dir3, const float* h_center3, float distance, float radius,
            int aperture, long long n_rays, unsigned long long seed, unsigned long long* h_counts,
            unsigned long long* d_stats, void* stream) {
  FFK_REQUIRE(s && h_dir3 && h_center3 && h_counts && n_rays >= 0, FFK_ERR_ARG, "ffk_sun: bad args");
  FFK_REQUIRE(s->has_grid && s->has_meshes, FFK_ERR_ARG, "ffk_sun: meshes and grid must be set");
  FFK_REQUIRE(aperture == 0 || aperture == 1, FFK_ERR_ARG, "ffk_sun: aperture %d (0 disk, 1 square)", aperture);
  const int nm = s->n_meshes;
  for (int i = 0; i < 2 * nm; ++i) h_counts[i] = 0ull;
  if (n_rays == 0) return FFK_OK;
  hipStream_t st = as_stream(stream);
  const v3 dir = mk(h_dir3[0], h_dir3[1], h_dir3[2]);
  rot33 rot;
  rotate_towards(mk(0.f, 0.f, 1.f), dir, rot.m);
  unsigned long long* d_buf = nullptr;
  int r;
  if ((r = dalloc(&d_buf, 2 * (size_t)nm)) < 0) return r;
  k_fill_counts<<<(2 * nm + 255) / 256, 256, 0, st>>>(d_buf, 2 * nm);
  const long long blocks_needed = (n_rays + 255) / 256;
  const int blocks = (int)std::max<long long>(1, std::min<long long>(blocks_needed, 64LL * device_cu_count()));
  const v3 center = mk(h_center3[0], h_center3[1], h_center3[2]);
  if (d_stats)
    k_sun<true><<<blocks, 256, 0, st>>>(s->desc(), dir, center, distance, radius, aperture, rot, n_rays,
                                        (uint32_t)seed, (uint32_t)(seed >> 32), d_buf, d_stats);
  else
    k_sun<false><<<blocks, 256, 0, st>>>(s->desc(), dir, center, distance, radius, aperture, rot, n_rays,
                                         (uint32_t)seed, (uint32_t)(seed >> 32), d_buf, nullptr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h_counts, d_buf, 16 * (size_t)nm, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipStreamSynchronize(st);
  dfree(d_buf);
  FFK_HIP(e);
  return FFK_OK;
}

int ffk_sun_count_dev(ffk_scene* s, const float* h_dir3, const float* h_center3, float distance, float radius,
                      int aperture, long long n_rays, unsigned long long seed, unsigned long long* d_counts,
                      void* stream) {
  FFK_REQUIRE(s && h_dir3 && h_center3 && d_counts && n_rays >= 0, FFK_ERR_ARG, "ffk_sun_count_dev: bad args");
  FFK_REQUIRE(s->has_grid && s->has_meshes, FFK_ERR_ARG, "ffk_sun_count_dev: meshes and grid must be set");
  FFK_REQUIRE(aperture == 0 || aperture == 1, FFK_ERR_ARG, "ffk_sun_count_dev: aperture %d (0 disk, 1 square)",
              aperture);
  const int nm = s->n_meshes;
  if (nm == 0) return FFK_OK;
  hipStream_t st = as_stream(stream);
  k_fill_counts<<<(2 * nm + 255) / 256, 256, 0, st>>>(d_counts, 2 * nm);
  if (n_rays > 0) {
    const v3 dir = mk(h_dir3[0], h_dir3[1], h_dir3[2]);
    rot33 rot;
    rotate_towards(mk(0.f, 0.f, 1.f), dir, rot.m);
    const long long blocks_needed = (n_rays + 255) / 256;
    const int blocks = (int)std::max<long long>(1, std::min<long long>(blocks_needed, 64LL * device_cu_count()));
    const v3 center = mk(h_center3[0], h_center3[1], h_center3[2]);
    k_sun<false><<<blocks, 256, 0, st>>>(s->desc(), dir, center, distance, radius, aperture, rot, n_rays,
                                         (uint32_t)seed, (uint32_t)(seed >> 32), d_counts, nullptr);
  }
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_sun_deposit_dev(int n_meshes, const unsigned long long* d_counts, const float* d_absorb2, float ray_power,
                        float* d_absorption, void* stream) {
  FFK_REQUIRE(n_meshes >= 0 && d_counts && d_absorb2 && d_absorption, FFK_ERR_ARG, "ffk_sun_deposit_dev: bad args");
  if (n_meshes == 0) return FFK_OK;
  k_sun_deposit_f<<<(n_meshes + 255) / 256, 256, 0, as_stream(stream)>>>(n_meshes, d_counts, d_absorb2, ray_power,
                                                                         d_absorption);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_sun_deposit(int n_meshes, const unsigned long long* h_counts, const float* h_absorb2, float ray_power,
                    double* h_absorption) {
  FFK_REQUIRE(n_meshes >= 0 && h_counts && h_absorb2 && h_absorption, FFK_ERR_ARG, "ffk_sun_deposit: bad args");
  for (int m = 0; m < n_meshes; ++m) {
    // the reference's per-ray float value, summed exactly
    const float vf = h_absorb2[2 * m] * ray_power, vr = h_absorb2[2 * m + 1] * ray_power;
    h_absorption[m] += (double)h_counts[2 * (size_t)m] * (double)vf + (double)h_counts[2 * (size_t)m + 1] * (double)vr;
  }
  return FFK_OK;
}

}  // extern "C"
