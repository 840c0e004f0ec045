// Nearest-hit casting of a device ray array (ray_caster::system_cast).
#include <algorithm>

#include "ffk_trace.hpp"

namespace ffk {

__device__ inline void wave_add_stats(unsigned long long* stats, uint32_t tests, uint32_t cells) {
  unsigned long long t = tests, c = cells;
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

__device__ inline void load_ray(const float* __restrict__ rays, long long i, v3* o, v3* dir) {
  const float* r = rays + 6 * i;
  *o = mk(r[0], r[1], r[2]);
  *dir = mk(r[3], r[4], r[5]) - *o;  // direction = end - origin (math/triangle.cpp:41)
}

__device__ inline void store_hit(const hit_rec& h, long long i, int* __restrict__ idx,
                                 float* __restrict__ pts) {
  idx[i] = h.f;
  if (pts && h.f >= 0) {
    pts[3 * i] = h.p.x;
    pts[3 * i + 1] = h.p.y;
    pts[3 * i + 2] = h.p.z;
  }
}

template <bool STATS>
__global__ void __launch_bounds__(256) k_cast_grid(scene_desc S, const float* __restrict__ rays,
                                                   long long n, int* __restrict__ idx,
                                                   float* __restrict__ pts,
                                                   unsigned long long* __restrict__ stats) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nt = 0, nc = 0;
  if (i < n) {
    v3 o, dir;
    load_ray(rays, i, &o, &dir);
    hit_rec h = {3.402823466e38f, -1, mk(0.f, 0.f, 0.f)};
    cast_grid<STATS>(S, o, dir, h, nt, nc);
    store_hit(h, i, idx, pts);
  }
  if (STATS) wave_add_stats(stats, nt, nc);
}

// The same cast with lane refill and pooled triangle tests (pooled_walk):
// block b casts rays [b * chunk, (b + 1) * chunk). The hit point of the
// winning face is recomputed with the same test, so points are those of
// k_cast_grid bit for bit.
template <bool STATS, int WALK>
__global__ void __launch_bounds__(256) k_cast_pooled(scene_desc S, const float* __restrict__ rays, long long n,
                                                     long long chunk, int* __restrict__ idx,
                                                     float* __restrict__ pts, unsigned long long* __restrict__ stats,
                                                     int refill, int pool) {
  __shared__ unsigned long long best[4 * kPoolWords];
  __shared__ int s_next;
  const long long r0 = (long long)blockIdx.x * chunk;
  const long long cnt = min(chunk, n - r0);
  if (threadIdx.x == 0) s_next = 0;
  __syncthreads();
  uint32_t nt = 0, nc = 0;
  pooled_walk<STATS, WALK>(
      S, best + kPoolWords * (threadIdx.x >> 6), &s_next, cnt, refill, pool,
      [&](long long i, v3& o, v3& dir) { load_ray(rays, r0 + i, &o, &dir); },
      [&](long long i, v3 o, v3 dir, unsigned long long key) {
        hit_rec h = {3.402823466e38f, -1, mk(0.f, 0.f, 0.f)};
        decode_key(key, h.d2, h.f);
        if (h.f >= 0 && pts) {
          const float4* q = reinterpret_cast<const float4*>(S.tri + h.f);
          float d2;
          (void)hit_terms(gld4(q), gld4(q + 1), gld4(q + 2), gld4(q + 3), o, dir, &h.p, &d2);
        }
        store_hit(h, r0 + i, idx, pts);
      },
      nt, nc);
  if (STATS) wave_add_stats(stats, nt, nc);
}

// Brute force for many rays: one ray per thread, faces streamed through LDS
// in tiles of 256 (16 KB of terms); every lane reads the same LDS word, a
// broadcast, so the inner loop is conflict free.
constexpr int kNaiveTile = 256;

__global__ void __launch_bounds__(256) k_cast_naive_lds(const tri_terms* __restrict__ tri, int n_faces,
                                                        const float* __restrict__ rays, long long n,
                                                        int* __restrict__ idx, float* __restrict__ pts) {
  __shared__ float4 sm[4 * kNaiveTile];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  v3 o = mk(0.f, 0.f, 0.f), dir = mk(0.f, 0.f, 1.f);
  if (live) load_ray(rays, i, &o, &dir);
  hit_rec h = {3.402823466e38f, -1, mk(0.f, 0.f, 0.f)};
  const float4* src = reinterpret_cast<const float4*>(tri);
  for (int base = 0; base < n_faces; base += kNaiveTile) {
    const int cnt = min(kNaiveTile, n_faces - base);
    __syncthreads();
    for (int k = threadIdx.x; k < 4 * cnt; k += blockDim.x) sm[k] = src[4 * (size_t)base + k];
    __syncthreads();
    if (live) {
      for (int j = 0; j < cnt; ++j) {
        v3 p;
        float d2;
        if (hit_terms(sm[4 * j], sm[4 * j + 1], sm[4 * j + 2], sm[4 * j + 3], o, dir, &p, &d2))
          consider(h, base + j, d2, p);
      }
    }
  }
  if (live) store_hit(h, i, idx, pts);
}

// Brute force for few rays: one wavefront per ray, lanes stride over the
// faces, then a cross-lane (distance, index) minimum resolves the nearest
// hit with the same tie rule as the sequential loop.
__global__ void __launch_bounds__(256) k_cast_naive_wave(const tri_terms* __restrict__ tri, int n_faces,
                                                         const float* __restrict__ rays, long long n,
                                                         int* __restrict__ idx, float* __restrict__ pts) {
  const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;  // whole wavefront leaves together
  v3 o, dir;
  load_ray(rays, i, &o, &dir);
  hit_rec h = {3.402823466e38f, -1, mk(0.f, 0.f, 0.f)};
  for (int f = lane; f < n_faces; f += 64) test_face(tri, f, o, dir, h);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    hit_rec o2;
    o2.d2 = __shfl_xor(h.d2, off, 64);
    o2.f = __shfl_xor(h.f, off, 64);
    o2.p.x = __shfl_xor(h.p.x, off, 64);
    o2.p.y = __shfl_xor(h.p.y, off, 64);
    o2.p.z = __shfl_xor(h.p.z, off, 64);
    if (o2.f >= 0) consider(h, o2.f, o2.d2, o2.p);
  }
  if (lane == 0) store_hit(h, i, idx, pts);
}

}  // namespace ffk

using namespace ffk;

extern "C" int ffk_cast(ffk_scene* s, int method, const float* d_rays, long long n_rays, int* d_hit_idx,
                        float* d_hit_pts, unsigned long long* d_stats, void* stream) {
  FFK_REQUIRE(s && s->has_faces, FFK_ERR_ARG, "ffk_cast: scene has no faces");
  FFK_REQUIRE(n_rays >= 0, FFK_ERR_ARG, "ffk_cast: negative ray count");
  if (n_rays == 0) return FFK_OK;
  FFK_REQUIRE(d_rays && d_hit_idx, FFK_ERR_ARG, "ffk_cast: NULL rays or hit buffer");
  hipStream_t st = as_stream(stream);
  if (method == FFK_CAST_GRID) {
    FFK_REQUIRE(s->has_grid, FFK_ERR_ARG, "ffk_cast: grid not built");
    if (s->tune_refill > 0 && s->tune_pool > 0) {
      // About eight blocks per CU, at least a ray per lane and at most 64.
      const long long target = 8LL * ffk::device_cu_count();
      const long long chunk = std::min<long long>(16384, std::max<long long>(256, (n_rays + target - 1) / target));
      const unsigned blocks = (unsigned)((n_rays + chunk - 1) / chunk);
      auto go = [&](auto kernel) {
        kernel<<<blocks, 256, 0, st>>>(s->desc(), d_rays, n_rays, chunk, d_hit_idx, d_hit_pts, d_stats,
                                       s->tune_refill, pool_arg(s->tune_pool, s->tune_walk));
      };
      const bool jump = (s->tune_walk & kWalkJump) != 0;
      if (d_stats) jump ? go(k_cast_pooled<true, 1>) : go(k_cast_pooled<true, 2>);
      else jump ? go(k_cast_pooled<false, 1>) : go(k_cast_pooled<false, 2>);
    } else {
      const unsigned blocks = (unsigned)((n_rays + 255) / 256);
      if (d_stats)
        k_cast_grid<true><<<blocks, 256, 0, st>>>(s->desc(), d_rays, n_rays, d_hit_idx, d_hit_pts, d_stats);
      else
        k_cast_grid<false><<<blocks, 256, 0, st>>>(s->desc(), d_rays, n_rays, d_hit_idx, d_hit_pts, nullptr);
    }
  } else if (method == FFK_CAST_NAIVE) {
    // Few rays: a wavefront per ray keeps every CU busy on the face loop.
    if (n_rays * 64 <= 4LL * ffk::device_cu_count() * 1024) {
      const unsigned blocks = (unsigned)((n_rays * 64 + 255) / 256);
      k_cast_naive_wave<<<blocks, 256, 0, st>>>(s->d_tri, s->n_faces, d_rays, n_rays, d_hit_idx, d_hit_pts);
    } else {
      const unsigned blocks = (unsigned)((n_rays + 255) / 256);
      k_cast_naive_lds<<<blocks, 256, 0, st>>>(s->d_tri, s->n_faces, d_rays, n_rays, d_hit_idx, d_hit_pts);
    }
    if (d_stats) {
      // Brute force tests every face; no cells are visited.
      unsigned long long add[2] = {(unsigned long long)n_rays * (unsigned long long)s->n_faces, 0ull};
      unsigned long long cur[2];
      FFK_HIP(hipMemcpyAsync(cur, d_stats, 16, hipMemcpyDeviceToHost, st));
      FFK_HIP(hipStreamSynchronize(st));
      cur[0] += add[0];
      FFK_HIP(hipMemcpyAsync(d_stats, cur, 16, hipMemcpyHostToDevice, st));
      FFK_HIP(hipStreamSynchronize(st));
    }
  } else {
    FFK_REQUIRE(false, FFK_ERR_ARG, "ffk_cast: unknown method %d", method);
  }
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}
