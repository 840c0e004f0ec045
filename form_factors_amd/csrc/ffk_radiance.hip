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
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

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
  // Each wavefront's hits waiting for their deposit: ray index in the block
  // and face, fewer than 64 before an iteration and at most 64 added by one.
  __shared__ int s_hit_ray[4][2 * 64], s_hit_face[4][2 * 64];
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
  int whint = s_walk[0];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int* const hit_ray = s_hit_ray[wave];
  int* const hit_face = s_hit_face[wave];
  int n_queued = 0;              // wave-uniform: changed only where every lane runs
  int pend_ray = 0, pend_face = -1;  // this lane's hit of this iteration, not yet queued
  // The face of the ray a refill gives this lane, and the ray's index in it,
  // found by the whole wavefront before the refill (prep below).
  int new_face = 0, new_j = 0;
  int window = s_faces[0];  // wave-uniform: first face of the offsets the wavefront loads
  pooled_walk<STATS, WALK>(
      S, best + kPoolWords * wave, &s_next, cnt, refill, pool,
      [&](long long, v3& o, v3& dir) {
        v3 e;
        malley_ray(S, new_face, new_j, k0, k1, &o, &e);
        dir = e - o;
      },
      // A ray's deposit needs a chain of dependent loads (its mesh, walk
      // entry, the entry's face and power); queued, 64 of them run side by
      // side at the end of an iteration instead of holding up the walk. The
      // sums are integers, so their order does not matter. A lane ends at
      // most one ray per iteration.
      [&](long long i, v3, v3, unsigned long long key) {
        float d2;
        decode_key(key, d2, pend_face);
        pend_ray = (int)i;
      },
      nt, nc,
      [&](bool last) {
        const unsigned long long hits = __ballot(pend_face >= 0);
        if (pend_face >= 0) {
          const int at = n_queued + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(hits >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)hits, 0u));
          hit_ray[at] = pend_ray;
          hit_face[at] = pend_face;
          pend_face = -1;
        }
        n_queued += __popcll(hits);
        if (n_queued < 64 && !(last && n_queued > 0)) return;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the queue as every lane wrote it
        while (n_queued >= 64 || (last && n_queued > 0)) {
          const int take = min(n_queued, 64);
          n_queued -= take;
          if (lane < take) {
            const int i = hit_ray[n_queued + lane], hf = hit_face[n_queued + lane];
            const int m = gld(S.face2mesh + hf);
            if (m >= 0) {
              const long long g = r0 + i;
              const int k = find_face_in(walk_start, s_walk, g, whint);
              whint = k;
              const int wf = walk_face[k];
              deposit(acc, S.n_meshes, m, q[2 * k + ((g - walk_start[k]) < gld(S.front + wf) ? 0 : 1)]);
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // read before the next writes
      },
      // A refill's rays are consecutive: the wavefront loads the first ray
      // offsets of 64 consecutive faces from `window` in one load, and each
      // new ray finds its face among them by a search over the lanes (the
      // last face whose first ray is at or before it), instead of a chain of
      // dependent loads per lane. When the 64 faces do not hold every new
      // ray, each lane searches the table as before.
      [&](int base, unsigned long long idle) {
        const int n_new = __popcll(idle);
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
        // Positions relative to the block's first ray, in 32 bits (clamped:
        // order is all the search needs, and the window is used only when its
        // first and last offsets bracket the new rays).
        const int last = (int)min((long long)base + n_new, cnt) - 1;  // wave-uniform
        const int gi = min(base + rank, last);
        const int wf = window + lane;
        const long long off64 = wf <= S.n_faces ? gld(S.face_off + wf) - r0 : 0x7fffffffLL;
        const int off = (int)max(min(off64, 0x7fffffffLL), -0x7fffffffLL);
        const int first_off = __shfl(off, 0, 64), end_off = __shfl(off, 63, 64);
        if (first_off <= base && last < end_off && first_off > -0x7fffffff) {
          int k = 0;  // the last lane whose offset is at or before gi
#pragma unroll
          for (int step = 32; step >= 1; step >>= 1) {
            const int o_c = __shfl(off, k + step, 64);
            if (o_c <= gi) k += step;
          }
          new_face = window + k;
          new_j = gi - __shfl(off, k, 64);
        } else {
          const long long g = r0 + gi;
          new_face = find_face_in(S.face_off, s_faces, g, window);
          new_j = (int)(g - gld(S.face_off + new_face));
        }
        window = __builtin_amdgcn_readlane(new_face, 63 - __builtin_clzll(idle));
      });
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

constexpr float kSigma = 5.670400e-8f;  // thermal_equation/radiance_cpu.cpp:29
constexpr int kRadThreads = 256;

// Binary exponent of a positive float (ilogb), from its bits.
__host__ __device__ inline int pow_exponent(float p) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint32_t u = __float_as_uint(p);
#else
  uint32_t u;
  std::memcpy(&u, &p, 4);
#endif
  const int e = (int)((u >> 23) & 0xff);
  return e ? e - 127 : ilogbf(p);
}

// Round half to even of a non-negative x < 2^60 (what llrint does in the
// default rounding mode; x - trunc(x) is exact, so the comparison is too).
__host__ __device__ inline unsigned long long round_even(double x) {
  const long long v = (long long)x;
  const double frac = x - (double)v;
  return (unsigned long long)v + (frac > 0.5 || (frac == 0.5 && (v & 1)) ? 1ull : 0ull);
}

// Deterministic block reduction (fixed shuffle tree, then the wavefronts in
// order); the result is valid in every thread. sh: blockDim.x / 64 slots.
template <class T, class Op>
__device__ inline T block_reduce(T v, Op op, T* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  T r = sh[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) r = op(r, sh[i]);
  __syncthreads();
  return r;
}

// Ray power of each walk entry's sides (radiance_cpu.cpp:148-176): the
// emission of the entry's face under the entry's mesh over the face's ray
// count, in the float expressions of the host powers_impl, so the values are
// its bits. (A face in several meshes has an entry in each, each with its
// mesh's temperature and emissivities, as in the reference's walk.)
// te: T[n_meshes] then eps[2 * n_meshes].
__global__ void __launch_bounds__(kRadThreads)
    k_rad_powers(int W, const int* __restrict__ walk_face, const int* __restrict__ walk_mesh,
                 const float* __restrict__ te, int nm, const float* __restrict__ area,
                 const long long* __restrict__ face_off, const int* __restrict__ front, float* __restrict__ p) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= W) return;
  const int f = walk_face[k], m = walk_mesh[k];
  float pf = 0.f, pr = 0.f;
  {
    const float T = te[m];
    const float power_density = kSigma * (T * T * T * T);
    const float front_density = power_density * te[nm + 2 * m];
    const float rear_density = power_density * te[nm + 2 * m + 1];
    const float a = area[f];
    const float front_emission = front_density * a;
    const float rear_emission = rear_density * a;
    const int cf = front[f], cr = (int)(face_off[f + 1] - face_off[f]) - cf;
    pf = cf > 0 ? front_emission / cf : 0.f;
    pr = cr > 0 ? rear_emission / cr : 0.f;
  }
  p[2 * k] = pf;
  p[2 * k + 1] = pr;
}

// The emission each mesh adds to the step's emission (radiance_cpu.cpp:
// 148-178): its faces' front and rear emission, added face by face in float
// to what the array holds, in the expressions and order of the host
// powers_impl, so the sums are its bits. te as in k_rad_powers; mesh m's
// faces are walk entries [mesh_walk[m], mesh_walk[m + 1]).
__global__ void __launch_bounds__(kRadThreads)
    k_rad_emission(int nm, const int* __restrict__ mesh_walk, const int* __restrict__ walk_face,
                   const float* __restrict__ te, const float* __restrict__ area, float* __restrict__ emission) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const float T = te[m];
  const float power_density = kSigma * (T * T * T * T);
  const float front_density = power_density * te[nm + 2 * m];
  const float rear_density = power_density * te[nm + 2 * m + 1];
  float e = emission[m];
  for (int k = mesh_walk[m]; k < mesh_walk[m + 1]; ++k) {
    const float a = area[walk_face[k]];
    const float front_emission = front_density * a;
    const float rear_emission = rear_density * a;
    e += front_emission + rear_emission;
  }
  emission[m] = e;
}

// The absorbed power of each mesh added, as a float, to the step's
// absorption (what the host form adds from its double array).
__global__ void __launch_bounds__(kRadThreads)
    k_rad_add_absorption(int nm, const double* __restrict__ abs, float* __restrict__ absorption) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nm) absorption[m] += (float)abs[m];
}

// Caller-given powers per face side, copied to the walk entries.
__global__ void __launch_bounds__(kRadThreads)
    k_rad_expand(int W, const int* __restrict__ walk_face, const float* __restrict__ pface, float* __restrict__ p) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= W) return;
  const int f = walk_face[k];
  p[2 * k] = pface[2 * f];
  p[2 * k + 1] = pface[2 * f + 1];
}

// Status words of a pass (d_rad_info): [0] bits of the largest ray power of
// a walked side that emits, [1] bands in use, [2] bad-power flag, [3] lowest
// face with a bad power, [4] blocks of k_rad_bands done.
constexpr int kInfoWords = 5;

// Rays of each walk entry's face; entry W (the end) has none. Also resets
// the status words for the pass.
__global__ void __launch_bounds__(kRadThreads)
    k_rad_counts(int W, const int* __restrict__ walk_face, const long long* __restrict__ face_off,
                 long long* __restrict__ cnt, int* __restrict__ info) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    info[0] = 0;
    info[1] = 0;
    info[2] = 0;
    info[3] = 0x7fffffff;
    info[4] = 0;
  }
  if (k > W) return;
  long long c = 0;
  if (k < W) {
    const int f = walk_face[k];
    c = face_off[f + 1] - face_off[f];
  }
  cnt[k] = c;
}

// Entries the walk reaches (first position < total; radiance_cpu.cpp:170):
// the power check (negative or not finite: status bit, lowest such face) and
// the largest power of a side that emits rays, as bits (positive floats order
// as their bits, so the maximum does not depend on the order of the atomics).
__global__ void __launch_bounds__(kRadThreads)
    k_rad_entries(int W, long long total, const long long* __restrict__ start, const int* __restrict__ walk_face,
                  const long long* __restrict__ face_off, const int* __restrict__ front, const float* __restrict__ p,
                  int* __restrict__ info) {
  __shared__ unsigned sh[kRadThreads / 64];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned vmax = 0u;
  if (k < W && start[k] < total) {
    const int f = walk_face[k];
    const float pf = p[2 * k], pr = p[2 * k + 1];
    if (!(pf >= 0.f && pr >= 0.f && isfinite(pf) && isfinite(pr))) {
      atomicOr(&info[2], 1);
      atomicMin(&info[3], f);
    }
    const int cf = front[f], cr = (int)(face_off[f + 1] - face_off[f]) - cf;
    if (cf > 0 && pf > 0.f) vmax = max(vmax, __float_as_uint(pf));
    if (cr > 0 && pr > 0.f) vmax = max(vmax, __float_as_uint(pr));
  }
  const unsigned r = block_reduce(vmax, [](unsigned a, unsigned b) { return max(a, b); }, sh);
  if (threadIdx.x == 0 && r) atomicMax(reinterpret_cast<unsigned*>(&info[0]), r);
}

// Band of each emitting side: its binary order below the largest power in
// steps of kBandOrders (the last band takes everything smaller); per block and
// band the sum of count * power and the largest power (per wavefront by
// shuffles, then the wavefronts in order). The last block to finish then
// fixes the scales: per band the largest possible sum (every ray of the band
// absorbed by one mesh) and the largest single power give a scale that keeps
// sums below 2^63 and a single ray's value below 2^60, clear of the band
// bits. A ray of power p in band b is then rounded by at most 0.5 / scale_b,
// i.e. relatively by at most 2^-55 times the rays in the band, or 2^-52
// (2^kBandOrders / 2^60). info[1] = bands in use (0: no side emits). The
// partial sums are added in a fixed order, so the scales are the same every
// run.
__global__ void __launch_bounds__(kRadThreads)
    k_rad_bands(int W, long long total, const long long* __restrict__ start, const int* __restrict__ walk_face,
                const long long* __restrict__ face_off, const int* __restrict__ front, const float* __restrict__ p,
                int* __restrict__ info, unsigned char* __restrict__ band, double* __restrict__ part_sum,
                unsigned* __restrict__ part_max, double* __restrict__ scale) {
  constexpr int kWaves = kRadThreads / 64;
  __shared__ double shd[kWaves][kPowerBands];
  __shared__ unsigned shu[kWaves][kPowerBands];
  __shared__ int last;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned pmax_bits = (unsigned)info[0];
  const int e_max = pmax_bits ? pow_exponent(__uint_as_float(pmax_bits)) : 0;
  int b2[2] = {-1, -1};
  double v2[2] = {0.0, 0.0};
  unsigned u2[2] = {0u, 0u};
  if (k < W && start[k] < total) {
    const int f = walk_face[k];
    const int cnt[2] = {front[f], (int)(face_off[f + 1] - face_off[f]) - front[f]};
    for (int side = 0; side < 2; ++side) {
      const float ps = p[2 * k + side];
      const bool emits = ps > 0.f && cnt[side] > 0;
      const int b = emits ? min(kPowerBands - 1, (e_max - pow_exponent(ps)) / kBandOrders) : 0;
      band[2 * k + side] = (unsigned char)b;
      if (!emits) continue;
      b2[side] = b;
      v2[side] = (double)cnt[side] * ps;
      u2[side] = __float_as_uint(ps);
    }
  }
  for (int b = 0; b < kPowerBands; ++b) {
    double v = (b2[0] == b ? v2[0] : 0.0) + (b2[1] == b ? v2[1] : 0.0);
    unsigned u = max(b2[0] == b ? u2[0] : 0u, b2[1] == b ? u2[1] : 0u);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v += __shfl_down(v, off, 64);
      u = max(u, __shfl_down(u, off, 64));
    }
    if (lane == 0) {
      shd[w][b] = v;
      shu[w][b] = u;
    }
  }
  __syncthreads();
  if (threadIdx.x < kPowerBands) {
    const int b = threadIdx.x;
    double v = shd[0][b];
    unsigned u = shu[0][b];
    for (int i = 1; i < kWaves; ++i) {
      v += shd[i][b];
      u = max(u, shu[i][b]);
    }
    part_sum[(size_t)blockIdx.x * kPowerBands + b] = v;
    part_max[(size_t)blockIdx.x * kPowerBands + b] = u;
  }
  // the last block to finish reduces the partials
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(&info[4], 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  // thread t sums band t % 16 over the partials i = t / 16, t / 16 + 16, ...
  // in order; then the 16 chunk sums of each band in order
  const int b = threadIdx.x % kPowerBands, c = threadIdx.x / kPowerBands;
  constexpr int kChunks = kRadThreads / kPowerBands;
  double v = 0.0;
  unsigned u = 0u;
#pragma unroll 8
  for (int i = c; i < (int)gridDim.x; i += kChunks) {
    v += __builtin_nontemporal_load(&part_sum[(size_t)i * kPowerBands + b]);
    u = max(u, __builtin_nontemporal_load(&part_max[(size_t)i * kPowerBands + b]));
  }
  __shared__ double chd[kChunks][kPowerBands];
  __shared__ unsigned chu[kChunks][kPowerBands];
  chd[c][b] = v;
  chu[c][b] = u;
  __syncthreads();
  __shared__ int used[kPowerBands];
  if (threadIdx.x < kPowerBands) {
    double up = chd[0][b];
    unsigned um = chu[0][b];
    for (int i = 1; i < kChunks; ++i) {
      up += chd[i][b];
      um = max(um, chu[i][b]);
    }
    int e_sum = 0, e_one = 0;
    (void)frexp(up, &e_sum);                            // up < 2^e_sum
    (void)frexp((double)__uint_as_float(um), &e_one);  // the largest power < 2^e_one
    scale[b] = up > 0.0 ? ldexp(1.0, min(63 - e_sum, kBandShift - e_one)) : 1.0;
    used[b] = up > 0.0 ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int n = 0;
    for (int i = 0; i < kPowerBands; ++i)
      if (used[i]) n = i + 1;
    info[1] = n;
  }
}

// Fixed-point word of each walked side that emits (band in the top bits);
// the walk positions capped at the walk's end (the entries past it are never
// reached: no ray position lies at or past their start). Also zeroes the
// accumulator rows of the bands in use.
__global__ void __launch_bounds__(kRadThreads)
    k_rad_q(int W, long long total, long long walk_total, long long* __restrict__ start,
            const int* __restrict__ walk_face, const long long* __restrict__ face_off, const int* __restrict__ front,
            const float* __restrict__ p, const unsigned char* __restrict__ band, const double* __restrict__ scale,
            unsigned long long* __restrict__ q, unsigned long long* __restrict__ acc, int nm,
            const int* __restrict__ info) {
  const size_t n_acc = (size_t)info[1] * nm;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_acc; i += (size_t)gridDim.x * blockDim.x)
    acc[i] = 0ull;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > W) return;
  const long long s0 = start[k];
  if (k < W) {
    const int f = walk_face[k];
    const int cnt[2] = {front[f], (int)(face_off[f + 1] - face_off[f]) - front[f]};
    for (int side = 0; side < 2; ++side) {
      const float ps = p[2 * k + side];
      const int b = band[2 * k + side];
      // sides without rays are never read; their power may lie outside the bands
      q[2 * k + side] = s0 < total && ps > 0.f && cnt[side] > 0
                            ? ((unsigned long long)b << kBandShift) | round_even((double)ps * scale[b])
                            : 0ull;
    }
  }
  start[k] = min(s0, walk_total);
}

// Absorbed power per mesh: the band rows' fixed-point sums over their scales.
__global__ void __launch_bounds__(kRadThreads)
    k_rad_absorb(int nm, const unsigned long long* __restrict__ acc, const double* __restrict__ scale,
                 const int* __restrict__ info, double* __restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int n_bands = info[1];
  double a = 0.0;
  for (int b = 0; b < n_bands; ++b) a += (double)acc[(size_t)b * nm + m] / scale[b];
  out[m] = a;
}

// Lays out the radiance work buffers for the scene's faces and meshes (once
// per faces / meshes): the walk entries (faces of each mesh, mesh by mesh)
// and their tables, the accumulators, the reduction partials, the hipCUB scan
// storage, and the pinned staging of the per-step uploads and downloads.
int rad_prepare(ffk_scene* s, hipStream_t st) {
  if (s->rad_W >= 0) return FFK_OK;
  const int nf = s->n_faces, nm = s->n_meshes;
  long long W64 = 0;
  for (int m = 0; m < nm; ++m) W64 += s->h_meshes[3 * m + 1];
  FFK_REQUIRE(W64 < (1LL << 30), FFK_ERR_LIMIT, "radiance: %lld mesh faces", W64);
  const int W = (int)W64;
  const int n_parts = (W + 1 + kRadThreads - 1) / kRadThreads;
  size_t scan_bytes = 0;
  FFK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (long long*)nullptr, (long long*)nullptr, W + 1,
                                           st));
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += up(b);
    return o;
  };
  const size_t o_face = take(4 * (size_t)W + 4), o_mesh = take(4 * (size_t)W + 4), o_mw = take(4 * ((size_t)nm + 1)),
               o_cnt = take(8 * ((size_t)W + 1)),
               o_start = take(8 * ((size_t)W + 1)), o_q = take(16 * (size_t)W + 16), o_band = take(2 * (size_t)W + 2),
               o_acc = take(8 * (size_t)kPowerBands * nm), o_abs = take(8 * (size_t)nm),
               o_scale = take(8 * kPowerBands), o_p = take(8 * (size_t)W + 8), o_pface = take(8 * (size_t)nf),
               o_te = take(12 * (size_t)nm),
               o_area = take(4 * (size_t)nf), o_pu = take(4 * (size_t)n_parts * kPowerBands),
               o_pd = take(8 * (size_t)n_parts * kPowerBands), o_info = take(4 * kInfoWords), o_scan = take(scan_bytes);
  free_radiance(s);
  int r;
  if ((r = dmalloc(&s->d_rad_block, off)) < 0) return r;
  char* d = static_cast<char*>(s->d_rad_block);
  s->d_walk_face = reinterpret_cast<int*>(d + o_face);
  s->d_walk_mesh = reinterpret_cast<int*>(d + o_mesh);
  s->d_mesh_walk = reinterpret_cast<int*>(d + o_mw);
  s->d_walk_cnt = reinterpret_cast<long long*>(d + o_cnt);
  s->d_walk_start = reinterpret_cast<long long*>(d + o_start);
  s->d_walk_q = reinterpret_cast<unsigned long long*>(d + o_q);
  s->d_walk_band = reinterpret_cast<unsigned char*>(d + o_band);
  s->d_rad_acc = reinterpret_cast<unsigned long long*>(d + o_acc);
  s->d_rad_abs = reinterpret_cast<double*>(d + o_abs);
  s->d_rad_scale = reinterpret_cast<double*>(d + o_scale);
  s->d_rad_p = reinterpret_cast<float*>(d + o_p);
  s->d_rad_pface = reinterpret_cast<float*>(d + o_pface);
  s->d_rad_te = reinterpret_cast<float*>(d + o_te);
  s->d_rad_area = reinterpret_cast<float*>(d + o_area);
  s->d_rad_part_u = reinterpret_cast<unsigned*>(d + o_pu);
  s->d_rad_part_d = reinterpret_cast<double*>(d + o_pd);
  s->d_rad_info = reinterpret_cast<int*>(d + o_info);
  s->d_rad_scan = d + o_scan;
  s->rad_scan_bytes = scan_bytes;
  s->rad_parts = n_parts;
  // pinned staging: T and eps (3 nm floats), ray powers (2 nf floats), then
  // the absorption (nm doubles) and the status words (4 ints)
  const size_t pin = up(12 * (size_t)nm) + up(8 * (size_t)nf) + up(8 * (size_t)nm) + 4 * kInfoWords;
  FFK_HIP(hipHostMalloc(&s->h_rad_pin, pin, hipHostMallocDefault));
  // the walk entries' faces and meshes and the face areas, once
  std::vector<int> wf((size_t)W + 1, 0), wm((size_t)W + 1, 0), mw((size_t)nm + 1, 0);
  int k = 0;
  for (int m = 0; m < nm; ++m) {
    mw[m] = k;
    for (int f = s->h_meshes[3 * m]; f < s->h_meshes[3 * m] + s->h_meshes[3 * m + 1]; ++f) {
      wf[k] = f;
      wm[k++] = m;
    }
  }
  mw[nm] = k;
  FFK_HIP(hipMemcpyAsync(s->d_walk_face, wf.data(), 4 * ((size_t)W + 1), hipMemcpyHostToDevice, st));
  FFK_HIP(hipMemcpyAsync(s->d_walk_mesh, wm.data(), 4 * ((size_t)W + 1), hipMemcpyHostToDevice, st));
  FFK_HIP(hipMemcpyAsync(s->d_mesh_walk, mw.data(), 4 * ((size_t)nm + 1), hipMemcpyHostToDevice, st));
  FFK_HIP(hipMemcpyAsync(s->d_rad_area, s->h_area.data(), 4 * (size_t)nf, hipMemcpyHostToDevice, st));
  FFK_HIP(hipStreamSynchronize(st));  // the vector above goes out of scope
  s->rad_W = W;
  return FFK_OK;
}

void free_radiance(ffk_scene* s) {
  if (s->rad_pin_ev) (void)hipEventSynchronize(s->rad_pin_ev);  // no transfer through the pinned block in flight
  dfree(s->d_rad_block);
  if (s->h_rad_pin) (void)hipHostFree(s->h_rad_pin);
  s->d_rad_block = nullptr;
  s->h_rad_pin = nullptr;
  s->rad_W = -1;
  s->rad_parts = 0;
  s->d_walk_face = nullptr;
  s->d_walk_mesh = nullptr;
  s->d_mesh_walk = nullptr;
  s->d_walk_cnt = nullptr;
  s->d_walk_start = nullptr;
  s->d_walk_q = nullptr;
  s->d_walk_band = nullptr;
  s->d_rad_acc = nullptr;
  s->d_rad_abs = nullptr;
  s->d_rad_scale = nullptr;
  s->d_rad_p = nullptr;
  s->d_rad_pface = nullptr;
  s->d_rad_te = nullptr;
  s->d_rad_area = nullptr;
  s->d_rad_part_u = nullptr;
  s->d_rad_part_d = nullptr;
  s->d_rad_info = nullptr;
  s->d_rad_scan = nullptr;
}

// Pinned staging regions of rad_prepare's layout.
inline float* pin_te(ffk_scene* s) { return static_cast<float*>(s->h_rad_pin); }
inline float* pin_p(ffk_scene* s) {
  return reinterpret_cast<float*>(static_cast<char*>(s->h_rad_pin) + ((12 * (size_t)s->n_meshes + 255) & ~(size_t)255));
}
inline double* pin_abs(ffk_scene* s) {
  const size_t o = ((12 * (size_t)s->n_meshes + 255) & ~(size_t)255) + ((8 * (size_t)s->n_faces + 255) & ~(size_t)255);
  return reinterpret_cast<double*>(static_cast<char*>(s->h_rad_pin) + o);
}
inline int* pin_info(ffk_scene* s) {
  return reinterpret_cast<int*>(reinterpret_cast<char*>(pin_abs(s)) + ((8 * (size_t)s->n_meshes + 255) & ~(size_t)255));
}

// The device pass over the walk entries' ray powers in d_rad_p: walk table, bands, scales
// and fixed-point words, the emit-cast-deposit kernel, the per-mesh sums in
// d_rad_abs; the status words are copied to the pinned block (read them once
// the stream has drained).
int rad_enqueue(ffk_scene* s, unsigned long long seed, unsigned long long* d_stats, hipStream_t st) {
  const int nm = s->n_meshes, W = s->rad_W, nb = (W + kRadThreads) / kRadThreads;
  const long long total = s->n_rays;
  const long long walk_total = s->h_mesh_off[nm];  // the walk's end: min(rays of the meshes' faces, total)
  int* info = pin_info(s);
  k_rad_counts<<<nb, kRadThreads, 0, st>>>(W, s->d_walk_face, s->d_face_off, s->d_walk_cnt, s->d_rad_info);
  size_t scan_bytes = s->rad_scan_bytes;
  FFK_HIP(hipcub::DeviceScan::ExclusiveSum(s->d_rad_scan, scan_bytes, s->d_walk_cnt, s->d_walk_start, W + 1, st));
  k_rad_entries<<<nb, kRadThreads, 0, st>>>(W, total, s->d_walk_start, s->d_walk_face, s->d_face_off, s->d_front,
                                            s->d_rad_p, s->d_rad_info);
  k_rad_bands<<<nb, kRadThreads, 0, st>>>(W, total, s->d_walk_start, s->d_walk_face, s->d_face_off, s->d_front,
                                          s->d_rad_p, s->d_rad_info, s->d_walk_band, s->d_rad_part_d,
                                          s->d_rad_part_u, s->d_rad_scale);
  k_rad_q<<<nb, kRadThreads, 0, st>>>(W, total, walk_total, s->d_walk_start, s->d_walk_face, s->d_face_off,
                                      s->d_front, s->d_rad_p, s->d_walk_band, s->d_rad_scale, s->d_walk_q,
                                      s->d_rad_acc, nm, s->d_rad_info);
  FFK_HIP(hipGetLastError());
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  if (walk_total > 0 && s->tune_refill > 0 && s->tune_pool > 0) {
    // Whole rounds of the blocks the device holds at once (the last round is
    // not left part empty), at most 16384 rays and at least a ray per lane
    // per block.
    auto go = [&](auto kernel) {
      static int per_cu_cached = 0;  // per kernel instance: resident blocks per CU (the device's one property)
      int per_cu = per_cu_cached;
      if (per_cu < 1) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1)
          per_cu = 1;
        per_cu_cached = per_cu;
      }
      const long long resident = (long long)per_cu * device_cu_count();
      const long long rounds = std::max<long long>(1, (walk_total + resident * 16384 - 1) / (resident * 16384));
      const long long chunk = std::max<long long>(256, (walk_total + resident * rounds - 1) / (resident * rounds));
      const unsigned blocks = (unsigned)((walk_total + chunk - 1) / chunk);
      kernel<<<blocks, 256, 0, st>>>(s->desc(), s->d_walk_start, s->d_walk_face, W, s->d_walk_q, walk_total, chunk,
                                     k0, k1, s->d_rad_acc, d_stats, s->tune_refill,
                                     pool_arg(s->tune_pool, s->tune_walk));
    };
    const bool jump = (s->tune_walk & kWalkJump) != 0;
    if (d_stats) jump ? go(k_radiance_pooled<true, 1>) : go(k_radiance_pooled<true, 2>);
    else jump ? go(k_radiance_pooled<false, 1>) : go(k_radiance_pooled<false, 2>);
  } else if (walk_total > 0) {
    const long long blocks_needed = (walk_total + 255) / 256;
    const int blocks = (int)std::max<long long>(1, std::min<long long>(blocks_needed, 64LL * device_cu_count()));
    if (d_stats)
      k_radiance<true><<<blocks, 256, 0, st>>>(s->desc(), s->d_walk_start, s->d_walk_face, W, s->d_walk_q,
                                               walk_total, k0, k1, s->d_rad_acc, d_stats);
    else
      k_radiance<false><<<blocks, 256, 0, st>>>(s->desc(), s->d_walk_start, s->d_walk_face, W, s->d_walk_q,
                                                walk_total, k0, k1, s->d_rad_acc, nullptr);
  }
  k_rad_absorb<<<(nm + kRadThreads - 1) / kRadThreads, kRadThreads, 0, st>>>(nm, s->d_rad_acc, s->d_rad_scale,
                                                                           s->d_rad_info, s->d_rad_abs);
  FFK_HIP(hipGetLastError());
  FFK_HIP(hipMemcpyAsync(info, s->d_rad_info, 4 * kInfoWords, hipMemcpyDeviceToHost, st));
  if (!s->rad_pin_ev) FFK_HIP(hipEventCreateWithFlags(&s->rad_pin_ev, hipEventDisableTiming));
  FFK_HIP(hipEventRecord(s->rad_pin_ev, st));
  return FFK_OK;
}

// Status of the last pass, once its stream has drained: FFK_ERR_ARG when a
// walked ray power was negative or not finite.
int rad_status(ffk_scene* s) {
  const int* info = pin_info(s);
  FFK_REQUIRE(info[2] == 0, FFK_ERR_ARG, "ffk_radiance: ray power of face %d is negative or not finite", info[3]);
  return FFK_OK;
}

// rad_enqueue, then the absorption copied to h_absorption (left untouched
// when a walked ray power is negative or not finite: FFK_ERR_ARG).
int rad_pass(ffk_scene* s, unsigned long long seed, double* h_absorption, unsigned long long* d_stats,
             hipStream_t st) {
  int r;
  if ((r = rad_enqueue(s, seed, d_stats, st)) < 0) return r;
  double* ab = pin_abs(s);
  FFK_HIP(hipMemcpyAsync(ab, s->d_rad_abs, 8 * (size_t)s->n_meshes, hipMemcpyDeviceToHost, st));
  FFK_HIP(hipStreamSynchronize(st));
  if ((r = rad_status(s)) < 0) return r;
  std::memcpy(h_absorption, ab, 8 * (size_t)s->n_meshes);
  return FFK_OK;
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

// thermal_equation/radiance_cpu.cpp:148-178: the emission added to each mesh
// (float, face by face), and, when h_ray_power is given, the power of one
// front / rear ray of each face (k_rad_powers computes the same on the device).
void powers_impl(const float* h_faces, const float* h_area, int n_faces, const int* h_meshes, int n_meshes,
                 const float* h_eps2, const float* h_T, const int* h_front, const int* h_rear, float* h_ray_power,
                 float* h_emission) {
  if (h_ray_power)
    for (int i = 0; i < 2 * n_faces; ++i) h_ray_power[i] = 0.f;
  for (int m = 0; m < n_meshes; ++m) {
    const float T = h_T[m];
    const float power_density = kSigma * (T * T * T * T);
    const float front_density = power_density * h_eps2[2 * m];
    const float rear_density = power_density * h_eps2[2 * m + 1];
    for (int f = h_meshes[3 * m]; f < h_meshes[3 * m] + h_meshes[3 * m + 1]; ++f) {
      const float a = area_of(h_area, h_faces, f);
      const float front_emission = front_density * a;
      const float rear_emission = rear_density * a;
      h_emission[m] += front_emission + rear_emission;
      if (!h_ray_power) continue;
      h_ray_power[2 * f] = h_front[f] > 0 ? front_emission / h_front[f] : 0.f;
      h_ray_power[2 * f + 1] = h_rear[f] > 0 ? rear_emission / h_rear[f] : 0.f;
    }
  }
}

// The host part of a radiance step (radiance_cpu.cpp:96-146): weights and
// their float total in mesh order, the emitted counts and, when they changed,
// the emission layout; then T and eps to the device (pinned staging, reused
// only once its last upload has run) and the walk entries' ray powers there.
int step_prepare(ffk_scene* s, const float* h_eps2, const float* h_T, int n_rays, int frame, hipStream_t st) {
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
    if ((r = set_emission(s, s->rad_front.data(), s->rad_rear.data(), frame, nullptr, st, true)) < 0) return r;
  }
  if ((r = rad_prepare(s, st)) < 0) return r;
  // T and eps to the device, where the ray powers are computed (their bits
  // are powers_impl's); the staging is rewritten only once the last transfer
  // through it (a device step's status copy, at the latest) has run
  if (!s->rad_pin_ev) FFK_HIP(hipEventCreateWithFlags(&s->rad_pin_ev, hipEventDisableTiming));
  FFK_HIP(hipEventSynchronize(s->rad_pin_ev));
  float* te = pin_te(s);
  std::memcpy(te, h_T, 4 * (size_t)nm);
  std::memcpy(te + nm, h_eps2, 8 * (size_t)nm);
  FFK_HIP(hipMemcpyAsync(s->d_rad_te, te, 12 * (size_t)nm, hipMemcpyHostToDevice, st));
  FFK_HIP(hipEventRecord(s->rad_pin_ev, st));
  if (s->rad_W > 0)
    k_rad_powers<<<(s->rad_W + kRadThreads - 1) / kRadThreads, kRadThreads, 0, st>>>(
        s->rad_W, s->d_walk_face, s->d_walk_mesh, s->d_rad_te, nm, s->d_rad_area, s->d_face_off, s->d_front,
        s->d_rad_p);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
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
  int r;
  if ((r = rad_prepare(s, st)) < 0) return r;
  float* pp = pin_p(s);
  std::memcpy(pp, h_ray_power, 8 * (size_t)s->n_faces);
  FFK_HIP(hipMemcpyAsync(s->d_rad_pface, pp, 8 * (size_t)s->n_faces, hipMemcpyHostToDevice, st));
  if (s->rad_W > 0)
    k_rad_expand<<<(s->rad_W + kRadThreads - 1) / kRadThreads, kRadThreads, 0, st>>>(s->rad_W, s->d_walk_face,
                                                                                    s->d_rad_pface, s->d_rad_p);
  FFK_HIP(hipGetLastError());
  return rad_pass(s, seed, h_absorption, d_stats, st);
}

int ffk_radiance_step(ffk_scene* s, const float* h_eps2, const float* h_T, int n_rays, int frame,
                      unsigned long long seed, float* h_emission, double* h_absorption,
                      unsigned long long* d_stats, void* stream) {
  FFK_REQUIRE(s && s->has_faces && s->has_meshes && h_eps2 && h_T && h_emission && h_absorption && n_rays >= 0,
              FFK_ERR_ARG, "ffk_radiance_step: bad args");
  FFK_REQUIRE(s->has_grid, FFK_ERR_ARG, "ffk_radiance_step: the grid must be built");
  hipStream_t st = as_stream(stream);
  int r;
  if ((r = step_prepare(s, h_eps2, h_T, n_rays, frame, st)) < 0) return r;
  // meanwhile the per-mesh emission is added to h_emission here
  // (radiance_cpu.cpp adds to the task's emission)
  powers_impl(s->h_faces.data(), s->h_area.data(), s->n_faces, s->h_meshes.data(), s->n_meshes, h_eps2, h_T,
              nullptr, nullptr, nullptr, h_emission);
  return rad_pass(s, seed, h_absorption, d_stats, st);
}

int ffk_radiance_step_dev(ffk_scene* s, const float* h_eps2, const float* h_T, int n_rays, int frame,
                          unsigned long long seed, float* d_emission, float* d_absorption, void* stream) {
  FFK_REQUIRE(s && s->has_faces && s->has_meshes && h_eps2 && h_T && d_emission && d_absorption && n_rays >= 0,
              FFK_ERR_ARG, "ffk_radiance_step_dev: bad args");
  FFK_REQUIRE(s->has_grid, FFK_ERR_ARG, "ffk_radiance_step_dev: the grid must be built");
  hipStream_t st = as_stream(stream);
  int r;
  if ((r = step_prepare(s, h_eps2, h_T, n_rays, frame, st)) < 0) return r;
  const int nm = s->n_meshes, nb = (nm + kRadThreads - 1) / kRadThreads;
  k_rad_emission<<<nb, kRadThreads, 0, st>>>(nm, s->d_mesh_walk, s->d_walk_face, s->d_rad_te, s->d_rad_area,
                                             d_emission);
  FFK_HIP(hipGetLastError());
  if ((r = rad_enqueue(s, seed, nullptr, st)) < 0) return r;
  k_rad_add_absorption<<<nb, kRadThreads, 0, st>>>(nm, s->d_rad_abs, d_absorption);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_radiance_status(ffk_scene* s) {
  FFK_REQUIRE(s && s->rad_W >= 0, FFK_ERR_ARG, "ffk_radiance_status: no radiance pass on this scene");
  return rad_status(s);
}

}  // extern "C"
