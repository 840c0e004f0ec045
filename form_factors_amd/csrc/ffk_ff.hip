// Emission, exact tallies, normalisation and the fused form-factor kernel.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ffk_trace.hpp"

namespace ffk {

// ------------------------------------------------------------- emission --

__global__ void __launch_bounds__(256) k_emit_malley(scene_desc S, uint32_t k0, uint32_t k1,
                                                     float* __restrict__ rays) {
  long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.n_rays) return;
  int f = find_face(S.face_off, S.n_faces, g, 0);
  int j = (int)(g - gld(S.face_off + f));
  v3 o, e;
  malley_ray(S, f, j, k0, k1, &o, &e);
  float* r = rays + 6 * g;
  r[0] = o.x; r[1] = o.y; r[2] = o.z;
  r[3] = e.x; r[4] = e.y; r[5] = e.z;
}

// emission/parallel_rays_cpu.cpp:88-106: a point on a disk of the given
// radius (rho = radius * sqrt(U), angle 2*pi*U'), rotated from z onto dir,
// placed at -distance along it around center; the end point is 0.1 along dir.
__global__ void __launch_bounds__(256) k_emit_parallel(v3 dir, v3 center, float distance, float radius,
                                                       const float* __restrict__ rot, long long n,
                                                       uint32_t k0, uint32_t k1, float* __restrict__ rays) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u4 r = philox(u4{(uint32_t)i, 0u, 1u, (uint32_t)(i >> 32)}, k0, k1);
  float rho = radius * sqrtf(u01(r.x));
  float s, c;
  sincos_turns(u01(r.y), &s, &c);
  float x = rho * c, y = rho * s;
  v3 o = mat_vec(rot, mk(x, y, -distance)) + center;
  v3 e = o + dir * 0.1f;
  float* q = rays + 6 * i;
  q[0] = o.x; q[1] = o.y; q[2] = o.z;
  q[3] = e.x; q[4] = e.y; q[5] = e.z;
}

// --------------------------------------------------------------- tally --

__global__ void k_fill_u64(unsigned long long* __restrict__ p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0ull;
}

__global__ void k_fill_i32(int* __restrict__ p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Mesh m whose tally range [mesh_off[m], mesh_off[m+1]) holds ray g.
__device__ inline int find_mesh(const long long* __restrict__ off, int n, long long g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (off[mid + 1] > g) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_tally(scene_desc S, const int* __restrict__ hit, long long n,
                                               unsigned long long* __restrict__ tally) {
  long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  if (g >= S.mesh_off[S.n_meshes]) return;  // rays past the mesh walk are not counted
  int h = hit[g];
  if (h < 0) return;
  int dst = gld(S.face2mesh + h);
  if (dst < 0) return;
  int src = find_mesh(S.mesh_off, S.n_meshes, g);
  atomicAdd(&tally[(size_t)src * S.n_meshes + dst], 1ull);
}

// F[m, j] = (float)count * (1.f / (float)rays_of_m). While a cell's count is
// below 2^24 this equals the reference's float accumulation `F += 1.f`
// (cpu_system.cpp:181-191) bit for bit. Above it the two part on purpose: the
// reference's sum stops at 2^24 (adding 1 no longer changes it), while here
// the exact integer count is rounded once to the nearest float.
__device__ inline float normalise_one(unsigned long long cnt, float ratio) { return (float)cnt * ratio; }

template <int HIST>
__device__ inline uint32_t hist_count(const uint32_t* hist, int j) {
  return HIST == 1 ? hist[j] : ((hist[j >> 1] >> ((j & 1) * 16)) & 0xffffu);
}

// Writes a tile's F row from its LDS histogram: four cells per lane in 16-byte
// non-temporal stores where the row is 16-byte aligned (the row is written
// once and not read back by this kernel, so it should not displace the scene
// tables from L2), the rest one cell per lane.
template <int HIST>
__device__ inline void write_row(float* F, const uint32_t* hist, int N, float ratio) {
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  int j0 = 0;
  if ((reinterpret_cast<uintptr_t>(F) & 15u) == 0) {
    const int n4 = N >> 2;
    f32x4* F4 = reinterpret_cast<f32x4*>(F);
    for (int k = threadIdx.x; k < n4; k += blockDim.x) {
      uint32_t c0, c1, c2, c3;
      if (HIST == 1) {
        const uint4 h = reinterpret_cast<const uint4*>(hist)[k];
        c0 = h.x, c1 = h.y, c2 = h.z, c3 = h.w;
      } else {
        const uint2 h = reinterpret_cast<const uint2*>(hist)[k];
        c0 = h.x & 0xffffu, c1 = h.x >> 16, c2 = h.y & 0xffffu, c3 = h.y >> 16;
      }
      f32x4 v;
      v.x = normalise_one(c0, ratio);
      v.y = normalise_one(c1, ratio);
      v.z = normalise_one(c2, ratio);
      v.w = normalise_one(c3, ratio);
      __builtin_nontemporal_store(v, F4 + k);
    }
    j0 = n4 << 2;
  }
  for (int j = j0 + threadIdx.x; j < N; j += blockDim.x) F[j] = normalise_one(hist_count<HIST>(hist, j), ratio);
}

__global__ void __launch_bounds__(256) k_normalise_rows(const long long* __restrict__ mesh_off, int N,
                                                        const int* __restrict__ rows, int n_rows,
                                                        const unsigned long long* __restrict__ tally,
                                                        float* __restrict__ F) {
  const int r = blockIdx.x;
  if (r >= n_rows) return;
  const int m = rows ? rows[r] : r;
  const float ratio = 1.f / (float)(int)(mesh_off[m + 1] - mesh_off[m]);
  for (int j = threadIdx.x; j < N; j += blockDim.x)
    F[(size_t)m * N + j] = normalise_one(tally[(size_t)m * N + j], ratio);
}

__global__ void k_zero_rows(int N, const int* __restrict__ rows, int n_rows,
                            unsigned long long* __restrict__ tally) {
  const int r = blockIdx.x;
  if (r >= n_rows) return;
  const int m = rows[r];
  for (int j = threadIdx.x; j < N; j += blockDim.x) tally[(size_t)m * N + j] = 0ull;
}

// ------------------------------------------------------- fused kernel --
// HIST: 0 = global uint64 atomics per hit; 1 = uint32 LDS histogram of the
// tile's row; 2 = two uint16 counters per LDS word (tiles <= 65535 rays).
// PERSIST: blocks pull tiles from a queue (batched launch) instead of one
// tile per block (plain launch).
// Idle lanes (of 64) that trigger a refill in the fused kernel; 0 = off.
constexpr int kDefaultRefill = 32;
// Pooled triangle tests run once waiting lanes x kDefaultPool >= active
// lanes; 0 = each lane tests its own cells.
constexpr int kDefaultPool = 2;
// Hit faces a pooled wavefront queues before counting them: two rounds of 64
// in a one-scene launch; one in a persistent launch, whose LDS classes would
// lose resident blocks to the larger queue.
__host__ __device__ constexpr int hit_queue(bool persist) { return persist ? 64 : 128; }

struct fused_args {
  const scene_desc* descs;
  const tile* tiles;
  int n_tiles;

  int* queue;
  const unsigned long long* seeds;
  unsigned long long* const* tallies;
  float* const* Fs;
  unsigned long long* stats;
  int refill;  // 0: one ray per lane at a time; else the idle-lane threshold of the refill loop
  int pool;    // 0: per-lane tests; else pool the tests once waiting lanes x pool >= active lanes
  int walk;    // bit 0 (kWalkJump): always the jump code, else a per-wavefront choice of the
               // empty-cell step; bit 1 (kWalkWhole): pooled tests dealt in whole chunks
};

// REFILL: lanes walk one grid cell per iteration and, once at least
// `refill` lanes of the wavefront are idle, the wavefront reserves that many
// rays of the tile with one LDS atomic and the idle lanes start them. Without
// it each lane carries one ray at a time and the wavefront waits for its
// slowest ray before the next round, which leaves most lanes idle on
// incoherent rays. Rays are tallied as integers, so the order in which they
// are cast does not change F.
// At least 6 wavefronts per SIMD (at most 80 VGPRs): three 512-lane blocks
// per CU, the most the row histograms of large scenes leave LDS for.
template <int HIST, bool PERSIST, bool STATS, int SCHED>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6)))
    k_ff_fused(const scene_desc* __restrict__ descs, const tile* __restrict__ tiles, int n_tiles,
               int* __restrict__ queue, const unsigned long long* __restrict__ seeds,
               unsigned long long* const* __restrict__ tallies, float* const* __restrict__ Fs,
               unsigned long long* __restrict__ stats, int refill, int pool) {
  // Read-only, restrict-qualified tables indexed by wave-uniform values are
  // read with scalar loads: the scene descriptor lives in SGPRs.
  // 16-byte aligned: the dynamic region follows the static words below, and
  // the pooled-test best-hit words after the histogram take 64-bit LDS
  // atomics, which need natural alignment.
  extern __shared__ __align__(16) uint32_t hist[];
  __shared__ int s_tile;
  __shared__ int s_next;      // next ray of the tile (REFILL)
  __shared__ int s_faces[2];  // faces of the tile's first and last rays
  uint32_t nt = 0, nc = 0;
  int t_idx = blockIdx.x;
  while (true) {
    if (PERSIST) {
      __syncthreads();
      if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1);
      __syncthreads();
      t_idx = __builtin_amdgcn_readfirstlane(s_tile);
    }
    if (t_idx >= n_tiles) break;
    const tile T = tiles[t_idx];
    const int sc = __builtin_amdgcn_readfirstlane(T.scene);
    const scene_desc S = descs[sc];
    const int N = S.n_meshes;
    const int words = HIST == 1 ? N : (N + 1) >> 1;
    if (HIST != 0)
      for (int w = threadIdx.x; w < words; w += blockDim.x) hist[w] = 0u;
    if (SCHED != 0 && threadIdx.x == 0) {
      s_next = 0;
      s_faces[0] = T.f_lo;
      s_faces[1] = T.f_hi;
    }
    if (HIST != 0 || SCHED != 0) __syncthreads();
    const unsigned long long seed = seeds[sc];
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    unsigned long long* tally = tallies ? tallies[sc] : nullptr;
    int hint = T.f_lo;
    auto count_hit = [&](const hit_rec& h) {
      if (h.f >= 0) {
        const int m = gld(gidx(S.face2mesh, (uint32_t)h.f));
        if (m >= 0) {
          if (HIST == 0) atomicAdd(&tally[(size_t)T.mesh * N + m], 1ull);
          else if (HIST == 1) atomicAdd(&hist[m], 1u);
          else atomicAdd(&hist[m >> 1], (m & 1) ? 65536u : 1u);
        }
      }
    };
    if (SCHED >= 2) {
      // As SCHED 1, with the triangle tests of the wavefront's waiting lanes
      // pooled over its 64 lanes (pooled_walk); SCHED 3 leaves every empty
      // cell by the jump code.
      unsigned long long* const best_all =
          reinterpret_cast<unsigned long long*>(hist + ((HIST == 0 ? 0 : (words + 1)) & ~1));
      const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
      unsigned long long* best = best_all + kPoolWords * wave;
      // Each wavefront's hit faces waiting to be counted, fewer than 64
      // between iterations. Counted 64 at a time, the mesh loads of 64 hits
      // are in flight together instead of holding up the walk once per
      // iteration; the counts are integers, so their order does not matter.
      constexpr int kQueue = hit_queue(PERSIST);
      int* const queue = reinterpret_cast<int*>(best_all + kPoolWords * (blockDim.x >> 6)) + kQueue * wave;
      int n_queued = 0;     // wave-uniform: changed only where every lane runs
      int pend_face = -1;  // the hit of the ray this lane ended in this iteration
      pooled_walk<STATS, SCHED == 3 ? 1 : 2>(
          S, best, &s_next, T.g1 - T.g0, refill, pool,
          [&](long long i, v3& o, v3& dir) {
            const long long g = T.g0 + i;
            // A tile within one face needs neither the search nor the face's
            // first ray (two loads fewer ahead of the ray's own loads).
            int f, j;
            if (T.f_lo == T.f_hi && T.j0 >= 0) {
              f = T.f_lo;
              j = (int)i + T.j0;
            } else {
              f = find_face_in(S.face_off, s_faces, g, hint);
              hint = f;
              j = (int)(g - gld(gidx(S.face_off, (uint32_t)f)));
            }
            v3 e;
            malley_ray(S, f, j, k0, k1, &o, &e);
            dir = e - o;
          },
          [&](long long, v3, v3, unsigned long long key) {
            float d2;
            decode_key(key, d2, pend_face);
          },
          nt, nc,
          [&](bool last) {
            const unsigned long long hits = __ballot(pend_face >= 0);
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(hits >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)hits, 0u));
            const int n_new = __popcll(hits);
            if constexpr (kQueue >= 128) {
              // Room for this iteration's hits beside fewer than 64 queued;
              // then whole rounds of 64 from the top (and the rest at the end).
              if (pend_face >= 0) queue[n_queued + rank] = pend_face;
              n_queued += n_new;
              if (n_queued < 64 && !(last && n_queued > 0)) {
                pend_face = -1;
                return;
              }
              __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the queue as every lane wrote it
              while (n_queued >= 64 || (last && n_queued > 0)) {
                const int take = min(n_queued, 64);
                n_queued -= take;
                if (lane < take) {
                  hit_rec h = {0.f, queue[n_queued + lane], mk(0.f, 0.f, 0.f)};
                  count_hit(h);
                }
              }
              __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // read before the next writes
            } else {
              // This iteration's hits fill the queue up to 64; a full queue
              // is counted and the remaining hits start the next one.
              const int room = kQueue - n_queued;
              if (pend_face >= 0 && rank < room) queue[n_queued + rank] = pend_face;
              if (n_new >= room || (last && n_queued + n_new > 0)) {
                const int take = min(n_queued + n_new, kQueue);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the queue as every lane wrote it
                if (lane < take) {
                  hit_rec h = {0.f, queue[lane], mk(0.f, 0.f, 0.f)};
                  count_hit(h);
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // read before the next writes
                if (pend_face >= 0 && rank >= room) queue[rank - room] = pend_face;
                n_queued = n_queued + n_new - take;
                if (last && n_queued > 0) {  // the hits that did not fit the last round
                  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                  if (lane < n_queued) {
                    hit_rec h = {0.f, queue[lane], mk(0.f, 0.f, 0.f)};
                    count_hit(h);
                  }
                  n_queued = 0;
                }
              } else {
                n_queued += n_new;
              }
            }
            pend_face = -1;
          });
    } else if (SCHED == 1) {
      const long long n = T.g1 - T.g0;
      const int lane = threadIdx.x & 63;
      bool have = false, more = true;  // `more` is wave-uniform
      v3 o = mk(0.f, 0.f, 0.f), dir = mk(0.f, 0.f, 0.f);
      float ix = 0.f, iy = 0.f, iz = 0.f, dlen = 0.f, tnx = 0.f, tny = 0.f, tnz = 0.f;
      int cx = 0, cy = 0, cz = 0, cell = 0;
      hit_rec h = {3.402823466e38f, -1, mk(0.f, 0.f, 0.f)};
      mailbox mb;
      mb.clear();
      while (true) {
        const unsigned long long idle = __ballot(!have);
        const int n_idle = __popcll(idle);
        if (more && n_idle >= refill) {
          int base = 0;
          if (lane == 0) base = atomicAdd(&s_next, n_idle);
          base = __shfl(base, 0, 64);
          if ((long long)base + n_idle >= n) more = false;
          if (!have) {
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const long long off = (long long)base + rank;
            if (off < n) {
              const long long g = T.g0 + off;
              const int f = T.f_lo == T.f_hi ? T.f_lo : find_face_in(S.face_off, s_faces, g, hint);
              hint = f;
              v3 e;
              malley_ray(S, f, (int)(g - gld(S.face_off + f)), k0, k1, &o, &e);
              dir = e - o;
              h.d2 = 3.402823466e38f;
              h.f = -1;
              mb.clear();
              have = walk_begin(S.g, o, dir, ix, iy, iz, dlen, cx, cy, cz, tnx, tny, tnz, cell);
            }
          }
        }
        if (!more && __ballot(have) == 0ull) break;
        if (have && !walk_cell<STATS>(S, o, dir, ix, iy, iz, dlen, cx, cy, cz, tnx, tny, tnz, cell, h, mb, nt, nc)) {
          count_hit(h);
          have = false;
        }
      }
    } else {
      for (long long g = T.g0 + threadIdx.x; g < T.g1; g += blockDim.x) {
        const int f = find_face(S.face_off, S.n_faces, g, hint);  // s_faces is not filled here
        hint = f;
        v3 o, e;
        malley_ray(S, f, (int)(g - gld(S.face_off + f)), k0, k1, &o, &e);
        hit_rec h = {3.402823466e38f, -1, mk(0.f, 0.f, 0.f)};
        cast_grid<STATS>(S, o, e - o, h, nt, nc);
        count_hit(h);
      }
    }
    if (HIST != 0) {
      __syncthreads();
      if (T.direct) {
        write_row<HIST>(Fs[sc] + (size_t)T.mesh * N, hist, N, 1.f / (float)(int)(T.g1 - T.g0));
      } else {
        unsigned long long* row = tally + (size_t)T.mesh * N;
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
          const uint32_t c = hist_count<HIST>(hist, j);
          if (c) atomicAdd(&row[j], (unsigned long long)c);
        }
      }
    }
    if (!PERSIST) break;
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

// ---------------------------------------------------------- planning --

// Dynamic LDS a fused launch may use: the block's LDS less the kernel's
// static words (s_tile, s_next), padded to the 16-byte alignment of the
// dynamic region.
constexpr int kStaticLds = 16;

static int max_lds_bytes() {
  static int cached = 0;
  if (!cached) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
      cached = (int)std::max<size_t>(p.sharedMemPerBlock, 65536) - kStaticLds;
    else
      cached = 65536 - kStaticLds;
  }
  return cached;
}

// Histogram mode for a row of N meshes and tiles of at most max_tile rays.
// Above this many bytes of 16-bit row histogram a block would hold more than
// a third of the CU's LDS, and fewer than three blocks of 512 fit: the global
// uint64 tally (N^2 words zeroed and normalised, one L2 atomic per hit) then
// runs faster than the LDS histogram (measured on the 64-scene set: scenes of
// 24,577-81,920 meshes 12-17 % faster, below that slower).
constexpr long long kMaxDefaultHist = 49152;

static int pick_hist(int N, long long max_tile, int requested) {
  const int lds = max_lds_bytes();
  if (requested == 0) return 0;
  if (requested == 1 && 4LL * N <= lds) return 1;
  if (requested == 2 && 2LL * ((N + 1) / 2) * 2 <= lds && max_tile <= 65535) return 2;
  if (4LL * N <= std::min(65536, lds)) return 1;
  if (4LL * ((N + 1) / 2) <= std::min<long long>(kMaxDefaultHist, lds) && max_tile <= 65535) return 2;
  return 0;
}

static size_t hist_bytes(int mode, int N) {
  if (mode == 1) return 4 * (size_t)N;
  if (mode == 2) return 4 * (size_t)((N + 1) / 2);
  return 0;
}

template <int H, bool P, bool S, int R>
static void set_lds_attr(size_t bytes) {
  if (bytes + kStaticLds > 65536)
    (void)hipFuncSetAttribute((const void*)&k_ff_fused<H, P, S, R>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Dynamic LDS of a launch: the row histogram, then (pooled tests) the
// wavefronts' kPoolWords 64-bit words each (best hits and owner slots) and
// their hit queues.
static size_t fused_lds(int hist, size_t hist_lds, int sched, int block, bool persist) {
  size_t b = hist ? hist_lds : 0;
  if (sched >= 2) b = ((b + 7) & ~(size_t)7) + (size_t)(block / 64) * (kPoolWords * 8 + hit_queue(persist) * 4);
  return b;
}

template <int H, bool P, bool S, int R>
static void launch_one(dim3 grid, int block, size_t lds, hipStream_t st, const fused_args& A) {
  const size_t bytes = fused_lds(H, lds, R, block, P);
  set_lds_attr<H, P, S, R>(bytes);
  k_ff_fused<H, P, S, R><<<grid, block, bytes, st>>>(A.descs, A.tiles, A.n_tiles, A.queue, A.seeds, A.tallies,
                                                     A.Fs, A.stats, A.refill, pool_arg(A.pool, A.walk));
}

template <bool P, bool S, int R>
static hipError_t launch_mode(int hist, dim3 grid, int block, size_t lds, hipStream_t st, const fused_args& A) {
  switch (hist) {
    case 1: launch_one<1, P, S, R>(grid, block, lds, st, A); break;
    case 2: launch_one<2, P, S, R>(grid, block, lds, st, A); break;
    default: launch_one<0, P, S, R>(grid, block, lds, st, A); break;
  }
  return hipGetLastError();
}

// Scheduling: 0 = one ray per lane, 1 = lane refill, 2 = lane refill with
// pooled tests (when the best-hit words fit beside the histogram), 3 = as 2
// with the jump code for every empty cell (walk bit 0). All cast the same rays
// and tally the same hits.
static int sched_of(int refill, int pool, int walk, int hist, size_t hist_lds, int block, bool persist) {
  if (refill <= 0) return 0;
  if (pool <= 0 || fused_lds(hist, hist_lds, 2, block, persist) > (size_t)max_lds_bytes()) return 1;
  return (walk & kWalkJump) ? 3 : 2;
}

template <bool P, bool S>
static hipError_t launch_refill(int hist, dim3 grid, int block, size_t lds, hipStream_t st, const fused_args& A) {
  const int sched = sched_of(A.refill, A.pool, A.walk, hist, lds, block, P);
  return sched == 3   ? launch_mode<P, S, 3>(hist, grid, block, lds, st, A)
         : sched == 2 ? launch_mode<P, S, 2>(hist, grid, block, lds, st, A)
         : sched == 1 ? launch_mode<P, S, 1>(hist, grid, block, lds, st, A)
                      : launch_mode<P, S, 0>(hist, grid, block, lds, st, A);
}

static hipError_t launch_fused(int hist, bool persist, bool stats, dim3 grid, int block, size_t lds,
                               hipStream_t st, const fused_args& A) {
  if (persist)
    return stats ? launch_refill<true, true>(hist, grid, block, lds, st, A)
                 : launch_refill<true, false>(hist, grid, block, lds, st, A);
  return stats ? launch_refill<false, true>(hist, grid, block, lds, st, A)
               : launch_refill<false, false>(hist, grid, block, lds, st, A);
}

static long long auto_rays_per_tile(long long total) {
  long long per = (total + 8LL * device_cu_count() - 1) / (8LL * device_cu_count());
  return std::max(1024LL, std::min(65535LL, per));
}

// Tiles of one scene: each mesh range split into pieces of at most
// rays_per_tile; a mesh covered by one tile writes its F row directly.
// The face emitting ray g (g < total rays): the first f with face_off[f + 1] > g.
static int host_face_of(const ffk_scene* s, long long g) {
  const auto& off = s->h_face_off;
  return (int)(std::upper_bound(off.begin(), off.end(), g) - off.begin()) - 1;
}

// Tiles ordered by size class, longest first: classes of half an octave in
// rays (a factor sqrt(2)), mesh order within a class. Neighbouring meshes run
// side by side and share the cells their rays cross in L2, while the launch
// still ends on short tiles (a strict sort by ray count scattered the
// near-equal tiles of a uniform mesh and quadrupled the table reads).
static int size_class(const tile& t) { return (int)std::floor(2.0 * std::log2((double)std::max(1LL, t.g1 - t.g0))); }
static bool longer_class(const tile& a, const tile& b) { return size_class(a) > size_class(b); }

static void scene_tiles(const ffk_scene* s, int scene_id, long long rpt, std::vector<tile>& out,
                        std::vector<int>& indirect) {
  auto add = [&](long long g0, long long g1, int m, int direct) {
    const int f_lo = host_face_of(s, g0);
    const long long j0 = g0 - s->h_face_off[f_lo];
    out.push_back(tile{g0, g1, m, scene_id, direct, f_lo, host_face_of(s, g1 - 1),
                       j0 + (g1 - g0) < (1LL << 31) ? (int)j0 : -1});
  };
  for (int m = 0; m < s->n_meshes; ++m) {
    long long a = s->h_mesh_off[m], b = s->h_mesh_off[m + 1];
    if (b <= a) {  // no rays: 0 * (1/0) in the reference, produced by the tally path
      indirect.push_back(m);
      continue;
    }
    if (b - a <= rpt) {
      add(a, b, m, 1);
      continue;
    }
    indirect.push_back(m);
    for (long long g = a; g < b; g += rpt) add(g, std::min(b, g + rpt), m, 0);
  }
}

int build_tiles(ffk_scene* s, int rays_per_tile, hipStream_t st) {
  dfree(s->d_tiles);
  dfree(s->d_indirect_rows);
  dfree(s->d_desc);
  s->d_tiles = nullptr;
  s->d_indirect_rows = nullptr;
  s->d_desc = nullptr;
  long long rpt = rays_per_tile > 0 ? rays_per_tile : auto_rays_per_tile(s->n_rays);
  rpt = std::min(rpt, 65535LL);
  std::vector<tile> tiles;
  std::vector<int> indirect;
  scene_tiles(s, 0, rpt, tiles, indirect);
  // Longest tiles first: blocks start in index order, so uneven tiles (meshes
  // of unequal area) in mesh order leave the last full-size tiles running
  // alone at the end of the launch; each tile writes its own row or adds
  // integers to the tally, so the order does not change F.
  std::stable_sort(tiles.begin(), tiles.end(), longer_class);
  s->n_tiles = (int)tiles.size();
  s->n_indirect_rows = (int)indirect.size();
  s->tiles_rays_per_tile = (int)rpt;
  int r;
  if ((r = dalloc(&s->d_tiles, std::max<size_t>(1, tiles.size()))) < 0) return r;
  if ((r = dalloc(&s->d_indirect_rows, std::max<size_t>(1, indirect.size()))) < 0) return r;
  if ((r = dalloc(&s->d_desc, 1)) < 0) return r;
  if (!tiles.empty())
    FFK_HIP(hipMemcpy(s->d_tiles, tiles.data(), tiles.size() * sizeof(tile), hipMemcpyHostToDevice));
  if (!indirect.empty())
    FFK_HIP(hipMemcpy(s->d_indirect_rows, indirect.data(), indirect.size() * 4, hipMemcpyHostToDevice));
  (void)st;
  s->tiles_dirty = false;
  return FFK_OK;
}

// Rebuilds the tile list after a layout change; the old list may still be
// read by earlier launches on the stream.
int ensure_tiles(ffk_scene* s, hipStream_t st) {
  if (!s->has_emission || !s->tiles_dirty) return FFK_OK;
  FFK_HIP(hipStreamSynchronize(st));
  return build_tiles(s, s->tune_rays_per_tile, st);
}

}  // namespace ffk

using namespace ffk;

// ------------------------------------------------------------ batches --

struct ffk_batch {
  std::vector<ffk_scene*> scenes;
  std::vector<unsigned long long> generations;  // each scene's generation when the batch was created
  tile* d_tiles = nullptr;
  int n_tiles = 0;
  scene_desc* d_descs = nullptr;
  int2* d_rows = nullptr;  // (scene, mesh) rows tallied through the global tally
  int n_rows = 0;
  // Scenes grouped by the LDS their row histogram needs: one persistent
  // launch per class, so small scenes are not run at the occupancy of the
  // largest one. Tiles of class c are d_tiles[first, first + n).
  struct klass {
    int id;  // 0..4: 16-bit histogram of at most 8, 24, 48, 96, 160 KB; 5: global tally
    int hist;
    size_t lds;
    int first, n;
    long long rays;
  };
  std::vector<klass> classes;
  std::vector<long long> tile_rays;  // rays of each queued tile, in queue order
  // Per-class launch settings (-1 = the batch-wide value), kept across
  // re-plans; set by ffk_batch_set_class_tuning / ffk_batch_autotune.
  struct class_tuning {
    int block = -1, refill = -1, pool = -1, walk = -1;
  };
  class_tuning ctune[6];
  std::vector<char> scene_global;  // scene tallied entirely through the global tally (hist 0)
  std::vector<char> scene_tally;   // scene needs a tally buffer (global, or rows over several tiles)
  int* d_queue = nullptr;          // one counter per class
  unsigned long long* d_seeds = nullptr;
  unsigned long long** d_tallies = nullptr;
  float** d_Fs = nullptr;
  long long total_rays = 0;
  int block = 0, rpt = 0, n_blocks = 0;
  size_t lds_limit = kMaxDefaultHist;  // scenes whose 16-bit row histogram is larger use the global tally
  bool concurrent = false;             // class launches on forked streams (else back to back)
  int order = 1;  // tile queue: 0 longest first, 1 by scene (most rays first), longest first within
  int refill = kDefaultRefill;
  int pool = kDefaultPool;
  int walk = kDefaultWalk;
  // One stream per LDS class (default priority): the class launches are
  // forked from the launch stream after the tallies are cleared and joined
  // before the normalisation, so that a class's blocks start while the
  // previous class drains (in a captured graph: parallel branches).
  std::vector<hipStream_t> fork;
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  // graph state
  hipStream_t cap = nullptr;  // capture stream (default priority)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned long long*> g_tallies;
  std::vector<float*> g_Fs;
  std::vector<unsigned long long> g_seeds;
  // Seeds reach d_seeds through a pinned ring: a slot is rewritten only after
  // the copy that last read it has completed (its event), so a run that only
  // changes the seeds neither re-captures the graph nor waits for the device.
  static constexpr int kSeedSlots = 4;
  unsigned long long* h_seed_ring = nullptr;  // kSeedSlots * scenes.size(), pinned
  hipEvent_t seed_ev[kSeedSlots] = {};
  int seed_slot = 0;
};

namespace ffk {

__global__ void k_zero_rows_batch(const int2* __restrict__ rows, int n_rows, const scene_desc* __restrict__ d,
                                  unsigned long long* const* __restrict__ tallies) {
  const int r = blockIdx.x;
  if (r >= n_rows) return;
  const int2 sm = rows[r];
  const int N = d[sm.x].n_meshes;
  unsigned long long* t = tallies[sm.x] + (size_t)sm.y * N;
  for (int j = threadIdx.x; j < N; j += blockDim.x) t[j] = 0ull;
}

__global__ void k_normalise_batch(const int2* __restrict__ rows, int n_rows, const scene_desc* __restrict__ d,
                                  unsigned long long* const* __restrict__ tallies, float* const* __restrict__ Fs) {
  const int r = blockIdx.x;
  if (r >= n_rows) return;
  const int2 sm = rows[r];
  const scene_desc& S = d[sm.x];
  const int N = S.n_meshes, m = sm.y;
  const float ratio = 1.f / (float)(int)(S.mesh_off[m + 1] - S.mesh_off[m]);
  const unsigned long long* t = tallies[sm.x] + (size_t)m * N;
  float* F = Fs[sm.x] + (size_t)m * N;
  for (int j = threadIdx.x; j < N; j += blockDim.x) F[j] = normalise_one(t[j], ratio);
}

static void free_graph(ffk_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Histogram class of a scene with N meshes (tiles of at most 65535 rays):
// 0..4 = packed 16-bit LDS histogram of at most 8, 24, 48, 96 or 160 KB,
// 5 = global uint64 tally.
static int hist_class(int N, size_t lds_limit) {
  static const size_t limits[5] = {8192, 24576, 49152, 98304, 163840};
  const size_t need = hist_bytes(2, N);
  if (need > lds_limit) return 5;
  for (int c = 0; c < 5; ++c)
    if (need <= limits[c] && (int)need <= max_lds_bytes()) return c;
  return 5;
}

static int batch_plan(ffk_batch* b) {
  free_graph(b);
  dfree(b->d_tiles);
  dfree(b->d_rows);
  dfree(b->d_queue);
  b->d_tiles = nullptr;
  b->d_rows = nullptr;
  b->d_queue = nullptr;
  long long total = 0;
  for (ffk_scene* s : b->scenes) total += s->n_rays;
  long long rpt = b->rpt > 0 ? b->rpt : auto_rays_per_tile(total);
  rpt = std::min(rpt, 65535LL);
  std::vector<std::vector<tile>> per_class(6);
  std::vector<size_t> class_lds(6, 0);
  std::vector<int2> rows;
  b->scene_global.assign(b->scenes.size(), 0);
  b->scene_tally.assign(b->scenes.size(), 0);
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    const int c = hist_class(b->scenes[i]->n_meshes, b->lds_limit);
    std::vector<int> ind;
    scene_tiles(b->scenes[i], (int)i, rpt, per_class[c], ind);
    b->scene_tally[i] = c == 5 || !ind.empty();
    if (c == 5) {
      b->scene_global[i] = 1;  // whole tally zeroed and normalised
    } else {
      for (int m : ind) rows.push_back(make_int2((int)i, m));
      class_lds[c] = std::max(class_lds[c], hist_bytes(2, b->scenes[i]->n_meshes));
    }
  }
  // Classes with the largest histograms first; inside a class the longest
  // tiles first, a greedy order that keeps the persistent blocks busy to the
  // end of the queue.
  std::vector<tile> tiles;
  b->classes.clear();
  std::vector<long long> scene_rays(b->scenes.size());
  for (size_t i = 0; i < b->scenes.size(); ++i) scene_rays[i] = b->scenes[i]->n_rays;
  for (int c = 5; c >= 0; --c) {
    if (per_class[c].empty()) continue;
    if (b->order == 1)  // one scene's tables stay hot in L2 while its tiles run
      std::stable_sort(per_class[c].begin(), per_class[c].end(), [&](const tile& x, const tile& y) {
        if (x.scene != y.scene)
          return scene_rays[x.scene] != scene_rays[y.scene] ? scene_rays[x.scene] > scene_rays[y.scene]
                                                              : x.scene < y.scene;
        return longer_class(x, y);
      });
    else
      std::stable_sort(per_class[c].begin(), per_class[c].end(), longer_class);
    ffk_batch::klass k;
    k.id = c;
    k.hist = c == 5 ? 0 : 2;
    k.lds = c == 5 ? 0 : class_lds[c];
    k.first = (int)tiles.size();
    k.n = (int)per_class[c].size();
    k.rays = 0;
    for (const tile& t : per_class[c]) k.rays += t.g1 - t.g0;
    tiles.insert(tiles.end(), per_class[c].begin(), per_class[c].end());
    b->classes.push_back(k);
  }
  b->tile_rays.resize(tiles.size());
  for (size_t i = 0; i < tiles.size(); ++i) b->tile_rays[i] = tiles[i].g1 - tiles[i].g0;
  b->n_tiles = (int)tiles.size();
  b->n_rows = (int)rows.size();
  b->total_rays = total;
  int r;
  if ((r = dalloc(&b->d_tiles, std::max<size_t>(1, tiles.size()))) < 0) return r;
  if ((r = dalloc(&b->d_rows, std::max<size_t>(1, rows.size()))) < 0) return r;
  if ((r = dalloc(&b->d_queue, std::max<size_t>(1, b->classes.size()))) < 0) return r;
  if (!tiles.empty())
    FFK_HIP(hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(tile), hipMemcpyHostToDevice));
  if (!rows.empty())
    FFK_HIP(hipMemcpy(b->d_rows, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice));
  return FFK_OK;
}

// Threads per block: a histogram above 80 KB leaves room for one block per
// CU, which then runs 16 wavefronts.
static int default_block(int hist, size_t lds) { return hist == 0 ? 256 : (lds > 81920 ? 1024 : 512); }

}  // namespace ffk

extern "C" {

int ffk_emit_malley(ffk_scene* s, unsigned long long seed, float* d_rays, void* stream) {
  FFK_REQUIRE(s && s->has_emission, FFK_ERR_ARG, "ffk_emit_malley: emission layout not set");
  if (s->n_rays == 0) return FFK_OK;
  FFK_REQUIRE(d_rays, FFK_ERR_ARG, "ffk_emit_malley: NULL ray buffer");
  hipStream_t st = as_stream(stream);
  k_emit_malley<<<(unsigned)((s->n_rays + 255) / 256), 256, 0, st>>>(s->desc(), (uint32_t)seed,
                                                                     (uint32_t)(seed >> 32), d_rays);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_emit_parallel(const float* h_dir3, const float* h_center3, float distance, float radius,
                      long long n_rays, unsigned long long seed, float* d_rays, void* stream) {
  FFK_REQUIRE(h_dir3 && h_center3 && n_rays >= 0, FFK_ERR_ARG, "ffk_emit_parallel: bad args");
  if (n_rays == 0) return FFK_OK;
  FFK_REQUIRE(d_rays, FFK_ERR_ARG, "ffk_emit_parallel: NULL ray buffer");
  hipStream_t st = as_stream(stream);
  v3 dir = mk(h_dir3[0], h_dir3[1], h_dir3[2]);
  v3 center = mk(h_center3[0], h_center3[1], h_center3[2]);
  float rot[9];
  rotate_towards(mk(0.f, 0.f, 1.f), dir, rot);
  float* d_rot = nullptr;
  int r;
  if ((r = dalloc(&d_rot, 9)) < 0) return r;
  FFK_HIP(hipMemcpyAsync(d_rot, rot, sizeof(rot), hipMemcpyHostToDevice, st));
  k_emit_parallel<<<(unsigned)((n_rays + 255) / 256), 256, 0, st>>>(
      dir, center, distance, radius, d_rot, n_rays, (uint32_t)seed, (uint32_t)(seed >> 32), d_rays);
  hipError_t e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(st);
  dfree(d_rot);
  FFK_HIP(e);
  FFK_HIP(e2);
  return FFK_OK;
}

int ffk_tally_zero(const ffk_scene* s, unsigned long long* d_tally, void* stream) {
  FFK_REQUIRE(s && s->has_meshes && d_tally, FFK_ERR_ARG, "ffk_tally_zero: bad args");
  long long n = (long long)s->n_meshes * s->n_meshes;
  int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  k_fill_u64<<<blocks, 256, 0, as_stream(stream)>>>(d_tally, n);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_tally(const ffk_scene* s, const int* d_hit_idx, long long n_rays, unsigned long long* d_tally,
              void* stream) {
  FFK_REQUIRE(s && s->has_emission, FFK_ERR_ARG, "ffk_tally: emission layout not set");
  FFK_REQUIRE(n_rays >= 0 && n_rays <= s->n_rays, FFK_ERR_ARG, "ffk_tally: %lld rays, layout has %lld",
              n_rays, s->n_rays);
  if (n_rays == 0) return FFK_OK;
  FFK_REQUIRE(d_hit_idx && d_tally, FFK_ERR_ARG, "ffk_tally: NULL buffer");
  k_tally<<<(unsigned)((n_rays + 255) / 256), 256, 0, as_stream(stream)>>>(s->desc(), d_hit_idx, n_rays,
                                                                          d_tally);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_normalise(const ffk_scene* s, const unsigned long long* d_tally, float* d_F, void* stream) {
  FFK_REQUIRE(s && s->has_emission && d_tally && d_F, FFK_ERR_ARG, "ffk_normalise: bad args");
  const int N = s->n_meshes;
  dim3 grid((unsigned)N);
  k_normalise_rows<<<grid, 256, 0, as_stream(stream)>>>(s->d_mesh_off, N, nullptr, N, d_tally, d_F);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_set_tuning(ffk_scene* s, int block_size, int rays_per_tile, int hist_mode) {
  FFK_REQUIRE(s, FFK_ERR_ARG, "scene is NULL");
  FFK_REQUIRE(block_size == 0 || (block_size % 64 == 0 && block_size >= 64 && block_size <= 1024),
              FFK_ERR_ARG, "block size %d is not a multiple of 64 in [64, 1024]", block_size);
  FFK_REQUIRE(rays_per_tile >= 0 && rays_per_tile <= 65535, FFK_ERR_LIMIT,
              "rays per tile %d outside [0, 65535]", rays_per_tile);
  FFK_REQUIRE(hist_mode >= -1 && hist_mode <= 2, FFK_ERR_ARG, "hist mode %d", hist_mode);
  s->tune_block = block_size;
  s->tune_hist = hist_mode;
  if (rays_per_tile != s->tune_rays_per_tile) {
    s->tune_rays_per_tile = rays_per_tile;
    s->tiles_dirty = true;
  }
  return FFK_OK;
}

int ffk_set_refill(ffk_scene* s, int lanes, int pool) {
  FFK_REQUIRE(s, FFK_ERR_ARG, "scene is NULL");
  FFK_REQUIRE(lanes >= 0 && lanes <= 64, FFK_ERR_ARG, "refill lanes %d outside [0, 64]", lanes);
  FFK_REQUIRE(pool >= 0 && pool <= 64, FFK_ERR_ARG, "pool ratio %d outside [0, 64]", pool);
  s->tune_refill = lanes;
  s->tune_pool = pool;
  return FFK_OK;
}

int ffk_get_refill(const ffk_scene* s, int* lanes, int* pool) {
  FFK_REQUIRE(s, FFK_ERR_ARG, "scene is NULL");
  if (lanes) *lanes = s->tune_refill;
  if (pool) *pool = s->tune_pool;
  return FFK_OK;
}

int ffk_set_walk(ffk_scene* s, int mode) {
  FFK_REQUIRE(s, FFK_ERR_ARG, "scene is NULL");
  FFK_REQUIRE(mode >= 0 && mode <= 7, FFK_ERR_ARG, "walk mode %d is not in 0..7", mode);
  s->tune_walk = mode;
  return FFK_OK;
}

int ffk_batch_set_walk(ffk_batch* b, int mode) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  FFK_REQUIRE(mode >= 0 && mode <= 7, FFK_ERR_ARG, "walk mode %d is not in 0..7", mode);
  b->walk = mode;
  free_graph(b);
  return FFK_OK;
}

int ffk_batch_set_refill(ffk_batch* b, int lanes, int pool) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  FFK_REQUIRE(lanes >= 0 && lanes <= 64, FFK_ERR_ARG, "refill lanes %d outside [0, 64]", lanes);
  FFK_REQUIRE(pool >= 0 && pool <= 64, FFK_ERR_ARG, "pool ratio %d outside [0, 64]", pool);
  b->refill = lanes;
  b->pool = pool;
  free_graph(b);
  return FFK_OK;
}

int ffk_ff_needs_tally(const ffk_scene* s) {
  if (!s || !s->has_emission) return 0;
  long long rpt = s->tiles_rays_per_tile;
  size_t n_indirect = (size_t)s->n_indirect_rows;
  if (s->tiles_dirty) {  // the tile list the next fused launch builds
    rpt = std::min(s->tune_rays_per_tile > 0 ? (long long)s->tune_rays_per_tile : auto_rays_per_tile(s->n_rays),
                   65535LL);
    std::vector<tile> tiles;
    std::vector<int> indirect;
    scene_tiles(s, 0, rpt, tiles, indirect);
    n_indirect = indirect.size();
  }
  const long long max_tile = std::min<long long>(rpt, 65535);
  const int hist = pick_hist(s->n_meshes, max_tile, s->tune_hist);
  return (hist == 0 || n_indirect > 0) ? 1 : 0;
}

int ffk_batch_tally_flags(const ffk_batch* b, int* h_flags) {
  FFK_REQUIRE(b && h_flags, FFK_ERR_ARG, "ffk_batch_tally_flags: bad args");
  for (size_t i = 0; i < b->scenes.size(); ++i) h_flags[i] = b->scene_tally[i] ? 1 : 0;
  return FFK_OK;
}

int ffk_ff_fused(ffk_scene* s, int method, unsigned long long seed, unsigned long long* d_tally, float* d_F,
                 unsigned long long* d_stats, void* stream) {
  FFK_REQUIRE(s && s->has_emission && s->has_grid, FFK_ERR_ARG,
              "ffk_ff_fused: emission layout and grid must be set");
  FFK_REQUIRE(method == FFK_CAST_GRID, FFK_ERR_ARG, "ffk_ff_fused: only the grid caster is fused");
  FFK_REQUIRE(d_F, FFK_ERR_ARG, "ffk_ff_fused: NULL F");
  hipStream_t st = as_stream(stream);
  int rt;
  if ((rt = ensure_tiles(s, st)) < 0) return rt;
  const int N = s->n_meshes;
  long long max_tile = std::min<long long>(s->tiles_rays_per_tile, 65535);
  int hist = pick_hist(N, max_tile, s->tune_hist);
  const bool needs_tally = hist == 0 || s->n_indirect_rows > 0;
  FFK_REQUIRE(!needs_tally || d_tally, FFK_ERR_ARG,
              "ffk_ff_fused: this scene needs the N*N uint64 tally buffer");
  // per-launch argument block in device memory
  scene_desc d = s->desc();
  FFK_HIP(hipMemcpyAsync(s->d_desc, &d, sizeof(d), hipMemcpyHostToDevice, st));
  if (!s->d_misc) {  // seed, tally pointer, F pointer
    int r;
    if ((r = dalloc(&s->d_misc, 4)) < 0) return r;
  }
  unsigned long long* d_misc = s->d_misc;
  unsigned long long misc[3] = {seed, (unsigned long long)(uintptr_t)d_tally,
                                (unsigned long long)(uintptr_t)d_F};
  FFK_HIP(hipMemcpyAsync(d_misc, misc, sizeof(misc), hipMemcpyHostToDevice, st));
  if (hist == 0) {
    long long n = (long long)N * N;
    k_fill_u64<<<(int)std::min<long long>((n + 255) / 256, 65536), 256, 0, st>>>(d_tally, n);
  } else if (s->n_indirect_rows > 0) {
    dim3 grid((unsigned)s->n_indirect_rows);
    k_zero_rows<<<grid, 256, 0, st>>>(N, s->d_indirect_rows, s->n_indirect_rows, d_tally);
  }
  FFK_HIP(hipGetLastError());
  fused_args A;
  A.descs = s->d_desc;
  A.tiles = s->d_tiles;
  A.n_tiles = s->n_tiles;
  A.queue = nullptr;
  A.seeds = d_misc;
  A.tallies = reinterpret_cast<unsigned long long* const*>(d_misc + 1);
  A.Fs = reinterpret_cast<float* const*>(d_misc + 2);
  A.stats = d_stats;
  A.refill = s->tune_refill;
  A.pool = s->tune_pool;
  A.walk = s->tune_walk;
  const int block = s->tune_block > 0 ? s->tune_block : default_block(hist, hist_bytes(hist, N));
  if (s->n_tiles > 0)
    FFK_HIP(launch_fused(hist, false, d_stats != nullptr, dim3((unsigned)s->n_tiles), block,
                         hist_bytes(hist, N), st, A));
  if (hist == 0) {
    dim3 grid((unsigned)N);
    k_normalise_rows<<<grid, 256, 0, st>>>(s->d_mesh_off, N, nullptr, N, d_tally, d_F);
  } else if (s->n_indirect_rows > 0) {
    dim3 grid((unsigned)s->n_indirect_rows);
    k_normalise_rows<<<grid, 256, 0, st>>>(s->d_mesh_off, N, s->d_indirect_rows, s->n_indirect_rows,
                                           d_tally, d_F);
  }
  FFK_HIP(hipGetLastError());
  // The argument block is reused by the next call on this thread.
  FFK_HIP(hipStreamSynchronize(st));
  return FFK_OK;
}

int ffk_batch_create(ffk_batch** out, ffk_scene* const* scenes, int n_scenes, void* stream) {
  FFK_REQUIRE(out && scenes && n_scenes > 0, FFK_ERR_ARG, "ffk_batch_create: bad args");
  for (int i = 0; i < n_scenes; ++i)
    FFK_REQUIRE(scenes[i] && scenes[i]->has_emission && scenes[i]->has_grid, FFK_ERR_ARG,
                "ffk_batch_create: scene %d lacks emission layout or grid", i);
  ffk_batch* b = new ffk_batch();
  b->scenes.assign(scenes, scenes + n_scenes);
  for (int i = 0; i < n_scenes; ++i) b->generations.push_back(scenes[i]->generation);
  int r;
  if ((r = dalloc(&b->d_descs, (size_t)n_scenes)) < 0 ||
      (r = dalloc(&b->d_seeds, (size_t)n_scenes)) < 0 || (r = dalloc(&b->d_tallies, (size_t)n_scenes)) < 0 ||
      (r = dalloc(&b->d_Fs, (size_t)n_scenes)) < 0) {
    ffk_batch_destroy(b);
    return r;
  }
  std::vector<scene_desc> descs(n_scenes);
  for (int i = 0; i < n_scenes; ++i) descs[i] = scenes[i]->desc();
  hipError_t e = hipMemcpy(b->d_descs, descs.data(), descs.size() * sizeof(scene_desc), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("hipMemcpy failed: %s", hipGetErrorString(e));
    ffk_batch_destroy(b);
    return FFK_ERR_HIP;
  }
  if ((r = batch_plan(b)) < 0) {
    ffk_batch_destroy(b);
    return r;
  }
  (void)stream;
  *out = b;
  return FFK_OK;
}

void ffk_batch_destroy(ffk_batch* b) {
  if (!b) return;
  (void)hipDeviceSynchronize();
  free_graph(b);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  for (hipStream_t f : b->fork) (void)hipStreamDestroy(f);
  for (hipEvent_t e : b->ev_join) (void)hipEventDestroy(e);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  for (hipEvent_t e : b->seed_ev)
    if (e) (void)hipEventDestroy(e);
  if (b->h_seed_ring) (void)hipHostFree(b->h_seed_ring);
  dfree(b->d_tiles);
  dfree(b->d_descs);
  dfree(b->d_rows);
  dfree(b->d_queue);
  dfree(b->d_seeds);
  dfree(b->d_tallies);
  dfree(b->d_Fs);
  delete b;
}

// The batch holds each scene's device pointers and ray layout as they were at
// ffk_batch_create; a scene whose grid, faces, meshes or emission layout has
// been set again since would be read through freed or stale buffers.
static int batch_scenes_current(const ffk_batch* b, const char* fn) {
  for (size_t i = 0; i < b->scenes.size(); ++i)
    FFK_REQUIRE(b->scenes[i]->generation == b->generations[i], FFK_ERR_ARG,
                "%s: scene %zu was rebuilt (grid, faces, meshes or emission set again) after "
                "ffk_batch_create; create the batch again",
                fn, i);
  return FFK_OK;
}

int ffk_batch_set_order(ffk_batch* b, int order) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  FFK_REQUIRE(order == 0 || order == 1, FFK_ERR_ARG, "tile order %d is not 0 or 1", order);
  int rc;
  if ((rc = batch_scenes_current(b, "ffk_batch_set_order")) < 0) return rc;
  if (order == b->order) return FFK_OK;
  b->order = order;
  return batch_plan(b);
}

int ffk_batch_set_concurrent(ffk_batch* b, int on) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  b->concurrent = on != 0;
  free_graph(b);
  return FFK_OK;
}

int ffk_batch_set_lds_limit(ffk_batch* b, int bytes) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  FFK_REQUIRE(bytes >= 0, FFK_ERR_ARG, "LDS limit %d", bytes);
  int rc;
  if ((rc = batch_scenes_current(b, "ffk_batch_set_lds_limit")) < 0) return rc;
  const size_t lim = bytes == 0 ? (size_t)kMaxDefaultHist : (size_t)bytes;
  if (lim == b->lds_limit) return FFK_OK;
  b->lds_limit = lim;
  return batch_plan(b);
}

int ffk_batch_set_tuning(ffk_batch* b, int block_size, int rays_per_tile, int n_blocks) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  int rc;
  if ((rc = batch_scenes_current(b, "ffk_batch_set_tuning")) < 0) return rc;
  FFK_REQUIRE(block_size == 0 || (block_size % 64 == 0 && block_size >= 64 && block_size <= 1024),
              FFK_ERR_ARG, "block size %d", block_size);
  FFK_REQUIRE(rays_per_tile >= 0 && rays_per_tile <= 65535, FFK_ERR_LIMIT, "rays per tile %d", rays_per_tile);
  FFK_REQUIRE(n_blocks >= 0, FFK_ERR_ARG, "n_blocks %d", n_blocks);
  b->block = block_size;
  b->n_blocks = n_blocks;
  if (rays_per_tile != b->rpt) {
    b->rpt = rays_per_tile;
    return batch_plan(b);
  }
  free_graph(b);
  return FFK_OK;
}

int ffk_batch_n_scenes(const ffk_batch* b) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  return (int)b->scenes.size();
}

int ffk_batch_get_tuning(const ffk_batch* b, int* block_size, int* rays_per_tile, int* n_blocks) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  if (block_size) *block_size = b->block;
  if (rays_per_tile) *rays_per_tile = b->rpt;
  if (n_blocks) *n_blocks = b->n_blocks;
  return FFK_OK;
}

}  // extern "C"

template <int R>
static hipError_t persistent_occupancy(int hist, int block, size_t lds, int* per_cu) {
  const size_t bytes = fused_lds(hist, lds, R, block, true);
  switch (hist) {
    case 1:
      set_lds_attr<1, true, false, R>(bytes);
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_ff_fused<1, true, false, R>, block, bytes);
    case 2:
      set_lds_attr<2, true, false, R>(bytes);
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_ff_fused<2, true, false, R>, block, bytes);
    default:
      set_lds_attr<0, true, false, R>(bytes);
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_ff_fused<0, true, false, R>, block, bytes);
  }
}

static int batch_blocks(const ffk_batch* b, int hist, size_t lds, int block, int n_tiles, int refill, int pool,
                        int walk);

extern "C" {

static int batch_blocks(const ffk_batch* b, int hist, size_t lds, int block, int n_tiles, int refill, int pool,
                        int walk) {
  int n_blocks = b->n_blocks;
  if (n_blocks <= 0) {
    int per_cu = 0;
    const int sched = sched_of(refill, pool, walk, hist, lds, block, true);
    const hipError_t e = sched == 3   ? persistent_occupancy<3>(hist, block, lds, &per_cu)
                         : sched == 2 ? persistent_occupancy<2>(hist, block, lds, &per_cu)
                         : sched == 1 ? persistent_occupancy<1>(hist, block, lds, &per_cu)
                                      : persistent_occupancy<0>(hist, block, lds, &per_cu);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    n_blocks = per_cu * device_cu_count();
  }
  return std::max(1, std::min(n_blocks, n_tiles));
}

// The persistent launch of class c over the first n_tiles of its queue
// (its queue counter must be zero).
static int launch_class(ffk_batch* b, int c, int n_tiles, bool stats, unsigned long long* d_stats, hipStream_t st) {
  const ffk_batch::klass& k = b->classes[c];
  const ffk_batch::class_tuning& ct = b->ctune[k.id];
  const int block = ct.block > 0 ? ct.block : b->block > 0 ? b->block : default_block(k.hist, k.lds);
  fused_args A;
  A.descs = b->d_descs;
  A.tiles = b->d_tiles + k.first;
  A.n_tiles = n_tiles;
  A.queue = b->d_queue + c;
  A.seeds = b->d_seeds;
  A.tallies = b->d_tallies;
  A.Fs = b->d_Fs;
  A.stats = d_stats;
  A.refill = ct.refill >= 0 ? ct.refill : b->refill;
  A.pool = ct.pool >= 0 ? ct.pool : b->pool;
  A.walk = ct.walk >= 0 ? ct.walk : b->walk;
  const int n_blocks = batch_blocks(b, k.hist, k.lds, block, n_tiles, A.refill, A.pool, A.walk);
  FFK_HIP(launch_fused(k.hist, true, stats, dim3((unsigned)n_blocks), block, k.lds, st, A));
  return FFK_OK;
}

static int batch_enqueue(ffk_batch* b, bool stats, unsigned long long* d_stats, hipStream_t st) {
  const int n_classes = (int)b->classes.size();
  k_fill_i32<<<1, 64, 0, st>>>(b->d_queue, std::max(1, n_classes), 0);
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    if (!b->scene_global[i]) continue;
    long long n = (long long)b->scenes[i]->n_meshes * b->scenes[i]->n_meshes;
    k_fill_u64<<<(int)std::min<long long>((n + 255) / 256, 65536), 256, 0, st>>>(b->g_tallies[i], n);
  }
  if (b->n_rows > 0) {
    dim3 grid((unsigned)b->n_rows);
    k_zero_rows_batch<<<grid, 256, 0, st>>>(b->d_rows, b->n_rows, b->d_descs, b->d_tallies);
  }
  const bool fork = n_classes > 1 && b->concurrent;
  if (fork) {
    while ((int)b->fork.size() < n_classes) {
      hipStream_t f = nullptr;
      hipEvent_t e = nullptr;
      FFK_HIP(hipStreamCreateWithFlags(&f, hipStreamNonBlocking));
      b->fork.push_back(f);
      FFK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->ev_join.push_back(e);
    }
    if (!b->ev_fork) FFK_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    FFK_HIP(hipEventRecord(b->ev_fork, st));
  }
  for (int c = 0; c < n_classes; ++c) {
    hipStream_t cs = st;
    if (fork) {
      cs = b->fork[c];
      FFK_HIP(hipStreamWaitEvent(cs, b->ev_fork, 0));
    }
    int rc;
    if ((rc = launch_class(b, c, b->classes[c].n, stats, d_stats, cs)) < 0) return rc;
    if (fork) {
      FFK_HIP(hipEventRecord(b->ev_join[c], cs));
      FFK_HIP(hipStreamWaitEvent(st, b->ev_join[c], 0));
    }
  }
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    if (!b->scene_global[i]) continue;
    const ffk_scene* s = b->scenes[i];
    const int N = s->n_meshes;
    dim3 grid((unsigned)N);
    k_normalise_rows<<<grid, 256, 0, st>>>(s->d_mesh_off, N, nullptr, N, b->g_tallies[i], b->g_Fs[i]);
  }
  if (b->n_rows > 0) {
    dim3 grid((unsigned)b->n_rows);
    k_normalise_batch<<<grid, 256, 0, st>>>(b->d_rows, b->n_rows, b->d_descs, b->d_tallies, b->d_Fs);
  }
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_batch_run(ffk_batch* b, const unsigned long long* seeds, unsigned long long* const* d_tallies,
                  float* const* d_Fs, int use_graph, unsigned long long* d_stats, void* stream) {
  FFK_REQUIRE(b && seeds && d_Fs, FFK_ERR_ARG, "ffk_batch_run: bad args");
  int rc;
  if ((rc = batch_scenes_current(b, "ffk_batch_run")) < 0) return rc;
  const size_t n = b->scenes.size();
  for (size_t i = 0; i < n; ++i) {
    FFK_REQUIRE(d_Fs[i], FFK_ERR_ARG, "ffk_batch_run: NULL F for scene %zu", i);
    FFK_REQUIRE(!b->scene_tally[i] || (d_tallies && d_tallies[i]), FFK_ERR_ARG,
                "ffk_batch_run: scene %zu needs a tally buffer", i);
  }
  hipStream_t st = as_stream(stream);
  std::vector<unsigned long long*> tal(n, nullptr);
  std::vector<float*> fs(d_Fs, d_Fs + n);
  if (d_tallies) tal.assign(d_tallies, d_tallies + n);
  std::vector<unsigned long long> sd(seeds, seeds + n);
  // The kernels read seeds, tally and F pointers from device arrays, so a
  // captured graph stays valid when only the seeds change: they are copied
  // in stream order ahead of the launch (after the previous run's kernels
  // have read the old ones). New tally or F pointers re-plan the graph.
  const bool same_bufs = b->exec && tal == b->g_tallies && fs == b->g_Fs;
  if (!same_bufs) {
    free_graph(b);
    FFK_HIP(hipMemcpyAsync(b->d_tallies, tal.data(), n * 8, hipMemcpyHostToDevice, st));
    FFK_HIP(hipMemcpyAsync(b->d_Fs, fs.data(), n * 8, hipMemcpyHostToDevice, st));
    b->g_tallies = tal;
    b->g_Fs = fs;
  }
  if (!same_bufs || sd != b->g_seeds) {
    if (!b->h_seed_ring)
      FFK_HIP(hipHostMalloc((void**)&b->h_seed_ring, ffk_batch::kSeedSlots * n * 8, hipHostMallocDefault));
    const int slot = b->seed_slot;
    if (b->seed_ev[slot]) FFK_HIP(hipEventSynchronize(b->seed_ev[slot]));
    else FFK_HIP(hipEventCreateWithFlags(&b->seed_ev[slot], hipEventDisableTiming));
    unsigned long long* src = b->h_seed_ring + (size_t)slot * n;
    std::copy(sd.begin(), sd.end(), src);
    FFK_HIP(hipMemcpyAsync(b->d_seeds, src, n * 8, hipMemcpyHostToDevice, st));
    FFK_HIP(hipEventRecord(b->seed_ev[slot], st));
    b->seed_slot = (slot + 1) % ffk_batch::kSeedSlots;
    b->g_seeds = sd;
  }
  if (!same_bufs) FFK_HIP(hipStreamSynchronize(st));
  if (d_stats || !use_graph) return batch_enqueue(b, d_stats != nullptr, d_stats, st);
  if (!b->exec) {
    // Record on the batch's own stream (the caller's may be the legacy null
    // stream, which cannot be captured); the graph is launched on the
    // caller's stream below.
    if (!b->cap) FFK_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    FFK_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    int r = batch_enqueue(b, false, nullptr, b->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (r < 0) {
      if (g) (void)hipGraphDestroy(g);
      return r;
    }
    FFK_HIP(e);
    b->graph = g;
    FFK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  FFK_HIP(hipGraphLaunch(b->exec, st));
  return FFK_OK;
}

int ffk_batch_signature(const ffk_batch* b, unsigned long long* key) {
  FFK_REQUIRE(b && key, FFK_ERR_ARG, "ffk_batch_signature: bad args");
  unsigned long long h = 1469598103934665603ull;
  auto mix = [&h](const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  };
  for (const ffk_scene* s : b->scenes) {
    mix(s->h_faces.data(), s->h_faces.size() * 4);
    mix(s->h_meshes.data(), s->h_meshes.size() * 4);
    mix(s->h_front.data(), s->h_front.size() * 4);
    mix(s->h_rear.data(), s->h_rear.size() * 4);
    mix(&s->frame, sizeof(s->frame));
    const int dims[3] = {s->grid.n[0], s->grid.n[1], s->grid.n[2]};
    mix(dims, sizeof(dims));
  }
  const long long settings[6] = {(long long)b->lds_limit, b->order, b->concurrent ? 1 : 0, b->rpt, b->block,
                                 b->n_blocks};
  mix(settings, sizeof(settings));
  *key = h;
  return FFK_OK;
}

int ffk_batch_classes(const ffk_batch* b, int* ids, long long* rays, int* n_tiles, int capacity) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  const int n = (int)b->classes.size();
  for (int c = 0; c < n && c < capacity; ++c) {
    if (ids) ids[c] = b->classes[c].id;
    if (rays) rays[c] = b->classes[c].rays;
    if (n_tiles) n_tiles[c] = b->classes[c].n;
  }
  return n;
}

int ffk_batch_set_class_tuning(ffk_batch* b, int class_id, int block_size, int refill, int pool, int walk) {
  FFK_REQUIRE(b && class_id >= 0 && class_id < 6, FFK_ERR_ARG, "ffk_batch_set_class_tuning: class %d", class_id);
  FFK_REQUIRE(block_size <= 0 || (block_size % 64 == 0 && block_size <= 1024), FFK_ERR_ARG, "block size %d",
              block_size);
  FFK_REQUIRE(refill <= 64 && pool <= 64 && walk <= 7, FFK_ERR_ARG, "refill %d / pool %d / walk %d", refill, pool,
              walk);
  ffk_batch::class_tuning& t = b->ctune[class_id];
  t.block = block_size > 0 ? block_size : -1;
  t.refill = refill;
  t.pool = pool;
  t.walk = walk;
  free_graph(b);
  return FFK_OK;
}

int ffk_batch_get_class_tuning(const ffk_batch* b, int class_id, int* block_size, int* refill, int* pool,
                               int* walk) {
  FFK_REQUIRE(b && class_id >= 0 && class_id < 6, FFK_ERR_ARG, "ffk_batch_get_class_tuning: class %d", class_id);
  const ffk_batch::class_tuning& t = b->ctune[class_id];
  if (block_size) *block_size = t.block;
  if (refill) *refill = t.refill;
  if (pool) *pool = t.pool;
  if (walk) *walk = t.walk;
  return FFK_OK;
}

int ffk_batch_run_class(ffk_batch* b, int class_id, long long max_rays, void* stream) {
  FFK_REQUIRE(b, FFK_ERR_ARG, "batch is NULL");
  int rc;
  if ((rc = batch_scenes_current(b, "ffk_batch_run_class")) < 0) return rc;
  FFK_REQUIRE(!b->g_Fs.empty(), FFK_ERR_ARG, "ffk_batch_run_class: run the batch once first (binds its buffers)");
  int c = -1;
  for (int i = 0; i < (int)b->classes.size(); ++i)
    if (b->classes[i].id == class_id) c = i;
  FFK_REQUIRE(c >= 0, FFK_ERR_ARG, "ffk_batch_run_class: the batch has no class %d", class_id);
  const ffk_batch::klass& k = b->classes[c];
  int n = 0;
  long long sum = 0;
  while (n < k.n && (n == 0 || max_rays <= 0 || sum < max_rays)) sum += b->tile_rays[k.first + n++];
  hipStream_t st = as_stream(stream);
  k_fill_i32<<<1, 64, 0, st>>>(b->d_queue + c, 1, 0);
  return launch_class(b, c, n, false, nullptr, st);
}

}  // extern "C"
