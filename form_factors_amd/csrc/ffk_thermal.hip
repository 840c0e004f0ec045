// Radiative exchange q = F^T E with F in BF16/FP16/FP32 storage and FP32
// accumulation, CSR conduction, and the FP64 Adams-Bashforth step.
//
// Exchange storage. A Monte Carlo form factor is F[m, j] = k / n_m with an
// integer count k and the rays n_m emitted by mesh m
// (cpuFactorsCalculator/cpu_system.cpp:193-195). Rounding F itself to BF16
// gives every cell of equal count the same relative error, so the errors add
// up coherently over a row. The packed storage therefore keeps the counts
// S = F[m, j] * mul_m (exact in BF16 up to 256 and in FP16 up to 2048) and
// the product applies scale_m = 1 / mul_m to E_m. S is stored transposed,
// St[j, m], so that q_j = sum_m St[j, m] * (E_m * scale_m) is one contiguous
// row per target mesh: a wavefront streams the row from HBM while E sits in
// LDS, and no partial sums leave the chip.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <mutex>

#include "ffk_internal.hpp"

namespace ffk {

// Stefan-Boltzmann constant of thermal_equation/form_factors.cpp:33.
constexpr double kSigma = 5.670400e-8;

// float -> bfloat16 bits, round to nearest even (NaN kept quiet).
__device__ inline uint16_t f2bf(float x) {
  uint32_t b = __float_as_uint(x);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline void store_as(void* out, size_t i, int dtype, float v) {
  if (dtype == FFK_BF16) reinterpret_cast<uint16_t*>(out)[i] = f2bf(v);
  else if (dtype == FFK_F16) reinterpret_cast<__half*>(out)[i] = __float2half_rn(v);
  else reinterpret_cast<float*>(out)[i] = v;
}

__global__ void k_convert(const float* __restrict__ in, long long n, int dtype, void* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) store_as(out, (size_t)i, dtype, in[i]);
}

// Block-wide reductions over 256 threads (4 wavefronts).
__device__ inline float block_max(float v, float* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}
__device__ inline float block_min(float v, float* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
}

// One block per row m of F: choose mul_m and scale_m = 1 / mul_m.
//  - row_mul given: mul_m = row_mul[m] (e.g. the rays of mesh m).
//  - otherwise a row is taken as count-valued when c = rint(1 / min nonzero)
//    makes every entry an integer to within 4e-6 relative (the smallest
//    count is then 1 and c = n_m); then mul_m = c, else 1.
//  - FP16: when the scaled row maximum would reach 32768 (or the row is not
//    count-valued), mul_m is the power of two that puts the row maximum in
//    (16384, 32768]: no overflow and no subnormals for small form factors.
//  - FP32 storage and all-zero rows keep mul_m = 1.
__global__ void __launch_bounds__(256) k_row_mul(const float* __restrict__ F, int n_cols, int dtype,
                                                 const double* __restrict__ row_mul, float* __restrict__ mul_out,
                                                 double* __restrict__ row_scale) {
  __shared__ float red[4];
  const int m = blockIdx.x;
  const float* row = F + (size_t)m * n_cols;
  float mx = 0.f, mn = __builtin_huge_valf();
  for (int j = threadIdx.x; j < n_cols; j += blockDim.x) {
    const float a = fabsf(row[j]);
    mx = fmaxf(mx, a);
    if (a > 0.f) mn = fminf(mn, a);
  }
  mx = block_max(mx, red);
  mn = block_min(mn, red);
  float mul = 1.f;
  bool counts = false;
  if (dtype != FFK_F32 && mx > 0.f) {
    if (row_mul) {
      mul = (float)row_mul[m];
      counts = true;
    } else {
      const float c = rintf(1.0f / mn);
      bool ok = c >= 1.f && c <= 16777216.f;
      if (ok) {
        float worst = 0.f;
        for (int j = threadIdx.x; j < n_cols; j += blockDim.x) {
          const float x = fabsf(row[j]) * c;
          worst = fmaxf(worst, fabsf(x - rintf(x)) - 4e-6f * x);
        }
        ok = block_max(worst, red) <= 1e-6f;
      }
      if (ok) {
        mul = c;
        counts = true;
      }
    }
    if (dtype == FFK_F16 && (!counts || mx * mul >= 32768.f)) {
      int e;
      frexpf(32768.f / mx, &e);  // 32768/mx = f * 2^e, f in [0.5, 1)
      mul = ldexpf(1.f, e - 1);
    }
  }
  if (threadIdx.x == 0) {
    mul_out[m] = mul;
    row_scale[m] = 1.0 / (double)mul;
  }
}

// St[j, m] = dtype(F[m, j] * mul_m) through 64 x 64 LDS tiles (coalesced
// reads along j, coalesced writes along m).
__global__ void __launch_bounds__(256) k_pack_t(const float* __restrict__ F, int n, int dtype,
                                                const float* __restrict__ mul, void* __restrict__ out) {
  __shared__ float tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int j0 = blockIdx.x * 64, m0 = blockIdx.y * 64;
  for (int r = ty; r < 64; r += 4) {
    const int m = m0 + r, j = j0 + tx;
    tile[r][tx] = (m < n && j < n) ? F[(size_t)m * n + j] * mul[m] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int j = j0 + r, m = m0 + tx;
    if (j < n && m < n) store_as(out, (size_t)j * n + m, dtype, tile[tx][r]);
  }
}

// E_m = (eps_f + eps_r) * sigma * T^4 * A_m in FP64, the reference's
// expression order (thermal_equation/form_factors.cpp:116-120).
__device__ inline double sb_power(double t, double area, double ef, double er) {
  const double power = kSigma * (t * t * t * t) * area;
  return ef * power + er * power;
}

// emission[m] += E_m and the scaled FP32 copy E_m * scale_m for the product.
__global__ void k_sb_emission(int n, const double* __restrict__ T, const double* __restrict__ area,
                              const double* __restrict__ ef, const double* __restrict__ er,
                              const double* __restrict__ scale, double* __restrict__ emission,
                              float* __restrict__ E32) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double e = sb_power(T[m], area[m], ef[m], er[m]);
  if (emission) emission[m] += e;
  E32[m] = (float)(scale ? e * scale[m] : e);
}

// Conduction over a CSR operator: flow_m = sum_k G_mk (T_k - T_m).
__device__ inline double conduction_flow(int m, const int* __restrict__ row_ptr, const int* __restrict__ cols,
                                         const double* __restrict__ G, const double* __restrict__ T) {
  const double tm = T[m];
  double flow = 0.0;
  for (int k = row_ptr[m]; k < row_ptr[m + 1]; ++k) flow += G[k] * (T[cols[k]] - tm);
  return flow;
}

// The reference splits each face pair's flow (thermal_equation/conductive_cpu.cpp:123-132):
// the mesh it flows into absorbs it, the other emits it. Per neighbour k of
// mesh m: f = G_mk (T_k - T_m) > 0 is absorbed by m, else -f is emitted.
__global__ void k_conduction(int n, const int* __restrict__ row_ptr, const int* __restrict__ cols,
                             const double* __restrict__ G, const double* __restrict__ T,
                             double* __restrict__ emission, double* __restrict__ absorption) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double tm = T[m];
  double ab = 0.0, em = 0.0;
  for (int k = row_ptr[m]; k < row_ptr[m + 1]; ++k) {
    const double f = G[k] * (T[cols[k]] - tm);
    if (f > 0) ab += f;
    else em -= f;
  }
  absorption[m] += ab;
  emission[m] += em;
}

// thermal_solution/cpu_adams.cpp:136-142 coefficients (AB1..AB5).
__constant__ double c_u[5][5] = {
    {1.0, 0, 0, 0, 0},
    {3.0 / 2, -1.0 / 2, 0, 0, 0},
    {23.0 / 12, -4.0 / 3, 5.0 / 12, 0, 0},
    {55.0 / 24, -59.0 / 24, 37.0 / 24, -3.0 / 8, 0},
    {1901.0 / 720, -1387.0 / 360, 109.0 / 30, -637.0 / 360, 251.0 / 720}};

// Records E_n = net / C in the ring and returns T_n + dt * sum u[k][l] E_{n-l}.
__device__ inline double adams_update(int n, int m, int n_step, double dt, double net, double C,
                                      double* __restrict__ E_hist, double tn) {
  E_hist[(size_t)(n_step % 5) * n + m] = net / C;
  const int k = n_step < 4 ? n_step : 4;
  double sum = 0.0;
  for (int l = 0; l <= k; ++l) sum += c_u[k][l] * E_hist[(size_t)((n_step - l) % 5) * n + m];
  return tn + dt * sum;
}

// Conduction from the step's float temperatures into its float arrays: the
// double per-pair sums of k_conduction (from zero), rounded to float and added.
__global__ void k_conduction_f(int n, const int* __restrict__ row_ptr, const int* __restrict__ cols,
                               const double* __restrict__ G, const float* __restrict__ T,
                               float* __restrict__ emission, float* __restrict__ absorption) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double tm = T[m];
  double ab = 0.0, em = 0.0;
  for (int k = row_ptr[m]; k < row_ptr[m + 1]; ++k) {
    const double f = G[k] * ((double)T[cols[k]] - tm);
    if (f > 0) ab += f;
    else em -= f;
  }
  absorption[m] += (float)(0.0 + ab);
  emission[m] += (float)(0.0 + em);
}

// Heat sources in list order (one thread: a handful of additions, some of
// which may hit the same mesh).
__global__ void k_heat_sources(int n_sources, const int* __restrict__ mesh, const float* __restrict__ power,
                               float* __restrict__ emission, float* __restrict__ absorption) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int i = 0; i < n_sources; ++i) {
    const float p = power[i];
    if (p > 0) absorption[mesh[i]] += p;
    else emission[mesh[i]] += -p;
  }
}

// SB_FF from the step's float temperatures: T promoted to double, the double
// emission and absorption zeroed (what the host form uploads).
__global__ void k_sbff_begin(int n, const float* __restrict__ T_f, double* __restrict__ T, double* __restrict__ em,
                             double* __restrict__ ab) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  T[m] = (double)T_f[m];
  em[m] = 0.0;
  ab[m] = 0.0;
}

// Double sums rounded to float and added to the step's float arrays.
__global__ void k_add_rounded(int n, const double* __restrict__ em, const double* __restrict__ ab,
                              float* __restrict__ em_f, float* __restrict__ ab_f) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  em_f[m] += (float)em[m];
  ab_f[m] += (float)ab[m];
}

__global__ void k_step_begin(int n, const double* __restrict__ T_n, float* __restrict__ T_f,
                             float* __restrict__ emission, float* __restrict__ absorption) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  T_f[m] = (float)T_n[m];
  emission[m] = 0.f;
  absorption[m] = 0.f;
}

template <class In>
__global__ void k_adams(int n, int n_step, double dt, const double* __restrict__ C,
                        const In* __restrict__ emission, const In* __restrict__ absorption,
                        double* __restrict__ E_hist, const double* __restrict__ T_n,
                        double* __restrict__ T_next, double* __restrict__ balance) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  double net = 0.0, stored = 0.0;
  if (m < n) {
    net = (double)absorption[m] - (double)emission[m];
    const double tn = T_n[m];
    const double tx = adams_update(n, m, n_step, dt, net, C[m], E_hist, tn);
    T_next[m] = tx;
    stored = C[m] * (tx - tn) / dt;
  }
  if (balance) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      net += __shfl_xor(net, off, 64);
      stored += __shfl_xor(stored, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&balance[0], net);
      atomicAdd(&balance[1], stored);
    }
  }
}

// ------------------------------------------------ the exchange product --
// q_j = sum_m St[j, m] * E32[m]: one wavefront per target row j at a time
// (rows j = wave, wave + n_waves, ...), 8 consecutive sources per lane and
// UNROLL such 16-byte (BF16/FP16) or 32-byte (FP32) loads in flight; St is
// streamed with non-temporal loads. E32 is staged in LDS when it fits
// (ELDS), else read through the caches. Each lane keeps 8 FP32 partial sums;
// the lane sums and the wavefront tree are in a fixed order (deterministic).
constexpr int kRowBlock = 1024;     // 16 wavefronts
constexpr int kRowWaves = kRowBlock / 64;
constexpr int kMaxLdsMeshes = 32768;  // E32 in LDS up to 128 KB
constexpr int kDefaultUnroll = 4;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int DT>
__device__ inline void load8(const void* S, size_t base, float* v) {
  if (DT == FFK_F32) {
    const f32x4* p = reinterpret_cast<const f32x4*>(reinterpret_cast<const float*>(S) + base);
    const f32x4 a = __builtin_nontemporal_load(p);
    const f32x4 b = __builtin_nontemporal_load(p + 1);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
    const u32x4 a = __builtin_nontemporal_load(
        reinterpret_cast<const u32x4*>(reinterpret_cast<const uint16_t*>(S) + base));
    const unsigned w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint16_t lo = (uint16_t)(w[k] & 0xffffu), hi = (uint16_t)(w[k] >> 16);
      if (DT == FFK_BF16) {
        v[2 * k] = bf2f(lo);
        v[2 * k + 1] = bf2f(hi);
      } else {
        v[2 * k] = __half2float(__ushort_as_half(lo));
        v[2 * k + 1] = __half2float(__ushort_as_half(hi));
      }
    }
  }
}

template <int DT>
__device__ inline float load1(const void* S, size_t i) {
  if (DT == FFK_F32) return reinterpret_cast<const float*>(S)[i];
  const uint16_t h = reinterpret_cast<const uint16_t*>(S)[i];
  return DT == FFK_BF16 ? bf2f(h) : __half2float(__ushort_as_half(h));
}

template <int DT, int U>
__device__ inline float row_dot(int n, int j, int lane, const void* __restrict__ S, const float* __restrict__ E) {
  float acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = 0.f;
  const size_t row = (size_t)j * n;
  if ((n & 7) == 0) {  // rows start on 16-byte boundaries
    int m = lane * 8;
    for (; m + (U - 1) * 512 < n; m += U * 512) {
      float v[U][8];
#pragma unroll
      for (int u = 0; u < U; ++u) load8<DT>(S, row + m + u * 512, v[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const f32x4 e0 = *reinterpret_cast<const f32x4*>(E + m + u * 512);
        const f32x4 e1 = *reinterpret_cast<const f32x4*>(E + m + u * 512 + 4);
        acc[0] += v[u][0] * e0.x; acc[1] += v[u][1] * e0.y; acc[2] += v[u][2] * e0.z; acc[3] += v[u][3] * e0.w;
        acc[4] += v[u][4] * e1.x; acc[5] += v[u][5] * e1.y; acc[6] += v[u][6] * e1.z; acc[7] += v[u][7] * e1.w;
      }
    }
    for (; m < n; m += 512) {
      float v[8];
      load8<DT>(S, row + m, v);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += v[k] * E[m + k];
    }
  } else {
    for (int m = lane; m < n; m += 64) acc[0] += load1<DT>(S, row + m) * E[m];
  }
  float q = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
  return q;
}

// absorption[j] += q_j (the one-shot ffk_sbff path).
template <int DT, int U, bool ELDS>
__global__ void __launch_bounds__(kRowBlock) k_exchange_rows(int n, const void* __restrict__ S,
                                                             const float* __restrict__ E32,
                                                             double* __restrict__ absorption) {
  extern __shared__ float e_lds[];
  if (ELDS) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) e_lds[i] = E32[i];
    __syncthreads();
  }
  const float* E = ELDS ? e_lds : E32;
  const int lane = threadIdx.x & 63;
  const int n_waves = gridDim.x * kRowWaves;
  for (int j = blockIdx.x * kRowWaves + (threadIdx.x >> 6); j < n; j += n_waves) {
    const float q = row_dot<DT, U>(n, j, lane, S, E);
    if (lane == 0) absorption[j] += (double)q;
  }
}

// ------------------------------------------------ device-resident run --

struct step_args {
  int n;
  const void* S;  // packed exchange storage (NULL: no radiative exchange)
  const float* E32;   // scaled FP32 emission of the current temperature
  float* E32_next;    // ... of the new temperature, for the next step
  const double *T, *C, *area, *ef, *er, *scale, *power, *G;
  const int *row_ptr, *cols;
  double *T_next, *E_hist, *balance;
  int* ctl;  // [0] step, [1] first step recorded by this run, [2] blocks done
  double dt;
  int bal_cap;
};

// Everything of one step for mesh j after the exchange product: the mesh's
// own emission, conduction, heat sources (positive power absorbed, negative
// emitted: thermal_equation/heat_source_cpu.cpp:48-61), the Adams update into
// the other temperature buffer and the scaled FP32 emission of the new
// temperature for the next step. Adds the net and stored power to the sums.
__device__ inline void mesh_update(const step_args& a, int j, int n_step, float q, double& bal_net,
                                   double& bal_stored) {
  const double tm = a.T[j];
  double ab = (double)q, em = sb_power(tm, a.area[j], a.ef[j], a.er[j]);
  if (a.row_ptr) {
    const double flow = conduction_flow(j, a.row_ptr, a.cols, a.G, a.T);
    if (flow > 0) ab += flow;
    else em -= flow;
  }
  if (a.power) {
    const double p = a.power[j];
    if (p > 0) ab += p;
    else if (p < 0) em -= p;
  }
  const double net = ab - em;
  const double Cj = a.C[j];
  const double tx = adams_update(a.n, j, n_step, a.dt, net, Cj, a.E_hist, tm);
  a.T_next[j] = tx;
  bal_net += net;
  bal_stored += Cj * (tx - tm) / a.dt;
  const double e1 = sb_power(tx, a.area[j], a.ef[j], a.er[j]);
  a.E32_next[j] = (float)(a.scale ? e1 * a.scale[j] : e1);
}

// One whole step. Each wavefront takes rows j = wave, wave + n_waves, ...;
// for each it computes the absorbed radiation q_j (exchange product) and
// parks it in lane r (the row's rank within the wavefront). The mesh updates
// of up to 64 parked rows then run together, one lane per mesh, so their
// dependent FP64 loads do not stall the HBM stream row by row. The last
// block to finish advances the step counter; the step's energy-balance
// record is optional.
template <int DT, int U, bool ELDS>
__global__ void __launch_bounds__(kRowBlock) k_step(step_args a) {
  extern __shared__ float e_lds[];
  __shared__ double bal_s[kRowWaves][2];
  const int n = a.n;
  if (ELDS && a.S) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) e_lds[i] = a.E32[i];
    __syncthreads();
  }
  const float* E = ELDS ? e_lds : a.E32;
  const int n_step = __builtin_nontemporal_load(&a.ctl[0]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n_waves = gridDim.x * kRowWaves;
  double bal_net = 0.0, bal_stored = 0.0;
  float my_q = 0.f;
  int my_j = -1, parked = 0;
  for (int j = blockIdx.x * kRowWaves + w; j < n; j += n_waves) {
    const float q = a.S ? row_dot<DT, U>(n, j, lane, a.S, E) : 0.f;
    if (lane == parked) {
      my_q = q;
      my_j = j;
    }
    if (++parked == 64) {
      if (my_j >= 0) mesh_update(a, my_j, n_step, my_q, bal_net, bal_stored);
      my_j = -1;
      parked = 0;
    }
  }
  if (my_j >= 0) mesh_update(a, my_j, n_step, my_q, bal_net, bal_stored);
  const int rec = n_step - a.ctl[1];
  if (a.balance && rec >= 0 && rec < a.bal_cap) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      bal_net += __shfl_xor(bal_net, off, 64);
      bal_stored += __shfl_xor(bal_stored, off, 64);
    }
    if (lane == 0) {
      bal_s[w][0] = bal_net;
      bal_s[w][1] = bal_stored;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double s0 = 0.0, s1 = 0.0;
      for (int k = 0; k < kRowWaves; ++k) {
        s0 += bal_s[k][0];
        s1 += bal_s[k][1];
      }
      atomicAdd(&a.balance[2 * (size_t)rec], s0);
      atomicAdd(&a.balance[2 * (size_t)rec + 1], s1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&a.ctl[2], 1) == (int)gridDim.x - 1) {
      a.ctl[2] = 0;
      a.ctl[0] = n_step + 1;
      __threadfence();
    }
  }
}

// Scaled FP32 emission of the current temperature (start of a run).
__global__ void k_emission32(int n, const double* __restrict__ T, const double* __restrict__ area,
                             const double* __restrict__ ef, const double* __restrict__ er,
                             const double* __restrict__ scale, float* __restrict__ E32) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double e = sb_power(T[m], area[m], ef[m], er[m]);
  E32[m] = (float)(scale ? e * scale[m] : e);
}

__global__ void k_fill_f64(double* __restrict__ p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_set_ctl(int* __restrict__ ctl, int step, int base, int reset_step) {
  if (threadIdx.x == 0) {
    if (reset_step) ctl[0] = step;
    ctl[1] = reset_step ? base : ctl[0];
    ctl[2] = 0;
  }
}

// Launch shape of the row kernels: E32 in LDS for n <= kMaxLdsMeshes, grid
// = blocks resident on the device (at most one wavefront per row).
struct row_launch {
  int blocks;
  size_t lds;
  bool elds;
};

template <class K>
static row_launch plan_rows(K kernel_lds, K kernel_l2, int n, int max_blocks) {
  row_launch r;
  r.elds = n <= kMaxLdsMeshes;
  r.lds = r.elds ? (size_t)n * 4 : 0;
  const void* k = reinterpret_cast<const void*>(r.elds ? kernel_lds : kernel_l2);
  if (r.lds > 65536) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)r.lds);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kRowBlock, r.lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  int blocks = per_cu * device_cu_count();
  if (max_blocks > 0) blocks = std::min(blocks, max_blocks);
  blocks = std::min(blocks, (n + kRowWaves - 1) / kRowWaves);
  r.blocks = std::max(blocks, 1);
  return r;
}

template <int DT, int U>
static row_launch plan_step_t(int n, int max_blocks) {
  return plan_rows(&k_step<DT, U, true>, &k_step<DT, U, false>, n, max_blocks);
}

template <int DT>
static row_launch plan_step_u(int unroll, int n, int max_blocks) {
  if (unroll == 2) return plan_step_t<DT, 2>(n, max_blocks);
  if (unroll == 8) return plan_step_t<DT, 8>(n, max_blocks);
  return plan_step_t<DT, 4>(n, max_blocks);
}

// Launch plan of the step kernel; computed outside stream capture.
static row_launch plan_step(int dtype, int unroll, int n, bool has_exchange, int max_blocks) {
  row_launch L = dtype == FFK_BF16  ? plan_step_u<FFK_BF16>(unroll, n, max_blocks)
                 : dtype == FFK_F16 ? plan_step_u<FFK_F16>(unroll, n, max_blocks)
                                    : plan_step_u<FFK_F32>(unroll, n, max_blocks);
  if (!has_exchange) {
    L.elds = false;
    L.lds = 0;
  }
  return L;
}

template <int DT, int U>
static void launch_step_t(const step_args& a, const row_launch& L, hipStream_t st) {
  if (L.elds) k_step<DT, U, true><<<L.blocks, kRowBlock, L.lds, st>>>(a);
  else k_step<DT, U, false><<<L.blocks, kRowBlock, 0, st>>>(a);
}

template <int DT>
static void launch_step_u(int unroll, const step_args& a, const row_launch& L, hipStream_t st) {
  if (unroll == 2) launch_step_t<DT, 2>(a, L, st);
  else if (unroll == 8) launch_step_t<DT, 8>(a, L, st);
  else launch_step_t<DT, 4>(a, L, st);
}

static void launch_step(int dtype, int unroll, const step_args& a, const row_launch& L, hipStream_t st) {
  if (dtype == FFK_BF16) launch_step_u<FFK_BF16>(unroll, a, L, st);
  else if (dtype == FFK_F16) launch_step_u<FFK_F16>(unroll, a, L, st);
  else launch_step_u<FFK_F32>(unroll, a, L, st);
}

template <int DT>
static void launch_rows(int n, const void* S, const float* E32, double* absorption, hipStream_t st) {
  const row_launch L =
      plan_rows(&k_exchange_rows<DT, kDefaultUnroll, true>, &k_exchange_rows<DT, kDefaultUnroll, false>, n, 0);
  if (L.elds) k_exchange_rows<DT, kDefaultUnroll, true><<<L.blocks, kRowBlock, L.lds, st>>>(n, S, E32, absorption);
  else k_exchange_rows<DT, kDefaultUnroll, false><<<L.blocks, kRowBlock, 0, st>>>(n, S, E32, absorption);
}

}  // namespace ffk

using namespace ffk;

struct ffk_thermal {
  int n = 0;
  double *d_C = nullptr, *d_area = nullptr, *d_ef = nullptr, *d_er = nullptr;
  double* d_T[2] = {nullptr, nullptr};  // ping-pong temperature buffers
  int cur = 0;                          // buffer holding the current temperature
  double* d_E = nullptr;                // Adams ring, 5 rows of n
  double* d_power = nullptr;
  bool has_sources = false;
  int *d_row = nullptr, *d_col = nullptr;
  double* d_G = nullptr;
  bool has_conduction = false;
  const void* d_F = nullptr;
  int dtype = FFK_F32;
  double* d_scale = nullptr;  // per-row scale of the packed exchange storage
  bool has_scale = false;
  float* d_E32[2] = {nullptr, nullptr};  // scaled FP32 emission, with the temperature's parity
  int unroll = kDefaultUnroll, max_blocks = 0;
  row_launch plan = {0, 0, false};
  int* d_ctl = nullptr;
  double* d_bal = nullptr;
  int bal_cap = 0;
  double dt = 0.0;
  hipStream_t cap = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
};

namespace ffk {

static void thermal_free_graph(ffk_thermal* t) {
  if (t->exec) (void)hipGraphExecDestroy(t->exec);
  if (t->graph) (void)hipGraphDestroy(t->graph);
  t->exec = nullptr;
  t->graph = nullptr;
}

// One step on stream st: one kernel.
static int thermal_enqueue_step(ffk_thermal* t, hipStream_t st) {
  step_args a;
  a.n = t->n;
  a.S = t->d_F;
  a.E32 = t->d_E32[t->cur];
  a.E32_next = t->d_E32[t->cur ^ 1];
  a.T = t->d_T[t->cur];
  a.T_next = t->d_T[t->cur ^ 1];
  a.C = t->d_C;
  a.area = t->d_area;
  a.ef = t->d_ef;
  a.er = t->d_er;
  a.scale = t->has_scale ? t->d_scale : nullptr;
  a.power = t->has_sources ? t->d_power : nullptr;
  a.row_ptr = t->has_conduction ? t->d_row : nullptr;
  a.cols = t->d_col;
  a.G = t->d_G;
  a.E_hist = t->d_E;
  a.balance = t->bal_cap > 0 ? t->d_bal : nullptr;
  a.bal_cap = t->bal_cap;
  a.ctl = t->d_ctl;
  a.dt = t->dt;
  launch_step(t->dtype, t->unroll, a, t->plan, st);
  FFK_HIP(hipGetLastError());
  t->cur ^= 1;
  return FFK_OK;
}

}  // namespace ffk

extern "C" {

int ffk_convert_f(const float* d_F, long long n_elems, int dtype, void* d_out, void* stream) {
  FFK_REQUIRE(d_F && d_out && n_elems >= 0, FFK_ERR_ARG, "ffk_convert_f: bad args");
  FFK_REQUIRE(dtype == FFK_F32 || dtype == FFK_BF16 || dtype == FFK_F16, FFK_ERR_ARG, "dtype %d", dtype);
  if (n_elems == 0) return FFK_OK;
  int blocks = (int)std::min<long long>((n_elems + 255) / 256, 65536);
  k_convert<<<blocks, 256, 0, as_stream(stream)>>>(d_F, n_elems, dtype, d_out);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_pack_exchange(const float* d_F, int n, int dtype, const double* d_row_mul, void* d_out,
                      double* d_row_scale, void* stream) {
  FFK_REQUIRE(d_F && d_out && d_row_scale && n > 0, FFK_ERR_ARG, "ffk_pack_exchange: bad args");
  FFK_REQUIRE(dtype == FFK_F32 || dtype == FFK_BF16 || dtype == FFK_F16, FFK_ERR_ARG, "dtype %d", dtype);
  hipStream_t st = as_stream(stream);
  float* d_mul = nullptr;
  int r;
  if ((r = dalloc(&d_mul, (size_t)n)) < 0) return r;
  k_row_mul<<<(unsigned)n, 256, 0, st>>>(d_F, n, dtype, d_row_mul, d_mul, d_row_scale);
  const unsigned tiles = (unsigned)((n + 63) / 64);
  k_pack_t<<<dim3(tiles, tiles), 256, 0, st>>>(d_F, n, dtype, d_mul, d_out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dfree(d_mul);
  FFK_HIP(e);
  return FFK_OK;
}

int ffk_sbff(int n, const void* d_S, int dtype, const double* d_row_scale, const double* d_T,
             const double* d_area, const double* d_eps_front, const double* d_eps_rear, double* d_emission,
             double* d_absorption, float* d_work, void* stream) {
  FFK_REQUIRE(n > 0 && d_S && d_T && d_area && d_eps_front && d_eps_rear && d_emission && d_absorption && d_work,
              FFK_ERR_ARG, "ffk_sbff: bad args");
  FFK_REQUIRE(dtype == FFK_F32 || dtype == FFK_BF16 || dtype == FFK_F16, FFK_ERR_ARG, "dtype %d", dtype);
  hipStream_t st = as_stream(stream);
  float* E32 = d_work;
  k_sb_emission<<<(n + 255) / 256, 256, 0, st>>>(n, d_T, d_area, d_eps_front, d_eps_rear, d_row_scale, d_emission,
                                                 E32);
  if (dtype == FFK_BF16) launch_rows<FFK_BF16>(n, d_S, E32, d_absorption, st);
  else if (dtype == FFK_F16) launch_rows<FFK_F16>(n, d_S, E32, d_absorption, st);
  else launch_rows<FFK_F32>(n, d_S, E32, d_absorption, st);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_sbff_dev(int n, const void* d_S, int dtype, const double* d_row_scale, const float* d_T,
                 const double* d_area, const double* d_eps_front, const double* d_eps_rear, double* d_T64,
                 double* d_em64, double* d_ab64, float* d_work, float* d_emission, float* d_absorption,
                 void* stream) {
  FFK_REQUIRE(n > 0 && d_T && d_T64 && d_em64 && d_ab64 && d_emission && d_absorption, FFK_ERR_ARG,
              "ffk_sbff_dev: bad args");
  hipStream_t st = as_stream(stream);
  k_sbff_begin<<<(n + 255) / 256, 256, 0, st>>>(n, d_T, d_T64, d_em64, d_ab64);
  FFK_HIP(hipGetLastError());
  int r;
  if ((r = ffk_sbff(n, d_S, dtype, d_row_scale, d_T64, d_area, d_eps_front, d_eps_rear, d_em64, d_ab64, d_work,
                    stream)) < 0)
    return r;
  k_add_rounded<<<(n + 255) / 256, 256, 0, st>>>(n, d_em64, d_ab64, d_emission, d_absorption);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

}  // extern "C"

// ---- allocation-free SB_FF ------------------------------------------------------
// ffk_sbff_noalloc's scratch: the scaled FP32 emission vector, in a static
// device array instead of caller or library allocations. Each device holds
// its own copy for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns on that device (the array's
// name in host code is a stand-in, not a device address).
static __device__ __align__(256) float scratch[FFK_NOALLOC_MAX_MESHES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call made on that device and kept.
int scratch_on_current_device(float** out) {
  int dev = 0;
  FFK_HIP(hipGetDevice(&dev));
  FFK_REQUIRE(dev >= 0 && dev < kMaxDevices, FFK_ERR_LIMIT, "device index %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    FFK_REQUIRE(e == hipSuccess && p, FFK_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FFK_OK;
}
}  // namespace

extern "C" {

int ffk_sbff_noalloc(int n, const void* d_S, int dtype, const double* d_row_scale, const double* d_T,
                     const double* d_area, const double* d_eps_front, const double* d_eps_rear, double* d_emission,
                     double* d_absorption, void* stream) {
  FFK_REQUIRE(n > 0, FFK_ERR_ARG, "ffk_sbff_noalloc: n = %d", n);
  FFK_REQUIRE(n <= FFK_NOALLOC_MAX_MESHES, FFK_ERR_LIMIT,
              "ffk_sbff_noalloc: %d meshes exceed the static scratch of %d meshes (use ffk_sbff)", n,
              FFK_NOALLOC_MAX_MESHES);
  float* work = nullptr;
  int r;
  if ((r = scratch_on_current_device(&work)) < 0) return r;
  // the same kernels as ffk_sbff; the scratch is written in full (n values)
  // before it is read, so nothing left by an earlier call is used
  return ffk_sbff(n, d_S, dtype, d_row_scale, d_T, d_area, d_eps_front, d_eps_rear, d_emission, d_absorption, work,
                  stream);
}

int ffk_conduction(int n, const int* d_row_ptr, const int* d_cols, const double* d_G, const double* d_T,
                   double* d_emission, double* d_absorption, void* stream) {
  FFK_REQUIRE(n > 0 && d_row_ptr && d_T && d_emission && d_absorption, FFK_ERR_ARG, "ffk_conduction: bad args");
  k_conduction<<<(n + 255) / 256, 256, 0, as_stream(stream)>>>(n, d_row_ptr, d_cols, d_G, d_T, d_emission,
                                                              d_absorption);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_adams_step(int n, int n_step, double dt, const double* d_C, const double* d_emission,
                   const double* d_absorption, double* d_E_hist, const double* d_T_n, double* d_T_next,
                   double* d_balance, void* stream) {
  FFK_REQUIRE(n > 0 && n_step >= 0 && d_C && d_emission && d_absorption && d_E_hist && d_T_n && d_T_next,
              FFK_ERR_ARG, "ffk_adams_step: bad args");
  k_adams<double><<<(n + 255) / 256, 256, 0, as_stream(stream)>>>(n, n_step, dt, d_C, d_emission, d_absorption,
                                                                 d_E_hist, d_T_n, d_T_next, d_balance);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}
int ffk_adams_step_f(int n, int n_step, double dt, const double* d_C, const float* d_emission,
                     const float* d_absorption, double* d_E_hist, const double* d_T_n, double* d_T_next,
                     double* d_balance, void* stream) {
  FFK_REQUIRE(n > 0 && n_step >= 0 && d_C && d_emission && d_absorption && d_E_hist && d_T_n && d_T_next,
              FFK_ERR_ARG, "ffk_adams_step_f: bad args");
  k_adams<float><<<(n + 255) / 256, 256, 0, as_stream(stream)>>>(n, n_step, dt, d_C, d_emission, d_absorption,
                                                                d_E_hist, d_T_n, d_T_next, d_balance);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}
int ffk_conduction_dev(int n, const int* d_row_ptr, const int* d_cols, const double* d_G, const float* d_T,
                       float* d_emission, float* d_absorption, void* stream) {
  FFK_REQUIRE(n > 0 && d_row_ptr && d_T && d_emission && d_absorption, FFK_ERR_ARG, "ffk_conduction_dev: bad args");
  k_conduction_f<<<(n + 255) / 256, 256, 0, as_stream(stream)>>>(n, d_row_ptr, d_cols, d_G, d_T, d_emission,
                                                                d_absorption);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}
int ffk_heat_sources_dev(int n_sources, const int* d_mesh, const float* d_power, float* d_emission,
                         float* d_absorption, void* stream) {
  FFK_REQUIRE(n_sources >= 0 && (n_sources == 0 || (d_mesh && d_power)) && d_emission && d_absorption, FFK_ERR_ARG,
              "ffk_heat_sources_dev: bad args");
  if (n_sources == 0) return FFK_OK;
  k_heat_sources<<<1, 64, 0, as_stream(stream)>>>(n_sources, d_mesh, d_power, d_emission, d_absorption);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}
int ffk_step_begin_dev(int n, const double* d_T_n, float* d_T_f, float* d_emission, float* d_absorption,
                       void* stream) {
  FFK_REQUIRE(n > 0 && d_T_n && d_T_f && d_emission && d_absorption, FFK_ERR_ARG, "ffk_step_begin_dev: bad args");
  k_step_begin<<<(n + 255) / 256, 256, 0, as_stream(stream)>>>(n, d_T_n, d_T_f, d_emission, d_absorption);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

int ffk_thermal_create(ffk_thermal** out, int n, const double* h_C, const double* h_area,
                       const double* h_eps_front, const double* h_eps_rear) {
  FFK_REQUIRE(out && n > 0 && h_C && h_area && h_eps_front && h_eps_rear, FFK_ERR_ARG,
              "ffk_thermal_create: bad args");
  ffk_thermal* t = new ffk_thermal();
  t->n = n;
  int r;
  if ((r = dalloc(&t->d_C, n)) < 0 || (r = dalloc(&t->d_area, n)) < 0 || (r = dalloc(&t->d_ef, n)) < 0 ||
      (r = dalloc(&t->d_er, n)) < 0 || (r = dalloc(&t->d_T[0], n)) < 0 || (r = dalloc(&t->d_T[1], n)) < 0 ||
      (r = dalloc(&t->d_E, 5 * (size_t)n)) < 0 || (r = dalloc(&t->d_power, n)) < 0 ||
      (r = dalloc(&t->d_scale, n)) < 0 || (r = dalloc(&t->d_E32[0], n)) < 0 ||
      (r = dalloc(&t->d_E32[1], n)) < 0 || (r = dalloc(&t->d_ctl, 4)) < 0) {
    ffk_thermal_destroy(t);
    return r;
  }
  if (hipMemcpy(t->d_C, h_C, 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(t->d_area, h_area, 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(t->d_ef, h_eps_front, 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(t->d_er, h_eps_rear, 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipStreamCreateWithFlags(&t->cap, hipStreamNonBlocking) != hipSuccess) {
    set_error("ffk_thermal_create: HIP copy failed");
    ffk_thermal_destroy(t);
    return FFK_ERR_HIP;
  }
  k_set_ctl<<<1, 64>>>(t->d_ctl, 0, 0, 1);
  if (hipDeviceSynchronize() != hipSuccess) {
    set_error("ffk_thermal_create: initialisation failed");
    ffk_thermal_destroy(t);
    return FFK_ERR_HIP;
  }
  *out = t;
  return FFK_OK;
}

void ffk_thermal_destroy(ffk_thermal* t) {
  if (!t) return;
  (void)hipDeviceSynchronize();
  thermal_free_graph(t);
  if (t->cap) (void)hipStreamDestroy(t->cap);
  dfree(t->d_C);
  dfree(t->d_area);
  dfree(t->d_ef);
  dfree(t->d_er);
  dfree(t->d_T[0]);
  dfree(t->d_T[1]);
  dfree(t->d_E);
  dfree(t->d_power);
  dfree(t->d_scale);
  dfree(t->d_E32[0]);
  dfree(t->d_E32[1]);
  dfree(t->d_ctl);
  dfree(t->d_row);
  dfree(t->d_col);
  dfree(t->d_G);
  dfree(t->d_bal);
  delete t;
}

int ffk_thermal_set_exchange(ffk_thermal* t, const void* d_F, int dtype, const double* d_row_scale) {
  FFK_REQUIRE(t, FFK_ERR_ARG, "thermal is NULL");
  FFK_REQUIRE(dtype == FFK_F32 || dtype == FFK_BF16 || dtype == FFK_F16, FFK_ERR_ARG, "dtype %d", dtype);
  t->d_F = d_F;
  t->dtype = dtype;
  t->has_scale = d_row_scale != nullptr;
  if (d_row_scale) FFK_HIP(hipMemcpy(t->d_scale, d_row_scale, 8 * (size_t)t->n, hipMemcpyDeviceToDevice));
  thermal_free_graph(t);
  return FFK_OK;
}

int ffk_thermal_set_tuning(ffk_thermal* t, int unroll, int max_blocks) {
  FFK_REQUIRE(t, FFK_ERR_ARG, "thermal is NULL");
  if (unroll == 0) unroll = kDefaultUnroll;
  FFK_REQUIRE(unroll == 2 || unroll == 4 || unroll == 8, FFK_ERR_ARG, "unroll %d: 2, 4 or 8", unroll);
  FFK_REQUIRE(max_blocks >= 0, FFK_ERR_ARG, "max_blocks %d", max_blocks);
  t->unroll = unroll;
  t->max_blocks = max_blocks;
  thermal_free_graph(t);
  return FFK_OK;
}

int ffk_thermal_set_conduction(ffk_thermal* t, const int* h_row_ptr, const int* h_cols, const double* h_G) {
  FFK_REQUIRE(t && h_row_ptr, FFK_ERR_ARG, "ffk_thermal_set_conduction: bad args");
  const int nnz = h_row_ptr[t->n];
  FFK_REQUIRE(nnz >= 0 && (nnz == 0 || (h_cols && h_G)), FFK_ERR_ARG, "ffk_thermal_set_conduction: bad CSR");
  for (int m = 0; m < t->n; ++m)
    FFK_REQUIRE(h_row_ptr[m] >= 0 && h_row_ptr[m] <= h_row_ptr[m + 1], FFK_ERR_ARG, "CSR row %d", m);
  for (int m = 0; m < nnz; ++m)
    FFK_REQUIRE(h_cols[m] >= 0 && h_cols[m] < t->n, FFK_ERR_ARG, "CSR column %d out of range", h_cols[m]);
  dfree(t->d_row);
  dfree(t->d_col);
  dfree(t->d_G);
  t->d_row = nullptr;
  t->d_col = nullptr;
  t->d_G = nullptr;
  int r;
  if ((r = dalloc(&t->d_row, (size_t)t->n + 1)) < 0 || (r = dalloc(&t->d_col, std::max(1, nnz))) < 0 ||
      (r = dalloc(&t->d_G, std::max(1, nnz))) < 0)
    return r;
  FFK_HIP(hipMemcpy(t->d_row, h_row_ptr, 4 * ((size_t)t->n + 1), hipMemcpyHostToDevice));
  if (nnz) {
    FFK_HIP(hipMemcpy(t->d_col, h_cols, 4 * (size_t)nnz, hipMemcpyHostToDevice));
    FFK_HIP(hipMemcpy(t->d_G, h_G, 8 * (size_t)nnz, hipMemcpyHostToDevice));
  }
  t->has_conduction = nnz > 0;
  thermal_free_graph(t);
  return FFK_OK;
}

int ffk_thermal_set_sources(ffk_thermal* t, int n_sources, const int* h_mesh, const double* h_power) {
  FFK_REQUIRE(t && n_sources >= 0 && (n_sources == 0 || (h_mesh && h_power)), FFK_ERR_ARG,
              "ffk_thermal_set_sources: bad args");
  std::vector<double> p(t->n, 0.0);
  for (int i = 0; i < n_sources; ++i) {
    FFK_REQUIRE(h_mesh[i] >= 0 && h_mesh[i] < t->n, FFK_ERR_ARG, "source %d on mesh %d", i, h_mesh[i]);
    p[h_mesh[i]] += h_power[i];
  }
  FFK_HIP(hipMemcpy(t->d_power, p.data(), 8 * (size_t)t->n, hipMemcpyHostToDevice));
  t->has_sources = n_sources > 0;
  thermal_free_graph(t);
  return FFK_OK;
}

int ffk_thermal_run(ffk_thermal* t, const double* h_T0, int n_steps, double dt, double* h_T_out,
                    double* h_net_stored, int use_graph, void* stream) {
  FFK_REQUIRE(t && n_steps >= 0 && dt > 0, FFK_ERR_ARG, "ffk_thermal_run: bad args");
  hipStream_t st = as_stream(stream);
  const int n = t->n;
  // Balance records for this run (a graph holds the buffer's address).
  if (h_net_stored && n_steps > t->bal_cap) {
    FFK_HIP(hipStreamSynchronize(st));
    thermal_free_graph(t);
    dfree(t->d_bal);
    t->d_bal = nullptr;
    t->bal_cap = 0;
    int r;
    if ((r = dalloc(&t->d_bal, 2 * (size_t)n_steps)) < 0) return r;
    t->bal_cap = n_steps;
  }
  if (t->bal_cap > 0)
    k_fill_f64<<<(unsigned)((2LL * t->bal_cap + 255) / 256), 256, 0, st>>>(t->d_bal, 2LL * t->bal_cap);
  // Runs start from buffer 0: an even number of steps per graph then leaves
  // the current buffer unchanged, so one graph serves every launch.
  if (t->cur != 0) {
    FFK_HIP(hipMemcpyAsync(t->d_T[0], t->d_T[1], 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
    t->cur = 0;
  }
  if (h_T0) {
    FFK_HIP(hipMemcpyAsync(t->d_T[t->cur], h_T0, 8 * (size_t)n, hipMemcpyHostToDevice, st));
    k_fill_f64<<<(unsigned)((5LL * n + 255) / 256), 256, 0, st>>>(t->d_E, 5LL * n);
  }
  // step counter: restart at 0 with a new T0, else continue; records start here
  k_set_ctl<<<1, 64, 0, st>>>(t->d_ctl, 0, 0, h_T0 ? 1 : 0);
  k_emission32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, t->d_T[t->cur], t->d_area, t->d_ef, t->d_er,
                                                            t->has_scale ? t->d_scale : nullptr,
                                                            t->d_E32[t->cur]);
  FFK_HIP(hipGetLastError());
  if (dt != t->dt) thermal_free_graph(t);
  t->dt = dt;
  t->plan = plan_step(t->dtype, t->unroll, n, t->d_F != nullptr, t->max_blocks);
  const int per_graph = 10;
  int done = 0;
  if (use_graph && n_steps >= per_graph) {
    if (!t->exec) {
      FFK_HIP(hipStreamSynchronize(st));
      FFK_HIP(hipStreamBeginCapture(t->cap, hipStreamCaptureModeThreadLocal));
      int r = FFK_OK;
      for (int k = 0; k < per_graph && r == FFK_OK; ++k) r = thermal_enqueue_step(t, t->cap);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(t->cap, &g);
      t->cur = 0;
      if (r < 0) {
        if (g) (void)hipGraphDestroy(g);
        return r;
      }
      FFK_HIP(e);
      t->graph = g;
      FFK_HIP(hipGraphInstantiate(&t->exec, t->graph, nullptr, nullptr, 0));
    }
    for (; done + per_graph <= n_steps; done += per_graph) FFK_HIP(hipGraphLaunch(t->exec, st));
  }
  for (; done < n_steps; ++done) {
    int r = thermal_enqueue_step(t, st);
    if (r < 0) return r;
  }
  if (h_T_out) FFK_HIP(hipMemcpyAsync(h_T_out, t->d_T[t->cur], 8 * (size_t)n, hipMemcpyDeviceToHost, st));
  if (h_net_stored && n_steps > 0)
    FFK_HIP(hipMemcpyAsync(h_net_stored, t->d_bal, 16 * (size_t)n_steps, hipMemcpyDeviceToHost, st));
  FFK_HIP(hipStreamSynchronize(st));
  return FFK_OK;
}

}  // extern "C"
