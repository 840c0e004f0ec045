// Radiative exchange q = F^T E with F in BF16/FP16/FP32 storage and FP32
// accumulation, CSR conduction, and the FP64 Adams-Bashforth step.
#include <hip/hip_fp16.h>

#include <algorithm>

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

__global__ void k_convert(const float* __restrict__ in, long long n, int dtype, void* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    float v = in[i];
    if (dtype == FFK_BF16) reinterpret_cast<uint16_t*>(out)[i] = f2bf(v);
    else if (dtype == FFK_F16) reinterpret_cast<__half*>(out)[i] = __float2half_rn(v);
    else reinterpret_cast<float*>(out)[i] = v;
  }
}

// E_m = (eps_f + eps_r) * sigma * T^4 * A_m in FP64 (emission[m] += E_m),
// plus an FP32 copy for the exchange product.
__global__ void k_sb_emission(int n, const double* __restrict__ T, const double* __restrict__ area,
                              const double* __restrict__ ef, const double* __restrict__ er,
                              double* __restrict__ emission, float* __restrict__ E32) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double t = T[m];
  double power = kSigma * (t * t * t * t) * area[m];
  double e = ef[m] * power + er[m] * power;
  emission[m] += e;
  E32[m] = (float)e;
}

// Partial column sums over a chunk of rows: part[chunk][col] = sum over the
// chunk's rows m of E[m] * F[m, col]. Each thread owns 8 adjacent columns
// and reads 16 bytes (BF16/FP16) or 32 bytes (FP32) per row.
constexpr int kRowsPerChunk = 128;
constexpr int kColsPerThread = 8;

template <int DT>
__global__ void __launch_bounds__(256) k_exchange_partial(int n, const void* __restrict__ F,
                                                          const float* __restrict__ E,
                                                          float* __restrict__ part) {
  const int col0 = (blockIdx.x * blockDim.x + threadIdx.x) * kColsPerThread;
  const int chunk = blockIdx.y;
  const int m0 = chunk * kRowsPerChunk;
  const int m1 = min(n, m0 + kRowsPerChunk);
  if (col0 >= n) return;
  float acc[kColsPerThread];
#pragma unroll
  for (int k = 0; k < kColsPerThread; ++k) acc[k] = 0.f;
  const bool full = col0 + kColsPerThread <= n && (n % kColsPerThread) == 0;
  for (int m = m0; m < m1; ++m) {
    const float e = E[m];
    const size_t base = (size_t)m * n + col0;
    float v[kColsPerThread];
    if (DT == FFK_F32) {
      const float* row = reinterpret_cast<const float*>(F) + base;
      if (full) {
        float4 a = *reinterpret_cast<const float4*>(row);
        float4 b = *reinterpret_cast<const float4*>(row + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
      } else {
#pragma unroll
        for (int k = 0; k < kColsPerThread; ++k) v[k] = col0 + k < n ? row[k] : 0.f;
      }
    } else {
      const uint16_t* row = reinterpret_cast<const uint16_t*>(F) + base;
      uint16_t hv[kColsPerThread];
      if (full) {
        uint4 a = *reinterpret_cast<const uint4*>(row);
        hv[0] = a.x & 0xffff; hv[1] = a.x >> 16; hv[2] = a.y & 0xffff; hv[3] = a.y >> 16;
        hv[4] = a.z & 0xffff; hv[5] = a.z >> 16; hv[6] = a.w & 0xffff; hv[7] = a.w >> 16;
      } else {
#pragma unroll
        for (int k = 0; k < kColsPerThread; ++k) hv[k] = col0 + k < n ? row[k] : (uint16_t)0;
      }
#pragma unroll
      for (int k = 0; k < kColsPerThread; ++k) {
        if (DT == FFK_BF16) v[k] = bf2f(hv[k]);
        else v[k] = __half2float(__ushort_as_half(hv[k]));
      }
    }
#pragma unroll
    for (int k = 0; k < kColsPerThread; ++k) acc[k] += e * v[k];
  }
  float* p = part + (size_t)chunk * n + col0;
#pragma unroll
  for (int k = 0; k < kColsPerThread; ++k)
    if (col0 + k < n) p[k] = acc[k];
}

// absorption[col] += sum over chunks, in chunk order (deterministic).
__global__ void k_exchange_reduce(int n, int n_chunks, const float* __restrict__ part,
                                  double* __restrict__ absorption) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float s = 0.f;
  for (int k = 0; k < n_chunks; ++k) s += part[(size_t)k * n + c];
  absorption[c] += (double)s;
}

// Conduction over a CSR operator: flow_m = sum_k G_mk (T_k - T_m).
__global__ void k_conduction(int n, const int* __restrict__ row_ptr, const int* __restrict__ cols,
                             const double* __restrict__ G, const double* __restrict__ T,
                             double* __restrict__ emission, double* __restrict__ absorption) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double tm = T[m], flow = 0.0;
  for (int k = row_ptr[m]; k < row_ptr[m + 1]; ++k) flow += G[k] * (T[cols[k]] - tm);
  if (flow > 0) absorption[m] += flow;
  else emission[m] -= flow;
}

// thermal_solution/cpu_adams.cpp:136-142 coefficients (AB1..AB5).
__constant__ double c_u[5][5] = {
    {1.0, 0, 0, 0, 0},
    {3.0 / 2, -1.0 / 2, 0, 0, 0},
    {23.0 / 12, -4.0 / 3, 5.0 / 12, 0, 0},
    {55.0 / 24, -59.0 / 24, 37.0 / 24, -3.0 / 8, 0},
    {1901.0 / 720, -1387.0 / 360, 109.0 / 30, -637.0 / 360, 251.0 / 720}};

__global__ void k_adams(int n, int n_step, double dt, const double* __restrict__ C,
                        const double* __restrict__ emission, const double* __restrict__ absorption,
                        double* __restrict__ E_hist, const double* __restrict__ T_n,
                        double* __restrict__ T_next, double* __restrict__ balance) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  double net = 0.0, stored = 0.0;
  if (m < n) {
    net = absorption[m] - emission[m];
    E_hist[(size_t)(n_step % 5) * n + m] = net / C[m];
    const int k = n_step < 4 ? n_step : 4;
    double sum = 0.0;
    for (int l = 0; l <= k; ++l) sum += c_u[k][l] * E_hist[(size_t)((n_step - l) % 5) * n + m];
    const double tn = T_n[m];
    const double tx = tn + dt * sum;
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

// ------------------------------------------------ device-resident run --

__global__ void k_zero_f64(double* __restrict__ a, double* __restrict__ b, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = 0.0;
    b[i] = 0.0;
  }
}

// Heat sources and sinks (thermal_equation/heat_source_cpu.cpp:48-61),
// combined per mesh on the host: positive power is absorbed, negative emitted.
__global__ void k_sources(int n, const double* __restrict__ power, double* __restrict__ emission,
                          double* __restrict__ absorption) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double p = power[m];
  if (p > 0) absorption[m] += p;
  else if (p < 0) emission[m] -= p;
}

// Adams step with the step number read from device memory, so that a
// captured graph can be replayed step after step. Records, per step, the
// net power sum(absorption - emission) and the stored power
// sum(C (T_next - T_n) / dt) for the energy-balance check.
__global__ void k_adams_dev(int n, const int* __restrict__ d_step, double dt, const double* __restrict__ C,
                            const double* __restrict__ emission, const double* __restrict__ absorption,
                            double* __restrict__ E_hist, const double* __restrict__ T_n,
                            double* __restrict__ T_next, double* __restrict__ balance, const int* __restrict__ d_base,
                            int max_rec) {
  const int n_step = *d_step;
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  double net = 0.0, stored = 0.0;
  if (m < n) {
    net = absorption[m] - emission[m];
    E_hist[(size_t)(n_step % 5) * n + m] = net / C[m];
    const int k = n_step < 4 ? n_step : 4;
    double sum = 0.0;
    for (int l = 0; l <= k; ++l) sum += c_u[k][l] * E_hist[(size_t)((n_step - l) % 5) * n + m];
    const double tn = T_n[m];
    const double tx = tn + dt * sum;
    T_next[m] = tx;
    stored = C[m] * (tx - tn) / dt;
  }
  const int rec = n_step - *d_base;
  if (balance && rec >= 0 && rec < max_rec) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      net += __shfl_xor(net, off, 64);
      stored += __shfl_xor(stored, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&balance[2 * rec], net);
      atomicAdd(&balance[2 * rec + 1], stored);
    }
  }
}

__global__ void k_advance(int n, double* __restrict__ T, const double* __restrict__ T_next, int* __restrict__ d_step) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) T[m] = T_next[m];
  if (m == 0) *d_step += 1;
}

__global__ void k_fill_f64(double* __restrict__ p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace ffk

using namespace ffk;

struct ffk_thermal {
  int n = 0;
  double *d_C = nullptr, *d_area = nullptr, *d_ef = nullptr, *d_er = nullptr;
  double *d_T = nullptr, *d_Tn = nullptr, *d_em = nullptr, *d_ab = nullptr, *d_E = nullptr;
  double* d_power = nullptr;
  bool has_sources = false;
  int *d_row = nullptr, *d_col = nullptr;
  double* d_G = nullptr;
  bool has_conduction = false;
  const void* d_F = nullptr;
  int dtype = FFK_F32;
  float *d_E32 = nullptr, *d_part = nullptr;
  int n_chunks = 0;
  int* d_step = nullptr;  // [0] = step, [1] = first step recorded by this run
  double* d_bal = nullptr;
  int bal_cap = 0;
  double dt = 0.0;
  hipStream_t cap = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_steps = 0;
  double graph_dt = 0.0;
  double* graph_bal = nullptr;
};

namespace ffk {

static void thermal_free_graph(ffk_thermal* t) {
  if (t->exec) (void)hipGraphExecDestroy(t->exec);
  if (t->graph) (void)hipGraphDestroy(t->graph);
  t->exec = nullptr;
  t->graph = nullptr;
}

// One step on stream st.
static int thermal_enqueue_step(ffk_thermal* t, hipStream_t st) {
  const int n = t->n;
  const unsigned nb = (unsigned)((n + 255) / 256);
  k_zero_f64<<<nb, 256, 0, st>>>(t->d_em, t->d_ab, n);
  if (t->d_F) {
    k_sb_emission<<<nb, 256, 0, st>>>(n, t->d_T, t->d_area, t->d_ef, t->d_er, t->d_em, t->d_E32);
    const int cols_per_block = 256 * kColsPerThread;
    dim3 grid((unsigned)((n + cols_per_block - 1) / cols_per_block), (unsigned)t->n_chunks);
    if (t->dtype == FFK_BF16) k_exchange_partial<FFK_BF16><<<grid, 256, 0, st>>>(n, t->d_F, t->d_E32, t->d_part);
    else if (t->dtype == FFK_F16) k_exchange_partial<FFK_F16><<<grid, 256, 0, st>>>(n, t->d_F, t->d_E32, t->d_part);
    else k_exchange_partial<FFK_F32><<<grid, 256, 0, st>>>(n, t->d_F, t->d_E32, t->d_part);
    k_exchange_reduce<<<nb, 256, 0, st>>>(n, t->n_chunks, t->d_part, t->d_ab);
  }
  if (t->has_conduction) k_conduction<<<nb, 256, 0, st>>>(n, t->d_row, t->d_col, t->d_G, t->d_T, t->d_em, t->d_ab);
  if (t->has_sources) k_sources<<<nb, 256, 0, st>>>(n, t->d_power, t->d_em, t->d_ab);
  k_adams_dev<<<nb, 256, 0, st>>>(n, t->d_step, t->dt, t->d_C, t->d_em, t->d_ab, t->d_E, t->d_T, t->d_Tn, t->d_bal,
                                  t->d_step + 1, t->bal_cap);
  k_advance<<<nb, 256, 0, st>>>(n, t->d_T, t->d_Tn, t->d_step);
  FFK_HIP(hipGetLastError());
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

int ffk_sbff(int n, const void* d_F, int dtype, const double* d_T, const double* d_area,
             const double* d_eps_front, const double* d_eps_rear, double* d_emission, double* d_absorption,
             void* stream) {
  FFK_REQUIRE(n > 0 && d_F && d_T && d_area && d_eps_front && d_eps_rear && d_emission && d_absorption,
              FFK_ERR_ARG, "ffk_sbff: bad args");
  FFK_REQUIRE(dtype == FFK_F32 || dtype == FFK_BF16 || dtype == FFK_F16, FFK_ERR_ARG, "dtype %d", dtype);
  hipStream_t st = as_stream(stream);
  // Scratch for E (FP32) and the per-chunk partial sums, kept per thread and
  // grown on demand.
  static thread_local float* d_scratch = nullptr;
  static thread_local size_t scratch_n = 0;
  const int n_chunks = (n + kRowsPerChunk - 1) / kRowsPerChunk;
  const size_t need = (size_t)n + (size_t)n_chunks * n;
  if (need > scratch_n) {
    FFK_HIP(hipStreamSynchronize(st));
    dfree(d_scratch);
    d_scratch = nullptr;
    scratch_n = 0;
    int r;
    if ((r = dalloc(&d_scratch, need)) < 0) return r;
    scratch_n = need;
  }
  float* E32 = d_scratch;
  float* part = d_scratch + n;
  k_sb_emission<<<(n + 255) / 256, 256, 0, st>>>(n, d_T, d_area, d_eps_front, d_eps_rear, d_emission, E32);
  const int cols_per_block = 256 * kColsPerThread;
  dim3 grid((unsigned)((n + cols_per_block - 1) / cols_per_block), (unsigned)n_chunks);
  if (dtype == FFK_BF16) k_exchange_partial<FFK_BF16><<<grid, 256, 0, st>>>(n, d_F, E32, part);
  else if (dtype == FFK_F16) k_exchange_partial<FFK_F16><<<grid, 256, 0, st>>>(n, d_F, E32, part);
  else k_exchange_partial<FFK_F32><<<grid, 256, 0, st>>>(n, d_F, E32, part);
  k_exchange_reduce<<<(n + 255) / 256, 256, 0, st>>>(n, n_chunks, part, d_absorption);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
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
  k_adams<<<(n + 255) / 256, 256, 0, as_stream(stream)>>>(n, n_step, dt, d_C, d_emission, d_absorption,
                                                         d_E_hist, d_T_n, d_T_next, d_balance);
  FFK_HIP(hipGetLastError());
  return FFK_OK;
}

}  // extern "C"

extern "C" {

int ffk_thermal_create(ffk_thermal** out, int n, const double* h_C, const double* h_area,
                       const double* h_eps_front, const double* h_eps_rear) {
  FFK_REQUIRE(out && n > 0 && h_C && h_area && h_eps_front && h_eps_rear, FFK_ERR_ARG,
              "ffk_thermal_create: bad args");
  ffk_thermal* t = new ffk_thermal();
  t->n = n;
  t->n_chunks = (n + kRowsPerChunk - 1) / kRowsPerChunk;
  int r;
  if ((r = dalloc(&t->d_C, n)) < 0 || (r = dalloc(&t->d_area, n)) < 0 || (r = dalloc(&t->d_ef, n)) < 0 ||
      (r = dalloc(&t->d_er, n)) < 0 || (r = dalloc(&t->d_T, n)) < 0 || (r = dalloc(&t->d_Tn, n)) < 0 ||
      (r = dalloc(&t->d_em, n)) < 0 || (r = dalloc(&t->d_ab, n)) < 0 || (r = dalloc(&t->d_E, 5 * (size_t)n)) < 0 ||
      (r = dalloc(&t->d_power, n)) < 0 || (r = dalloc(&t->d_E32, n)) < 0 ||
      (r = dalloc(&t->d_part, (size_t)t->n_chunks * n)) < 0 || (r = dalloc(&t->d_step, 2)) < 0) {
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
  dfree(t->d_T);
  dfree(t->d_Tn);
  dfree(t->d_em);
  dfree(t->d_ab);
  dfree(t->d_E);
  dfree(t->d_power);
  dfree(t->d_E32);
  dfree(t->d_part);
  dfree(t->d_step);
  dfree(t->d_row);
  dfree(t->d_col);
  dfree(t->d_G);
  dfree(t->d_bal);
  delete t;
}

int ffk_thermal_set_exchange(ffk_thermal* t, const void* d_F, int dtype) {
  FFK_REQUIRE(t, FFK_ERR_ARG, "thermal is NULL");
  FFK_REQUIRE(dtype == FFK_F32 || dtype == FFK_BF16 || dtype == FFK_F16, FFK_ERR_ARG, "dtype %d", dtype);
  t->d_F = d_F;
  t->dtype = dtype;
  thermal_free_graph(t);
  return FFK_OK;
}

int ffk_thermal_set_conduction(ffk_thermal* t, const int* h_row_ptr, const int* h_cols, const double* h_G) {
  FFK_REQUIRE(t && h_row_ptr, FFK_ERR_ARG, "ffk_thermal_set_conduction: bad args");
  const int nnz = h_row_ptr[t->n];
  FFK_REQUIRE(nnz >= 0 && (nnz == 0 || (h_cols && h_G)), FFK_ERR_ARG, "ffk_thermal_set_conduction: bad CSR");
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
  if (h_T0) {
    std::vector<double> zero(5 * (size_t)n, 0.0);
    FFK_HIP(hipMemcpyAsync(t->d_T, h_T0, 8 * (size_t)n, hipMemcpyHostToDevice, st));
    FFK_HIP(hipMemcpyAsync(t->d_E, zero.data(), 40 * (size_t)n, hipMemcpyHostToDevice, st));
    int z = 0;
    FFK_HIP(hipMemcpyAsync(t->d_step, &z, 4, hipMemcpyHostToDevice, st));
    FFK_HIP(hipStreamSynchronize(st));
  }
  // balance records for this run
  if (h_net_stored && n_steps > t->bal_cap) {
    thermal_free_graph(t);
    dfree(t->d_bal);
    t->d_bal = nullptr;
    t->bal_cap = 0;
    int r;
    if ((r = dalloc(&t->d_bal, 2 * (size_t)n_steps)) < 0) return r;
    t->bal_cap = n_steps;
  }
  if (t->bal_cap > 0) {
    k_fill_f64<<<(unsigned)((2LL * t->bal_cap + 255) / 256), 256, 0, st>>>(t->d_bal, 2LL * t->bal_cap);
  }
  int base = 0;
  FFK_HIP(hipMemcpyAsync(&base, t->d_step, 4, hipMemcpyDeviceToHost, st));
  FFK_HIP(hipStreamSynchronize(st));
  FFK_HIP(hipMemcpyAsync(t->d_step + 1, &base, 4, hipMemcpyHostToDevice, st));
  if (dt != t->dt) thermal_free_graph(t);
  t->dt = dt;
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
  if (h_T_out) FFK_HIP(hipMemcpyAsync(h_T_out, t->d_T, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
  if (h_net_stored && n_steps > 0)
    FFK_HIP(hipMemcpyAsync(h_net_stored, t->d_bal, 16 * (size_t)n_steps, hipMemcpyDeviceToHost, st));
  FFK_HIP(hipStreamSynchronize(st));
  return FFK_OK;
}

}  // extern "C"
