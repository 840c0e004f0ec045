// Scene state on the device: per-face intersection terms, face->mesh map,
// emission layout, fused-work tiles and the uniform grid built on the GPU.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "ffk_internal.hpp"

namespace ffk {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int dmalloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return FFK_ERR_NOMEM;
  }
  return FFK_OK;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

int device_cu_count() {
  static int cached = 0;
  if (!cached) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cached = prop.multiProcessorCount;
    else
      cached = 256;
  }
  return cached;
}

// ------------------------------------------------------------- kernels --

__global__ void k_face_terms(const float* __restrict__ faces, int n, tri_terms* __restrict__ out) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  out[f] = make_terms(faces + 9 * (size_t)f);
}

// Emission frame per face: rotate_towards(z, n) with the unnormalised normal
// (emission/malley_cpu.cpp:125-129) or with n / |n|.
__global__ void k_face_rotation(const float* __restrict__ faces, int n, int frame,
                                float* __restrict__ rot) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const float* p = faces + 9 * (size_t)f;
  v3 p0 = mk(p[0], p[1], p[2]), p1 = mk(p[3], p[4], p[5]), p2 = mk(p[6], p[7], p[8]);
  v3 nn = cross(p1 - p0, p2 - p0);
  if (frame == FFK_FRAME_ORTHONORMAL) {
    float len = sqrtf(dot(nn, nn));
    nn = mk(nn.x / len, nn.y / len, nn.z / len);
  }
  rotate_towards(mk(0.f, 0.f, 1.f), nn, rot + 9 * (size_t)f);
}

// Order-preserving float <-> uint map for atomic min/max of floats.
__device__ inline uint32_t ord(float f) {
  uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ inline float unord(uint32_t u) {
  uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  memcpy(&f, &b, 4);
  return f;
}

__global__ void k_bounds(const float* __restrict__ faces, int n, uint32_t* __restrict__ box) {
  __shared__ uint32_t smin[3], smax[3];
  if (threadIdx.x < 3) {
    smin[threadIdx.x] = 0xffffffffu;
    smax[threadIdx.x] = 0u;
  }
  __syncthreads();
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n) {
    const float* p = faces + 9 * (size_t)f;
    for (int a = 0; a < 3; ++a) {
      float lo = fminf(fminf(p[a], p[3 + a]), p[6 + a]);
      float hi = fmaxf(fmaxf(p[a], p[3 + a]), p[6 + a]);
      atomicMin(&smin[a], ord(lo));
      atomicMax(&smax[a], ord(hi));
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    atomicMin(&box[threadIdx.x], smin[threadIdx.x]);
    atomicMax(&box[3 + threadIdx.x], smax[threadIdx.x]);
  }
}

// Conservative cell range of a face: its AABB grown by `pad` on every side.
__device__ inline void cell_range(const float* p, const grid_desc& g, float pad, int* c0, int* c1) {
  for (int a = 0; a < 3; ++a) {
    float lo = fminf(fminf(p[a], p[3 + a]), p[6 + a]) - pad;
    float hi = fmaxf(fmaxf(p[a], p[3 + a]), p[6 + a]) + pad;
    int i0 = (int)floorf((lo - g.lo[a]) * g.inv_h[a]);
    int i1 = (int)floorf((hi - g.lo[a]) * g.inv_h[a]);
    c0[a] = min(max(i0, 0), g.n[a] - 1);
    c1[a] = min(max(i1, 0), g.n[a] - 1);
  }
}

// Separating-axis test of a triangle against an axis-aligned box (centre c,
// half size h), in double: the 3 box axes, the triangle normal and the 9
// edge x axis products (Akenine-Moller, "Fast 3D triangle-box overlap
// testing", JGT 2001). Returns true when no axis separates them.
__device__ inline bool tri_box_overlap(const float* p, const double c[3], const double h[3]) {
  double v[3][3];
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) v[k][a] = (double)p[3 * k + a] - c[a];
  for (int a = 0; a < 3; ++a) {  // box axes
    double lo = fmin(fmin(v[0][a], v[1][a]), v[2][a]), hi = fmax(fmax(v[0][a], v[1][a]), v[2][a]);
    if (lo > h[a] || hi < -h[a]) return false;
  }
  double e[3][3];
  for (int a = 0; a < 3; ++a) {
    e[0][a] = v[1][a] - v[0][a];
    e[1][a] = v[2][a] - v[1][a];
    e[2][a] = v[0][a] - v[2][a];
  }
  double n[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                 e[0][0] * e[1][1] - e[0][1] * e[1][0]};
  {  // triangle plane
    double d = n[0] * v[0][0] + n[1] * v[0][1] + n[2] * v[0][2];
    double r = h[0] * fabs(n[0]) + h[1] * fabs(n[1]) + h[2] * fabs(n[2]);
    if (fabs(d) > r) return false;
  }
  for (int i = 0; i < 3; ++i)      // edge
    for (int a = 0; a < 3; ++a) {  // box axis
      // axis = e_i x unit_a
      double ax[3] = {0, 0, 0};
      const int b = (a + 1) % 3, c2 = (a + 2) % 3;
      ax[b] = e[i][c2];
      ax[c2] = -e[i][b];
      double p0 = ax[0] * v[0][0] + ax[1] * v[0][1] + ax[2] * v[0][2];
      double p1 = ax[0] * v[1][0] + ax[1] * v[1][1] + ax[2] * v[1][2];
      double p2 = ax[0] * v[2][0] + ax[1] * v[2][1] + ax[2] * v[2][2];
      double lo = fmin(fmin(p0, p1), p2), hi = fmax(fmax(p0, p1), p2);
      double r = h[0] * fabs(ax[0]) + h[1] * fabs(ax[1]) + h[2] * fabs(ax[2]);
      if (lo > r || hi < -r) return false;
    }
  return true;
}

// Degenerate faces ((double)dot(n,n) < 1e-8, math/triangle.cpp:38) are never
// hit, so they are not referenced.
__device__ inline bool degenerate(const float* p) {
  v3 p0 = mk(p[0], p[1], p[2]), p1 = mk(p[3], p[4], p[5]), p2 = mk(p[6], p[7], p[8]);
  v3 n = cross(p1 - p0, p2 - p0);
  return !(dot(n, n) > kEps8);
}

// Visit every cell whose box grown by `pad` the face overlaps.
template <class F>
__device__ inline void for_face_cells(const float* p, const grid_desc& g, float pad, F fn) {
  if (degenerate(p)) return;
  int c0[3], c1[3];
  cell_range(p, g, pad, c0, c1);
  const double h[3] = {0.5 * g.h[0] + pad, 0.5 * g.h[1] + pad, 0.5 * g.h[2] + pad};
  for (int z = c0[2]; z <= c1[2]; ++z)
    for (int y = c0[1]; y <= c1[1]; ++y)
      for (int x = c0[0]; x <= c1[0]; ++x) {
        const double c[3] = {g.lo[0] + (x + 0.5) * (double)g.h[0], g.lo[1] + (y + 0.5) * (double)g.h[1],
                             g.lo[2] + (z + 0.5) * (double)g.h[2]};
        if (tri_box_overlap(p, c, h)) fn((uint32_t)(x + g.n[0] * (y + g.n[1] * z)));
      }
}

__global__ void k_ref_count(const float* __restrict__ faces, int n, grid_desc g, float pad,
                            uint32_t* __restrict__ cnt) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  uint32_t k = 0;
  for_face_cells(faces + 9 * (size_t)f, g, pad, [&](uint32_t) { ++k; });
  cnt[f] = k;
}

__global__ void k_ref_emit(const float* __restrict__ faces, int n, grid_desc g, float pad,
                           const uint32_t* __restrict__ off, uint32_t* __restrict__ keys,
                           int* __restrict__ vals) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  uint32_t o = off[f];
  const uint32_t end = off[f + 1];
  for_face_cells(faces + 9 * (size_t)f, g, pad, [&](uint32_t cell) {
    if (o < end) {  // the count pass visited the same cells
      keys[o] = cell;
      vals[o] = f;
      ++o;
    }
  });
}

__global__ void k_cell_count(const uint32_t* __restrict__ keys, long long n, int* __restrict__ cnt) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(&cnt[keys[i]], 1);
}

// Empty boxes per ray octant. For octant o (bit 0: x decreasing, bit 1: y,
// bit 2: z) nibble o of a cell's word is the largest k <= 15 such that the k
// cells per axis from this cell toward that octant, [c, c + s*(k-1)] clipped
// to the grid, are all empty; 0 for a non-empty cell. A ray whose direction
// lies in octant o leaves such a box only through its three far faces, so it
// crosses no occupied cell before leaving it. Cells outside the grid count as
// empty. Relaxed from 15 down: k = 1 + min over the 7 neighbours toward the
// octant, exact after 16 passes (the values are capped at 15).
__global__ void k_empty_init(const int* __restrict__ cs, long long nc, uint32_t* __restrict__ d) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nc) d[c] = cs[c + 1] > cs[c] ? 0u : 0xffffffffu;
}

__global__ void k_empty_relax(grid_desc g, const uint32_t* __restrict__ in, uint32_t* __restrict__ out) {
  const long long nc = (long long)g.n[0] * g.n[1] * g.n[2];
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint32_t self = in[c];
  if (self == 0u) {  // non-empty (an empty cell has k >= 1 in every octant)
    out[c] = 0u;
    return;
  }
  const int x = (int)(c % g.n[0]);
  const int y = (int)((c / g.n[0]) % g.n[1]);
  const int z = (int)(c / ((long long)g.n[0] * g.n[1]));
  uint32_t w[27];  // the 3x3x3 neighbourhood, index (dz+1)*9 + (dy+1)*3 + (dx+1)
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int X = x + dx, Y = y + dy, Z = z + dz;
        const bool in_grid = X >= 0 && Y >= 0 && Z >= 0 && X < g.n[0] && Y < g.n[1] && Z < g.n[2];
        w[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] =
            in_grid ? in[X + (long long)g.n[0] * (Y + (long long)g.n[1] * Z)] : 0xffffffffu;
      }
  uint32_t res = 0u;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    // index of the neighbour toward the octant along each axis (1 = this cell)
    const int fx = (o & 1) ? 0 : 2, fy = (o & 2) ? 0 : 2, fz = (o & 4) ? 0 : 2;
    uint32_t m = 15u;
#pragma unroll
    for (int e = 1; e < 8; ++e) {
      const int ix = (e & 1) ? fx : 1, iy = (e & 2) ? fy : 1, iz = (e & 4) ? fz : 1;
      m = min(m, (w[iz * 9 + iy * 3 + ix] >> (4 * o)) & 15u);
    }
    res |= min(m + 1u, 15u) << (4 * o);
  }
  out[c] = res;
}

__global__ void k_cell_info(const int* __restrict__ cs, const uint32_t* __restrict__ d, long long nc,
                            int2* __restrict__ info) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int a = cs[c], b = cs[c + 1];
  info[c] = b > a ? make_int2(a, b - a) : make_int2((int)d[c], 0);
}

}  // namespace ffk

using namespace ffk;

ffk::scene_desc ffk_scene::desc() const {
  scene_desc d;
  d.tri = d_tri;
  d.rot = d_rot;
  d.face_off = d_face_off;
  d.front = d_front;
  d.face2mesh = d_face2mesh;
  d.mesh_off = d_mesh_off;
  d.cell_start = d_cell_start;
  d.refs = d_refs;
  d.cell_info = d_cell_info;
  d.g = grid;
  d.n_faces = n_faces;
  d.n_meshes = n_meshes;
  d.n_rays = n_rays;
  return d;
}

static void free_grid(ffk_scene* s) {
  ++s->generation;
  dfree(s->d_cell_start);
  dfree(s->d_refs);
  dfree(s->d_cell_info);
  s->d_cell_start = nullptr;
  s->d_refs = nullptr;
  s->d_cell_info = nullptr;
  s->n_refs = 0;
  s->has_grid = false;
}

static void free_emission(ffk_scene* s) {
  ++s->generation;
  dfree(s->d_rot);
  dfree(s->d_face_off);
  dfree(s->d_front);
  dfree(s->d_mesh_off);
  dfree(s->d_tiles);
  dfree(s->d_indirect_rows);
  s->d_rot = nullptr;
  s->d_face_off = nullptr;
  s->d_front = nullptr;
  s->d_mesh_off = nullptr;
  s->d_tiles = nullptr;
  s->d_indirect_rows = nullptr;
  s->n_tiles = 0;
  s->tiles_dirty = false;
  s->rot_frame = -1;
  s->has_emission = false;
}

extern "C" {

const char* ffk_last_error(void) { return g_err; }

int ffk_device_info(int* n_cu, int* clock_khz, int* mem_clock_khz, int* bus_width_bits,
                    int* l2_bytes, char* name, int name_len) {
  int dev = 0;
  FFK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  FFK_HIP(hipGetDeviceProperties(&p, dev));
  if (n_cu) *n_cu = p.multiProcessorCount;
  if (clock_khz) *clock_khz = p.clockRate;
  if (mem_clock_khz) *mem_clock_khz = p.memoryClockRate;
  if (bus_width_bits) *bus_width_bits = p.memoryBusWidth;
  if (l2_bytes) *l2_bytes = p.l2CacheSize;
  if (name && name_len > 0) {
    snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  return FFK_OK;
}

int ffk_scene_create(ffk_scene** out) {
  FFK_REQUIRE(out, FFK_ERR_ARG, "ffk_scene_create: out is NULL");
  *out = new ffk_scene();
  return FFK_OK;
}

void ffk_scene_destroy(ffk_scene* s) {
  if (!s) return;
  (void)hipDeviceSynchronize();
  free_grid(s);
  free_emission(s);
  dfree(s->d_faces);
  dfree(s->d_tri);
  dfree(s->d_face2mesh);
  dfree(s->d_desc);
  dfree(s->d_misc);
  free_radiance(s);
  if (s->h_layout_pin) (void)hipHostFree(s->h_layout_pin);
  if (s->rad_pin_ev) (void)hipEventDestroy(s->rad_pin_ev);
  if (s->layout_pin_ev) (void)hipEventDestroy(s->layout_pin_ev);
  delete s;
}

int ffk_scene_set_faces(ffk_scene* s, const float* h_faces, int n_faces, void* stream) {
  FFK_REQUIRE(s, FFK_ERR_ARG, "scene is NULL");
  FFK_REQUIRE(n_faces > 0 && h_faces, FFK_ERR_EMPTY, "ffk_scene_set_faces: empty scene");
  hipStream_t st = as_stream(stream);
  FFK_HIP(hipStreamSynchronize(st));
  free_grid(s);
  free_emission(s);
  free_radiance(s);
  dfree(s->d_faces);
  dfree(s->d_tri);
  dfree(s->d_face2mesh);
  s->d_faces = nullptr;
  s->d_tri = nullptr;
  s->d_face2mesh = nullptr;
  s->has_meshes = false;
  s->n_faces = n_faces;
  s->h_faces.assign(h_faces, h_faces + 9 * (size_t)n_faces);
  // math::triangle_area (math/triangle.cpp:86-92) in float, as the radiance
  // weights and powers use it every step
  s->h_area.resize(n_faces);
  for (int f = 0; f < n_faces; ++f) {
    const float* p = h_faces + 9 * (size_t)f;
    const v3 u = mk(p[3], p[4], p[5]) - mk(p[0], p[1], p[2]);
    const v3 w = mk(p[6], p[7], p[8]) - mk(p[0], p[1], p[2]);
    const v3 n = cross(u, w);
    s->h_area[f] = sqrtf(dot(n, n)) / 2;
  }
  int r;
  if ((r = dalloc(&s->d_faces, 9 * (size_t)n_faces)) < 0) return r;
  if ((r = dalloc(&s->d_tri, (size_t)n_faces)) < 0) return r;
  FFK_HIP(hipMemcpyAsync(s->d_faces, h_faces, 36 * (size_t)n_faces, hipMemcpyHostToDevice, st));
  k_face_terms<<<(n_faces + 255) / 256, 256, 0, st>>>(s->d_faces, n_faces, s->d_tri);
  FFK_HIP(hipGetLastError());
  s->has_faces = true;
  return FFK_OK;
}

int ffk_scene_set_meshes(ffk_scene* s, const int* h_meshes, int n_meshes, void* stream) {
  FFK_REQUIRE(s && s->has_faces, FFK_ERR_ARG, "ffk_scene_set_meshes: faces not set");
  FFK_REQUIRE(n_meshes > 0 && h_meshes, FFK_ERR_EMPTY, "ffk_scene_set_meshes: no meshes");
  hipStream_t st = as_stream(stream);
  std::vector<int> f2m(s->n_faces, -1);
  for (int m = 0; m < n_meshes; ++m) {
    int first = h_meshes[3 * m], cnt = h_meshes[3 * m + 1];
    FFK_REQUIRE(first >= 0 && cnt >= 0 && (long long)first + cnt <= s->n_faces, FFK_ERR_ARG,
                "mesh %d (first %d, n %d) outside %d faces", m, first, cnt, s->n_faces);
    for (int f = 0; f < cnt; ++f) f2m[first + f] = m;
  }
  FFK_HIP(hipStreamSynchronize(st));
  free_emission(s);
  free_radiance(s);
  dfree(s->d_face2mesh);
  s->d_face2mesh = nullptr;
  int r;
  if ((r = dalloc(&s->d_face2mesh, (size_t)s->n_faces)) < 0) return r;
  FFK_HIP(hipMemcpy(s->d_face2mesh, f2m.data(), 4 * (size_t)s->n_faces, hipMemcpyHostToDevice));
  s->h_meshes.assign(h_meshes, h_meshes + 3 * (size_t)n_meshes);
  s->n_meshes = n_meshes;
  s->has_meshes = true;
  return FFK_OK;
}

}  // extern "C"

namespace ffk {
// The emission layout of ffk_scene_set_emission. one_stream: the caller
// orders every use of the scene on this stream (the radiance step does), so
// stream order replaces the two waits; the pinned staging is then reused only
// once its last upload has run.
int set_emission(ffk_scene* s, const int* h_front, const int* h_rear, int frame, long long* n_rays_out,
                 hipStream_t st, bool one_stream) {
  FFK_REQUIRE(s && s->has_faces && s->has_meshes, FFK_ERR_ARG,
              "ffk_scene_set_emission: faces and meshes must be set first");
  FFK_REQUIRE(frame == FFK_FRAME_REFERENCE || frame == FFK_FRAME_ORTHONORMAL, FFK_ERR_ARG,
              "unknown frame %d", frame);
  // The layout buffers are reused (their sizes follow the faces and meshes,
  // whose setters free them): wait for the work that may still read them.
  if (!one_stream) FFK_HIP(hipStreamSynchronize(st));
  s->has_emission = false;
  ++s->generation;  // ray counts, offsets and tiles change in place
  const int nf = s->n_faces, nm = s->n_meshes;
  s->h_front.assign(h_front, h_front + nf);
  s->h_rear.assign(h_rear, h_rear + nf);
  // Face-order layout: face f owns rays [face_off[f], face_off[f+1]), front
  // block first (emission/malley_cpu.cpp:150-186).
  s->h_face_off.assign(nf + 1, 0);
  for (int f = 0; f < nf; ++f) {
    FFK_REQUIRE(h_front[f] >= 0 && h_rear[f] >= 0, FFK_ERR_ARG, "negative ray count at face %d", f);
    s->h_face_off[f + 1] = s->h_face_off[f] + h_front[f] + h_rear[f];
  }
  const long long total = s->h_face_off[nf];
  // Mesh-order walk of the tally (cpu_system.cpp:165-191): mesh m consumes
  // the next sum of its faces' ray counts, capped at the total.
  s->h_mesh_off.assign(nm + 1, 0);
  for (int m = 0; m < nm; ++m) {
    long long c = 0;
    for (int f = 0; f < s->h_meshes[3 * m + 1]; ++f) {
      int fi = s->h_meshes[3 * m] + f;
      c += h_front[fi] + h_rear[fi];
    }
    s->h_mesh_off[m + 1] = std::min(total, s->h_mesh_off[m] + c);
  }
  s->n_rays = total;
  s->frame = frame;
  int r;
  if (!s->d_rot && (r = dalloc(&s->d_rot, 9 * (size_t)nf)) < 0) return r;
  if (!s->d_face_off && (r = dalloc(&s->d_face_off, (size_t)nf + 1)) < 0) return r;
  if (!s->d_front && (r = dalloc(&s->d_front, (size_t)nf)) < 0) return r;
  if (!s->d_mesh_off && (r = dalloc(&s->d_mesh_off, (size_t)nm + 1)) < 0) return r;
  // Uploads through pinned staging, rewritten once its last upload has run.
  const size_t b_face = 8 * ((size_t)nf + 1), b_front = 4 * (size_t)nf, b_mesh = 8 * ((size_t)nm + 1);
  if (!s->layout_pin_ev) FFK_HIP(hipEventCreateWithFlags(&s->layout_pin_ev, hipEventDisableTiming));
  FFK_HIP(hipEventSynchronize(s->layout_pin_ev));
  if (s->layout_pin_bytes < b_face + b_front + b_mesh) {
    if (s->h_layout_pin) (void)hipHostFree(s->h_layout_pin);
    s->h_layout_pin = nullptr;
    s->layout_pin_bytes = 0;
    FFK_HIP(hipHostMalloc(&s->h_layout_pin, b_face + b_front + b_mesh, hipHostMallocDefault));
    s->layout_pin_bytes = b_face + b_front + b_mesh;
  }
  char* pin = static_cast<char*>(s->h_layout_pin);
  std::memcpy(pin, s->h_face_off.data(), b_face);
  std::memcpy(pin + b_face, h_front, b_front);
  std::memcpy(pin + b_face + b_front, s->h_mesh_off.data(), b_mesh);
  FFK_HIP(hipMemcpyAsync(s->d_face_off, pin, b_face, hipMemcpyHostToDevice, st));
  FFK_HIP(hipMemcpyAsync(s->d_front, pin + b_face, b_front, hipMemcpyHostToDevice, st));
  FFK_HIP(hipMemcpyAsync(s->d_mesh_off, pin + b_face + b_front, b_mesh, hipMemcpyHostToDevice, st));
  FFK_HIP(hipEventRecord(s->layout_pin_ev, st));
  // one wait for the three copies: callers may go on with another stream
  if (!one_stream) FFK_HIP(hipStreamSynchronize(st));
  if (s->rot_frame != frame) {  // the rotations depend on the faces and the frame only
    k_face_rotation<<<(nf + 255) / 256, 256, 0, st>>>(s->d_faces, nf, frame, s->d_rot);
    FFK_HIP(hipGetLastError());
    s->rot_frame = frame;
  }
  s->has_emission = true;
  s->tiles_dirty = true;  // the fused kernel's tile list is built when a fused launch needs it
  if (n_rays_out) *n_rays_out = total;
  return FFK_OK;
}
}  // namespace ffk

extern "C" {

int ffk_scene_set_emission(ffk_scene* s, const int* h_front, const int* h_rear, int frame,
                           long long* n_rays_out, void* stream) {
  return set_emission(s, h_front, h_rear, frame, n_rays_out, as_stream(stream), false);
}

int ffk_scene_layout(const ffk_scene* s, float* h_faces, int* h_meshes, int* h_front, int* h_rear, int* frame) {
  FFK_REQUIRE(s && s->has_faces && s->has_meshes && s->has_emission, FFK_ERR_ARG,
              "ffk_scene_layout: faces, meshes and emission must be set");
  if (h_faces) std::copy(s->h_faces.begin(), s->h_faces.end(), h_faces);
  if (h_meshes) std::copy(s->h_meshes.begin(), s->h_meshes.end(), h_meshes);
  if (h_front) std::copy(s->h_front.begin(), s->h_front.end(), h_front);
  if (h_rear) std::copy(s->h_rear.begin(), s->h_rear.end(), h_rear);
  if (frame) *frame = s->frame;
  return FFK_OK;
}

int ffk_scene_counts(const ffk_scene* s, int* n_faces, int* n_meshes, long long* n_rays) {
  FFK_REQUIRE(s, FFK_ERR_ARG, "scene is NULL");
  if (n_faces) *n_faces = s->n_faces;
  if (n_meshes) *n_meshes = s->n_meshes;
  if (n_rays) *n_rays = s->n_rays;
  return FFK_OK;
}

int ffk_scene_build_grid(ffk_scene* s, int nx, int ny, int nz, void* stream) {
  FFK_REQUIRE(s && s->has_faces, FFK_ERR_ARG, "ffk_scene_build_grid: faces not set");
  FFK_REQUIRE(nx >= 0 && ny >= 0 && nz >= 0 && nx <= 1024 && ny <= 1024 && nz <= 1024, FFK_ERR_LIMIT,
              "grid dims %dx%dx%d outside [0,1024]", nx, ny, nz);
  hipStream_t st = as_stream(stream);
  const int nf = s->n_faces;
  free_grid(s);
  int r;
  // 1. bounds on the device
  uint32_t* d_box = nullptr;
  if ((r = dalloc(&d_box, 6)) < 0) return r;
  uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0, 0, 0};
  FFK_HIP(hipMemcpyAsync(d_box, init, sizeof(init), hipMemcpyHostToDevice, st));
  k_bounds<<<(nf + 255) / 256, 256, 0, st>>>(s->d_faces, nf, d_box);
  FFK_HIP(hipGetLastError());
  uint32_t box[6];
  FFK_HIP(hipMemcpyAsync(box, d_box, sizeof(box), hipMemcpyDeviceToHost, st));
  FFK_HIP(hipStreamSynchronize(st));
  dfree(d_box);
  float lo[3], hi[3], ext[3];
  float maxext = 0.f;
  for (int a = 0; a < 3; ++a) {
    lo[a] = unord(box[a]);
    hi[a] = unord(box[3 + a]);
    ext[a] = hi[a] - lo[a];
    maxext = std::max(maxext, ext[a]);
  }
  FFK_REQUIRE(std::isfinite(maxext), FFK_ERR_ARG, "non-finite scene bounds");
  if (maxext <= 0.f) maxext = 1.f;
  // Grow the box so that every vertex and every hit point computed on a face
  // lies strictly inside; flat scenes get a thickness.
  const float grow = 1e-4f * maxext;
  for (int a = 0; a < 3; ++a) {
    lo[a] -= grow;
    hi[a] += grow;
    ext[a] = hi[a] - lo[a];
  }
  int dims[3] = {nx, ny, nz};
  if (nx == 0 || ny == 0 || nz == 0) {
    // Cells as close to cubes as the box allows: about 12 per face (empty
    // cells cost little with the octant jumps, smaller cells mean fewer
    // tests), but at most 2^18 cells (2 MB of cell words, beside the face
    // terms in L2) unless that leaves fewer than 2 per face.
    double vol = (double)ext[0] * ext[1] * ext[2];
    double target = std::min(12.0 * nf, std::max(2.0 * nf, 262144.0));
    double cell = std::cbrt(vol / target);
    for (int a = 0; a < 3; ++a)
      dims[a] = (int)std::min(512.0, std::max(1.0, std::ceil(ext[a] / cell)));
  }
  grid_desc g;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = dims[a];
    g.lo[a] = lo[a];
    g.hi[a] = hi[a];
    g.h[a] = ext[a] / (float)dims[a];
    g.inv_h[a] = (float)dims[a] / ext[a];
  }
  const float pad = 1e-3f * std::min(g.h[0], std::min(g.h[1], g.h[2])) + 1e-6f * maxext;
  const long long nc = (long long)dims[0] * dims[1] * dims[2];
  FFK_REQUIRE(nc < (1LL << 31), FFK_ERR_LIMIT, "grid has %lld cells", nc);
  // 2. per-face reference counts and their exclusive scan
  uint32_t *d_cnt = nullptr, *d_off = nullptr;
  if ((r = dalloc(&d_cnt, (size_t)nf + 1)) < 0) return r;
  if ((r = dalloc(&d_off, (size_t)nf + 1)) < 0) return r;
  FFK_HIP(hipMemsetAsync(d_cnt, 0, 4 * ((size_t)nf + 1), st));
  k_ref_count<<<(nf + 255) / 256, 256, 0, st>>>(s->d_faces, nf, g, pad, d_cnt);
  FFK_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  FFK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_cnt, d_off, nf + 1, st));
  void* d_tmp = nullptr;
  if ((r = dmalloc(&d_tmp, tmp_bytes)) < 0) return r;
  FFK_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_cnt, d_off, nf + 1, st));
  uint32_t n_refs = 0;
  FFK_HIP(hipMemcpyAsync(&n_refs, d_off + nf, 4, hipMemcpyDeviceToHost, st));
  FFK_HIP(hipStreamSynchronize(st));
  dfree(d_tmp);
  // 3. (cell, face) pairs in face order, stably sorted by cell
  uint32_t *d_keys = nullptr, *d_keys2 = nullptr;
  int* d_vals = nullptr;
  if ((r = dalloc(&d_keys, n_refs)) < 0) return r;
  if ((r = dalloc(&d_keys2, n_refs)) < 0) return r;
  if ((r = dalloc(&d_vals, n_refs)) < 0) return r;
  if ((r = dalloc(&s->d_refs, n_refs)) < 0) return r;
  k_ref_emit<<<(nf + 255) / 256, 256, 0, st>>>(s->d_faces, nf, g, pad, d_off, d_keys, d_vals);
  FFK_HIP(hipGetLastError());
  int end_bit = 1;
  while ((1LL << end_bit) < nc) ++end_bit;
  tmp_bytes = 0;
  FFK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_keys, d_keys2, d_vals, s->d_refs,
                                             (int)n_refs, 0, end_bit, st));
  if ((r = dmalloc(&d_tmp, tmp_bytes)) < 0) return r;
  FFK_HIP(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_bytes, d_keys, d_keys2, d_vals, s->d_refs,
                                             (int)n_refs, 0, end_bit, st));
  // 4. cell starts = exclusive scan of per-cell counts
  int* d_ccnt = nullptr;
  if ((r = dalloc(&d_ccnt, (size_t)nc + 1)) < 0) return r;
  if ((r = dalloc(&s->d_cell_start, (size_t)nc + 1)) < 0) return r;
  FFK_HIP(hipMemsetAsync(d_ccnt, 0, 4 * ((size_t)nc + 1), st));
  if (n_refs) {
    k_cell_count<<<(unsigned)((n_refs + 255) / 256), 256, 0, st>>>(d_keys2, n_refs, d_ccnt);
    FFK_HIP(hipGetLastError());
  }
  size_t tmp2 = 0;
  FFK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, d_ccnt, s->d_cell_start, nc + 1, st));
  void* d_tmp2 = nullptr;
  if ((r = dmalloc(&d_tmp2, tmp2)) < 0) return r;
  FFK_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp2, tmp2, d_ccnt, s->d_cell_start, nc + 1, st));
  // 5. the per-cell (range | empty boxes per octant) words
  if ((r = dalloc(&s->d_cell_info, (size_t)nc)) < 0) return r;
  uint32_t *d_e0 = nullptr, *d_e1 = nullptr;
  if ((r = dalloc(&d_e0, (size_t)nc)) < 0) return r;
  if ((r = dalloc(&d_e1, (size_t)nc)) < 0) return r;
  const unsigned cblocks = (unsigned)((nc + 255) / 256);
  k_empty_init<<<cblocks, 256, 0, st>>>(s->d_cell_start, nc, d_e0);
  for (int p = 0; p < 16; ++p) {
    k_empty_relax<<<cblocks, 256, 0, st>>>(g, d_e0, d_e1);
    std::swap(d_e0, d_e1);
  }
  k_cell_info<<<cblocks, 256, 0, st>>>(s->d_cell_start, d_e0, nc, s->d_cell_info);
  FFK_HIP(hipGetLastError());
  FFK_HIP(hipStreamSynchronize(st));
  dfree(d_e0);
  dfree(d_e1);
  dfree(d_tmp);
  dfree(d_tmp2);
  dfree(d_ccnt);
  dfree(d_keys);
  dfree(d_keys2);
  dfree(d_vals);
  dfree(d_cnt);
  dfree(d_off);
  s->grid = g;
  s->n_refs = n_refs;
  s->has_grid = true;
  return FFK_OK;
}

int ffk_scene_grid_info(const ffk_scene* s, int* dims3, float* lo3, float* hi3, long long* n_refs) {
  FFK_REQUIRE(s && s->has_grid, FFK_ERR_ARG, "grid not built");
  for (int a = 0; a < 3; ++a) {
    if (dims3) dims3[a] = s->grid.n[a];
    if (lo3) lo3[a] = s->grid.lo[a];
    if (hi3) hi3[a] = s->grid.hi[a];
  }
  if (n_refs) *n_refs = s->n_refs;
  return FFK_OK;
}

int ffk_scene_grid_cells(const ffk_scene* s, int* h_cells) {
  FFK_REQUIRE(s && s->has_grid && h_cells, FFK_ERR_ARG, "ffk_scene_grid_cells: grid not built or NULL output");
  const size_t nc = (size_t)s->grid.n[0] * s->grid.n[1] * s->grid.n[2];
  FFK_HIP(hipDeviceSynchronize());
  FFK_HIP(hipMemcpy(h_cells, s->d_cell_info, 8 * nc, hipMemcpyDeviceToHost));
  return FFK_OK;
}

int ffk_area_weights(const float* h_faces, int n_faces, float* h_weights, float* total) {
  FFK_REQUIRE(h_faces && h_weights && total && n_faces >= 0, FFK_ERR_ARG, "ffk_area_weights: bad args");
  float t = 0;
  for (int i = 0; i < n_faces; ++i) {
    const float* p = h_faces + 9 * (size_t)i;
    v3 u = mk(p[3], p[4], p[5]) - mk(p[0], p[1], p[2]);
    v3 w = mk(p[6], p[7], p[8]) - mk(p[0], p[1], p[2]);
    v3 n = cross(u, w);
    float a = sqrtf(dot(n, n)) / 2;  // math/triangle.cpp:86-92
    t += 2.f * a;
    h_weights[2 * i] = a;
    h_weights[2 * i + 1] = a;
  }
  *total = t;
  return FFK_OK;
}

int ffk_emitted_counts(const float* w, int n_faces, int n_rays, float total, int* front, int* rear) {
  FFK_REQUIRE(w && front && rear && n_faces >= 0, FFK_ERR_ARG, "ffk_emitted_counts: bad args");
  for (int f = 0; f < n_faces; ++f) {
    float a = w[2 * f], b = w[2 * f + 1];
    front[f] = a > 0 ? std::max<int>(1, (int)(n_rays * (a / total))) : 0;
    rear[f] = b > 0 ? std::max<int>(1, (int)(n_rays * (b / total))) : 0;
  }
  return FFK_OK;
}

}  // extern "C"
