// Internal definitions shared by the ffk translation units (gfx950 only).
//
// Every file of the library is compiled with -ffp-contract=off and IEEE
// division/sqrt, so the float expressions below round exactly like the
// reference's SSE code (math/triangle.cpp, math/mat.cpp, emission/*.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "ffk.h"

namespace ffk {

// ---------------------------------------------------------------- errors --
void set_error(const char* fmt, ...);

#define FFK_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::ffk::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                       \
      return FFK_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define FFK_REQUIRE(cond, code, ...) \
  do {                               \
    if (!(cond)) {                   \
      ::ffk::set_error(__VA_ARGS__); \
      return (code);                 \
    }                                \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Device allocation that records failures as FFK_ERR_NOMEM.
int dmalloc(void** p, size_t bytes);
template <class T>
int dalloc(T** p, size_t n) {
  return dmalloc(reinterpret_cast<void**>(p), n * sizeof(T));
}
void dfree(void* p);

// ------------------------------------------------------------ constants --
// (double)x < 1e-8  <=>  x <= 9.99999993922529e-09f for every float x: the
// reference compares float values against the double literal _FLT_EPSILON
// (math/operations.h:31, math/triangle.cpp:38,45).
constexpr float kEps8 = 9.99999993922529e-09f;

// ---------------------------------------------------------- vector math --
struct v3 {
  float x, y, z;
};
__host__ __device__ inline v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__host__ __device__ inline v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Per-face intersection terms (64 B, four float4):
//   q0 = {p0, uu}  q1 = {u, uv}  q2 = {v, vv}  q3 = {n, 1/D}
// u = p1-p0, v = p2-p0, n = u x v, uu/uv/vv = dots, D = uv*uv - uu*vv, all
// with the reference's expressions (math/triangle.cpp:35-38,57-63), so the
// same bits; D is recomputed from uu, uv, vv where the test needs it. A
// degenerate face ((double)dot(n,n) < 1e-8, never hit: math/triangle.cpp:38)
// stores n = 0, which every ray meets as parallel (|dot(n, dir)| = 0), so it
// is never hit either; such faces are never put in the grid.
struct tri_terms {
  float4 q0, q1, q2, q3;
};

__host__ __device__ inline tri_terms make_terms(const float* f) {
  v3 p0 = mk(f[0], f[1], f[2]), p1 = mk(f[3], f[4], f[5]), p2 = mk(f[6], f[7], f[8]);
  v3 u = p1 - p0, v = p2 - p0, n = cross(u, v);
  const bool deg = !(dot(n, n) > kEps8);
  float uu = dot(u, u), uv = dot(u, v), vv = dot(v, v);
  float D = uv * uv - uu * vv;
  tri_terms t;
  t.q0 = make_float4(p0.x, p0.y, p0.z, uu);
  t.q1 = make_float4(u.x, u.y, u.z, uv);
  t.q2 = make_float4(v.x, v.y, v.z, vv);
  t.q3 = deg ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(n.x, n.y, n.z, 1.0f / D);
  return t;
}

// The reference test (math/triangle.cpp:28-79) on precomputed terms, for a
// ray with origin o and direction dir = end - origin. Returns true for a
// unique hit and writes the point and its squared distance to o.
//
// The inside test needs s = sn / D and t = tn / D only for their comparisons
// with 0 and 1. s' = sn * (1/D) and t' = tn * (1/D) are within 2.4e-7
// (relative) of the exact quotients, so a face with s' or t' beyond the
// rejection margins below fails the exact test, a face with s', t', s' + t'
// inside by 1e-5 passes it, and only the faces in between take the two
// correctly rounded divisions: the result is the reference's bit for bit,
// and a wavefront runs the divisions only when one of its lanes is that close
// to an edge.
__host__ __device__ inline bool hit_terms(const float4 q0, const float4 q1, const float4 q2,
                                          const float4 q3, v3 o, v3 dir, v3* point, float* d2) {
  const v3 p0 = mk(q0.x, q0.y, q0.z), u = mk(q1.x, q1.y, q1.z), v = mk(q2.x, q2.y, q2.z);
  const v3 n = mk(q3.x, q3.y, q3.z);
  const v3 w0 = o - p0;
  const float a = -dot(n, w0);
  const float b = dot(n, dir);
  if (fabsf(b) <= kEps8) return false;
  const float r = a / b;
  if (r < 0.0f) return false;
  const v3 p = o + dir * r;
  const v3 w = p - p0;
  const float uu = q0.w, uv = q1.w, vv = q2.w;
  const float wu = dot(w, u), wv = dot(w, v);
  const float sn = uv * wv - vv * wu;
  const float tn = uv * wu - uu * wv;
  const float s1 = sn * q3.w, t1 = tn * q3.w;
  if (s1 < -1e-30f || s1 > 1.000001f || t1 < -1e-30f || s1 + t1 > 1.00001f) return false;
  if (!(s1 > 1e-5f && s1 < 0.99999f && t1 > 1e-5f && s1 + t1 < 0.99999f)) {
    const float D = uv * uv - uu * vv;
    const float s = sn / D;
    if (s < 0.0f || s > 1.0f) return false;
    const float t = tn / D;
    if (t < 0.0f || (s + t) > 1.0f) return false;
  }
  const v3 sd = p - o;
  *point = p;
  *d2 = dot(sd, sd);
  return true;
}

// rotate_towards (math/mat.cpp:25-49) with the element placement of the
// make_mat33-based helpers in math/mat.h:64-97; m[row*3+col] = p[row][col].
__host__ __device__ inline void rotate_towards(v3 s, v3 to, float* m) {
  v3 v = cross(s, to);
  float s2 = dot(v, v);
  float c = dot(s, to);
  float r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (s2 > 1.17549435e-38f /* FLT_MIN */) {
    const float a[9] = {0, v.z, -v.y, -v.z, 0, v.x, v.y, -v.x, 0};
#define A(i, j) a[(i)*3 + (j)]
    float q[9] = {A(0, 1) * A(1, 0) + A(0, 2) * A(2, 0), A(1, 2) * A(2, 0), A(2, 1) * A(1, 0),
                  A(0, 1) * A(1, 1) + A(0, 2) * A(2, 1), A(1, 0) * A(0, 1) + A(1, 2) * A(2, 1),
                  A(2, 0) * A(0, 1), A(0, 1) * A(1, 2), A(1, 0) * A(0, 2),
                  A(2, 0) * A(0, 2) + A(2, 1) * A(1, 2)};
#undef A
    float k = (1 - c) / s2;
    float qs[9];  // (q * k) stored transposed, as operator*(mat33, float) builds it
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) qs[i * 3 + j] = q[j * 3 + i] * k;
    float s1[9];  // (I + a) transposed, as operator+ builds it
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) s1[i * 3 + j] = r[j * 3 + i] + a[j * 3 + i];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r[i * 3 + j] = s1[j * 3 + i] + qs[j * 3 + i];
  } else {
    v3 f = s + to;
    float l = dot(f, f);
    if (l < 1.1920929e-07f /* FLT_EPSILON */) {
      r[4] = -1;
      r[8] = -1;
    }
  }
  for (int i = 0; i < 9; ++i) m[i] = r[i];
}

// Loads through the global address space. The scene tables are reached
// through a descriptor read from memory, so the compiler cannot prove their
// pointers global and emits flat instructions, which count against both the
// vector-memory and the LDS wait counters: every table load then also waits
// for the walk's cross-lane LDS shuffles. The casts make them global loads.
#define FFK_GMEM __attribute__((address_space(1)))
#ifdef __HIP_DEVICE_COMPILE__
template <class T>
__device__ inline T gld(const T* p) {
  return *(const FFK_GMEM T*)p;
}
__device__ inline int2 gld2(const int2* p) {
  typedef int i2v __attribute__((ext_vector_type(2)));
  const i2v v = *(const FFK_GMEM i2v*)p;
  return make_int2(v.x, v.y);
}
__device__ inline float4 gld4(const float4* p) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v v = *(const FFK_GMEM f4v*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
// Element i of a table as a 32-bit unsigned byte offset from its base: with
// a wave-uniform base the load takes the scalar-base form (no 64-bit address
// arithmetic per lane). The tables here stay far below 4 GiB.
template <class T>
__device__ inline const T* gidx(const T* base, uint32_t i) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)(i * (uint32_t)sizeof(T)));
}
#else
template <class T>
__host__ __device__ inline T gld(const T* p) {
  return *p;
}
__host__ __device__ inline int2 gld2(const int2* p) { return *p; }
__host__ __device__ inline float4 gld4(const float4* p) { return *p; }
template <class T>
__host__ __device__ inline const T* gidx(const T* base, uint32_t i) {
  return base + i;
}
#endif

// mat33 * vec3 with the reference's column access (math/mat.h:30-37).
__host__ __device__ inline v3 mat_vec(const float* m, v3 v) {
  return mk((m[0] * v.x) + (m[3] * v.y) + (m[6] * v.z), (m[1] * v.x) + (m[4] * v.y) + (m[7] * v.z),
            (m[2] * v.x) + (m[5] * v.y) + (m[8] * v.z));
}

// --------------------------------------------------- counter-based RNG --
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), written from the paper.
struct u4 {
  uint32_t x, y, z, w;
};
__host__ __device__ inline u4 philox(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    u4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
__host__ __device__ inline float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }

// sin and cos of 2*pi*u, u in [0,1): exact octant reduction, then Taylor
// polynomials on [0, pi/4] in Horner form (truncation error < 2e-9).
__host__ __device__ inline void sincos_turns(float u, float* s_out, float* c_out) {
  float u8 = u * 8.0f;
  int o = (int)u8;
  float f = u8 - (float)o;
  if (o & 1) f = 1.0f - f;
  float x = f * 0.785398163397448f;
  float x2 = x * x;
  float sp = 1.0f + x2 * (-1.66666667e-1f +
                          x2 * (8.33333333e-3f + x2 * (-1.98412698e-4f + x2 * 2.75573192e-6f)));
  float s = x * sp;
  float c = 1.0f + x2 * (-0.5f + x2 * (4.16666667e-2f +
                                       x2 * (-1.38888889e-3f +
                                             x2 * (2.48015873e-5f + x2 * -2.75573192e-7f))));
  int k = ((o + 1) >> 1) & 3;
  float bs = (o & 1) ? -s : s;
  float bc = c;
  float rs, rc;
  if (k == 0) {
    rs = bs; rc = bc;
  } else if (k == 1) {
    rs = bc; rc = -bs;
  } else if (k == 2) {
    rs = -bs; rc = -bc;
  } else {
    rs = -bc; rc = bs;
  }
  *s_out = rs;
  *c_out = rc;
}

// ---------------------------------------------------------- scene state --
// Grid over the scene AABB: cell (ix,iy,iz) -> c = ix + nx*(iy + ny*iz);
// refs[cell_start[c] .. cell_start[c+1]) are face indices in ascending order.
struct grid_desc {
  int n[3];
  float lo[3];
  float hi[3];
  float h[3];      // cell size
  float inv_h[3];  // 1 / cell size
};

// Everything a kernel needs to emit, cast and tally one scene (POD; lives in
// device memory for the batched kernel).
struct scene_desc {
  const tri_terms* tri;
  const float* rot;            // 9 per face, emission frame
  const long long* face_off;   // n_faces+1: first ray of each face
  const int* front;            // front rays per face
  const int* face2mesh;        // -1 for faces outside every mesh
  const long long* mesh_off;   // n_meshes+1: mesh-order tally ranges
  const int* cell_start;
  const int* refs;             // face of each reference, in cell order
  const int2* cell_info;       // (first ref, count) or, for an empty cell, (boxes, 0):
                               // nibble o of boxes = size of the empty box toward octant o
  grid_desc g;
  int n_faces;
  int n_meshes;
  long long n_rays;
};

// One unit of fused work: rays [g0, g1) tallied for mesh `mesh` of scene
// `scene`. `direct` = the tile covers the whole mesh range, so the row of F
// is written from the tile's histogram without the global tally. The tile's
// rays are emitted by faces f_lo..f_hi (face order), which bounds the face
// lookup of each ray.
struct tile {
  long long g0, g1;
  int mesh;
  int scene;
  int direct;
  int f_lo, f_hi;  // faces of the first and the last ray
  int j0;          // the first ray's index within face f_lo, or -1 when it is too large for int
};

}  // namespace ffk

struct ffk_scene {
  int n_faces = 0, n_meshes = 0;
  long long n_rays = 0;
  int frame = FFK_FRAME_REFERENCE;
  bool has_faces = false, has_meshes = false, has_emission = false, has_grid = false;
  // Bumped whenever a device buffer named by desc() is freed or the emission
  // layout changes (grid rebuilt, faces/meshes/emission set again): a batch
  // records it at creation and refuses to run on a scene that has moved on.
  unsigned long long generation = 0;
  // device buffers
  float* d_faces = nullptr;
  ffk::tri_terms* d_tri = nullptr;
  float* d_rot = nullptr;
  int* d_face2mesh = nullptr;
  long long* d_face_off = nullptr;
  int* d_front = nullptr;
  long long* d_mesh_off = nullptr;
  int* d_cell_start = nullptr;
  int* d_refs = nullptr;
  int2* d_cell_info = nullptr;
  long long n_refs = 0;
  ffk::grid_desc grid = {};
  // fused-kernel work list
  ffk::tile* d_tiles = nullptr;
  int n_tiles = 0;
  int n_direct_rows = 0;
  int* d_indirect_rows = nullptr;  // meshes tallied through the global tally
  int n_indirect_rows = 0;
  ffk::scene_desc* d_desc = nullptr;
  // per-launch arguments of the fused kernel (seed, tally and F pointers),
  // one block per scene so that launches of different scenes on different
  // streams do not share it
  unsigned long long* d_misc = nullptr;
  // radiance work (ffk_radiance.hip): one device block and one pinned host
  // block, laid out for the current faces and meshes on the first pass and
  // released when either is set again (rad_W < 0: not built)
  int rad_W = -1;         // walk entries: the faces of each mesh, mesh by mesh
  int rad_parts = 0;      // blocks of the per-entry reductions
  size_t rad_scan_bytes = 0;
  void* d_rad_block = nullptr;
  void* h_rad_pin = nullptr;
  hipEvent_t rad_pin_ev = nullptr;  // the last transfer through h_rad_pin (T / eps up, status down)
  int* d_walk_face = nullptr;               // W: face of each walk entry
  int* d_walk_mesh = nullptr;               // W: mesh of each walk entry
  int* d_mesh_walk = nullptr;               // n_meshes + 1: first walk entry of each mesh
  long long* d_walk_cnt = nullptr;          // W + 1: rays of each entry's face (then 0)
  long long* d_walk_start = nullptr;        // W + 1: first walk position of each entry
  unsigned long long* d_walk_q = nullptr;   // 2W: fixed-point front / rear ray power
  unsigned char* d_walk_band = nullptr;     // 2W: power band of each side
  unsigned long long* d_rad_acc = nullptr;  // kPowerBands x n_meshes fixed-point sums
  double* d_rad_abs = nullptr;              // n_meshes: absorbed power
  double* d_rad_scale = nullptr;            // kPowerBands fixed-point scales
  float* d_rad_p = nullptr;                 // 2W: ray power per walk entry side
  float* d_rad_pface = nullptr;             // 2 n_faces: caller-given power per face side
  float* d_rad_te = nullptr;                // T (n_meshes), eps (2 n_meshes)
  float* d_rad_area = nullptr;              // n_faces: h_area
  unsigned* d_rad_part_u = nullptr;         // reduction partials (see ffk_radiance.hip)
  double* d_rad_part_d = nullptr;
  int* d_rad_info = nullptr;                // e_max, n_bands, status, first bad face
  void* d_rad_scan = nullptr;               // hipCUB scan storage
  // layout upload staging (pinned): face offsets, front counts, mesh offsets
  void* h_layout_pin = nullptr;
  size_t layout_pin_bytes = 0;
  hipEvent_t layout_pin_ev = nullptr;  // the last upload from h_layout_pin
  // host copies
  std::vector<float> h_faces;
  std::vector<int> h_meshes, h_front, h_rear;
  std::vector<long long> h_face_off, h_mesh_off;
  // tuning (0 = automatic)
  int tune_block = 0, tune_rays_per_tile = 0, tune_hist = -1;
  int tune_refill = 32;  // idle lanes that trigger a refill in the fused kernel (0 = off)
  int tune_pool = 2;     // pooled tests once waiting lanes x tune_pool >= active lanes (0 = off)
  int tune_walk = FFK_WALK_DEFAULT;  // walk mode: bit 0 = always the jump code for empty cells, bit 1 = whole chunks of pooled tests
  int tiles_rays_per_tile = 0;  // value the tile list was built with
  bool tiles_dirty = false;     // the emission layout changed since the tile list was built
  int rot_frame = -1;           // frame of d_rot (-1: not computed)
  // face areas in float (math::triangle_area), computed once with the faces
  std::vector<float> h_area;
  // host work arrays of the radiance step
  std::vector<float> rad_w, rad_p;
  std::vector<int> rad_front, rad_rear;

  ffk::scene_desc desc() const;
};

namespace ffk {
// Releases the radiance work buffers (ffk_radiance.hip); they are laid out
// again for the current faces and meshes on the next radiance pass.
void free_radiance(ffk_scene* s);
int set_emission(ffk_scene* s, const int* h_front, const int* h_rear, int frame, long long* n_rays_out,
                 hipStream_t st, bool one_stream);
int build_tiles(ffk_scene* s, int rays_per_tile, hipStream_t st);
int ensure_tiles(ffk_scene* s, hipStream_t st);
int device_cu_count();
}  // namespace ffk
