/*
 * CPU oracle for the Monte Carlo form-factor path -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement, written for this project, of the reference's CPU algorithm.
 * Each function names the reference lines it follows. The operation order
 * matches the reference's expressions so that, compiled with
 * -ffp-contract=off, results are bit-identical to the compiled reference.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this library; the product library never does.
 */
#include "ff_oracle.h"

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

namespace {

struct v3 { float x, y, z; };
struct tri { v3 p[3]; };
struct ray { v3 o, d; };        // d is the END POINT, as in math/types.h:72-76
struct m33 { float p[3][3]; };

inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline v3 mul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline v3 mulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline v3 divv(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
inline v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
// math/operations.h:44-47: left-to-right sum of products
inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// math/operations.h:50-53
inline v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline v3 vabs(v3 a) { return mk(a.x < 0 ? -a.x : a.x, a.y < 0 ? -a.y : a.y, a.z < 0 ? -a.z : a.z); }
inline v3 frac(v3 a) { return sub(a, mk((float)(int)a.x, (float)(int)a.y, (float)(int)a.z)); }
// std::min / std::max semantics
inline float fmin_(float a, float b) { return (b < a) ? b : a; }
inline float fmax_(float a, float b) { return (a < b) ? b : a; }
inline v3 vmin(v3 a, v3 b) { return mk(fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)); }
inline v3 vmax(v3 a, v3 b) { return mk(fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)); }

inline tri load_tri(const float* f) {
  tri t;
  for (int k = 0; k < 3; ++k) t.p[k] = mk(f[3 * k], f[3 * k + 1], f[3 * k + 2]);
  return t;
}
inline ray load_ray(const float* r) {
  ray q; q.o = mk(r[0], r[1], r[2]); q.d = mk(r[3], r[4], r[5]); return q;
}

const double kEps = 0.00000001;  // math/operations.h:31 (a double literal)

// math/triangle.cpp:28-79. Returns 0 (unique hit) or a negative code:
// -1 disjoint, -2 degenerate, -3 ray in the triangle plane.
int intersect(const ray& r, const tri& T, v3* point) {
  v3 u = sub(T.p[1], T.p[0]);
  v3 v = sub(T.p[2], T.p[0]);
  v3 n = cross(u, v);
  if ((double)dot(n, n) < kEps) return -2;
  v3 dir = sub(r.d, r.o);
  v3 w0 = sub(r.o, T.p[0]);
  float a = -dot(n, w0);
  float b = dot(n, dir);
  if ((double)std::fabs(b) < kEps) return a == 0 ? -3 : -1;
  float rr = a / b;
  if (rr < 0.0) return -1;
  *point = add(r.o, mul(dir, rr));  // ray.origin + r * dir
  float uu = dot(u, u), uv = dot(u, v), vv = dot(v, v);
  v3 w = sub(*point, T.p[0]);
  float wu = dot(w, u), wv = dot(w, v);
  float D = uv * uv - uu * vv;
  float s = (uv * wv - vv * wu) / D;
  if (s < 0.0 || s > 1.0) return -1;
  float t = (uv * wu - uu * wv) / D;
  if (t < 0.0 || (s + t) > 1.0) return -1;
  return 0;
}

// math/triangle.cpp:86-92
float area(const tri& T) {
  v3 v = sub(T.p[1], T.p[0]);
  v3 w = sub(T.p[2], T.p[0]);
  v3 n = cross(v, w);
  return std::sqrt(dot(n, n)) / 2;
}
float area_xy(tri T) { T.p[0].z = 0; T.p[1].z = 0; T.p[2].z = 0; return area(T); }
v3 normal(const tri& T) { return cross(sub(T.p[1], T.p[0]), sub(T.p[2], T.p[0])); }

// math/mat.h:30-37: note the column-major access p[col][row]
v3 mat_vec(const m33& m, v3 v) {
  return mk((m.p[0][0] * v.x) + (m.p[1][0] * v.y) + (m.p[2][0] * v.z),
            (m.p[0][1] * v.x) + (m.p[1][1] * v.y) + (m.p[2][1] * v.z),
            (m.p[0][2] * v.x) + (m.p[1][2] * v.y) + (m.p[2][2] * v.z));
}
m33 mat(float a00, float a01, float a02, float a10, float a11, float a12, float a20,
        float a21, float a22) {
  m33 m = {{{a00, a01, a02}, {a10, a11, a12}, {a20, a21, a22}}};
  return m;
}

// math/mat.cpp:25-49 with the element placement of math/mat.h:64-97
// (ssc_mul, operator* by a scalar and operator+ each build their result
// through make_mat33 in the order written there).
m33 rotate_towards(v3 subject, v3 to) {
  v3 v = cross(subject, to);
  float s2 = dot(v, v);
  float c = dot(subject, to);
  m33 rot = mat(1, 0, 0, 0, 1, 0, 0, 0, 1);
  if (s2 > FLT_MIN) {
    m33 a = mat(0, v.z, -v.y, -v.z, 0, v.x, v.y, -v.x, 0);
    const m33& b = a;
    m33 q = mat(a.p[0][1] * b.p[1][0] + a.p[0][2] * b.p[2][0], a.p[1][2] * b.p[2][0],
                a.p[2][1] * b.p[1][0], a.p[0][1] * b.p[1][1] + a.p[0][2] * b.p[2][1],
                a.p[1][0] * b.p[0][1] + a.p[1][2] * b.p[2][1], a.p[2][0] * b.p[0][1],
                a.p[0][1] * b.p[1][2], a.p[1][0] * b.p[0][2],
                a.p[2][0] * b.p[0][2] + a.p[2][1] * b.p[1][2]);
    float k = (1 - c) / s2;
    q = mat(q.p[0][0] * k, q.p[1][0] * k, q.p[2][0] * k, q.p[0][1] * k, q.p[1][1] * k,
            q.p[2][1] * k, q.p[0][2] * k, q.p[1][2] * k, q.p[2][2] * k);
    m33 s1 = mat(rot.p[0][0] + a.p[0][0], rot.p[1][0] + a.p[1][0], rot.p[2][0] + a.p[2][0],
                 rot.p[0][1] + a.p[0][1], rot.p[1][1] + a.p[1][1], rot.p[2][1] + a.p[2][1],
                 rot.p[0][2] + a.p[0][2], rot.p[1][2] + a.p[1][2], rot.p[2][2] + a.p[2][2]);
    rot = mat(s1.p[0][0] + q.p[0][0], s1.p[1][0] + q.p[1][0], s1.p[2][0] + q.p[2][0],
              s1.p[0][1] + q.p[0][1], s1.p[1][1] + q.p[1][1], s1.p[2][1] + q.p[2][1],
              s1.p[0][2] + q.p[0][2], s1.p[1][2] + q.p[1][2], s1.p[2][2] + q.p[2][2]);
  } else {
    v3 f = add(subject, to);
    float l = dot(f, f);
    if (l < FLT_EPSILON) rot = mat(1, 0, 0, 0, -1, 0, 0, 0, -1);
  }
  return rot;
}

// Unit normal for the orthonormal frame: n / |n| component-wise.
v3 unit(v3 n) {
  float len = std::sqrt(dot(n, n));
  return mk(n.x / len, n.y / len, n.z / len);
}

/* ---------------- Philox4x32-10 and the turn-based sin/cos -------------- */
// Philox4x32 with 10 rounds, as published by Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3" (SC'11), Table 2 constants.
void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
  uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
  uint32_t k0 = k_in[0], k1 = k_in[1];
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
inline float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// sin and cos of 2*pi*u for u in [0, 1): octant reduction that is exact in
// binary32 (u*8 and 1-f are exact), then Taylor polynomials on [0, pi/4]
// evaluated in Horner form. Truncation error < 2e-9 on that interval.
void sincos_turns(float u, float* s_out, float* c_out) {
  float u8 = u * 8.0f;
  int o = (int)u8;
  float f = u8 - (float)o;
  if (o & 1) f = 1.0f - f;
  float x = f * 0.785398163397448f;  // f * pi/4
  float x2 = x * x;
  float sp = 1.0f + x2 * (-1.66666667e-1f + x2 * (8.33333333e-3f + x2 * (-1.98412698e-4f + x2 * 2.75573192e-6f)));
  float s = x * sp;
  float c = 1.0f + x2 * (-0.5f + x2 * (4.16666667e-2f + x2 * (-1.38888889e-3f + x2 * (2.48015873e-5f + x2 * -2.75573192e-7f))));
  // angle = k*90deg + x (even octant) or k*90deg - x (odd octant)
  int k = ((o + 1) >> 1) & 3;
  float bs = (o & 1) ? -s : s;
  float bc = c;
  float rs, rc;
  switch (k) {
    case 0: rs = bs; rc = bc; break;
    case 1: rs = bc; rc = -bs; break;
    case 2: rs = -bs; rc = -bc; break;
    default: rs = -bc; rc = bs; break;
  }
  *s_out = rs; *c_out = rc;
}

/* ---------------- casters ------------------------------------------------ */

// ray_caster/naive_cpu.cpp:65-97
void naive_range(const tri* T, int n_faces, const float* rays, long lo, long hi, int* idx,
                 float* pts) {
  for (long t = lo; t < hi; ++t) {
    ray r = load_ray(rays + 6 * t);
    float best = std::numeric_limits<float>::max();
    int hit = -1;
    v3 hp;
    for (int f = 0; f != n_faces; ++f) {
      v3 p;
      if (intersect(r, T[f], &p) == 0) {
        v3 sd = sub(p, r.o);
        float d = dot(sd, sd);
        if (d < best) { best = d; hit = f; hp = p; }
      }
    }
    idx[t] = hit;
    if (hit >= 0) { pts[3 * t] = hp.x; pts[3 * t + 1] = hp.y; pts[3 * t + 2] = hp.z; }
  }
}

template <class F>
void parallel_slices(long n, int n_threads, F fn) {
  if (n_threads <= 1 || n < 2 * n_threads) { fn(0L, n); return; }
  std::vector<std::thread> th;
  long chunk = (n + n_threads - 1) / n_threads;
  for (int i = 0; i < n_threads; ++i) {
    long lo = i * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=] { fn(lo, hi); });
  }
  for (auto& t : th) t.join();
}

/* ---- 2-D "zgrid" (math/grid.cpp, math/triangle.cpp:412-450) ------------- */
struct grid2 { v3 base, side, size; int nx, ny; };
const int kContour = 41;  // math/grid.h:90

grid2 make_grid(v3 base, v3 size, int nx, int ny) {  // math/grid.h:55-58
  grid2 g; g.base = base; g.size = size;
  g.side = divv(size, mk((float)nx, (float)ny, 0));
  g.nx = nx; g.ny = ny;
  return g;
}

// math/ray.cpp:23-35
bool ray_segment(const ray& p, const ray& q, float& s) {
  v3 v = sub(q.d, q.o);
  v3 w = sub(p.o, q.o);
  v3 u = sub(p.d, p.o);
  float d = (v.x * u.y - v.y * u.x);
  s = (v.y * w.x - v.x * w.y) / d;
  float t = (u.x * w.y - u.y * w.x) / -d;
  return t >= 0 && t <= 1 && s >= 0;
}

// math/grid.cpp:29-64
bool rebase(const grid2& g, ray& r) {
  v3 u = sub(r.o, g.base);
  v3 h = sub(u, g.size);
  v3 v = sub(r.d, r.o);
  bool out = u.x < 0 || h.x > 0 || u.y < 0 || h.y > 0;
  if (!out) return true;
  ray pr = {u, sub(r.d, g.base)};
  float s1, s2, s3, s4;
  ray e1 = {mk(0, 0, 0), mk(g.size.x, 0, 0)};
  ray e2 = {mk(0, 0, 0), mk(0, g.size.y, 0)};
  ray e3 = {mk(0, g.size.y, 0), mk(g.size.x, g.size.y, 0)};
  ray e4 = {mk(g.size.x, 0, 0), mk(g.size.x, g.size.y, 0)};
  bool i1 = ray_segment(pr, e1, s1);
  bool i2 = ray_segment(pr, e2, s2);
  bool i3 = ray_segment(pr, e3, s3);
  bool i4 = ray_segment(pr, e4, s4);
  if (!i1 && !i2 && !i3 && !i4) return false;
  float s = FLT_MAX;
  if (i1) s = fmin_(s, s1);
  if (i2) s = fmin_(s, s2);
  if (i3) s = fmin_(s, s3);
  if (i4) s = fmin_(s, s4);
  v3 shift = mul(v, s + 0.00001f);
  r.o = add(r.o, shift);
  r.d = add(r.d, shift);
  return true;
}

// math/grid.cpp:66-115. The callback returns true to stop. Cells outside
// [0,nx)x[0,ny) are not visited (the reference would index out of bounds).
template <class CB>
void traverse(const grid2& g, ray r, bool is_segment, CB cb) {
  if (!rebase(g, r)) return;
  v3 u = sub(r.o, g.base);
  v3 v = sub(r.d, r.o);
  int px = (int)(u.x / g.side.x), py = (int)(u.y / g.side.y);
  v3 t_delta = vabs(divv(g.side, v));
  v3 base_distance = mk(v.x < 0 ? 0 : 1.f, v.y < 0 ? 0 : 1.f, 0);
  v3 t = vabs(mulv(t_delta, sub(base_distance, frac(divv(u, g.side)))));
  int step_x = v.x < 0 ? -1 : 1, step_y = v.y < 0 ? -1 : 1;
  int stop_x = v.x < 0 ? -1 : g.nx, stop_y = v.y < 0 ? -1 : g.ny;
  for (long guard = 0; guard < (1L << 26); ++guard) {
    if (px < 0 || py < 0 || px >= g.nx || py >= g.ny) return;
    if (cb(px, py)) return;
    if (is_segment && (t.x > 1 && t.y > 1)) return;
    const bool do_x = t.x <= t.y;
    const bool do_y = t.y <= t.x;
    if (do_x) { px += step_x; if (px == stop_x) return; t.x += t_delta.x; }
    if (do_y) { py += step_y; if (py == stop_y) return; t.y += t_delta.y; }
    if (!do_x && !do_y) return;  // NaN steps: the reference would loop forever
  }
}

// math/grid.cpp:127-205: contour rows kept sorted by y, capped at 41 rows.
struct contour { int size; int pivot[kContour], mn[kContour], mx[kContour]; };

void contour_add(contour& c, int x, int y) {
  int lo = 0;
  while (lo < c.size && c.pivot[lo] < y) ++lo;
  if (lo < c.size && c.pivot[lo] == y) {
    c.mn[lo] = std::min(c.mn[lo], x);
    c.mx[lo] = std::max(c.mx[lo], x);
    return;
  }
  if (c.size == kContour) return;  // row dropped, as in the reference
  for (int i = c.size; i > lo; --i) {
    c.pivot[i] = c.pivot[i - 1]; c.mn[i] = c.mn[i - 1]; c.mx[i] = c.mx[i - 1];
  }
  c.pivot[lo] = y; c.mn[lo] = x; c.mx[lo] = x;
  ++c.size;
}

template <class CB>
void rasterize(const grid2& g, const tri& T, CB cb) {
  contour c; c.size = 0;
  auto col = [&](int x, int y) { contour_add(c, x, y); return false; };
  ray e01 = {T.p[0], T.p[1]}, e02 = {T.p[0], T.p[2]}, e12 = {T.p[1], T.p[2]};
  traverse(g, e01, true, col);
  traverse(g, e02, true, col);
  traverse(g, e12, true, col);
  for (int i = 0; i != c.size; ++i)
    for (int x = c.mn[i]; x <= c.mx[i]; ++x)
      if (cb(x, c.pivot[i])) return;
}

}  // namespace

struct or_zgrid {
  grid2 g;
  std::vector<int> count, offset, refs;
  std::vector<tri> faces;
};

extern "C" {

int or_triangle_intersect(const float* ray6, const float* tri9, float* point3) {
  v3 p = mk(0, 0, 0);
  int r = intersect(load_ray(ray6), load_tri(tri9), &p);
  if (r == 0 && point3) { point3[0] = p.x; point3[1] = p.y; point3[2] = p.z; }
  return r;
}

float or_triangle_area(const float* tri9) { return area(load_tri(tri9)); }

void or_rotate_towards(const float* a, const float* b, float* m9) {
  m33 m = rotate_towards(mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]));
  std::memcpy(m9, m.p, sizeof(m.p));
}

void or_mat_vec(const float* m9, const float* v, float* out) {
  m33 m; std::memcpy(m.p, m9, sizeof(m.p));
  v3 r = mat_vec(m, mk(v[0], v[1], v[2]));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void or_philox4x32_10(const uint32_t* c, const uint32_t* k, uint32_t* o) { philox(c, k, o); }
void or_sincos_turns(float u, float* s, float* c) { sincos_turns(u, s, c); }

void or_naive_cast(const float* faces, int n_faces, const float* rays, long n_rays, int* idx,
                   float* pts, int n_threads) {
  std::vector<tri> T(n_faces);
  for (int f = 0; f < n_faces; ++f) T[f] = load_tri(faces + 9 * f);
  parallel_slices(n_rays, n_threads, [&](long lo, long hi) {
    naive_range(T.data(), n_faces, rays, lo, hi, idx, pts);
  });
}

// ray_caster/zgrid_cpu.cpp:57-68: triangles_analyze -> grid_deduce_optimal
// -> grid_make_index -> grid_index_triangles.
or_zgrid* or_zgrid_build(const float* faces, int n_faces) {
  if (n_faces <= 0) return nullptr;
  or_zgrid* z = new or_zgrid;
  z->faces.resize(n_faces);
  for (int f = 0; f < n_faces; ++f) z->faces[f] = load_tri(faces + 9 * f);
  const std::vector<tri>& T = z->faces;
  // triangles_analyze (math/triangle.cpp:412-450), including its use of
  // max_area where max_area_xy is meant (:432)
  v3 amin = vmin(vmin(T[0].p[0], T[0].p[1]), T[0].p[2]);
  v3 amax = vmax(vmax(T[0].p[0], T[0].p[1]), T[0].p[2]);
  float max_area = area(T[0]);
  float max_area_xy = area_xy(T[0]);
  float average_area = max_area, average_area_xy = max_area_xy;
  for (int i = 1; i != n_faces; ++i) {
    v3 tmin = vmin(vmin(T[i].p[0], T[i].p[1]), T[i].p[2]);
    v3 tmax = vmax(vmax(T[i].p[0], T[i].p[1]), T[i].p[2]);
    amin = vmin(amin, tmin);
    amax = vmax(amax, tmax);
    float a = area(T[i]), axy = area_xy(T[i]);
    max_area = fmax_(max_area, a);
    average_area += a;
    max_area_xy = fmax_(max_area, axy);
    average_area_xy += axy;
  }
  amax = mul(amax, 1.0001f);
  average_area /= (float)n_faces;
  average_area_xy /= (float)n_faces;
  // grid_deduce_optimal (math/grid.cpp:376-386)
  v3 volume = sub(amax, amin);
  float total_area = volume.x * volume.y;
  int n_avg_max_side = (int)std::sqrt(kContour * total_area / (2.5f * average_area_xy));
  int n_max_side = (int)std::sqrt(kContour * total_area / (2.5f * max_area_xy));
  int side = std::min(n_max_side, n_avg_max_side);
  if (side < 1) side = 1;
  z->g = make_grid(amin, volume, side, side);
  // grid_index_triangles (math/grid.cpp:291-306) + linearize (:231-259)
  int nc = side * side;
  std::vector<std::vector<int>> lists(nc);
  for (int t = 0; t != n_faces; ++t)
    rasterize(z->g, T[t], [&](int x, int y) { lists[x + side * y].push_back(t); return false; });
  z->count.assign(nc, 0);
  z->offset.assign(nc, 0);
  int l = 0;
  for (int i = 0; i < nc; ++i) {
    if (lists[i].empty()) continue;
    z->offset[i] = l;
    z->count[i] = (int)lists[i].size();
    z->refs.insert(z->refs.end(), lists[i].begin(), lists[i].end());
    l += (int)lists[i].size();
  }
  return z;
}

void or_zgrid_free(or_zgrid* g) { delete g; }

void or_zgrid_info(const or_zgrid* z, int* nx, int* ny, float* base3, float* side3, float* size3,
                   int* n_refs) {
  *nx = z->g.nx; *ny = z->g.ny;
  base3[0] = z->g.base.x; base3[1] = z->g.base.y; base3[2] = z->g.base.z;
  side3[0] = z->g.side.x; side3[1] = z->g.side.y; side3[2] = z->g.side.z;
  size3[0] = z->g.size.x; size3[1] = z->g.size.y; size3[2] = z->g.size.z;
  *n_refs = (int)z->refs.size();
}

void or_zgrid_cells(const or_zgrid* z, int* counts, int* offsets, int* refs) {
  std::memcpy(counts, z->count.data(), z->count.size() * sizeof(int));
  std::memcpy(offsets, z->offset.data(), z->offset.size() * sizeof(int));
  if (!z->refs.empty()) std::memcpy(refs, z->refs.data(), z->refs.size() * sizeof(int));
}

// ray_caster/zgrid_cpu.cpp:79-132: stop at the first cell with a hit whose
// point lies in that cell (locator check), nearest by squared distance.
void or_zgrid_cast(const or_zgrid* z, const float* rays, long n_rays, int* idx, float* pts,
                   int n_threads) {
  parallel_slices(n_rays, n_threads, [&](long lo, long hi) {
    for (long t = lo; t < hi; ++t) {
      ray r = load_ray(rays + 6 * t);
      float best = std::numeric_limits<float>::max();
      int hit = -1;
      v3 hp;
      traverse(z->g, r, false, [&](int px, int py) {
        int c = px + z->g.nx * py;
        for (int i = 0; i != z->count[c]; ++i) {
          int f = z->refs[z->offset[c] + i];
          v3 p;
          if (intersect(r, z->faces[f], &p) == 0) {
            v3 loc = divv(sub(p, z->g.base), z->g.side);
            if (px != (int)loc.x || py != (int)loc.y) continue;
            v3 sd = sub(p, r.o);
            float d = dot(sd, sd);
            if (d < best) { best = d; hit = f; hp = p; }
          }
        }
        return hit != -1;
      });
      idx[t] = hit;
      if (hit >= 0) { pts[3 * t] = hp.x; pts[3 * t + 1] = hp.y; pts[3 * t + 2] = hp.z; }
    }
  });
}

int or_grid_traverse(const float* base3, const float* size3, int nx, int ny, const float* ray6,
                     int is_segment, int* cells, int max_cells) {
  grid2 g = make_grid(mk(base3[0], base3[1], base3[2]), mk(size3[0], size3[1], size3[2]), nx, ny);
  int n = 0;
  traverse(g, load_ray(ray6), is_segment != 0, [&](int x, int y) {
    if (n < max_cells) { cells[2 * n] = x; cells[2 * n + 1] = y; }
    ++n;
    return n >= max_cells;
  });
  return n;
}

int or_grid_rasterize(const float* base3, const float* size3, int nx, int ny, const float* tri9,
                      int* cells, int max_cells) {
  grid2 g = make_grid(mk(base3[0], base3[1], base3[2]), mk(size3[0], size3[1], size3[2]), nx, ny);
  int n = 0;
  rasterize(g, load_tri(tri9), [&](int x, int y) {
    if (n < max_cells) { cells[2 * n] = x; cells[2 * n + 1] = y; }
    ++n;
    return n >= max_cells;
  });
  return n;
}

int or_ray_intersect_segment(const float* p6, const float* q6, float* s) {
  return ray_segment(load_ray(p6), load_ray(q6), *s) ? 1 : 0;
}

/* ---------------- emission ---------------------------------------------- */

// emission/malley_emission.cpp:50-60: (int)(n_rays * (w / total)) in float,
// at least one ray for a positive weight.
static inline int emitted(float w, int n_rays, float total) {
  return w > 0 ? std::max<int>(1, (int)(n_rays * (w / total))) : 0;
}

void or_emitted_counts(const float* weights, int n_faces, int n_rays, float total, int* front,
                       int* rear) {
  for (int f = 0; f < n_faces; ++f) {
    front[f] = emitted(weights[2 * f], n_rays, total);
    rear[f] = emitted(weights[2 * f + 1], n_rays, total);
  }
}

float or_area_weights(const float* faces, int n_faces, float* weights) {
  float total = 0;
  for (int i = 0; i != n_faces; ++i) {
    float a = area(load_tri(faces + 9 * i));
    total += 2.f * a;
    weights[2 * i] = a;
    weights[2 * i + 1] = a;
  }
  return total;
}

long or_malley_count(const float* w, int n_faces, int n_rays, float total) {
  long n = 0;  // emission/malley_cpu.cpp:96-106 (an int there)
  for (int f = 0; f != n_faces; ++f)
    n += emitted(w[2 * f], n_rays, total) + emitted(w[2 * f + 1], n_rays, total);
  return n;
}

// emission/malley_cpu.cpp:109-189. In MT19937 mode the four generators and
// distributions are those of malley_cpu.cpp:55-65 (fresh per call); in
// PHILOX mode each ray draws one Philox block with counter (j, face, 0, 0)
// and key (seed), and cos/sin of 2*pi*u come from sincos_turns.
long or_malley_emit(const float* faces, int n_faces, const float* weights, float total,
                    int n_rays, int rng, uint64_t seed, int frame, float* out) {
  long n_real = or_malley_count(weights, n_faces, n_rays, total);
  std::mt19937 gA(1), gB(2), gTheta(3), gR(4);
  std::uniform_real_distribution<float> d01(0, 1);
  std::uniform_real_distribution<float> d2pi(0, float(6.28318530717958647692528676655900576));
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  long n_ray = 0;
  for (int f = 0; f != n_faces; ++f) {
    tri T = load_tri(faces + 9 * f);
    int fr = emitted(weights[2 * f], n_rays, total);
    int rr = emitted(weights[2 * f + 1], n_rays, total);
    int face_rays = fr + rr;
    v3 n = normal(T);
    m33 R = rotate_towards(mk(0, 0, 1), frame == OR_FRAME_ORTHONORMAL ? unit(n) : n);
    for (int j = 0; j != face_rays && n_ray < n_real; ++j, ++n_ray) {
      float a, b, r, cphi, sphi;
      if (rng == OR_RNG_MT19937) {
        a = d01(gA);
        b = d01(gB);
        r = d01(gR);
        float phi = d2pi(gTheta);
        cphi = std::cos(phi);
        sphi = std::sin(phi);
      } else {
        uint32_t ctr[4] = {(uint32_t)j, (uint32_t)f, 0u, 0u}, o[4];
        philox(ctr, key, o);
        a = u01(o[0]); b = u01(o[1]); r = u01(o[2]);
        sincos_turns(u01(o[3]), &sphi, &cphi);
      }
      if (a + b > 1) { a = 1.f - a; b = 1.f - b; }
      v3 v0 = sub(T.p[1], T.p[0]);
      v3 v1 = sub(T.p[2], T.p[0]);
      v3 origin = add(add(T.p[0], mul(v0, a)), mul(v1, b));
      float rad = std::sqrt(r);
      v3 m = mk(rad * cphi, rad * sphi, std::sqrt(1 - r));
      if (j >= fr) m.z = -m.z;
      v3 d = mat_vec(R, m);
      v3 ro = add(origin, mul(d, 0.0001f));
      v3 re = add(origin, d);
      float* o6 = out + 6 * n_ray;
      o6[0] = ro.x; o6[1] = ro.y; o6[2] = ro.z; o6[3] = re.x; o6[4] = re.y; o6[5] = re.z;
    }
  }
  return n_real;
}

// emission/parallel_rays_cpu.cpp:88-106 (disk aperture of the given radius
// at -distance along dir around center). PHILOX mode: counter (r, 0, 1, 0).
void or_parallel_emit2(const float* dir3, const float* center3, float distance, float radius,
                       int aperture, int n_rays, int rng, uint64_t seed, float* out);

void or_parallel_emit(const float* dir3, const float* center3, float distance, float radius,
                      int n_rays, int rng, uint64_t seed, float* out) {
  or_parallel_emit2(dir3, center3, distance, radius, 0, n_rays, rng, seed, out);
}

// aperture 0: disk of the given radius (emission/parallel_rays_cpu.cpp:88-106);
// aperture 1: square of half side `radius` (Philox only), x = radius * (2u - 1),
// y = radius * (2v - 1) in the frame rotated towards dir.
void or_parallel_emit2(const float* dir3, const float* center3, float distance, float radius,
                       int aperture, int n_rays, int rng, uint64_t seed, float* out) {
  v3 dir = mk(dir3[0], dir3[1], dir3[2]);
  v3 center = mk(center3[0], center3[1], center3[2]);
  m33 R = rotate_towards(mk(0, 0, 1), dir);
  std::mt19937 gR(1), gTheta(2);
  std::uniform_real_distribution<float> dR(0, 1);
  std::uniform_real_distribution<float> dTheta(0, 2 * M_PI);
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int r = 0; r != n_rays; ++r) {
    float rad, c, s;
    if (rng == OR_RNG_MT19937) {
      rad = radius * std::sqrt(dR(gR));
      float theta = dTheta(gTheta);
      c = std::cos(theta);
      s = std::sin(theta);
    } else {
      uint32_t ctr[4] = {(uint32_t)r, 0u, 1u, 0u}, o[4];
      philox(ctr, key, o);
      if (aperture == 1) {
        rad = 1.f;
        c = radius * (2.f * u01(o[0]) - 1.f);
        s = radius * (2.f * u01(o[1]) - 1.f);
      } else {
        rad = radius * std::sqrt(u01(o[0]));
        sincos_turns(u01(o[1]), &s, &c);
      }
    }
    float x = rad * c, y = rad * s;
    v3 origin = add(mat_vec(R, mk(x, y, -distance)), center);
    v3 end = add(origin, mul(dir, 0.1f));
    float* o6 = out + 6 * (long)r;
    o6[0] = origin.x; o6[1] = origin.y; o6[2] = origin.z;
    o6[3] = end.x; o6[4] = end.y; o6[5] = end.z;
  }
}

/* ---------------- form factors ------------------------------------------ */

// cpuFactorsCalculator/cpu_system.cpp:163-195: rays are walked in mesh order,
// counted as floats, each row scaled by 1/(rays leaving that mesh).
int or_ff_tally(const int* meshes, int n_meshes, int n_faces, const int* front, const int* rear,
                long n_total, const int* hit_idx, float* F) {
  std::vector<int> f2m(n_faces, -1);
  for (int m = 0; m != n_meshes; ++m)
    for (int f = 0; f != meshes[3 * m + 1]; ++f) f2m[meshes[3 * m] + f] = m;
  std::memset(F, 0, sizeof(float) * (size_t)n_meshes * n_meshes);
  long n_ray = 0;
  for (int m = 0; m != n_meshes; ++m) {
    int outgoing = 0;
    for (int f = 0; f != meshes[3 * m + 1]; ++f) {
      int fi = meshes[3 * m] + f;
      int face_rays = front[fi] + rear[fi];
      for (int j = 0; j != face_rays && n_ray < n_total; ++j, ++n_ray) {
        ++outgoing;
        int h = hit_idx[n_ray];
        if (h >= 0 && f2m[h] >= 0) F[(size_t)m * n_meshes + f2m[h]] += 1.f;
      }
    }
    float ratio = 1.f / (float)outgoing;
    for (int j = 0; j != n_meshes; ++j) F[(size_t)m * n_meshes + j] *= ratio;
  }
  return 0;
}

int or_ff_calculate(const float* faces, int n_faces, const int* meshes, int n_meshes, int n_rays,
                    int rng, uint64_t seed, int frame, int caster, int n_threads, float* F) {
  if (n_faces == 0 || n_meshes == 0) return -101;  // FORM_FACTORS_EMPTY_SCENE
  std::vector<float> w(2 * (size_t)n_faces);
  float total = or_area_weights(faces, n_faces, w.data());
  if (total < FLT_EPSILON) return -102;  // FORM_FACTORS_SCENE_TOO_SMALL
  std::vector<int> fr(n_faces), rr(n_faces);
  or_emitted_counts(w.data(), n_faces, n_rays, total, fr.data(), rr.data());
  long n = or_malley_count(w.data(), n_faces, n_rays, total);
  std::vector<float> rays(6 * (size_t)n);
  std::vector<float> pts(3 * (size_t)n);
  std::vector<int> idx(n);
  or_malley_emit(faces, n_faces, w.data(), total, n_rays, rng, seed, frame, rays.data());
  if (caster == 0) {
    or_naive_cast(faces, n_faces, rays.data(), n, idx.data(), pts.data(), n_threads);
  } else {
    or_zgrid* z = or_zgrid_build(faces, n_faces);
    or_zgrid_cast(z, rays.data(), n, idx.data(), pts.data(), n_threads);
    or_zgrid_free(z);
  }
  return or_ff_tally(meshes, n_meshes, n_faces, fr.data(), rr.data(), n, idx.data(), F);
}

/* ---------------- thermal ----------------------------------------------- */

// math::triangles_bsphere (math/triangle.cpp:361-383): the centre is the mean
// of the triangle centres ((p0 + p1 + p2) / 3, summed in order, times 1/n);
// the radius is the largest vertex distance over triangles 1..n-1 (the
// reference's loop starts at 1, so triangle 0's vertices are not measured).
void or_bsphere(const float* faces, int n_faces, float* center3, float* radius) {
  v3 gc = mk(0, 0, 0);
  for (int i = 0; i != n_faces; ++i) {
    const float* t = faces + 9 * (size_t)i;
    v3 c = add(add(mk(t[0], t[1], t[2]), mk(t[3], t[4], t[5])), mk(t[6], t[7], t[8]));
    c = mk(c.x / 3.f, c.y / 3.f, c.z / 3.f);
    gc = add(gc, c);
  }
  const float inv = 1.f / (float)n_faces;
  gc = mk(gc.x * inv, gc.y * inv, gc.z * inv);
  float r = 0;
  for (int i = 1; i < n_faces; ++i) {
    const float* t = faces + 9 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      v3 d = sub(gc, mk(t[3 * k], t[3 * k + 1], t[3 * k + 2]));
      float n = std::sqrt(dot(d, d));
      if (n > r) r = n;
    }
  }
  center3[0] = gc.x;
  center3[1] = gc.y;
  center3[2] = gc.z;
  *radius = r;
}

// parallel_rays_cpu::calculate (thermal_equation/parallel_rays_cpu.cpp:92-131):
// rays from a disk (aperture 0) or square (1) of radius 1.1 * bsphere radius,
// distance 1.1 + radius before the bsphere centre; each hit deposits
// absorbance * ray_power on the hit face's mesh, front absorbance when
// dot(normal, dir) < 0, ray_power = power * (pi * r * r) / n_rays. absorption
// accumulates in float in ray order, absorption64 (may be NULL) in double.
void or_sun(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* absorb2,
            float power, const float* dir3, int n_rays, int rng, uint64_t seed, int aperture, int caster,
            int n_threads, float* absorption, double* absorption64) {
  std::vector<int> f2m(n_faces, -1);
  for (int m = 0; m != n_meshes; ++m)
    for (int f = 0; f != meshes[3 * m + 1]; ++f) f2m[meshes[3 * m] + f] = m;
  float center[3], r0;
  or_bsphere(faces, n_faces, center, &r0);
  const float radius = r0 * 1.1f;
  const float distance = 1.1f + radius;
  std::vector<float> rays(6 * (size_t)std::max(n_rays, 1)), pts(3 * (size_t)std::max(n_rays, 1));
  std::vector<int> idx(std::max(n_rays, 1));
  or_parallel_emit2(dir3, center, distance, radius, aperture, n_rays, rng, seed, rays.data());
  if (caster == 0) {
    or_naive_cast(faces, n_faces, rays.data(), n_rays, idx.data(), pts.data(), n_threads);
  } else {
    or_zgrid* z = or_zgrid_build(faces, n_faces);
    or_zgrid_cast(z, rays.data(), n_rays, idx.data(), pts.data(), n_threads);
    or_zgrid_free(z);
  }
  const float ray_power = power * ((float)M_PI * radius * radius) / n_rays;
  const v3 dir = mk(dir3[0], dir3[1], dir3[2]);
  for (int r = 0; r != n_rays; ++r) {
    const int h = idx[r];
    if (h < 0 || f2m[h] < 0) continue;
    const float* t = faces + 9 * (size_t)h;
    v3 nrm = cross(sub(mk(t[3], t[4], t[5]), mk(t[0], t[1], t[2])), sub(mk(t[6], t[7], t[8]), mk(t[0], t[1], t[2])));
    const float side = dot(nrm, dir);
    const int m = f2m[h];
    const float v = (side < 0 ? absorb2[2 * m] : absorb2[2 * m + 1]) * ray_power;
    absorption[m] += v;
    if (absorption64) absorption64[m] += (double)v;
  }
}

// thermal_equation/radiance_cpu.cpp:96-122: emission weights T * eps * A per
// face side, in float, accumulated in mesh order; faces outside every mesh
// keep weight 0.
float or_radiance_weights(const float* faces, int n_faces, const int* meshes, int n_meshes,
                          const float* eps2, const float* T, float* weights) {
  std::memset(weights, 0, sizeof(float) * 2 * (size_t)n_faces);
  float total = 0;
  for (int m = 0; m != n_meshes; ++m) {
    const float density = T[m];
    const float front_density = density * eps2[2 * m];
    const float rear_density = density * eps2[2 * m + 1];
    for (int f = meshes[3 * m]; f != meshes[3 * m] + meshes[3 * m + 1]; ++f) {
      const float a = or_triangle_area(faces + 9 * (size_t)f);
      const float front = front_density * a;
      const float rear = rear_density * a;
      total += front + rear;
      weights[2 * f] = front;
      weights[2 * f + 1] = rear;
    }
  }
  return total;
}

// thermal_equation/radiance_cpu.cpp:132-190 over given hit indices of the
// Malley rays emitted with those weights (emission order = face order). The
// walk goes mesh by mesh, face by face, consuming front then rear rays of
// each face; a ray that hits deposits front_emission / front_rays (or the
// rear analogue) on the hit face's mesh. absorption is accumulated in float
// in ray order, as the reference does; absorption64 (may be NULL) holds the
// same sum in double. Hits on faces outside every mesh are dropped.
void or_radiance_walk(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* eps2,
                      const float* T, const int* front, const int* rear, long n_total, const int* hit_idx,
                      float* emission, float* absorption, double* absorption64) {
  const float sigma = 5.670400e-8f;
  std::vector<int> f2m(n_faces, -1);
  for (int m = 0; m != n_meshes; ++m)
    for (int f = 0; f != meshes[3 * m + 1]; ++f) f2m[meshes[3 * m] + f] = m;
  long n_ray = 0;
  for (int m = 0; m != n_meshes; ++m) {
    const float t = T[m];
    const float power_density = sigma * (t * t * t * t);
    const float front_density = power_density * eps2[2 * m];
    const float rear_density = power_density * eps2[2 * m + 1];
    for (int f = meshes[3 * m]; f != meshes[3 * m] + meshes[3 * m + 1]; ++f) {
      const float a = or_triangle_area(faces + 9 * (size_t)f);
      const float front_emission = front_density * a;
      const float rear_emission = rear_density * a;
      emission[m] += front_emission + rear_emission;
      const int face_rays = front[f] + rear[f];
      for (int j = 0; j != face_rays && n_ray < n_total; ++j, ++n_ray) {
        const int h = hit_idx[n_ray];
        if (h < 0 || f2m[h] < 0) continue;
        const float ray_power = j < front[f] ? (front_emission / front[f]) : (rear_emission / rear[f]);
        absorption[f2m[h]] += ray_power;
        if (absorption64) absorption64[f2m[h]] += (double)ray_power;
      }
    }
  }
}

// The whole radiance_equation::calculate: weights, emitted counts, Malley
// rays (rng/seed/frame as or_malley_emit), cast (0 naive, 1 zgrid), walk.
int or_radiance(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* eps2,
                const float* T, int n_rays, int rng, uint64_t seed, int frame, int caster, int n_threads,
                float* emission, float* absorption, double* absorption64) {
  std::vector<float> w(2 * (size_t)n_faces);
  const float total = or_radiance_weights(faces, n_faces, meshes, n_meshes, eps2, T, w.data());
  std::vector<int> fr(n_faces), rr(n_faces);
  or_emitted_counts(w.data(), n_faces, n_rays, total, fr.data(), rr.data());
  const long n = or_malley_count(w.data(), n_faces, n_rays, total);
  std::vector<float> rays(6 * (size_t)std::max(n, 1L));
  std::vector<float> pts(3 * (size_t)std::max(n, 1L));
  std::vector<int> idx(std::max(n, 1L));
  or_malley_emit(faces, n_faces, w.data(), total, n_rays, rng, seed, frame, rays.data());
  if (caster == 0) {
    or_naive_cast(faces, n_faces, rays.data(), n, idx.data(), pts.data(), n_threads);
  } else {
    or_zgrid* z = or_zgrid_build(faces, n_faces);
    or_zgrid_cast(z, rays.data(), n, idx.data(), pts.data(), n_threads);
    or_zgrid_free(z);
  }
  or_radiance_walk(faces, n_faces, meshes, n_meshes, eps2, T, fr.data(), rr.data(), n, idx.data(), emission,
                   absorption, absorption64);
  return 0;
}

// thermal_equation/form_factors.cpp:109-128, in double precision.
void or_sbff_f64(int n, const double* F, const double* T, const double* area,
                 const double* ef, const double* er, double* emission, double* absorption) {
  const double sigma = 5.670400e-8;
  for (int m = 0; m != n; ++m) {
    double t = T[m];
    double power = sigma * (t * t * t * t) * area[m];
    double e = ef[m] * power + er[m] * power;
    emission[m] += e;
    for (int k = 0; k != n; ++k) absorption[k] += e * F[(size_t)m * n + k];
  }
}

// thermal_solution/cpu_adams.cpp:136-175 in double precision.
void or_adams_combine_f64(int n, int n_step, double dt, const double* T_n, const double* E_hist,
                          double* T_next) {
  static const double u[5][5] = {
      {1.0, 0, 0, 0, 0},
      {3.0 / 2, -1.0 / 2, 0, 0, 0},
      {23.0 / 12, -4.0 / 3, 5.0 / 12, 0, 0},
      {55.0 / 24, -59.0 / 24, 37.0 / 24, -3.0 / 8, 0},
      {1901.0 / 720, -1387.0 / 360, 109.0 / 30, -637.0 / 360, 251.0 / 720}};
  int k = n_step < 4 ? n_step : 4;
  for (int m = 0; m != n; ++m) {
    double sum = 0;
    for (int l = 0; l <= k; ++l) sum += u[k][l] * E_hist[(size_t)((n_step - l) % 5) * n + m];
    T_next[m] = T_n[m] + dt * sum;
  }
}

}  // extern "C"
