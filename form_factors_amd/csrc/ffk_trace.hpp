// Device functions for emission and nearest-hit casting (gfx950).
#pragma once

#include "ffk_internal.hpp"

namespace ffk {

struct hit_rec {
  float d2;
  int f;
  v3 p;
};

// Nearest-hit update with the naive caster's order: strict "<" on the
// squared distance, equal distances resolved to the lower face index
// (ray_caster/naive_cpu.cpp:82-92 visits faces in ascending order).
__device__ inline void consider(hit_rec& h, int f, float d2, v3 p) {
  if (d2 < h.d2 || (d2 == h.d2 && f < h.f)) {
    h.d2 = d2;
    h.f = f;
    h.p = p;
  }
}

__device__ inline void test_face(const tri_terms* __restrict__ tri, int f, v3 o, v3 dir, hit_rec& h) {
  const float4* q = reinterpret_cast<const float4*>(tri + f);
  float4 q0 = gld4(q), q1 = gld4(q + 1), q2 = gld4(q + 2), q3 = gld4(q + 3);
  v3 p;
  float d2;
  if (hit_terms(q0, q1, q2, q3, o, dir, &p, &d2)) consider(h, f, d2, p);
}

// Faces tested recently by this ray: a face met again in a later cell gives
// the same result, so it is not tested twice.
struct mailbox {
  int f0, f1, f2, f3;
  __device__ void clear() { f0 = f1 = f2 = f3 = -1; }
  __device__ bool seen(int f) const { return f == f0 || f == f1 || f == f2 || f == f3; }
  __device__ void put(int f) {
    f3 = f2;
    f2 = f1;
    f1 = f0;
    f0 = f;
  }
};

// Reference k of the grid (a face index, in cell order): the reference test
// on that face's terms, unless the ray met the face in an earlier cell.
// Returns 1 if a test was run.
__device__ inline uint32_t test_ref(const scene_desc& S, int k, v3 o, v3 dir, hit_rec& h, mailbox& mb) {
  const int f = gld(S.refs + k);
  if (mb.seen(f)) return 0;
  mb.put(f);
  test_face(S.tri, f, o, dir, h);
  return 1;
}

// Walk the uniform grid front to back (3-D DDA). Every face whose padded
// footprint overlaps a visited cell is tested with the reference test; the
// walk stops once the best squared distance is below the squared distance
// to the current cell's exit (with a 1e-4 relative margin, far above the
// rounding of both), so the result equals testing every face. An empty cell
// carries, for each of the 8 direction octants, the size k of the empty box
// [c, c + s*(k-1)] toward that octant (4 bits each); the walk leaves that box
// in one step through one of its far faces.
//
// The walk is split in two so that a caller can interleave many rays per
// lane: walk_begin clips the ray to the grid box and finds its first cell,
// walk_cell visits one cell (tests, then one step or jump). All walk state
// is kept in plain scalars passed by reference: selecting among struct
// members here makes the compiler address the struct in scratch memory.

// Bit offset of the direction's octant nibble in an empty cell's word (bit 0
// of the octant: x decreasing, bit 1: y, bit 2: z). A zero component counts as
// increasing: the ray crosses no face of that axis, so either box side holds.
__device__ inline int octant_shift(v3 dir) {
  return (dir.x < 0.f ? 4 : 0) | (dir.y < 0.f ? 8 : 0) | (dir.z < 0.f ? 16 : 0);
}

// Size k >= 1 of an empty cell's box toward the octant at bit offset `shift`.
__device__ inline int empty_box(int word, int shift) { return (int)(((uint32_t)word >> shift) & 15u); }

// Ray parameter of the next boundary of cell c along one axis (s = step sign).
__device__ inline float walk_boundary(float lo, float h, int c, int s, float o, float inv) {
  return s ? (lo + (float)(c + (s > 0)) * h - o) * inv : 3.4e38f;
}

__device__ inline int walk_sign(float d) { return d > 0.f ? 1 : (d < 0.f ? -1 : 0); }

// Returns false when the ray misses the grid box (nothing to walk).
__device__ inline bool walk_begin(const grid_desc& g, v3 o, v3 dir, float& ix, float& iy, float& iz, float& dlen,
                                  int& cx, int& cy, int& cz, float& tnx, float& tny, float& tnz, int& cell) {
  // Clip to the grid box (slabs); t is in units of dir.
  float t0 = 0.f, t1 = 3.4e38f;
  ix = 0.f;
  iy = 0.f;
  iz = 0.f;
  if (dir.x != 0.f) {
    ix = 1.0f / dir.x;
    float ta = (g.lo[0] - o.x) * ix, tb = (g.hi[0] - o.x) * ix;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.x < g.lo[0] || o.x > g.hi[0]) {
    return false;
  }
  if (dir.y != 0.f) {
    iy = 1.0f / dir.y;
    float ta = (g.lo[1] - o.y) * iy, tb = (g.hi[1] - o.y) * iy;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.y < g.lo[1] || o.y > g.hi[1]) {
    return false;
  }
  if (dir.z != 0.f) {
    iz = 1.0f / dir.z;
    float ta = (g.lo[2] - o.z) * iz, tb = (g.hi[2] - o.z) * iz;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.z < g.lo[2] || o.z > g.hi[2]) {
    return false;
  }
  if (!(t0 <= t1)) return false;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  cx = min(max((int)floorf((o.x + dir.x * t0 - g.lo[0]) * g.inv_h[0]), 0), nx - 1);
  cy = min(max((int)floorf((o.y + dir.y * t0 - g.lo[1]) * g.inv_h[1]), 0), ny - 1);
  cz = min(max((int)floorf((o.z + dir.z * t0 - g.lo[2]) * g.inv_h[2]), 0), nz - 1);
  tnx = walk_boundary(g.lo[0], g.h[0], cx, walk_sign(dir.x), o.x, ix);
  tny = walk_boundary(g.lo[1], g.h[1], cy, walk_sign(dir.y), o.y, iy);
  tnz = walk_boundary(g.lo[2], g.h[2], cz, walk_sign(dir.z), o.z, iz);
  // Distances along the ray are compared as (t * |dir|)^2 so that neither
  // factor overflows for very short or very long direction vectors.
  dlen = sqrtf(dot(dir, dir));
  cell = cx + nx * cy + nx * ny * cz;
  return true;
}

// Visits the current cell: tests its references, then moves to the next
// cell (or past an empty box). Returns false when the walk is over.
template <bool STATS>
__device__ inline bool walk_cell(const scene_desc& S, v3 o, v3 dir, float ix, float iy, float iz, float dlen,
                                 int& cx, int& cy, int& cz, float& tnx, float& tny, float& tnz, int& cell,
                                 hit_rec& h, mailbox& mb, uint32_t& n_tests, uint32_t& n_cells) {
  const grid_desc& g = S.g;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int nxy = nx * ny;
  const int sx = walk_sign(dir.x), sy = walk_sign(dir.y), sz = walk_sign(dir.z);
  const int2 ci = gld2(S.cell_info + cell);
  if (STATS) n_cells += 1;
  for (int k = ci.x; k < ci.x + ci.y; ++k) {
    const uint32_t t = test_ref(S, k, o, dir, h, mb);
    if (STATS) n_tests += t;
  }
  const int R = ci.y > 0 ? 0 : empty_box(ci.x, octant_shift(dir)) - 1;
  if (R >= 1) {
    // The box [c, c + s*R] toward the ray's octant (clipped to the grid) is
    // empty: leave it in one step.
    const int hx = sx > 0 ? min(cx + R, nx - 1) : max(cx - R, 0);
    const int hy = sy > 0 ? min(cy + R, ny - 1) : max(cy - R, 0);
    const int hz = sz > 0 ? min(cz + R, nz - 1) : max(cz - R, 0);
    const float bx = walk_boundary(g.lo[0], g.h[0], hx, sx, o.x, ix);
    const float by = walk_boundary(g.lo[1], g.h[1], hy, sy, o.y, iy);
    const float bz = walk_boundary(g.lo[2], g.h[2], hz, sz, o.z, iz);
    const bool ex = (bx <= by) && (bx <= bz);
    const bool ey = !ex && (by <= bz);
    const float tb = ex ? bx : (ey ? by : bz);
    if (!(tb < 3.4e38f)) return false;
    const float de = tb * dlen;
    if (h.f >= 0 && h.d2 < de * de * 0.9999f) return false;
    // Cell just past the exit face; the other two coordinates come from
    // the exit point, kept inside the box and never moved backwards.
    const float px = o.x + dir.x * tb, py = o.y + dir.y * tb, pz = o.z + dir.z * tb;
    int nxc = min(max((int)floorf((px - g.lo[0]) * g.inv_h[0]), max(cx - R, 0)), min(cx + R, nx - 1));
    int nyc = min(max((int)floorf((py - g.lo[1]) * g.inv_h[1]), max(cy - R, 0)), min(cy + R, ny - 1));
    int nzc = min(max((int)floorf((pz - g.lo[2]) * g.inv_h[2]), max(cz - R, 0)), min(cz + R, nz - 1));
    nxc = sx > 0 ? max(nxc, cx) : (sx < 0 ? min(nxc, cx) : cx);
    nyc = sy > 0 ? max(nyc, cy) : (sy < 0 ? min(nyc, cy) : cy);
    nzc = sz > 0 ? max(nzc, cz) : (sz < 0 ? min(nzc, cz) : cz);
    if (ex) nxc = hx + sx;
    else if (ey) nyc = hy + sy;
    else nzc = hz + sz;
    if (nxc < 0 || nxc >= nx || nyc < 0 || nyc >= ny || nzc < 0 || nzc >= nz) return false;
    cx = nxc;
    cy = nyc;
    cz = nzc;
    tnx = walk_boundary(g.lo[0], g.h[0], cx, sx, o.x, ix);
    tny = walk_boundary(g.lo[1], g.h[1], cy, sy, o.y, iy);
    tnz = walk_boundary(g.lo[2], g.h[2], cz, sz, o.z, iz);
    cell = cx + nx * cy + nxy * cz;
    return true;
  }
  const bool ax = (tnx <= tny) && (tnx <= tnz);
  const bool ay = !ax && (tny <= tnz);
  const float t_exit = ax ? tnx : (ay ? tny : tnz);
  const float de = t_exit * dlen;
  if (h.f >= 0 && h.d2 < de * de * 0.9999f) return false;
  const int st = ax ? sx : (ay ? sy : sz);
  if (st == 0) return false;  // zero or NaN direction: the start cell was all
  if (ax) {
    cx += sx;
    if (cx < 0 || cx >= nx) return false;
    cell += sx;
    tnx = walk_boundary(g.lo[0], g.h[0], cx, sx, o.x, ix);
  } else if (ay) {
    cy += sy;
    if (cy < 0 || cy >= ny) return false;
    cell += sy * nx;
    tny = walk_boundary(g.lo[1], g.h[1], cy, sy, o.y, iy);
  } else {
    cz += sz;
    if (cz < 0 || cz >= nz) return false;
    cell += sz * nxy;
    tnz = walk_boundary(g.lo[2], g.h[2], cz, sz, o.z, iz);
  }
  return true;
}

// Leaves the current cell after its tests (the second half of walk_cell):
// past the empty box [c, c + s*R] toward the ray's octant (R >= 1) or to the
// next cell of the DDA (R = 0). The box code clamps the exit cell to
// [c-R, c+R] and then to the ray's side of c, which is that box. Returns
// false when the walk is over. UNIFIED takes the box code
// for every R >= 0: with R = 0 the box is the cell itself and the next cell,
// the exit test and the new boundaries are those of the DDA step, so a
// caller may run one code path for lanes that jump and lanes that step. The
// grid fields are read into scalars and the moving axis is chosen with
// selects: written with arrays or a branch per axis, the compiler keeps the
// scene descriptor and the walk state in scratch memory.
template <bool UNIFIED = false>
__device__ inline bool walk_leave(const grid_desc& g, v3 o, v3 dir, float ix, float iy, float iz, float dlen, int R,
                                  int& cx, int& cy, int& cz, float& tnx, float& tny, float& tnz, int& cell,
                                  float best_d2, int best_f) {
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int nxy = nx * ny;
  const int sx = walk_sign(dir.x), sy = walk_sign(dir.y), sz = walk_sign(dir.z);
  const float lo0 = g.lo[0], lo1 = g.lo[1], lo2 = g.lo[2];
  const float h0 = g.h[0], h1 = g.h[1], h2 = g.h[2];
  const float ih0 = g.inv_h[0], ih1 = g.inv_h[1], ih2 = g.inv_h[2];
  if (UNIFIED || R >= 1) {
    R = R < 0 ? 0 : R;
    const int hx = sx > 0 ? min(cx + R, nx - 1) : max(cx - R, 0);
    const int hy = sy > 0 ? min(cy + R, ny - 1) : max(cy - R, 0);
    const int hz = sz > 0 ? min(cz + R, nz - 1) : max(cz - R, 0);
    const float bx = walk_boundary(lo0, h0, hx, sx, o.x, ix);
    const float by = walk_boundary(lo1, h1, hy, sy, o.y, iy);
    const float bz = walk_boundary(lo2, h2, hz, sz, o.z, iz);
    const bool ex = (bx <= by) && (bx <= bz);
    const bool ey = !ex && (by <= bz);
    const float tb = ex ? bx : (ey ? by : bz);
    if (!(tb < 3.4e38f)) return false;
    const float de = tb * dlen;
    if (best_f >= 0 && best_d2 < de * de * 0.9999f) return false;
    const float px = o.x + dir.x * tb, py = o.y + dir.y * tb, pz = o.z + dir.z * tb;
    int nxc = min(max((int)floorf((px - lo0) * ih0), max(cx - R, 0)), min(cx + R, nx - 1));
    int nyc = min(max((int)floorf((py - lo1) * ih1), max(cy - R, 0)), min(cy + R, ny - 1));
    int nzc = min(max((int)floorf((pz - lo2) * ih2), max(cz - R, 0)), min(cz + R, nz - 1));
    nxc = sx > 0 ? max(nxc, cx) : (sx < 0 ? min(nxc, cx) : cx);
    nyc = sy > 0 ? max(nyc, cy) : (sy < 0 ? min(nyc, cy) : cy);
    nzc = sz > 0 ? max(nzc, cz) : (sz < 0 ? min(nzc, cz) : cz);
    // (selects, not branches: a branch per axis is turned into an indexed
    // store to a scratch array)
    nxc = ex ? hx + sx : nxc;
    nyc = (!ex && ey) ? hy + sy : nyc;
    nzc = (!ex && !ey) ? hz + sz : nzc;
    if (nxc < 0 || nxc >= nx || nyc < 0 || nyc >= ny || nzc < 0 || nzc >= nz) return false;
    cx = nxc;
    cy = nyc;
    cz = nzc;
    tnx = walk_boundary(lo0, h0, cx, sx, o.x, ix);
    tny = walk_boundary(lo1, h1, cy, sy, o.y, iy);
    tnz = walk_boundary(lo2, h2, cz, sz, o.z, iz);
    cell = cx + nx * cy + nxy * cz;
    return true;
  }
  const bool ax = (tnx <= tny) && (tnx <= tnz);
  const bool ay = !ax && (tny <= tnz);
  const float t_exit = ax ? tnx : (ay ? tny : tnz);
  const float de = t_exit * dlen;
  if (best_f >= 0 && best_d2 < de * de * 0.9999f) return false;
  const int st = ax ? sx : (ay ? sy : sz);
  if (st == 0) return false;
  // One axis moves; its operands are selected rather than branched on (a
  // branch per axis is turned into an indexed store to a scratch array).
  const int c_new = (ax ? cx : (ay ? cy : cz)) + st;
  if (c_new < 0 || c_new >= (ax ? nx : (ay ? ny : nz))) return false;
  const float tb_new = walk_boundary(ax ? lo0 : (ay ? lo1 : lo2), ax ? h0 : (ay ? h1 : h2), c_new, st,
                                     ax ? o.x : (ay ? o.y : o.z), ax ? ix : (ay ? iy : iz));
  cx = ax ? c_new : cx;
  cy = (!ax && ay) ? c_new : cy;
  cz = (!ax && !ay) ? c_new : cz;
  tnx = ax ? tb_new : tnx;
  tny = (!ax && ay) ? tb_new : tny;
  tnz = (!ax && !ay) ? tb_new : tnz;
  cell += ax ? st : (ay ? st * nx : st * nxy);
  return true;
}

// The nearest-hit rule of `consider` as one 64-bit key: the squared distance's
// bits (non-negative floats order as unsigned integers) above the face index,
// so the smallest key is the nearest hit with ties to the lower face.
__device__ inline unsigned long long hit_key(float d2, int f) {
  return ((unsigned long long)__float_as_uint(d2) << 32) | (uint32_t)f;
}

constexpr unsigned long long kNoHit = ~0ull;  // above every hit_key

__device__ inline void decode_key(unsigned long long key, float& d2, int& f) {
  f = key == kNoHit ? -1 : (int)(uint32_t)key;
  d2 = key == kNoHit ? 3.402823466e38f : __uint_as_float((uint32_t)(key >> 32));
}

// The walk of pooled_walk, on per-ray coefficients and in mirrored cell
// coordinates: along an axis where the ray moves toward lower cells, cell c
// is walked as c' = n - 1 - c, so on every axis the ray moves toward higher
// c'. The ray parameter of the far boundary of cell c' is c'*A + B (A >= 0),
// and the ray's mirrored grid coordinate at parameter t is P + t*D (D >= 0),
// so a boundary or an exit cell costs one fused multiply-add, and the exit's
// clamps and range checks need no direction signs. The grid cell of (c'x,
// c'y, c'z) is c0 + gx*c'x + gy*c'y + gz*c'z (signed strides per ray). For a
// zero direction component A = 0, B = 3.4e38 (the ray crosses no face of that
// axis) and P is the middle of the ray's cell. These values round
// differently from walk_leave's expressions by a few units in the last place;
// the walk tolerates that: faces are binned into cells grown by 1e-3 of a
// cell, and the walk stops only with a 1e-4 relative margin on the exit
// distance. Returns false when the ray misses the grid box.
__device__ inline bool coef_begin(const grid_desc& g, v3 o, v3 dir, float& ax, float& ay, float& az, float& bx,
                                  float& by, float& bz, float& px, float& py, float& pz, float& dx, float& dy,
                                  float& dz, float& dlen, int& cx, int& cy, int& cz, int& c0, int& gx, int& gy,
                                  int& gz) {
  // Hardware reciprocals and square root (about 1 ulp): these values only
  // place the walk, which tolerates far larger errors (see above).
  const float ix = dir.x != 0.f ? __builtin_amdgcn_rcpf(dir.x) : 0.f;
  const float iy = dir.y != 0.f ? __builtin_amdgcn_rcpf(dir.y) : 0.f;
  const float iz = dir.z != 0.f ? __builtin_amdgcn_rcpf(dir.z) : 0.f;
  // Clip to the grid box (slabs); t is in units of dir.
  float t0 = 0.f, t1 = 3.4e38f;
  if (dir.x != 0.f) {
    const float ta = (g.lo[0] - o.x) * ix, tb = (g.hi[0] - o.x) * ix;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.x < g.lo[0] || o.x > g.hi[0]) {
    return false;
  }
  if (dir.y != 0.f) {
    const float ta = (g.lo[1] - o.y) * iy, tb = (g.hi[1] - o.y) * iy;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.y < g.lo[1] || o.y > g.hi[1]) {
    return false;
  }
  if (dir.z != 0.f) {
    const float ta = (g.lo[2] - o.z) * iz, tb = (g.hi[2] - o.z) * iz;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.z < g.lo[2] || o.z > g.hi[2]) {
    return false;
  }
  if (!(t0 <= t1)) return false;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const bool fx = dir.x < 0.f, fy = dir.y < 0.f, fz = dir.z < 0.f;
  // Mirrored grid coordinates of the origin, measured from the face the ray
  // moves away from, and the entry cell.
  px = (fx ? g.hi[0] - o.x : o.x - g.lo[0]) * g.inv_h[0];
  py = (fy ? g.hi[1] - o.y : o.y - g.lo[1]) * g.inv_h[1];
  pz = (fz ? g.hi[2] - o.z : o.z - g.lo[2]) * g.inv_h[2];
  dx = fabsf(dir.x) * g.inv_h[0];
  dy = fabsf(dir.y) * g.inv_h[1];
  dz = fabsf(dir.z) * g.inv_h[2];
  cx = min(max((int)floorf(fmaf(t0, dx, px)), 0), nx - 1);
  cy = min(max((int)floorf(fmaf(t0, dy, py)), 0), ny - 1);
  cz = min(max((int)floorf(fmaf(t0, dz, pz)), 0), nz - 1);
  gx = fx ? -1 : 1;
  gy = fy ? -nx : nx;
  gz = fz ? -nx * ny : nx * ny;
  c0 = (fx ? nx - 1 : 0) + (fy ? (ny - 1) * nx : 0) + (fz ? (nz - 1) * nx * ny : 0);
  // A cell is one unit of the grid coordinate: the far boundary of cell c'
  // is reached at t = (c' + 1 - P) * A.
  ax = fabsf(g.h[0] * ix);
  ay = fabsf(g.h[1] * iy);
  az = fabsf(g.h[2] * iz);
  bx = dir.x != 0.f ? fmaf(-px, ax, ax) : 3.4e38f;
  by = dir.y != 0.f ? fmaf(-py, ay, ay) : 3.4e38f;
  bz = dir.z != 0.f ? fmaf(-pz, az, az) : 3.4e38f;
  if (dir.x == 0.f) px = (float)cx + 0.5f;
  if (dir.y == 0.f) py = (float)cy + 0.5f;
  if (dir.z == 0.f) pz = (float)cz + 0.5f;
  dlen = __builtin_amdgcn_sqrtf(dot(dir, dir));
  return true;
}

// Leaves the current cell: past the empty box [c', c' + R] (mirrored
// coordinates; R >= 1) or to the next cell of the DDA (R = 0: the box is the
// cell itself). The exit cell's other two coordinates come from the exit
// point, clamped to the box. BOX = false assumes R = 0 for the lane and skips
// the exit point. Returns false when the walk is over (past the grid, or the
// best hit is nearer than the box's exit).
template <bool BOX>
__device__ inline bool coef_leave(const grid_desc& g, float ax, float ay, float az, float bx, float by, float bz,
                                  float px, float py, float pz, float dx, float dy, float dz, float dlen, int R,
                                  int& cx, int& cy, int& cz, float best_d2, int best_f) {
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  if (!BOX) R = 0;
  const int hx = min(cx + R, nx - 1), hy = min(cy + R, ny - 1), hz = min(cz + R, nz - 1);
  const float tx = fmaf((float)hx, ax, bx), ty = fmaf((float)hy, ay, by), tz = fmaf((float)hz, az, bz);
  const bool ex = (tx <= ty) && (tx <= tz);
  const bool ey = !ex && (ty <= tz);
  const float tb = ex ? tx : (ey ? ty : tz);
  const float de = tb * dlen;
  bool on = tb < 3.4e38f && !(best_f >= 0 && best_d2 < de * de * 0.9999f);
  int nxc = cx, nyc = cy, nzc = cz;
  if (BOX) {
    // the exit point's cell, clamped to [c', h'] on each axis
    nxc = min(max((int)floorf(fmaf(tb, dx, px)), cx), hx);
    nyc = min(max((int)floorf(fmaf(tb, dy, py)), cy), hy);
    nzc = min(max((int)floorf(fmaf(tb, dz, pz)), cz), hz);
  }
  const int he = ex ? hx : (ey ? hy : hz);
  const int ne = ex ? nx : (ey ? ny : nz);
  nxc = ex ? hx + 1 : nxc;
  nyc = (!ex && ey) ? hy + 1 : nyc;
  nzc = (!ex && !ey) ? hz + 1 : nzc;
  on = on && he + 1 < ne;
  cx = on ? nxc : cx;
  cy = on ? nyc : cy;
  cz = on ? nzc : cz;
  return on;
}

// LDS words per wavefront of pooled_walk: 64 best-hit words (hit_key).
constexpr int kPoolWords = 64;

// Walk mode bits (ffk_set_walk): kWalkJump = the jump code for every empty
// cell (template WALK 1, else WALK 2); kWalkWhole = pooled tests dealt in
// whole chunks, passed to pooled_walk as kPoolWhole in its pool argument.
constexpr int kWalkJump = FFK_WALK_JUMP, kWalkWhole = FFK_WALK_WHOLE, kWalkStaged = FFK_WALK_STAGED;
constexpr int kPoolWhole = 0x100, kPoolStaged = 0x200;
constexpr int kDefaultWalk = FFK_WALK_DEFAULT;
__host__ __device__ inline int pool_arg(int pool, int walk) {
  return pool | ((walk & kWalkWhole) ? kPoolWhole : 0) | ((walk & kWalkStaged) ? kPoolStaged : 0);
}

// hit_terms on a face's terms in memory, loaded in two stages: the plane
// terms (p0, uu | n, 1/D) first, the edge terms (u, uv | v, vv) only in the
// lanes whose ray meets the face's plane in front of it. The arithmetic and
// the outcome are hit_terms' (the hit point itself is not kept).
__device__ inline bool hit_staged(const float4* q, v3 o, v3 dir, float* d2) {
  const float4 q0 = gld4(q), q3 = gld4(q + 3);
  const v3 p0 = mk(q0.x, q0.y, q0.z), n = mk(q3.x, q3.y, q3.z);
  const v3 w0 = o - p0;
  const float a = -dot(n, w0);
  const float b = dot(n, dir);
  if (fabsf(b) <= kEps8) return false;
  const float r = a / b;
  if (r < 0.0f) return false;
  const float4 q1 = gld4(q + 1), q2 = gld4(q + 2);
  const v3 u = mk(q1.x, q1.y, q1.z), v = mk(q2.x, q2.y, q2.z);
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
  *d2 = dot(sd, sd);
  return true;
}

// Inclusive prefix sum over the wavefront's 64 lanes: DPP row shifts within
// each row of 16 lanes (lanes shifted in from outside a row read 0), then
// row_bcast:15 adds lane 15 of rows 0 and 2 to rows 1 and 3, and
// row_bcast:31 adds lane 31 to rows 2 and 3. Every lane must be active.
__device__ inline int wave_incl_sum(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

// Inclusive prefix maximum over the wavefront's 64 lanes of non-negative
// values: DPP row shifts within each row of 16 lanes, then row_bcast:15
// (lane 15 of each row into the next row, rows 1 and 3) and row_bcast:31
// (lane 31 into rows 2 and 3). Every lane must be active.
__device__ inline int wave_incl_max(int x) {
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));  // row_shr:1
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));  // row_shr:2
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));  // row_shr:4
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));  // row_shr:8
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return x;
}

// Casts rays 0..n-1 of a work list with a wavefront's 64 lanes working as a
// pool. Each lane walks its own ray one grid cell per iteration; a lane that
// reaches an occupied cell waits, and once waiting lanes x `pool` >= active
// lanes the wavefront spreads the (ray, reference) pairs of all waiting
// lanes over its 64 lanes: each lane tests one pair and folds the result
// into the owning ray's best hit with a 64-bit LDS atomic minimum on hit_key
// (the rule of `consider`). Once at least `refill` lanes are idle, they take
// the next rays of the list (one LDS atomic on *next per refill). The hits
// are those of cast_grid, whatever the order. Lanes in an empty cell leave
// it by the jump past its empty box or, next to occupied cells, by one DDA
// step (coef_leave); WALK 2 runs the jump code for all of a wavefront's
// empty-cell lanes when any of them can jump and the DDA step otherwise
// (never both), WALK 1 always runs the jump code (the same cells).
//   best:  this wavefront's kPoolWords words of LDS (8-byte aligned)
//   next:  the list's LDS ray counter, zero at the start, shared by the
//          block's wavefronts
//   pool:  the pooling ratio in the low byte; with kPoolWhole set, a round of
//          at least 64 pairs deals whole chunks of 64 only, and the lanes
//          whose references lie past the last whole chunk stay waiting with
//          the rest for the next round (their best hit so far stays in LDS)
//   start(i, o, dir):            sets up ray i
//   finish(i, o, dir, key):      ray i is done; key = hit_key of its hit or kNoHit
//                                (called by the lanes whose rays end)
//   after(last):                 called by every lane of the wavefront together
//                                at the end of each iteration (last = false)
//                                and once after the last one (last = true)
//   prep(base, idle):            called by every lane together before a refill
//                                starts rays base + (rank of the lane among the
//                                idle lanes) in the idle lanes
// Every lane of the wavefront must call this together.
struct walk_no_after {
  __device__ void operator()(bool) const {}
};
struct walk_no_prep {
  __device__ void operator()(int, unsigned long long) const {}
};
template <bool STATS, int WALK, class Start, class Finish, class After = walk_no_after, class Prep = walk_no_prep>
__device__ inline void pooled_walk(const scene_desc& S, unsigned long long* best, int* next, long long n, int refill,
                                   int pool, Start start, Finish finish, uint32_t& n_tests, uint32_t& n_cells,
                                   After after = After(), Prep prep = Prep()) {
  const int lane = threadIdx.x & 63;
  const bool whole = (pool & kPoolWhole) != 0, staged = (pool & kPoolStaged) != 0;
  pool &= 0xff;
  float d2 = 0.f;
  bool have = false, more = true, pend = false;  // `more` is wave-uniform
  long long cur = 0;
  int ref0 = 0, nref = 0;
  v3 o = mk(0.f, 0.f, 0.f), dir = mk(0.f, 0.f, 0.f);
  float ax = 0.f, ay = 0.f, az = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
  float px = 0.f, py = 0.f, pz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, dlen = 0.f;
  int cx = 0, cy = 0, cz = 0, c0 = 0, gx = 0, gy = 0, gz = 0, osh = 0;
  while (true) {
    const unsigned long long idle = __ballot(!have);
    const int n_idle = __popcll(idle);
    if (more && n_idle >= refill) {
      int base = 0;
      if (lane == 0) base = atomicAdd(next, n_idle);
      base = __shfl(base, 0, 64);
      if ((long long)base + n_idle >= n) more = false;
      prep(base, idle);
      if (!have) {
        const int rank =
            (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
        const long long i = (long long)base + rank;
        if (i < n) {
          cur = i;
          start(i, o, dir);
          osh = octant_shift(dir);
          best[lane] = kNoHit;
          pend = false;
          have = coef_begin(S.g, o, dir, ax, ay, az, bx, by, bz, px, py, pz, dx, dy, dz, dlen, cx, cy, cz, c0, gx,
                            gy, gz);
          if (!have) finish(i, o, dir, kNoHit);
        }
      }
    }
    if (!more && __ballot(have) == 0ull) break;
    // Lanes that are not waiting for tests advance by one cell.
    int2 ci = make_int2(0, 0);
    const bool step = have && !pend;
    if (step) {
      ci = gld2(gidx(S.cell_info, (uint32_t)(c0 + gx * cx + gy * cy + gz * cz)));
      if (STATS) n_cells += 1;
      if (ci.y > 0) {
        pend = true;
        ref0 = ci.x;
        nref = ci.y;
      }
    }
    // Lanes in an empty cell leave it below, together with the lanes whose
    // pooled tests finish in this iteration: one exit computation per
    // iteration for all of them (R = 0 for a tested cell: the box code then
    // takes exactly the DDA step).
    bool leave = step && ci.y <= 0;  // then ci.x holds the empty boxes
    int R = leave ? empty_box(ci.x, osh) - 1 : 0;
    const unsigned long long waiting = __ballot(have && pend);
    const unsigned long long active = __ballot(have);
    if (waiting != 0ull && __popcll(waiting) * pool >= __popcll(active)) {
      // The waiting lanes' (ray, reference) pairs, numbered in lane order:
      // lane w owns positions [excl, incl) of the prefix sum of the counts.
      const int cnt = (have && pend) ? nref : 0;
      const int incl = wave_incl_sum(cnt);
      const int excl = incl - cnt;
      const int total = __builtin_amdgcn_readlane(incl, 63);
      const int base = ref0 - excl;  // reference index of position q of this lane's range: base + q
      const int deal = whole && total >= 64 ? (total & ~63) : total;  // pairs tested in this round
      for (int p0 = 0; p0 < deal; p0 += 64) {
        const int p = p0 + lane;
        // Owner of position p: the waiting lane whose range [excl, incl)
        // holds it. Each lane whose range starts inside this chunk sends
        // lane + 1 to the chunk position of its start (ds_permute: a forward
        // permute, where the highest sending lane wins and a position no lane
        // sends to reads 0); the lane holding p0 and every other lane send
        // the same value, that lane + 1, to position 0. A prefix maximum then
        // carries each start's lane over its range. Lanes past the last range
        // get its lane and test nothing (p >= deal).
        const int first = __builtin_ctzll(__ballot(incl > p0));  // wave-uniform
        const bool starts = cnt > 0 && excl > p0 && excl < p0 + 64;
        const int sent = __builtin_amdgcn_ds_permute((starts ? excl - p0 : 0) << 2, starts ? lane + 1 : first + 1);
        const int owner = wave_incl_max(sent) - 1;
        const int ref_o = __shfl(base, owner, 64) + p;
        const v3 oo = mk(__shfl(o.x, owner, 64), __shfl(o.y, owner, 64), __shfl(o.z, owner, 64));
        const v3 dd = mk(__shfl(dir.x, owner, 64), __shfl(dir.y, owner, 64), __shfl(dir.z, owner, 64));
        if (p < deal) {
          const int f = gld(gidx(S.refs, (uint32_t)ref_o));
          const float4* q = reinterpret_cast<const float4*>(gidx(S.tri, (uint32_t)f));
          if (staged) {
            if (hit_staged(q, oo, dd, &d2) && d2 < 3.402823466e38f) atomicMin(&best[owner], hit_key(d2, f));
          } else {
            v3 pt;
            if (hit_terms(gld4(q), gld4(q + 1), gld4(q + 2), gld4(q + 3), oo, dd, &pt, &d2) && d2 < 3.402823466e38f)
              atomicMin(&best[owner], hit_key(d2, f));
          }
          if (STATS) n_tests += 1;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (have && pend) {
        if (incl <= deal) {  // every reference of the cell tested
          pend = false;
          leave = true;
          R = 0;
        } else if (excl < deal) {  // the rest waits for the next round
          ref0 += deal - excl;
          nref -= deal - excl;
        }
      }
    }
    const bool jump = WALK == 1 || __ballot(leave && R >= 1) != 0ull;
    if (leave) {
      const unsigned long long key = best[lane];
      float bd;
      int bf;
      decode_key(key, bd, bf);
      const bool on = jump ? coef_leave<true>(S.g, ax, ay, az, bx, by, bz, px, py, pz, dx, dy, dz, dlen, R, cx, cy,
                                              cz, bd, bf)
                           : coef_leave<false>(S.g, ax, ay, az, bx, by, bz, px, py, pz, dx, dy, dz, dlen, 0, cx, cy,
                                               cz, bd, bf);
      if (!on) {
        finish(cur, o, dir, key);
        have = false;
      }
    }
    after(false);
  }
  after(true);
}

// One ray, walked to its end in a single loop (the same steps as
// walk_begin + walk_cell). Callers that carry one ray per lane use this form:
// written as begin + a loop over walk_cell, the one-ray-per-lane kernels ran
// at half the speed on gfx950; the interleaved (refill) scheduling uses the
// split form. Both give the same hits (tested against the oracle).
template <bool STATS>
__device__ inline void cast_grid(const scene_desc& S, v3 o, v3 dir, hit_rec& h, uint32_t& n_tests,
                                 uint32_t& n_cells) {
  const grid_desc& g = S.g;
  // Clip to the grid box (slabs); t is in units of dir.
  float t0 = 0.f, t1 = 3.4e38f;
  float ix = 0.f, iy = 0.f, iz = 0.f;
  if (dir.x != 0.f) {
    ix = 1.0f / dir.x;
    float ta = (g.lo[0] - o.x) * ix, tb = (g.hi[0] - o.x) * ix;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.x < g.lo[0] || o.x > g.hi[0]) {
    return;
  }
  if (dir.y != 0.f) {
    iy = 1.0f / dir.y;
    float ta = (g.lo[1] - o.y) * iy, tb = (g.hi[1] - o.y) * iy;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.y < g.lo[1] || o.y > g.hi[1]) {
    return;
  }
  if (dir.z != 0.f) {
    iz = 1.0f / dir.z;
    float ta = (g.lo[2] - o.z) * iz, tb = (g.hi[2] - o.z) * iz;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
  } else if (o.z < g.lo[2] || o.z > g.hi[2]) {
    return;
  }
  if (!(t0 <= t1)) return;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  int cx = min(max((int)floorf((o.x + dir.x * t0 - g.lo[0]) * g.inv_h[0]), 0), nx - 1);
  int cy = min(max((int)floorf((o.y + dir.y * t0 - g.lo[1]) * g.inv_h[1]), 0), ny - 1);
  int cz = min(max((int)floorf((o.z + dir.z * t0 - g.lo[2]) * g.inv_h[2]), 0), nz - 1);
  const int sx = dir.x > 0.f ? 1 : (dir.x < 0.f ? -1 : 0);
  const int sy = dir.y > 0.f ? 1 : (dir.y < 0.f ? -1 : 0);
  const int sz = dir.z > 0.f ? 1 : (dir.z < 0.f ? -1 : 0);
  // Ray parameter of the next cell boundary along each axis.
  auto bx_ = [&](int c) { return sx ? (g.lo[0] + (float)(c + (sx > 0)) * g.h[0] - o.x) * ix : 3.4e38f; };
  auto by_ = [&](int c) { return sy ? (g.lo[1] + (float)(c + (sy > 0)) * g.h[1] - o.y) * iy : 3.4e38f; };
  auto bz_ = [&](int c) { return sz ? (g.lo[2] + (float)(c + (sz > 0)) * g.h[2] - o.z) * iz : 3.4e38f; };
  float tnx = bx_(cx), tny = by_(cy), tnz = bz_(cz);
  // Distances along the ray are compared as (t * |dir|)^2 so that neither
  // factor overflows for very short or very long direction vectors.
  const float dlen = sqrtf(dot(dir, dir));
  const int nxy = nx * ny;
  int cell = cx + nx * cy + nxy * cz;
  const int osh = octant_shift(dir);
  mailbox mb;
  mb.clear();
  while (true) {
    const int2 ci = gld2(S.cell_info + cell);
    if (STATS) n_cells += 1;
    for (int k = ci.x; k < ci.x + ci.y; ++k) {
      const uint32_t t = test_ref(S, k, o, dir, h, mb);
      if (STATS) n_tests += t;
    }
    const int R = ci.y > 0 ? 0 : empty_box(ci.x, osh) - 1;
    if (R >= 1) {
      // The box [c, c + s*R] toward the ray's octant (clipped to the grid)
      // is empty: leave it in one step.
      const int hx = sx > 0 ? min(cx + R, nx - 1) : max(cx - R, 0);
      const int hy = sy > 0 ? min(cy + R, ny - 1) : max(cy - R, 0);
      const int hz = sz > 0 ? min(cz + R, nz - 1) : max(cz - R, 0);
      const float bx = bx_(hx), by = by_(hy), bz = bz_(hz);
      const bool ex = (bx <= by) && (bx <= bz);
      const bool ey = !ex && (by <= bz);
      const float tb = ex ? bx : (ey ? by : bz);
      if (!(tb < 3.4e38f)) break;
      const float de = tb * dlen;
      if (h.f >= 0 && h.d2 < de * de * 0.9999f) break;
      // Cell just past the exit face; the other two coordinates come from
      // the exit point, kept inside the box and never moved backwards.
      const float px = o.x + dir.x * tb, py = o.y + dir.y * tb, pz = o.z + dir.z * tb;
      int nxc = min(max((int)floorf((px - g.lo[0]) * g.inv_h[0]), max(cx - R, 0)), min(cx + R, nx - 1));
      int nyc = min(max((int)floorf((py - g.lo[1]) * g.inv_h[1]), max(cy - R, 0)), min(cy + R, ny - 1));
      int nzc = min(max((int)floorf((pz - g.lo[2]) * g.inv_h[2]), max(cz - R, 0)), min(cz + R, nz - 1));
      nxc = sx > 0 ? max(nxc, cx) : (sx < 0 ? min(nxc, cx) : cx);
      nyc = sy > 0 ? max(nyc, cy) : (sy < 0 ? min(nyc, cy) : cy);
      nzc = sz > 0 ? max(nzc, cz) : (sz < 0 ? min(nzc, cz) : cz);
      if (ex) nxc = hx + sx;
      else if (ey) nyc = hy + sy;
      else nzc = hz + sz;
      if (nxc < 0 || nxc >= nx || nyc < 0 || nyc >= ny || nzc < 0 || nzc >= nz) break;
      cx = nxc;
      cy = nyc;
      cz = nzc;
      tnx = bx_(cx);
      tny = by_(cy);
      tnz = bz_(cz);
      cell = cx + nx * cy + nxy * cz;
      continue;
    }
    const bool ax = (tnx <= tny) && (tnx <= tnz);
    const bool ay = !ax && (tny <= tnz);
    const float t_exit = ax ? tnx : (ay ? tny : tnz);
    const float de = t_exit * dlen;
    if (h.f >= 0 && h.d2 < de * de * 0.9999f) break;
    const int st = ax ? sx : (ay ? sy : sz);
    if (st == 0) break;  // zero or NaN direction: the start cell was all
    if (ax) {
      cx += sx;
      if (cx < 0 || cx >= nx) break;
      cell += sx;
      tnx = bx_(cx);
    } else if (ay) {
      cy += sy;
      if (cy < 0 || cy >= ny) break;
      cell += sy * nx;
      tny = by_(cy);
    } else {
      cz += sz;
      if (cz < 0 || cz >= nz) break;
      cell += sz * nxy;
      tnz = bz_(cz);
    }
  }
}

// First face f with face_off[f+1] > g, starting the search at `hint`.
__device__ inline int find_face(const long long* __restrict__ off, int n_faces, long long g, int hint) {
  const long long oh = gld(gidx(off, (uint32_t)hint));
  if (oh <= g && g < gld(gidx(off, (uint32_t)hint + 1u))) return hint;
  int lo = (oh <= g) ? hint + 1 : 0;
  int hi = (oh <= g) ? n_faces - 1 : hint;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (gld(gidx(off, (uint32_t)mid + 1u)) > g) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// find_face over the faces bounds[0]..bounds[1] only, which hold ray g
// (off[bounds[0]] <= g < off[bounds[1] + 1]); hint lies between them. Returns
// the same face as find_face: the first f with off[f + 1] > g. A block whose
// rays span a few faces searches those instead of the scene's whole table
// when the hint misses. The bounds are read (from LDS) only on a miss, so
// they hold no registers through the caller's ray loop.
__device__ inline int find_face_in(const long long* __restrict__ off, const int* bounds, long long g, int hint) {
  int lo, hi;
  if (gld(gidx(off, (uint32_t)hint)) <= g) {
    if (g < gld(gidx(off, (uint32_t)hint + 1u))) return hint;
    lo = hint + 1;
    hi = bounds[1];
  } else {
    lo = bounds[0];
    hi = hint - 1;
  }
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (gld(gidx(off, (uint32_t)mid + 1u)) > g) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Malley ray j of face f (emission/malley_cpu.cpp:109-186): uniform point on
// the face, cosine-weighted direction about the face frame, rear rays with
// the z component negated, origin pushed 1e-4 along the direction. The four
// uniforms come from one Philox4x32-10 block, counter (j, f, 0, 0).
__device__ inline void malley_ray(const scene_desc& S, int f, int j, uint32_t k0, uint32_t k1,
                                  v3* origin_out, v3* end_out) {
  // The key is wave-uniform. Made opaque here, its ten round keys are
  // derived with scalar adds at each call instead of being hoisted out of
  // the caller's loop into registers that spill to vector-register lanes.
  asm volatile("" : "+s"(k0), "+s"(k1));
  u4 r = philox(u4{(uint32_t)j, (uint32_t)f, 0u, 0u}, k0, k1);
  float a = u01(r.x), b = u01(r.y), rr = u01(r.z);
  float sphi, cphi;
  sincos_turns(u01(r.w), &sphi, &cphi);
  if (a + b > 1) {
    a = 1.f - a;
    b = 1.f - b;
  }
  const float4* q = reinterpret_cast<const float4*>(gidx(S.tri, (uint32_t)f));
  float4 q0 = gld4(q), q1 = gld4(q + 1), q2 = gld4(q + 2);
  v3 p0 = mk(q0.x, q0.y, q0.z), v0 = mk(q1.x, q1.y, q1.z), v1 = mk(q2.x, q2.y, q2.z);
  v3 origin = (p0 + v0 * a) + v1 * b;
  float rad = sqrtf(rr);
  v3 m = mk(rad * cphi, rad * sphi, sqrtf(1 - rr));
  if (j >= gld(gidx(S.front, (uint32_t)f))) m.z = -m.z;
  float R[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) R[i] = gld(gidx(S.rot, 9u * (uint32_t)f + (uint32_t)i));
  v3 d = mat_vec(R, m);
  *origin_out = origin + d * 0.0001f;
  *end_out = origin + d;
}

}  // namespace ffk
