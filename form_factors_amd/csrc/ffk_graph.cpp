// Mesh adjacency for conduction (host side, set-up time).
//
// thermal_equation/conductive_cpu.cpp:83-141 walks every pair of faces that
// share an edge (subject/mesh_graph.cpp:116-279, each unordered pair once)
// and, for faces of different meshes, exchanges
//   flow = (T_r - T_l) / R,   R = (d_l / k_l + d_r / k_r) / (min(thickness) * edge)
// where d is the distance from the shared edge's midpoint to each mesh's
// centre (mean of its face centres). Here each such pair contributes
// G = 1 / R to a symmetric CSR operator over meshes; the device applies
// flow_m = sum_k G_mk (T_k - T_m) in FP64 (ffk_conduction).
//
// Vertices are matched with the reference's tolerance: two points are the
// same when every coordinate differs by at most FLT_EPSILON relative to the
// larger magnitude (math/operations.cpp:22-38).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "ffk.h"

namespace {

struct p3 {
  float x, y, z;
};

inline bool same_coord(float a, float b) {
  const float size = std::max(std::fabs(a), std::fabs(b));
  if (size == 0.f) return true;
  return std::fabs((a - b) / size) <= FLT_EPSILON;
}

// Union of vertices that the reference's fuzzy comparison treats as equal,
// found by sorting on x and comparing within the tolerance window.
std::vector<int> unify(const std::vector<p3>& v) {
  const int n = (int)v.size();
  std::vector<int> order(n), id(n, -1);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return v[a].x < v[b].x; });
  int next = 0;
  for (int a = 0; a < n; ++a) {
    const int i = order[a];
    if (id[i] >= 0) continue;
    id[i] = next;
    for (int b = a + 1; b < n; ++b) {
      const int j = order[b];
      const float span = std::max(std::fabs(v[i].x), std::fabs(v[j].x)) * 2 * FLT_EPSILON;
      if (v[j].x - v[i].x > span && v[j].x - v[i].x > 0) break;
      if (id[j] < 0 && same_coord(v[i].x, v[j].x) && same_coord(v[i].y, v[j].y) && same_coord(v[i].z, v[j].z))
        id[j] = next;
    }
    ++next;
  }
  return id;
}

inline p3 sub(p3 a, p3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline p3 add(p3 a, p3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float norm(p3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

}  // namespace

extern "C" int ffk_conduction_graph(const float* h_faces, int n_faces, const int* h_meshes, int n_meshes,
                                    const float* h_conductivity, const float* h_thickness, int* h_row_ptr,
                                    int* h_cols, double* h_G, int capacity, int* nnz_out) {
  if (!h_faces || !h_meshes || !h_conductivity || !h_thickness || !nnz_out || n_faces < 0 || n_meshes < 0)
    return FFK_ERR_ARG;
  std::vector<int> f2m(n_faces, -1);
  for (int m = 0; m < n_meshes; ++m)
    for (int f = 0; f < h_meshes[3 * m + 1]; ++f) {
      const int fi = h_meshes[3 * m] + f;
      if (fi < 0 || fi >= n_faces) return FFK_ERR_ARG;
      f2m[fi] = m;
    }
  // mesh centres: mean of face centres, in float (conductive_cpu.cpp:72-80)
  std::vector<p3> centre(n_meshes, p3{0, 0, 0});
  auto face_centre = [&](int f) {
    const float* p = h_faces + 9 * (size_t)f;
    p3 s = add(add(p3{p[0], p[1], p[2]}, p3{p[3], p[4], p[5]}), p3{p[6], p[7], p[8]});
    return p3{s.x / 3.f, s.y / 3.f, s.z / 3.f};
  };
  for (int m = 0; m < n_meshes; ++m) {
    const int first = h_meshes[3 * m], cnt = h_meshes[3 * m + 1];
    if (cnt <= 0) continue;
    p3 c = face_centre(first);
    for (int f = 1; f < cnt; ++f) c = add(c, face_centre(first + f));
    const float inv = (float)cnt;
    centre[m] = p3{c.x / inv, c.y / inv, c.z / inv};
  }
  std::vector<p3> verts(3 * (size_t)n_faces);
  for (int f = 0; f < n_faces; ++f)
    for (int k = 0; k < 3; ++k) {
      const float* p = h_faces + 9 * (size_t)f + 3 * k;
      verts[3 * (size_t)f + k] = p3{p[0], p[1], p[2]};
    }
  const std::vector<int> vid = unify(verts);
  // edges keyed by unified vertex ids -> faces using them
  std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> edges;  // key -> (face, edge k)
  for (int f = 0; f < n_faces; ++f)
    for (int k = 0; k < 3; ++k) {
      int a = vid[3 * (size_t)f + k], b = vid[3 * (size_t)f + (k + 1) % 3];
      if (a == b) continue;
      edges[std::make_pair(std::min(a, b), std::max(a, b))].push_back(std::make_pair(f, k));
    }
  std::map<std::pair<int, int>, double> G;  // (mesh, mesh) -> conductance, both directions
  for (auto& kv : edges) {
    auto& fs = kv.second;
    for (size_t i = 0; i < fs.size(); ++i)
      for (size_t j = i + 1; j < fs.size(); ++j) {
        int l = fs[i].first, r = fs[j].first, kl = fs[i].second;
        if (l > r) {
          std::swap(l, r);
          kl = fs[j].second;
        }
        const int ml = f2m[l], mr = f2m[r];
        if (ml < 0 || mr < 0 || ml == mr) continue;
        const float* lp = h_faces + 9 * (size_t)l;
        const p3 a{lp[3 * kl], lp[3 * kl + 1], lp[3 * kl + 2]};
        const int kb = (kl + 1) % 3;
        const p3 b{lp[3 * kb], lp[3 * kb + 1], lp[3 * kb + 2]};
        const float edge = norm(sub(a, b));
        const p3 s = add(a, b);
        const p3 median{s.x / 2.f, s.y / 2.f, s.z / 2.f};
        const float dl = norm(sub(median, centre[ml]));
        const float dr = norm(sub(median, centre[mr]));
        const float thickness = std::fmin(h_thickness[ml], h_thickness[mr]);
        const float area = thickness * edge;
        const float R = (dl / h_conductivity[ml] + dr / h_conductivity[mr]) / area;
        const double g = 1.0 / (double)R;
        G[std::make_pair(ml, mr)] += g;
        G[std::make_pair(mr, ml)] += g;
      }
  }
  const int nnz = (int)G.size();
  *nnz_out = nnz;
  if (!h_row_ptr) return FFK_OK;  // size query
  if (capacity < nnz || !h_cols || !h_G) return FFK_ERR_LIMIT;
  std::fill(h_row_ptr, h_row_ptr + n_meshes + 1, 0);
  int k = 0;
  for (auto& kv : G) {  // ordered by (row, col)
    h_row_ptr[kv.first.first + 1]++;
    h_cols[k] = kv.first.second;
    h_G[k] = kv.second;
    ++k;
  }
  for (int m = 0; m < n_meshes; ++m) h_row_ptr[m + 1] += h_row_ptr[m];
  return FFK_OK;
}
