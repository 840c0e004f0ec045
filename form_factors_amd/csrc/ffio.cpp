// Scene, material and frame files around the form-factor path (host code):
// OBJ + thermal MTL import, .task frame import, the OBJ-like export and the
// CSV export of F, with the reference's line rules (import_export/*.cpp).
#include "ff_io.h"

#include <sys/types.h>

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace {

// The fourteen properties of a material block (obj_import.cpp:84-98) and
// where each lands in the reference-layout material.
struct mtl_prop {
  const char* key;
  size_t offset;
};

const mtl_prop kProps[14] = {
    {"shell.density", offsetof(ffp_material, density)},
    {"shell.heat_capacity", offsetof(ffp_material, heat_capacity)},
    {"shell.thermal_conductivity", offsetof(ffp_material, thermal_conductivity)},
    {"shell.thickness", offsetof(ffp_material, thickness)},
    {"front.specular_reflectance", offsetof(ffp_material, front) + 0 * sizeof(float)},
    {"front.diffuse_reflectance", offsetof(ffp_material, front) + 1 * sizeof(float)},
    {"front.absorbance", offsetof(ffp_material, front) + 2 * sizeof(float)},
    {"front.transmittance", offsetof(ffp_material, front) + 3 * sizeof(float)},
    {"front.emissivity", offsetof(ffp_material, front) + 4 * sizeof(float)},
    {"rear.specular_reflectance", offsetof(ffp_material, rear) + 0 * sizeof(float)},
    {"rear.diffuse_reflectance", offsetof(ffp_material, rear) + 1 * sizeof(float)},
    {"rear.absorbance", offsetof(ffp_material, rear) + 2 * sizeof(float)},
    {"rear.transmittance", offsetof(ffp_material, rear) + 3 * sizeof(float)},
    {"rear.emissivity", offsetof(ffp_material, rear) + 4 * sizeof(float)},
};

int find_prop(const std::string& key) {
  for (int k = 0; k < 14; ++k)
    if (key == kProps[k].key) return k;
  return -1;
}

float* prop_ptr(ffp_material& m, int k) {
  return reinterpret_cast<float*>(reinterpret_cast<char*>(&m) + kProps[k].offset);
}

// subject::black_body (subject/objects.cpp:157-188): unit shell, absorbance
// and emissivity 1 on both sides.
ffp_material black_body() {
  ffp_material m;
  std::memset(&m, 0, sizeof(m));
  m.density = m.heat_capacity = m.thermal_conductivity = m.thickness = 1.f;
  m.front[2] = m.front[4] = 1.f;
  m.rear[2] = m.rear[4] = 1.f;
  std::strcpy(m.name, "black_body");
  return m;
}

// One material block whose "newmtl" line is `s` (obj_import.cpp:65-187): a
// block ends at an empty line, or once all fourteen properties are set --
// the loop then has already read the next line, which is dropped
// (obj_import.cpp:102).
int read_block(std::istream& in, std::string& s, unsigned& line, std::vector<ffp_material>& mats,
               std::map<std::string, int>& index) {
  if (s.compare(0, 6, "newmtl") != 0 || s.size() < 7) {
    std::fprintf(stderr, "ffio: material line %u: newmtl without a name\n", line);
    return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
  }
  const std::string name = s.substr(7);
  ffp_material m;
  std::memset(&m, 0, sizeof(m));
  if (name.size() >= sizeof(m.name)) {
    std::fprintf(stderr, "ffio: material line %u: name of %zu characters, 127 at most\n", line, name.size());
    return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
  }
  std::memcpy(m.name, name.data(), name.size());
  std::set<int> set;
  while (std::getline(in, s) && set.size() != 14) {
    ++line;
    if (s.empty()) break;
    if (s[0] == '#') continue;
    const size_t pos = s.find_first_of(" \t");
    // (a line without a separator reads as one unknown key: npos + 1 == 0)
    if (pos + 1 >= s.size()) {
      std::fprintf(stderr, "ffio: material line %u: expected a property and a value\n", line);
      return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
    }
    const std::string key = s.substr(0, pos), value = s.substr(pos + 1);
    const int k = find_prop(key);
    if (k < 0) {
      std::fprintf(stderr, "ffio: material line %u: property '%s' ignored\n", line, key.c_str());
      continue;
    }
    if (std::sscanf(value.c_str(), "%f", prop_ptr(m, k)) != 1) {
      std::fprintf(stderr, "ffio: material line %u: no number in '%s' for '%s'\n", line, value.c_str(), key.c_str());
      return -FFIO_OBJ_IMPORT_MATERIAL_INVALID_PARAMETER_VALUE;
    }
    set.insert(k);
  }
  if (set.size() != 14) {
    std::fprintf(stderr, "ffio: material line %u: '%s' sets %zu of the 14 properties\n", line, name.c_str(),
                 set.size());
    return -FFIO_OBJ_IMPORT_MATERIAL_NOT_ENOUGH_PARAMETERS;
  }
  mats.push_back(m);
  index[name] = (int)mats.size() - 1;
  return FFIO_OK;
}

// obj_import::material_lib (obj_import.cpp:42-193): top-level lines other
// than "newmtl ..." are ignored.
int read_materials(const std::string& path, std::vector<ffp_material>& mats, std::map<std::string, int>& index) {
  std::ifstream in(path);
  if (!in.is_open()) {
    std::fprintf(stderr, "ffio: cannot open the material library %s\n", path.c_str());
    return -FFIO_OBJ_IMPORT_FILE_ERROR;
  }
  std::string s;
  unsigned line = 0;
  while (std::getline(in, s)) {
    ++line;
    if (s.empty() || s[0] != 'n') continue;
    if (int r = read_block(in, s, line, mats, index)) return r;
  }
  return FFIO_OK;
}

// The material library path of an "mtllib" line: relative to the OBJ file's
// directory unless the name contains the two characters "/\" (the
// reference's test, obj_import.cpp:306-317).
std::string library_path(const char* obj_file, const std::string& name) {
  if (name.find("/\\") != std::string::npos) return name;
  std::string base(obj_file);
  const size_t dir = base.find_last_of("/\\");
  base = dir != std::string::npos ? base.substr(0, dir + 1) : std::string();
  return base + name;
}

template <class T>
T* malloc_copy(const std::vector<T>& v, size_t min_n) {
  T* p = static_cast<T*>(std::malloc(sizeof(T) * std::max(v.size(), min_n)));
  if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

}  // namespace

extern "C" {

int ffio_obj_import(const char* filename, ffp_subject_scene** scene) {
  return ffio_obj_import_ex(filename, 0, scene);
}

int ffio_obj_import_ex(const char* filename, int flags, ffp_subject_scene** scene) {
  if (!filename || !scene) return -FFIO_OBJ_IMPORT_FILE_ERROR;
  *scene = nullptr;
  std::ifstream in(filename);
  if (!in.is_open()) {
    std::fprintf(stderr, "ffio: cannot open the scene %s\n", filename);
    return -FFIO_OBJ_IMPORT_FILE_ERROR;
  }
  std::vector<ffp_vec3> verts;
  std::vector<std::array<int, 3>> tris;
  std::vector<ffp_subject_mesh> meshes;
  std::vector<ffp_material> mats;
  std::map<std::string, int> index;
  std::string s;
  unsigned line = 0;
  while (std::getline(in, s)) {
    ++line;
    if (s.empty() || s[0] == '#' || s.size() < 3) continue;
    switch (s[0]) {
      case 'v': {
        if (s[1] == 'n' || s[1] == 't') continue;
        ffp_vec3 v;
        if (std::sscanf(s.c_str() + 1, "%f %f %f", &v.x, &v.y, &v.z) != 3) return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        verts.push_back(v);
        break;
      }
      case 'f': {
        int a = 0, b = 0, c = 0, d;
        const char* p = s.c_str() + 1;
        if (std::sscanf(p, "%d %d %d", &a, &b, &c) != 3 &&
            std::sscanf(p, "%d//%d %d//%d %d//%d", &a, &d, &b, &d, &c, &d) != 6 &&
            std::sscanf(p, "%d/%d/%d %d/%d/%d %d/%d/%d", &a, &d, &d, &b, &d, &d, &c, &d, &d) != 9) {
          std::fprintf(stderr, "ffio: scene line %u: no three vertex indices in '%s'\n", line, s.c_str());
          return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        }
        tris.push_back({a - 1, b - 1, c - 1});
        break;
      }
      case 's':
        continue;
      case 'o':
      case 'g': {
        if (!meshes.empty()) meshes.back().n_faces = (int)tris.size() - meshes.back().first_idx;
        meshes.push_back({(int)tris.size(), 0, 0});
        break;
      }
      case 'u': {
        const size_t pos = s.find_first_of(" \t");
        if (pos == std::string::npos || s.substr(0, pos) != "usemtl") {
          std::fprintf(stderr, "ffio: scene line %u: '%s' is not a usemtl line\n", line, s.c_str());
          return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        }
        const auto it = index.find(s.substr(pos + 1));
        if (it == index.end()) {
          std::fprintf(stderr, "ffio: scene line %u: usemtl names '%s', which no library defines\n", line,
                       s.substr(pos + 1).c_str());
          return -FFIO_OBJ_IMPORT_MATERIAL_NOT_DEFINED;
        }
        if (meshes.empty()) {
          std::fprintf(stderr, "ffio: scene line %u: usemtl before the first group\n", line);
          return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        }
        meshes.back().material_idx = it->second;
        break;
      }
      case 'n':
        if ((flags & FFIO_IMPORT_INLINE_MATERIALS) && s.compare(0, 6, "newmtl") == 0) {
          if (int r = read_block(in, s, line, mats, index)) return r;
          break;
        }
        std::fprintf(stderr, "ffio: scene line %u: '%s' skipped\n", line, s.c_str());
        continue;
      case 'm': {
        if (s.compare(0, 6, "mtllib") != 0 || s.size() < 7) {
          std::fprintf(stderr, "ffio: scene line %u: mtllib without a file name\n", line);
          return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        }
        const std::string path = library_path(filename, s.substr(7));
        std::fprintf(stderr, "ffio: reading materials from %s\n", path.c_str());
        if (int r = read_materials(path, mats, index)) return r;
        break;
      }
      default:
        std::fprintf(stderr, "ffio: scene line %u: '%s' skipped\n", line, s.c_str());
        continue;
    }
  }
  const int nv = (int)verts.size();
  for (const auto& t : tris)
    for (int k = 0; k < 3; ++k)
      if (t[k] < 0 || t[k] >= nv) {
        std::fprintf(stderr, "ffio: a face names vertex %d of %d\n", t[k] + 1, nv);
        return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
      }
  ffp_subject_scene* out = static_cast<ffp_subject_scene*>(std::malloc(sizeof(ffp_subject_scene)));
  if (!out) return -FFIO_OBJ_IMPORT_FILE_ERROR;
  std::memset(out, 0, sizeof(*out));
  out->n_faces = (int)tris.size();
  out->faces = static_cast<ffp_face*>(std::malloc(sizeof(ffp_face) * std::max<size_t>(1, tris.size())));
  for (size_t i = 0; out->faces && i < tris.size(); ++i)
    for (int k = 0; k < 3; ++k) out->faces[i].points[k] = verts[tris[i][k]];
  if (meshes.empty()) {
    meshes.push_back({0, (int)tris.size(), 0});
  } else {
    meshes.back().n_faces = (int)tris.size() - meshes.back().first_idx;
  }
  out->n_meshes = (int)meshes.size();
  out->meshes = malloc_copy(meshes, 1);
  if (mats.empty()) mats.push_back(black_body());
  out->n_materials = (int)mats.size();
  out->materials = malloc_copy(mats, 1);
  if (!out->faces || !out->meshes || !out->materials) {
    ffio_scene_free(out);
    return -FFIO_OBJ_IMPORT_FILE_ERROR;
  }
  *scene = out;
  return FFIO_OK;
}

void ffio_scene_free(ffp_subject_scene* scene) {
  if (!scene) return;
  std::free(scene->faces);
  std::free(scene->meshes);
  std::free(scene->materials);
  std::free(scene);
}

int ffio_task_import(FILE* in, int n_meshes, ffp_ts_task* t, ffp_heat_source_params* heat_source,
                     ffp_source** distant_source) {
  if (!in || !t) return -FFIO_OBJ_IMPORT_FILE_ERROR;
  char* line = nullptr;
  size_t cap = 0;
  // n: temperatures read into this frame; left: room left in t->temperatures
  // (grown 8 at a time; a new frame makes the previous frame's room
  // available again, obj_import.cpp:472-498).
  size_t n = 0, left = 0;
  while (n_meshes > 0 && getline(&line, &cap, in) != -1) {
    switch (line[0]) {
      case 'n':
        left += n;
        n = 0;
        break;
      case 's':
        if (std::sscanf(line, "step %f", &t->time_delta) != 1) {
          std::free(line);
          return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        }
        break;
      case 't': {
        if (left == 0) {
          left = 8;
          float* grown = static_cast<float*>(std::realloc(t->temperatures, (n + left) * sizeof(float)));
          if (!grown) {
            std::free(line);
            return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
          }
          t->temperatures = grown;
        }
        if (std::sscanf(line, "tmprt %f", t->temperatures + n++) != 1) {
          std::free(line);
          return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        }
        --n_meshes;
        --left;
        break;
      }
      case 'h': {
        ffp_heat_source src;
        if (std::sscanf(line, "htsrc %d %f", &src.mesh_idx, &src.power) != 2) {
          std::free(line);
          return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        }
        if (heat_source) {
          ffp_heat_source* grown = static_cast<ffp_heat_source*>(
              std::realloc(heat_source->sources, (1 + (size_t)heat_source->n_sources) * sizeof(src)));
          if (!grown) {
            std::free(line);
            return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
          }
          heat_source->sources = grown;
          heat_source->sources[heat_source->n_sources++] = src;
        }
        break;
      }
      case 'd': {
        ffp_source src;
        if (std::sscanf(line, "dstsrc %f %f %f %f", &src.power, &src.direction.x, &src.direction.y,
                        &src.direction.z) != 4) {
          std::free(line);
          return -FFIO_OBJ_IMPORT_FORMAT_ERROR;
        }
        if (distant_source) {
          if (!*distant_source) *distant_source = static_cast<ffp_source*>(std::malloc(sizeof(ffp_source)));
          if (*distant_source) **distant_source = src;
        }
        break;
      }
      default:
        break;
    }
  }
  std::free(line);
  // Skip the line breaks after the frame so that the next call starts at
  // its first line.
  while (!std::feof(in)) {
    const int c = std::fgetc(in);
    if (c != '\n' && c != '\r') {
      std::ungetc(c, in);
      break;
    }
  }
  return (int)n;
}

int ffio_obj_export_scene(FILE* out, const ffp_subject_scene* scene) {
  if (!out || !scene) return -FFIO_OBJ_IMPORT_FILE_ERROR;
  for (int i = 0; i < scene->n_faces; ++i)
    for (int k = 0; k < 3; ++k) {
      const ffp_vec3& p = scene->faces[i].points[k];
      std::fprintf(out, "v %f %f %f\n", p.x, p.y, p.z);
    }
  std::fprintf(out, "\n");
  static const char* const kFmt = "%s %f\n";
  for (int i = 0; i < scene->n_materials; ++i) {
    ffp_material m = scene->materials[i];
    std::fprintf(out, "newmtl %s\n", m.name);
    for (int k = 0; k < 14; ++k) std::fprintf(out, kFmt, kProps[k].key, *prop_ptr(m, k));
    std::fprintf(out, "\n");
  }
  for (int i = 0; i < scene->n_meshes; ++i) {
    const ffp_subject_mesh& m = scene->meshes[i];
    std::fprintf(out, "g mesh_%d\n", i);
    std::fprintf(out, "usemtl %s\n", scene->materials[m.material_idx].name);
    for (int j = 0; j < m.n_faces; ++j) {
      const int v = 1 + 3 * (m.first_idx + j);
      std::fprintf(out, "f %d %d %d\n", v, v + 1, v + 2);
    }
    std::fprintf(out, "\n");
  }
  return FFIO_OK;
}

int ffio_obj_export_task(FILE* out, int n_meshes, const ffp_ts_task* task) {
  if (!out || !task) return -FFIO_OBJ_IMPORT_FILE_ERROR;
  std::fprintf(out, "newfrm frame_%d\n", task->n_step);
  std::fprintf(out, "step %f\n", task->time_delta);
  for (int m = 0; m < n_meshes; ++m) std::fprintf(out, "tmprt %f\n", task->temperatures[m]);
  std::fprintf(out, "\n");
  std::fflush(out);
  return FFIO_OK;
}

int ffio_obj_export_task_binary(FILE* out, int n_meshes, const ffp_ts_task* task) {
  if (!out || !task) return -FFIO_OBJ_IMPORT_FILE_ERROR;
  std::fprintf(out, "newfrm frame_%d\n", task->n_step);
  std::fprintf(out, "step %f\n", task->time_delta);
  std::fprintf(out, "btmprt %d\n", n_meshes);
  std::fwrite(task->temperatures, 1, sizeof(float) * (size_t)n_meshes, out);
  std::fprintf(out, "\n");
  std::fflush(out);
  return FFIO_OK;
}

int ffio_csv_export(const char* filename, const ffp_ff_scene* scene, const ffp_ff_task* task) {
  if (!filename || !scene || !task) return -FFIO_CSV_EXPORT_FILE_ERROR;
  std::ofstream out(filename);
  if (!out.is_open()) return -FFIO_CSV_EXPORT_FILE_ERROR;
  const int n = scene->n_meshes;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      if (j != 0) out << ";";
      out << task->form_factors[(size_t)i * n + j];
    }
    out << std::endl;
  }
  return FFIO_OK;
}

}  // extern "C"
