// OBJ scene loading with the parse rules the reference inherits from
// tinyobjloader 0.9.6 (util/tinyobj/tiny_obj_loader.cc):
//   - coordinates read with (float)atof (:91-97);
//   - face entries "v", "v/vt", "v//vn", "v/vt/vn", 1-based, negative = relative
//     to the current count (:66-78, :115-147);
//   - polygons fan-triangulated from their first vertex (:218-241);
//   - vertices de-duplicated per shape by their (v, vt, vn) triple in first-use
//     order (:149-193);
//   - "o" and "g" start a new shape; a referenced material library must exist
//     (:284-295, :582-595).
// Face normals are computed as scene.cpp:36-64 does. The reference keeps only
// the last shape's vertices and indices while accumulating every shape's
// normals (scene.cpp:23-64); here all shapes are concatenated instead.
#include "scene.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <tuple>

namespace {

struct vidx {
  int v = -1, vt = -1, vn = -1;
  bool operator<(const vidx& o) const {
    return std::tie(v, vt, vn) < std::tie(o.v, o.vt, o.vn);
  }
};

struct shape {
  std::vector<float> positions;
  std::vector<unsigned int> indices;
};

inline bool is_space(char c) { return c == ' ' || c == '\t'; }
inline bool is_newline(char c) { return c == '\r' || c == '\n' || c == '\0'; }

int fix_index(int idx, int n) {
  if (idx > 0) return idx - 1;
  if (idx == 0) return 0;
  return n + idx;
}

float parse_float(const char*& t) {
  t += std::strspn(t, " \t");
  float f = (float)std::atof(t);
  t += std::strcspn(t, " \t\r");
  return f;
}

vidx parse_triple(const char*& t, int nv, int nvn, int nvt) {
  vidx vi;
  vi.v = fix_index(std::atoi(t), nv);
  t += std::strcspn(t, "/ \t\r");
  if (t[0] != '/') return vi;
  ++t;
  if (t[0] == '/') {
    ++t;
    vi.vn = fix_index(std::atoi(t), nvn);
    t += std::strcspn(t, "/ \t\r");
    return vi;
  }
  vi.vt = fix_index(std::atoi(t), nvt);
  t += std::strcspn(t, "/ \t\r");
  if (t[0] != '/') return vi;
  ++t;
  vi.vn = fix_index(std::atoi(t), nvn);
  t += std::strcspn(t, "/ \t\r");
  return vi;
}

bool flush(std::vector<shape>& shapes, const std::vector<float>& v,
           std::vector<std::vector<vidx>>& faces, std::string& err) {
  if (faces.empty()) return true;
  shape sh;
  std::map<vidx, unsigned int> cache;
  auto id = [&](const vidx& k) -> long {
    auto it = cache.find(k);
    if (it != cache.end()) return it->second;
    if (k.v < 0 || (size_t)(3 * k.v + 2) >= v.size()) return -1;
    sh.positions.push_back(v[3 * k.v + 0]);
    sh.positions.push_back(v[3 * k.v + 1]);
    sh.positions.push_back(v[3 * k.v + 2]);
    unsigned int n = (unsigned int)(sh.positions.size() / 3 - 1);
    cache[k] = n;
    return n;
  };
  for (const auto& f : faces) {
    if (f.size() < 3) continue;
    for (size_t k = 2; k < f.size(); ++k) {
      long a = id(f[0]), b = id(f[k - 1]), c = id(f[k]);
      if (a < 0 || b < 0 || c < 0) {
        err = "face refers to a vertex that does not exist";
        return false;
      }
      sh.indices.push_back((unsigned int)a);
      sh.indices.push_back((unsigned int)b);
      sh.indices.push_back((unsigned int)c);
    }
  }
  faces.clear();
  shapes.push_back(std::move(sh));
  return true;
}

bool load_obj(const std::string& path, const std::string& basepath, std::vector<shape>& shapes,
              std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "Cannot open file [" + path + "]";
    return false;
  }
  std::vector<float> v;
  int nvn = 0, nvt = 0;
  std::vector<std::vector<vidx>> faces;
  std::string line;
  while (std::getline(in, line)) {
    const char* t = line.c_str();
    t += std::strspn(t, " \t");
    if (t[0] == '\0' || t[0] == '#') continue;
    if (t[0] == 'v' && is_space(t[1])) {
      t += 2;
      float x = parse_float(t), y = parse_float(t), z = parse_float(t);
      v.push_back(x);
      v.push_back(y);
      v.push_back(z);
    } else if (t[0] == 'v' && t[1] == 'n' && is_space(t[2])) {
      ++nvn;
    } else if (t[0] == 'v' && t[1] == 't' && is_space(t[2])) {
      ++nvt;
    } else if (t[0] == 'f' && is_space(t[1])) {
      t += 2;
      t += std::strspn(t, " \t");
      std::vector<vidx> f;
      while (!is_newline(t[0])) {
        f.push_back(parse_triple(t, (int)(v.size() / 3), nvn, nvt));
        t += std::strspn(t, " \t\r");
      }
      faces.push_back(std::move(f));
    } else if (std::strncmp(t, "mtllib", 6) == 0 && is_space(t[6])) {
      char name[4096] = {0};
      if (std::sscanf(t + 7, "%4095s", name) == 1) {
        std::ifstream mtl(basepath + name);
        if (!mtl) {
          err = "Cannot open file [" + basepath + name + "]";
          return false;
        }
      }
    } else if ((t[0] == 'g' || t[0] == 'o') && is_space(t[1])) {
      if (!flush(shapes, v, faces, err)) return false;
    }
  }
  return flush(shapes, v, faces, err);
}

}  // namespace

bool scene::load(std::string filename) {
  return load_path(std::string("scenes/") + filename);
}

bool scene::load_path(const std::string& path) {
  std::vector<shape> shapes;
  std::string err;
  std::string base;
  size_t slash = path.find_last_of('/');
  if (slash != std::string::npos) base = path.substr(0, slash + 1);
  if (!load_obj(path, base, shapes, err)) {
    std::cerr << err << std::endl;
    return false;
  }
  std::cout << "Scene Loading - number of shapes in file [" << path << "]: " << shapes.size()
            << std::endl;
  face_count = 0;
  face_normals.clear();
  vertices.clear();
  indices.clear();
  for (const shape& sh : shapes) {
    if (sh.indices.size() % 3 != 0) {
      std::cerr << "Meshes must be made of triangles only" << std::endl;
      return false;
    }
    unsigned int base_vertex = (unsigned int)(vertices.size() / 3);
    vertices.insert(vertices.end(), sh.positions.begin(), sh.positions.end());
    for (size_t f = 0; f < sh.indices.size(); f += 3) {
      const float* a = &sh.positions[3 * sh.indices[f + 0]];
      const float* b = &sh.positions[3 * sh.indices[f + 1]];
      const float* c = &sh.positions[3 * sh.indices[f + 2]];
      float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
      float vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
      float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
      float len = std::sqrt(nx * nx + ny * ny + nz * nz);
      face_normals.push_back(nx / len);
      face_normals.push_back(ny / len);
      face_normals.push_back(nz / len);
      for (int k = 0; k < 3; ++k) indices.push_back(base_vertex + sh.indices[f + k]);
    }
    face_count += (unsigned int)(sh.indices.size() / 3);
  }
  return true;
}
