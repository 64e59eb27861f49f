// Drop-in counterpart of libclsph/scene.h (:7-15).
#ifndef CLSPH_API_SCENE_H_
#define CLSPH_API_SCENE_H_

#include <string>
#include <vector>

class scene {
 public:
  // Loads "scenes/<filename>" (scene.cpp:9-67): an OBJ file whose faces are
  // fan-triangulated, with unit face normals (B - A) x (C - A) per triangle.
  // Returns false when the file cannot be read or a face is not made of
  // triangles. Several shapes in one file are concatenated.
  bool load(std::string filename);
  // Same, from an explicit path (no "scenes/" prefix).
  bool load_path(const std::string& path);

  unsigned int face_count = 0;
  std::vector<float> face_normals;
  std::vector<float> vertices;
  std::vector<unsigned int> indices;
};

#endif  // CLSPH_API_SCENE_H_
