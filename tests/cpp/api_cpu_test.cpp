// Host-only checks of the drop-in C++ API (no device is touched).
// Prints "ok <name>" per check; exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../libclsph_amd/api/common/util.h"
#include "../../libclsph_amd/api/file_save_delegates/houdini_file_saver.h"
#include "../../libclsph_amd/api/sph_simulation.h"
#include "../../libclsph_amd/api/util/cl_boilerplate.h"

static int fails = 0;
#define EXPECT(c)                                                      \
  do {                                                                 \
    if (!(c)) {                                                        \
      std::cerr << "FAIL " << __LINE__ << ": " #c << std::endl;        \
      ++fails;                                                         \
    }                                                                  \
  } while (0)

static void write(const std::string& path, const std::string& text) {
  std::ofstream f(path);
  f << text;
}

// The text HoudiniFileDumpParticles::dump produces, written with ostream <<
// exactly as that writer formats it (a float through operator<<).
static std::string ostream_geo(const particle* p, unsigned n, float mass) {
  std::ostringstream s;
  s << "PGEOMETRY V5" << std::endl;
  s << "NPoints " << n << " NPrims 1" << std::endl;
  s << "NPointGroups 0 NPrimGroups 1" << std::endl;
  s << "NPointAttrib " << 3 << " NVertexAttrib 0 NPrimAttrib 2 NAttrib 0" << std::endl;
  s << "PointAttrib" << std::endl;
  s << "v 3 float 1 1 1" << std::endl << "color 3 float 1 1 1" << std::endl
    << "mass 1 float 1" << std::endl;
  for (unsigned i = 0; i < n; ++i) {
    float d = p[i].density;
    float r = d > 1000.f && d <= 2000.f ? (d - 1000.f) / 1000.f : 0.f;
    float g = d >= 0.f && d < 1000.f ? 1.f - d / 1000.f : 0.f;
    float b = d >= 500.f && d <= 1000.f ? (d - 500.f) / 500.f
              : d >= 1000.f && d <= 1500.f ? 1.f - (d - 1000.f) / 500.f : 0.f;
    float pw = 0;
    s << p[i].position.s[0] << " " << p[i].position.s[1] << " " << p[i].position.s[2] << " "
      << pw << " ";
    s << "(" << p[i].velocity.s[0] << " " << p[i].velocity.s[1] << " " << p[i].velocity.s[2]
      << "\t" << r << " " << g << " " << b << "\t" << mass << ")" << std::endl;
  }
  s << "PrimitiveAttrib" << std::endl << "generator 1 index 1 location1" << std::endl
    << "dopobject 1 index 1 /obj/AutoDopNetwork:1" << std::endl;
  s << "Part " << n;
  for (unsigned i = 0; i < n; ++i) s << " " << i;
  s << " [0\t0]" << std::endl;
  s << "box_object1 unordered" << std::endl << "1 1" << std::endl;
  s << "beginExtra" << std::endl << "endExtra" << std::endl;
  return s.str();
}

int main(int argc, char** argv) {
  // `read <file>`: prints readKernelFile's result (its error exits are
  // checked by the caller from the exit status and the output)
  if (argc > 2 && std::string(argv[1]) == "read") {
    std::cout << readKernelFile(argv[2]);
    return 0;
  }
  std::string dir = argc > 1 ? argv[1] : ".";

  // ---- load_settings: fields and derived values (sph_simulation.cpp:405-506)
  write(dir + "/fluid.json",
        "{ \"fluid_density\" : 998.29, \"dynamic_viscosity\" : 3.5, \"restitution\" : 0.25,\n"
        "  \"k\" : 100, \"surface_tension_threshold\" : 7.065, \"surface_tension\" : 0.0728,\n"
        "  \"particles_inside_influence_radius\" : 20.7 };");
  write(dir + "/sim.json",
        "{ \"particles_count\" : 32000, \"particle_mass\" : 0.05, \"simulation_time\" : 10,\n"
        "  \"target_fps\" : 60, \"simulation_scale\" : 0.1, \"write_all_frames\" : false,\n"
        "  \"serialize\" : true, \"constant_acceleration\" : { \"x\" : 0, \"y\" : -9.8, \"z\" : 0.5 } };");
  sph_simulation sim;
  sim.load_settings(dir + "/fluid.json", dir + "/sim.json");
  simulation_parameters& p = sim.parameters;
  EXPECT(p.fluid_density == 998.29f && p.dynamic_viscosity == 3.5f && p.restitution == 0.25f);
  EXPECT(p.K == 100.f && p.surface_tension_threshold == 7.065f && p.surface_tension == 0.0728f);
  EXPECT(p.particles_count == 32000 && p.particle_mass == 0.05f && p.target_fps == 60.f);
  EXPECT(p.constant_acceleration.s[1] == -9.8f && p.constant_acceleration.s[2] == 0.5f);
  EXPECT(sim.serialize && !sim.write_intermediate_frames);
  double V = 32000 * 0.05 / 998.29;
  // particles_inside_influence_radius is truncated to an int (20)
  double h = std::cbrt(3.0 * 20 * (V / 32000) / (4 * M_PI));
  EXPECT(std::fabs(p.h - h) < 1e-6 * h);
  EXPECT(std::fabs(sim.initial_volume - V) < 1e-6 * V);
  EXPECT(p.time_delta == 1.f / 60.f);
  EXPECT(std::fabs(p.max_velocity - 0.8 * h * 60) < 1e-5 * p.max_velocity);
  double hf = p.h;
  EXPECT(std::fabs(sim.precomputed_terms.poly_6 / (315 / (64 * M_PI * std::pow(hf, 9))) - 1) < 1e-6);
  EXPECT(std::fabs(sim.precomputed_terms.spiky / (-45 / (M_PI * std::pow(hf, 6))) - 1) < 1e-6);
  std::cout << "ok load_settings" << std::endl;

  // ---- error paths
  write(dir + "/bad_e.json",
        "{ \"fluid_density\" : 1000, \"dynamic_viscosity\" : 1, \"restitution\" : 1.5, \"k\" : 1,"
        "  \"surface_tension_threshold\" : 1, \"surface_tension\" : 1,"
        "  \"particles_inside_influence_radius\" : 20 }");
  try {
    sph_simulation s2;
    s2.load_settings(dir + "/bad_e.json", dir + "/sim.json");
    EXPECT(false);
  } catch (const std::runtime_error& e) {
    EXPECT(std::string(e.what()) == "Restitution has an invalid value!");
  }
  write(dir + "/missing.json", "{ \"fluid_density\" : 1000 }");
  try {
    sph_simulation s3;
    s3.load_settings(dir + "/missing.json", dir + "/sim.json");
    EXPECT(false);
  } catch (const std::runtime_error& e) {
    EXPECT(std::string(e.what()).find("dynamic_viscosity") != std::string::npos);
  }
  std::cout << "ok settings_errors" << std::endl;

  // ---- OBJ loading: quads fan-triangulated, v/vt/vn triples, negative indices,
  //      two shapes concatenated, unit normals (B - A) x (C - A)
  write(dir + "/m.mtl", "newmtl None\n");
  write(dir + "/t.obj",
        "# test\nmtllib m.mtl\no A\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
        "f 1/1/1 2/1/1 3/1/1 4/1/1\ng B\nv 0 0 2\nv 0 0 3\r\nv 0 1 2\nf -3 -2 -1\n");
  scene sc;
  EXPECT(sc.load_path(dir + "/t.obj"));
  EXPECT(sc.face_count == 3);
  EXPECT(sc.indices.size() == 9 && sc.vertices.size() == 7 * 3);
  // quad (0,1,2,3) -> (0,1,2), (0,2,3)
  EXPECT(sc.indices[0] == 0 && sc.indices[1] == 1 && sc.indices[2] == 2);
  EXPECT(sc.indices[3] == 0 && sc.indices[4] == 2 && sc.indices[5] == 3);
  EXPECT(sc.indices[6] == 4 && sc.indices[7] == 5 && sc.indices[8] == 6);
  EXPECT(sc.face_normals[2] == 1.f && sc.face_normals[5] == 1.f);
  // (0,0,2) (0,0,3) (0,1,2): (0,0,1) x (0,1,0) = (-1,0,0)
  EXPECT(sc.face_normals[6] == -1.f && sc.face_normals[7] == 0.f);
  EXPECT(sc.vertices[3 * 5 + 2] == 3.f);
  write(dir + "/nomtl.obj", "mtllib absent.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n");
  scene sc2;
  EXPECT(!sc2.load_path(dir + "/nomtl.obj"));
  scene sc3;
  EXPECT(!sc3.load_path(dir + "/does_not_exist.obj"));
  std::cout << "ok scene" << std::endl;

  // ---- Houdini text, byte for byte against the ostream formatting
  for (unsigned n : {1u, 7u, 70000u}) {
    std::vector<particle> v(n);
    std::memset(v.data(), 0, sizeof(particle) * n);
    unsigned s = 12345;
    auto rnd = [&]() {
      s = s * 1103515245u + 12345u;
      return (float)((s >> 8) & 0xFFFF) / 65536.f;
    };
    for (unsigned i = 0; i < n; ++i) {
      for (int a = 0; a < 3; ++a) {
        v[i].position.s[a] = (rnd() - 0.5f) * (i % 5 == 0 ? 1e-7f : 3.f);
        v[i].velocity.s[a] = (rnd() - 0.5f) * 20.f;
      }
      v[i].density = rnd() * 2500.f - 100.f;
    }
    if (n > 3) {
      v[0].position.s[0] = -0.f;
      v[1].velocity.s[1] = std::numeric_limits<float>::quiet_NaN();
      v[2].density = 1000.f;
      v[3].position.s[2] = 123456789.f;
    }
    std::string a = clsph_houdini_geo_text(v.data(), n, 0.0078125f);
    std::string b = ostream_geo(v.data(), n, 0.0078125f);
    EXPECT(a == b);
  }
  EXPECT(clsph_zero_pad(7) == "0000007" && clsph_zero_pad(123456789) == "3456789");
  {
    houdini_file_saver saver(dir + "/out_");
    std::vector<particle> v(3);
    std::memset(v.data(), 0, sizeof(particle) * 3);
    simulation_parameters q = p;
    q.particles_count = 3;
    std::string cmd = "mkdir -p " + dir + "/out_frames";
    EXPECT(std::system(cmd.c_str()) == 0);
    saver.writeFrameToFile(v.data(), q);
    saver.writeFrameToFile(v.data(), q);
    std::ifstream f1(dir + "/out_frames/frame0000002.geo");
    EXPECT((bool)f1);
  }
  std::cout << "ok houdini" << std::endl;

  // ---- Morton helpers of common/util.h
  for (unsigned x = 0; x < 1024; x += 7)
    for (unsigned y = 0; y < 1024; y += 131)
      for (unsigned z = 0; z < 1024; z += 257) {
        cl_uint k = get_grid_index_z_curve(x, y, z);
        cl_uint3 c = get_cell_coords_z_curve(k);
        EXPECT(c.s[0] == x && c.s[1] == y && c.s[2] == z);
      }
  EXPECT(get_grid_index_z_curve(1, 0, 0) == 1 && get_grid_index_z_curve(0, 1, 0) == 2 &&
         get_grid_index_z_curve(0, 0, 1) == 4 && get_grid_index_z_curve(2, 0, 0) == 8);
  std::cout << "ok util" << std::endl;
  return fails ? 1 : 0;
}
