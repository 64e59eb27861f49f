// Device-side checks of the drop-in C++ API: callback protocol, agreement of
// simulate() with the raw C ABI substeps, checkpoint resume.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/clsph_hip.h"
#include "../../libclsph_amd/api/sph_simulation.h"

static int fails = 0;
#define EXPECT(c)                                               \
  do {                                                          \
    if (!(c)) {                                                 \
      std::cerr << "FAIL " << __LINE__ << ": " #c << std::endl; \
      ++fails;                                                  \
    }                                                           \
  } while (0)

static void setup(sph_simulation& s, unsigned n, const std::string& dir) {
  // water-like settings for a small cube of fluid in a 1 m box
  s.parameters.particles_count = n;
  s.parameters.fluid_density = 1000.f;
  s.parameters.particle_mass = 1000.f * 0.01f * 0.01f * 0.01f;
  s.parameters.total_mass = n * s.parameters.particle_mass;
  s.initial_volume = s.parameters.total_mass / s.parameters.fluid_density;
  s.parameters.dynamic_viscosity = 3.5e-3f;
  s.parameters.K = 3.f;
  s.parameters.surface_tension = 0.0728f;
  s.parameters.surface_tension_threshold = 7.065f;
  s.parameters.restitution = 0.f;
  s.parameters.constant_acceleration.s[1] = -9.81f;
  s.parameters.h = 0.02f;
  s.parameters.target_fps = 1000.f;
  s.parameters.time_delta = 1.f / s.parameters.target_fps;
  s.parameters.simulation_scale = 0.5f;  // two substeps per frame
  s.parameters.simulation_time = 0.005f;
  s.parameters.max_velocity = 0.8f * s.parameters.h / s.parameters.time_delta;
  clsph_kernel_terms(s.parameters.h, &s.precomputed_terms);
  std::ofstream obj(dir + "/box.obj");
  obj << "v -0.5 -0.3 -0.5\nv 0.5 -0.3 -0.5\nv 0.5 0.7 -0.5\nv -0.5 0.7 -0.5\n"
         "v -0.5 -0.3 0.5\nv 0.5 -0.3 0.5\nv 0.5 0.7 0.5\nv -0.5 0.7 0.5\n"
         "f 1 2 3 4\nf 5 8 7 6\nf 1 5 6 2\nf 4 3 7 8\nf 1 4 8 5\nf 2 6 7 3\n";
  obj.close();
  if (!s.current_scene.load_path(dir + "/box.obj")) ++fails;
  s.checkpoint_file = dir + "/last_frame.bin";
}

int main(int argc, char** argv) {
  std::string dir = argc > 1 ? argv[1] : ".";
  const unsigned n = 4096;
  std::remove((dir + "/last_frame.bin").c_str());

  // ---- callback protocol and cube initialisation (sph_simulation.cpp:48-94, :346-403)
  std::vector<particle> first, last;
  int pre_full = 0, pre_sub = 0, post_full = 0, post_sub = 0;
  {
    sph_simulation s;
    setup(s, n, dir);
    s.pre_frame = [&](particle* p, const simulation_parameters& prm, bool full) {
      if (full) {
        if (pre_full == 0) first.assign(p, p + prm.particles_count);
        ++pre_full;
      } else {
        ++pre_sub;
      }
      std::ofstream out(dir + "/last_frame.bin", std::ios::binary | std::ios::trunc);
      out.write(reinterpret_cast<const char*>(p), sizeof(particle) * prm.particles_count);
    };
    s.post_frame = [&](particle* p, const simulation_parameters& prm, bool full) {
      (full ? post_full : post_sub)++;
      last.assign(p, p + prm.particles_count);
      EXPECT(prm.grid_cell_count > 0);
    };
    s.simulate();  // ceil(0.005 * 1000) = 5 frames of 2 substeps
  }
  EXPECT(pre_full == 5 && post_full == 5 && pre_sub == 10 && post_sub == 10);
  int side = (int)std::ceil(std::cbrt((float)n));
  float side_length = std::cbrt(n * 1e-6f);
  float spacing = side_length / side;
  EXPECT(first.size() == n);
  for (unsigned i = 0; i < n; i += 97) {
    EXPECT(first[i].position.s[0] == (float)(i % side) * spacing - side_length / 2.f);
    EXPECT(first[i].position.s[1] == (float)((i / side) % side) * spacing);
    EXPECT(first[i].position.s[2] == (float)(i / (side * side)) * spacing - side_length / 2.f);
  }
  std::cout << "ok callbacks" << std::endl;

  // ---- the same 10 substeps through the raw C ABI give the same particles
  {
    sph_simulation s;
    setup(s, n, dir);
    clsph_ctx* ctx;
    clsph_state* st;
    EXPECT(clsph_init(0, &ctx) == 0);
    EXPECT(clsph_state_alloc(ctx, n, 0, &st) == 0);
    EXPECT(clsph_set_params(st, &s.parameters, &s.precomputed_terms, 64) == 0);
    EXPECT(clsph_scene_upload(st, s.current_scene.face_normals.data(),
                              s.current_scene.vertices.data(),
                              (uint32_t)(s.current_scene.vertices.size() / 3),
                              s.current_scene.indices.data(), s.current_scene.face_count) == 0);
    EXPECT(clsph_upload_aos(st, first.data(), n) == 0);
    EXPECT(clsph_step(st, 10) == 0);
    std::vector<particle> out(n);
    EXPECT(clsph_download_aos(st, out.data(), n) == 0);
    EXPECT(std::memcmp(out.data(), last.data(), sizeof(particle) * n) == 0);
    clsph_state_free(st);
    clsph_free(ctx);
  }
  std::cout << "ok abi_agreement" << std::endl;

  // ---- resume: init_particles reads the checkpoint the callbacks wrote
  {
    std::vector<particle> ck(n);
    std::ifstream in(dir + "/last_frame.bin", std::ios::binary);
    in.read(reinterpret_cast<char*>(ck.data()), sizeof(particle) * n);
    EXPECT((bool)in);
    sph_simulation s;
    setup(s, n, dir);
    std::vector<particle> got;
    s.pre_frame = [&](particle* p, const simulation_parameters& prm, bool full) {
      if (full && got.empty()) got.assign(p, p + prm.particles_count);
    };
    s.simulate(1);
    EXPECT(got.size() == n && std::memcmp(got.data(), ck.data(), sizeof(particle) * n) == 0);
  }
  std::cout << "ok resume" << std::endl;

  // ---- read-only callbacks: the overlapped path delivers the same particles
  // and parameters, in the same order, as the synchronous one
  {
    std::remove((dir + "/last_frame.bin").c_str());
    std::vector<std::vector<particle>> seen[2];
    std::vector<std::pair<int, unsigned>> meta[2];
    for (int mode = 0; mode < 2; ++mode) {
      sph_simulation s;
      setup(s, n, dir);
      s.callbacks_read_only = mode == 1;
      s.pre_frame = [&](particle* p, const simulation_parameters& prm, bool full) {
        seen[mode].emplace_back(p, p + prm.particles_count);
        meta[mode].emplace_back(full ? 2 : 3, prm.grid_cell_count);
      };
      s.post_frame = [&](particle* p, const simulation_parameters& prm, bool full) {
        seen[mode].emplace_back(p, p + prm.particles_count);
        meta[mode].emplace_back(full ? 1 : 0, prm.grid_cell_count);
      };
      s.simulate();
    }
    EXPECT(seen[0].size() == 30 && seen[1].size() == seen[0].size());
    EXPECT(meta[0] == meta[1]);
    bool same = seen[0].size() == seen[1].size();
    for (size_t k = 0; same && k < seen[0].size(); ++k)
      same = std::memcmp(seen[0][k].data(), seen[1][k].data(), sizeof(particle) * n) == 0;
    EXPECT(same);
  }
  std::cout << "ok overlapped_callbacks" << std::endl;
  return fails ? 1 : 0;
}
