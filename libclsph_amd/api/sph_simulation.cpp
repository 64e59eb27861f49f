// sph_simulation on the clsph HIP C ABI.
//
// Behaviour follows libclsph/sph_simulation.cpp:
//   load_settings  :405-506  JSON fields, restitution check, derived h, dt,
//                            max_velocity and kernel constants
//   init_particles :48-94    resume from the checkpoint file or a cube lattice
//   simulate       :346-403  frame x substep loop with pre/post callbacks
// The substep itself (:173-344) runs on the device. The particle array is
// handed to the callbacks in the substep's sorted order, as the reference
// does; it is downloaded only when a callback will see it and uploaded again
// only when a callback has had it since the last upload, so a run without
// callbacks stays resident in device memory.
#include "sph_simulation.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "../../include/clsph_hip.h"
#include "json.h"

struct clsph_sim_impl {
  clsph_ctx* ctx = nullptr;
  clsph_state* st = nullptr;
};

namespace {

const unsigned kPreferredWorkGroupSizeMultiple = 64;

// Reference behaviour on a device error: print where and the code, exit(-1)
// (util/cl_boilerplate.h:22-34 with EXIT_ON_CL_ERROR).
void check(int rc, clsph_ctx* ctx, const char* what, int line) {
  if (rc == CLSPH_OK) return;
  std::cerr << "A clsph error occured (" << __FILE__ << ":" << line << " " << what << ")-> " << rc
            << " " << clsph_last_error(ctx) << std::endl;
  std::exit(-1);
}
#define CLSPH_CHECK(expr, ctx) check((expr), (ctx), #expr, __LINE__)

float f(const clsph_json::value& o, const char* key) {
  return (float)o.at(key).number(key);
}

}  // namespace

sph_simulation::sph_simulation()
    : parameters(), precomputed_terms(), write_intermediate_frames(false), serialize(false),
      initial_volume(0.f), impl_(new clsph_sim_impl) {}

sph_simulation::~sph_simulation() {
  if (impl_->st) clsph_state_free(impl_->st);
  if (impl_->ctx) clsph_free(impl_->ctx);
  delete impl_;
}

void sph_simulation::load_settings(std::string fluid_file_name, std::string parameters_file_name) {
  int particles_inside_influence_radius = 0;
  {
    clsph_json::value fl = clsph_json::parse_file(fluid_file_name);
    parameters.fluid_density = f(fl, "fluid_density");
    parameters.dynamic_viscosity = f(fl, "dynamic_viscosity");
    parameters.restitution = f(fl, "restitution");
    if (parameters.restitution < 0 || parameters.restitution > 1)
      throw std::runtime_error("Restitution has an invalid value!");
    parameters.K = f(fl, "k");
    parameters.surface_tension_threshold = f(fl, "surface_tension_threshold");
    parameters.surface_tension = f(fl, "surface_tension");
    particles_inside_influence_radius =
        (int)fl.at("particles_inside_influence_radius").number("particles_inside_influence_radius");
  }
  {
    clsph_json::value sp = clsph_json::parse_file(parameters_file_name);
    parameters.particles_count = (unsigned int)sp.at("particles_count").number("particles_count");
    if (parameters.particles_count % kPreferredWorkGroupSizeMultiple != 0) {
      std::cout << std::endl
                << "\033[1;31m You should choose a number of particles that is divisble by the "
                   "preferred work group size.\033[0m";
      std::cout << std::endl << "\033[1;31m Performances will be sub-optimal.\033[0m" << std::endl;
    }
    parameters.particle_mass = f(sp, "particle_mass");
    parameters.simulation_time = f(sp, "simulation_time");
    parameters.target_fps = f(sp, "target_fps");
    parameters.simulation_scale = f(sp, "simulation_scale");
    const clsph_json::value& g = sp.at("constant_acceleration");
    parameters.constant_acceleration.s[0] = f(g, "x");
    parameters.constant_acceleration.s[1] = f(g, "y");
    parameters.constant_acceleration.s[2] = f(g, "z");
    parameters.constant_acceleration.s[3] = 0.f;
    write_intermediate_frames = sp.at("write_all_frames").boolean("write_all_frames");
    serialize = sp.at("serialize").boolean("serialize");
  }
  const double pi = 3.14159265358979323846;
  parameters.total_mass = parameters.particles_count * parameters.particle_mass;
  initial_volume = parameters.total_mass / parameters.fluid_density;
  parameters.h = cbrtf(3.f * (particles_inside_influence_radius *
                              (initial_volume / parameters.particles_count)) /
                       (4.f * pi));
  parameters.time_delta = 1.f / parameters.target_fps;
  parameters.max_velocity = 0.8f * parameters.h / parameters.time_delta;
  clsph_kernel_terms(parameters.h, &precomputed_terms);
}

void sph_simulation::init_particles(particle* buffer, const simulation_parameters& p) {
  int per_side = (int)ceil(cbrtf((float)p.particles_count));
  float side_length = cbrtf(initial_volume);
  float spacing = side_length / (float)per_side;
  std::cout << "volume: " << initial_volume << " side_length: " << side_length
            << " spacing: " << spacing << std::endl;

  std::ifstream in(checkpoint_file, std::ios::in | std::ios::binary);
  if (in) {
    // Raw particle[N] bytes, as cereal's saveBinary/loadBinary store them.
    in.read(reinterpret_cast<char*>(buffer), (std::streamsize)(sizeof(particle) * p.particles_count));
    return;
  }
  std::memset(buffer, 0, sizeof(particle) * p.particles_count);
  for (unsigned int i = 0; i < p.particles_count; ++i) {
    buffer[i].position.s[0] = (float)(i % per_side) * spacing - side_length / 2.f;
    buffer[i].position.s[1] = ((float)((i / per_side) % per_side) * spacing);
    buffer[i].position.s[2] = (float)(i / (per_side * per_side)) * spacing - side_length / 2.f;
  }
}

void sph_simulation::simulate(int frame_count) {
  if (frame_count == 0) frame_count = (int)ceil(parameters.simulation_time * parameters.target_fps);
  const unsigned int n = parameters.particles_count;

  if (!impl_->ctx) {
    int rc = clsph_init(device, &impl_->ctx);
    if (rc != CLSPH_OK) {
      std::cerr << "No usable GPU: " << clsph_last_error(nullptr) << std::endl;
      std::exit(-1);
    }
    char name[256];
    if (clsph_device_name(impl_->ctx, name, sizeof(name)) == CLSPH_OK)
      std::cout << "Using device: " << name << std::endl;
  }
  clsph_ctx* ctx = impl_->ctx;
  if (impl_->st) {
    clsph_state_free(impl_->st);
    impl_->st = nullptr;
  }
  CLSPH_CHECK(clsph_state_alloc(ctx, n, 0, &impl_->st), ctx);
  clsph_state* st = impl_->st;
  CLSPH_CHECK(clsph_set_params(st, &parameters, &precomputed_terms, max_bounces), ctx);
  CLSPH_CHECK(clsph_scene_upload(st, current_scene.face_normals.data(),
                                 current_scene.vertices.data(),
                                 (uint32_t)(current_scene.vertices.size() / 3),
                                 current_scene.indices.data(), current_scene.face_count),
              ctx);

  std::vector<particle> host(n);
  particle* particles = host.data();
  init_particles(particles, parameters);
  CLSPH_CHECK(clsph_upload_aos(st, particles, n), ctx);

  // Grid fields the callbacks see are the last substep's (sph_simulation.cpp:229-252).
  auto refresh_params = [&]() {
    simulation_parameters dev;
    CLSPH_CHECK(clsph_get_grid(st, &dev), ctx);
    parameters.min_point = dev.min_point;
    parameters.max_point = dev.max_point;
    parameters.grid_size_x = dev.grid_size_x;
    parameters.grid_size_y = dev.grid_size_y;
    parameters.grid_size_z = dev.grid_size_z;
    parameters.grid_cell_count = dev.grid_cell_count;
  };

  // host_current: the host array equals the device state.
  // host_dirty: a callback has had the array since the last upload and may
  // have changed it; the reference re-uploads the host array before every
  // substep (sph_simulation.cpp:195-197), so such changes must reach the device.
  if (callbacks_read_only && (pre_frame || post_frame)) {
    simulate_overlapped(frame_count, particles);
    return;
  }

  bool host_current = true;
  bool host_dirty = false;
  auto sync_host = [&]() {
    if (!host_current) {
      CLSPH_CHECK(clsph_download_aos(st, particles, n), ctx);
      refresh_params();
      host_current = true;
    }
  };
  for (int i = 0; i < frame_count; ++i) {
    if (pre_frame) {
      sync_host();
      pre_frame(particles, parameters, true);
      host_dirty = true;
    }
    for (int j = 0; (float)j < (1.f / parameters.simulation_scale); ++j) {
      if (pre_frame) {
        sync_host();
        pre_frame(particles, parameters, false);
        host_dirty = true;
      }
      if (host_dirty) {
        CLSPH_CHECK(clsph_upload_aos(st, particles, n), ctx);
        host_dirty = false;
      }
      CLSPH_CHECK(clsph_step(st, 1), ctx);
      host_current = false;
      if (post_frame) {
        sync_host();
        post_frame(particles, parameters, false);
        host_dirty = true;
      }
    }
    if (post_frame) {
      sync_host();
      post_frame(particles, parameters, true);
      host_dirty = true;
    }
  }
  CLSPH_CHECK(clsph_sync(ctx), ctx);
}

// simulate() with read-only callbacks. With S_k the state after k substeps
// and m substeps per frame, the reference's calls (sph_simulation.cpp:384-401)
// see, in order: for each k, post_frame(false) and, at a frame end,
// post_frame(true) on S_k (k >= 1), then pre_frame(true) at a frame start
// and pre_frame(false) on S_k (k < total). S_0 is the initial host array;
// each later S_k is downloaded into one of two page-locked buffers on a copy
// stream, and its callbacks run while the device computes S_k+1.
void sph_simulation::simulate_overlapped(int frame_count, particle* initial) {
  clsph_ctx* ctx = impl_->ctx;
  clsph_state* st = impl_->st;
  const unsigned int n = parameters.particles_count;
  int m = 0;
  while ((float)m < (1.f / parameters.simulation_scale)) ++m;
  const long long total = (long long)frame_count * m;
  particle* buf[2] = {nullptr, nullptr};
  for (int b = 0; b < 2; ++b)
    CLSPH_CHECK(clsph_host_alloc(ctx, sizeof(particle) * n, reinterpret_cast<void**>(&buf[b])),
                ctx);
  auto callbacks = [&](particle* p, long long k) {
    if (k >= 1 && post_frame) {
      post_frame(p, parameters, false);
      if (k % m == 0) post_frame(p, parameters, true);
    }
    if (k < total && pre_frame) {
      if (k % m == 0) pre_frame(p, parameters, true);
      pre_frame(p, parameters, false);
    }
  };
  auto receive = [&](int slot) {
    simulation_parameters dev;
    CLSPH_CHECK(clsph_download_wait(st, (uint32_t)slot, &dev), ctx);
    parameters.min_point = dev.min_point;
    parameters.max_point = dev.max_point;
    parameters.grid_size_x = dev.grid_size_x;
    parameters.grid_size_y = dev.grid_size_y;
    parameters.grid_size_z = dev.grid_size_z;
    parameters.grid_cell_count = dev.grid_cell_count;
  };
  callbacks(initial, 0);
  int pending = -1, slot = 0;
  for (long long k = 0; k < total; ++k) {
    CLSPH_CHECK(clsph_step(st, 1), ctx);
    CLSPH_CHECK(clsph_download_begin(st, (uint32_t)slot, buf[slot], n), ctx);
    if (pending >= 0) {
      receive(pending);
      callbacks(buf[pending], k);
    }
    pending = slot;
    slot ^= 1;
  }
  if (pending >= 0) {
    receive(pending);
    callbacks(buf[pending], total);
  }
  CLSPH_CHECK(clsph_sync(ctx), ctx);
  for (int b = 0; b < 2; ++b) clsph_host_free(ctx, buf[b]);
}
