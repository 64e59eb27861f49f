// Drop-in counterpart of libclsph/sph_simulation.h (:8-52). The public
// members are the reference's; the device state behind them lives in the
// clsph_hip C ABI (include/clsph_hip.h) and is hidden behind a pointer, so
// this header pulls in no GPU headers.
#ifndef CLSPH_API_SPH_SIMULATION_H_
#define CLSPH_API_SPH_SIMULATION_H_

#include <functional>
#include <string>

#include "common/structures.h"
#include "scene.h"

struct clsph_sim_impl;

class sph_simulation {
 public:
  sph_simulation();
  ~sph_simulation();
  sph_simulation(const sph_simulation&) = delete;
  sph_simulation& operator=(const sph_simulation&) = delete;

  // Runs frame_count frames (0: ceil(simulation_time * target_fps)), each of
  // ceil(1 / simulation_scale) substeps, calling pre_frame / post_frame
  // around frames (full_frame = true) and substeps (false), as
  // sph_simulation.cpp:346-403 does.
  void simulate(int frame_count = 0);

  simulation_parameters parameters;
  precomputed_kernel_values precomputed_terms;

  std::function<void(particle*, const simulation_parameters&, bool)> pre_frame;
  std::function<void(particle*, const simulation_parameters&, bool)> post_frame;

  // Reads the fluid and simulation JSON files (sph_simulation.cpp:405-506).
  // Throws std::runtime_error on an invalid restitution, a missing key or an
  // unreadable file.
  void load_settings(std::string fluid_file_name, std::string parameters_file_name);

  bool write_intermediate_frames;
  bool serialize;
  float initial_volume;
  scene current_scene;

  // Device to run on (-1: the first one), and the collision-loop cap.
  int device = -1;
  unsigned int max_bounces = 64;
  // Checkpoint file read at the start of simulate() (sph_simulation.cpp:57-67).
  std::string checkpoint_file = "last_frame.bin";
  // Set when the callbacks only read the particles (frame writers, probes):
  // simulate() then downloads each substep asynchronously into one of two
  // page-locked buffers and runs the callbacks on a substep while the device
  // computes the next one. The callbacks see the same particles and
  // parameters, in the same order; changes they make are not uploaded.
  bool callbacks_read_only = false;

 private:
  void init_particles(particle* buffer, const simulation_parameters&);
  void simulate_overlapped(int frame_count, particle* initial);
  clsph_sim_impl* impl_;
};

#endif  // CLSPH_API_SPH_SIMULATION_H_
