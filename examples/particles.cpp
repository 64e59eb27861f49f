// Command-line driver with the flow of the reference's example/particles.cpp:
//   clsph_particles <fluid> <simulation_properties> <scene> <frames_prefix> [--yes]
// reads fluid_properties/<fluid>.json, simulation_properties/<sim>.json and
// scenes/<scene>, writes <prefix>frames/frameNNNNNNN.geo per frame (every
// substep when write_all_frames is set) and, when serialize is set, the raw
// particle array to last_frame.bin, which the next run resumes from.
// Without --yes it asks for confirmation on stdin like the reference.
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "../libclsph_amd/api/file_save_delegates/houdini_file_saver.h"
#include "../libclsph_amd/api/sph_simulation.h"

int main(int argc, char** argv) {
  if (argc < 5) {
    std::cout << "Too few arguments" << std::endl
              << "Usage: ./clsph_particles <fluid_name> <simulation_properties_name> "
                 "<scene_name> <frames_folder_prefix> [--yes]"
              << std::endl;
    return -1;
  }
  bool yes = argc > 5 && std::strcmp(argv[5], "--yes") == 0;

  sph_simulation simulation;
  houdini_file_saver saver = houdini_file_saver(std::string(argv[4]));
  try {
    simulation.load_settings(std::string("fluid_properties/") + argv[1] + ".json",
                             std::string("simulation_properties/") + argv[2] + ".json");
  } catch (const std::exception& ex) {
    std::cerr << ex.what() << std::endl;
    std::exit(-1);
  }

  int i = 0;
  // the callbacks only write files: the frame downloads can overlap the
  // following substeps
  simulation.callbacks_read_only = true;
  simulation.pre_frame = [&](particle* particles, const simulation_parameters& params,
                             bool full_frame) {
    if (simulation.write_intermediate_frames != full_frame) saver.writeFrameToFile(particles, params);
    if (simulation.serialize) {
      std::ofstream out("last_frame.bin", std::ios::binary | std::ios::trunc);
      out.write(reinterpret_cast<const char*>(particles),
                (std::streamsize)(sizeof(particle) * params.particles_count));
    }
    ++i;
    int num_frames = params.simulation_time / (params.time_delta * params.simulation_scale);
    int progress = num_frames ? i * 80 / num_frames : 80;
    std::cout << "[";
    for (int j = 0; j < 80; ++j) std::cout << (j < progress ? "-" : " ");
    std::cout << "] " << i << "/" << num_frames << std::endl;
  };

  std::cout << std::endl
            << "Loaded parameters          " << std::endl
            << "-----------------          " << std::endl
            << "Simulation time:           " << simulation.parameters.simulation_time << std::endl
            << "Target FPS:                " << simulation.parameters.target_fps << std::endl
            << "Time delta:                " << simulation.parameters.time_delta << std::endl
            << "Simulation scale:          " << simulation.parameters.simulation_scale << std::endl
            << "Write intermediate frames: "
            << (simulation.write_intermediate_frames ? "true" : "false") << std::endl
            << "Serialize frames:          " << (simulation.serialize ? "true" : "false")
            << std::endl
            << std::endl
            << "Particle count:            " << simulation.parameters.particles_count << std::endl
            << "Particle mass:             " << simulation.parameters.particle_mass << std::endl
            << "Total mass:                " << simulation.parameters.total_mass << std::endl
            << "Initial volume:            " << simulation.initial_volume << std::endl
            << std::endl
            << "Fluid density:             " << simulation.parameters.fluid_density << std::endl
            << "Dynamic viscosity:         " << simulation.parameters.dynamic_viscosity << std::endl
            << "Surface tension threshold: " << simulation.parameters.surface_tension_threshold
            << std::endl
            << "Surface tension:           " << simulation.parameters.surface_tension << std::endl
            << "Stiffness (k):             " << simulation.parameters.K << std::endl
            << "Restitution:               " << simulation.parameters.restitution << std::endl
            << std::endl
            << "Kernel support radius (h): " << simulation.parameters.h << std::endl
            << std::endl
            << "Saving to folder:          " << saver.frames_folder_prefix + "frames/" << std::endl;

  if (!simulation.current_scene.load(argv[3])) {
    std::cerr << "Unable to load scene: " << argv[3] << std::endl;
    return -1;
  }

  // A checkpoint of the wrong size means a different particle count or an
  // interrupted write: refuse to resume from it.
  std::ifstream ck("last_frame.bin", std::ios::binary | std::ios::ate);
  if (ck) {
    size_t size = (size_t)ck.tellg();
    if (size == simulation.parameters.particles_count * sizeof(particle)) {
      std::cout << std::endl
                << "\033[1;32m Serialized frame found.  Simulation will pick up where last run "
                   "left off.\033[0m"
                << std::endl
                << "\033[1;32m To start a new simulation, delete last_frame.bin. \033[0m"
                << std::endl;
    } else {
      std::cout << std::endl
                << "\033[1;31m Serialized frame of incorrect size found. Revert to last know "
                   "settings or delete it, then try again. \033[0m"
                << std::endl;
      return 0;
    }
  }

  std::cout << std::endl
            << "Revise simulation parameters.  Press q to quit, any other key to proceed with "
               "simulation"
            << std::endl;
  char response = 'y';
  if (!yes) std::cin >> response;
  if (response != 'q') simulation.simulate();
  return 0;
}
