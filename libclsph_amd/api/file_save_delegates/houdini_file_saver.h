// Drop-in counterpart of libclsph/file_save_delegates/houdini_file_saver.h.
#ifndef CLSPH_API_HOUDINI_FILE_SAVER_H_
#define CLSPH_API_HOUDINI_FILE_SAVER_H_

#include <string>

#include "../common/structures.h"

class houdini_file_saver {
 public:
  houdini_file_saver(std::string frames_folder_prefix)
      : frames_folder_prefix(frames_folder_prefix), frame_count(0) {}

  // Writes <prefix>frames/frameNNNNNNN.geo (PGEOMETRY V5 ASCII: position,
  // v, density-derived color, mass), as houdini_file_saver.cpp:25-92 does.
  int writeFrameToFile(particle* particles, const simulation_parameters& parameters);

  std::string frames_folder_prefix;

 private:
  int frame_count;
};

// The ASCII geometry text for n particles (what writeFrameToFile writes).
std::string clsph_houdini_geo_text(const particle* particles, unsigned int n, float mass);
// 7-digit zero-padded frame number (houdini_file_saver.cpp:15-23).
std::string clsph_zero_pad(int num);

#endif  // CLSPH_API_HOUDINI_FILE_SAVER_H_
