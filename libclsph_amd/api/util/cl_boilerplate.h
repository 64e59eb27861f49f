// Drop-in counterpart of the part of util/cl_boilerplate.h that survives the
// move to ahead-of-time HIP kernels: readKernelFile (:43), a plain text-file
// reader that the reference's GL example also uses for its shaders
// (example/opengl.cpp:19). Device selection and program building
// (init_cl_single_device, make_program) are clsph_init in clsph_hip.h.
#ifndef CLSPH_API_CL_BOILERPLATE_H_
#define CLSPH_API_CL_BOILERPLATE_H_

#include <string>

// Whole contents of the file at `kernelFileName`. A file that cannot be
// opened prints "<name> file not found." and an empty one "<name> is empty."
// on standard output, then the process exits with status 1
// (util/cl_boilerplate.cpp:90-112).
std::string readKernelFile(std::string kernelFileName);

#endif  // CLSPH_API_CL_BOILERPLATE_H_
