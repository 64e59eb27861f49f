// Drop-in counterpart of libclsph/common/structures.h (:16-54): the particle,
// simulation_parameters and precomputed_kernel_values types with the same
// names, members and byte layout, without any OpenCL header.
#ifndef CLSPH_API_STRUCTURES_H_
#define CLSPH_API_STRUCTURES_H_

#include "../../../include/clsph_types.h"

// The reference spells its members with OpenCL host types; provide them when
// no OpenCL header is around so that user code using them still compiles.
#ifndef __CL_PLATFORM_H
typedef float cl_float;
typedef uint32_t cl_uint;
typedef int32_t cl_int;
typedef clsph_float3 cl_float3;
typedef struct {
  cl_uint s[4];
} __attribute__((aligned(16))) cl_uint3;
#endif

typedef clsph_particle particle;
typedef clsph_simulation_parameters simulation_parameters;
typedef clsph_precomputed_kernel_values precomputed_kernel_values;

#define COLLISION_VOLUMES_COUNT 3

#endif  // CLSPH_API_STRUCTURES_H_
