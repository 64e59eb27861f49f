/*
 * Plain-C layout of the three structures that cross the libclsph boundary.
 *
 * These are byte-for-byte the layouts of the reference's shared structs
 * (libclsph/common/structures.h:16-54) as seen by a 64-bit host compiler with
 * cl_float3 == 16-byte aligned float[4]:
 *
 *   particle                  80 B  position@0 velocity@16
 *                                   intermediate_velocity@32 acceleration@48
 *                                   density@64 pressure@68 grid_index@72
 *   simulation_parameters    128 B  constant_acceleration@64 grid_size_x@80
 *                                   grid_cell_count@92 min_point@96
 *                                   max_point@112
 *   precomputed_kernel_values 20 B
 *
 * Keeping the layout identical is what lets `last_frame.bin` checkpoints and
 * the pre_frame/post_frame callbacks exchange raw particle arrays unchanged.
 */
#ifndef CLSPH_TYPES_H_
#define CLSPH_TYPES_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 16-byte aligned 3-vector with a padding lane, like OpenCL's float3. */
typedef struct clsph_float3 {
  float s[4];
} __attribute__((aligned(16))) clsph_float3;

typedef struct clsph_particle {
  clsph_float3 position;
  clsph_float3 velocity;
  clsph_float3 intermediate_velocity;
  clsph_float3 acceleration;
  float density;
  float pressure;
  uint32_t grid_index;
} clsph_particle;

typedef struct clsph_simulation_parameters {
  uint32_t particles_count;
  float max_velocity;
  float fluid_density;
  float total_mass;
  float particle_mass;
  float dynamic_viscosity;
  float simulation_time;
  float target_fps;
  float h;
  float simulation_scale;
  float time_delta;
  float surface_tension_threshold;
  float surface_tension;
  float restitution;
  float K;
  clsph_float3 constant_acceleration;
  int32_t grid_size_x;
  int32_t grid_size_y;
  int32_t grid_size_z;
  uint32_t grid_cell_count;
  clsph_float3 min_point;
  clsph_float3 max_point;
} clsph_simulation_parameters;

typedef struct clsph_precomputed_kernel_values {
  float poly_6;
  float poly_6_gradient;
  float poly_6_laplacian;
  float spiky;
  float viscosity;
} clsph_precomputed_kernel_values;

#ifdef __cplusplus
}
#define CLSPH_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define CLSPH_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

CLSPH_STATIC_ASSERT(sizeof(clsph_particle) == 80, "particle is 80 bytes");
CLSPH_STATIC_ASSERT(sizeof(clsph_simulation_parameters) == 128,
                    "simulation_parameters is 128 bytes");
CLSPH_STATIC_ASSERT(sizeof(clsph_precomputed_kernel_values) == 20,
                    "precomputed_kernel_values is 20 bytes");

#endif /* CLSPH_TYPES_H_ */
