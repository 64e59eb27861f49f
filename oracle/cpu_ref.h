/*
 * CPU oracle for the libclsph per-step SPH pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * algorithm, used as the checker by tests/, by the smoke check and by the
 * bench's cpu_baseline leg. The product path (libclsph_amd) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates. The FP32
 * functions follow the reference's single-precision arithmetic (OpenCL C
 * float, no contraction); the FP64 functions run the same algorithm in
 * double and are used to report the FP32 error of each quantity.
 *
 * Implementation choices where the reference leaves arithmetic to the OpenCL
 * runtime (parity unpinned at the ulp level, absorbed by the tolerances):
 *   length(v)    = sqrt(x*x + y*y + z*z)  (correctly rounded sqrt)
 *   normalize(v) = v / length(v)
 *   pown(x, n)   = x * x * ... * x, left to right
 * The collision do-while of advection_collision (libclsph/kernels/sph.cl:82-101)
 * has no iteration cap in the reference; here it is capped by `max_bounces`
 * and the number of particles that hit the cap is returned.
 */
#ifndef CLSPH_ORACLE_CPU_REF_H_
#define CLSPH_ORACLE_CPU_REF_H_

#include <stdint.h>
#include "../include/clsph_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef clsph_particle particle;
typedef clsph_simulation_parameters simulation_parameters;
typedef clsph_precomputed_kernel_values precomputed_kernel_values;

/* Double-precision twin of `particle` for the FP64 oracle. */
typedef struct orc_particle64 {
  double position[4];
  double velocity[4];
  double intermediate_velocity[4];
  double acceleration[4];
  double density;
  double pressure;
  uint32_t grid_index;
} orc_particle64;

/* ---- integer / grid part (FP32 arithmetic where the reference has it) ---- */
uint32_t orc_morton(uint32_t x, uint32_t y, uint32_t z);
void orc_morton_decode(uint32_t key, uint32_t out[3]);

/* AABB, padding and grid dimensions: sph_simulation.cpp:201-252.
 * Returns 0, or -1 where the reference's assert (dimension < 1024) fires. */
int orc_grid_params(const particle* in, uint32_t n,
                    simulation_parameters* params);

/* locate_in_grid: grid.cl:43-67 */
void orc_locate_in_grid(const particle* in, particle* out, uint32_t n,
                        const simulation_parameters* params);

/* sort_particles passes: sort.cl:3-59 with the host scan of
 * sph_simulation.cpp:114-156 (prefix entries sized to the host buffer). */
void orc_sort_particles(const particle* in, particle* out, particle* scratch,
                        uint32_t n);

/* cell table: sph_simulation.cpp:162-170 */
void orc_cell_table(const particle* sorted, uint32_t n, uint32_t cell_count,
                    uint32_t* table);

/* ---- physics, FP32 ---- */
void orc_density_pressure_f32(const particle* in, particle* out, uint32_t n,
                              const simulation_parameters* params,
                              const precomputed_kernel_values* terms,
                              const uint32_t* cell_table);
void orc_forces_f32(const particle* in, particle* out, uint32_t n,
                    const simulation_parameters* params,
                    const precomputed_kernel_values* terms,
                    const uint32_t* cell_table);
/* Returns the number of particles whose collision loop hit max_bounces. */
uint32_t orc_advection_collision_f32(
    const particle* in, particle* out, uint32_t n,
    const simulation_parameters* params, const float* face_normals,
    const float* vertices, const uint32_t* indices, uint32_t face_count,
    uint32_t max_bounces);

/* ---- physics, FP64 (same algorithm, double arithmetic) ---- */
void orc_to_f64(const particle* in, orc_particle64* out, uint32_t n);
void orc_density_pressure_f64(const orc_particle64* in, orc_particle64* out,
                              uint32_t n, const simulation_parameters* params,
                              const precomputed_kernel_values* terms,
                              const uint32_t* cell_table);
void orc_forces_f64(const orc_particle64* in, orc_particle64* out, uint32_t n,
                    const simulation_parameters* params,
                    const precomputed_kernel_values* terms,
                    const uint32_t* cell_table);
uint32_t orc_advection_collision_f64(
    const orc_particle64* in, orc_particle64* out, uint32_t n,
    const simulation_parameters* params, const float* face_normals,
    const float* vertices, const uint32_t* indices, uint32_t face_count,
    uint32_t max_bounces);

/* One substep, sph_simulation.cpp:173-344: `state` holds the particles on
 * entry and the advected particles (in the new sorted order) on exit.
 * `params` receives the step's grid fields. `work` and `work2` hold n
 * particles each; `table` holds at least `table_cap` entries.
 * Returns 0, -1 on a grid-dimension overflow, -2 if table_cap is too small.
 * `capped` (may be NULL) receives the bounce-cap count. */
int orc_step_f32(particle* state, particle* work, particle* work2,
                 uint32_t* table, uint32_t table_cap,
                 simulation_parameters* params,
                 const precomputed_kernel_values* terms,
                 const float* face_normals, const float* vertices,
                 const uint32_t* indices, uint32_t face_count,
                 uint32_t max_bounces, uint32_t* capped);

/* Runs `steps` substeps and records, after each, the mean density and the
 * kinetic energy 0.5*m*sum|velocity|^2 (both accumulated in double). */
int orc_run_f32(particle* state, uint32_t steps,
                simulation_parameters* params,
                const precomputed_kernel_values* terms,
                const float* face_normals, const float* vertices,
                const uint32_t* indices, uint32_t face_count,
                uint32_t max_bounces, double* mean_density, double* kinetic);

/* Number of OpenMP threads the oracle runs on (1 when built without it). */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif /* CLSPH_ORACLE_CPU_REF_H_ */
