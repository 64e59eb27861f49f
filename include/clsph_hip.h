/*
 * clsph_hip: C ABI of the MI355X (gfx950) SPH substep pipeline.
 *
 * This is the layer that replaces the reference's OpenCL boundary
 * (util/cl_boilerplate.{h,cpp} and the cl::CommandQueue calls of
 * libclsph/sph_simulation.cpp). Plain pointers and sizes only; every call
 * returns 0 on success or a negative CLSPH_E* code, with a message available
 * from clsph_last_error(). Calls on one context are ordered on its stream.
 *
 * Ownership: the caller owns host arrays; the library owns device buffers.
 * Threading: one context per host thread; a state belongs to one context.
 */
#ifndef CLSPH_HIP_H_
#define CLSPH_HIP_H_

#include <stddef.h>
#include <stdint.h>

#include "clsph_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CLSPH_OK = 0,
  CLSPH_EINVAL = -1,      /* bad argument */
  CLSPH_EDEVICE = -2,     /* no usable GPU / HIP runtime error */
  CLSPH_ENOMEM = -3,      /* device allocation failed */
  CLSPH_EGRID = -4,       /* a grid dimension reached 1024 (sph_simulation.cpp:247-249) */
  CLSPH_ETABLE = -5,      /* grid_cell_count exceeded the cell table capacity */
};

typedef struct clsph_ctx clsph_ctx;
typedef struct clsph_state clsph_state;

/* ---- context: replaces init_cl_single_device (util/cl_boilerplate.h:41,
 * .cpp:3-63) and make_program (util/cl_boilerplate.h:42, .cpp:65-88); the
 * kernels are compiled ahead of time for gfx950, so there is no runtime
 * program build. `device` < 0 picks device 0, like the reference's
 * "first platform, first device" default. */
int clsph_init(int device, clsph_ctx** out);
int clsph_free(clsph_ctx* ctx);
/* Message of the last failing call on this context ("" if none). Replaces the
 * check_cl_error print (util/cl_boilerplate.h:22-34). */
const char* clsph_last_error(const clsph_ctx* ctx);
/* Name of the device the context runs on. */
int clsph_device_name(const clsph_ctx* ctx, char* buf, size_t len);
/* Waits for the context's stream and reports kernel-raised errors. */
int clsph_sync(clsph_ctx* ctx);

/* ---- simulation state: replaces the front/back/sort_count cl::Buffers
 * (sph_simulation.cpp:373-380). A state holds up to n_cap particles of one
 * scene. `table_cap` is the cell-table capacity in cells of side 2h
 * (0 = max(2^20, 4 n_cap)); a substep whose grid has more cells fails with
 * CLSPH_ETABLE. */
int clsph_state_alloc(clsph_ctx* ctx, uint32_t n_cap, uint32_t table_cap,
                      clsph_state** out);
int clsph_state_free(clsph_state* st);

/* Host parameters (the fields load_settings fills, sph_simulation.cpp:405-506);
 * the grid fields are ignored and recomputed on the device each substep.
 * `max_bounces` caps the collision loop (0 = 64). */
int clsph_set_params(clsph_state* st, const clsph_simulation_parameters* params,
                     const clsph_precomputed_kernel_values* terms,
                     uint32_t max_bounces);

/* Particle transfers in the reference's 80-byte layout. Upload replaces the
 * write of sph_simulation.cpp:195-197 and computes the bounding box for the
 * next substep; download replaces the read of :339-341 (particles come back in
 * the last substep's sorted order with acceleration 0, as the reference's
 * advection writes them). */
int clsph_upload_aos(clsph_state* st, const clsph_particle* particles, uint32_t n);
int clsph_download_aos(clsph_state* st, clsph_particle* particles, uint32_t n);

/* Collision mesh: replaces the per-substep buffer creation and writes of
 * sph_simulation.cpp:296-319 (uploaded once here). `indices` holds 3 vertex
 * indices per face, `vertices` 3 floats per vertex, `face_normals` 3 floats
 * per face. */
int clsph_scene_upload(clsph_state* st, const float* face_normals,
                       const float* vertices, uint32_t vertex_count,
                       const uint32_t* indices, uint32_t face_count);

/* Collision acceleration structure for the uploaded mesh. `cell_size` > 0
 * builds a uniform grid of that cell size over the mesh (each cell lists the
 * triangles whose bounding box, padded by the rounding margin, meets it);
 * 0 picks a size from the mean triangle extent; < 0 removes the grid so
 * every substep tests every triangle, as advection_collision does
 * (collisions.cl:21-87). clsph_scene_upload builds a grid by itself for
 * meshes of more than 64 triangles. Hits are the same either way: the
 * nearest hit, ties going to the highest triangle index. */
int clsph_scene_grid(clsph_state* st, float cell_size);

/* ---- stages of one substep (simulate_single_frame, sph_simulation.cpp:173-344) */
/* AABB + padding + grid dimensions: sph_simulation.cpp:199-252 */
int clsph_grid(clsph_state* st);
/* locate_in_grid: grid.cl:43-67, launched at sph_simulation.cpp:257-262 */
int clsph_hash(clsph_state* st);
/* sort_particles' radix passes (sph_simulation.cpp:107-160, sort.cl:23-59);
 * also gathers the state into sorted order */
int clsph_sort(clsph_state* st);
/* cell table: sph_simulation.cpp:162-170, 270-276 */
int clsph_cell_table(clsph_state* st);
/* density_pressure: sph.cl:9-40, launched at sph_simulation.cpp:283-294 */
int clsph_density_pressure(clsph_state* st);
/* forces: sph.cl:42-62, launched at sph_simulation.cpp:322-327 */
int clsph_forces(clsph_state* st);
/* advection_collision: sph.cl:64-112, launched at sph_simulation.cpp:330-337 */
int clsph_advect_collide(clsph_state* st);
/* `nsteps` whole substeps (all stages above), replayed from one captured
 * HIP graph; replaces simulate_single_frame's body. */
int clsph_step(clsph_state* st, uint32_t nsteps);

/* ---- batches: many independent scenes advanced together. No counterpart
 * in the reference, which runs one scene per process; every stage runs as one
 * launch over all scenes. Scene s holds exactly counts[s] particles; each
 * scene has its own parameters, mesh, grid and cell table, and its results
 * are those of a state holding that scene alone. `table_cap` counts cells of
 * all scenes together (0 = max(2^20, 4 * sum(counts))). The single-scene
 * calls above act on scene 0. */
int clsph_batch_alloc(clsph_ctx* ctx, uint32_t n_scenes, const uint32_t* counts,
                      uint32_t table_cap, clsph_state** out);
int clsph_batch_set_params(clsph_state* st, uint32_t scene,
                           const clsph_simulation_parameters* params,
                           const clsph_precomputed_kernel_values* terms, uint32_t max_bounces);
int clsph_batch_upload(clsph_state* st, uint32_t scene, const clsph_particle* particles,
                       uint32_t n);
int clsph_batch_download(clsph_state* st, uint32_t scene, clsph_particle* particles,
                         uint32_t n);
int clsph_batch_scene_upload(clsph_state* st, uint32_t scene, const float* face_normals,
                             const float* vertices, uint32_t vertex_count,
                             const uint32_t* indices, uint32_t face_count);
int clsph_batch_get_grid(clsph_state* st, uint32_t scene, clsph_simulation_parameters* params);
int clsph_batch_bounce_capped(clsph_state* st, uint32_t scene, uint32_t* count);
/* Work-group size of each stage's per-particle kernels (stage order as in
 * clsph_time_stage; 64, 128, 256 or 512; 0 keeps the current one). The
 * autotuner sets these; results do not depend on them. */
int clsph_set_block_sizes(clsph_state* st, const uint32_t sizes[7]);
int clsph_get_block_sizes(const clsph_state* st, uint32_t sizes[7]);

/* ---- readback of stage results (for checks and the save delegates) ---- */
/* Device-computed grid fields of the current substep, written into *params. */
int clsph_get_grid(clsph_state* st, clsph_simulation_parameters* params);
/* The kernels' own particle order of the current substep (half-cell order:
 * by cell id, then octant, stable): for internal position k, slot[k] is its
 * position in the reference's sorted order (clsph_read_sorted,
 * clsph_download_sorted_aos) and perm[k] its canonical index. The learned
 * surrogate's graph, outputs and targets are indexed this way. */
int clsph_read_order(clsph_state* st, uint32_t* slot, uint32_t* perm, uint32_t n);
/* Cell ids in canonical (pre-sort) order after clsph_hash. */
int clsph_read_keys(clsph_state* st, uint32_t* keys, uint32_t n);
/* Sorted cell ids and the permutation (sorted slot -> canonical index). */
int clsph_read_sorted(clsph_state* st, uint32_t* keys, uint32_t* perm, uint32_t n);
/* Cell table entries [0, count). */
int clsph_read_cell_table(clsph_state* st, uint32_t* table, uint32_t count);
/* The sorted particles as the reference's front/back buffer holds them after
 * the last stage run: position, velocity, intermediate_velocity, acceleration
 * (from forces), density, pressure, grid_index. */
int clsph_download_sorted_aos(clsph_state* st, clsph_particle* particles, uint32_t n);
/* Particles whose collision loop reached max_bounces in the last substep. */
int clsph_bounce_capped(clsph_state* st, uint32_t* count);

/* ---- timing on the context's stream (HIP events) ---- */
/* Times `nsteps` graph-replayed substeps; writes elapsed milliseconds. */
int clsph_time_steps(clsph_state* st, uint32_t nsteps, float* ms);
/* Times one stage kernel sequence `reps` times in a row (stage: 0 grid,
 * 1 hash, 2 sort, 3 cell table, 4 density, 5 forces, 6 advection); writes
 * the average milliseconds per repetition. Stages are idempotent on the
 * state except advection, which advances it. */
int clsph_time_stage(clsph_state* st, int stage, uint32_t reps, float* ms);

/* Runs `nsteps` whole substeps stage by stage with HIP events recorded
 * between stages on the context stream; writes the average milliseconds
 * per substep of each stage into stage_ms[0..6] (stage order as in
 * clsph_time_stage; stage 2 includes the state gather into sorted order). */
int clsph_profile_steps(clsph_state* st, uint32_t nsteps, float* stage_ms);

/* Number of ordered neighbour pairs (r < h, self excluded) the last density
 * stage found, summed over all particles: the pair interactions of density
 * and forces. A diagnostic with no counterpart in the reference. */
int clsph_neighbour_pairs(clsph_state* st, uint64_t* pairs);

/* ---- asynchronous frame downloads ----
 * The reference reads the particle buffer back after every substep
 * (sph_simulation.cpp:339-341) synchronously. These copy the particles, in
 * the layout and order of clsph_download_aos, into page-locked host memory
 * on a second stream: the kernels of the following substeps run while the
 * copy is in flight and while the host works on an earlier copy. Two slots
 * (0 and 1) double-buffer the device staging. Single-scene states; the host
 * array must stay valid until clsph_download_wait returns. */
int clsph_host_alloc(clsph_ctx* ctx, size_t bytes, void** out);
int clsph_host_free(clsph_ctx* ctx, void* p);
int clsph_download_begin(clsph_state* st, uint32_t slot, clsph_particle* out, uint32_t n);
/* Waits for the download of `slot` and writes the simulation_parameters of
 * that substep (as clsph_get_grid) to *grid (optional); returns the device
 * error of that substep, if any. */
int clsph_download_wait(clsph_state* st, uint32_t slot, clsph_simulation_parameters* grid);

/* ---- learned force surrogate ----
 * No counterpart in the reference, whose only acceleration model is the
 * `forces` kernel (kernels/sph.cl:42-62). A one-round message-passing
 * network over the neighbour graph of the density stage (r < h, at most 128
 * neighbours per particle) corrects the forces stage's acceleration a0 so
 * that it stands for the mean acceleration over a macro step of 4 substeps:
 * a = a0 + 2 out[0:3] (m/s^2). Its products run on bf16 matrix cores with
 * float32 accumulation; with all parameters zero it returns a0 unchanged.
 * Parameters are one float32 vector of clsph_nn_param_count() values
 * (row-major inputs x outputs matrices and biases in a fixed order).
 * Single-scene states only. */
typedef struct clsph_nn clsph_nn;
uint32_t clsph_nn_param_count(void);
/* Workspace for states of up to n_cap particles and graphs of up to e_cap
 * edges (0: 64 n_cap); parameters start at zero. */
int clsph_nn_create(clsph_ctx* ctx, uint32_t n_cap, uint32_t e_cap, clsph_nn** out);
int clsph_nn_free(clsph_nn* nn);
/* Sets the parameters (and clears the optimiser moments) / reads them. */
int clsph_nn_set_params(clsph_nn* nn, const float* params, uint32_t count);
int clsph_nn_get_params(clsph_nn* nn, float* params, uint32_t count);
/* Training evaluation on a state whose last stage was forces: builds the
 * graph, stores every intermediate the backward pass needs and the output
 * (clsph_nn_read_output), writes the edge count to *edges (optional). The
 * state is not changed. */
int clsph_nn_forward(clsph_nn* nn, clsph_state* st, uint32_t* edges);
/* Inference on a state whose last stage was forces, with one fused kernel
 * (no per-edge storage, any edge count, no host synchronisation): adds the
 * correction to the forces stage's accelerations, which the following
 * advection uses, and stores the output. Once per forces stage. */
int clsph_nn_infer(clsph_nn* nn, clsph_state* st);
/* Times the inference `reps` times with HIP events on the context stream,
 * on a copy of the accelerations (the state is not changed): ms[0] =
 * average milliseconds per evaluation; *edges (optional) = the edges
 * evaluated per pass. The last stage must be forces. */
int clsph_nn_profile(clsph_nn* nn, clsph_state* st, uint32_t reps, float* ms, uint64_t* edges);
/* The graph of the last training evaluation (dst-major CSR: off[n + 1],
 * src[e]) and the raw output of the last evaluation (n x 32 floats; columns
 * 0..2 are (a - a0) / 2). */
int clsph_nn_read_graph(clsph_nn* nn, uint32_t* off, uint32_t* src, uint32_t n, uint32_t e);
int clsph_nn_read_output(clsph_nn* nn, float* out, uint32_t n);
/* Training target (n x 3, in the output's units) from the host, or from two
 * states: `learner` after its forces stage (a0) and `teacher` started as a
 * copy of the learner's particles with identities tracked, then advanced by
 * the plain method over `dt` seconds: ((v_teacher - v_learner) / dt - a0) / 2. */
int clsph_nn_set_target(clsph_nn* nn, const float* target, uint32_t n);
int clsph_nn_read_target(clsph_nn* nn, float* target, uint32_t n);
int clsph_nn_target_from(clsph_nn* nn, clsph_state* learner, clsph_state* teacher, float dt);
/* Target = (mean over `nsub` plain substeps of each particle's forces-stage
 * acceleration - a0) / 2, `learner` after its forces stage: `teacher` (a
 * copy of the learner's particles, identities tracked) is advanced by nsub
 * substeps stage by stage. Collision impulses, which advection applies, are
 * left out. */
int clsph_nn_target_mean_acc(clsph_nn* nn, clsph_state* learner, clsph_state* teacher,
                             uint32_t nsub);
/* Mean squared error of the last evaluation against the target and its
 * gradient with respect to the parameters (*loss optional), then one Adam
 * step (bias-corrected). */
int clsph_nn_backward(clsph_nn* nn, float* loss);
int clsph_nn_read_grad(clsph_nn* nn, float* grad, uint32_t count);
int clsph_nn_adam(clsph_nn* nn, float lr, float beta1, float beta2, float eps);
/* `nsteps` substeps at the state's time_delta, each the plain substep with
 * the network's correction added to the forces stage's accelerations. */
int clsph_learned_step(clsph_state* st, clsph_nn* nn, uint32_t nsteps);
/* `macro_steps` learned steps of `learner` (time_delta = the macro step,
 * 4 plain substeps), judged every `judge_every` steps (0: never) by `judge`
 * (time_delta = the plain substep, identities tracked in both): the judge
 * starts from the learner's particles and runs 4 plain substeps, both get
 * their densities, and when max |rho - rho_judge| / fluid_density exceeds
 * `tol` the learner takes the judge's particles (a fallback). Writes the
 * counts and, per judged step (one float each, optional), the maximum and
 * the mean of |rho - rho_judge| / fluid_density. One host synchronisation
 * per judged step. No counterpart in the reference. */
int clsph_learned_run(clsph_state* learner, clsph_state* judge, clsph_nn* nn,
                      uint32_t macro_steps, uint32_t judge_every, float tol, uint32_t* judged,
                      uint32_t* fallbacks, float* judge_err, float* judge_mean);
/* Density difference of the same particles in two states after both ran
 * their density stage (identities tracked in both): max and mean of
 * |rho_a - rho_b| / fluid_density. */
int clsph_nn_compare_density(clsph_nn* nn, clsph_state* a, clsph_state* b, float* max_rel,
                             float* mean_rel);
/* Copies the particles (and the per-particle fields) of single-scene `src`
 * into `dst`, and their identities when both track them; parameters and
 * meshes of `dst` stay its own. */
int clsph_state_copy(clsph_state* dst, clsph_state* src);
/* Particle identities: enable resets them to the current canonical index
 * (0..n-1) and keeps them up to date through every substep; 0 disables. */
int clsph_track_ids(clsph_state* st, int enable);
int clsph_read_ids(clsph_state* st, uint32_t* ids, uint32_t n);
/* The network's bf16 matrix-core product on host arrays (for testing):
 * c (m x n float32, row-major) = op(a) op(b) [+ bias], or += with
 * `accumulate`. a is m x k bf16 (k x m with trans_a), b is k x n (n x k
 * with trans_b), bf16 given as raw 16-bit patterns; the contiguous
 * dimension of each operand must be a multiple of 8. */
int clsph_nn_gemm(clsph_ctx* ctx, int trans_a, int trans_b, int accumulate, const uint16_t* a,
                  const uint16_t* b, float* c, const float* bias, uint32_t m, uint32_t n,
                  uint32_t k);
/* clsph_nn_gemm without any device allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (harnesses that
 * forbid hipMalloc, tight loops that must not allocate). The device copies of
 * a, b, c and bias go to a static workspace of 8 MiB per GPU, held until the
 * process ends; a call whose operands (each at a 256-byte aligned offset)
 * need more fails with CLSPH_EINVAL. Results equal clsph_nn_gemm's. Every
 * context on a GPU shares that GPU's workspace: calls on different contexts
 * of one GPU that overlap in time race; calls on one context do not (each
 * waits for its stream). */
int clsph_nn_gemm_noalloc(clsph_ctx* ctx, int trans_a, int trans_b, int accumulate,
                          const uint16_t* a, const uint16_t* b, float* c, const float* bias,
                          uint32_t m, uint32_t n, uint32_t k);

/* ---- two-round network ----
 * No counterpart in the reference, whose only acceleration model is the
 * `forces` kernel (kernels/sph.cl:42-62). The full-size network a learned
 * force surrogate can use: two rounds of message passing over the neighbour
 * graph of the density stage (r < h, at most 128 neighbours per particle),
 * three 128-wide SiLU layers per edge and per node network, output
 * a = g + 10 out[0:3] (m/s^2). Its products run on bf16 matrix cores with
 * float32 accumulation. Parameters are one float32 vector of
 * clsph_gnn_param_count() values (row-major inputs x outputs matrices and
 * biases in a fixed order). Single-scene states only. */
typedef struct clsph_gnn clsph_gnn;
uint32_t clsph_gnn_param_count(void);
/* Workspace for states of up to n_cap particles and graphs of up to e_cap
 * edges (0: 64 n_cap); parameters start at zero. */
int clsph_gnn_create(clsph_ctx* ctx, uint32_t n_cap, uint32_t e_cap, clsph_gnn** out);
int clsph_gnn_free(clsph_gnn* nn);
/* Sets the parameters (and clears the optimiser moments) / reads them. */
int clsph_gnn_set_params(clsph_gnn* nn, const float* params, uint32_t count);
int clsph_gnn_get_params(clsph_gnn* nn, float* params, uint32_t count);
/* Evaluates the network on a state whose last stage was density: builds the
 * graph, writes the accelerations where the forces stage would (the
 * following advection uses them) and the edge count to *edges (optional). */
int clsph_gnn_forward(clsph_gnn* nn, clsph_state* st, uint32_t* edges);
/* The same evaluation with the fused inference kernels: no per-edge
 * storage (any edge count), no host synchronisation, no gradients. */
int clsph_gnn_infer(clsph_gnn* nn, clsph_state* st);
/* Runs clsph_gnn_infer `reps` times with HIP events on the context stream;
 * ms[0..3] = average milliseconds of edge round 1, the node part of round 1,
 * edge round 2 and the node part of round 2; *edges (optional) = the edges
 * evaluated per pass. */
int clsph_gnn_profile(clsph_gnn* nn, clsph_state* st, uint32_t reps, float* ms, uint64_t* edges);
/* The graph of the last evaluation (dst-major CSR: off[n + 1], src[e]) and
 * the raw network output (n x 32 floats; columns 0..2 are (a - g) / 10). */
int clsph_gnn_read_graph(clsph_gnn* nn, uint32_t* off, uint32_t* src, uint32_t n, uint32_t e);
int clsph_gnn_read_output(clsph_gnn* nn, float* out, uint32_t n);
/* Training target (n x 3, in the output's units). */
int clsph_gnn_set_target(clsph_gnn* nn, const float* target, uint32_t n);
int clsph_gnn_read_target(clsph_gnn* nn, float* target, uint32_t n);
/* Mean squared error of the last clsph_gnn_forward against the target and
 * its gradient with respect to the parameters (*loss optional), then one
 * Adam step (bias-corrected). */
int clsph_gnn_backward(clsph_gnn* nn, float* loss);
int clsph_gnn_read_grad(clsph_gnn* nn, float* grad, uint32_t count);
int clsph_gnn_adam(clsph_gnn* nn, float lr, float beta1, float beta2, float eps);

/* ---- scene builders (host only; no device needed) ---- */
/* Dam-break block: a cubic lattice of `n` particles at `spacing` with
 * round(extent/spacing) sites along x and z, stacked layer by layer from
 * origin upward in y (x fastest, then z, then y) until n are placed; site
 * (i, j, k) sits at origin + (i + 0.5, j + 0.5, k + 0.5) * spacing, each
 * coordinate jittered by U(-jitter, +jitter) * spacing from a counter-based
 * generator seeded by `seed` (extent[1] is not used). Velocities and all
 * other fields are zero. Returns n. */
uint32_t clsph_make_block(clsph_particle* out, uint32_t n, float spacing,
                          const float origin[3], const float extent[3],
                          float jitter, uint64_t seed);
/* Smoothing-kernel constants for support radius h, with the reference's
 * expressions (sph_simulation.cpp:499-505). There the unqualified
 * pow(h, 9.f) / pow(h, 6.f) resolve to ::pow(double, double) under
 * g++/libstdc++, so h^9, h^6, M_PI and every quotient are double and each
 * constant is rounded once to float on store. */
void clsph_kernel_terms(float h, clsph_precomputed_kernel_values* out);
/* Heightfield over [x0, x0+size] x [z0, z0+size]: res x res quads, 2
 * triangles each (2 res^2 faces, (res+1)^2 vertices), heights
 * y = amplitude * N(x, z) with N in [0, 1) a `octaves`-octave value noise
 * (lattice values from a counter-based generator seeded by `seed`, smooth
 * bilinear interpolation, frequency doubling and amplitude halving per
 * octave, base lattice of 4 cells across). Normals as scene.cpp:36-64
 * computes them, oriented upward. Writes (res+1)^2*3 vertex floats, 6 res^2
 * indices and 6 res^2 normal floats. */
void clsph_make_heightfield(uint32_t res, float x0, float z0, float size, float amplitude,
                            uint32_t octaves, uint64_t seed, float* vertices,
                            uint32_t* indices, float* face_normals);
/* Column of n particles on a cubic lattice at `spacing`: round(extent/spacing)
 * sites along x and z from origin, stacked upward in y, jittered by
 * U(+-jitter)*spacing, each with velocity (U(-vh, vh), -U(v0, v1), U(-vh, vh))
 * from a counter-based generator seeded by `seed` (intermediate_velocity
 * equal to velocity). Returns n. */
uint32_t clsph_make_column(clsph_particle* out, uint32_t n, float spacing, const float origin[3],
                           const float extent[3], float jitter, float vh, float v0, float v1,
                           uint64_t seed);
/* Axis-aligned box [lo, hi] as 12 triangles with inward unit normals:
 * fills 8*3 vertex floats, 12*3 indices and 12*3 normal floats. */
void clsph_make_box(const float lo[3], const float hi[3], float* vertices,
                    uint32_t* indices, float* face_normals);

#ifdef __cplusplus
}
#endif
#endif /* CLSPH_HIP_H_ */
