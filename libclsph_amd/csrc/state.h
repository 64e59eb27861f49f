// Host-side view of a simulation's device state and the stage launchers.
#pragma once

#include <string>

#include "common.h"

struct clsph_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  int cus = 256;  // compute units on the device
};

struct clsph_state {
  clsph_ctx* ctx = nullptr;
  uint32_t n_cap = 0;
  uint32_t n = 0;
  uint32_t table_cap = 0;

  // Canonical particle arrays, in the order the reference's host array holds
  // them at the start of a substep (the previous substep's sorted order).
  float4* pos = nullptr;  // x, y, z, 0
  float4* vel = nullptr;  // velocity
  float4* vh = nullptr;   // intermediate_velocity

  // Sorted working arrays for the current substep.
  float4* pos_s = nullptr;
  float4* vel_s = nullptr;
  float4* vh_s = nullptr;
  float4* acc_s = nullptr;  // acceleration from forces
  float* rho_s = nullptr;   // density
  float* prs_s = nullptr;   // pressure
  float4* fa = nullptr;     // force inputs: (x, y, z, p / rho^2)
  float4* fb = nullptr;     // force inputs: (vx, vy, vz, m / rho)
  uint4* nbr = nullptr;           // neighbour lists, 4 per entry: nbr[q * n + i]
  uint32_t* nbr_count = nullptr;  // neighbours per particle

  // Radix sort buffers: keys/values ping-pong; after kSortPasses (odd) passes
  // the sorted keys and permutation sit in key_b / val_b.
  uint32_t* key_a = nullptr;
  uint32_t* val_a = nullptr;
  uint32_t* key_b = nullptr;
  uint32_t* val_b = nullptr;
  uint32_t* hist = nullptr;      // per-tile digit counts, digit-major
  uint32_t* digit_base = nullptr;  // exclusive prefix of digit totals
  uint32_t n_tiles_cap = 0;

  // Half-cell table (8 * table_cap + 1 entries): table[f] = #{keys < f}; the
  // reference's cell table is table[8c]. slot[k]: the canonical position
  // (stable order by cell id) of the particle at sorted position k.
  uint32_t* table = nullptr;
  uint32_t* slot = nullptr;

  // Per-block bounding boxes (6 floats per 256-particle block) written by the
  // bounding-box kernel or by advection, reduced by the grid-parameter kernel.
  float* aabb_partial = nullptr;

  // Collision mesh (uploaded once, sph_simulation.cpp:296-319 re-uploads it
  // every substep).
  float* face_normals = nullptr;
  float* vertices = nullptr;
  uint32_t* indices = nullptr;
  uint32_t face_count = 0;
  uint32_t vert_floats = 0;
  // Uniform triangle grid (CSR: tgrid_start[cell]..tgrid_start[cell+1] index
  // tgrid_tris), see clsph_scene_grid.
  clsph::TriGrid tgrid{};
  uint32_t* tgrid_start = nullptr;
  uint32_t* tgrid_tris = nullptr;

  clsph::DevParams* dp = nullptr;  // device parameter block
  clsph::DevParams host_dp{};      // host copy of the host-set fields

  hipGraphExec_t graph = nullptr;  // one captured substep
  bool graph_dirty = true;

  // Last stage run (-1 after an upload) and whether the canonical
  // acceleration is the zero vector advection leaves behind.
  int stage = -1;
  bool acc_zero = false;
  // True while rho_s / prs_s / key_b / acc_s still hold the uploaded
  // particles in canonical order (no substep has run since the upload).
  bool fields_canonical = true;
};

namespace clsph {

// grid.hip
void launch_aabb(clsph_state* s);
void launch_grid_params(clsph_state* s);
void launch_hash(clsph_state* s);
void launch_permute(clsph_state* s);
void launch_cell_table(clsph_state* s);
// sort.hip
void launch_sort(clsph_state* s);
uint32_t sort_tiles_for(uint32_t n);
// sph.hip
void launch_density(clsph_state* s);
void launch_forces(clsph_state* s);
// advect.hip
void launch_advect(clsph_state* s);

}  // namespace clsph
