// Host-side view of a simulation's device state and the stage launchers.
//
// A state holds one or more independent scenes (a batch). Scene s owns the
// particles [begin_s, begin_s + n_s) of every particle array, in every
// order; its half-cell keys occupy [key_base_s, key_base_s + 8 G_s) of one
// shared key space, so the sort, the tables and the canonical slots serve
// all scenes at once. Per-particle kernels run over a table of work-groups,
// each inside one scene.
#pragma once

#include <string>
#include <vector>

#include "common.h"

struct clsph_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  int cus = 256;  // compute units on the device
  // Device address of the static workspace of clsph_nn_gemm_noalloc on this
  // context's device (shared by every context on the device).
  void* scratch = nullptr;
};

// Stage indices (clsph_time_stage order).
enum { kStageGrid, kStageHash, kStageSort, kStageTable, kStageDensity, kStageForces,
       kStageAdvect, kNumStages };

struct clsph_state {
  clsph_ctx* ctx = nullptr;
  uint32_t n_scenes = 1;
  bool fixed_counts = false;  // batches keep the per-scene counts given at allocation
  uint32_t n_cap = 0;         // particles the arrays hold
  uint32_t n = 0;             // particles in all scenes
  std::vector<uint32_t> scene_begin, scene_n, scene_cap;
  uint32_t table_cap = 0;     // half-cell keys the table holds (all scenes)

  // Canonical particle arrays, in the order the reference's host array holds
  // them at the start of a substep (the previous substep's sorted order).
  float4* pos = nullptr;  // x, y, z, 0
  float4* vel = nullptr;  // velocity
  float4* vh = nullptr;   // intermediate_velocity

  // Sorted working arrays for the current substep (half-cell order).
  float4* pos_s = nullptr;
  float4* vel_s = nullptr;
  float4* vh_s = nullptr;
  float4* acc_s = nullptr;  // acceleration from forces
  float* rho_s = nullptr;   // density
  float* prs_s = nullptr;   // pressure
  // force inputs, interleaved so that a neighbour's two records share a
  // cache line: fab[2i] = (x, y, z, p / rho^2), fab[2i + 1] = (vx, vy, vz, m / rho)
  float4* fab = nullptr;
  uint4* nbr = nullptr;           // neighbour lists, 4 per entry: nbr[q * n + i]
  uint32_t* nbr_count = nullptr;  // neighbours per particle

  // Sort buffers: half-cell keys in canonical order (key_a); the LSD passes'
  // outputs (key_t / val_t after odd passes, key_u / val_u after even ones,
  // values = canonical indices); the sorted keys and permutation (key_b /
  // val_b: sorted position -> canonical index).
  uint32_t* key_a = nullptr;
  uint32_t* key_t = nullptr;
  uint32_t* val_t = nullptr;
  uint32_t* key_u = nullptr;
  uint32_t* val_u = nullptr;
  uint32_t* key_b = nullptr;
  uint32_t* val_b = nullptr;
  uint32_t sort_passes = 2;           // LSD passes, from the table capacity
  uint32_t* sort_hist = nullptr;      // per-tile digit counts, then offsets (tiles x kMaxBins)
  uint32_t* sort_totals = nullptr;    // digit totals of the current pass (kMaxBins)
  // long runs of empty keys of the half-cell table, listed by the cell pass
  // and written by the fill pass ({first key, end key, value, 0}); the list's
  // length (emptied before each cell pass) and a spare word
  uint4* sort_runs = nullptr;
  uint32_t* sort_runs_ctl = nullptr;
  uint32_t sort_runs_cap = 0;
  uint32_t n_tiles_cap = 0;

  // Half-cell table (table_cap + 1 entries): table[f] = #{keys < f}; the
  // reference's cell table of scene s is table[key_base_s + 8c] - begin_s.
  // slot[k]: the canonical position (stable order by cell id) of the
  // particle at sorted position k.
  uint32_t* table = nullptr;
  uint32_t* slot = nullptr;

  // Optional particle identities in canonical order (clsph_track_ids): after
  // each substep ids[slot[k]] takes the id of the particle sorted to k.
  uint32_t* ids = nullptr;
  uint32_t* ids_tmp = nullptr;

  // Per-work-group bounding boxes written by the bounding-box kernel or by
  // advection over the advection work-group table, reduced per scene by the
  // grid-parameter kernel: bound k (min x, y, z, max x, y, z) of work-group b
  // at aabb_partial[k * aabb_stride + b].
  float* aabb_partial = nullptr;
  uint32_t aabb_stride = 0;

  // Work-group tables, one per block size (kBlockSizes), and the block size
  // each stage runs with. Advection and the bounding-box kernel share one.
  clsph::BlockDesc* blocks[clsph::kNumBlockSizes] = {};
  uint32_t nblocks[clsph::kNumBlockSizes] = {};
  // indices into kBlockSizes: 256 threads, 512 for density (two 80 KB
  // halo-staging work-groups per CU) and forces, measured fastest at 1M
  // particles
  int stage_bs[kNumStages] = {2, 2, 2, 2, 3, 3, 2};

  // Collision meshes of all scenes, concatenated (uploaded once;
  // sph_simulation.cpp:296-319 re-uploads it every substep).
  float* face_normals = nullptr;
  float* vertices = nullptr;
  uint32_t* indices = nullptr;
  uint32_t* tgrid_start = nullptr;  // triangle-grid cells of all scenes (CSR)
  uint32_t* tgrid_tris = nullptr;
  struct Mesh {
    std::vector<float> normals, vertices;
    std::vector<uint32_t> indices;
    float grid_cell = -1.f;  // < 0: no grid; 0: automatic; > 0: cell size
  };
  std::vector<Mesh> meshes;

  clsph::DevParams* dps = nullptr;   // device parameters, one per scene
  clsph::BatchParams* bp = nullptr;  // device batch parameters
  std::vector<clsph::DevParams> host_dps;

  hipGraphExec_t graph = nullptr;  // one captured substep
  bool graph_dirty = true;
  // One captured learned macro step (clsph_learned_step), for the network
  // with serial lgraph_serial at lgraph_n particles; parameters are baked in.
  hipGraphExec_t lgraph = nullptr;
  bool lgraph_dirty = true;
  uint64_t lgraph_serial = 0;
  uint32_t lgraph_n = 0;

  // Asynchronous downloads (clsph_download_begin / clsph_download_wait):
  // per slot, a device staging buffer of particle records, a device and a
  // pinned host snapshot of the scene's DevParams and the BatchParams, and
  // the events ordering pack -> copy -> host; one copy stream. Allocated on
  // first use.
  clsph_particle* stage_aos[2] = {};
  uint8_t* stage_meta[2] = {};
  uint8_t* host_meta[2] = {};
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_pack[2] = {};
  hipEvent_t ev_done[2] = {};
  bool slot_used[2] = {};

  // Last stage run (-1 after an upload) and whether the canonical
  // acceleration is the zero vector advection leaves behind.
  int stage = -1;
  bool acc_zero = false;
  // True once the learned surrogate has added its correction to the forces
  // stage's accelerations (cleared by every stage).
  bool acc_learned = false;
  // True while rho_s / prs_s / key_b / acc_s still hold the uploaded
  // particles in canonical order (no substep has run since the upload).
  bool fields_canonical = true;

  int bs_of(int stage_index) const { return clsph::kBlockSizes[stage_bs[stage_index]]; }
  const clsph::BlockDesc* table_of(int stage_index) const { return blocks[stage_bs[stage_index]]; }
  uint32_t nblocks_of(int stage_index) const { return nblocks[stage_bs[stage_index]]; }
};

namespace clsph {

// grid.hip
void launch_aabb(clsph_state* s);
void launch_grid_params(clsph_state* s);
void launch_hash(clsph_state* s);
void launch_track_ids(clsph_state* s);
// Particle records (reference layout, canonical order) of a single-scene
// state into out (device, n records), as the host download assembles them.
void launch_pack_aos(clsph_state* s, clsph_particle* out);
// sort.hip
// The sort: sorted keys and permutation, the half-cell table, the canonical
// slots and the state gathered into sorted order.
void launch_sort(clsph_state* s);
// Its last pass alone (table, slots and gather again from the same grouping).
void launch_cell_table(clsph_state* s);
uint32_t sort_tiles_for(uint32_t n);
uint32_t sort_runs_cap_for(uint32_t table_cap);
// sph.hip
// With `nu`, also the learned surrogate's node inputs of every particle.
void launch_density(clsph_state* s, float2* nu = nullptr);
void launch_forces(clsph_state* s);
// The forces stage for the particles whose neighbour list overflowed only.
void launch_forces_overflow(clsph_state* s);
// advect.hip
void launch_advect(clsph_state* s);

}  // namespace clsph
