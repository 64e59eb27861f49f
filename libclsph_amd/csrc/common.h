// Device-side shared definitions for the clsph HIP kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/clsph_types.h"

namespace clsph {

// Error bits raised by kernels into DevParams::err / BatchParams::err; read
// back by the host at synchronisation points.
enum : uint32_t {
  kErrGridDim = 1u,        // a grid dimension reached 1024 (sph_simulation.cpp:247-249)
  kErrTableCap = 2u,       // the cells of all scenes exceed the table allocation
  kErrFineKey = 4u,        // half-cell keys need more than 32 bits
};

// Particles are sorted by the Morton key of the half-size cell (side h) they
// fall in. Its top bits are the reference's cell id (side 2h), so the order
// groups each cell's particles by octant; the reference's stable order by
// cell id is recovered from it per particle. The scenes of a batch occupy
// consecutive key ranges, so one sort orders every scene.
constexpr int kMaxDigitBits = 11;   // LSD digits of the sort: at most 2048 bins
constexpr int kMaxBins = 1 << kMaxDigitBits;
constexpr int kMaxSortPasses = 3;   // 3 x 11 bits cover any cell id
// Neighbour-list slots per particle; a particle with more neighbours is
// handled by forces with a cell walk instead.
constexpr uint32_t kMaxNeighbours = 128;
// Work-group sizes a per-particle kernel can be launched with.
constexpr int kNumBlockSizes = 4;
constexpr int kBlockSizes[kNumBlockSizes] = {64, 128, 256, 512};
constexpr int kMaxBlock = 512;

// Uniform grid over a scene's collision mesh (cells x fastest, then y, then
// z); cell_base offsets into the batch's concatenated cell lists.
struct TriGrid {
  float origin[3];
  float inv_cell;
  uint32_t dims[3];
  uint32_t enabled;
  uint32_t cell_base;
};

// Device-resident parameters of one scene. `p` mirrors the reference's
// simulation_parameters; its grid fields (min/max point, grid sizes,
// grid_cell_count) are written on the device every substep.
struct DevParams {
  clsph_simulation_parameters p;   // 128 B
  clsph_precomputed_kernel_values t;
  uint32_t begin, n;     // the scene's particles: [begin, begin + n) in every order
  uint32_t key_base;     // first half-cell key of the scene (a multiple of 8)
  uint32_t key_end;      // key_base + 8 * grid_cell_count
  uint32_t err;          // kErr* bits of this scene, sticky until cleared by the host
  uint32_t capped;       // particles whose collision loop hit max_bounces (last substep)
  uint32_t max_bounces;
  uint32_t face_begin, face_count;  // the scene's triangles in the batch mesh
  uint32_t block_begin, block_end;  // the scene's work-groups in the bounding-box table
  TriGrid tgrid;
};

// Batch-wide values, computed on the device each substep.
struct BatchParams {
  uint32_t n_scenes;
  uint32_t n;            // particles in all scenes
  uint32_t key_total;    // half-cell keys of all scenes: 8 * sum(grid_cell_count + 1)
  uint32_t table_cap;    // half-cell table capacity in keys
  uint32_t key_bits;     // bitlen(key_total)
  uint32_t sort_passes;  // LSD passes over the cell-id bits (set by the host)
  uint32_t digit_bits;   // bits per pass: ceil((key_bits - 3) / lsd_passes), at most 11
  uint32_t err;
  uint32_t ticket;       // grid-parameter work-groups done this substep (zero between)
  uint32_t lsd_passes;   // LSD passes run this substep: sort_passes, or 2 when the cell
                         // ids fit two (the graph's third pass then does nothing)
};

// LSD passes of this substep (sort_passes before the first layout).
__device__ inline uint32_t lsd_passes_of(const BatchParams* __restrict__ bp) {
  return bp->lsd_passes ? bp->lsd_passes : bp->sort_passes;
}

// Work-group descriptor for per-particle kernels: scene and first particle.
struct BlockDesc {
  uint32_t scene;
  uint32_t first;
};

__host__ __device__ inline uint32_t spread10(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
// 10-bit-per-axis Morton code, x at bit 3k (libclsph/common/util.h:41-62).
__host__ __device__ inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
}

__host__ __device__ inline uint32_t bitlen(uint32_t v) {
  uint32_t b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

// locate_in_grid (grid.cl:43-67): the reference's cell id is
// Morton((uint)((x - min) / (h * 2))); the half-cell key adds the octant and
// the scene's key base.
__device__ inline uint32_t hash_key(float4 p, const DevParams* __restrict__ dp) {
#pragma clang fp contract(off)
  float L = dp->p.h * 2;
  float u[3] = {(p.x - dp->p.min_point.s[0]) / L, (p.y - dp->p.min_point.s[1]) / L,
                (p.z - dp->p.min_point.s[2]) / L};
  uint32_t c[3], oct = 0;
  for (int a = 0; a < 3; ++a) {
    c[a] = (uint32_t)u[a];
    oct |= (((uint32_t)(2.f * u[a])) & 1u) << a;
  }
  return dp->key_base + ((morton3(c[0], c[1], c[2]) << 3) | oct);
}

// The learned surrogate's node inputs of a particle: nu = [rho / rho0 - 1,
// asinh(p / K) / 4].
__device__ inline float2 node_in(float rho, float p, float rho0, float K) {
  return make_float2(rho / rho0 - 1.f, asinhf(p / K) * 0.25f);
}

// The chip deals a launch's work-groups to its kXcds XCDs round-robin by
// index (XCD = blockIdx.x % kXcds), and each XCD has an L2 of its own. This
// maps launch index b of G to an entry of the work list such that each XCD
// takes runs of SPAN consecutive entries: launch indices
// [g, g + kXcds SPAN) take the list's entries of that range, XCD x the x-th
// run of them (a last partial range keeps the launch order). Work-groups
// that gather each other's particles then share an L2, while every XCD still
// samples the whole list (one contiguous share per XCD unbalances them:
// particles' work differs across a scene).
constexpr uint32_t kXcds = 8;
template <uint32_t SPAN>
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t G) {
  constexpr uint32_t group = kXcds * SPAN;
  const uint32_t g0 = b - b % group;
  if (g0 + group > G) return b;
  const uint32_t w = b - g0;
  return g0 + (w % kXcds) * SPAN + w / kXcds;
}

// The particle of this thread (0xFFFFFFFF past its scene's end) and its
// scene, from the launch's work list; SPAN > 0: the list taken in
// xcd_block<SPAN> order.
template <uint32_t SPAN = 0>
__device__ inline uint32_t particle_of(const BlockDesc* __restrict__ blocks,
                                       const DevParams* __restrict__ dps, uint32_t& scene) {
  uint32_t e = blockIdx.x;
  if constexpr (SPAN != 0) e = xcd_block<SPAN>(blockIdx.x, gridDim.x);
  BlockDesc b = blocks[e];
  scene = b.scene;
  uint32_t i = b.first + threadIdx.x;
  const DevParams& d = dps[b.scene];
  return i < d.begin + d.n ? i : 0xFFFFFFFFu;
}

}  // namespace clsph
