// Device-side shared definitions for the clsph HIP kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/clsph_types.h"

namespace clsph {

// Error bits raised by kernels into DevParams::err; read back by the host at
// synchronisation points.
enum : uint32_t {
  kErrGridDim = 1u,        // a grid dimension reached 1024 (sph_simulation.cpp:247-249)
  kErrTableCap = 2u,       // grid_cell_count exceeds the cell table allocation
  kErrFineKey = 4u,        // half-cell keys need more than 32 bits (grid_cell_count >= 2^29)
};

// Particles are sorted by the Morton key of the half-size cell (side h) they
// fall in. Its top bits are the reference's cell id (side 2h), so the order
// groups each cell's particles by octant; the reference's stable order by
// cell id is recovered from it per particle.
constexpr int kSortPasses = 3;       // LSD passes; digit width adapts to the key width
constexpr int kMaxDigitBits = 11;    // 3 x 11 bits covers any 32-bit key
constexpr int kMaxBins = 1 << kMaxDigitBits;
// Neighbour-list slots per particle; a particle with more neighbours is
// handled by forces with a cell walk instead.
constexpr uint32_t kMaxNeighbours = 128;

// Device-resident parameter block. `p` mirrors the reference's
// simulation_parameters; its grid fields (min/max point, grid sizes,
// grid_cell_count) are written on the device every step.
struct DevParams {
  clsph_simulation_parameters p;   // 128 B
  clsph_precomputed_kernel_values t;
  uint32_t n;            // particle count
  uint32_t table_cap;    // cell capacity of the tables (half-cell table holds 8x + 1)
  uint32_t key_bits;     // width of the half-cell keys: bitlen(grid_cell_count) + 3
  uint32_t digit_bits;   // per LSD pass
  uint32_t err;          // kErr* bits, sticky until cleared by the host
  uint32_t capped;       // particles whose collision loop hit max_bounces (last step)
  uint32_t max_bounces;
};

// Uniform grid over the collision mesh (cells x fastest, then y, then z).
struct TriGrid {
  float origin[3];
  float inv_cell;
  uint32_t dims[3];
  uint32_t enabled;
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

}  // namespace clsph
