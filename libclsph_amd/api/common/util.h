// Drop-in counterpart of libclsph/common/util.h (:4-62): the 10-bit-per-axis
// Morton (Z-curve) encoding of grid cells and its inverse.
#ifndef CLSPH_API_UTIL_H_
#define CLSPH_API_UTIL_H_

#include <stdint.h>

#include "structures.h"

// Spreads the low 10 bits of v so bit k lands at bit 3k.
inline uint32_t clsph_spread10(uint32_t v) {
  v &= 0x3FFu;
  uint32_t r = 0;
  for (int k = 0; k < 10; ++k) r |= ((v >> k) & 1u) << (3 * k);
  return r;
}

// Collects bits 0, 3, 6, ..., 27 of v into bits 0..9.
inline cl_uint uninterleave(cl_uint value) {
  uint32_t r = 0;
  for (int k = 0; k < 10; ++k) r |= ((value >> (3 * k)) & 1u) << k;
  return r;
}

inline cl_uint3 get_cell_coords_z_curve(cl_uint index) {
  cl_uint3 c;
  c.s[0] = uninterleave(index & 0x9249249u);
  c.s[1] = uninterleave((index >> 1) & 0x9249249u);
  c.s[2] = uninterleave((index >> 2) & 0x9249249u);
  c.s[3] = 0;
  return c;
}

inline cl_uint get_grid_index_z_curve(cl_uint in_x, cl_uint in_y, cl_uint in_z) {
  return clsph_spread10(in_x) | (clsph_spread10(in_y) << 1) | (clsph_spread10(in_z) << 2);
}

#endif  // CLSPH_API_UTIL_H_
