// Wave64 and work-group reductions shared by the kernels.
#pragma once

#include "common.h"

namespace clsph {

__device__ inline float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Work-group min/max of (x, y, z) into partial[k * stride + blockIdx.x]
// (k = 0-2 the minima, 3-5 the maxima: one array per bound, so that the
// reader loads consecutive words); NaN coordinates are skipped, like the
// reference's `<` / `>` tests (sph_simulation.cpp:207-217). Every thread of
// the work-group must call it.
__device__ inline void block_aabb_partial(float4 p, bool valid, float* __restrict__ partial,
                                          uint32_t stride) {
  __shared__ float red[6][kMaxBlock / 64];
  const float inf = __builtin_huge_valf();
  float v[6];
  v[0] = (valid && p.x == p.x) ? p.x : inf;
  v[1] = (valid && p.y == p.y) ? p.y : inf;
  v[2] = (valid && p.z == p.z) ? p.z : inf;
  v[3] = (valid && p.x == p.x) ? p.x : -inf;
  v[4] = (valid && p.y == p.y) ? p.y : -inf;
  v[5] = (valid && p.z == p.z) ? p.z : -inf;
  for (int k = 0; k < 3; ++k) v[k] = wave_min(v[k]);
  for (int k = 3; k < 6; ++k) v[k] = wave_max(v[k]);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0)
    for (int k = 0; k < 6; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    int nw = (blockDim.x + 63) >> 6;
    float r = red[k][0];
    for (int q = 1; q < nw; ++q) r = (k < 3) ? fminf(r, red[k][q]) : fmaxf(r, red[k][q]);
    partial[k * stride + blockIdx.x] = r;
  }
}

}  // namespace clsph
