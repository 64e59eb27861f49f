// Grid stages: bounding box, grid parameters, cell-id hashing, state
// permutation and the cell-start table.
//
// The bounding box / grid arithmetic must reproduce the reference's host FP32
// sequence bit for bit (sph_simulation.cpp:201-252), and the cell id the
// kernel's (grid.cl:43-67): no contraction, correctly rounded division.
#include "state.h"

namespace clsph {

constexpr int kBlock = 256;

__device__ inline float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-level min/max of (x, y, z) written to partial[blockIdx.x * 6 + k].
// NaN coordinates are skipped, like the reference's `<` / `>` tests.
__device__ inline void block_aabb_partial(float4 p, bool valid, float* partial) {
  __shared__ float red[6][kBlock / 64];
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
    float r = red[k][0];
    for (int q = 1; q < kBlock / 64; ++q)
      r = (k < 3) ? fminf(r, red[k][q]) : fmaxf(r, red[k][q]);
    partial[blockIdx.x * 6 + k] = r;
  }
}

__global__ __launch_bounds__(kBlock) void k_aabb(const float4* __restrict__ pos,
                                                 uint32_t n,
                                                 float* __restrict__ partial) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  float4 p = (i < n) ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  block_aabb_partial(p, i < n, partial);
}

// One block: reduces the per-block partial boxes, then restates
// sph_simulation.cpp:201-252 in FP32.
__global__ __launch_bounds__(1024) void k_grid_params(DevParams* __restrict__ dp,
                                                      const float* __restrict__ partial,
                                                      uint32_t n_partial) {
#pragma clang fp contract(off)
  __shared__ float red[6][16];
  const float inf = __builtin_huge_valf();
  float v[6] = {inf, inf, inf, -inf, -inf, -inf};
  for (uint32_t b = threadIdx.x; b < n_partial; b += blockDim.x) {
    for (int k = 0; k < 3; ++k) v[k] = fminf(v[k], partial[b * 6 + k]);
    for (int k = 3; k < 6; ++k) v[k] = fmaxf(v[k], partial[b * 6 + k]);
  }
  for (int k = 0; k < 3; ++k) v[k] = wave_min(v[k]);
  for (int k = 3; k < 6; ++k) v[k] = wave_max(v[k]);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0)
    for (int k = 0; k < 6; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  int nw = blockDim.x >> 6;
  for (int k = 0; k < 6; ++k) {
    float r = red[k][0];
    for (int q = 1; q < nw; ++q) r = (k < 3) ? fminf(r, red[k][q]) : fmaxf(r, red[k][q]);
    v[k] = r;
  }
  // Initial values are numeric_limits<cl_int>::max()/min() as float; a
  // position replaces them only when strictly smaller / larger.
  const float init_min = 2147483648.0f, init_max = -2147483648.0f;
  float L = dp->p.h * 2;
  float mn[3], mx[3];
  uint32_t gs[3];
  uint32_t err = 0;
  for (int a = 0; a < 3; ++a) {
    mn[a] = (v[a] < init_min) ? v[a] : init_min;
    mx[a] = (v[3 + a] > init_max) ? v[3 + a] : init_max;
    mn[a] -= L * 2;
    mx[a] += L * 2;
    gs[a] = (uint32_t)((mx[a] - mn[a]) / L);
    if (gs[a] >= 1024u) err |= kErrGridDim;
    dp->p.min_point.s[a] = mn[a];
    dp->p.max_point.s[a] = mx[a];
  }
  dp->p.min_point.s[3] = 0.f;
  dp->p.max_point.s[3] = 0.f;
  dp->p.grid_size_x = (int32_t)gs[0];
  dp->p.grid_size_y = (int32_t)gs[1];
  dp->p.grid_size_z = (int32_t)gs[2];
  uint32_t G = morton3(gs[0], gs[1], gs[2]);
  dp->p.grid_cell_count = G;
  if (G > dp->table_cap) err |= kErrTableCap;
  uint32_t kb = bitlen(G);
  dp->key_bits = kb;
  uint32_t db = (kb + kSortPasses - 1) / kSortPasses;
  dp->digit_bits = db < 1 ? 1 : db;
  dp->capped = 0;
  if (err) atomicOr(&dp->err, err);
}

// locate_in_grid (grid.cl:43-67): key = Morton((uint)((x - min) / (h * 2))).
__global__ __launch_bounds__(kBlock) void k_hash(const float4* __restrict__ pos,
                                                 const DevParams* __restrict__ dp,
                                                 uint32_t* __restrict__ key,
                                                 uint32_t* __restrict__ val,
                                                 uint32_t n) {
#pragma clang fp contract(off)
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  float L = dp->p.h * 2;
  uint32_t cx = (uint32_t)((p.x - dp->p.min_point.s[0]) / L);
  uint32_t cy = (uint32_t)((p.y - dp->p.min_point.s[1]) / L);
  uint32_t cz = (uint32_t)((p.z - dp->p.min_point.s[2]) / L);
  key[i] = morton3(cx, cy, cz);
  val[i] = i;
}

// Gathers the canonical state into sorted order: s[i] = canonical[perm[i]].
__global__ __launch_bounds__(kBlock) void k_permute(
    const uint32_t* __restrict__ perm, const float4* __restrict__ pos,
    const float4* __restrict__ vel, const float4* __restrict__ vh,
    float4* __restrict__ pos_s, float4* __restrict__ vel_s,
    float4* __restrict__ vh_s, uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t j = perm[i];
  pos_s[i] = pos[j];
  vel_s[i] = vel[j];
  vh_s[i] = vh[j];
}

// Cell-start table T[c] = #{sorted keys < c} for c < grid_cell_count
// (sph_simulation.cpp:162-170): lower_bound over the sorted keys.
__global__ __launch_bounds__(kBlock) void k_cell_table(const uint32_t* __restrict__ keys,
                                                       uint32_t n,
                                                       const DevParams* __restrict__ dp,
                                                       uint32_t* __restrict__ table) {
  uint32_t G = dp->p.grid_cell_count;
  if (G > dp->table_cap) G = dp->table_cap;
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < G; c += gridDim.x * kBlock) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (keys[mid] < c)
        lo = mid + 1;
      else
        hi = mid;
    }
    table[c] = lo;
  }
}

static inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

void launch_aabb(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_aabb, dim3(blocks_for(s->n)), dim3(kBlock), 0, s->ctx->stream,
                     s->pos, s->n, s->aabb_partial);
}

void launch_grid_params(clsph_state* s) {
  hipLaunchKernelGGL(k_grid_params, dim3(1), dim3(1024), 0, s->ctx->stream, s->dp,
                     s->aabb_partial, blocks_for(s->n));
}

void launch_hash(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_hash, dim3(blocks_for(s->n)), dim3(kBlock), 0, s->ctx->stream,
                     s->pos, s->dp, s->key_a, s->val_a, s->n);
}

void launch_permute(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_permute, dim3(blocks_for(s->n)), dim3(kBlock), 0, s->ctx->stream,
                     s->val_b, s->pos, s->vel, s->vh, s->pos_s, s->vel_s, s->vh_s, s->n);
}

void launch_cell_table(clsph_state* s) {
  uint32_t blocks = (uint32_t)s->ctx->cus * 8;
  hipLaunchKernelGGL(k_cell_table, dim3(blocks), dim3(kBlock), 0, s->ctx->stream,
                     s->key_b, s->n, s->dp, s->table);
}

}  // namespace clsph
