// Grid stages: bounding box, grid parameters, cell-id hashing, state
// permutation, cell tables and the reference's stable order.
//
// The bounding box / grid arithmetic must reproduce the reference's host FP32
// sequence bit for bit (sph_simulation.cpp:201-252), and the cell id the
// kernel's (grid.cl:43-67): no contraction, correctly rounded division.
//
// Particles are hashed to the half-size cell (side h) they fall in. With
// u = (x - min) / (2h) per axis, the cell coordinate of the reference is
// (uint)u and the half-cell coordinate is (uint)(2u) = 2 (uint)u + [frac(u) >=
// 1/2], exact because 2u is exact. The Morton key of the half-cell is then
// (reference key << 3) | octant, so sorting by it keeps the reference's cells
// contiguous and in order, each grouped by octant.
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
  uint32_t kb = bitlen(G) + 3;
  if (kb > 32) err |= kErrFineKey;
  dp->key_bits = kb;
  uint32_t db = (kb + kSortPasses - 1) / kSortPasses;
  dp->digit_bits = db > (uint32_t)kMaxDigitBits ? (uint32_t)kMaxDigitBits : db;
  dp->capped = 0;
  if (err) atomicOr(&dp->err, err);
}

// locate_in_grid (grid.cl:43-67): the reference's cell id is
// Morton((uint)((x - min) / (h * 2))); the half-cell key adds the octant.
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
  float u[3] = {(p.x - dp->p.min_point.s[0]) / L, (p.y - dp->p.min_point.s[1]) / L,
                (p.z - dp->p.min_point.s[2]) / L};
  uint32_t c[3], oct = 0;
  for (int a = 0; a < 3; ++a) {
    c[a] = (uint32_t)u[a];
    oct |= (((uint32_t)(2.f * u[a])) & 1u) << a;
  }
  key[i] = (morton3(c[0], c[1], c[2]) << 3) | oct;
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

// First index in [lo, hi) whose key is >= k (keys sorted ascending).
__device__ inline uint32_t lower_bound(const uint32_t* __restrict__ keys, uint32_t lo,
                                       uint32_t hi, uint32_t k) {
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Half-cell table: ht[f] = #{keys < f} for f in [0, 8G], one thread per cell
// of side 2h. The reference's cell table (sph_simulation.cpp:162-170),
// T[c] = #{cell ids < c}, is ht[8c].
__global__ __launch_bounds__(kBlock) void k_cell_table(const uint32_t* __restrict__ keys,
                                                       uint32_t n,
                                                       const DevParams* __restrict__ dp,
                                                       uint32_t* __restrict__ ht) {
  uint32_t G = dp->p.grid_cell_count;
  if (G > dp->table_cap || (dp->err & kErrFineKey)) return;
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < G; c += gridDim.x * kBlock) {
    uint32_t start = lower_bound(keys, 0, n, c << 3);
    // end of the cell: gallop from start, then bisect
    uint32_t nxt = (c + 1) << 3;
    uint32_t lo = start, step = 64, hi = start;
    while (hi < n && keys[hi] < nxt) {
      lo = hi + 1;
      hi = (n - hi > step) ? hi + step : n;
      step <<= 1;
    }
    uint32_t end = lower_bound(keys, lo, hi, nxt);
    ht[8u * c] = start;
    uint32_t b = start;
    for (uint32_t o = 1; o < 8; ++o) {
      b = lower_bound(keys, b, end, (c << 3) | o);
      ht[8u * c + o] = b;
    }
    if (c == G - 1) ht[8u * G] = n;
  }
}

// Slot of each sorted particle in the reference's order (stable by cell id).
// Within a cell the half-cell order is octant-major with the canonical
// (input) order kept inside each octant, so the particle's rank in its cell
// is its index in its octant plus the members of the other octants that came
// before it in the input.
__global__ __launch_bounds__(kBlock) void k_canonical_slot(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ perm,
                                                           const uint32_t* __restrict__ ht,
                                                           const DevParams* __restrict__ dp,
                                                           uint32_t* __restrict__ slot,
                                                           uint32_t n) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  uint32_t G = dp->p.grid_cell_count;
  if (G > dp->table_cap || (dp->err & kErrFineKey)) {
    slot[k] = k;
    return;
  }
  uint32_t f = keys[k], c = f >> 3, o = f & 7u;
  uint32_t me = perm[k];
  uint32_t rank = k - ht[f];
  for (uint32_t q = 0; q < 8; ++q) {
    if (q == o) continue;
    uint32_t lo = ht[8u * c + q], hi = ht[8u * c + q + 1];
    uint32_t a = lo, b = hi;
    while (a < b) {
      uint32_t mid = (a + b) >> 1;
      if (perm[mid] < me)
        a = mid + 1;
      else
        b = mid;
    }
    rank += a - lo;
  }
  slot[k] = ht[8u * c] + rank;
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
  if (s->n == 0) return;
  uint32_t blocks = (uint32_t)s->ctx->cus * 8;
  hipLaunchKernelGGL(k_cell_table, dim3(blocks), dim3(kBlock), 0, s->ctx->stream,
                     s->key_b, s->n, s->dp, s->table);
  hipLaunchKernelGGL(k_canonical_slot, dim3(blocks_for(s->n)), dim3(kBlock), 0,
                     s->ctx->stream, s->key_b, s->val_b, s->table, s->dp, s->slot, s->n);
}

}  // namespace clsph
