// Grid stages: bounding box, grid parameters and cell-id hashing, for every
// scene of a batch at once (the sort, the cell table and the reference's
// stable order are in sort.hip).
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
// contiguous and in order, each grouped by octant. A scene's keys are offset
// by its key_base.
#include "reduce.h"
#include "state.h"

#pragma clang fp contract(off)

namespace clsph {

__global__ __launch_bounds__(kMaxBlock) void k_aabb(const float4* __restrict__ pos,
                                                    const BlockDesc* __restrict__ blocks,
                                                    const DevParams* __restrict__ dps,
                                                    float* __restrict__ partial, uint32_t stride) {
  uint32_t scene;
  uint32_t i = particle_of(blocks, dps, scene);
  bool valid = i != 0xFFFFFFFFu;
  float4 p = valid ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  block_aabb_partial(p, valid, partial, stride);
}

// The batch's key total, key width and sort digit width for `base` keys in
// all (one thread); `err` holds the error bits raised while laying them out.
__device__ void batch_layout(BatchParams* __restrict__ bp, uint64_t base, uint32_t err,
                             uint32_t table_cap, uint32_t sort_passes) {
  if (base > table_cap) err |= kErrTableCap;
  bp->key_total = (uint32_t)base;
  uint32_t kb = bitlen((uint32_t)base);
  bp->key_bits = kb;
  // the sort orders by cell id (key >> 3) in sort_passes LSD passes; a state
  // sized for three runs two while the cell ids fit them
  uint32_t cb = kb > 3u ? kb - 3u : 1u;
  uint32_t passes = sort_passes ? sort_passes : 1u;
  if (passes > 2u && cb <= 2u * (uint32_t)kMaxDigitBits) passes = 2u;
  bp->lsd_passes = passes;
  uint32_t db = (cb + passes - 1u) / passes;
  bp->digit_bits = db < 1u ? 1u : (db > (uint32_t)kMaxDigitBits ? (uint32_t)kMaxDigitBits : db);
  if (err) atomicOr(&bp->err, err);
}

// Lays the scenes' half-cell key ranges end to end (each scene gets
// 8 (G + 1) keys, its base clamped at 0xFFFFFFF0 with kErrFineKey raised
// once the running total passes it) and sizes the sort digits for the total.
// One wave, 64 scenes at a time: each scene's base is the exclusive sum of
// the keys before it, a wave scan.
__device__ void batch_keys(DevParams* __restrict__ dps, BatchParams* __restrict__ bp) {
  constexpr uint64_t kTop = 0xFFFFFFF0ull;
  const uint32_t lane = threadIdx.x & 63u, ns = bp->n_scenes;
  uint64_t carry = 0;
  for (uint32_t s0 = 0; s0 < ns; s0 += 64u) {
    const uint32_t s = s0 + lane;
    uint64_t G = 0, keys = 0;
    if (s < ns) {
      G = (dps[s].err & kErrFineKey) ? 0u : dps[s].p.grid_cell_count;
      keys = 8ull * (G + 1);
    }
    uint64_t incl = keys;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t lo = __shfl_up((uint32_t)incl, o, 64), hi = __shfl_up((uint32_t)(incl >> 32), o, 64);
      if (lane >= o) incl += ((uint64_t)hi << 32) | lo;
    }
    if (s < ns) {
      const uint64_t ex = carry + incl - keys;
      const uint64_t base = ex < kTop ? ex : kTop;
      dps[s].key_base = (uint32_t)base;
      dps[s].key_end = (uint32_t)(base + 8ull * G);
    }
    const uint32_t tlo = __shfl((uint32_t)incl, 63, 64), thi = __shfl((uint32_t)(incl >> 32), 63, 64);
    carry += ((uint64_t)thi << 32) | tlo;
  }
  if (lane == 0)
    batch_layout(bp, carry < kTop ? carry : kTop, carry > kTop ? kErrFineKey : 0u, bp->table_cap,
                 bp->sort_passes);
}

// One work-group per scene: reduces the scene's partial boxes, then restates
// sph_simulation.cpp:201-252 in FP32. The last work-group to finish lays out
// the batch's keys (batch_keys); a single scene does so from the values in
// its registers. The parameters thread 0 needs are loaded before the
// reduction, so that their latency overlaps it.
__device__ uint32_t scene_grid(DevParams* __restrict__ dp, const float v[6], float h, uint32_t n,
                               uint32_t& G);

__global__ __launch_bounds__(1024) void k_grid_params(DevParams* __restrict__ dps,
                                                      const float* __restrict__ partial,
                                                      uint32_t stride,
                                                      BatchParams* __restrict__ bp) {
  __shared__ float red[6][16];
  DevParams* dp = dps + blockIdx.x;
  const float h = dp->p.h;
  const uint32_t np = dp->n, err0 = dp->err;
  const uint32_t table_cap = bp->table_cap, sort_passes = bp->sort_passes;
  const float inf = __builtin_huge_valf();
  float v[6] = {inf, inf, inf, -inf, -inf, -inf};
  // eight partial boxes per thread in flight at a time
  const uint32_t b0 = dp->block_begin, b1 = dp->block_end;
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += 8u * blockDim.x) {
    float q[8][6];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t bu = b + (uint32_t)u * blockDim.x;
#pragma unroll
      for (int k = 0; k < 6; ++k) q[u][k] = bu < b1 ? partial[k * stride + bu] : (k < 3 ? inf : -inf);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      for (int k = 0; k < 3; ++k) v[k] = fminf(v[k], q[u][k]);
      for (int k = 3; k < 6; ++k) v[k] = fmaxf(v[k], q[u][k]);
    }
  }
  for (int k = 0; k < 3; ++k) v[k] = wave_min(v[k]);
  for (int k = 3; k < 6; ++k) v[k] = wave_max(v[k]);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0)
    for (int k = 0; k < 6; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x >= 64u) return;  // the first wave joins the waves' boxes
  // lanes 0-15 take one wave's box each; a butterfly over 16 lanes leaves
  // the scene's box in every lane
  static_assert(sizeof(red[0]) / sizeof(float) == 16, "16 waves per work-group");
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    float r = red[k][l & 15];
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) {
      const float y = __shfl_xor(r, o, 64);
      r = (k < 3) ? fminf(r, y) : fmaxf(r, y);
    }
    v[k] = r;
  }
  if (threadIdx.x != 0 && gridDim.x == 1u) return;
  uint32_t G = 0, err = 0;
  if (threadIdx.x == 0) err = scene_grid(dp, v, h, np, G);
  if (gridDim.x == 1u) {
    // one scene: batch_keys with the scene's values in registers
    const uint64_t Gk = ((err0 | err) & kErrFineKey) ? 0u : G;
    dp->key_base = 0;
    dp->key_end = (uint32_t)(8ull * Gk);
    uint64_t base = 8ull * (Gk + 1);
    uint32_t berr = 0;
    if (base > 0xFFFFFFF0ull) {
      berr = kErrFineKey;
      base = 0xFFFFFFF0ull;
    }
    batch_layout(bp, base, berr, table_cap, sort_passes);
    return;
  }
  // the scenes' grids are in place once every work-group has passed here
  uint32_t last = 0;
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(&bp->ticket, 1u) == gridDim.x - 1u ? 1u : 0u;
  }
  if (!__shfl(last, 0, 64)) return;
  __threadfence();
  if (threadIdx.x == 0) bp->ticket = 0;
  batch_keys(dps, bp);
}

// The grid of one scene from its bounding box v = (min xyz, max xyz), its
// smoothing length h and particle count n; returns the error bits it raised
// and the cell count in G.
__device__ uint32_t scene_grid(DevParams* __restrict__ dp, const float v[6], float h, uint32_t n,
                               uint32_t& G) {
  dp->capped = 0;
  G = 0;
  if (n == 0) {
    dp->p.grid_cell_count = 0;
    return 0u;
  }
  // Initial values are numeric_limits<cl_int>::max()/min() as float; a
  // position replaces them only when strictly smaller / larger.
  const float init_min = 2147483648.0f, init_max = -2147483648.0f;
  float L = h * 2;
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
  G = morton3(gs[0], gs[1], gs[2]);
  dp->p.grid_cell_count = G;
  if (G >= (1u << 29)) err |= kErrFineKey;
  if (err) atomicOr(&dp->err, err);
  return err;
}

// Half-cell keys of every particle (hash_key).
__global__ __launch_bounds__(kMaxBlock) void k_hash(const float4* __restrict__ pos,
                                                    const BlockDesc* __restrict__ blocks,
                                                    const DevParams* __restrict__ dps,
                                                    uint32_t* __restrict__ key) {
  uint32_t scene;
  uint32_t i = particle_of(blocks, dps, scene);
  if (i == 0xFFFFFFFFu) return;
  key[i] = hash_key(pos[i], dps + scene);
}

void launch_aabb(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_aabb, dim3(s->nblocks_of(kStageAdvect)), dim3(s->bs_of(kStageAdvect)), 0,
                     s->ctx->stream, s->pos, s->table_of(kStageAdvect), s->dps, s->aabb_partial,
                     s->aabb_stride);
}

void launch_grid_params(clsph_state* s) {
  hipLaunchKernelGGL(k_grid_params, dim3(s->n_scenes), dim3(1024), 0, s->ctx->stream, s->dps,
                     s->aabb_partial, s->aabb_stride, s->bp);
}

void launch_hash(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_hash, dim3(s->nblocks_of(kStageHash)), dim3(s->bs_of(kStageHash)), 0,
                     s->ctx->stream, s->pos, s->table_of(kStageHash), s->dps, s->key_a);
}

}  // namespace clsph

// Identity tracking across a substep: the particle at sorted position k came
// from canonical index perm[k] and is written back to canonical slot[k].
__global__ __launch_bounds__(256) void k_track_ids(const uint32_t* __restrict__ perm,
                                                   const uint32_t* __restrict__ slot,
                                                   const uint32_t* __restrict__ ids,
                                                   uint32_t* __restrict__ out, uint32_t n) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  uint32_t s = slot[k], p = perm[k];
  if (s < n && p < n) out[s] = ids[p];
}

__global__ __launch_bounds__(256) void k_copy_u32(const uint32_t* __restrict__ in,
                                                  uint32_t* __restrict__ out, uint32_t n) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) out[k] = in[k];
}

namespace clsph {
void launch_track_ids(clsph_state* s) {
  if (!s->n) return;
  dim3 g((s->n + 255) / 256);
  hipLaunchKernelGGL(k_track_ids, g, dim3(256), 0, s->ctx->stream, s->val_b, s->slot, s->ids,
                     s->ids_tmp, s->n);
  hipLaunchKernelGGL(k_copy_u32, g, dim3(256), 0, s->ctx->stream, s->ids_tmp, s->ids, s->n);
}
}  // namespace clsph

namespace clsph {

// Particle records in the reference's layout (80 B: position, velocity,
// intermediate velocity, acceleration as float4 with w = 0, then density,
// pressure, grid_index and padding) in canonical order, as the host download
// assembles them: positions and velocities are canonical already; density,
// pressure, cell id and, while the fields are canonical, the acceleration
// are in sorted order and go to canonical slot[k]. Single-scene states.
__global__ __launch_bounds__(256) void k_pack_aos(
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float4* __restrict__ vh, const float4* __restrict__ acc_s,
    const float* __restrict__ rho_s, const float* __restrict__ prs_s,
    const uint32_t* __restrict__ key, const uint32_t* __restrict__ slot,
    const DevParams* __restrict__ dps, uint32_t n, int reorder, int want_acc,
    uint4* __restrict__ out) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  auto bits = [](float a, float b, float c) {
    return make_uint4(__float_as_uint(a), __float_as_uint(b), __float_as_uint(c), 0u);
  };
  const float4 p = pos[k], v = vel[k], h = vh[k];
  uint4* r = out + 5 * (size_t)k;
  r[0] = bits(p.x, p.y, p.z);
  r[1] = bits(v.x, v.y, v.z);
  r[2] = bits(h.x, h.y, h.z);
  const uint32_t o = reorder ? slot[k] : k;
  if (o >= n) return;
  uint4* q = out + 5 * (size_t)o;
  uint4 a = make_uint4(0u, 0u, 0u, 0u);
  if (want_acc && !reorder) {
    const float4 ak = acc_s[k];
    a = bits(ak.x, ak.y, ak.z);
  }
  q[3] = a;
  const uint32_t gi = reorder ? (key[k] - dps[0].key_base) >> 3 : key[k];
  q[4] = make_uint4(__float_as_uint(rho_s[k]), __float_as_uint(prs_s[k]), gi, 0u);
}

void launch_pack_aos(clsph_state* s, clsph_particle* out) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_pack_aos, dim3((s->n + 255) / 256), dim3(256), 0, s->ctx->stream, s->pos,
                     s->vel, s->vh, s->acc_s, s->rho_s, s->prs_s, s->key_b, s->slot, s->dps, s->n,
                     s->fields_canonical ? 0 : 1, s->acc_zero ? 0 : 1,
                     reinterpret_cast<uint4*>(out));
}

}  // namespace clsph
