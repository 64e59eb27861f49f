// Density/pressure and internal forces over the grid neighbourhood.
//
// Reference: density_pressure (sph.cl:9-40, forces.cl:15-43) and forces
// (sph.cl:42-62, forces.cl:45-112) with the kernels of smoothing.cl. The
// reference visits all 27 cells (side 2h) around a particle's cell and
// multiplies every candidate's term by chi = 1 - clamp(floor(r/h), 0, 1).
// Particles here are ordered by half-cell (side h), so every particle within
// h of particle i lies in the 3x3x3 block of half-cells around i's own:
// visiting that block drops exactly the terms chi zeroes, and the sums are
// the reference's up to FP32 rounding (summation order differs).
#include "state.h"

namespace clsph {

// Half-cell coordinates, with the hash's arithmetic (grid.cl:55-61):
// (uint)(2 (x - min) / (2h)).
__device__ inline void half_cell(float4 x, const DevParams* __restrict__ dp, uint32_t hc[3]) {
#pragma clang fp contract(off)
  float L = dp->p.h * 2;
  hc[0] = (uint32_t)(2.f * ((x.x - dp->p.min_point.s[0]) / L));
  hc[1] = (uint32_t)(2.f * ((x.y - dp->p.min_point.s[1]) / L));
  hc[2] = (uint32_t)(2.f * ((x.z - dp->p.min_point.s[2]) / L));
}

// Morton key of a half-cell within its scene: (cell id << 3) | octant.
__device__ inline uint32_t half_key(uint32_t hx, uint32_t hy, uint32_t hz) {
  return (morton3(hx >> 1, hy >> 1, hz >> 1) << 3) | ((hz & 1u) << 2) | ((hy & 1u) << 1) |
         (hx & 1u);
}

// Particles of the half-cells [f, f + len) (Morton-adjacent, keys of the
// batch), from the half-cell table ht; keys at or beyond F (the end of the
// scene's keys) hold none of its particles. The clamps keep a table from a
// stage run out of order from indexing past the arrays.
__device__ inline void half_range(uint32_t f, uint32_t len, const uint32_t* __restrict__ ht,
                                  uint32_t F, uint32_t n, uint32_t& s, uint32_t& e) {
  if (f >= F) {
    s = e = 0;
    return;
  }
  uint32_t g = f + len;
  g = g > F ? F : g;
  s = ht[f];
  e = ht[g];
  e = e < n ? e : n;
  s = s < e ? s : e;
}

// End of the scene's keys usable for table lookups; 0 when the table
// overflowed (an error bit is raised in that case).
__device__ inline uint32_t usable_key_end(const DevParams* __restrict__ dp,
                                          const BatchParams* __restrict__ bp) {
  return (bp->err & (kErrTableCap | kErrFineKey)) ? 0u : dp->key_end;
}

// Visits the particles of the 3x3x3 half-cells around hc in z, y, x order.
// Along x the three cells form two Morton-contiguous runs: the pair starting
// at the even coordinate, and the remaining one.
template <typename F>
__device__ inline void for_each_candidate(const uint32_t hc[3], uint32_t base,
                                          const uint32_t* __restrict__ ht, uint32_t Fc,
                                          uint32_t n, F&& fn) {
  const uint32_t x0 = hc[0] - 1u;
  const bool even = (x0 & 1u) == 0u;
  for (uint32_t dz = 0; dz < 3; ++dz)
    for (uint32_t dy = 0; dy < 3; ++dy) {
      uint32_t y = hc[1] + dy - 1u, z = hc[2] + dz - 1u;
      uint32_t a = base + half_key(x0, y, z);
      uint32_t b = base + half_key(even ? x0 + 2u : x0 + 1u, y, z);
      uint32_t s, e;
      half_range(a, even ? 2u : 1u, ht, Fc, n, s, e);
      for (uint32_t j = s; j < e; ++j) fn(j);
      half_range(b, even ? 1u : 2u, ht, Fc, n, s, e);
      for (uint32_t j = s; j < e; ++j) fn(j);
    }
}

// Half-cell coordinates and the particle's fractional position inside its
// half-cell (in units of the half-cell side h), with the hash's arithmetic.
__device__ inline void half_cell_frac(float4 x, const DevParams* __restrict__ dp, uint32_t hc[3],
                                      float fr[3]) {
#pragma clang fp contract(off)
  float L = dp->p.h * 2;
  float u[3] = {2.f * ((x.x - dp->p.min_point.s[0]) / L),
                2.f * ((x.y - dp->p.min_point.s[1]) / L),
                2.f * ((x.z - dp->p.min_point.s[2]) / L)};
  for (int a = 0; a < 3; ++a) {
    hc[a] = (uint32_t)u[a];
    fr[a] = u[a] - (float)hc[a];
  }
}

// Squared reach, in half-cell units, beyond which a half-cell cannot hold a
// particle within h: (1.01)^2. The 1 % margin exceeds the FP32 rounding of
// the scaled coordinates (a few ulp of at most 2048 half-cells, < 1e-3).
constexpr float kReach2 = 1.0201f;

// Visits the particle ranges of the half-cells around hc that a sphere of
// radius h about the particle can reach: a (dy, dz) row is skipped when its
// nearest face is out of reach, and its x-range keeps the neighbours on
// either side only when they are in reach. Within a row the half-cells
// 2k, 2k+1 share a cell and are Morton-adjacent; each maximal such piece
// of the x-range is one run (at most two per row).
template <typename R>
__device__ inline void for_each_run(const uint32_t hc[3], const float fr[3], uint32_t base,
                                    const uint32_t* __restrict__ ht, uint32_t Fc, uint32_t n,
                                    R&& run) {
  const float lo_x = fr[0] * fr[0], hi_x = (1.f - fr[0]) * (1.f - fr[0]);
  for (int dz = -1; dz <= 1; ++dz) {
    const float az = dz < 0 ? fr[2] : (dz > 0 ? 1.f - fr[2] : 0.f);
    for (int dy = -1; dy <= 1; ++dy) {
      const float ay = dy < 0 ? fr[1] : (dy > 0 ? 1.f - fr[1] : 0.f);
      const float ryz = az * az + ay * ay;
      if (ryz >= kReach2) continue;
      const uint32_t y = hc[1] + (uint32_t)dy, z = hc[2] + (uint32_t)dz;
      const uint32_t xa = lo_x + ryz < kReach2 ? hc[0] - 1u : hc[0];
      const uint32_t xb = hi_x + ryz < kReach2 ? hc[0] + 1u : hc[0];
      uint32_t x = xa;
      for (int piece = 0; piece < 2 && x <= xb; ++piece) {
        const uint32_t xe = (x | 1u) < xb ? (x | 1u) : xb;
        uint32_t s, e;
        half_range(base + half_key(x, y, z), xe - x + 1u, ht, Fc, n, s, e);
        run(s, e);
        x = xe + 1u;
      }
    }
  }
}

// Two- and four-lane unsigned vectors, held in aligned register pairs and
// quads; the neighbour list's rows are addressed as global memory.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x4 global_u32x4;

// (p0, p1, p2, p3) <- (p1, p2, p3, tail.y): the quad as two register pairs,
// each filled by one 64-bit move that takes one 32-bit half from each of two
// pairs (dst.lo = src0 half op_sel[0], dst.hi = src1 half op_sel[1]).
__device__ inline void shift_in(u32x4& reg, u32x2 tail) {
  u32x2 lo = reg.xy, hi = reg.zw;
  asm("v_pk_mov_b32 %0, %0, %1 op_sel:[1,0]" : "+v"(lo) : "v"(hi));
  asm("v_pk_mov_b32 %0, %0, %1 op_sel:[1,1]" : "+v"(hi) : "v"(tail));
  reg = (u32x4){lo.x, lo.y, hi.x, hi.y};
}

// Squared distance of a candidate, with every rounding written out so that
// both walks (and any inlining of them) compute the same bits.
__device__ inline float dist2(float4 xi, float4 c) {
#pragma clang fp contract(off)
  const float dx = xi.x - c.x, dy = xi.y - c.y, dz = xi.z - c.z;
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Density sum and neighbour list of one particle i. The sum starts with i's
// own term (r = 0: (h^2)^3) and adds the other candidates within h in
// visiting order. The list holds those neighbours (self excluded) in
// visiting order, four to a uint4, at nbr[q * n + i] (q < kMaxNeighbours / 4);
// pending entries wait in the shift register `reg` (oldest in .x, newest in
// .w), and `row` points at the list row the next four go to (nbr + i, then
// one row of n entries further per store).
struct DensityWalk {
  float acc;
  u32x4 reg = {0u, 0u, 0u, 0u};
  uint32_t cnt = 0;
  global_u32x4* row;
  uint64_t n;

  __device__ DensityWalk(uint4* __restrict__ nbr, uint32_t i, uint32_t n_, float h2)
      : acc((h2 * h2) * h2), row((global_u32x4*)(nbr + i)), n(n_) {}

  // Candidate j = tail.y at squared distance d2; in_run is false for the
  // padding entries of a group of four. Testing j != i with the distance
  // keeps the append below unconditional for the lanes that take it.
  __device__ inline void visit(bool in_run, float d2, float h2, u32x2 tail, uint32_t i) {
    static_assert((kMaxNeighbours & (kMaxNeighbours - 1u)) == 0u && kMaxNeighbours >= 4u,
                  "the list-room test below needs a power of two");
    if (in_run && d2 < h2 && tail.y != i) {
      const float t = h2 - d2;
      acc = __builtin_fmaf(t, t * t, acc);
      // the register shifts on every neighbour and is stored after each
      // fourth one while the list has room: cnt % 4 == 3 and cnt < kMaxNeighbours
      const bool full = (cnt & (3u - kMaxNeighbours)) == 3u;
      shift_in(reg, tail);
      if (full) {
        *row = reg;
        // the next row, after the store has read `row`, in place
        asm volatile("v_lshl_add_u64 %0, %1, 4, %0" : "+v"(row) : "s"(n) : "memory");
      }
      ++cnt;
    }
  }

  // Entries still pending (cnt % 4 of them, the newest of the register) go
  // to row cnt / 4 while the list has room.
  __device__ inline void finish() {
    const uint32_t r = cnt & 3u;
    if (cnt < kMaxNeighbours && r)
      *row = r == 1 ? (u32x4){reg.w, 0u, 0u, 0u}
                    : (r == 2 ? (u32x4){reg.z, reg.w, 0u, 0u} : (u32x4){reg.y, reg.z, reg.w, 0u});
  }
};

// The walk over the sorted positions in HBM: the reach-pruned half-cell runs
// of for_each_run, candidates loaded four at a time so that their latencies
// overlap (pos is padded by 4 entries past the particles, so a group needs no
// clamp).
__device__ inline void density_walk_global(DensityWalk& w, uint32_t i, float4 xi,
                                           const uint32_t hc[3], const float fr[3],
                                           const DevParams* __restrict__ dp, uint32_t F,
                                           const float4* __restrict__ pos,
                                           const uint32_t* __restrict__ table, float h2,
                                           uint32_t n) {
  for_each_run(hc, fr, dp->key_base, table, F, n, [&](uint32_t s, uint32_t e) {
    for (uint32_t j0 = s; j0 < e; j0 += 4) {
      const float4* pj = pos + j0;
      float4 c[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u] = pj[u];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t j = j0 + u;
        w.visit(j < e, dist2(xi, c[u]), h2, (u32x2){j, j}, i);
      }
    }
  });
}

// The same walk over a cell's halo staged in LDS: the 4x4x4 half-cells
// around the particle's cell (side 2h), half-cell q = x + 4y + 16z at
// cand[off[q], off[q + 1]) in sorted order, each entry (x, y, z, sorted
// index). The particle sits at local coordinates 1 + (hc & 1), so each row
// (dy, dz) of its 3x3x3 block is one contiguous range. Visiting order, and
// so every sum and list, is that of density_walk_global.
__device__ inline void density_walk_halo(DensityWalk& w, uint32_t i, float4 xi,
                                         const uint32_t hc[3], const float fr[3],
                                         const float4* cand, const uint32_t* off, float h2) {
  const uint32_t lx = 1u + (hc[0] & 1u), ly = 1u + (hc[1] & 1u), lz = 1u + (hc[2] & 1u);
  const float lo_x = fr[0] * fr[0], hi_x = (1.f - fr[0]) * (1.f - fr[0]);
  for (int dz = -1; dz <= 1; ++dz) {
    const float az = dz < 0 ? fr[2] : (dz > 0 ? 1.f - fr[2] : 0.f);
    for (int dy = -1; dy <= 1; ++dy) {
      const float ay = dy < 0 ? fr[1] : (dy > 0 ? 1.f - fr[1] : 0.f);
      const float ryz = az * az + ay * ay;
      if (ryz >= kReach2) continue;
      const uint32_t row = 4u * (ly + (uint32_t)dy) + 16u * (lz + (uint32_t)dz);
      const uint32_t xa = lo_x + ryz < kReach2 ? lx - 1u : lx;
      const uint32_t xb = hi_x + ryz < kReach2 ? lx + 1u : lx;
      // byte offsets into cand: the group at a holds entry u inside the
      // range when a < end[u] = 16 (e - u)
      const int s = (int)off[row + xa] * 16, e = (int)off[row + xb + 1u] * 16;
      const int end[4] = {e, e - 16, e - 32, e - 48};
      for (int a = s; a < e; a += 64) {
        const float4* g = (const float4*)((const char*)cand + a);
        float4 c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = g[u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          // (z, index): the candidate's last register pair
          w.visit(u == 0 || a < end[u], dist2(xi, c[u]), h2,
                  (u32x2){__float_as_uint(c[u].z), __float_as_uint(c[u].w)}, i);
        }
      }
    }
  }
}

// Inverse of spread10: the 10 bits at positions 3k.
__device__ inline uint32_t compact10(uint32_t v) {
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030C30C3u;
  v = (v | (v >> 4)) & 0x0300F00Fu;
  v = (v | (v >> 8)) & 0x030000FFu;
  v = (v | (v >> 16)) & 0x000003FFu;
  return v;
}

// Halo staging of the density kernel, per work-group of WAVES waves: up to
// kHalos cells (their 4x4x4 half-cell halos) and kCap candidates (x, y, z,
// index: 16 B each). A bulk cell holds about 64 particles, so the
// work-group's particles span about WAVES + 1 cells of about 512 halo
// candidates each. Up to 4 waves, 768 candidates per wave (three 4-wave
// work-groups per CU); at 8 waves, 12 cells and 4860 candidates keep a
// work-group within 80 KB, so that two fit per CU.
template <int WAVES>
struct DensityHalos {
  static constexpr uint32_t kHalos = WAVES < 8 ? 2u * WAVES : 12u;
  static constexpr uint32_t kCap = WAVES < 8 ? 768u * WAVES - 4u : 4860u;
  float4 cand[kCap + 4];               // + the padding of a group of four
  uint32_t off[kHalos * 64u + 1u];     // halo h, half-cell q: off[64 h + q]
  uint32_t cell[kHalos];               // first key of each cell (a multiple of 8)
  uint32_t wave_sum[WAVES];
  uint32_t wave_cells[WAVES];
};

// Density and pressure (sph.cl:9-40, forces.cl:15-43) of the particles
// [first, first + 64 WAVES) of one scene, with each particle's neighbour
// list (self excluded) in visiting order into nbr[q * n + i], its count into
// nbr_count[i], the force inputs fab[2i] = (x, p / rho^2), fab[2i + 1] =
// (v, m / rho), and with `nu` the learned surrogate's node inputs. The
// work-group first stages, for each distinct cell (side 2h) its particles
// fall in, the cell's 4x4x4 half-cell halo in LDS (coalesced runs
// of the sorted positions); particles then walk their 3x3x3 half-cells
// there. Cells beyond the staging capacity, and every particle when the
// table is unusable, walk the sorted positions in HBM instead, with the same
// visiting order.
template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_density(const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel,
                                                        const uint32_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ table,
                                                        const BlockDesc* __restrict__ blocks,
                                                        const DevParams* __restrict__ dps,
                                                        const BatchParams* __restrict__ bp,
                                                        float* __restrict__ rho,
                                                        float* __restrict__ prs,
                                                        float4* __restrict__ fab,
                                                        uint4* __restrict__ nbr,
                                                        uint32_t* __restrict__ nbr_count,
                                                        float2* __restrict__ nu, uint32_t n) {
  using Halos = DensityHalos<WAVES>;
  static_assert(Halos::kHalos <= 2u * WAVES, "at most two (halo, half-cell) entries per thread");
  __shared__ Halos sh;
  // runs of 16 work-groups per XCD: neighbouring halos share an L2
  const BlockDesc b = blocks[xcd_block<16>(blockIdx.x, gridDim.x)];
  const DevParams* dp = dps + b.scene;
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  const uint32_t se = dp->begin + dp->n;
  const uint32_t i = b.first + t;
  const bool live = i < se;
  const float h = dp->p.h, h2 = h * h;
  const uint32_t F = usable_key_end(dp, bp);

  // Cells of the work-group's particles, in order: a particle starts one
  // when it is the first particle or its cell differs from its predecessor's.
  const uint32_t key = live ? keys[i] : 0u;
  const uint32_t prev = (live && t > 0) ? keys[i - 1] : 0u;
  const bool starts = live && (t == 0 || (key >> 3) != (prev >> 3));
  const uint64_t mask = __ballot(starts);
  if (lane == 0) sh.wave_cells[wv] = (uint32_t)__popcll(mask);
  __syncthreads();
  uint32_t before = 0, cells = 0;
  for (int q = 0; q < WAVES; ++q) {
    const uint32_t c = sh.wave_cells[q];
    before += (uint32_t)q < wv ? c : 0u;
    cells += c;
  }
  // index of the particle's cell among the work-group's cells
  const uint32_t ci = before + (uint32_t)__popcll(mask & (~0ull >> (63u - lane))) - 1u;
  if (starts && ci < Halos::kHalos) sh.cell[ci] = key & ~7u;
  const uint32_t halos = F == 0u ? 0u : (cells < Halos::kHalos ? cells : Halos::kHalos);
  __syncthreads();

  // Halo half-cell runs: thread t takes the entries 2t and 2t + 1 of the
  // (halo, half-cell) list.
  uint32_t rs[2] = {0u, 0u}, rc[2] = {0u, 0u};
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const uint32_t q = 2u * t + (uint32_t)u, hh = q >> 6, l = q & 63u;
    if (hh < halos) {
      const uint32_t c8 = sh.cell[hh];
      const uint32_t cl = (c8 - dp->key_base) >> 3;
      const uint32_t hx = 2u * compact10(cl) - 1u + (l & 3u);
      const uint32_t hy = 2u * compact10(cl >> 1) - 1u + ((l >> 2) & 3u);
      const uint32_t hz = 2u * compact10(cl >> 2) - 1u + (l >> 4);
      if (hx < 2048u && hy < 2048u && hz < 2048u) {
        const uint32_t f = dp->key_base + half_key(hx, hy, hz);
        if (f < F) {
          uint32_t e = table[f + 1u];
          e = e < n ? e : n;
          uint32_t s = table[f];
          s = s < e ? s : e;
          rs[u] = s;
          rc[u] = e - s;
        }
      }
    }
  }
  // exclusive scan of the run lengths over the work-group
  const uint32_t mine = rc[0] + rc[1];
  uint32_t incl = mine;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) sh.wave_sum[wv] = incl;
  __syncthreads();
  uint32_t base = 0;
  for (int q = 0; q < WAVES; ++q) base += (uint32_t)q < wv ? sh.wave_sum[q] : 0u;
  const uint32_t ex = base + incl - mine;
  if (2u * t < Halos::kHalos * 64u) {  // threads past the last halo hold no runs
    sh.off[2u * t] = ex;
    sh.off[2u * t + 1u] = ex + rc[0];
  }
  if (t == WAVES * 64 - 1) sh.off[Halos::kHalos * 64u] = ex + mine;
  __syncthreads();
  // halos staged: the leading ones whose candidates fit
  uint32_t staged = 0;
  while (staged < halos && sh.off[64u * (staged + 1u)] <= Halos::kCap) ++staged;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const uint32_t q = 2u * t + (uint32_t)u;
    if ((q >> 6) < staged) {
      const uint32_t o = u ? ex + rc[0] : ex;
      for (uint32_t k = 0; k < rc[u]; ++k) {
        const uint32_t j = rs[u] + k;
        const float4 p = pos[j];
        sh.cand[o + k] = make_float4(p.x, p.y, p.z, __uint_as_float(j));
      }
    }
  }
  __syncthreads();

  if (!live) return;
  const float4 xi = pos[i];
  uint32_t hc[3];
  float fr[3];
  half_cell_frac(xi, dp, hc, fr);
  DensityWalk w(nbr, i, n, h2);
  if (ci < staged)
    density_walk_halo(w, i, xi, hc, fr, sh.cand, sh.off + 64u * ci, h2);
  else
    density_walk_global(w, i, xi, hc, fr, dp, F, pos, table, h2, n);
  w.finish();
  nbr_count[i] = w.cnt;
  // density = sum m * poly6 * (h^2 - r^2)^3 ; Tait pressure (sph.cl:36-39)
  float m = dp->p.particle_mass;
  float d = m * (dp->t.poly_6 * w.acc);
  float q = d / dp->p.fluid_density;
  float q2 = q * q, q4 = q2 * q2;
  float p = dp->p.K * (q4 * q2 * q - 1.f);
  rho[i] = d;
  prs[i] = p;
  float4 vi = vel[i];
  fab[2u * i] = make_float4(xi.x, xi.y, xi.z, p / (d * d));
  fab[2u * i + 1u] = make_float4(vi.x, vi.y, vi.z, m / d);
  if (nu) nu[i] = node_in(d, p, dp->p.fluid_density, dp->p.K);
}

// Pair terms of compute_internal_forces_with_grid (forces.cl:60-101) for one
// neighbour j != i at squared distance d2 < h^2; a = (x, p/rho^2) and
// b = (v, m/rho) of each particle.
struct ForceSums {
  float3 pt, vt, nrm;
  float lap;
};

struct ForceConsts {
  float h, h2, m, c_spiky, c_visc, c_grad, c_lap, spiky0;
};

__device__ inline void force_pair(ForceSums& f, const ForceConsts& k, float4 ai, float4 bi,
                                  float4 aj, float4 bj) {
  float dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
  float d2 = dx * dx + dy * dy + dz * dz;
  float r = __builtin_amdgcn_sqrtf(d2);
  float t = k.h2 - d2;
  float hr = k.h - r;
  // pressure (forces.cl:66-73): (p_j/rho_j^2 + p_i/rho_i^2) m grad W_spiky
  float ps = (aj.w + ai.w) * k.m;
  const bool coincident = r < 1e-7f;
  float g = coincident ? ps * k.spiky0 : ps * k.c_spiky * hr * hr * __builtin_amdgcn_rcpf(r);
  f.pt.x += g * (coincident ? 1.f : dx);
  f.pt.y += g * (coincident ? 1.f : dy);
  f.pt.z += g * (coincident ? 1.f : dz);
  // viscosity (forces.cl:75-80): (v_j - v_i) (m/rho_j) lap W_visc
  float vw = bj.w * k.c_visc * hr;
  f.vt.x += (bj.x - bi.x) * vw;
  f.vt.y += (bj.y - bi.y) * vw;
  f.vt.z += (bj.z - bi.z) * vw;
  // colour field gradient and laplacian (forces.cl:83-91)
  float gw = bj.w * k.c_grad * t * t;
  f.nrm.x += gw * dx;
  f.nrm.y += gw * dy;
  f.nrm.z += gw * dz;
  f.lap += bj.w * k.c_lap * t * (3.f * k.h2 - 7.f * d2);
}

// OVERFLOW_ONLY: only the particles whose neighbours overflow the list
// (more than kMaxNeighbours), for callers that sum the listed ones
// themselves.
template <bool OVERFLOW_ONLY>
__global__ __launch_bounds__(kMaxBlock) void k_forces(const float4* __restrict__ fab,
                                                      const float* __restrict__ rho,
                                                      const uint32_t* __restrict__ table,
                                                      const uint4* __restrict__ nbr,
                                                      const uint32_t* __restrict__ nbr_count,
                                                      const BlockDesc* __restrict__ blocks,
                                                      const DevParams* __restrict__ dps,
                                                      const BatchParams* __restrict__ bp,
                                                      float4* __restrict__ acc_out, uint32_t n) {
  uint32_t scene;
  // runs of 4 work-groups per XCD: neighbour records shared in an L2
  uint32_t i = particle_of<4>(blocks, dps, scene);
  if (i == 0xFFFFFFFFu) return;
  if (OVERFLOW_ONLY && nbr_count[i] <= kMaxNeighbours) return;
  const DevParams* dp = dps + scene;
  ForceConsts k;
  k.h = dp->p.h;
  k.h2 = k.h * k.h;
  k.m = dp->p.particle_mass;
  k.c_spiky = dp->t.spiky;
  k.c_visc = dp->t.viscosity;
  k.c_grad = dp->t.poly_6_gradient;
  k.c_lap = dp->t.poly_6_laplacian;
  // Coincident-particle branch of spiky_gradient (smoothing.cl:23-25).
  const float h3 = k.h2 * k.h;
  k.spiky0 = -45.f / (3.14159274f * (h3 * h3));
  const uint32_t F = usable_key_end(dp, bp);
  float4 ai = fab[2u * i], bi = fab[2u * i + 1u];
  float rho_i = rho[i];
  ForceSums fs;
  fs.pt = make_float3(0.f, 0.f, 0.f);
  fs.vt = fs.pt;
  fs.nrm = fs.pt;
  // Self term of the colour-field laplacian: (m/rho_i) poly6_lap h^2 (3 h^2).
  fs.lap = bi.w * k.c_lap * k.h2 * (3.f * k.h2);
  uint32_t cnt = nbr_count[i];
  if (cnt <= kMaxNeighbours) {
    // groups of four neighbours, software-pipelined: while a group's pair
    // terms are summed, the next group's records and the indices of the one
    // after it are in flight
    const uint32_t groups = cnt >> 2;
    auto gather = [&](uint4 js, float4 ra[4], float4 rb[4]) {
      const uint32_t jj[4] = {js.x, js.y, js.z, js.w};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ra[u] = fab[2u * jj[u]];
        rb[u] = fab[2u * jj[u] + 1u];
      }
    };
    float4 ca[4], cb[4];
    uint4 jn = make_uint4(0u, 0u, 0u, 0u);
    if (groups) {
      gather(nbr[i], ca, cb);
      if (groups > 1) jn = nbr[(size_t)n + i];
    }
    for (uint32_t gq = 0; gq < groups; ++gq) {
      float4 na[4], nb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) na[u] = ca[u], nb[u] = cb[u];
      if (gq + 1 < groups) gather(jn, na, nb);
      if (gq + 2 < groups) jn = nbr[(size_t)(gq + 2) * n + i];
#pragma unroll
      for (int u = 0; u < 4; ++u) force_pair(fs, k, ai, bi, ca[u], cb[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u) ca[u] = na[u], cb[u] = nb[u];
    }
    const uint32_t q = groups << 2;
    if (q < cnt) {
      uint4 js = nbr[(size_t)(q >> 2) * n + i];
      uint32_t r = cnt - q;
      force_pair(fs, k, ai, bi, fab[2u * js.x], fab[2u * js.x + 1u]);
      if (r > 1) force_pair(fs, k, ai, bi, fab[2u * js.y], fab[2u * js.y + 1u]);
      if (r > 2) force_pair(fs, k, ai, bi, fab[2u * js.z], fab[2u * js.z + 1u]);
    }
  } else {
    // More neighbours than the list holds: walk the half-cells again.
    uint32_t hc[3];
    half_cell(ai, dp, hc);
    const float h2 = k.h2;
    for_each_candidate(hc, dp->key_base, table, F, n, [&](uint32_t j) {
      float4 aj = fab[2u * j];
      float dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
      float d2 = dx * dx + dy * dy + dz * dz;
      if (!(d2 < h2) || j == i) return;
      force_pair(fs, k, ai, bi, aj, fab[2u * j + 1u]);
    });
  }
  float3 pt = fs.pt, vt = fs.vt, nrm = fs.nrm;
  float lap = fs.lap;
  // forces.cl:95-111 and sph.cl:53-58
  float mu = dp->p.dynamic_viscosity;
  float3 f = make_float3(-rho_i * pt.x + vt.x * mu, -rho_i * pt.y + vt.y * mu,
                         -rho_i * pt.z + vt.z * mu);
  float nl = sqrtf(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
  if (nl > dp->p.surface_tension_threshold) {
    float kk = -dp->p.surface_tension * lap / nl;
    f.x += kk * nrm.x;
    f.y += kk * nrm.y;
    f.z += kk * nrm.z;
  }
  const clsph_float3& g = dp->p.constant_acceleration;
  acc_out[i] = make_float4(f.x / rho_i + g.s[0], f.y / rho_i + g.s[1], f.z / rho_i + g.s[2], 0.f);
}

void launch_density(clsph_state* s, float2* nu) {
  if (s->n == 0) return;
  const int bs = s->bs_of(kStageDensity);
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(s->nblocks_of(kStageDensity)), dim3(bs), 0, s->ctx->stream,
                       s->pos_s, s->vel_s, s->key_b, s->table, s->table_of(kStageDensity), s->dps,
                       s->bp, s->rho_s, s->prs_s, s->fab, s->nbr, s->nbr_count, nu, s->n);
  };
  switch (bs) {
    case 64: go(k_density<1>); break;
    case 128: go(k_density<2>); break;
    case 256: go(k_density<4>); break;
    default: go(k_density<8>); break;
  }
}

void launch_forces(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL((k_forces<false>), dim3(s->nblocks_of(kStageForces)),
                     dim3(s->bs_of(kStageForces)), 0, s->ctx->stream, s->fab, s->rho_s,
                     s->table, s->nbr, s->nbr_count, s->table_of(kStageForces), s->dps, s->bp,
                     s->acc_s, s->n);
}

void launch_forces_overflow(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL((k_forces<true>), dim3(s->nblocks_of(kStageForces)),
                     dim3(s->bs_of(kStageForces)), 0, s->ctx->stream, s->fab, s->rho_s,
                     s->table, s->nbr, s->nbr_count, s->table_of(kStageForces), s->dps, s->bp,
                     s->acc_s, s->n);
}

}  // namespace clsph
