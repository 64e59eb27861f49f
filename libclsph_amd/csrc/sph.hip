// Density/pressure and internal forces over the grid neighbourhood.
//
// Reference: density_pressure (sph.cl:9-40, forces.cl:15-43) and forces
// (sph.cl:42-62, forces.cl:45-112) with the kernels of smoothing.cl. The
// reference visits all 27 cells around a particle's cell (side 2h) and
// multiplies every candidate's term by chi = 1 - clamp(floor(r/h), 0, 1).
// A radius-h sphere inside a cell of side 2h overlaps at most 2 cells per
// axis: the half of the cell the particle sits in picks them. Visiting only
// those 8 cells (in the reference's z, y, x order) drops exactly the terms
// that chi zeroes, so the sum is the reference's up to FP32 rounding.
#include "state.h"

namespace clsph {

constexpr int kBlock = 256;

struct CellWalk {
  uint32_t c0[3];  // lower corner of the 2x2x2 block of cells
};

// Cell coordinates with the hash's arithmetic (grid.cl:55-61), then the
// 2x2x2 block that contains the particle's h-sphere.
__device__ inline CellWalk octant_block(float4 x, const DevParams* __restrict__ dp) {
#pragma clang fp contract(off)
  CellWalk w;
  float L = dp->p.h * 2;
  float u[3] = {(x.x - dp->p.min_point.s[0]) / L, (x.y - dp->p.min_point.s[1]) / L,
                (x.z - dp->p.min_point.s[2]) / L};
  for (int a = 0; a < 3; ++a) {
    uint32_t c = (uint32_t)u[a];
    w.c0[a] = (u[a] - (float)c >= 0.5f) ? c : c - 1u;
  }
  return w;
}

// get_start_end_indices_for_cell (grid.cl:22-32). Cells at or beyond G hold
// no particles (every particle's key is below G); the test also keeps a
// corrupt position (NaN) from indexing past the table.
__device__ inline void cell_range(uint32_t cell, const uint32_t* __restrict__ table,
                                  uint32_t G, uint32_t n, uint32_t& s, uint32_t& e) {
  if (cell >= G) {
    s = e = 0;
    return;
  }
  s = table[cell];
  e = (G > cell + 1) ? table[cell + 1] : n;
  // A table from a stage run out of order must not index past the arrays.
  e = e < n ? e : n;
  s = s < e ? s : e;
}

// Grid cell count usable for table lookups; 0 when the table overflowed
// (kErrTableCap is raised by the grid-parameter kernel in that case).
__device__ inline uint32_t usable_cells(const DevParams* __restrict__ dp) {
  uint32_t G = dp->p.grid_cell_count;
  return G > dp->table_cap ? 0u : G;
}

// Besides density and pressure, records the particle's neighbours (r < h,
// self excluded) in visiting order into nbr[k * n + i], k < kMaxNeighbours,
// with the full count in nbr_count[i]; forces walks this list.
__global__ __launch_bounds__(kBlock) void k_density(const float4* __restrict__ pos,
                                                    const uint32_t* __restrict__ table,
                                                    const DevParams* __restrict__ dp,
                                                    float* __restrict__ rho,
                                                    float* __restrict__ prs,
                                                    float2* __restrict__ qm,
                                                    uint32_t* __restrict__ nbr,
                                                    uint32_t* __restrict__ nbr_count,
                                                    uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float h = dp->p.h, h2 = h * h;
  const uint32_t G = usable_cells(dp);
  float4 xi = pos[i];
  CellWalk w = octant_block(xi, dp);
  float acc = 0.f;
  uint32_t cnt = 0;
  for (uint32_t dz = 0; dz < 2; ++dz)
    for (uint32_t dy = 0; dy < 2; ++dy)
      for (uint32_t dx0 = 0; dx0 < 2; ++dx0) {
        uint32_t s, e;
        cell_range(morton3(w.c0[0] + dx0, w.c0[1] + dy, w.c0[2] + dz), table, G, n, s, e);
        for (uint32_t j = s; j < e; ++j) {
          float4 xj = pos[j];
          float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
          float d2 = dx * dx + dy * dy + dz * dz;
          if (d2 < h2) {
            float t = h2 - d2;
            acc += t * t * t;
            if (j != i) {
              if (cnt < kMaxNeighbours) nbr[(size_t)cnt * n + i] = j;
              ++cnt;
            }
          }
        }
      }
  nbr_count[i] = cnt;
  // density = sum m * poly6 * (h^2 - r^2)^3 ; Tait pressure (sph.cl:36-39)
  float m = dp->p.particle_mass;
  float d = m * (dp->t.poly_6 * acc);
  float q = d / dp->p.fluid_density;
  float q2 = q * q, q4 = q2 * q2;
  float p = dp->p.K * (q4 * q2 * q - 1.f);
  rho[i] = d;
  prs[i] = p;
  qm[i] = make_float2(p / (d * d), m / d);
}

// Pair terms of compute_internal_forces_with_grid (forces.cl:60-101) for one
// neighbour j != i at squared distance d2 < h^2.
struct ForceSums {
  float3 pt, vt, nrm;
  float lap;
};

__device__ inline void force_pair(ForceSums& f, float dx, float dy, float dz, float d2,
                                  float4 vi, float4 vj, float2 qmi, float2 qmj, float h, float h2,
                                  float m, float c_spiky, float c_visc, float c_grad,
                                  float c_lap, float spiky0) {
  float r = __builtin_amdgcn_sqrtf(d2);
  float t = h2 - d2;
  float hr = h - r;
  // pressure (forces.cl:66-73): (p_j/rho_j^2 + p_i/rho_i^2) m grad W_spiky
  float ps = (qmj.x + qmi.x) * m;
  if (r < 1e-7f) {
    float g = ps * spiky0;
    f.pt.x += g;
    f.pt.y += g;
    f.pt.z += g;
  } else {
    float g = ps * c_spiky * hr * hr * __builtin_amdgcn_rcpf(r);
    f.pt.x += g * dx;
    f.pt.y += g * dy;
    f.pt.z += g * dz;
  }
  // viscosity (forces.cl:75-80): (v_j - v_i) (m/rho_j) lap W_visc
  float vw = qmj.y * c_visc * hr;
  f.vt.x += (vj.x - vi.x) * vw;
  f.vt.y += (vj.y - vi.y) * vw;
  f.vt.z += (vj.z - vi.z) * vw;
  // colour field gradient and laplacian (forces.cl:83-91)
  float gw = qmj.y * c_grad * t * t;
  f.nrm.x += gw * dx;
  f.nrm.y += gw * dy;
  f.nrm.z += gw * dz;
  f.lap += qmj.y * c_lap * t * (3.f * h2 - 7.f * d2);
}

__global__ __launch_bounds__(kBlock) void k_forces(const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel,
                                                   const float* __restrict__ rho,
                                                   const float2* __restrict__ qm,
                                                   const uint32_t* __restrict__ table,
                                                   const uint32_t* __restrict__ nbr,
                                                   const uint32_t* __restrict__ nbr_count,
                                                   const DevParams* __restrict__ dp,
                                                   float4* __restrict__ acc_out, uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float h = dp->p.h, h2 = h * h, m = dp->p.particle_mass;
  const float c_spiky = dp->t.spiky, c_visc = dp->t.viscosity;
  const float c_grad = dp->t.poly_6_gradient, c_lap = dp->t.poly_6_laplacian;
  // Coincident-particle branch of spiky_gradient (smoothing.cl:23-25).
  const float h3 = h2 * h;
  const float spiky0 = -45.f / (3.14159274f * (h3 * h3));
  const uint32_t G = usable_cells(dp);
  float4 xi = pos[i], vi = vel[i];
  float2 qmi = qm[i];
  float rho_i = rho[i];
  ForceSums fs;
  fs.pt = make_float3(0.f, 0.f, 0.f);
  fs.vt = fs.pt;
  fs.nrm = fs.pt;
  // Self term of the colour-field laplacian: (m/rho_i) poly6_lap h^2 (3 h^2).
  fs.lap = qmi.y * c_lap * h2 * (3.f * h2);
  uint32_t cnt = nbr_count[i];
  if (cnt <= kMaxNeighbours) {
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = nbr[(size_t)k * n + i];
      float4 xj = pos[j], vj = vel[j];
      float2 qmj = qm[j];
      float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      float d2 = dx * dx + dy * dy + dz * dz;
      force_pair(fs, dx, dy, dz, d2, vi, vj, qmi, qmj, h, h2, m, c_spiky, c_visc, c_grad, c_lap,
                 spiky0);
    }
  } else {
    // More neighbours than the list holds: walk the cells again.
    CellWalk w = octant_block(xi, dp);
    for (uint32_t dz = 0; dz < 2; ++dz)
      for (uint32_t dy = 0; dy < 2; ++dy)
        for (uint32_t dx0 = 0; dx0 < 2; ++dx0) {
          uint32_t s, e;
          cell_range(morton3(w.c0[0] + dx0, w.c0[1] + dy, w.c0[2] + dz), table, G, n, s, e);
          for (uint32_t j = s; j < e; ++j) {
            float4 xj = pos[j];
            float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            float d2 = dx * dx + dy * dy + dz * dz;
            if (!(d2 < h2) || j == i) continue;
            force_pair(fs, dx, dy, dz, d2, vi, vel[j], qmi, qm[j], h, h2, m, c_spiky, c_visc,
                       c_grad, c_lap, spiky0);
          }
        }
  }
  float3 pt = fs.pt, vt = fs.vt, nrm = fs.nrm;
  float lap = fs.lap;
  // forces.cl:95-111 and sph.cl:53-58
  float mu = dp->p.dynamic_viscosity;
  float3 f = make_float3(-rho_i * pt.x + vt.x * mu, -rho_i * pt.y + vt.y * mu,
                         -rho_i * pt.z + vt.z * mu);
  float nl = sqrtf(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
  if (nl > dp->p.surface_tension_threshold) {
    float k = -dp->p.surface_tension * lap / nl;
    f.x += k * nrm.x;
    f.y += k * nrm.y;
    f.z += k * nrm.z;
  }
  const clsph_float3& g = dp->p.constant_acceleration;
  acc_out[i] = make_float4(f.x / rho_i + g.s[0], f.y / rho_i + g.s[1], f.z / rho_i + g.s[2], 0.f);
}

static inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

void launch_density(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_density, dim3(blocks_for(s->n)), dim3(kBlock), 0, s->ctx->stream,
                     s->pos_s, s->table, s->dp, s->rho_s, s->prs_s, s->qm_s, s->nbr, s->nbr_count,
                     s->n);
}

void launch_forces(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_forces, dim3(blocks_for(s->n)), dim3(kBlock), 0, s->ctx->stream,
                     s->pos_s, s->vel_s, s->rho_s, s->qm_s, s->table, s->nbr, s->nbr_count, s->dp,
                     s->acc_s, s->n);
}

}  // namespace clsph
