/*
 * CPU oracle for the libclsph per-step SPH pipeline. TEST INFRASTRUCTURE ONLY:
 * see cpu_ref.h. Built with -ffp-contract=off so that no multiply-add is
 * fused, matching OpenCL C float semantics of the reference kernels.
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* get_grid_index_z_curve: libclsph/common/util.h:41-62 (10 bits per axis,
 * x at bit 3k, y at 3k+1, z at 3k+2). */
static uint32_t spread10(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
uint32_t orc_morton(uint32_t x, uint32_t y, uint32_t z) {
  return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
}

/* uninterleave + get_cell_coords_z_curve: libclsph/common/util.h:4-38.
 * Gathers bit 3k of the masked word into bit k for k < 10. */
static uint32_t gather10(uint32_t v) {
  uint32_t r = 0;
  for (int k = 0; k < 10; ++k) r |= ((v >> (3 * k)) & 1u) << k;
  return r;
}
void orc_morton_decode(uint32_t key, uint32_t out[3]) {
  const uint32_t mask = 0x9249249u;
  out[0] = gather10(key & mask);
  out[1] = gather10((key >> 1) & mask);
  out[2] = gather10((key >> 2) & mask);
}

/* sph_simulation.cpp:201-252 */
int orc_grid_params(const particle* in, uint32_t n,
                    simulation_parameters* params) {
  float L = params->h * 2;
  float mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = (float)2147483647;  /* numeric_limits<cl_int>::max() as float */
    mx[a] = (float)(-2147483647 - 1);
  }
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      float p = in[i].position.s[a];
      if (p < mn[a]) mn[a] = p;
      if (p > mx[a]) mx[a] = p;
    }
  int32_t gs[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] -= L * 2;
    mx[a] += L * 2;
    params->min_point.s[a] = mn[a];
    params->max_point.s[a] = mx[a];
    params->min_point.s[3] = 0.f;
    params->max_point.s[3] = 0.f;
    gs[a] = (int32_t)(uint32_t)((mx[a] - mn[a]) / L);
  }
  params->grid_size_x = gs[0];
  params->grid_size_y = gs[1];
  params->grid_size_z = gs[2];
  for (int a = 0; a < 3; ++a)
    if ((uint32_t)gs[a] >= 1024u) return -1;
  params->grid_cell_count =
      orc_morton((uint32_t)gs[0], (uint32_t)gs[1], (uint32_t)gs[2]);
  return 0;
}

/* locate_in_grid: grid.cl:43-67 */
void orc_locate_in_grid(const particle* in, particle* out, uint32_t n,
                        const simulation_parameters* params) {
  float L = params->h * 2;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    out[i] = in[i];
    uint32_t c[3];
    for (int a = 0; a < 3; ++a)
      c[a] = (uint32_t)((in[i].position.s[a] - params->min_point.s[a]) / L);
    out[i].grid_index = orc_morton(c[0], c[1], c[2]);
  }
}

/* sort_particles: sph_simulation.cpp:107-156, sort_count/sort: sort.cl:3-59.
 * 128 "work items", each owning a contiguous chunk (the last one takes the
 * remainder, sort.cl:9-20); per pass a digit histogram laid out
 * bucket-major (counts[bucket*128 + t], sort.cl:35), a host exclusive scan
 * in that flat order (sph_simulation.cpp:133-138) and a scatter in chunk
 * order (sort.cl:50-58). Prefix entries are held at the host's 4-byte width,
 * which is the evident intent of the kernel's size_t* declaration. */
#define ORC_SORT_THREADS 128
#define ORC_BUCKETS 256
void orc_sort_particles(const particle* in, particle* out, particle* scratch,
                        uint32_t n) {
  uint32_t* counts =
      (uint32_t*)malloc(sizeof(uint32_t) * ORC_SORT_THREADS * ORC_BUCKETS);
  const particle* src = in;
  particle* bufs[2] = {scratch, out};
  /* 4 passes: pass 0 writes scratch, 1 writes out, 2 scratch, 3 out. */
  for (int pass = 0; pass < 4; ++pass) {
    particle* dst = bufs[pass & 1];
    uint32_t shift = 8u * (uint32_t)pass;
    memset(counts, 0, sizeof(uint32_t) * ORC_SORT_THREADS * ORC_BUCKETS);
    uint32_t per = n / ORC_SORT_THREADS;
    for (uint32_t t = 0; t < ORC_SORT_THREADS; ++t) {
      uint32_t s = per * t, e = (t == ORC_SORT_THREADS - 1) ? n : s + per;
      for (uint32_t i = s; i < e; ++i)
        counts[((src[i].grid_index >> shift) & 0xFFu) * ORC_SORT_THREADS + t]++;
    }
    uint32_t running = 0;
    for (uint32_t k = 0; k < ORC_SORT_THREADS * ORC_BUCKETS; ++k) {
      uint32_t tmp = counts[k];
      counts[k] = running;
      running += tmp;
    }
    for (uint32_t t = 0; t < ORC_SORT_THREADS; ++t) {
      uint32_t s = per * t, e = (t == ORC_SORT_THREADS - 1) ? n : s + per;
      for (uint32_t i = s; i < e; ++i) {
        uint32_t b = (src[i].grid_index >> shift) & 0xFFu;
        dst[counts[b * ORC_SORT_THREADS + t]++] = src[i];
      }
    }
    src = dst;
  }
  free(counts);
}

/* cell table: sph_simulation.cpp:162-170 */
void orc_cell_table(const particle* sorted, uint32_t n, uint32_t cell_count,
                    uint32_t* table) {
  uint32_t cur = 0;
  for (uint32_t c = 0; c < cell_count; ++c) {
    table[c] = cur;
    while (cur != n && sorted[cur].grid_index == c) cur++;
  }
}

/* ---------------- physics: FP32 and FP64 instantiations ---------------- */
#define REAL float
#define PART particle
#define SFX(x) x##_f32
#define F3(p, f) ((p).f.s)
#define SQRT sqrtf
#define FLOOR floorf
#define EPS7 0.0000001f
#define PI_R ((float)M_PI)
#include "cpu_ref_phys.inc"
#undef REAL
#undef PART
#undef SFX
#undef SQRT
#undef FLOOR
#undef EPS7
#undef PI_R
#undef F3

#define REAL double
#define PART orc_particle64
#define SFX(x) x##_f64
#define F3(p, f) ((p).f)
#define SQRT sqrt
#define FLOOR floor
#define EPS7 1e-7
#define PI_R M_PI
#include "cpu_ref_phys.inc"
#undef REAL
#undef PART
#undef SFX
#undef SQRT
#undef FLOOR
#undef EPS7
#undef PI_R
#undef F3

void orc_to_f64(const particle* in, orc_particle64* out, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    for (int a = 0; a < 4; ++a) {
      out[i].position[a] = in[i].position.s[a];
      out[i].velocity[a] = in[i].velocity.s[a];
      out[i].intermediate_velocity[a] = in[i].intermediate_velocity.s[a];
      out[i].acceleration[a] = in[i].acceleration.s[a];
    }
    out[i].density = in[i].density;
    out[i].pressure = in[i].pressure;
    out[i].grid_index = in[i].grid_index;
  }
}

int orc_step_f32(particle* state, particle* work, particle* work2,
                 uint32_t* table, uint32_t table_cap,
                 simulation_parameters* params,
                 const precomputed_kernel_values* terms,
                 const float* face_normals, const float* vertices,
                 const uint32_t* indices, uint32_t face_count,
                 uint32_t max_bounces, uint32_t* capped) {
  uint32_t n = params->particles_count;
  if (orc_grid_params(state, n, params) != 0) return -1;
  if (params->grid_cell_count > table_cap) return -2;
  orc_locate_in_grid(state, work, n, params);
  orc_sort_particles(work, state, work2, n); /* sorted -> state */
  orc_cell_table(state, n, params->grid_cell_count, table);
  orc_density_pressure_f32(state, work, n, params, terms, table);
  orc_forces_f32(work, work2, n, params, terms, table);
  uint32_t c = orc_advection_collision_f32(work2, state, n, params,
                                           face_normals, vertices, indices,
                                           face_count, max_bounces);
  if (capped) *capped = c;
  return 0;
}

int orc_run_f32(particle* state, uint32_t steps,
                simulation_parameters* params,
                const precomputed_kernel_values* terms,
                const float* face_normals, const float* vertices,
                const uint32_t* indices, uint32_t face_count,
                uint32_t max_bounces, double* mean_density, double* kinetic) {
  uint32_t n = params->particles_count;
  particle* work = (particle*)malloc(sizeof(particle) * (size_t)n);
  particle* work2 = (particle*)malloc(sizeof(particle) * (size_t)n);
  uint32_t cap = 1u << 20;
  uint32_t* table = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
  int rc = 0;
  for (uint32_t s = 0; s < steps && rc == 0; ++s) {
    if (orc_grid_params(state, n, params) == 0 &&
        params->grid_cell_count > cap) {
      free(table);
      cap = params->grid_cell_count;
      table = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
    }
    rc = orc_step_f32(state, work, work2, table, cap, params, terms,
                      face_normals, vertices, indices, face_count, max_bounces,
                      NULL);
    double rho = 0.0, ke = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      rho += state[i].density;
      double vx = state[i].velocity.s[0], vy = state[i].velocity.s[1],
             vz = state[i].velocity.s[2];
      ke += vx * vx + vy * vy + vz * vz;
    }
    if (mean_density) mean_density[s] = rho / (double)n;
    if (kinetic) kinetic[s] = 0.5 * (double)params->particle_mass * ke;
  }
  free(work);
  free(work2);
  free(table);
  return rc;
}
