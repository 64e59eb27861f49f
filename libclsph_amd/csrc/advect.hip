// Leapfrog advection with continuous collision against the scene triangles.
//
// Reference: advection_collision (sph.cl:64-112), advect (advection.cl:6-23),
// detect_collision / respond / handle_collisions (collisions.cl:15-129).
// Collision outcomes hinge on threshold tests (0 <= r <= 1, barycentric
// s, t >= 0), so this kernel keeps the reference's FP32 operation order with
// no contraction and IEEE division / square root. The reference's collision
// do-while has no iteration cap; here it stops after max_bounces passes and
// the particles that hit the cap are counted in DevParams::capped.
//
// The kernel also writes the per-block bounding box of the new positions, so
// the next substep's grid parameters need no separate pass over the state.
// Results go to each particle's slot in the reference's order (stable by
// cell id), which is the order the next substep starts from.
#include "reduce.h"
#include "state.h"

// No multiply-add fusion anywhere in this file, helpers included: the
// reference evaluates these expressions as separate float operations.
#pragma clang fp contract(off)

namespace clsph {

struct V3 {
  float x, y, z;
};
__device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ inline V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
__device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float len(V3 a) { return sqrtf(dot(a, a)); }

// The batch's concatenated meshes; a scene uses faces
// [face_begin, face_begin + face_count) and its own triangle grid.
struct Mesh {
  const float* __restrict__ normals;
  const float* __restrict__ verts;
  const uint32_t* __restrict__ idx;
  const uint32_t* __restrict__ cell_start;
  const uint32_t* __restrict__ cell_tris;
};

struct Hit {
  bool hit;
  uint32_t face;
  V3 point, normal;
  float depth;
};

// Segment p0 -> p1 against triangle f (collisions.cl:21-76): on a hit,
// returns true with the intersection, the normal oriented along the motion
// and the penetration depth |p1 - intersection|.
//
// Most segments are far from most triangles' planes, and those are rejected
// first without the divisions and square roots. Negating the normal negates
// both dot products below exactly, so r = dot(n, v0 - p0) / dot(n, seg) does
// not depend on the orientation step; the reference rejects whenever its
// quotient is outside [0, 1], which |num| > 2 |denom| guarantees (the rounded
// quotient is then at least 2 in magnitude), and so do opposite signs while
// the quotient is far above the underflow range (a quotient that rounds to
// -0 would pass the reference's 0 <= r). Every other segment takes the
// reference's full test.
__device__ inline bool face_hit(uint32_t f, V3 p0, V3 p1, V3 seg, float seg_len, const Mesh& m,
                                V3& isect, V3& nrm, float& depth) {
  nrm = v3(m.normals[3 * f + 0], m.normals[3 * f + 1], m.normals[3 * f + 2]);
  const float* a = m.verts + 3 * m.idx[3 * f + 0];
  {
    const float den0 = dot(nrm, seg);
    const float num0 = dot(nrm, v3(a[0], a[1], a[2]) - p0);
    if (den0 == 0.f) return false;
    const float an = fabsf(num0), ad = fabsf(den0);
    if (an > 2.f * ad) return false;
    if ((num0 < 0.f) != (den0 < 0.f) && an > ad * 0x1p-60f) return false;
  }
  if (dot(nrm, seg) / (len(nrm) * seg_len) <= 0.f) nrm = v3(-nrm.x, -nrm.y, -nrm.z);
  const float* b = m.verts + 3 * m.idx[3 * f + 1];
  const float* c = m.verts + 3 * m.idx[3 * f + 2];
  V3 v0 = v3(a[0], a[1], a[2]);
  V3 u = v3(b[0], b[1], b[2]) - v0;
  V3 v = v3(c[0], c[1], c[2]) - v0;
  float denom = dot(nrm, seg);
  if (denom == 0.f) return false;
  float r = dot(nrm, v0 - p0) / denom;
  if (!(0.f <= r && r <= 1.f)) return false;
  isect = p0 + r * seg;
  V3 w = isect - v0;
  float uv = dot(u, v), wv = dot(w, v), vv = dot(v, v), wu = dot(w, u), uu = dot(u, u);
  float d2 = uv * uv - uu * vv;
  float s = (uv * wv - vv * wu) / d2;
  float t = (uv * wu - uu * wv) / d2;
  if (!(s >= 0.f && t >= 0.f && s + t <= 1.f)) return false;
  depth = len(p1 - isect);
  return true;
}

// detect_collision (collisions.cl:15-89): the nearest hit along p0 -> p1; a
// hit at equal distance replaces the earlier one, so the highest index among
// the nearest wins. Without a grid every triangle is tested in index order
// with the reference's test; with a grid only the triangles listed in the
// cells the segment's box meets are tested (the others cannot be hit), in
// cell order, and "nearest, then highest index" is applied explicitly.
__device__ inline Hit detect(V3 p0, V3 p1, const Mesh& m, const DevParams* __restrict__ dp) {
  const TriGrid& grid = dp->tgrid;
  Hit h;
  h.hit = false;
  h.face = 0;
  h.depth = 0.f;
  h.point = h.normal = v3(0.f, 0.f, 0.f);
  V3 seg = p1 - p0;
  float seg_len = len(seg);
  if (!grid.enabled) {
    for (uint32_t f = dp->face_begin; f < dp->face_begin + dp->face_count; ++f) {
      V3 isect, nrm;
      float depth;
      if (!face_hit(f, p0, p1, seg, seg_len, m, isect, nrm, depth)) continue;
      if (h.hit && len(p0 - isect) > len(p0 - h.point)) continue;
      h.hit = true;
      h.face = f;
      h.point = isect;
      h.normal = nrm;
      h.depth = depth;
    }
    return h;
  }
  float lo[3] = {fminf(p0.x, p1.x), fminf(p0.y, p1.y), fminf(p0.z, p1.z)};
  float hi[3] = {fmaxf(p0.x, p1.x), fmaxf(p0.y, p1.y), fmaxf(p0.z, p1.z)};
  uint32_t clo[3], chi[3];
  for (int a = 0; a < 3; ++a) {
    float cl = floorf((lo[a] - grid.origin[a]) * grid.inv_cell);
    float ch = floorf((hi[a] - grid.origin[a]) * grid.inv_cell);
    float top = (float)(grid.dims[a] - 1);
    // the grid covers every triangle with margin: a box outside it hits nothing
    if (!(ch >= 0.f) || !(cl <= top)) return h;
    clo[a] = (uint32_t)fmaxf(cl, 0.f);
    chi[a] = (uint32_t)fminf(ch, top);
  }
  for (uint32_t z = clo[2]; z <= chi[2]; ++z)
    for (uint32_t y = clo[1]; y <= chi[1]; ++y)
      for (uint32_t x = clo[0]; x <= chi[0]; ++x) {
        uint32_t c = grid.cell_base + (z * grid.dims[1] + y) * grid.dims[0] + x;
        for (uint32_t q = m.cell_start[c]; q < m.cell_start[c + 1]; ++q) {
          uint32_t f = m.cell_tris[q];
          V3 isect, nrm;
          float depth;
          if (!face_hit(f, p0, p1, seg, seg_len, m, isect, nrm, depth)) continue;
          if (h.hit) {
            float d = len(p0 - isect), db = len(p0 - h.point);
            if (d > db || (d == db && f <= h.face)) continue;
          }
          h.hit = true;
          h.face = f;
          h.point = isect;
          h.normal = nrm;
          h.depth = depth;
        }
      }
  return h;
}

__global__ __launch_bounds__(kMaxBlock) void k_advect(
    const float4* __restrict__ pos_s, const float4* __restrict__ vh_s,
    const float4* __restrict__ acc_s, const uint32_t* __restrict__ slot, Mesh mesh,
    const BlockDesc* __restrict__ blocks, DevParams* __restrict__ dps,
    float4* __restrict__ pos, float4* __restrict__ vel, float4* __restrict__ vh,
    float* __restrict__ aabb_partial, uint32_t aabb_stride, uint32_t n) {
  uint32_t scene;
  uint32_t i = particle_of(blocks, dps, scene);
  bool valid = i != 0xFFFFFFFFu;
  DevParams* dp = dps + scene;
  float4 outp = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t capped = 0;
  if (valid) {
    const float vmax = dp->p.max_velocity, e = dp->p.restitution;
    const uint32_t max_b = dp->max_bounces;
    float time_to_go = dp->p.time_delta * dp->p.simulation_scale;
    float4 p4 = pos_s[i], vh4 = vh_s[i], a4 = acc_s[i];
    V3 p = v3(p4.x, p4.y, p4.z), v = v3(vh4.x, vh4.y, vh4.z), a = v3(a4.x, a4.y, a4.z);
    V3 rpos = p, rvel = v;
    bool happened;
    uint32_t it = 0;
    do {
      // advect (advection.cl:6-23)
      V3 nv = v + a * time_to_go;
      float nl = len(nv);
      if (nl > vmax) nv = (nv / nl) * vmax;
      V3 np = p + nv * time_to_go;
      // handle_collisions (collisions.cl:109-129)
      rpos = np;
      rvel = nv;
      float elapsed = time_to_go;
      Hit hh = detect(p, np, mesh, dp);
      happened = hh.hit;
      if (happened) {
        // respond (collisions.cl:91-104)
        rpos = hh.point - hh.normal * 0.001f;
        float k = (1.f + e * hh.depth / (time_to_go * len(rvel))) * dot(rvel, hh.normal);
        rvel = rvel - k * hh.normal;
        elapsed = time_to_go * (len(rpos - p) / len(np - p));
      }
      p = rpos;
      v = rvel;
      time_to_go -= elapsed;
      a = v3(0.f, 0.f, 0.f);
      ++it;
    } while (happened && it < max_b);
    capped = happened ? 1u : 0u;
    V3 vo = (v3(vh4.x, vh4.y, vh4.z) + rvel) / 2.f;
    outp = make_float4(rpos.x, rpos.y, rpos.z, 0.f);
    uint32_t o = slot[i];
    if (o < n) {  // a valid permutation always passes; this guards misuse
      pos[o] = outp;
      vel[o] = make_float4(vo.x, vo.y, vo.z, 0.f);
      vh[o] = make_float4(rvel.x, rvel.y, rvel.z, 0.f);
    }
  }
  // Work-group bounding box of the new positions for the next substep's grid
  // (the work-group lies in one scene), and the count of particles that
  // reached the bounce cap.
  block_aabb_partial(outp, valid, aabb_partial, aabb_stride);
  __shared__ uint32_t cap_red[kMaxBlock / 64];
  uint64_t cb = __builtin_amdgcn_ballot_w64(capped != 0);
  if ((threadIdx.x & 63) == 0) cap_red[threadIdx.x >> 6] = (uint32_t)__builtin_popcountll(cb);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    for (uint32_t q = 0; q < (blockDim.x + 63) / 64; ++q) c += cap_red[q];
    if (c) atomicAdd(&dp->capped, c);
  }
}

void launch_advect(clsph_state* s) {
  if (s->n == 0) return;
  Mesh m{s->face_normals, s->vertices, s->indices, s->tgrid_start, s->tgrid_tris};
  hipLaunchKernelGGL(k_advect, dim3(s->nblocks_of(kStageAdvect)), dim3(s->bs_of(kStageAdvect)), 0,
                     s->ctx->stream, s->pos_s, s->vh_s, s->acc_s, s->slot, m,
                     s->table_of(kStageAdvect), s->dps, s->pos, s->vel, s->vh, s->aabb_partial, s->aabb_stride,
                     s->n);
}

}  // namespace clsph
