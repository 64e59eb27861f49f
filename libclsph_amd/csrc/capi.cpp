// C ABI of the clsph HIP pipeline (include/clsph_hip.h).
#include <math.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/clsph_hip.h"
#include "hostutil.h"
#include "nn.h"
#include "state.h"

using namespace clsph;
using namespace clsph::host;

namespace {

int read_dps(clsph_state* st, std::vector<DevParams>& out, BatchParams* bp = nullptr) {
  out.resize(st->n_scenes);
  hipStream_t s = st->ctx->stream;
  HIPCK(st->ctx, hipMemcpyAsync(out.data(), st->dps, sizeof(DevParams) * st->n_scenes,
                                hipMemcpyDeviceToHost, s));
  if (bp) HIPCK(st->ctx, hipMemcpyAsync(bp, st->bp, sizeof(BatchParams), hipMemcpyDeviceToHost, s));
  HIPCK(st->ctx, hipStreamSynchronize(s));
  return CLSPH_OK;
}

void free_state_buffers(clsph_state* st) {
  dfree(st->pos), dfree(st->vel), dfree(st->vh);
  dfree(st->pos_s), dfree(st->vel_s), dfree(st->vh_s), dfree(st->acc_s);
  dfree(st->rho_s), dfree(st->prs_s), dfree(st->fab);
  dfree(st->nbr), dfree(st->nbr_count);
  dfree(st->key_a), dfree(st->key_t), dfree(st->val_t), dfree(st->key_u), dfree(st->val_u);
  dfree(st->key_b), dfree(st->val_b);
  dfree(st->sort_hist), dfree(st->sort_totals), dfree(st->sort_runs), dfree(st->sort_runs_ctl);
  dfree(st->table), dfree(st->slot);
  dfree(st->ids), dfree(st->ids_tmp);
  dfree(st->aabb_partial);
  dfree(st->face_normals), dfree(st->vertices), dfree(st->indices);
  dfree(st->tgrid_start), dfree(st->tgrid_tris);
  dfree(st->dps), dfree(st->bp);
  for (auto& b : st->blocks) dfree(b);
  if (st->graph) (void)hipGraphExecDestroy(st->graph);
  st->graph = nullptr;
  if (st->lgraph) (void)hipGraphExecDestroy(st->lgraph);
  st->lgraph = nullptr;
  if (st->copy_stream) {
    (void)hipStreamSynchronize(st->copy_stream);
    (void)hipStreamDestroy(st->copy_stream);
    st->copy_stream = nullptr;
  }
  for (int k = 0; k < 2; ++k) {
    dfree(st->stage_aos[k]), dfree(st->stage_meta[k]);
    if (st->host_meta[k]) (void)hipHostFree(st->host_meta[k]);
    st->host_meta[k] = nullptr;
    if (st->ev_pack[k]) (void)hipEventDestroy(st->ev_pack[k]);
    if (st->ev_done[k]) (void)hipEventDestroy(st->ev_done[k]);
    st->ev_pack[k] = st->ev_done[k] = nullptr;
    st->slot_used[k] = false;
  }
}

// Work-group tables for every block size, and each scene's range in the
// advection table (where the bounding-box partials live).
int build_blocks(clsph_state* st) {
  for (int b = 0; b < kNumBlockSizes; ++b) {
    uint32_t bs = (uint32_t)kBlockSizes[b];
    std::vector<BlockDesc> t;
    for (uint32_t s = 0; s < st->n_scenes; ++s) {
      if (b == st->stage_bs[kStageAdvect]) st->host_dps[s].block_begin = (uint32_t)t.size();
      for (uint32_t f = 0; f < st->scene_n[s]; f += bs) t.push_back({s, st->scene_begin[s] + f});
      if (b == st->stage_bs[kStageAdvect]) st->host_dps[s].block_end = (uint32_t)t.size();
    }
    dfree(st->blocks[b]);
    int rc = dalloc(st->ctx, &st->blocks[b], t.size());
    if (rc) return rc;
    if (!t.empty())
      HIPCK(st->ctx, hipMemcpy(st->blocks[b], t.data(), sizeof(BlockDesc) * t.size(),
                               hipMemcpyHostToDevice));
    st->nblocks[b] = (uint32_t)t.size();
  }
  st->graph_dirty = st->lgraph_dirty = true;
  return CLSPH_OK;
}

int upload_params(clsph_state* st) {
  st->lgraph_dirty = true;  // the learned step's graph holds parameters by value
  for (uint32_t s = 0; s < st->n_scenes; ++s) {
    st->host_dps[s].begin = st->scene_begin[s];
    st->host_dps[s].n = st->scene_n[s];
    st->host_dps[s].err = 0;
    st->host_dps[s].capped = 0;
  }
  BatchParams b{};
  b.n_scenes = st->n_scenes;
  b.n = st->n;
  b.table_cap = st->table_cap;
  b.sort_passes = st->sort_passes;
  HIPCK(st->ctx, hipMemcpyAsync(st->dps, st->host_dps.data(), sizeof(DevParams) * st->n_scenes,
                                hipMemcpyHostToDevice, st->ctx->stream));
  HIPCK(st->ctx, hipMemcpyAsync(st->bp, &b, sizeof(b), hipMemcpyHostToDevice, st->ctx->stream));
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  return CLSPH_OK;
}

// Uniform grid over one scene's mesh for collision queries: cell c lists
// every triangle whose bounding box, padded by a rounding margin, meets it.
// Face indices are global (face_begin added); cell starts are global offsets.
void build_tri_grid(const clsph_state::Mesh& m, uint32_t face_begin, float cell_size,
                    TriGrid& g, std::vector<uint32_t>& starts, std::vector<uint32_t>& tris) {
  g = TriGrid{};
  uint32_t faces = (uint32_t)(m.indices.size() / 3);
  if (faces == 0 || cell_size < 0.f) return;
  const float* verts = m.vertices.data();
  const uint32_t* idx = m.indices.data();
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  double ext_sum = 0.0;
  std::vector<float> box(6ull * faces);
  for (uint32_t f = 0; f < faces; ++f) {
    float* tb = &box[6ull * f];
    for (int a = 0; a < 3; ++a) tb[a] = 1e30f, tb[3 + a] = -1e30f;
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) {
        float v = verts[3 * idx[3 * f + k] + a];
        tb[a] = std::min(tb[a], v);
        tb[3 + a] = std::max(tb[3 + a], v);
      }
    float e = 0.f;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], tb[a]);
      hi[a] = std::max(hi[a], tb[3 + a]);
      e = std::max(e, tb[3 + a] - tb[a]);
    }
    ext_sum += e;
  }
  if (!(cell_size > 0.f)) cell_size = (float)(2.0 * ext_sum / faces);
  float span = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  if (!(cell_size > 0.f)) cell_size = span > 0.f ? span : 1.f;
  // Rounding margin: hit points are computed in FP32 and may stray from the
  // exact triangle by a few ulps of the coordinates involved.
  float mag = 0.f;
  for (int a = 0; a < 3; ++a) mag = std::max(mag, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
  float pad = 1e-5f * (mag + cell_size) + 1e-6f;
  uint32_t dims[3];
  for (;;) {
    uint64_t total = 1;
    for (int a = 0; a < 3; ++a) {
      double ext = (double)(hi[a] - lo[a]) + 2.0 * (pad + cell_size);
      dims[a] = (uint32_t)std::max(1.0, std::ceil(ext / (double)cell_size));
      total *= dims[a];
    }
    if (total <= (1ull << 25)) break;
    cell_size *= 1.5f;
  }
  float origin[3];
  for (int a = 0; a < 3; ++a) origin[a] = lo[a] - pad - cell_size;
  float inv = 1.f / cell_size;
  auto cell_of = [&](float v, int a) {
    float c = std::floor((v - origin[a]) * inv);
    c = std::max(0.f, std::min(c, (float)(dims[a] - 1)));
    return (uint32_t)c;
  };
  uint64_t ncells = (uint64_t)dims[0] * dims[1] * dims[2];
  std::vector<uint32_t> count(ncells, 0);
  std::vector<uint32_t> range(6ull * faces);
  for (uint32_t f = 0; f < faces; ++f) {
    for (int a = 0; a < 3; ++a) {
      range[6 * f + a] = cell_of(box[6ull * f + a] - pad, a);
      range[6 * f + 3 + a] = cell_of(box[6ull * f + 3 + a] + pad, a);
    }
    for (uint32_t z = range[6 * f + 2]; z <= range[6 * f + 5]; ++z)
      for (uint32_t y = range[6 * f + 1]; y <= range[6 * f + 4]; ++y)
        for (uint32_t x = range[6 * f + 0]; x <= range[6 * f + 3]; ++x)
          count[((uint64_t)z * dims[1] + y) * dims[0] + x]++;
  }
  g.cell_base = (uint32_t)starts.size();
  uint32_t run = (uint32_t)tris.size();
  std::vector<uint32_t> fill(ncells);
  for (uint64_t c = 0; c < ncells; ++c) {
    starts.push_back(run);
    fill[c] = run;
    run += count[c];
  }
  starts.push_back(run);
  tris.resize(run);
  for (uint32_t f = 0; f < faces; ++f)
    for (uint32_t z = range[6 * f + 2]; z <= range[6 * f + 5]; ++z)
      for (uint32_t y = range[6 * f + 1]; y <= range[6 * f + 4]; ++y)
        for (uint32_t x = range[6 * f + 0]; x <= range[6 * f + 3]; ++x)
          tris[fill[((uint64_t)z * dims[1] + y) * dims[0] + x]++] = face_begin + f;
  for (int a = 0; a < 3; ++a) {
    g.origin[a] = origin[a];
    g.dims[a] = dims[a];
  }
  g.inv_cell = inv;
  g.enabled = 1;
}

// Concatenates every scene's mesh (and triangle grid) into the device arrays.
int rebuild_meshes(clsph_state* st) {
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  std::vector<float> normals, verts;
  std::vector<uint32_t> idx, starts, tris;
  for (uint32_t s = 0; s < st->n_scenes; ++s) {
    const clsph_state::Mesh& m = st->meshes[s];
    uint32_t vbase = (uint32_t)(verts.size() / 3);
    uint32_t fbase = (uint32_t)(idx.size() / 3);
    normals.insert(normals.end(), m.normals.begin(), m.normals.end());
    verts.insert(verts.end(), m.vertices.begin(), m.vertices.end());
    for (uint32_t k : m.indices) idx.push_back(vbase + k);
    st->host_dps[s].face_begin = fbase;
    st->host_dps[s].face_count = (uint32_t)(m.indices.size() / 3);
    build_tri_grid(m, fbase, m.grid_cell, st->host_dps[s].tgrid, starts, tris);
  }
  dfree(st->face_normals), dfree(st->vertices), dfree(st->indices);
  dfree(st->tgrid_start), dfree(st->tgrid_tris);
  int rc = dalloc(st->ctx, &st->face_normals, normals.size());
  if (!rc) rc = dalloc(st->ctx, &st->vertices, verts.size());
  if (!rc) rc = dalloc(st->ctx, &st->indices, idx.size());
  if (!rc) rc = dalloc(st->ctx, &st->tgrid_start, starts.size());
  if (!rc) rc = dalloc(st->ctx, &st->tgrid_tris, tris.size());
  if (rc) return rc;
  auto up = [&](void* d, const void* h, size_t bytes) -> int {
    if (bytes) HIPCK(st->ctx, hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return CLSPH_OK;
  };
  rc = up(st->face_normals, normals.data(), 4 * normals.size());
  if (!rc) rc = up(st->vertices, verts.data(), 4 * verts.size());
  if (!rc) rc = up(st->indices, idx.data(), 4 * idx.size());
  if (!rc) rc = up(st->tgrid_start, starts.data(), 4 * starts.size());
  if (!rc) rc = up(st->tgrid_tris, tris.data(), 4 * tris.size());
  if (rc) return rc;
  st->graph_dirty = st->lgraph_dirty = true;
  // Only the mesh fields change; the rest of the device parameters is kept.
  for (uint32_t s = 0; s < st->n_scenes; ++s) {
    const DevParams& h = st->host_dps[s];
    size_t off = offsetof(DevParams, face_begin);
    HIPCK(st->ctx, hipMemcpy(reinterpret_cast<char*>(st->dps + s) + off,
                             reinterpret_cast<const char*>(&h) + off,
                             offsetof(DevParams, block_begin) - off, hipMemcpyHostToDevice));
    off = offsetof(DevParams, tgrid);
    HIPCK(st->ctx, hipMemcpy(reinterpret_cast<char*>(st->dps + s) + off,
                             reinterpret_cast<const char*>(&h) + off, sizeof(TriGrid),
                             hipMemcpyHostToDevice));
  }
  return CLSPH_OK;
}

// Launches the kernels of one whole substep on the context stream.
void enqueue_step(clsph_state* st) {
  launch_grid_params(st);
  launch_hash(st);
  launch_sort(st);
  launch_density(st);
  launch_forces(st);
  launch_advect(st);
  if (st->ids) launch_track_ids(st);
}

int run_stage(clsph_state* st, int stage) {
  switch (stage) {
    case kStageGrid:
      launch_grid_params(st);
      break;
    case kStageHash:
      launch_hash(st);
      break;
    case kStageSort:
      // sorting also builds the tables and the canonical slots, so that a
      // slot is always current once particles have been reordered
      launch_sort(st);
      st->fields_canonical = false;
      break;
    case kStageTable:
      launch_cell_table(st);
      break;
    case kStageDensity:
      launch_density(st);
      break;
    case kStageForces:
      launch_forces(st);
      break;
    case kStageAdvect:
      launch_advect(st);
      if (st->ids) launch_track_ids(st);
      st->acc_zero = true;
      break;
    default:
      return fail(st->ctx, CLSPH_EINVAL, "unknown stage " + std::to_string(stage));
  }
  st->stage = stage;
  st->acc_learned = false;
  return check_launch(st->ctx);
}

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

int alloc_state(clsph_ctx* ctx, uint32_t n_scenes, const uint32_t* caps, bool fixed,
                uint32_t table_cap, clsph_state** out) {
  if (!ctx || !out || n_scenes == 0 || !caps)
    return fail(ctx, CLSPH_EINVAL, "state allocation: bad argument");
  *out = nullptr;
  HIPCK(ctx, hipSetDevice(ctx->device));
  clsph_state* st = new clsph_state();
  st->ctx = ctx;
  st->n_scenes = n_scenes;
  st->fixed_counts = fixed;
  uint64_t total = 0, blocks64 = 0;
  for (uint32_t s = 0; s < n_scenes; ++s) {
    st->scene_begin.push_back((uint32_t)total);
    st->scene_cap.push_back(caps[s]);
    st->scene_n.push_back(fixed ? caps[s] : 0);
    total += caps[s];
    blocks64 += (caps[s] + 63) / 64;
  }
  if (total > 0x7FFFFFFFull) {
    delete st;
    return fail(ctx, CLSPH_EINVAL, "too many particles for one state");
  }
  st->n_cap = (uint32_t)total;
  st->n = fixed ? (uint32_t)total : 0;
  uint64_t cells = table_cap ? table_cap : std::max<uint64_t>(1u << 20, 4ull * total);
  if (cells > (1ull << 28)) {
    delete st;
    return fail(ctx, CLSPH_EINVAL, "table_cap above 2^28 cells is not supported");
  }
  st->table_cap = (uint32_t)(8ull * cells);
  // LSD passes of at most 11 bits over the widest cell id the table admits
  st->sort_passes = bitlen((uint32_t)(cells - 1)) <= 2u * kMaxDigitBits ? 2u : 3u;
  uint32_t n_cap = st->n_cap;
  st->n_tiles_cap = sort_tiles_for(n_cap ? n_cap : 1);
  int rc = CLSPH_OK;
  float4** f4[] = {&st->pos, &st->vel, &st->vh, &st->vel_s, &st->vh_s,
                   &st->acc_s};
  for (auto p : f4)
    if (rc == CLSPH_OK) rc = dalloc(ctx, p, n_cap);
  // sorted positions carry 4 padding entries: density loads candidates in
  // groups of 4 without clamping the last group to the run's end
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->pos_s, (size_t)n_cap + 4);
  if (rc == CLSPH_OK)
    HIPCK(ctx, hipMemset(st->pos_s + n_cap, 0, 4 * sizeof(float4)));
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->fab, 2 * (size_t)n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->rho_s, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->prs_s, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->nbr, (size_t)(kMaxNeighbours / 4) * n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->nbr_count, n_cap);
  uint32_t** u4[] = {&st->key_a, &st->key_t, &st->val_t, &st->key_u, &st->val_u,
                     &st->key_b, &st->val_b, &st->slot};
  for (auto p : u4)
    if (rc == CLSPH_OK) rc = dalloc(ctx, p, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->sort_hist, (size_t)st->n_tiles_cap * kMaxBins);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->sort_totals, (size_t)kMaxBins);
  st->sort_runs_cap = sort_runs_cap_for(st->table_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->sort_runs, (size_t)st->sort_runs_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->sort_runs_ctl, 2);
  if (rc == CLSPH_OK) HIPCK(ctx, hipMemset(st->sort_runs_ctl, 0, 2 * sizeof(uint32_t)));
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->table, (size_t)st->table_cap + 1);
  st->aabb_stride = (uint32_t)(blocks64 + 1);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->aabb_partial, (size_t)st->aabb_stride * 6);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->dps, n_scenes);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->bp, 1);
  if (rc != CLSPH_OK) {
    free_state_buffers(st);
    delete st;
    return rc;
  }
  st->host_dps.assign(n_scenes, DevParams{});
  for (auto& d : st->host_dps) d.max_bounces = 64;
  st->meshes.resize(n_scenes);
  if (!rc) rc = build_blocks(st);
  if (!rc) rc = upload_params(st);
  if (!rc) rc = rebuild_meshes(st);
  if (rc) {
    free_state_buffers(st);
    delete st;
    return rc;
  }
  *out = st;
  return CLSPH_OK;
}

int set_params(clsph_state* st, uint32_t scene, const clsph_simulation_parameters* params,
               const clsph_precomputed_kernel_values* terms, uint32_t max_bounces) {
  if (!st || !params || !terms) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (scene >= st->n_scenes) return fail(st->ctx, CLSPH_EINVAL, "scene index out of range");
  if (!(params->h > 0.f)) return fail(st->ctx, CLSPH_EINVAL, "h must be positive");
  DevParams& d = st->host_dps[scene];
  d.p = *params;
  d.t = *terms;
  d.max_bounces = max_bounces ? max_bounces : 64;
  return upload_params(st);
}

int upload(clsph_state* st, uint32_t scene, const clsph_particle* p, uint32_t n) {
  if (!st || (!p && n)) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (scene >= st->n_scenes) return fail(st->ctx, CLSPH_EINVAL, "scene index out of range");
  if (n > st->scene_cap[scene]) return fail(st->ctx, CLSPH_EINVAL, "more particles than the scene holds");
  if (st->fixed_counts && n != st->scene_n[scene])
    return fail(st->ctx, CLSPH_EINVAL, "a batch scene takes exactly its allocated count");
  std::vector<float4> pos(n), vel(n), vh(n), acc(n);
  std::vector<float> rho(n), prs(n);
  std::vector<uint32_t> key(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* a = p[i].position.s;
    const float* b = p[i].velocity.s;
    const float* c = p[i].intermediate_velocity.s;
    const float* d = p[i].acceleration.s;
    pos[i] = make_float4(a[0], a[1], a[2], 0.f);
    vel[i] = make_float4(b[0], b[1], b[2], 0.f);
    vh[i] = make_float4(c[0], c[1], c[2], 0.f);
    acc[i] = make_float4(d[0], d[1], d[2], 0.f);
    rho[i] = p[i].density;
    prs[i] = p[i].pressure;
    key[i] = p[i].grid_index;
  }
  uint32_t b0 = st->scene_begin[scene];
  hipStream_t s = st->ctx->stream;
  auto put = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes) HIPCK(st->ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return CLSPH_OK;
  };
  int rc = put(st->pos + b0, pos.data(), sizeof(float4) * n);
  if (!rc) rc = put(st->vel + b0, vel.data(), sizeof(float4) * n);
  if (!rc) rc = put(st->vh + b0, vh.data(), sizeof(float4) * n);
  if (!rc) rc = put(st->acc_s + b0, acc.data(), sizeof(float4) * n);
  if (!rc) rc = put(st->rho_s + b0, rho.data(), sizeof(float) * n);
  if (!rc) rc = put(st->prs_s + b0, prs.data(), sizeof(float) * n);
  if (!rc) rc = put(st->key_b + b0, key.data(), sizeof(uint32_t) * n);
  if (rc) return rc;
  if (n != st->scene_n[scene]) {
    st->scene_n[scene] = n;
    st->n = n;  // single-scene states only (batches keep their counts)
    rc = build_blocks(st);
    if (!rc) rc = upload_params(st);
    if (rc) return rc;
  }
  launch_aabb(st);
  st->acc_zero = false;
  st->fields_canonical = true;
  st->stage = -1;
  rc = check_launch(st->ctx);
  if (rc) return rc;
  HIPCK(st->ctx, hipStreamSynchronize(s));
  return CLSPH_OK;
}

// Copies scene `scene` to host particles. Fields computed in half-cell order
// go to each particle's canonical slot; `sorted` returns the reference's
// sorted buffer (stable by cell id) of the last substep instead of the state.
int download(clsph_state* st, uint32_t scene, clsph_particle* out, uint32_t n, bool sorted) {
  if (!st || (!out && n)) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (scene >= st->n_scenes) return fail(st->ctx, CLSPH_EINVAL, "scene index out of range");
  uint32_t sn = st->scene_n[scene];
  if (n > sn) return fail(st->ctx, CLSPH_EINVAL, "asked for more particles than are held");
  if (n == 0) return device_errors(st);
  if (n != sn && !st->fields_canonical)
    return fail(st->ctx, CLSPH_EINVAL, "partial downloads need the whole particle count");
  uint32_t b0 = st->scene_begin[scene];
  uint32_t cnt = n;
  std::vector<float4> pos(cnt), vel(cnt), vh(cnt), acc(cnt);
  std::vector<float> rho(cnt), prs(cnt);
  std::vector<uint32_t> key(cnt), slot(cnt);
  hipStream_t s = st->ctx->stream;
  const float4* P = (sorted ? st->pos_s : st->pos) + b0;
  const float4* V = (sorted ? st->vel_s : st->vel) + b0;
  const float4* H = (sorted ? st->vh_s : st->vh) + b0;
  auto get = [&](void* dst, const void* src, size_t bytes) -> int {
    HIPCK(st->ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    return CLSPH_OK;
  };
  int rc = get(pos.data(), P, sizeof(float4) * cnt);
  if (!rc) rc = get(vel.data(), V, sizeof(float4) * cnt);
  if (!rc) rc = get(vh.data(), H, sizeof(float4) * cnt);
  bool want_acc = sorted || !st->acc_zero;
  if (!rc && want_acc) rc = get(acc.data(), st->acc_s + b0, sizeof(float4) * cnt);
  if (!rc) rc = get(rho.data(), st->rho_s + b0, sizeof(float) * cnt);
  if (!rc) rc = get(prs.data(), st->prs_s + b0, sizeof(float) * cnt);
  if (!rc) rc = get(key.data(), st->key_b + b0, sizeof(uint32_t) * cnt);
  bool reorder = !st->fields_canonical;
  if (!rc && reorder) rc = get(slot.data(), st->slot + b0, sizeof(uint32_t) * cnt);
  if (rc) return rc;
  std::vector<DevParams> d;
  rc = read_dps(st, d);  // also waits for the copies above
  if (rc) return rc;
  uint32_t kbase = reorder ? d[scene].key_base : 0;
  for (uint32_t k = 0; k < cnt; ++k) {
    uint32_t o = reorder ? slot[k] - b0 : k;
    if (o >= n) return fail(st->ctx, CLSPH_EDEVICE, "canonical slots are not a permutation");
    clsph_particle& q = out[o];
    memset(&q, 0, sizeof(q));
    q.density = rho[k];
    q.pressure = prs[k];
    q.grid_index = reorder ? (key[k] - kbase) >> 3 : key[k];
    if (want_acc && (sorted || !reorder)) {
      q.acceleration.s[0] = acc[k].x, q.acceleration.s[1] = acc[k].y, q.acceleration.s[2] = acc[k].z;
    }
  }
  for (uint32_t k = 0; k < cnt; ++k) {
    uint32_t o = (sorted && reorder) ? slot[k] - b0 : k;
    clsph_particle& q = out[o];
    q.position.s[0] = pos[k].x, q.position.s[1] = pos[k].y, q.position.s[2] = pos[k].z;
    q.velocity.s[0] = vel[k].x, q.velocity.s[1] = vel[k].y, q.velocity.s[2] = vel[k].z;
    q.intermediate_velocity.s[0] = vh[k].x, q.intermediate_velocity.s[1] = vh[k].y,
    q.intermediate_velocity.s[2] = vh[k].z;
  }
  return device_errors(st);
}

int scene_upload(clsph_state* st, uint32_t scene, const float* face_normals,
                 const float* vertices, uint32_t vertex_count, const uint32_t* indices,
                 uint32_t face_count) {
  if (!st) return fail(nullptr, CLSPH_EINVAL, "NULL state");
  if (scene >= st->n_scenes) return fail(st->ctx, CLSPH_EINVAL, "scene index out of range");
  if (face_count && (!face_normals || !vertices || !indices))
    return fail(st->ctx, CLSPH_EINVAL, "NULL mesh array");
  for (uint64_t k = 0; k < 3ull * face_count; ++k)
    if (indices[k] >= vertex_count)
      return fail(st->ctx, CLSPH_EINVAL, "triangle index out of range of the vertex array");
  clsph_state::Mesh& m = st->meshes[scene];
  m.normals.assign(face_normals, face_normals + 3ull * face_count);
  m.vertices.assign(vertices, vertices + 3ull * vertex_count);
  m.indices.assign(indices, indices + 3ull * face_count);
  m.grid_cell = face_count > 64 ? 0.f : -1.f;
  return rebuild_meshes(st);
}

int read_u32(clsph_state* st, const uint32_t* src, uint32_t* dst, size_t count) {
  if (count == 0) return CLSPH_OK;
  HIPCK(st->ctx, hipMemcpyAsync(dst, src, sizeof(uint32_t) * count, hipMemcpyDeviceToHost,
                                st->ctx->stream));
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  return CLSPH_OK;
}

int ensure_graph(clsph_state* st) {
  if (!st->graph_dirty && st->graph) return CLSPH_OK;
  hipStream_t s = st->ctx->stream;
  if (st->graph) {
    // Replays of the old executable may still be queued on the stream: let
    // them finish before it is destroyed.
    HIPCK(st->ctx, hipStreamSynchronize(s));
    (void)hipGraphExecDestroy(st->graph);
    st->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCK(st->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  enqueue_step(st);
  hipError_t le = hipGetLastError();
  HIPCK(st->ctx, hipStreamEndCapture(s, &g));
  if (le != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(st->ctx, CLSPH_EDEVICE, std::string("capture: ") + hipGetErrorString(le));
  }
  hipError_t ie = hipGraphInstantiate(&st->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    st->graph = nullptr;
    return fail(st->ctx, CLSPH_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
  }
  st->graph_dirty = false;
  return CLSPH_OK;
}

}  // namespace

namespace clsph {
namespace host {

int device_errors(clsph_state* st) {
  std::vector<DevParams> d;
  BatchParams b;
  int rc = read_dps(st, d, &b);
  if (rc) return rc;
  uint32_t err = b.err;
  for (const DevParams& x : d) err |= x.err;
  if (err & kErrGridDim)
    return fail(st->ctx, CLSPH_EGRID,
                "grid dimension >= 1024 cells on an axis (particles spread too far)");
  if (err & kErrTableCap)
    return fail(st->ctx, CLSPH_ETABLE, "grid cells exceed the cell table capacity");
  if (err & kErrFineKey)
    return fail(st->ctx, CLSPH_ETABLE, "grid_cell_count reached 2^29 cells");
  return CLSPH_OK;
}

int resize_single(clsph_state* st, uint32_t n) {
  if (st->n_scenes != 1 || st->fixed_counts || n > st->scene_cap[0])
    return fail(st->ctx, CLSPH_EINVAL, "only a single-scene state can change its count");
  st->scene_n[0] = n;
  st->n = n;
  int rc = build_blocks(st);
  if (!rc) rc = upload_params(st);
  return rc;
}

}  // namespace host
}  // namespace clsph

extern "C" {

int clsph_init(int device, clsph_ctx** out) {
  if (!out) return fail(nullptr, CLSPH_EINVAL, "clsph_init: out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(nullptr, CLSPH_EDEVICE,
                std::string("no HIP device available: ") + hipGetErrorString(e));
  if (device < 0) device = 0;
  if (device >= count) return fail(nullptr, CLSPH_EINVAL, "device index out of range");
  clsph_ctx* ctx = new clsph_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return fail(nullptr, CLSPH_EDEVICE, "cannot create a stream on the device");
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cus = prop.multiProcessorCount;
  hipError_t se = clsph::nn::scratch_address(device, &ctx->scratch);
  if (se != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return fail(nullptr, CLSPH_EDEVICE,
                std::string("device workspace address: ") + hipGetErrorString(se));
  }
  *out = ctx;
  return CLSPH_OK;
}

int clsph_free(clsph_ctx* ctx) {
  if (!ctx) return CLSPH_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CLSPH_OK;
}

const char* clsph_last_error(const clsph_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : g_err.c_str();
}

int clsph_device_name(const clsph_ctx* ctx, char* buf, size_t len) {
  if (!ctx || !buf || len == 0) return CLSPH_EINVAL;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) return CLSPH_EDEVICE;
  snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return CLSPH_OK;
}

int clsph_sync(clsph_ctx* ctx) {
  if (!ctx) return fail(nullptr, CLSPH_EINVAL, "clsph_sync: NULL context");
  HIPCK(ctx, hipStreamSynchronize(ctx->stream));
  return CLSPH_OK;
}

int clsph_state_alloc(clsph_ctx* ctx, uint32_t n_cap, uint32_t table_cap, clsph_state** out) {
  return alloc_state(ctx, 1, &n_cap, false, table_cap, out);
}

int clsph_batch_alloc(clsph_ctx* ctx, uint32_t n_scenes, const uint32_t* counts,
                      uint32_t table_cap, clsph_state** out) {
  return alloc_state(ctx, n_scenes, counts, true, table_cap, out);
}

int clsph_state_free(clsph_state* st) {
  if (!st) return CLSPH_OK;
  (void)hipSetDevice(st->ctx->device);
  (void)hipStreamSynchronize(st->ctx->stream);
  free_state_buffers(st);
  delete st;
  return CLSPH_OK;
}

int clsph_set_params(clsph_state* st, const clsph_simulation_parameters* params,
                     const clsph_precomputed_kernel_values* terms, uint32_t max_bounces) {
  return set_params(st, 0, params, terms, max_bounces);
}

int clsph_batch_set_params(clsph_state* st, uint32_t scene,
                           const clsph_simulation_parameters* params,
                           const clsph_precomputed_kernel_values* terms, uint32_t max_bounces) {
  return set_params(st, scene, params, terms, max_bounces);
}

int clsph_upload_aos(clsph_state* st, const clsph_particle* p, uint32_t n) {
  return upload(st, 0, p, n);
}

int clsph_batch_upload(clsph_state* st, uint32_t scene, const clsph_particle* p, uint32_t n) {
  return upload(st, scene, p, n);
}

int clsph_download_aos(clsph_state* st, clsph_particle* p, uint32_t n) {
  return download(st, 0, p, n, false);
}

int clsph_batch_download(clsph_state* st, uint32_t scene, clsph_particle* p, uint32_t n) {
  return download(st, scene, p, n, false);
}

int clsph_download_sorted_aos(clsph_state* st, clsph_particle* p, uint32_t n) {
  return download(st, 0, p, n, true);
}

int clsph_scene_upload(clsph_state* st, const float* face_normals, const float* vertices,
                       uint32_t vertex_count, const uint32_t* indices, uint32_t face_count) {
  return scene_upload(st, 0, face_normals, vertices, vertex_count, indices, face_count);
}

int clsph_batch_scene_upload(clsph_state* st, uint32_t scene, const float* face_normals,
                             const float* vertices, uint32_t vertex_count,
                             const uint32_t* indices, uint32_t face_count) {
  return scene_upload(st, scene, face_normals, vertices, vertex_count, indices, face_count);
}

int clsph_scene_grid(clsph_state* st, float cell_size) {
  if (!st) return CLSPH_EINVAL;
  st->meshes[0].grid_cell = cell_size;
  return rebuild_meshes(st);
}

int clsph_set_block_sizes(clsph_state* st, const uint32_t sizes[7]) {
  if (!st || !sizes) return CLSPH_EINVAL;
  int idx[kNumStages];
  for (int s = 0; s < kNumStages; ++s) {
    idx[s] = st->stage_bs[s];
    if (sizes[s] == 0) continue;
    int found = -1;
    for (int b = 0; b < kNumBlockSizes; ++b)
      if ((uint32_t)kBlockSizes[b] == sizes[s]) found = b;
    if (found < 0) return fail(st->ctx, CLSPH_EINVAL, "block size must be 64, 128, 256 or 512");
    idx[s] = found;
  }
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  bool advect_changed = idx[kStageAdvect] != st->stage_bs[kStageAdvect];
  for (int s = 0; s < kNumStages; ++s) st->stage_bs[s] = idx[s];
  st->graph_dirty = st->lgraph_dirty = true;
  if (advect_changed) {
    // the bounding-box partials are laid out per advection work-group:
    // rebuild the scene ranges and recompute them from the current state
    std::vector<DevParams> d;
    int rc = read_dps(st, d);
    if (!rc) rc = build_blocks(st);
    if (rc) return rc;
    for (uint32_t s = 0; s < st->n_scenes; ++s) {
      d[s].block_begin = st->host_dps[s].block_begin;
      d[s].block_end = st->host_dps[s].block_end;
    }
    HIPCK(st->ctx, hipMemcpy(st->dps, d.data(), sizeof(DevParams) * st->n_scenes,
                             hipMemcpyHostToDevice));
    launch_aabb(st);
    rc = check_launch(st->ctx);
    if (rc) return rc;
    HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  }
  return CLSPH_OK;
}

int clsph_get_block_sizes(const clsph_state* st, uint32_t sizes[7]) {
  if (!st || !sizes) return CLSPH_EINVAL;
  for (int s = 0; s < kNumStages; ++s) sizes[s] = (uint32_t)st->bs_of(s);
  return CLSPH_OK;
}

int clsph_grid(clsph_state* st) { return st ? run_stage(st, kStageGrid) : CLSPH_EINVAL; }
int clsph_hash(clsph_state* st) { return st ? run_stage(st, kStageHash) : CLSPH_EINVAL; }
int clsph_sort(clsph_state* st) { return st ? run_stage(st, kStageSort) : CLSPH_EINVAL; }
int clsph_cell_table(clsph_state* st) { return st ? run_stage(st, kStageTable) : CLSPH_EINVAL; }
int clsph_density_pressure(clsph_state* st) {
  return st ? run_stage(st, kStageDensity) : CLSPH_EINVAL;
}
int clsph_forces(clsph_state* st) { return st ? run_stage(st, kStageForces) : CLSPH_EINVAL; }
int clsph_advect_collide(clsph_state* st) {
  return st ? run_stage(st, kStageAdvect) : CLSPH_EINVAL;
}

int clsph_step(clsph_state* st, uint32_t nsteps) {
  if (!st) return CLSPH_EINVAL;
  if (st->n == 0 || nsteps == 0) return CLSPH_OK;
  int rc = ensure_graph(st);
  if (rc) return rc;
  for (uint32_t k = 0; k < nsteps; ++k) HIPCK(st->ctx, hipGraphLaunch(st->graph, st->ctx->stream));
  st->acc_zero = true;
  st->fields_canonical = false;
  st->stage = kStageAdvect;
  return CLSPH_OK;
}

int clsph_batch_get_grid(clsph_state* st, uint32_t scene, clsph_simulation_parameters* params) {
  if (!st || !params) return CLSPH_EINVAL;
  if (scene >= st->n_scenes) return fail(st->ctx, CLSPH_EINVAL, "scene index out of range");
  std::vector<DevParams> d;
  int rc = read_dps(st, d);
  if (rc) return rc;
  *params = d[scene].p;
  return CLSPH_OK;
}

int clsph_get_grid(clsph_state* st, clsph_simulation_parameters* params) {
  return clsph_batch_get_grid(st, 0, params);
}

int clsph_neighbour_pairs(clsph_state* st, uint64_t* pairs) {
  if (!st || !pairs) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (st->stage < kStageDensity && st->stage != kStageAdvect)
    return fail(st->ctx, CLSPH_EINVAL, "no density stage has run on this state");
  std::vector<uint32_t> cnt(st->n);
  int rc = read_u32(st, st->nbr_count, cnt.data(), st->n);
  if (rc) return rc;
  uint64_t e = 0;
  for (uint32_t c : cnt) e += c;
  *pairs = e;
  return CLSPH_OK;
}

int clsph_host_alloc(clsph_ctx* ctx, size_t bytes, void** out) {
  if (!ctx || !out || bytes == 0) return fail(ctx, CLSPH_EINVAL, "clsph_host_alloc: bad argument");
  *out = nullptr;
  HIPCK(ctx, hipSetDevice(ctx->device));
  HIPCK(ctx, hipHostMalloc(out, bytes, hipHostMallocDefault));
  return CLSPH_OK;
}

int clsph_host_free(clsph_ctx* ctx, void* p) {
  if (!p) return CLSPH_OK;
  HIPCK(ctx, hipHostFree(p));
  return CLSPH_OK;
}

static constexpr size_t kMetaBytes = sizeof(DevParams) + sizeof(BatchParams);

int clsph_download_begin(clsph_state* st, uint32_t slot, clsph_particle* out, uint32_t n) {
  if (!st || !out || slot > 1) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL,
                                          "clsph_download_begin: bad argument");
  if (st->n_scenes != 1 || n != st->n || n == 0)
    return fail(st->ctx, CLSPH_EINVAL,
                "asynchronous downloads take the whole particle count of a single-scene state");
  clsph_ctx* c = st->ctx;
  if (!st->copy_stream) {
    // a second stream at the default priority, so that copies overlap kernels
    HIPCK(c, hipStreamCreateWithFlags(&st->copy_stream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
      HIPCK(c, hipEventCreateWithFlags(&st->ev_pack[k], hipEventDisableTiming));
      HIPCK(c, hipEventCreateWithFlags(&st->ev_done[k], hipEventDisableTiming));
    }
  }
  if (!st->stage_aos[slot]) {
    int rc = dalloc(c, &st->stage_aos[slot], st->n_cap);
    if (!rc) rc = dalloc(c, &st->stage_meta[slot], kMetaBytes);
    if (rc) return rc;
    HIPCK(c, hipHostMalloc((void**)&st->host_meta[slot], kMetaBytes, hipHostMallocDefault));
  }
  hipStream_t s = c->stream;
  // the staging buffer is free once the copy of its previous download ended
  if (st->slot_used[slot]) HIPCK(c, hipStreamWaitEvent(s, st->ev_done[slot], 0));
  launch_pack_aos(st, st->stage_aos[slot]);
  int rc = check_launch(c);
  if (rc) return rc;
  HIPCK(c, hipMemcpyAsync(st->stage_meta[slot], st->dps, sizeof(DevParams),
                          hipMemcpyDeviceToDevice, s));
  HIPCK(c, hipMemcpyAsync(st->stage_meta[slot] + sizeof(DevParams), st->bp, sizeof(BatchParams),
                          hipMemcpyDeviceToDevice, s));
  HIPCK(c, hipEventRecord(st->ev_pack[slot], s));
  HIPCK(c, hipStreamWaitEvent(st->copy_stream, st->ev_pack[slot], 0));
  HIPCK(c, hipMemcpyAsync(out, st->stage_aos[slot], sizeof(clsph_particle) * (size_t)n,
                          hipMemcpyDeviceToHost, st->copy_stream));
  HIPCK(c, hipMemcpyAsync(st->host_meta[slot], st->stage_meta[slot], kMetaBytes,
                          hipMemcpyDeviceToHost, st->copy_stream));
  HIPCK(c, hipEventRecord(st->ev_done[slot], st->copy_stream));
  st->slot_used[slot] = true;
  return CLSPH_OK;
}

int clsph_download_wait(clsph_state* st, uint32_t slot, clsph_simulation_parameters* grid) {
  if (!st || slot > 1) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "clsph_download_wait: bad argument");
  if (!st->slot_used[slot])
    return fail(st->ctx, CLSPH_EINVAL, "no asynchronous download was started in this slot");
  HIPCK(st->ctx, hipEventSynchronize(st->ev_done[slot]));
  DevParams d;
  BatchParams b;
  memcpy(&d, st->host_meta[slot], sizeof(DevParams));
  memcpy(&b, st->host_meta[slot] + sizeof(DevParams), sizeof(BatchParams));
  if (grid) *grid = d.p;
  const uint32_t err = b.err | d.err;
  if (err & kErrGridDim)
    return fail(st->ctx, CLSPH_EGRID,
                "grid dimension >= 1024 cells on an axis (particles spread too far)");
  if (err & (kErrTableCap | kErrFineKey))
    return fail(st->ctx, CLSPH_ETABLE, "grid cells exceed the cell table capacity");
  return CLSPH_OK;
}

int clsph_read_order(clsph_state* st, uint32_t* slot, uint32_t* perm, uint32_t n) {
  if (!st) return CLSPH_EINVAL;
  if (n != st->scene_n[0]) return fail(st->ctx, CLSPH_EINVAL, "n must be the particle count");
  uint32_t b0 = st->scene_begin[0];
  int rc = CLSPH_OK;
  if (slot) rc = read_u32(st, st->slot + b0, slot, n);
  if (!rc && perm) rc = read_u32(st, st->val_b + b0, perm, n);
  for (uint32_t i = 0; !rc && i < n; ++i) {
    if (slot) slot[i] -= b0;
    if (perm) perm[i] -= b0;
  }
  return rc;
}

int clsph_read_keys(clsph_state* st, uint32_t* keys, uint32_t n) {
  if (!st) return CLSPH_EINVAL;
  if (n > st->scene_n[0]) return fail(st->ctx, CLSPH_EINVAL, "n exceeds the particle count");
  std::vector<DevParams> d;
  int rc = read_dps(st, d);
  if (!rc) rc = read_u32(st, st->key_a + st->scene_begin[0], keys, n);
  for (uint32_t i = 0; !rc && i < n; ++i) keys[i] = (keys[i] - d[0].key_base) >> 3;
  return rc;
}

int clsph_read_sorted(clsph_state* st, uint32_t* keys, uint32_t* perm, uint32_t n) {
  if (!st) return CLSPH_EINVAL;
  if (n != st->scene_n[0]) return fail(st->ctx, CLSPH_EINVAL, "n must be the particle count");
  std::vector<uint32_t> k(n), v(n), slot(n);
  std::vector<DevParams> d;
  int rc = read_dps(st, d);
  uint32_t b0 = st->scene_begin[0];
  if (!rc) rc = read_u32(st, st->key_b + b0, k.data(), n);
  if (!rc) rc = read_u32(st, st->val_b + b0, v.data(), n);
  if (!rc) rc = read_u32(st, st->slot + b0, slot.data(), n);
  if (rc) return rc;
  // Half-cell order -> the reference's order (stable by cell id).
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t o = slot[i] - b0;
    if (o >= n) return fail(st->ctx, CLSPH_EDEVICE, "canonical slots are not a permutation");
    if (keys) keys[o] = (k[i] - d[0].key_base) >> 3;
    if (perm) perm[o] = v[i] - b0;
  }
  return CLSPH_OK;
}

int clsph_read_cell_table(clsph_state* st, uint32_t* table, uint32_t count) {
  if (!st) return CLSPH_EINVAL;
  if (count == 0) return device_errors(st);
  if (!table) return fail(st->ctx, CLSPH_EINVAL, "NULL output");
  std::vector<DevParams> d;
  int rc = read_dps(st, d);
  if (rc) return rc;
  if ((uint64_t)d[0].key_base + 8ull * count > st->table_cap)
    return fail(st->ctx, CLSPH_EINVAL, "count exceeds the table");
  std::vector<uint32_t> ht(8ull * count);
  rc = read_u32(st, st->table + d[0].key_base, ht.data(), ht.size());
  if (rc) return rc;
  for (uint32_t c = 0; c < count; ++c) table[c] = ht[8ull * c] - st->scene_begin[0];
  return device_errors(st);
}

int clsph_batch_bounce_capped(clsph_state* st, uint32_t scene, uint32_t* count) {
  if (!st || !count) return CLSPH_EINVAL;
  if (scene >= st->n_scenes) return fail(st->ctx, CLSPH_EINVAL, "scene index out of range");
  std::vector<DevParams> d;
  int rc = read_dps(st, d);
  if (!rc) *count = d[scene].capped;
  return rc;
}

int clsph_bounce_capped(clsph_state* st, uint32_t* count) {
  return clsph_batch_bounce_capped(st, 0, count);
}

int clsph_time_steps(clsph_state* st, uint32_t nsteps, float* ms) {
  if (!st || !ms) return CLSPH_EINVAL;
  int rc = ensure_graph(st);
  if (rc) return rc;
  hipEvent_t a, b;
  HIPCK(st->ctx, hipEventCreate(&a));
  HIPCK(st->ctx, hipEventCreate(&b));
  HIPCK(st->ctx, hipEventRecord(a, st->ctx->stream));
  for (uint32_t k = 0; k < nsteps; ++k) HIPCK(st->ctx, hipGraphLaunch(st->graph, st->ctx->stream));
  HIPCK(st->ctx, hipEventRecord(b, st->ctx->stream));
  HIPCK(st->ctx, hipEventSynchronize(b));
  HIPCK(st->ctx, hipEventElapsedTime(ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  st->acc_zero = true;
  st->fields_canonical = false;
  st->stage = kStageAdvect;
  return device_errors(st);
}

int clsph_time_stage(clsph_state* st, int stage, uint32_t reps, float* ms) {
  if (!st || !ms || reps == 0) return CLSPH_EINVAL;
  hipEvent_t a, b;
  HIPCK(st->ctx, hipEventCreate(&a));
  HIPCK(st->ctx, hipEventCreate(&b));
  HIPCK(st->ctx, hipEventRecord(a, st->ctx->stream));
  for (uint32_t k = 0; k < reps; ++k) {
    int rc = run_stage(st, stage);
    if (rc) return rc;
  }
  HIPCK(st->ctx, hipEventRecord(b, st->ctx->stream));
  HIPCK(st->ctx, hipEventSynchronize(b));
  float t = 0.f;
  HIPCK(st->ctx, hipEventElapsedTime(&t, a, b));
  *ms = t / (float)reps;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return device_errors(st);
}

int clsph_profile_steps(clsph_state* st, uint32_t nsteps, float* stage_ms) {
  if (!st || !stage_ms) return CLSPH_EINVAL;
  std::vector<hipEvent_t> ev((size_t)nsteps * (kNumStages + 1));
  for (auto& e : ev) HIPCK(st->ctx, hipEventCreate(&e));
  for (uint32_t k = 0; k < nsteps; ++k) {
    hipEvent_t* e = ev.data() + (size_t)k * (kNumStages + 1);
    HIPCK(st->ctx, hipEventRecord(e[0], st->ctx->stream));
    for (int s = 0; s < kNumStages; ++s) {
      int rc = run_stage(st, s);
      if (rc) return rc;
      HIPCK(st->ctx, hipEventRecord(e[s + 1], st->ctx->stream));
    }
  }
  if (!ev.empty()) HIPCK(st->ctx, hipEventSynchronize(ev.back()));
  for (int s = 0; s < kNumStages; ++s) stage_ms[s] = 0.f;
  for (uint32_t k = 0; k < nsteps; ++k) {
    hipEvent_t* e = ev.data() + (size_t)k * (kNumStages + 1);
    for (int s = 0; s < kNumStages; ++s) {
      float t = 0.f;
      HIPCK(st->ctx, hipEventElapsedTime(&t, e[s], e[s + 1]));
      stage_ms[s] += t;
    }
  }
  for (int s = 0; s < kNumStages; ++s) stage_ms[s] /= (float)(nsteps ? nsteps : 1);
  for (auto& e : ev) (void)hipEventDestroy(e);
  return device_errors(st);
}

uint32_t clsph_make_block(clsph_particle* out, uint32_t n, float spacing, const float origin[3],
                          const float extent[3], float jitter, uint64_t seed) {
  if (!out || !(spacing > 0.f)) return 0;
  // x and z hold round(extent / spacing) sites; layers stack up in y until n
  // particles are placed (the top layer may be partial).
  uint32_t dims[3];
  for (int a = 0; a < 3; a += 2) {
    double q = floor((double)extent[a] / (double)spacing + 0.5);
    dims[a] = q < 1.0 ? 1u : (uint32_t)q;
  }
  dims[1] = (uint32_t)((n + dims[0] * dims[2] - 1) / (dims[0] * dims[2]));
  uint32_t count = n;
  for (uint32_t k = 0; k < count; ++k) {
    uint32_t ix = k % dims[0];
    uint32_t iz = (k / dims[0]) % dims[2];
    uint32_t iy = k / (dims[0] * dims[2]);
    uint32_t id[3] = {ix, iy, iz};
    clsph_particle& p = out[k];
    memset(&p, 0, sizeof(p));
    for (int a = 0; a < 3; ++a) {
      uint64_t r = splitmix64(seed ^ (3ull * k + (uint64_t)a));
      double u = (double)(r >> 40) * (1.0 / 16777216.0);  // [0, 1)
      double off = (2.0 * u - 1.0) * (double)jitter * (double)spacing;
      p.position.s[a] =
          (float)((double)origin[a] + ((double)id[a] + 0.5) * (double)spacing + off);
    }
  }
  return count;
}

// Value noise: lattice values in [0, 1) from the counter-based generator,
// blended with the smoothstep of the fractional position.
static double lattice_value(uint64_t seed, uint32_t octave, int64_t ix, int64_t iz) {
  uint64_t key = seed ^ (0x9E3779B97F4A7C15ull * (octave + 1));
  key = splitmix64(key ^ ((uint64_t)(ix & 0xFFFFFFFF) << 32 | (uint64_t)(iz & 0xFFFFFFFF)));
  return (double)(key >> 11) * (1.0 / 9007199254740992.0);
}

static double value_noise(double x, double z, uint32_t octaves, uint64_t seed) {
  double sum = 0.0, amp = 1.0, norm = 0.0, freq = 4.0;
  for (uint32_t o = 0; o < octaves; ++o) {
    double fx = x * freq, fz = z * freq;
    double x0 = std::floor(fx), z0 = std::floor(fz);
    double tx = fx - x0, tz = fz - z0;
    tx = tx * tx * (3.0 - 2.0 * tx);
    tz = tz * tz * (3.0 - 2.0 * tz);
    int64_t ix = (int64_t)x0, iz = (int64_t)z0;
    double a = lattice_value(seed, o, ix, iz), b = lattice_value(seed, o, ix + 1, iz);
    double c = lattice_value(seed, o, ix, iz + 1), d = lattice_value(seed, o, ix + 1, iz + 1);
    double v = (a * (1 - tx) + b * tx) * (1 - tz) + (c * (1 - tx) + d * tx) * tz;
    sum += amp * v;
    norm += amp;
    amp *= 0.5;
    freq *= 2.0;
  }
  return norm > 0 ? sum / norm : 0.0;
}

void clsph_make_heightfield(uint32_t res, float x0, float z0, float size, float amplitude,
                            uint32_t octaves, uint64_t seed, float* vertices, uint32_t* indices,
                            float* face_normals) {
  uint32_t nv = res + 1;
  for (uint32_t j = 0; j < nv; ++j)
    for (uint32_t i = 0; i < nv; ++i) {
      double u = (double)i / res, w = (double)j / res;
      float* v = vertices + 3 * (j * nv + i);
      v[0] = (float)(x0 + u * size);
      v[2] = (float)(z0 + w * size);
      v[1] = (float)(amplitude * value_noise(u, w, octaves, seed));
    }
  uint32_t f = 0;
  for (uint32_t j = 0; j < res; ++j)
    for (uint32_t i = 0; i < res; ++i) {
      uint32_t a = j * nv + i, b = a + 1, c = a + nv, d = c + 1;
      // (a, c, b) and (b, c, d): (B - A) x (C - A) points up (+y)
      uint32_t tri[2][3] = {{a, c, b}, {b, c, d}};
      for (int t = 0; t < 2; ++t, ++f) {
        const float* A = vertices + 3 * tri[t][0];
        const float* B = vertices + 3 * tri[t][1];
        const float* C = vertices + 3 * tri[t][2];
        float ux = B[0] - A[0], uy = B[1] - A[1], uz = B[2] - A[2];
        float vx = C[0] - A[0], vy = C[1] - A[1], vz = C[2] - A[2];
        float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
        float l = sqrtf(nx * nx + ny * ny + nz * nz);
        face_normals[3 * f + 0] = nx / l;
        face_normals[3 * f + 1] = ny / l;
        face_normals[3 * f + 2] = nz / l;
        for (int k = 0; k < 3; ++k) indices[3 * f + k] = tri[t][k];
      }
    }
}

uint32_t clsph_make_column(clsph_particle* out, uint32_t n, float spacing, const float origin[3],
                           const float extent[3], float jitter, float vh, float v0, float v1,
                           uint64_t seed) {
  uint32_t got = clsph_make_block(out, n, spacing, origin, extent, jitter, seed);
  for (uint32_t k = 0; k < got; ++k) {
    double u[3];
    for (int a = 0; a < 3; ++a) {
      uint64_t r = splitmix64((seed * 0x2545F4914F6CDD1Dull) ^ (3ull * k + (uint64_t)a + 1));
      u[a] = (double)(r >> 40) * (1.0 / 16777216.0);
    }
    float vel[3] = {(float)((2 * u[0] - 1) * vh), (float)(-(v0 + u[1] * (v1 - v0))),
                    (float)((2 * u[2] - 1) * vh)};
    for (int a = 0; a < 3; ++a) {
      out[k].velocity.s[a] = vel[a];
      out[k].intermediate_velocity.s[a] = vel[a];
    }
  }
  return got;
}

void clsph_kernel_terms(float h, clsph_precomputed_kernel_values* out) {
  if (!out) return;
  // sph_simulation.cpp:499-505 writes pow(h, 9.f) unqualified after
  // <cmath>; under g++/libstdc++ that names ::pow(double, double), so the
  // powers and every quotient are double, rounded once to float on store.
  const double pi = 3.14159265358979323846;
  const double hd = (double)h;
  const double h9 = ::pow(hd, 9.0), h6 = ::pow(hd, 6.0);
  out->poly_6 = (float)(315.0 / (64.0 * pi * h9));
  out->poly_6_gradient = (float)(-945.0 / (32.0 * pi * h9));
  out->poly_6_laplacian = (float)(-945.0 / (32.0 * pi * h9));
  out->spiky = (float)(-45.0 / (pi * h6));
  out->viscosity = (float)(45.0 / (pi * h6));
}

void clsph_make_box(const float lo[3], const float hi[3], float* vertices, uint32_t* indices,
                    float* face_normals) {
  for (uint32_t v = 0; v < 8; ++v) {
    vertices[3 * v + 0] = (v & 1) ? hi[0] : lo[0];
    vertices[3 * v + 1] = (v & 2) ? hi[1] : lo[1];
    vertices[3 * v + 2] = (v & 4) ? hi[2] : lo[2];
  }
  // Each face as a quad (a, b, c, d) split into (a, b, c) and (a, c, d).
  static const uint32_t quads[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1},
                                       {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
  float cx = 0.5f * (lo[0] + hi[0]), cy = 0.5f * (lo[1] + hi[1]), cz = 0.5f * (lo[2] + hi[2]);
  for (int q = 0; q < 6; ++q) {
    uint32_t tri[2][3] = {{quads[q][0], quads[q][1], quads[q][2]},
                          {quads[q][0], quads[q][2], quads[q][3]}};
    for (int t = 0; t < 2; ++t) {
      uint32_t f = 2 * q + t;
      const float* A = vertices + 3 * tri[t][0];
      const float* B = vertices + 3 * tri[t][1];
      const float* C = vertices + 3 * tri[t][2];
      // Face normal as scene.cpp:36-64 computes it: (B - A) x (C - A), unit.
      float ux = B[0] - A[0], uy = B[1] - A[1], uz = B[2] - A[2];
      float vx = C[0] - A[0], vy = C[1] - A[1], vz = C[2] - A[2];
      float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
      float fx = (A[0] + B[0] + C[0]) / 3.f, fy = (A[1] + B[1] + C[1]) / 3.f,
            fz = (A[2] + B[2] + C[2]) / 3.f;
      if (nx * (cx - fx) + ny * (cy - fy) + nz * (cz - fz) < 0.f) {
        uint32_t tmp = tri[t][1];  // wind the other way: normal points inward
        tri[t][1] = tri[t][2];
        tri[t][2] = tmp;
        nx = -nx, ny = -ny, nz = -nz;
      }
      float l = sqrtf(nx * nx + ny * ny + nz * nz);
      face_normals[3 * f + 0] = nx / l;
      face_normals[3 * f + 1] = ny / l;
      face_normals[3 * f + 2] = nz / l;
      for (int k = 0; k < 3; ++k) indices[3 * f + k] = tri[t][k];
    }
  }
}

}  // extern "C"
