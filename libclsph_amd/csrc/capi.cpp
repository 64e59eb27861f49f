// C ABI of the clsph HIP pipeline (include/clsph_hip.h).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/clsph_hip.h"
#include "state.h"

using namespace clsph;

namespace {

thread_local std::string g_err;  // errors raised before a context exists

int fail(clsph_ctx* ctx, int code, const std::string& msg) {
  if (ctx)
    ctx->last_error = msg;
  else
    g_err = msg;
  return code;
}

#define HIPCK(ctx, expr)                                                        \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      return fail((ctx), CLSPH_EDEVICE,                                         \
                  std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

template <typename T>
int dalloc(clsph_ctx* ctx, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(ctx, CLSPH_ENOMEM,
                std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) +
                    " bytes: " + hipGetErrorString(e));
  }
  return CLSPH_OK;
}

int check_launch(clsph_ctx* ctx) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(ctx, CLSPH_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return CLSPH_OK;
}

int device_errors(clsph_state* st) {
  uint32_t err = 0;
  HIPCK(st->ctx, hipMemcpyAsync(&err, &st->dp->err, sizeof(err), hipMemcpyDeviceToHost,
                                st->ctx->stream));
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  if (err & kErrGridDim)
    return fail(st->ctx, CLSPH_EGRID,
                "grid dimension >= 1024 cells on an axis (particles spread too far)");
  if (err & kErrTableCap)
    return fail(st->ctx, CLSPH_ETABLE, "grid_cell_count exceeds the cell table capacity");
  return CLSPH_OK;
}

void free_state_buffers(clsph_state* st) {
  void* bufs[] = {st->tgrid_start, st->tgrid_tris,
                  st->pos,  st->vel,   st->vh,    st->pos_s, st->vel_s, st->vh_s,
                  st->acc_s, st->rho_s, st->prs_s, st->fa, st->fb, st->key_a, st->val_a,
                  st->key_b, st->val_b, st->hist,  st->digit_base, st->table,
                  st->aabb_partial, st->face_normals, st->vertices, st->indices, st->dp,
                  st->nbr, st->nbr_count, st->slot};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (st->graph) (void)hipGraphExecDestroy(st->graph);
}

// Launches the kernels of one whole substep on the context stream.
void enqueue_step(clsph_state* st) {
  launch_grid_params(st);
  launch_hash(st);
  launch_sort(st);
  launch_permute(st);
  launch_cell_table(st);
  launch_density(st);
  launch_forces(st);
  launch_advect(st);
}

int run_stage(clsph_state* st, int stage) {
  switch (stage) {
    case 0:
      launch_grid_params(st);
      break;
    case 1:
      launch_hash(st);
      break;
    case 2:
      // sorting also builds the tables and the canonical slots, so that a
      // slot is always current once particles have been reordered
      launch_sort(st);
      launch_permute(st);
      launch_cell_table(st);
      st->fields_canonical = false;
      break;
    case 3:
      launch_cell_table(st);
      break;
    case 4:
      launch_density(st);
      break;
    case 5:
      launch_forces(st);
      break;
    case 6:
      launch_advect(st);
      st->acc_zero = true;
      break;
    default:
      return fail(st->ctx, CLSPH_EINVAL, "unknown stage " + std::to_string(stage));
  }
  st->stage = stage;
  return check_launch(st->ctx);
}

// Uniform grid over the mesh for collision queries: cell c lists every
// triangle whose bounding box, padded by a rounding margin, meets it. Built
// on the host once per mesh (the mesh is static).
int build_tri_grid(clsph_state* st, const float* verts, const uint32_t* idx, uint32_t faces,
                   float cell_size) {
  clsph_ctx* ctx = st->ctx;
  if (st->tgrid_start) (void)hipFree(st->tgrid_start);
  if (st->tgrid_tris) (void)hipFree(st->tgrid_tris);
  st->tgrid_start = nullptr;
  st->tgrid_tris = nullptr;
  st->tgrid.enabled = 0;
  if (faces == 0 || cell_size < 0.f) return CLSPH_OK;
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  double ext_sum = 0.0;
  for (uint32_t f = 0; f < faces; ++f) {
    float tlo[3] = {1e30f, 1e30f, 1e30f}, thi[3] = {-1e30f, -1e30f, -1e30f};
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) {
        float v = verts[3 * idx[3 * f + k] + a];
        tlo[a] = std::min(tlo[a], v);
        thi[a] = std::max(thi[a], v);
      }
    float e = 0.f;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], tlo[a]);
      hi[a] = std::max(hi[a], thi[a]);
      e = std::max(e, thi[a] - tlo[a]);
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
      float ext = (hi[a] - lo[a]) + 2.f * (pad + cell_size);
      dims[a] = (uint32_t)std::max(1.0, std::ceil((double)ext / (double)cell_size));
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
  std::vector<uint32_t> count(ncells + 1, 0);
  std::vector<uint32_t> range(6ull * faces);
  for (uint32_t f = 0; f < faces; ++f) {
    float tlo[3] = {1e30f, 1e30f, 1e30f}, thi[3] = {-1e30f, -1e30f, -1e30f};
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) {
        float v = verts[3 * idx[3 * f + k] + a];
        tlo[a] = std::min(tlo[a], v);
        thi[a] = std::max(thi[a], v);
      }
    for (int a = 0; a < 3; ++a) {
      range[6 * f + a] = cell_of(tlo[a] - pad, a);
      range[6 * f + 3 + a] = cell_of(thi[a] + pad, a);
    }
    for (uint32_t z = range[6 * f + 2]; z <= range[6 * f + 5]; ++z)
      for (uint32_t y = range[6 * f + 1]; y <= range[6 * f + 4]; ++y)
        for (uint32_t x = range[6 * f + 0]; x <= range[6 * f + 3]; ++x)
          count[((uint64_t)z * dims[1] + y) * dims[0] + x]++;
  }
  std::vector<uint32_t> start(ncells + 1, 0);
  for (uint64_t c = 0; c < ncells; ++c) start[c + 1] = start[c] + count[c];
  std::vector<uint32_t> tris(std::max<uint32_t>(start[ncells], 1));
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  for (uint32_t f = 0; f < faces; ++f)
    for (uint32_t z = range[6 * f + 2]; z <= range[6 * f + 5]; ++z)
      for (uint32_t y = range[6 * f + 1]; y <= range[6 * f + 4]; ++y)
        for (uint32_t x = range[6 * f + 0]; x <= range[6 * f + 3]; ++x)
          tris[fill[((uint64_t)z * dims[1] + y) * dims[0] + x]++] = f;
  int rc = dalloc(ctx, &st->tgrid_start, ncells + 1);
  if (!rc) rc = dalloc(ctx, &st->tgrid_tris, tris.size());
  if (rc) return rc;
  HIPCK(ctx, hipMemcpy(st->tgrid_start, start.data(), 4 * (ncells + 1), hipMemcpyHostToDevice));
  HIPCK(ctx, hipMemcpy(st->tgrid_tris, tris.data(), 4 * tris.size(), hipMemcpyHostToDevice));
  for (int a = 0; a < 3; ++a) {
    st->tgrid.origin[a] = origin[a];
    st->tgrid.dims[a] = dims[a];
  }
  st->tgrid.inv_cell = inv;
  st->tgrid.enabled = 1;
  st->graph_dirty = true;
  return CLSPH_OK;
}

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

}  // namespace

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
  if (!ctx || !out) return fail(ctx, CLSPH_EINVAL, "clsph_state_alloc: NULL argument");
  *out = nullptr;
  if (table_cap == 0) {
    uint64_t c = 8ull * n_cap;
    table_cap = (uint32_t)(c < (1u << 20) ? (1u << 20) : (c > (1u << 28) ? (1u << 28) : c));
  }
  if (table_cap > (1u << 29))
    return fail(ctx, CLSPH_EINVAL, "table_cap above 2^29 cells is not supported");
  HIPCK(ctx, hipSetDevice(ctx->device));
  clsph_state* st = new clsph_state();
  st->ctx = ctx;
  st->n_cap = n_cap;
  st->table_cap = table_cap;
  st->n_tiles_cap = sort_tiles_for(n_cap ? n_cap : 1);
  uint32_t nb = (n_cap + 255) / 256 + 1;
  int rc = CLSPH_OK;
  float4** f4[] = {&st->pos, &st->vel, &st->vh, &st->pos_s, &st->vel_s, &st->vh_s, &st->acc_s};
  for (auto p : f4)
    if (rc == CLSPH_OK) rc = dalloc(ctx, p, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->rho_s, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->prs_s, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->fa, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->fb, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->nbr, (size_t)(kMaxNeighbours / 4) * n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->nbr_count, n_cap);
  uint32_t** u4[] = {&st->key_a, &st->val_a, &st->key_b, &st->val_b};
  for (auto p : u4)
    if (rc == CLSPH_OK) rc = dalloc(ctx, p, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->hist, (size_t)kMaxBins * st->n_tiles_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->digit_base, 2 * (size_t)kMaxBins);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->table, 8ull * table_cap + 1);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->slot, n_cap);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->aabb_partial, (size_t)nb * 6);
  if (rc == CLSPH_OK) rc = dalloc(ctx, &st->dp, 1);
  if (rc != CLSPH_OK) {
    free_state_buffers(st);
    delete st;
    return rc;
  }
  memset(&st->host_dp, 0, sizeof(st->host_dp));
  st->host_dp.table_cap = table_cap;
  st->host_dp.max_bounces = 64;
  HIPCK(ctx, hipMemcpy(st->dp, &st->host_dp, sizeof(DevParams), hipMemcpyHostToDevice));
  *out = st;
  return CLSPH_OK;
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
  if (!st || !params || !terms) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (!(params->h > 0.f)) return fail(st->ctx, CLSPH_EINVAL, "h must be positive");
  st->host_dp.p = *params;
  st->host_dp.t = *terms;
  st->host_dp.n = st->n;
  st->host_dp.table_cap = st->table_cap;
  st->host_dp.max_bounces = max_bounces ? max_bounces : 64;
  st->host_dp.err = 0;
  st->host_dp.capped = 0;
  HIPCK(st->ctx, hipMemcpyAsync(st->dp, &st->host_dp, sizeof(DevParams), hipMemcpyHostToDevice,
                                st->ctx->stream));
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  return CLSPH_OK;
}

int clsph_upload_aos(clsph_state* st, const clsph_particle* p, uint32_t n) {
  if (!st || (!p && n)) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (n > st->n_cap) return fail(st->ctx, CLSPH_EINVAL, "more particles than the state holds");
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
  hipStream_t s = st->ctx->stream;
  HIPCK(st->ctx, hipMemcpyAsync(st->pos, pos.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s));
  HIPCK(st->ctx, hipMemcpyAsync(st->vel, vel.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s));
  HIPCK(st->ctx, hipMemcpyAsync(st->vh, vh.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s));
  HIPCK(st->ctx, hipMemcpyAsync(st->acc_s, acc.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s));
  HIPCK(st->ctx, hipMemcpyAsync(st->rho_s, rho.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
  HIPCK(st->ctx, hipMemcpyAsync(st->prs_s, prs.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
  HIPCK(st->ctx, hipMemcpyAsync(st->key_b, key.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
  if (n != st->n) {
    st->n = n;
    st->graph_dirty = true;
  }
  st->host_dp.n = n;
  HIPCK(st->ctx, hipMemcpyAsync(&st->dp->n, &st->host_dp.n, sizeof(uint32_t), hipMemcpyHostToDevice, s));
  launch_aabb(st);
  st->acc_zero = false;
  st->fields_canonical = true;
  st->stage = -1;
  int rc = check_launch(st->ctx);
  if (rc) return rc;
  HIPCK(st->ctx, hipStreamSynchronize(s));
  return CLSPH_OK;
}

static int download_common(clsph_state* st, clsph_particle* out, uint32_t n, bool sorted) {
  if (!st || (!out && n)) return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (n > st->n) return fail(st->ctx, CLSPH_EINVAL, "asked for more particles than are held");
  if (n == 0) return device_errors(st);
  if (n != st->n && !st->fields_canonical)
    return fail(st->ctx, CLSPH_EINVAL, "partial downloads need the whole particle count");
  std::vector<float4> pos(n), vel(n), vh(n), acc(n);
  std::vector<float> rho(n), prs(n);
  std::vector<uint32_t> key(n), slot(n);
  hipStream_t s = st->ctx->stream;
  const float4* P = sorted ? st->pos_s : st->pos;
  const float4* V = sorted ? st->vel_s : st->vel;
  const float4* H = sorted ? st->vh_s : st->vh;
  HIPCK(st->ctx, hipMemcpyAsync(pos.data(), P, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
  HIPCK(st->ctx, hipMemcpyAsync(vel.data(), V, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
  HIPCK(st->ctx, hipMemcpyAsync(vh.data(), H, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
  bool want_acc = sorted || !st->acc_zero;
  if (want_acc)
    HIPCK(st->ctx, hipMemcpyAsync(acc.data(), st->acc_s, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
  HIPCK(st->ctx, hipMemcpyAsync(rho.data(), st->rho_s, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  HIPCK(st->ctx, hipMemcpyAsync(prs.data(), st->prs_s, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  HIPCK(st->ctx, hipMemcpyAsync(key.data(), st->key_b, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  bool reorder = !st->fields_canonical;
  if (reorder)
    HIPCK(st->ctx, hipMemcpyAsync(slot.data(), st->slot, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  HIPCK(st->ctx, hipStreamSynchronize(s));
  // Fields computed in half-cell order go to each particle's canonical slot;
  // the sorted-order arrays themselves are also in half-cell order.
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t o = reorder ? slot[k] : k;
    if (o >= n) return fail(st->ctx, CLSPH_EDEVICE, "canonical slots are not a permutation");
    clsph_particle& q = out[o];
    memset(&q, 0, sizeof(q));
    q.density = rho[k];
    q.pressure = prs[k];
    q.grid_index = reorder ? key[k] >> 3 : key[k];
    if (want_acc && (sorted || !reorder)) {
      q.acceleration.s[0] = acc[k].x, q.acceleration.s[1] = acc[k].y, q.acceleration.s[2] = acc[k].z;
    }
  }
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t o = (sorted && reorder) ? slot[k] : k;
    clsph_particle& q = out[o];
    q.position.s[0] = pos[k].x, q.position.s[1] = pos[k].y, q.position.s[2] = pos[k].z;
    q.velocity.s[0] = vel[k].x, q.velocity.s[1] = vel[k].y, q.velocity.s[2] = vel[k].z;
    q.intermediate_velocity.s[0] = vh[k].x, q.intermediate_velocity.s[1] = vh[k].y,
    q.intermediate_velocity.s[2] = vh[k].z;
  }
  return device_errors(st);
}

int clsph_download_aos(clsph_state* st, clsph_particle* p, uint32_t n) {
  return download_common(st, p, n, false);
}

int clsph_download_sorted_aos(clsph_state* st, clsph_particle* p, uint32_t n) {
  return download_common(st, p, n, true);
}

int clsph_scene_upload(clsph_state* st, const float* face_normals, const float* vertices,
                       uint32_t vertex_count, const uint32_t* indices, uint32_t face_count) {
  if (!st) return fail(nullptr, CLSPH_EINVAL, "NULL state");
  if (face_count && (!face_normals || !vertices || !indices))
    return fail(st->ctx, CLSPH_EINVAL, "NULL mesh array");
  for (uint64_t k = 0; k < 3ull * face_count; ++k)
    if (indices[k] >= vertex_count)
      return fail(st->ctx, CLSPH_EINVAL, "triangle index out of range of the vertex array");
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  if (st->face_normals) (void)hipFree(st->face_normals);
  if (st->vertices) (void)hipFree(st->vertices);
  if (st->indices) (void)hipFree(st->indices);
  st->face_normals = nullptr;
  st->vertices = nullptr;
  st->indices = nullptr;
  int rc = dalloc(st->ctx, &st->face_normals, 3ull * face_count);
  if (!rc) rc = dalloc(st->ctx, &st->vertices, 3ull * vertex_count);
  if (!rc) rc = dalloc(st->ctx, &st->indices, 3ull * face_count);
  if (rc) return rc;
  if (face_count) {
    HIPCK(st->ctx, hipMemcpy(st->face_normals, face_normals, 12ull * face_count, hipMemcpyHostToDevice));
    HIPCK(st->ctx, hipMemcpy(st->indices, indices, 12ull * face_count, hipMemcpyHostToDevice));
  }
  if (vertex_count)
    HIPCK(st->ctx, hipMemcpy(st->vertices, vertices, 12ull * vertex_count, hipMemcpyHostToDevice));
  st->face_count = face_count;
  st->vert_floats = 3 * vertex_count;
  st->graph_dirty = true;
  return build_tri_grid(st, vertices, indices, face_count, face_count > 64 ? 0.f : -1.f);
}

int clsph_scene_grid(clsph_state* st, float cell_size) {
  if (!st) return CLSPH_EINVAL;
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  std::vector<float> verts(st->vert_floats);
  std::vector<uint32_t> idx(3ull * st->face_count);
  if (st->vert_floats)
    HIPCK(st->ctx, hipMemcpy(verts.data(), st->vertices, 4ull * st->vert_floats, hipMemcpyDeviceToHost));
  if (st->face_count)
    HIPCK(st->ctx, hipMemcpy(idx.data(), st->indices, 12ull * st->face_count, hipMemcpyDeviceToHost));
  st->graph_dirty = true;
  return build_tri_grid(st, verts.data(), idx.data(), st->face_count, cell_size);
}

int clsph_grid(clsph_state* st) { return st ? run_stage(st, 0) : CLSPH_EINVAL; }
int clsph_hash(clsph_state* st) { return st ? run_stage(st, 1) : CLSPH_EINVAL; }
int clsph_sort(clsph_state* st) { return st ? run_stage(st, 2) : CLSPH_EINVAL; }
int clsph_cell_table(clsph_state* st) { return st ? run_stage(st, 3) : CLSPH_EINVAL; }
int clsph_density_pressure(clsph_state* st) { return st ? run_stage(st, 4) : CLSPH_EINVAL; }
int clsph_forces(clsph_state* st) { return st ? run_stage(st, 5) : CLSPH_EINVAL; }
int clsph_advect_collide(clsph_state* st) { return st ? run_stage(st, 6) : CLSPH_EINVAL; }

static int ensure_graph(clsph_state* st) {
  if (!st->graph_dirty && st->graph) return CLSPH_OK;
  if (st->graph) {
    (void)hipGraphExecDestroy(st->graph);
    st->graph = nullptr;
  }
  hipStream_t s = st->ctx->stream;
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

int clsph_step(clsph_state* st, uint32_t nsteps) {
  if (!st) return CLSPH_EINVAL;
  if (st->n == 0 || nsteps == 0) return CLSPH_OK;
  int rc = ensure_graph(st);
  if (rc) return rc;
  for (uint32_t k = 0; k < nsteps; ++k) HIPCK(st->ctx, hipGraphLaunch(st->graph, st->ctx->stream));
  st->acc_zero = true;
  st->fields_canonical = false;
  st->stage = 6;
  return CLSPH_OK;
}

int clsph_get_grid(clsph_state* st, clsph_simulation_parameters* params) {
  if (!st || !params) return CLSPH_EINVAL;
  DevParams d;
  HIPCK(st->ctx, hipMemcpyAsync(&d, st->dp, sizeof(d), hipMemcpyDeviceToHost, st->ctx->stream));
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  *params = d.p;
  return CLSPH_OK;
}

static int read_u32(clsph_state* st, const uint32_t* src, uint32_t* dst, size_t count) {
  if (count == 0) return CLSPH_OK;
  if (!dst) return fail(st->ctx, CLSPH_EINVAL, "NULL output");
  HIPCK(st->ctx, hipMemcpyAsync(dst, src, sizeof(uint32_t) * count, hipMemcpyDeviceToHost,
                                st->ctx->stream));
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  return CLSPH_OK;
}

int clsph_read_keys(clsph_state* st, uint32_t* keys, uint32_t n) {
  if (!st) return CLSPH_EINVAL;
  if (n > st->n) return fail(st->ctx, CLSPH_EINVAL, "n exceeds the particle count");
  int rc = read_u32(st, st->key_a, keys, n);
  for (uint32_t i = 0; !rc && i < n; ++i) keys[i] >>= 3;  // half-cell key -> cell id
  return rc;
}

int clsph_read_sorted(clsph_state* st, uint32_t* keys, uint32_t* perm, uint32_t n) {
  if (!st) return CLSPH_EINVAL;
  if (n != st->n) return fail(st->ctx, CLSPH_EINVAL, "n must be the particle count");
  std::vector<uint32_t> k(n), v(n), slot(n);
  int rc = read_u32(st, st->key_b, k.data(), n);
  if (!rc) rc = read_u32(st, st->val_b, v.data(), n);
  if (!rc) rc = read_u32(st, st->slot, slot.data(), n);
  if (rc) return rc;
  // Half-cell order -> the reference's order (stable by cell id).
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t o = slot[i];
    if (o >= n) return fail(st->ctx, CLSPH_EDEVICE, "canonical slots are not a permutation");
    if (keys) keys[o] = k[i] >> 3;
    if (perm) perm[o] = v[i];
  }
  return CLSPH_OK;
}

int clsph_read_cell_table(clsph_state* st, uint32_t* table, uint32_t count) {
  if (!st) return CLSPH_EINVAL;
  if (count > st->table_cap) return fail(st->ctx, CLSPH_EINVAL, "count exceeds table capacity");
  if (count == 0) return device_errors(st);
  std::vector<uint32_t> ht(8ull * count);
  int rc = read_u32(st, st->table, ht.data(), ht.size());
  if (rc) return rc;
  for (uint32_t c = 0; c < count; ++c) table[c] = ht[8ull * c];  // T[c] = ht[8c]
  return device_errors(st);
}

int clsph_bounce_capped(clsph_state* st, uint32_t* count) {
  if (!st || !count) return CLSPH_EINVAL;
  return read_u32(st, &st->dp->capped, count, 1);
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
  st->stage = 6;
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
  constexpr int kStages = 7;
  std::vector<hipEvent_t> ev((size_t)nsteps * (kStages + 1));
  for (auto& e : ev) HIPCK(st->ctx, hipEventCreate(&e));
  for (uint32_t k = 0; k < nsteps; ++k) {
    hipEvent_t* e = ev.data() + (size_t)k * (kStages + 1);
    HIPCK(st->ctx, hipEventRecord(e[0], st->ctx->stream));
    for (int s = 0; s < kStages; ++s) {
      int rc = run_stage(st, s);
      if (rc) return rc;
      HIPCK(st->ctx, hipEventRecord(e[s + 1], st->ctx->stream));
    }
  }
  HIPCK(st->ctx, hipEventSynchronize(ev.back()));
  for (int s = 0; s < kStages; ++s) stage_ms[s] = 0.f;
  for (uint32_t k = 0; k < nsteps; ++k) {
    hipEvent_t* e = ev.data() + (size_t)k * (kStages + 1);
    for (int s = 0; s < kStages; ++s) {
      float t = 0.f;
      HIPCK(st->ctx, hipEventElapsedTime(&t, e[s], e[s + 1]));
      stage_ms[s] += t;
    }
  }
  for (int s = 0; s < kStages; ++s) stage_ms[s] /= (float)(nsteps ? nsteps : 1);
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
  const double pi = 3.14159265358979323846;
  float h9 = powf(h, 9.f), h6 = powf(h, 6.f);
  out->poly_6 = (float)(315.f / (64.f * pi * h9));
  out->poly_6_gradient = (float)(-945.f / (32.f * pi * h9));
  out->poly_6_laplacian = (float)(-945.f / (32.f * pi * h9));
  out->spiky = (float)(-45.f / (pi * h6));
  out->viscosity = (float)(45.f / (pi * h6));
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
