// C ABI of the two-round message-passing network (include/clsph_hip.h,
// "two-round network" section): the network of the learned-surrogate
// configuration at full size, beside the corrective surrogate of nn_capi.cpp.
#include <string.h>

#include <string>
#include <vector>

#include "../../include/clsph_hip.h"
#include "hostutil.h"
#include "gnn.h"
#include "state.h"

using namespace clsph;
using namespace clsph::host;

namespace {

void free_gnn(clsph_gnn* m) {
  dfree(m->params), dfree(m->wb), dfree(m->grad), dfree(m->adam_m), dfree(m->adam_v);
  dfree(m->off), dfree(m->src), dfree(m->dst), dfree(m->scan_part);
  dfree(m->E1), dfree(m->nu), dfree(m->degc);
  for (int r = 0; r < 2; ++r) dfree(m->z1[r]), dfree(m->s1[r]), dfree(m->z2[r]), dfree(m->S[r]);
  dfree(m->tmp), dfree(m->u1), dfree(m->p1), dfree(m->y1), dfree(m->p2), dfree(m->y2);
  dfree(m->Hb), dfree(m->PA), dfree(m->PB), dfree(m->u2), dfree(m->q1), dfree(m->w1);
  dfree(m->q2), dfree(m->w2), dfree(m->out);
  dfree(m->target), dfree(m->loss), dfree(m->dout), dfree(m->dY), dfree(m->dpb), dfree(m->dS);
  dfree(m->dz2), dfree(m->ds1), dfree(m->dz1), dfree(m->dPA), dfree(m->dPB), dfree(m->dH);
  dfree(m->dPA_b), dfree(m->dPB_b);
}

int gnn_check_state(clsph_gnn* m, clsph_state* st) {
  if (!m || !st) return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (st->n_scenes != 1)
    return fail(m->ctx, CLSPH_EINVAL, "the two-round network runs on single-scene states");
  if (st->n > m->n_cap)
    return fail(m->ctx, CLSPH_EINVAL, "state holds more particles than the network workspace");
  return CLSPH_OK;
}

// Builds the neighbour graph of the state's last density stage and the edge
// inputs; reads the edge count back (one synchronisation).
int gnn_build_graph(clsph_gnn* m, clsph_state* st) {
  gnn::launch_graph(m, st);
  HIPCK(m->ctx, hipMemcpyAsync(&m->e, m->off + st->n, sizeof(uint32_t), hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  if (m->e > m->e_cap)
    return fail(m->ctx, CLSPH_ENOMEM,
                "neighbour graph has " + std::to_string(m->e) +
                    " edges, more than the network workspace holds (" +
                    std::to_string(m->e_cap) + ")");
  gnn::launch_edge_inputs(m, st);
  return check_launch(m->ctx);
}

void gnn_gravity(const clsph_state* st, float g[3]) {
  for (int c = 0; c < 3; ++c) g[c] = st->host_dps[0].p.constant_acceleration.s[c];
}


}  // namespace

extern "C" {

uint32_t clsph_gnn_param_count(void) { return gnn::kParamTotal; }

int clsph_gnn_create(clsph_ctx* ctx, uint32_t n_cap, uint32_t e_cap, clsph_gnn** out) {
  if (!ctx || !out || n_cap == 0) return fail(ctx, CLSPH_EINVAL, "clsph_gnn_create: bad argument");
  *out = nullptr;
  HIPCK(ctx, hipSetDevice(ctx->device));
  clsph_gnn* m = new clsph_gnn();
  m->ctx = ctx;
  m->n_cap = n_cap;
  m->e_cap = e_cap ? e_cap : 64u * n_cap;
  size_t N = n_cap, E = m->e_cap, Wd = gnn::kWidth, P = gnn::kParamTotal;
  int rc = CLSPH_OK;
#define AL(ptr, count) \
  if (rc == CLSPH_OK) rc = dalloc(ctx, &(ptr), (count))
  AL(m->params, P); AL(m->wb, P); AL(m->grad, P); AL(m->adam_m, P); AL(m->adam_v, P);
  AL(m->off, N + 1); AL(m->src, E); AL(m->dst, E);
  AL(m->scan_part, N / 1024 + 2);
  AL(m->E1, E * gnn::kFeat); AL(m->nu, 2 * N); AL(m->degc, N);
  for (int r = 0; r < 2; ++r) {
    AL(m->z1[r], E * Wd); AL(m->s1[r], E * Wd); AL(m->z2[r], E * Wd); AL(m->S[r], N * Wd);
  }
  AL(m->tmp, N * gnn::kIn2); AL(m->u1, N * gnn::kIn1); AL(m->p1, N * Wd); AL(m->y1, N * Wd);
  AL(m->p2, N * Wd); AL(m->y2, N * Wd); AL(m->Hb, N * Wd); AL(m->PA, N * Wd); AL(m->PB, N * Wd);
  AL(m->u2, N * gnn::kIn2); AL(m->q1, N * Wd); AL(m->w1, N * Wd); AL(m->q2, N * Wd);
  AL(m->w2, N * Wd); AL(m->out, N * gnn::kOut);
  AL(m->target, 3 * N); AL(m->loss, 1); AL(m->dout, N * gnn::kOut); AL(m->dY, N * gnn::kIn2);
  AL(m->dpb, N * gnn::kIn2); AL(m->dS, N * Wd); AL(m->dz2, E * Wd); AL(m->ds1, E * Wd);
  AL(m->dz1, E * Wd); AL(m->dPA, N * Wd); AL(m->dPB, N * Wd); AL(m->dH, N * Wd);
  AL(m->dPA_b, N * Wd); AL(m->dPB_b, N * Wd);
#undef AL
  if (rc) {
    free_gnn(m);
    delete m;
    return rc;
  }
  gnn::fill(m->params, P, 0.f, ctx->stream);
  gnn::fill(m->adam_m, P, 0.f, ctx->stream);
  gnn::fill(m->adam_v, P, 0.f, ctx->stream);
  gnn::cast_params(m);
  rc = check_launch(ctx);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
    rc = fail(ctx, CLSPH_EDEVICE, "clsph_gnn_create: synchronisation failed");
  if (rc) {
    free_gnn(m);
    delete m;
    return rc;
  }
  *out = m;
  return CLSPH_OK;
}

int clsph_gnn_free(clsph_gnn* m) {
  if (!m) return CLSPH_OK;
  (void)hipStreamSynchronize(m->ctx->stream);
  free_gnn(m);
  delete m;
  return CLSPH_OK;
}

int clsph_gnn_set_params(clsph_gnn* m, const float* params, uint32_t count) {
  if (!m || !params || count != gnn::kParamTotal)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_gnn_set_params: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(m->params, params, sizeof(float) * count, hipMemcpyHostToDevice,
                               m->ctx->stream));
  gnn::cast_params(m);
  gnn::fill(m->adam_m, count, 0.f, m->ctx->stream);
  gnn::fill(m->adam_v, count, 0.f, m->ctx->stream);
  m->adam_step = 0;
  int rc = check_launch(m->ctx);
  if (rc) return rc;
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_gnn_get_params(clsph_gnn* m, float* params, uint32_t count) {
  if (!m || !params || count != gnn::kParamTotal)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_gnn_get_params: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(params, m->params, sizeof(float) * count, hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_gnn_forward(clsph_gnn* m, clsph_state* st, uint32_t* edges) {
  int rc = gnn_check_state(m, st);
  if (rc) return rc;
  if (st->stage != kStageDensity && st->stage != kStageForces)
    return fail(m->ctx, CLSPH_EINVAL, "clsph_gnn_forward needs the density stage of the substep");
  rc = gnn_build_graph(m, st);
  if (rc) return rc;
  float g[3];
  gnn_gravity(st, g);
  gnn::forward(m, nullptr, st->acc_s, g[0], g[1], g[2]);
  rc = check_launch(m->ctx);
  if (rc) return rc;
  st->stage = kStageForces;
  if (edges) *edges = m->e;
  return CLSPH_OK;
}

int clsph_gnn_infer(clsph_gnn* m, clsph_state* st) {
  int rc = gnn_check_state(m, st);
  if (rc) return rc;
  if (st->stage != kStageDensity && st->stage != kStageForces)
    return fail(m->ctx, CLSPH_EINVAL, "clsph_gnn_infer needs the density stage of the substep");
  float g[3];
  gnn_gravity(st, g);
  gnn::forward_fused(m, st, st->acc_s, g[0], g[1], g[2]);
  rc = check_launch(m->ctx);
  if (rc) return rc;
  st->stage = kStageForces;
  return CLSPH_OK;
}

int clsph_gnn_profile(clsph_gnn* m, clsph_state* st, uint32_t reps, float* ms,
                     uint64_t* edges) {
  int rc = gnn_check_state(m, st);
  if (rc) return rc;
  if (!ms || reps == 0) return fail(m->ctx, CLSPH_EINVAL, "clsph_gnn_profile: bad argument");
  if (st->stage != kStageDensity && st->stage != kStageForces)
    return fail(m->ctx, CLSPH_EINVAL, "clsph_gnn_profile needs the density stage of the substep");
  float g[3];
  gnn_gravity(st, g);
  hipEvent_t ev[5];
  for (auto& e : ev) HIPCK(m->ctx, hipEventCreate(&e));
  double acc[4] = {0, 0, 0, 0};
  for (uint32_t r = 0; r < reps && !rc; ++r) {
    gnn::forward_fused(m, st, st->acc_s, g[0], g[1], g[2], ev);
    rc = check_launch(m->ctx);
    if (rc) break;
    if (hipEventSynchronize(ev[4]) != hipSuccess) {
      rc = fail(m->ctx, CLSPH_EDEVICE, "clsph_gnn_profile: event wait failed");
      break;
    }
    for (int k = 0; k < 4; ++k) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, ev[k], ev[k + 1]);
      acc[k] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  for (int k = 0; k < 4; ++k) ms[k] = (float)(acc[k] / reps);
  st->stage = kStageForces;
  if (edges) {
    std::vector<uint32_t> cnt(st->n);
    HIPCK(m->ctx, hipMemcpy(cnt.data(), st->nbr_count, sizeof(uint32_t) * st->n,
                            hipMemcpyDeviceToHost));
    uint64_t e = 0;
    for (uint32_t c : cnt) e += c < gnn::kMaxEdges ? c : gnn::kMaxEdges;
    *edges = e;
  }
  return CLSPH_OK;
}

int clsph_gnn_read_graph(clsph_gnn* m, uint32_t* off, uint32_t* src, uint32_t n, uint32_t e) {
  if (!m || n != m->n || e != m->e || (!off && n) || (!src && e))
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_gnn_read_graph: sizes do not match");
  hipStream_t s = m->ctx->stream;
  HIPCK(m->ctx, hipMemcpyAsync(off, m->off, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, s));
  if (e) HIPCK(m->ctx, hipMemcpyAsync(src, m->src, sizeof(uint32_t) * e, hipMemcpyDeviceToHost, s));
  HIPCK(m->ctx, hipStreamSynchronize(s));
  return CLSPH_OK;
}

int clsph_gnn_read_output(clsph_gnn* m, float* out, uint32_t n) {
  if (!m || !out || n != m->n)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_gnn_read_output: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(out, m->out, sizeof(float) * n * gnn::kOut, hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_gnn_set_target(clsph_gnn* m, const float* target, uint32_t n) {
  if (!m || !target || n != m->n)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_gnn_set_target: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(m->target, target, sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_gnn_read_target(clsph_gnn* m, float* target, uint32_t n) {
  if (!m || !target || n > m->n_cap)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_gnn_read_target: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(target, m->target, sizeof(float) * 3 * n, hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_gnn_backward(clsph_gnn* m, float* loss) {
  if (!m || m->n == 0) return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "no forward pass to differentiate");
  gnn::backward(m);
  int rc = check_launch(m->ctx);
  if (rc) return rc;
  if (loss) {
    HIPCK(m->ctx, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost,
                                 m->ctx->stream));
    HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  }
  return CLSPH_OK;
}

int clsph_gnn_read_grad(clsph_gnn* m, float* grad, uint32_t count) {
  if (!m || !grad || count != gnn::kParamTotal)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_gnn_read_grad: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(grad, m->grad, sizeof(float) * count, hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_gnn_adam(clsph_gnn* m, float lr, float beta1, float beta2, float eps) {
  if (!m) return fail(nullptr, CLSPH_EINVAL, "NULL network");
  gnn::adam(m, lr, beta1, beta2, eps);
  return check_launch(m->ctx);
}

}  // extern "C"
