// C ABI of the learned force surrogate and the state helpers its training
// and judging use (include/clsph_hip.h, "learned surrogate" section).
#include <math.h>
#include <string.h>

#include <atomic>

#include <string>
#include <vector>

#include "../../include/clsph_hip.h"
#include "hostutil.h"
#include "nn.h"
#include "state.h"

using namespace clsph;
using namespace clsph::host;

namespace {

void free_nn(clsph_nn* m) {
  dfree(m->params), dfree(m->wb), dfree(m->grad), dfree(m->adam_m), dfree(m->adam_v);
  dfree(m->off), dfree(m->src), dfree(m->dst), dfree(m->scan_part);
  dfree(m->E1), dfree(m->nu), dfree(m->degc), dfree(m->z), dfree(m->S), dfree(m->u);
  dfree(m->p1), dfree(m->y1), dfree(m->p2), dfree(m->y2), dfree(m->out);
  dfree(m->target), dfree(m->loss), dfree(m->dout), dfree(m->dout_b), dfree(m->dY);
  dfree(m->dpb), dfree(m->dz), dfree(m->sum), dfree(m->loc), dfree(m->inv), dfree(m->cmp);
}

int check_state(clsph_nn* m, clsph_state* st) {
  if (!m || !st) return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "NULL argument");
  if (st->n_scenes != 1)
    return fail(m->ctx, CLSPH_EINVAL, "the learned surrogate runs on single-scene states");
  if (st->n > m->n_cap)
    return fail(m->ctx, CLSPH_EINVAL, "state holds more particles than the network workspace");
  return CLSPH_OK;
}

// Builds the neighbour graph of the state's last density stage and the edge
// inputs; reads the edge count back (one synchronisation).
int build_graph(clsph_nn* m, clsph_state* st) {
  nn::launch_graph(m, st);
  HIPCK(m->ctx, hipMemcpyAsync(&m->e, m->off + st->n, sizeof(uint32_t), hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  if (m->e > m->e_cap)
    return fail(m->ctx, CLSPH_ENOMEM,
                "neighbour graph has " + std::to_string(m->e) +
                    " edges, more than the network workspace holds (" +
                    std::to_string(m->e_cap) + ")");
  nn::launch_edge_inputs(m, st);
  return check_launch(m->ctx);
}

void gravity(const clsph_state* st, float g[3]) {
  for (int c = 0; c < 3; ++c) g[c] = st->host_dps[0].p.constant_acceleration.s[c];
}

int copy_d(clsph_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes)
    HIPCK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return CLSPH_OK;
}

// Copies the particles and per-particle fields of single-scene `src` into
// `dst` on the context stream (identities when both track them).
int copy_state_async(clsph_state* dst, clsph_state* src) {
  if (!dst || !src) return fail(nullptr, CLSPH_EINVAL, "NULL state");
  if (dst->n_scenes != 1 || src->n_scenes != 1 || dst->n_cap < src->n)
    return fail(dst->ctx, CLSPH_EINVAL, "state copy needs single-scene states of enough capacity");
  clsph_ctx* c = dst->ctx;
  size_t n = src->n;
  int rc = copy_d(c, dst->pos, src->pos, sizeof(float4) * n);
  if (!rc) rc = copy_d(c, dst->vel, src->vel, sizeof(float4) * n);
  if (!rc) rc = copy_d(c, dst->vh, src->vh, sizeof(float4) * n);
  if (!rc) rc = copy_d(c, dst->acc_s, src->acc_s, sizeof(float4) * n);
  if (!rc) rc = copy_d(c, dst->rho_s, src->rho_s, sizeof(float) * n);
  if (!rc) rc = copy_d(c, dst->prs_s, src->prs_s, sizeof(float) * n);
  if (!rc) rc = copy_d(c, dst->key_b, src->key_b, sizeof(uint32_t) * n);
  if (!rc) rc = copy_d(c, dst->slot, src->slot, sizeof(uint32_t) * n);
  if (!rc && dst->ids && src->ids) rc = copy_d(c, dst->ids, src->ids, sizeof(uint32_t) * n);
  if (rc) return rc;
  if (dst->n != src->n) {
    rc = resize_single(dst, src->n);
    if (rc) return rc;
  }
  launch_aabb(dst);
  dst->acc_zero = src->acc_zero;
  dst->acc_learned = false;
  dst->fields_canonical = src->fields_canonical;
  dst->stage = -1;
  return check_launch(c);
}

// The stages of a substep up to density, as clsph_grid .. clsph_density_pressure.
void enqueue_density_stages(clsph_state* st) {
  launch_grid_params(st);
  launch_hash(st);
  launch_sort(st);
  launch_density(st);
  st->acc_learned = false;
  st->fields_canonical = false;
  st->stage = kStageDensity;
}

// The learned macro step (the plain substep's stages, the network's
// correction after forces, advection) captured as one HIP graph per state;
// recaptured when the network, the particle count, the parameters or the
// work-group tables change.
int ensure_learned_graph(clsph_state* st, clsph_nn* m) {
  if (st->lgraph && !st->lgraph_dirty && st->lgraph_serial == m->serial && st->lgraph_n == st->n)
    return CLSPH_OK;
  hipStream_t s = st->ctx->stream;
  if (st->lgraph) {
    // Queued replays of the old executable finish before it is destroyed.
    HIPCK(st->ctx, hipStreamSynchronize(s));
    (void)hipGraphExecDestroy(st->lgraph);
    st->lgraph = nullptr;
  }
  hipGraph_t graph = nullptr;
  HIPCK(st->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  launch_grid_params(st);
  launch_hash(st);
  launch_sort(st);
  launch_density(st, (float2*)m->nu);
  launch_forces_overflow(st);
  nn::forward_fused(m, st, st->acc_s, nullptr, true, nullptr, true);
  launch_advect(st);
  if (st->ids) launch_track_ids(st);
  hipError_t le = hipGetLastError();
  HIPCK(st->ctx, hipStreamEndCapture(s, &graph));
  if (le != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(st->ctx, CLSPH_EDEVICE, std::string("capture: ") + hipGetErrorString(le));
  }
  hipError_t ie = hipGraphInstantiate(&st->lgraph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) {
    st->lgraph = nullptr;
    return fail(st->ctx, CLSPH_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
  }
  st->lgraph_dirty = false;
  st->lgraph_serial = m->serial;
  st->lgraph_n = st->n;
  return CLSPH_OK;
}


}  // namespace

extern "C" {

uint32_t clsph_nn_param_count(void) { return nn::kParamTotal; }

int clsph_nn_create(clsph_ctx* ctx, uint32_t n_cap, uint32_t e_cap, clsph_nn** out) {
  if (!ctx || !out || n_cap == 0) return fail(ctx, CLSPH_EINVAL, "clsph_nn_create: bad argument");
  *out = nullptr;
  HIPCK(ctx, hipSetDevice(ctx->device));
  clsph_nn* m = new clsph_nn();
  static std::atomic<uint64_t> next_serial{1};
  m->ctx = ctx;
  m->serial = next_serial++;
  m->n_cap = n_cap;
  m->e_cap = e_cap ? e_cap : 64u * n_cap;
  size_t N = n_cap, E = m->e_cap, P = nn::kParamTotal;
  size_t EW = nn::kEdgeW, NW = nn::kNodeW;
  int rc = CLSPH_OK;
#define AL(ptr, count) \
  if (rc == CLSPH_OK) rc = dalloc(ctx, &(ptr), (count))
  AL(m->params, P); AL(m->wb, P); AL(m->grad, P); AL(m->adam_m, P); AL(m->adam_v, P);
  AL(m->off, N + 1); AL(m->src, E); AL(m->dst, E);
  AL(m->scan_part, N / 1024 + 2);
  AL(m->E1, E * nn::kFeat); AL(m->nu, 2 * N); AL(m->degc, N);
  AL(m->z, E * EW); AL(m->S, N * EW); AL(m->u, N * nn::kIn);
  AL(m->p1, N * NW); AL(m->y1, N * NW); AL(m->p2, N * NW); AL(m->y2, N * NW);
  AL(m->out, N * nn::kOut);
  AL(m->target, 3 * N); AL(m->loss, 1); AL(m->dout, N * nn::kOut); AL(m->dout_b, N * nn::kOut);
  AL(m->dY, N * NW); AL(m->dpb, N * NW); AL(m->dz, E * EW); AL(m->sum, 3 * N);
  AL(m->loc, N); AL(m->inv, N); AL(m->cmp, 2);
#undef AL
  if (rc) {
    free_nn(m);
    delete m;
    return rc;
  }
  nn::fill(m->params, P, 0.f, ctx->stream);
  nn::fill(m->adam_m, P, 0.f, ctx->stream);
  nn::fill(m->adam_v, P, 0.f, ctx->stream);
  nn::cast_params(m);
  rc = check_launch(ctx);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
    rc = fail(ctx, CLSPH_EDEVICE, "clsph_nn_create: synchronisation failed");
  if (rc) {
    free_nn(m);
    delete m;
    return rc;
  }
  *out = m;
  return CLSPH_OK;
}

int clsph_nn_free(clsph_nn* m) {
  if (!m) return CLSPH_OK;
  (void)hipStreamSynchronize(m->ctx->stream);
  free_nn(m);
  delete m;
  return CLSPH_OK;
}

int clsph_nn_set_params(clsph_nn* m, const float* params, uint32_t count) {
  if (!m || !params || count != nn::kParamTotal)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_nn_set_params: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(m->params, params, sizeof(float) * count, hipMemcpyHostToDevice,
                               m->ctx->stream));
  nn::cast_params(m);
  nn::fill(m->adam_m, count, 0.f, m->ctx->stream);
  nn::fill(m->adam_v, count, 0.f, m->ctx->stream);
  m->adam_step = 0;
  int rc = check_launch(m->ctx);
  if (rc) return rc;
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_nn_get_params(clsph_nn* m, float* params, uint32_t count) {
  if (!m || !params || count != nn::kParamTotal)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_nn_get_params: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(params, m->params, sizeof(float) * count, hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_nn_forward(clsph_nn* m, clsph_state* st, uint32_t* edges) {
  int rc = check_state(m, st);
  if (rc) return rc;
  if (st->stage != kStageForces || st->acc_learned)
    return fail(m->ctx, CLSPH_EINVAL, "clsph_nn_forward needs the forces stage of the substep");
  rc = build_graph(m, st);
  if (rc) return rc;
  float g[3];
  gravity(st, g);
  nn::forward(m, st->acc_s, g[0], g[1], g[2]);
  rc = check_launch(m->ctx);
  if (rc) return rc;
  if (edges) *edges = m->e;
  return CLSPH_OK;
}

int clsph_nn_infer(clsph_nn* m, clsph_state* st) {
  int rc = check_state(m, st);
  if (rc) return rc;
  if (st->stage != kStageForces || st->acc_learned)
    return fail(m->ctx, CLSPH_EINVAL,
                "clsph_nn_infer needs the forces stage of the substep, once");
  nn::forward_fused(m, st, st->acc_s, m->out, false);
  rc = check_launch(m->ctx);
  if (rc) return rc;
  st->acc_learned = true;
  return CLSPH_OK;
}

int clsph_nn_profile(clsph_nn* m, clsph_state* st, uint32_t reps, float* ms,
                     uint64_t* edges) {
  int rc = check_state(m, st);
  if (rc) return rc;
  if (!ms || reps == 0) return fail(m->ctx, CLSPH_EINVAL, "clsph_nn_profile: bad argument");
  if (st->stage != kStageForces)
    return fail(m->ctx, CLSPH_EINVAL, "clsph_nn_profile needs the forces stage of the substep");
  // the accelerations are copied into dY (n x 64 floats holds n float4s)
  float4* acc = reinterpret_cast<float4*>(m->dY);
  HIPCK(m->ctx, hipMemcpyAsync(acc, st->acc_s, sizeof(float4) * st->n, hipMemcpyDeviceToDevice,
                               m->ctx->stream));
  hipEvent_t ev[2];
  for (auto& e : ev) HIPCK(m->ctx, hipEventCreate(&e));
  double total = 0.0;
  for (uint32_t r = 0; r < reps && !rc; ++r) {
    nn::forward_fused(m, st, acc, nullptr, true, ev);
    rc = check_launch(m->ctx);
    if (rc) break;
    if (hipEventSynchronize(ev[1]) != hipSuccess) {
      rc = fail(m->ctx, CLSPH_EDEVICE, "clsph_nn_profile: event wait failed");
      break;
    }
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ev[0], ev[1]);
    total += t;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  ms[0] = (float)(total / reps);
  if (edges) {
    std::vector<uint32_t> cnt(st->n);
    HIPCK(m->ctx, hipMemcpy(cnt.data(), st->nbr_count, sizeof(uint32_t) * st->n,
                            hipMemcpyDeviceToHost));
    uint64_t e = 0;
    for (uint32_t c : cnt) e += c < nn::kMaxEdges ? c : nn::kMaxEdges;
    *edges = e;
  }
  return CLSPH_OK;
}

int clsph_nn_read_graph(clsph_nn* m, uint32_t* off, uint32_t* src, uint32_t n, uint32_t e) {
  if (!m || n != m->n || e != m->e || (!off && n) || (!src && e))
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_nn_read_graph: sizes do not match");
  hipStream_t s = m->ctx->stream;
  HIPCK(m->ctx, hipMemcpyAsync(off, m->off, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, s));
  if (e) HIPCK(m->ctx, hipMemcpyAsync(src, m->src, sizeof(uint32_t) * e, hipMemcpyDeviceToHost, s));
  HIPCK(m->ctx, hipStreamSynchronize(s));
  return CLSPH_OK;
}

int clsph_nn_read_output(clsph_nn* m, float* out, uint32_t n) {
  if (!m || !out || n != m->n)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_nn_read_output: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(out, m->out, sizeof(float) * n * nn::kOut, hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_nn_set_target(clsph_nn* m, const float* target, uint32_t n) {
  if (!m || !target || n != m->n)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_nn_set_target: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(m->target, target, sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_nn_read_target(clsph_nn* m, float* target, uint32_t n) {
  if (!m || !target || n > m->n_cap)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_nn_read_target: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(target, m->target, sizeof(float) * 3 * n, hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_nn_target_from(clsph_nn* m, clsph_state* learner, clsph_state* teacher, float dt) {
  int rc = check_state(m, learner);
  if (rc) return rc;
  if (!teacher || teacher->n != learner->n || !teacher->ids)
    return fail(m->ctx, CLSPH_EINVAL,
                "the teacher state must hold the learner's particles and track identities");
  if (learner->stage != kStageForces || learner->acc_learned || !(dt > 0.f))
    return fail(m->ctx, CLSPH_EINVAL,
                "the learner must have run its forces stage and dt must be positive");
  nn::launch_target(m, learner, teacher, dt);
  return check_launch(m->ctx);
}

int clsph_nn_target_mean_acc(clsph_nn* m, clsph_state* learner, clsph_state* teacher,
                             uint32_t nsub) {
  int rc = check_state(m, learner);
  if (rc) return rc;
  if (!teacher || teacher->n != learner->n || !teacher->ids || teacher->n_scenes != 1)
    return fail(m->ctx, CLSPH_EINVAL,
                "the teacher state must hold the learner's particles and track identities");
  if (learner->stage != kStageForces || learner->acc_learned || nsub == 0)
    return fail(m->ctx, CLSPH_EINVAL,
                "the learner must have run its forces stage and nsub must be positive");
  nn::launch_target_mean(m, learner, teacher, nsub);
  rc = check_launch(m->ctx);
  if (rc) return rc;
  teacher->acc_zero = true;
  teacher->fields_canonical = false;
  teacher->stage = kStageAdvect;
  return device_errors(teacher);
}

int clsph_nn_backward(clsph_nn* m, float* loss) {
  if (!m || m->n == 0) return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "no forward pass to differentiate");
  nn::backward(m);
  int rc = check_launch(m->ctx);
  if (rc) return rc;
  if (loss) {
    HIPCK(m->ctx, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost,
                                 m->ctx->stream));
    HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  }
  return CLSPH_OK;
}

int clsph_nn_read_grad(clsph_nn* m, float* grad, uint32_t count) {
  if (!m || !grad || count != nn::kParamTotal)
    return fail(m ? m->ctx : nullptr, CLSPH_EINVAL, "clsph_nn_read_grad: bad argument");
  HIPCK(m->ctx, hipMemcpyAsync(grad, m->grad, sizeof(float) * count, hipMemcpyDeviceToHost,
                               m->ctx->stream));
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CLSPH_OK;
}

int clsph_nn_adam(clsph_nn* m, float lr, float beta1, float beta2, float eps) {
  if (!m) return fail(nullptr, CLSPH_EINVAL, "NULL network");
  nn::adam(m, lr, beta1, beta2, eps);
  return check_launch(m->ctx);
}

int clsph_state_copy(clsph_state* dst, clsph_state* src) {
  int rc = copy_state_async(dst, src);
  if (rc) return rc;
  HIPCK(dst->ctx, hipStreamSynchronize(dst->ctx->stream));
  return CLSPH_OK;
}

int clsph_track_ids(clsph_state* st, int enable) {
  if (!st) return fail(nullptr, CLSPH_EINVAL, "NULL state");
  clsph_ctx* c = st->ctx;
  if (!enable) {
    if (st->ids) {
      HIPCK(c, hipStreamSynchronize(c->stream));
      dfree(st->ids), dfree(st->ids_tmp);
      st->graph_dirty = st->lgraph_dirty = true;
    }
    return CLSPH_OK;
  }
  if (!st->ids) {
    int rc = dalloc(c, &st->ids, st->n_cap);
    if (!rc) rc = dalloc(c, &st->ids_tmp, st->n_cap);
    if (rc) {
      dfree(st->ids), dfree(st->ids_tmp);
      return rc;
    }
    st->graph_dirty = st->lgraph_dirty = true;
  }
  std::vector<uint32_t> iota(st->n_cap);
  for (uint32_t k = 0; k < st->n_cap; ++k) iota[k] = k;
  HIPCK(c, hipMemcpyAsync(st->ids, iota.data(), sizeof(uint32_t) * st->n_cap,
                          hipMemcpyHostToDevice, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return CLSPH_OK;
}

int clsph_read_ids(clsph_state* st, uint32_t* ids, uint32_t n) {
  if (!st || !ids || n > st->n || !st->ids)
    return fail(st ? st->ctx : nullptr, CLSPH_EINVAL, "clsph_read_ids: no identities tracked");
  HIPCK(st->ctx, hipMemcpyAsync(ids, st->ids, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                                st->ctx->stream));
  HIPCK(st->ctx, hipStreamSynchronize(st->ctx->stream));
  return CLSPH_OK;
}

int clsph_learned_step(clsph_state* st, clsph_nn* m, uint32_t nsteps) {
  int rc = check_state(m, st);
  if (rc) return rc;
  if (st->n == 0 || nsteps == 0) return CLSPH_OK;
  rc = ensure_learned_graph(st, m);
  if (rc) return rc;
  for (uint32_t k = 0; k < nsteps; ++k)
    HIPCK(st->ctx, hipGraphLaunch(st->lgraph, st->ctx->stream));
  st->acc_zero = true;
  st->acc_learned = false;
  st->fields_canonical = false;
  st->stage = kStageAdvect;
  return device_errors(st);
}

int clsph_learned_run(clsph_state* learner, clsph_state* judge, clsph_nn* m,
                      uint32_t macro_steps, uint32_t judge_every, float tol, uint32_t* judged,
                      uint32_t* fallbacks, float* judge_err, float* judge_mean) {
  int rc = check_state(m, learner);
  if (!rc && judge_every) rc = check_state(m, judge);
  if (rc) return rc;
  if (judge_every && (judge == learner || judge->n_cap < learner->n || !judge->ids ||
                      !learner->ids))
    return fail(m->ctx, CLSPH_EINVAL,
                "a judged run needs a second state of enough capacity and identities in both");
  if (judge_every) {
    const float dl = learner->host_dps[0].p.time_delta, dj = judge->host_dps[0].p.time_delta;
    if (!(fabsf(dl - (float)nn::kMacro * dj) <= 1e-5f * dl))
      return fail(m->ctx, CLSPH_EINVAL,
                  "the learner's time_delta must be 4 times the judge's (one macro step)");
  }
  uint32_t nj = 0, nf = 0;
  const float rho0 = learner->host_dps[0].p.fluid_density;
  for (uint32_t k = 0; k < macro_steps; ++k) {
    const bool judge_now = judge_every && k % judge_every == 0;
    if (judge_now) {
      rc = copy_state_async(judge, learner);
      if (rc) return rc;
    }
    rc = ensure_learned_graph(learner, m);
    if (rc) return rc;
    HIPCK(m->ctx, hipGraphLaunch(learner->lgraph, m->ctx->stream));
    learner->acc_zero = true;
    learner->fields_canonical = false;
    learner->stage = kStageAdvect;
    if (!judge_now) continue;
    rc = clsph_step(judge, nn::kMacro);
    if (rc) return rc;
    enqueue_density_stages(learner);
    enqueue_density_stages(judge);
    float res[2];
    nn::launch_compare(m, learner, judge, res);
    rc = check_launch(m->ctx);
    if (rc) return rc;
    HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
    const float err = res[0] / rho0;
    if (judge_err) judge_err[nj] = err;
    if (judge_mean) judge_mean[nj] = res[1] / rho0 / (float)(learner->n ? learner->n : 1);
    ++nj;
    if (!(err <= tol)) {
      rc = copy_state_async(learner, judge);
      if (rc) return rc;
      ++nf;
    }
  }
  if (judged) *judged = nj;
  if (fallbacks) *fallbacks = nf;
  rc = device_errors(learner);
  if (!rc && judge_every) rc = device_errors(judge);
  return rc;
}

int clsph_nn_compare_density(clsph_nn* m, clsph_state* a, clsph_state* b, float* max_rel,
                             float* mean_rel) {
  int rc = check_state(m, a);
  if (!rc) rc = check_state(m, b);
  if (rc) return rc;
  if (a->n != b->n || !a->ids || !b->ids)
    return fail(m->ctx, CLSPH_EINVAL, "density comparison needs equal counts and identities");
  if (a->stage != kStageDensity || b->stage != kStageDensity)
    return fail(m->ctx, CLSPH_EINVAL, "both states must have run their density stage");
  float res[2];
  nn::launch_compare(m, a, b, res);
  rc = check_launch(m->ctx);
  if (rc) return rc;
  HIPCK(m->ctx, hipStreamSynchronize(m->ctx->stream));
  float rho0 = a->host_dps[0].p.fluid_density;
  if (max_rel) *max_rel = res[0] / rho0;
  if (mean_rel) *mean_rel = res[1] / rho0 / (float)(a->n ? a->n : 1);
  return CLSPH_OK;
}

namespace {

// Checks the shape of a clsph_nn_gemm call.
int gemm_check(clsph_ctx* ctx, const char* who, int trans_a, int trans_b, const uint16_t* a,
               const uint16_t* b, float* c, uint32_t m, uint32_t n, uint32_t k) {
  if (!ctx || !a || !b || !c || !m || !n || !k)
    return fail(ctx, CLSPH_EINVAL, std::string(who) + ": bad argument");
  // contiguous dimensions must hold whole 8-element (16-byte) vectors
  if ((trans_a ? m : k) % 8 || (trans_b ? k : n) % 8)
    return fail(ctx, CLSPH_EINVAL,
                std::string(who) + ": contiguous dimensions must be multiples of 8");
  return CLSPH_OK;
}

// Copies the host operands into the device buffers da, db, dc (and dbias),
// runs the product there and copies c back; waits for the stream.
int gemm_staged(clsph_ctx* ctx, const char* who, int trans_a, int trans_b, int accumulate,
                const uint16_t* a, const uint16_t* b, float* c, const float* bias, uint32_t m,
                uint32_t n, uint32_t k, uint16_t* da, uint16_t* db, float* dc, float* dbias) {
  size_t sa = (size_t)m * k, sb = (size_t)k * n, sc = (size_t)m * n;
  hipStream_t s = ctx->stream;
  if (hipMemcpyAsync(da, a, 2 * sa, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(db, b, 2 * sb, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dc, c, 4 * sc, hipMemcpyHostToDevice, s) != hipSuccess ||
      (bias && hipMemcpyAsync(dbias, bias, 4 * n, hipMemcpyHostToDevice, s) != hipSuccess))
    return fail(ctx, CLSPH_EDEVICE, std::string(who) + ": copy failed");
  int lda = trans_a ? (int)m : (int)k, ldb = trans_b ? (int)k : (int)n;
  // a non-accumulating product over a long K runs as one chunk
  nn::gemm_public(s, trans_a, trans_b, accumulate, da, lda, db, ldb, dc, (int)n, (int)m, (int)n,
                  (int)k, 1.f, bias ? dbias : nullptr, nullptr);
  int rc = check_launch(ctx);
  if (rc) return rc;
  if (hipMemcpyAsync(c, dc, 4 * sc, hipMemcpyDeviceToHost, s) != hipSuccess)
    return fail(ctx, CLSPH_EDEVICE, std::string(who) + ": copy back failed");
  HIPCK(ctx, hipStreamSynchronize(s));
  return CLSPH_OK;
}

size_t round256(size_t bytes) { return (bytes + 255) & ~size_t(255); }

}  // namespace

int clsph_nn_gemm(clsph_ctx* ctx, int trans_a, int trans_b, int accumulate, const uint16_t* a,
                  const uint16_t* b, float* c, const float* bias, uint32_t m, uint32_t n,
                  uint32_t k) {
  int rc = gemm_check(ctx, "clsph_nn_gemm", trans_a, trans_b, a, b, c, m, n, k);
  if (rc) return rc;
  size_t sa = (size_t)m * k, sb = (size_t)k * n, sc = (size_t)m * n;
  uint16_t *da = nullptr, *db = nullptr;
  float *dc = nullptr, *dbias = nullptr;
  rc = dalloc(ctx, &da, sa);
  if (!rc) rc = dalloc(ctx, &db, sb);
  if (!rc) rc = dalloc(ctx, &dc, sc);
  if (!rc && bias) rc = dalloc(ctx, &dbias, n);
  if (!rc)
    rc = gemm_staged(ctx, "clsph_nn_gemm", trans_a, trans_b, accumulate, a, b, c, bias, m, n, k,
                     da, db, dc, dbias);
  (void)hipStreamSynchronize(ctx->stream);
  dfree(da), dfree(db), dfree(dc), dfree(dbias);
  return rc;
}

int clsph_nn_gemm_noalloc(clsph_ctx* ctx, int trans_a, int trans_b, int accumulate,
                          const uint16_t* a, const uint16_t* b, float* c, const float* bias,
                          uint32_t m, uint32_t n, uint32_t k) {
  int rc = gemm_check(ctx, "clsph_nn_gemm_noalloc", trans_a, trans_b, a, b, c, m, n, k);
  if (rc) return rc;
  if (!ctx->scratch) return fail(ctx, CLSPH_EDEVICE, "clsph_nn_gemm_noalloc: no workspace");
  // a, b, c and bias at 256-byte aligned offsets of the static workspace
  size_t oa = 0;
  size_t ob = oa + round256(2 * (size_t)m * k);
  size_t oc = ob + round256(2 * (size_t)k * n);
  size_t obias = oc + round256(4 * (size_t)m * n);
  size_t total = obias + (bias ? 4 * (size_t)n : 0);
  if (total > nn::kScratchBytes)
    return fail(ctx, CLSPH_EINVAL,
                "clsph_nn_gemm_noalloc: the operands need " + std::to_string(total) +
                    " bytes of device workspace; it holds " + std::to_string(nn::kScratchBytes));
  uint8_t* w = (uint8_t*)ctx->scratch;
  return gemm_staged(ctx, "clsph_nn_gemm_noalloc", trans_a, trans_b, accumulate, a, b, c, bias,
                     m, n, k, (uint16_t*)(w + oa), (uint16_t*)(w + ob), (float*)(w + oc),
                     (float*)(w + obias));
}

}  // extern "C"
