// Learned force surrogate: a one-round message-passing network over the
// particles' neighbour graph (r < h) that corrects the forces stage's
// acceleration so that it stands for the mean acceleration over a macro step
// of 4 substeps. Evaluated and trained with bfloat16 matrix-core products
// accumulated in float32.
//
// Network (per particle i, neighbours j; all products bf16 x bf16 -> f32):
//   nu_i  = [rho_i / rho0 - 1, asinh(p_i / K) / 4]
//   e_ij  = [(x_i - x_j) / h, (v_i - v_j) / V, q, (1 - q^2)^3, nu_i, nu_j,
//            pressure and viscosity pair accelerations / A, colour-field
//            gradient term * h and laplacian term * h^2, the rate of the
//            pressure pair acceleration along the pair's motion, the
//            density-rate pair term, (x_i - x_j).(v_i - v_j), 0 x 5]
//   edges: S_i = bf16(sum_j relu(e_ij W1 + b1) / C)                (32 wide)
//   node:  u_i = [S_i, nu_i, (a0_i - g) / A, deg_i / C, 0 x 26]     (64 wide)
//          out_i = 3-layer MLP of u_i, SiLU between layers (64 -> 64 -> 64 -> 32)
//   acceleration a_i = a0_i + R out_i[0:3]
// a0 is the forces stage's acceleration of the same substep; a zero network
// leaves it unchanged.
#pragma once

#include <stdint.h>

#include "state.h"

namespace clsph {
namespace nn {

constexpr int kEdgeW = 32;   // edge layer width
constexpr int kNodeW = 64;   // node MLP width
constexpr int kFeat = 32;    // edge inputs (27 used)
constexpr int kFeatUsed = 27;
constexpr int kIn = 64;      // node MLP input (38 used)
constexpr int kOut = 32;     // network output (3 used)
constexpr float kAggC = 32.f;
constexpr float kVelScale = 1.f;
constexpr float kAccScale = 10.f;  // A
constexpr float kResScale = 2.f;   // R
constexpr uint32_t kMaxEdges = 128;  // neighbours per particle the network sees
constexpr uint32_t kMacro = 4;       // substeps per learned macro step

// Flat parameter vector: row-major (inputs x outputs) matrices and biases.
enum Param { E_W1, E_B1, N_W1, N_B1, N_W2, N_B2, N_W3, N_B3, kNumParams };
constexpr uint32_t kParamSize[kNumParams] = {
    kFeat * kEdgeW, kEdgeW, kIn * kNodeW, kNodeW, kNodeW * kNodeW, kNodeW, kNodeW * kOut, kOut};

constexpr uint32_t param_offset(int p) {
  uint32_t o = 0;
  for (int i = 0; i < p; ++i) o += kParamSize[i];
  return o;
}
constexpr uint32_t kParamTotal = param_offset(kNumParams);

}  // namespace nn
}  // namespace clsph

struct clsph_nn {
  clsph_ctx* ctx = nullptr;
  uint64_t serial = 0;  // unique per workspace (graphs captured for it name it)
  uint32_t n_cap = 0;  // particles
  uint32_t e_cap = 0;  // edges whose activations the training buffers hold

  // parameters (float32 masters), their bf16 copies, gradients, Adam moments
  float* params = nullptr;
  uint16_t* wb = nullptr;
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  uint32_t adam_step = 0;

  // neighbour graph of the current evaluation (dst-major CSR)
  uint32_t n = 0, e = 0;
  uint32_t* off = nullptr;      // n_cap + 1
  uint32_t* src = nullptr;      // e_cap
  uint32_t* dst = nullptr;      // e_cap
  uint32_t* scan_part = nullptr;

  // forward intermediates of the training path
  uint16_t* E1 = nullptr;       // e x 32 edge inputs
  float* nu = nullptr;          // n x 2
  float* degc = nullptr;        // deg / C
  float* z = nullptr;           // e x 32 edge pre-activations
  uint16_t* S = nullptr;        // n x 32: bf16(sum_j relu(z) / C)
  uint16_t* u = nullptr;        // n x 64 node inputs
  float* p1 = nullptr;          // n x 64
  uint16_t* y1 = nullptr;
  float* p2 = nullptr;
  uint16_t* y2 = nullptr;
  float* out = nullptr;         // n x 32

  // training
  float* target = nullptr;      // n x 3
  float* loss = nullptr;        // 1
  float* dout = nullptr;        // n x 32
  uint16_t* dout_b = nullptr;   // n x 32
  float* dY = nullptr;          // n x 64
  uint16_t* dpb = nullptr;      // n x 64
  uint16_t* dz = nullptr;       // e x 32
  float* sum = nullptr;         // n x 3 acceleration sums of the mean target
  uint32_t* loc = nullptr;      // n: inverse of a state's ids
  uint32_t* inv = nullptr;      // n: inverse of a sort permutation
  float* cmp = nullptr;         // 2: max and sum of |rho_a - rho_b|
};

namespace clsph {
namespace nn {

// Graph + edge inputs from a state whose density stage has run.
void launch_graph(clsph_nn* m, clsph_state* s);
void launch_edge_inputs(clsph_nn* m, clsph_state* s);
// Training forward (graph built): reads the forces stage's acceleration from
// `acc0` and leaves the output in m->out; the state is not changed.
void forward(clsph_nn* m, const float4* acc0, float gx, float gy, float gz);
// Inference forward, one fused kernel over the state's neighbour lists (no
// per-edge storage, no edge-count readback): acc[i] = a0_i + R out_i[0:3].
// a0 is acc[i] as the forces stage left it, or with `own_forces` the forces
// stage's acceleration summed in the same kernel from the listed neighbours
// (then acc must hold the forces stage's result for the particles whose list
// overflowed: launch_forces_overflow). With `out` non-null the full output
// is written there as well. With `ev` (2 events) the stream gets events
// before and after the kernel. With `nu_ready` the node inputs m->nu are
// already current (the density stage wrote them).
void forward_fused(clsph_nn* m, clsph_state* st, float4* acc, float* out, bool own_forces,
                   hipEvent_t* ev = nullptr, bool nu_ready = false);
// Loss against m->target and gradients into m->grad (zeroed first).
void backward(clsph_nn* m);
void adam(clsph_nn* m, float lr, float beta1, float beta2, float eps);
void cast_params(clsph_nn* m);
// target[i] = ((vh_T[loc[perm_S[i]]] - vh_S[perm_S[i]]) / dt - a0_i) / R,
// a0 the learner's forces-stage acceleration at sorted i.
void launch_target(clsph_nn* m, clsph_state* learner, clsph_state* teacher, float dt);
// Advances `teacher` (a copy of the learner's particles with identities
// reset) by nsub plain substeps stage by stage, summing each particle's
// forces-stage acceleration by identity; target[i] = (mean - a0_i) / R for
// the learner's sorted particle i, a0 its forces-stage acceleration.
// Collision impulses (applied by advection) are not part of the target.
void launch_target_mean(clsph_nn* m, clsph_state* learner, clsph_state* teacher, uint32_t nsub);
// Per-particle density difference of two states holding the same particles
// (matched through their identities); res (host, read after a
// synchronisation) = {max |rho_a - rho_b|, sum |rho_a - rho_b|}.
void launch_compare(clsph_nn* m, clsph_state* a, clsph_state* b, float res[2]);
void fill(float* x, size_t count, float v, hipStream_t s);
// Bytes of the static device workspace of clsph_nn_gemm_noalloc.
constexpr size_t kScratchBytes = size_t(8) << 20;
// Device address of that workspace on `device` (which must be current),
// looked up on the first call for the device.
hipError_t scratch_address(int device, void** ptr);
// The bf16 matrix-core GEMM (see nn.hip) for direct use and testing.
void gemm_public(hipStream_t s, bool ta, bool tb, bool add, const uint16_t* A, int lda,
                 const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K, float alpha,
                 const float* bias, const float* rowscale);

}  // namespace nn
}  // namespace clsph
