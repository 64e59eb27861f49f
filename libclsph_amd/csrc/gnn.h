// Two-round message-passing network over the particles' neighbour graph
// (r < h), evaluated and trained with bfloat16 matrix-core products
// accumulated in float32: the full-size network a learned force surrogate can
// use in place of the corrective one-round network of nn.h.
//
// Network (per particle i, neighbours j; all products bf16 x bf16 -> f32):
//   nu_i  = [rho_i / rho0 - 1, asinh(p_i / K) / 4]
//   e_ij  = [(x_i - x_j) / h, (v_i - v_j) / V, q, (1 - q^2)^3, nu_i, nu_j, 0 x 4]
//   round 1: s1 = silu(e W1 + b1), s2 = silu(s1 W2 + b2),
//            M1_i = bf16(sum_j s2_ij / C) W3 + (deg_i / C) b3
//            H_i  = 3-layer SiLU MLP of [M1_i, nu_i, 0 x 14]      (144 -> 128)
//   round 2: z1 = e W1' + b1' + bf16(H_i Wa) + bf16(H_j Wb), then as round 1
//            out_i = 3-layer SiLU MLP of [H_i, M2_i]                (256 -> 32)
//   acceleration a_i = g + A out_i[0:3]
// The third edge layer is linear and is applied after the neighbour sum.
#pragma once

#include <stdint.h>

#include "state.h"

namespace clsph {
namespace gnn {

constexpr int kWidth = 128;
constexpr int kFeat = 16;
constexpr int kIn1 = 144;
constexpr int kIn2 = 256;
constexpr int kOut = 32;
constexpr float kAggC = 32.f;
constexpr float kVelScale = 1.f;
constexpr float kAccScale = 10.f;
constexpr uint32_t kMaxEdges = 128;  // neighbours per particle the network sees

// Flat parameter vector: row-major (inputs x outputs) matrices and biases.
enum Param {
  E1_W1, E1_B1, E1_W2, E1_B2, E1_W3, E1_B3,
  N1_W1, N1_B1, N1_W2, N1_B2, N1_W3, N1_B3,
  E2_W1, E2_B1, E2_WA, E2_WB, E2_W2, E2_B2, E2_W3, E2_B3,
  N2_W1, N2_B1, N2_W2, N2_B2, N2_W3, N2_B3,
  kNumParams
};
constexpr uint32_t kParamSize[kNumParams] = {
    kFeat * kWidth, kWidth, kWidth * kWidth, kWidth, kWidth * kWidth, kWidth,
    kIn1 * kWidth, kWidth, kWidth * kWidth, kWidth, kWidth * kWidth, kWidth,
    kFeat * kWidth, kWidth, kWidth * kWidth, kWidth * kWidth, kWidth * kWidth, kWidth,
    kWidth * kWidth, kWidth,
    kIn2 * kWidth, kWidth, kWidth * kWidth, kWidth, kWidth * kOut, kOut};

constexpr uint32_t param_offset(int p) {
  uint32_t o = 0;
  for (int i = 0; i < p; ++i) o += kParamSize[i];
  return o;
}
constexpr uint32_t kParamTotal = param_offset(kNumParams);

}  // namespace gnn
}  // namespace clsph

struct clsph_gnn {
  clsph_ctx* ctx = nullptr;
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

  // forward intermediates
  uint16_t* E1 = nullptr;       // e x 16 edge inputs
  float* nu = nullptr;          // n x 2
  float* degc = nullptr;        // deg / C
  float* z1[2] = {};            // e x 128, per round
  uint16_t* s1[2] = {};
  float* z2[2] = {};
  uint16_t* S[2] = {};          // n x 128: bf16(sum_j s2 / C)
  float* tmp = nullptr;         // n x 256 float scratch
  uint16_t* u1 = nullptr;       // n x 144
  float* p1 = nullptr;
  uint16_t* y1 = nullptr;
  float* p2 = nullptr;
  uint16_t* y2 = nullptr;
  uint16_t* Hb = nullptr;
  uint16_t* PA = nullptr;
  uint16_t* PB = nullptr;
  uint16_t* u2 = nullptr;       // n x 256
  float* q1 = nullptr;
  uint16_t* w1 = nullptr;
  float* q2 = nullptr;
  uint16_t* w2 = nullptr;
  float* out = nullptr;         // n x 32

  // training
  float* target = nullptr;      // n x 3
  float* loss = nullptr;        // 1
  float* dout = nullptr;        // n x 32
  float* dY = nullptr;          // n x 256
  uint16_t* dpb = nullptr;      // n x 256 bf16 scratch
  float* dS = nullptr;          // n x 128
  uint16_t* dz2 = nullptr;      // e x 128
  float* ds1 = nullptr;         // e x 128
  uint16_t* dz1 = nullptr;      // e x 128
  float* dPA = nullptr;
  float* dPB = nullptr;
  float* dH = nullptr;
  uint16_t* dPA_b = nullptr;
  uint16_t* dPB_b = nullptr;
};

namespace clsph {
namespace gnn {

// Graph + edge inputs from a state whose density stage has run.
void launch_graph(clsph_gnn* m, clsph_state* s);
void launch_edge_inputs(clsph_gnn* m, clsph_state* s);
// Forward pass of the toolbox kernels; `acc` (float4 per particle) receives
// the accelerations when non-null.
void forward(clsph_gnn* m, const DevParams* dp_host, float4* acc, float gx, float gy, float gz);
// Inference forward with the fused edge kernels (reads the state's neighbour
// lists directly; no per-edge storage, no edge-count readback).
// With `ev` (5 events) the stream gets events before edge round 1, after it,
// after the node part of round 1, after edge round 2 and at the end.
void forward_fused(clsph_gnn* m, clsph_state* st, float4* acc, float gx, float gy, float gz,
                   hipEvent_t* ev = nullptr);
// Loss against m->target and gradients into m->grad (zeroed first).
void backward(clsph_gnn* m);
void adam(clsph_gnn* m, float lr, float beta1, float beta2, float eps);
void cast_params(clsph_gnn* m);
void fill(float* x, size_t count, float v, hipStream_t s);

}  // namespace gnn
}  // namespace clsph
