// Kernels of the two-round message-passing network (gfx950): the
// neighbour-graph and edge-input builders, the fused edge-round and node-MLP
// inference kernels, activation / aggregation / loss kernels and Adam, with
// the bf16 matrix-core GEMM of nn.hip for the training products. The
// forward and backward passes are sequences of these launches on the
// context stream.
#include <hip/hip_runtime.h>

#include "gnn.h"
#include "nn.h"

namespace clsph {
namespace gnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t to_bf16(float x) {
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x));
}
__device__ inline float from_bf16(uint16_t b) {
  return __builtin_bit_cast(float, (uint32_t)b << 16);
}
// Logistic function from the hardware exp2 and reciprocal (1 ulp each): the
// activations it feeds are rounded to bf16, so IEEE division would only add
// instructions.
__device__ inline float sigmoid(float x) {
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * -1.44269504f));
}
__device__ inline float silu(float x) { return x * sigmoid(x); }
__device__ inline float dsilu(float x) {
  float s = sigmoid(x);
  return s * (1.f + x * (1.f - s));
}

// The bf16 matrix-core GEMM of the corrective surrogate (nn.hip):
// C (f32, M x N, ldc) = alpha * op(A) op(B) [+ rowscale[m] * bias[n]], or
// added to C with `add`.
inline void gemm(hipStream_t s, bool ta, bool tb, bool add, const uint16_t* A, int lda,
                 const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K, float alpha,
                 const float* bias = nullptr, const float* rowscale = nullptr) {
  nn::gemm_public(s, ta, tb, add, A, lda, B, ldb, C, ldc, M, N, K, alpha, bias, rowscale);
}

// ---------------------------------------------------------------------------
// Exclusive scan of min(count, kMaxEdges) into off[0..n], off[n] = total.
constexpr int kScanItems = 4;
constexpr int kScanBlock = 256;
constexpr int kScanChunk = kScanItems * kScanBlock;

__device__ inline uint32_t block_scan_256(uint32_t v, uint32_t* lds, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  uint32_t base = 0;
  for (int w = 0; w < wave; ++w) base += lds[w];
  total = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return base + x - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_local(const uint32_t* __restrict__ count,
                                                           uint32_t n, uint32_t* __restrict__ off,
                                                           uint32_t* __restrict__ part) {
  __shared__ uint32_t lds[4];
  uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  uint32_t v[kScanItems], sum = 0;
  for (int t = 0; t < kScanItems; ++t) {
    uint32_t i = base + t;
    v[t] = i < n ? min(count[i], kMaxEdges) : 0u;
    sum += v[t];
  }
  uint32_t total;
  uint32_t ex = block_scan_256(sum, lds, total);
  for (int t = 0; t < kScanItems; ++t) {
    uint32_t i = base + t;
    if (i < n) off[i] = ex;
    ex += v[t];
  }
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// One work-group scans the chunk totals (any count) in place.
__global__ __launch_bounds__(kScanBlock) void k_scan_parts(uint32_t* __restrict__ part,
                                                           uint32_t nparts,
                                                           uint32_t* __restrict__ total_out) {
  __shared__ uint32_t lds[4];
  uint32_t carry = 0;
  for (uint32_t b = 0; b < nparts; b += kScanBlock) {
    uint32_t i = b + threadIdx.x;
    uint32_t v = i < nparts ? part[i] : 0u;
    uint32_t total;
    uint32_t ex = block_scan_256(v, lds, total);
    if (i < nparts) part[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ off,
                                                  const uint32_t* __restrict__ part, uint32_t n) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) off[i] += part[i / kScanChunk];
}

// ---------------------------------------------------------------------------
// src / dst of every edge (dst-major, the density kernel's neighbour order).
__global__ __launch_bounds__(256) void k_graph(const uint4* __restrict__ nbr,
                                               const uint32_t* __restrict__ count,
                                               const uint32_t* __restrict__ off, uint32_t n,
                                               uint32_t* __restrict__ src,
                                               uint32_t* __restrict__ dst,
                                               float* __restrict__ degc) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t deg = min(count[i], kMaxEdges), o = off[i];
  for (uint32_t q = 0; q < deg; q += 4) {
    uint4 js = nbr[(size_t)(q >> 2) * n + i];
    uint32_t jj[4] = {js.x, js.y, js.z, js.w};
    for (uint32_t t = 0; t < 4 && q + t < deg; ++t) {
      src[o + q + t] = jj[t];
      dst[o + q + t] = i;
    }
  }
  degc[i] = (float)deg / kAggC;
}

// The 16 edge inputs of edge (i <- j) in float32 (12 used, 4 zero).
__device__ inline void edge_features(float4 xi, float4 xj, float4 vi, float4 vj, float2 ni,
                                     float2 nj, float inv_h, float f[kFeat]) {
  float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
  float q = sqrtf(dx * dx + dy * dy + dz * dz) * inv_h;
  float w = 1.f - q * q;
  f[0] = dx * inv_h, f[1] = dy * inv_h, f[2] = dz * inv_h;
  f[3] = (vi.x - vj.x) / kVelScale, f[4] = (vi.y - vj.y) / kVelScale;
  f[5] = (vi.z - vj.z) / kVelScale;
  f[6] = q, f[7] = w * w * w;
  f[8] = ni.x, f[9] = ni.y, f[10] = nj.x, f[11] = nj.y;
  f[12] = f[13] = f[14] = f[15] = 0.f;
}

// Edge inputs e (16 bf16 per edge) and node inputs nu.
__global__ __launch_bounds__(256) void k_edge_inputs(
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float* __restrict__ rho, const float* __restrict__ prs,
    const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint32_t e, uint32_t n,
    float h, float rho0, float K, uint16_t* __restrict__ E1, float* __restrict__ nu) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) {
    float2 a = node_in(rho[k], prs[k], rho0, K);
    nu[2 * k] = a.x;
    nu[2 * k + 1] = a.y;
  }
  if (k >= e) return;
  uint32_t i = dst[k], j = src[k];
  float f[kFeat];
  edge_features(pos[i], pos[j], vel[i], vel[j], node_in(rho[i], prs[i], rho0, K),
                node_in(rho[j], prs[j], rho0, K), 1.f / h, f);
  uint32_t packed[8];
  for (int t = 0; t < 8; ++t)
    packed[t] = (uint32_t)to_bf16(f[2 * t]) | ((uint32_t)to_bf16(f[2 * t + 1]) << 16);
  u32x4* o = (u32x4*)(E1 + (size_t)k * kFeat);
  o[0] = u32x4{packed[0], packed[1], packed[2], packed[3]};
  o[1] = u32x4{packed[4], packed[5], packed[6], packed[7]};
}

// ---------------------------------------------------------------------------
// Fused edge round for inference: one wave per particle i walks its
// neighbour list (the density kernel's, at most kMaxEdges) in tiles of 16
// edges and computes S_i = bf16(sum_j silu(silu(e_ij W1 + b1 [+ PA_i + PB_j])
// W2 + b2) / C) without storing per-edge values.
//
// Both layers run transposed on 16x16x32 bf16 matrix cores: z^T = W^T e^T,
// so the output's columns are edges and a lane holds, for edge l % 16, the
// features 16 ob + 4 (l / 16) + r of output block ob (r = 0..3). Layer 2's
// B operand (lane: 8 consecutive K slots of its edge) is taken straight from
// those registers by permuting the K order: slot t of k-step ks and lane
// group g = l / 16 is feature 32 ks + 16 (t / 4) + 4 g + t % 4, and the
// weights are staged in LDS in that order. Tiles accumulate per lane; the
// 16 edge lanes are summed once per particle.
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline u32x4 pack8(const float v[8]) {
  u32x4 r;
  for (int t = 0; t < 4; ++t)
    r[t] = (uint32_t)to_bf16(v[2 * t]) | ((uint32_t)to_bf16(v[2 * t + 1]) << 16);
  return r;
}

template <bool ROUND2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_edge_fused(
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float* __restrict__ rho, const float* __restrict__ prs,
    const uint4* __restrict__ nbr, const uint32_t* __restrict__ count, uint32_t n, float h,
    float rho0, float K, const uint16_t* __restrict__ W1, const float* __restrict__ b1,
    const uint16_t* __restrict__ W2, const float* __restrict__ b2,
    const uint16_t* __restrict__ PA, const uint16_t* __restrict__ PB,
    uint16_t* __restrict__ S, float* __restrict__ degc, float* __restrict__ nu) {
  __shared__ u32x4 w1s[8 * 64];       // [ob][lane]: W1^T rows 16 ob + l % 16, k 8 g + t
  __shared__ u32x4 w2s[8 * 4 * 64];   // [ob][ks][lane]: W2^T in the permuted K order
  __shared__ float bs[2 * kWidth];
  for (int idx = threadIdx.x; idx < 8 * 64; idx += 256) {
    int ob = idx >> 6, l = idx & 63, g = l >> 4, row = 16 * ob + (l & 15);
    uint32_t w[4];
    for (int t = 0; t < 4; ++t) {
      int k0 = 8 * g + 2 * t;
      uint32_t lo = k0 < kFeat ? W1[k0 * kWidth + row] : 0u;
      uint32_t hi = k0 + 1 < kFeat ? W1[(k0 + 1) * kWidth + row] : 0u;
      w[t] = lo | (hi << 16);
    }
    w1s[idx] = u32x4{w[0], w[1], w[2], w[3]};
  }
  for (int idx = threadIdx.x; idx < 8 * 4 * 64; idx += 256) {
    int ob = idx >> 8, ks = (idx >> 6) & 3, l = idx & 63, g = l >> 4, row = 16 * ob + (l & 15);
    uint32_t w[4];
    for (int t = 0; t < 4; ++t) {
      int ta = 2 * t, tb = 2 * t + 1;
      int ka = 32 * ks + 16 * (ta >> 2) + 4 * g + (ta & 3);
      int kb = 32 * ks + 16 * (tb >> 2) + 4 * g + (tb & 3);
      w[t] = (uint32_t)W2[ka * kWidth + row] | ((uint32_t)W2[kb * kWidth + row] << 16);
    }
    w2s[idx] = u32x4{w[0], w[1], w[2], w[3]};
  }
  for (int idx = threadIdx.x; idx < kWidth; idx += 256) {
    bs[idx] = b1[idx];
    bs[kWidth + idx] = b2[idx];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const uint32_t waves = gridDim.x * 4;
  const float inv_h = 1.f / h;
  for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += waves) {
    const uint32_t deg = min(count[i], kMaxEdges);
    const float4 xi = pos[i], vi = vel[i];
    const float2 ni = node_in(rho[i], prs[i], rho0, K);
    float pa[8][4];
    for (int ob = 0; ob < 8; ++ob)
      for (int r = 0; r < 4; ++r) pa[ob][r] = bs[16 * ob + 4 * g + r];
    if (ROUND2) {
      for (int ob = 0; ob < 8; ++ob) {
        uint2 v = *(const uint2*)(PA + (size_t)i * kWidth + 16 * ob + 4 * g);
        pa[ob][0] += from_bf16((uint16_t)(v.x & 0xFFFFu));
        pa[ob][1] += from_bf16((uint16_t)(v.x >> 16));
        pa[ob][2] += from_bf16((uint16_t)(v.y & 0xFFFFu));
        pa[ob][3] += from_bf16((uint16_t)(v.y >> 16));
      }
    }
    f32x4 acc[8];
    for (int ob = 0; ob < 8; ++ob) acc[ob] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t q0 = 0; q0 < deg; q0 += 16) {
      const uint32_t q = q0 + c;
      const bool valid = q < deg;
      uint32_t j = i;
      if (valid) {
        uint4 js = nbr[(size_t)(q >> 2) * n + i];
        uint32_t sel = q & 3u;
        j = sel == 0 ? js.x : sel == 1 ? js.y : sel == 2 ? js.z : js.w;
      }
      // B operand of layer 1: features 8 g .. 8 g + 7 of this lane's edge
      float f[kFeat];
      edge_features(xi, pos[j], vi, vel[j], ni, node_in(rho[j], prs[j], rho0, K), inv_h, f);
      float sel8[8];
      for (int t = 0; t < 8; ++t) sel8[t] = g == 0 ? f[t] : (g == 1 ? f[8 + t] : 0.f);
      const bf16x8 eb = __builtin_bit_cast(bf16x8, pack8(sel8));
      // layer 1 (+ bias, node terms), SiLU, packed into layer 2's B operands
      float pb[8][4];
      if (ROUND2) {
        for (int ob = 0; ob < 8; ++ob) {
          uint2 v = *(const uint2*)(PB + (size_t)j * kWidth + 16 * ob + 4 * g);
          pb[ob][0] = from_bf16((uint16_t)(v.x & 0xFFFFu));
          pb[ob][1] = from_bf16((uint16_t)(v.x >> 16));
          pb[ob][2] = from_bf16((uint16_t)(v.y & 0xFFFFu));
          pb[ob][3] = from_bf16((uint16_t)(v.y >> 16));
        }
      }
      float s1[8][4];
      for (int ob = 0; ob < 8; ++ob) {
        f32x4 z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8, w1s[ob * 64 + lane]), eb, f32x4{0.f, 0.f, 0.f, 0.f}, 0,
            0, 0);
        for (int r = 0; r < 4; ++r) {
          float v = z[r] + pa[ob][r];
          if (ROUND2) v += pb[ob][r];
          s1[ob][r] = silu(v);
        }
      }
      bf16x8 bop[4];
      for (int ks = 0; ks < 4; ++ks) {
        float v8[8];
        for (int t = 0; t < 8; ++t) v8[t] = s1[2 * ks + (t >> 2)][t & 3];
        bop[ks] = __builtin_bit_cast(bf16x8, pack8(v8));
      }
      // layer 2, SiLU, masked accumulation over the tile's edges
      for (int ob = 0; ob < 8; ++ob) {
        f32x4 y = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 4; ++ks)
          y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(bf16x8, w2s[(ob * 4 + ks) * 64 + lane]), bop[ks], y, 0, 0, 0);
        for (int r = 0; r < 4; ++r) {
          float v = silu(y[r] + bs[kWidth + 16 * ob + 4 * g + r]);
          acc[ob][r] += valid ? v : 0.f;
        }
        __builtin_amdgcn_sched_barrier(0);  // one output block's reads at a time
      }
    }
    // sum over the 16 edge lanes of each group
    for (int ob = 0; ob < 8; ++ob)
      for (int r = 0; r < 4; ++r) {
        float v = acc[ob][r];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        acc[ob][r] = v;
      }
    // lane c < 8 of group g writes features 16 c + 4 g .. + 3
    for (int ob = 0; ob < 8; ++ob)
      if (c == ob) {
        uint32_t lo = (uint32_t)to_bf16(acc[ob][0] * (1.f / kAggC)) |
                      ((uint32_t)to_bf16(acc[ob][1] * (1.f / kAggC)) << 16);
        uint32_t hi = (uint32_t)to_bf16(acc[ob][2] * (1.f / kAggC)) |
                      ((uint32_t)to_bf16(acc[ob][3] * (1.f / kAggC)) << 16);
        *(uint2*)(S + (size_t)i * kWidth + 16 * ob + 4 * g) = make_uint2(lo, hi);
      }
    if (lane == 0) {
      degc[i] = (float)deg / kAggC;
      nu[2 * i] = ni.x;
      nu[2 * i + 1] = ni.y;
    }
  }
}

// ---------------------------------------------------------------------------
// Fused node MLPs for inference. A wave handles tiles of 16 particles; each
// layer runs transposed on 16x16x32 bf16 matrix cores (outputs as rows,
// particles as columns), chaining in registers as the edge kernel does. B
// operands come either from memory in the natural K order (slot t of k-step
// ks, lane group g: feature 32 ks + 8 g + t) or from the previous layer's
// registers in the permuted order (32 ks + 16 (t / 4) + 4 g + t % 4); each
// layer's weights are staged once per work-group in LDS in the matching
// order. Rounding points are those of the GEMM path (bf16 layer inputs,
// float32 accumulation).
constexpr int kNodeTile = 16;

// Stages W rows for k-steps [ks0, ks0 + nks) of a layer with `ks_total`
// k-steps and `nob` 16-wide output blocks: dst[(ob * ks_total + ks) * 64 +
// lane] holds W[k][16 ob + lane % 16] for the lane's 8 K slots; rows at or
// beyond `kvalid` are zero. `krow0` is the W row of the first staged k-step.
__device__ __attribute__((always_inline)) inline void stage_layer(u32x4* dst, int ks_total, int ks0, int nks, int nob,
                                   const uint16_t* __restrict__ Wm, int ld, int krow0, int kvalid,
                                   bool perm) {
  const int total = nob * nks * 64;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    int l = idx & 63, ks = (idx >> 6) % nks, ob = (idx >> 6) / nks, g = l >> 4;
    int col = 16 * ob + (l & 15);
    uint32_t w[4];
    for (int t2 = 0; t2 < 4; ++t2) {
      uint32_t half[2];
      for (int h = 0; h < 2; ++h) {
        int t = 2 * t2 + h;
        int k = krow0 + 32 * ks + (perm ? 16 * (t >> 2) + 4 * g + (t & 3) : 8 * g + t);
        half[h] = k < kvalid ? Wm[(size_t)k * ld + col] : 0u;
      }
      w[t2] = half[0] | (half[1] << 16);
    }
    dst[(ob * ks_total + ks0 + ks) * 64 + l] = u32x4{w[0], w[1], w[2], w[3]};
  }
}

template <int NOB, int NKS>
__device__ __attribute__((always_inline)) inline void mfma_layer(const u32x4* __restrict__ w, const bf16x8* bop, f32x4* y) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int ob = 0; ob < NOB; ++ob) {
    f32x4 a = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
      a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, w[(ob * NKS + ks) * 64 + lane]),
                                                   bop[ks], a, 0, 0, 0);
    y[ob] = a;
    // keep each block's weight reads next to its products: hoisting all of
    // a layer's LDS reads would need 4 NOB NKS registers more
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Permuted-order B operands (4 k-steps, 128 features) from a 128-wide
// layer's registers, optionally through SiLU.
__device__ __attribute__((always_inline)) inline void to_operands(const f32x4* y, bool act, bf16x8* bop) {
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    float v8[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      float v = y[2 * ks + (t >> 2)][t & 3];
      v8[t] = act ? silu(v) : v;
    }
    bop[ks] = __builtin_bit_cast(bf16x8, pack8(v8));
  }
}

// Natural-order B operands: features 32 ks + 8 g .. + 7 of a bf16 row.
__device__ __attribute__((always_inline)) inline void load_operands(const uint16_t* __restrict__ row, bool valid, int nks,
                                     bf16x8* bop) {
  const int g = (threadIdx.x & 63) >> 4;
  for (int ks = 0; ks < nks; ++ks) {
    u32x4 v = {0u, 0u, 0u, 0u};
    if (valid) v = *(const u32x4*)(row + 32 * ks + 8 * g);
    bop[ks] = __builtin_bit_cast(bf16x8, v);
  }
}

// Adds bias[16 ob + 4 g + r] (+ scale * bias2) to a 128-wide layer.
__device__ __attribute__((always_inline)) inline void add_bias(f32x4* y, int nob, const float* __restrict__ b,
                                const float* __restrict__ b2, float scale) {
  const int g = (threadIdx.x & 63) >> 4;
  for (int ob = 0; ob < nob; ++ob)
    for (int r = 0; r < 4; ++r) {
      int f = 16 * ob + 4 * g + r;
      y[ob][r] += b[f] + (b2 ? scale * b2[f] : 0.f);
    }
}

// Writes features 16 ob + 4 g .. + 3 of a particle's bf16 row.
__device__ __attribute__((always_inline)) inline void store_row_bf16(uint16_t* __restrict__ row, const f32x4* y) {
  const int g = (threadIdx.x & 63) >> 4;
  for (int ob = 0; ob < 8; ++ob) {
    uint32_t lo = (uint32_t)to_bf16(y[ob][0]) | ((uint32_t)to_bf16(y[ob][1]) << 16);
    uint32_t hi = (uint32_t)to_bf16(y[ob][2]) | ((uint32_t)to_bf16(y[ob][3]) << 16);
    *(uint2*)(row + 16 * ob + 4 * g) = make_uint2(lo, hi);
  }
}

// Round-1 node part: M1 = S0 W3 + (deg / C) b3; H = MLP([M1, nu, 0]); Hb.
__global__ __launch_bounds__(512) void k_node1(const uint16_t* __restrict__ S0,
                                               const float* __restrict__ degc,
                                               const float* __restrict__ nu, uint32_t n,
                                               const uint16_t* __restrict__ wb,
                                               const float* __restrict__ params,
                                               uint16_t* __restrict__ Hb) {
  extern __shared__ u32x4 lds[];
  u32x4* w3 = lds;                 // 8 x 4
  u32x4* wa = w3 + 8 * 4 * 64;     // 8 x 5
  u32x4* wb2 = wa + 8 * 5 * 64;    // 8 x 4
  u32x4* wc = wb2 + 8 * 4 * 64;    // 8 x 4
  float* bias = (float*)(wc + 8 * 4 * 64);  // b3, ba, bb, bc
  stage_layer(w3, 4, 0, 4, 8, wb + param_offset(E1_W3), kWidth, 0, kWidth, false);
  stage_layer(wa, 5, 0, 4, 8, wb + param_offset(N1_W1), kWidth, 0, kIn1, true);
  stage_layer(wa, 5, 4, 1, 8, wb + param_offset(N1_W1), kWidth, kWidth, kWidth + 2, false);
  stage_layer(wb2, 4, 0, 4, 8, wb + param_offset(N1_W2), kWidth, 0, kWidth, true);
  stage_layer(wc, 4, 0, 4, 8, wb + param_offset(N1_W3), kWidth, 0, kWidth, true);
  for (int k = threadIdx.x; k < kWidth; k += blockDim.x) {
    bias[k] = params[param_offset(E1_B3) + k];
    bias[kWidth + k] = params[param_offset(N1_B1) + k];
    bias[2 * kWidth + k] = params[param_offset(N1_B2) + k];
    bias[3 * kWidth + k] = params[param_offset(N1_B3) + k];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const uint32_t tiles = (n + kNodeTile - 1) / kNodeTile;
  for (uint32_t tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < tiles;
       tile += waves) {
    const uint32_t i = tile * kNodeTile + c;
    const bool valid = i < n;
    bf16x8 bop[5];
    f32x4 y[8];
    load_operands(S0 + (size_t)(valid ? i : 0) * kWidth, valid, 4, bop);
    mfma_layer<8, 4>(w3, bop, y);
    {
      const float dc = valid ? degc[i] : 0.f;  // M1 = S0 W3 + (deg / C) b3
      for (int ob = 0; ob < 8; ++ob)
        for (int r = 0; r < 4; ++r) y[ob][r] += dc * bias[16 * ob + 4 * g + r];
    }
    to_operands(y, false, bop);
    {
      float v8[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (g == 0 && valid) {
        v8[0] = nu[2 * i];
        v8[1] = nu[2 * i + 1];
      }
      bop[4] = __builtin_bit_cast(bf16x8, pack8(v8));
    }
    mfma_layer<8, 5>(wa, bop, y);
    add_bias(y, 8, bias + kWidth, nullptr, 0.f);
    to_operands(y, true, bop);
    mfma_layer<8, 4>(wb2, bop, y);
    add_bias(y, 8, bias + 2 * kWidth, nullptr, 0.f);
    to_operands(y, true, bop);
    mfma_layer<8, 4>(wc, bop, y);
    add_bias(y, 8, bias + 3 * kWidth, nullptr, 0.f);
    if (valid) store_row_bf16(Hb + (size_t)i * kWidth, y);
  }
}

// Round-2 node terms: PA = bf16(Hb Wa), PB = bf16(Hb Wb).
__global__ __launch_bounds__(512) void k_node_terms(const uint16_t* __restrict__ Hb, uint32_t n,
                                                    const uint16_t* __restrict__ wb,
                                                    uint16_t* __restrict__ PA,
                                                    uint16_t* __restrict__ PB) {
  extern __shared__ u32x4 lds[];
  u32x4* wa = lds;
  u32x4* wbb = wa + 8 * 4 * 64;
  stage_layer(wa, 4, 0, 4, 8, wb + param_offset(E2_WA), kWidth, 0, kWidth, false);
  stage_layer(wbb, 4, 0, 4, 8, wb + param_offset(E2_WB), kWidth, 0, kWidth, false);
  __syncthreads();
  const int c = threadIdx.x & 15;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const uint32_t tiles = (n + kNodeTile - 1) / kNodeTile;
  for (uint32_t tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < tiles;
       tile += waves) {
    const uint32_t i = tile * kNodeTile + c;
    const bool valid = i < n;
    bf16x8 bop[4];
    f32x4 y[8];
    load_operands(Hb + (size_t)(valid ? i : 0) * kWidth, valid, 4, bop);
    mfma_layer<8, 4>(wa, bop, y);
    if (valid) store_row_bf16(PA + (size_t)i * kWidth, y);
    mfma_layer<8, 4>(wbb, bop, y);
    if (valid) store_row_bf16(PB + (size_t)i * kWidth, y);
  }
}

// Round-2 node part: M2 = S1 W3' + (deg / C) b3'; out = MLP([H, M2]);
// acc = g + A out[0:3]. Writes out (32 floats per particle) as well.
__global__ __launch_bounds__(512) void k_node2(const uint16_t* __restrict__ S1,
                                               const uint16_t* __restrict__ Hb,
                                               const float* __restrict__ degc, uint32_t n,
                                               const uint16_t* __restrict__ wb,
                                               const float* __restrict__ params,
                                               float* __restrict__ out, float4* __restrict__ acc,
                                               float gx, float gy, float gz) {
  extern __shared__ u32x4 lds[];
  u32x4* w3 = lds;                 // 8 x 4
  u32x4* wa = w3 + 8 * 4 * 64;     // 8 x 8
  u32x4* wb2 = wa + 8 * 8 * 64;    // 8 x 4
  u32x4* wc = wb2 + 8 * 4 * 64;    // 2 x 4
  float* bias = (float*)(wc + 2 * 4 * 64);  // b3, ba, bb (128 each), bc (32)
  stage_layer(w3, 4, 0, 4, 8, wb + param_offset(E2_W3), kWidth, 0, kWidth, false);
  stage_layer(wa, 8, 0, 4, 8, wb + param_offset(N2_W1), kWidth, 0, kWidth, false);
  stage_layer(wa, 8, 4, 4, 8, wb + param_offset(N2_W1), kWidth, kWidth, kIn2, true);
  stage_layer(wb2, 4, 0, 4, 8, wb + param_offset(N2_W2), kWidth, 0, kWidth, true);
  stage_layer(wc, 4, 0, 4, 2, wb + param_offset(N2_W3), kOut, 0, kWidth, true);
  for (int k = threadIdx.x; k < kWidth; k += blockDim.x) {
    bias[k] = params[param_offset(E2_B3) + k];
    bias[kWidth + k] = params[param_offset(N2_B1) + k];
    bias[2 * kWidth + k] = params[param_offset(N2_B2) + k];
    if (k < kOut) bias[3 * kWidth + k] = params[param_offset(N2_B3) + k];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const uint32_t tiles = (n + kNodeTile - 1) / kNodeTile;
  for (uint32_t tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < tiles;
       tile += waves) {
    const uint32_t i = tile * kNodeTile + c;
    const bool valid = i < n;
    bf16x8 bop[8];
    f32x4 y[8];
    load_operands(S1 + (size_t)(valid ? i : 0) * kWidth, valid, 4, bop);
    mfma_layer<8, 4>(w3, bop, y);
    {
      const float dc = valid ? degc[i] : 0.f;
      for (int ob = 0; ob < 8; ++ob)
        for (int r = 0; r < 4; ++r) y[ob][r] += dc * bias[16 * ob + 4 * g + r];
    }
    load_operands(Hb + (size_t)(valid ? i : 0) * kWidth, valid, 4, bop);
    to_operands(y, false, bop + 4);
    mfma_layer<8, 8>(wa, bop, y);
    add_bias(y, 8, bias + kWidth, nullptr, 0.f);
    to_operands(y, true, bop);
    mfma_layer<8, 4>(wb2, bop, y);
    add_bias(y, 8, bias + 2 * kWidth, nullptr, 0.f);
    to_operands(y, true, bop);
    f32x4 o[2];
    mfma_layer<2, 4>(wc, bop, o);
    for (int ob = 0; ob < 2; ++ob)
      for (int r = 0; r < 4; ++r) o[ob][r] += bias[3 * kWidth + 16 * ob + 4 * g + r];
    if (valid) {
      float* orow = out + (size_t)i * kOut;
      for (int ob = 0; ob < 2; ++ob)
        *(float4*)(orow + 16 * ob + 4 * g) = make_float4(o[ob][0], o[ob][1], o[ob][2], o[ob][3]);
      if (g == 0)
        acc[i] = make_float4(gx + kAccScale * o[0][0], gy + kAccScale * o[0][1],
                             gz + kAccScale * o[0][2], 0.f);
    }
  }
}

// ---------------------------------------------------------------------------
// Elementwise and row kernels (grid-stride loops).
__global__ void k_silu_bf16(const float* __restrict__ z, uint16_t* __restrict__ s, size_t count) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x)
    s[k] = to_bf16(silu(z[k]));
}

__global__ void k_cast_2d(const float* __restrict__ x, int ldx, uint16_t* __restrict__ y, int ldy,
                          size_t rows, int cols) {
  size_t count = rows * cols;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t r = k / cols;
    int c = (int)(k % cols);
    y[r * ldy + c] = to_bf16(x[r * ldx + c]);
  }
}

__global__ void k_copy_2d(const float* __restrict__ x, int ldx, float* __restrict__ y, int ldy,
                          size_t rows, int cols) {
  size_t count = rows * cols;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t r = k / cols;
    int c = (int)(k % cols);
    y[r * ldy + c] = x[r * ldx + c];
  }
}

// dz = bf16(dy * silu'(p)); dy has row stride ldy, p and dz are dense.
__global__ void k_dsilu(const float* __restrict__ dy, int ldy, const float* __restrict__ p,
                        uint16_t* __restrict__ dz, size_t rows, int cols) {
  size_t count = rows * cols;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t r = k / cols;
    int c = (int)(k % cols);
    dz[k] = to_bf16(dy[r * ldy + c] * dsilu(p[k]));
  }
}

// dz2[e] = bf16(dS[dst[e]] * silu'(z2[e])), 128 features per edge.
__global__ void k_edge_dz(const float* __restrict__ dS, const float* __restrict__ z2,
                          const uint32_t* __restrict__ dst, uint16_t* __restrict__ dz,
                          size_t e) {
  size_t count = e * kWidth;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t ed = k / kWidth;
    int c = (int)(k % kWidth);
    dz[k] = to_bf16(dS[(size_t)dst[ed] * kWidth + c] * dsilu(z2[k]));
  }
}

// z1[e] += PA[dst[e]] + PB[src[e]] (bf16 node terms of round 2).
__global__ void k_edge_add(float* __restrict__ z1, const uint16_t* __restrict__ PA,
                           const uint16_t* __restrict__ PB, const uint32_t* __restrict__ src,
                           const uint32_t* __restrict__ dst, size_t e) {
  size_t count = e * kWidth;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t ed = k / kWidth;
    int c = (int)(k % kWidth);
    float z = z1[k] + from_bf16(PA[(size_t)dst[ed] * kWidth + c]);
    z1[k] = z + from_bf16(PB[(size_t)src[ed] * kWidth + c]);
  }
}

// S[i] = bf16(scale * sum over i's edges of silu(z2)); one wave per node.
__global__ __launch_bounds__(256) void k_segsum_silu(const float* __restrict__ z2,
                                                     const uint32_t* __restrict__ off, uint32_t n,
                                                     float scale, uint16_t* __restrict__ S) {
  uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (i >= n) return;
  float a0 = 0.f, a1 = 0.f;
  for (uint32_t k = off[i]; k < off[i + 1]; ++k) {
    float2 z = *(const float2*)(z2 + (size_t)k * kWidth + 2 * lane);
    a0 += silu(z.x);
    a1 += silu(z.y);
  }
  S[(size_t)i * kWidth + 2 * lane] = to_bf16(a0 * scale);
  S[(size_t)i * kWidth + 2 * lane + 1] = to_bf16(a1 * scale);
}

// out[i] = sum over i's edges of x (bf16 rows, 128 wide); one wave per node.
__global__ __launch_bounds__(256) void k_segsum_rows(const uint16_t* __restrict__ x,
                                                     const uint32_t* __restrict__ off, uint32_t n,
                                                     float* __restrict__ out) {
  uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (i >= n) return;
  float a0 = 0.f, a1 = 0.f;
  for (uint32_t k = off[i]; k < off[i + 1]; ++k) {
    uint32_t v = *(const uint32_t*)(x + (size_t)k * kWidth + 2 * lane);
    a0 += from_bf16((uint16_t)(v & 0xFFFFu));
    a1 += from_bf16((uint16_t)(v >> 16));
  }
  out[(size_t)i * kWidth + 2 * lane] = a0;
  out[(size_t)i * kWidth + 2 * lane + 1] = a1;
}

// out[src[e]] += x[e] (atomics; 128 wide).
__global__ void k_scatter_rows(const uint16_t* __restrict__ x, const uint32_t* __restrict__ src,
                               size_t e, float* __restrict__ out) {
  size_t count = e * (kWidth / 2);
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t ed = k / (kWidth / 2);
    int c = 2 * (int)(k % (kWidth / 2));
    uint32_t v = *(const uint32_t*)(x + ed * kWidth + c);
    float* o = out + (size_t)src[ed] * kWidth + c;
    atomicAdd(o, from_bf16((uint16_t)(v & 0xFFFFu)));
    atomicAdd(o + 1, from_bf16((uint16_t)(v >> 16)));
  }
}

// Column sums (optionally row-weighted) atomically added into out[0..cols).
// One thread per column over a chunk of 512 rows.
template <typename T>
__device__ inline float load_val(const T* p);
template <>
__device__ inline float load_val<float>(const float* p) { return *p; }
template <>
__device__ inline float load_val<uint16_t>(const uint16_t* p) { return from_bf16(*p); }

template <typename T>
__global__ void k_colsum(const T* __restrict__ x, int ldx, size_t rows, int cols,
                         const float* __restrict__ rowscale, float* __restrict__ out) {
  int c = threadIdx.x;
  if (c >= cols) return;
  size_t r0 = (size_t)blockIdx.x * 512, r1 = r0 + 512 < rows ? r0 + 512 : rows;
  float a = 0.f;
  for (size_t r = r0; r < r1; ++r) {
    float v = load_val<T>(x + r * ldx + c);
    a += rowscale ? v * rowscale[r] : v;
  }
  atomicAdd(out + c, a);
}

// u1 = bf16([M1, nu, 0...]) (144 wide).
__global__ void k_pack_u1(const float* __restrict__ M, const float* __restrict__ nu, uint32_t n,
                          uint16_t* __restrict__ u1) {
  size_t count = (size_t)n * kIn1;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t i = k / kIn1;
    int c = (int)(k % kIn1);
    float v = c < kWidth ? M[i * kWidth + c] : (c < kWidth + 2 ? nu[2 * i + c - kWidth] : 0.f);
    u1[k] = to_bf16(v);
  }
}

// u2 = [H (bf16), bf16(M2)] (256 wide).
__global__ void k_pack_u2(const uint16_t* __restrict__ Hb, const float* __restrict__ M2,
                          uint32_t n, uint16_t* __restrict__ u2) {
  size_t count = (size_t)n * kIn2;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t i = k / kIn2;
    int c = (int)(k % kIn2);
    u2[k] = c < kWidth ? Hb[i * kWidth + c] : to_bf16(M2[i * kWidth + c - kWidth]);
  }
}

// Loss = sum over particles and the 3 used outputs of (out - target)^2 / n;
// dout = d loss / d out.
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ out,
                                              const float* __restrict__ target, uint32_t n,
                                              float* __restrict__ dout, float* __restrict__ loss) {
  __shared__ float red[4];
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (i < n) {
    for (int c = 0; c < kOut; ++c) {
      float d = c < 3 ? out[(size_t)i * kOut + c] - target[3 * (size_t)i + c] : 0.f;
      dout[(size_t)i * kOut + c] = 2.f * d / (float)n;
      l += d * d;
    }
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) / (float)n);
}

__global__ void k_fill(float* __restrict__ x, size_t count, float v) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x)
    x[k] = v;
}

// Adam (Kingma & Ba 2015) with bias correction; refreshes the bf16 copy.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t count, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, uint16_t* __restrict__ wb) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    float gk = g[k];
    float mk = b1 * m[k] + (1.f - b1) * gk;
    float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    float pk = p[k] - lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
    p[k] = pk;
    wb[k] = to_bf16(pk);
  }
}

__global__ void k_cast(const float* __restrict__ x, uint16_t* __restrict__ y, size_t count) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x)
    y[k] = to_bf16(x[k]);
}

// acc[i] = g + A * out[i, 0:3]
__global__ void k_write_acc(const float* __restrict__ out, uint32_t n, float gx, float gy,
                            float gz, float4* __restrict__ acc) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* o = out + (size_t)i * kOut;
  acc[i] = make_float4(gx + kAccScale * o[0], gy + kAccScale * o[1], gz + kAccScale * o[2], 0.f);
}

// ---------------------------------------------------------------------------
inline dim3 grid_for(size_t count, size_t per_block = 256) {
  size_t b = (count + per_block - 1) / per_block;
  if (b > 65536) b = 65536;
  if (b == 0) b = 1;
  return dim3((uint32_t)b);
}

void fill(float* x, size_t count, float v, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_fill, grid_for(count), dim3(256), 0, s, x, count, v);
}

void launch_graph(clsph_gnn* m, clsph_state* st) {
  hipStream_t s = st->ctx->stream;
  uint32_t n = st->n;
  m->n = n;
  uint32_t chunks = (n + kScanChunk - 1) / kScanChunk;
  if (!chunks) chunks = 1;
  hipLaunchKernelGGL(k_scan_local, dim3(chunks), dim3(kScanBlock), 0, s, st->nbr_count, n, m->off,
                     m->scan_part);
  hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kScanBlock), 0, s, m->scan_part, chunks,
                     m->off + n);
  hipLaunchKernelGGL(k_scan_add, grid_for(n), dim3(256), 0, s, m->off, m->scan_part, n);
}

void launch_edge_inputs(clsph_gnn* m, clsph_state* st) {
  hipStream_t s = st->ctx->stream;
  uint32_t n = m->n, e = m->e;
  const DevParams& d = st->host_dps[0];
  hipLaunchKernelGGL(k_graph, grid_for(n), dim3(256), 0, s, st->nbr, st->nbr_count, m->off, n,
                     m->src, m->dst, m->degc);
  uint32_t cnt = e > n ? e : n;
  hipLaunchKernelGGL(k_edge_inputs, grid_for(cnt), dim3(256), 0, s, st->pos_s, st->vel_s,
                     st->rho_s, st->prs_s, m->src, m->dst, e, n, d.p.h, d.p.fluid_density,
                     d.p.K, m->E1, m->nu);
}

static inline const float* P(const clsph_gnn* m, int id) { return m->params + param_offset(id); }
static inline const uint16_t* W(const clsph_gnn* m, int id) { return m->wb + param_offset(id); }
static inline float* G(clsph_gnn* m, int id) { return m->grad + param_offset(id); }

// Edge MLP of one round: z1 (already holding e W1 + b1 [+ node terms]) ->
// s1 -> z2 -> S = bf16(sum silu(z2) / C).
static void edge_tail(clsph_gnn* m, int r, int w2, int b2, hipStream_t s) {
  size_t ew = (size_t)m->e * kWidth;
  hipLaunchKernelGGL(k_silu_bf16, grid_for(ew), dim3(256), 0, s, m->z1[r], m->s1[r], ew);
  gemm(s, false, false, false, m->s1[r], kWidth, W(m, w2), kWidth, m->z2[r], kWidth, (int)m->e,
       kWidth, kWidth, 1.f, P(m, b2));
  hipLaunchKernelGGL(k_segsum_silu, dim3((m->n + 3) / 4), dim3(256), 0, s, m->z2[r], m->off,
                     m->n, 1.f / kAggC, m->S[r]);
}

// Three-layer SiLU MLP over n rows: input u (bf16, kin wide) ->
// pre1/act1 -> pre2/act2 -> y (f32, kout wide).
static void mlp3(clsph_gnn* m, const uint16_t* u, int kin, int w1, float* pre1, uint16_t* act1,
                 float* pre2, uint16_t* act2, float* y, int kout, hipStream_t s) {
  int n = (int)m->n;
  size_t nw = (size_t)n * kWidth;
  gemm(s, false, false, false, u, kin, W(m, w1), kWidth, pre1, kWidth, n, kWidth, kin, 1.f,
       P(m, w1 + 1));
  hipLaunchKernelGGL(k_silu_bf16, grid_for(nw), dim3(256), 0, s, pre1, act1, nw);
  gemm(s, false, false, false, act1, kWidth, W(m, w1 + 2), kWidth, pre2, kWidth, n, kWidth,
       kWidth, 1.f, P(m, w1 + 3));
  hipLaunchKernelGGL(k_silu_bf16, grid_for(nw), dim3(256), 0, s, pre2, act2, nw);
  gemm(s, false, false, false, act2, kWidth, W(m, w1 + 4), kout, y, kout, n, kout, kWidth, 1.f,
       P(m, w1 + 5));
}

// Node part of round 1: M1 from S[0], node MLP 1 -> H (bf16), and the
// round-2 node terms PA = bf16(H Wa), PB = bf16(H Wb).
static void node_round1(clsph_gnn* m, hipStream_t s) {
  int n = (int)m->n;
  size_t nw = (size_t)n * kWidth;
  gemm(s, false, false, false, m->S[0], kWidth, W(m, E1_W3), kWidth, m->tmp, kWidth, n, kWidth,
       kWidth, 1.f, P(m, E1_B3), m->degc);
  hipLaunchKernelGGL(k_pack_u1, grid_for((size_t)n * kIn1), dim3(256), 0, s, m->tmp, m->nu,
                     m->n, m->u1);
  mlp3(m, m->u1, kIn1, N1_W1, m->p1, m->y1, m->p2, m->y2, m->tmp, kWidth, s);
  hipLaunchKernelGGL(k_cast, grid_for(nw), dim3(256), 0, s, m->tmp, m->Hb, nw);
  gemm(s, false, false, false, m->Hb, kWidth, W(m, E2_WA), kWidth, m->tmp, kWidth, n, kWidth,
       kWidth, 1.f);
  hipLaunchKernelGGL(k_cast, grid_for(nw), dim3(256), 0, s, m->tmp, m->PA, nw);
  gemm(s, false, false, false, m->Hb, kWidth, W(m, E2_WB), kWidth, m->tmp, kWidth, n, kWidth,
       kWidth, 1.f);
  hipLaunchKernelGGL(k_cast, grid_for(nw), dim3(256), 0, s, m->tmp, m->PB, nw);
}

// Node part of round 2: M2 from S[1], node MLP 2 -> out, accelerations.
static void node_round2(clsph_gnn* m, float4* acc, float gx, float gy, float gz, hipStream_t s) {
  int n = (int)m->n;
  gemm(s, false, false, false, m->S[1], kWidth, W(m, E2_W3), kWidth, m->tmp, kWidth, n, kWidth,
       kWidth, 1.f, P(m, E2_B3), m->degc);
  hipLaunchKernelGGL(k_pack_u2, grid_for((size_t)n * kIn2), dim3(256), 0, s, m->Hb, m->tmp,
                     m->n, m->u2);
  mlp3(m, m->u2, kIn2, N2_W1, m->q1, m->w1, m->q2, m->w2, m->out, kOut, s);
  if (acc)
    hipLaunchKernelGGL(k_write_acc, grid_for(n), dim3(256), 0, s, m->out, m->n, gx, gy, gz, acc);
}

void forward(clsph_gnn* m, const DevParams* dp, float4* acc, float gx, float gy, float gz) {
  (void)dp;
  hipStream_t s = m->ctx->stream;
  int e = (int)m->e;
  // round 1 edges
  gemm(s, false, false, false, m->E1, kFeat, W(m, E1_W1), kWidth, m->z1[0], kWidth, e, kWidth,
       kFeat, 1.f, P(m, E1_B1));
  edge_tail(m, 0, E1_W2, E1_B2, s);
  node_round1(m, s);
  // round 2 edges
  gemm(s, false, false, false, m->E1, kFeat, W(m, E2_W1), kWidth, m->z1[1], kWidth, e, kWidth,
       kFeat, 1.f, P(m, E2_B1));
  size_t ew = (size_t)e * kWidth;
  hipLaunchKernelGGL(k_edge_add, grid_for(ew), dim3(256), 0, s, m->z1[1], m->PA, m->PB, m->src,
                     m->dst, (size_t)e);
  edge_tail(m, 1, E2_W2, E2_B2, s);
  node_round2(m, acc, gx, gy, gz, s);
}

void forward_fused(clsph_gnn* m, clsph_state* st, float4* acc, float gx, float gy, float gz,
                   hipEvent_t* ev) {
  hipStream_t s = m->ctx->stream;
  uint32_t n = st->n;
  m->n = n;
  const DevParams& d = st->host_dps[0];
  // persistent grid: the work-groups that fit at once (three per compute
  // unit at the kernels' register and LDS use) stage the weights once each
  uint32_t blocks = (uint32_t)st->ctx->cus * 3;
  if (blocks * 4 > n) blocks = (n + 3) / 4;
  if (blocks == 0) blocks = 1;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL((k_edge_fused<false>), dim3(blocks), dim3(256), 0, s, st->pos_s, st->vel_s,
                     st->rho_s, st->prs_s, st->nbr, st->nbr_count, n, d.p.h, d.p.fluid_density,
                     d.p.K, W(m, E1_W1), P(m, E1_B1), W(m, E1_W2), P(m, E1_B2),
                     (const uint16_t*)nullptr, (const uint16_t*)nullptr, m->S[0], m->degc, m->nu);
  if (ev) (void)hipEventRecord(ev[1], s);
  // node part of round 1 (fused)
  constexpr size_t kLdsNode1 = (size_t)(8 * 4 + 8 * 5 + 8 * 4 + 8 * 4) * 64 * 16 + 4 * kWidth * 4;
  constexpr size_t kLdsTerms = (size_t)(8 * 4 + 8 * 4) * 64 * 16;
  constexpr size_t kLdsNode2 =
      (size_t)(8 * 4 + 8 * 8 + 8 * 4 + 2 * 4) * 64 * 16 + (3 * kWidth + kOut) * 4;
  static bool lds_set = false;
  if (!lds_set) {
    (void)hipFuncSetAttribute((const void*)k_node1, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kLdsNode1);
    (void)hipFuncSetAttribute((const void*)k_node_terms,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsTerms);
    (void)hipFuncSetAttribute((const void*)k_node2, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kLdsNode2);
    lds_set = true;
  }
  uint32_t tiles = (n + kNodeTile - 1) / kNodeTile;
  uint32_t nblocks = (uint32_t)st->ctx->cus;
  if (nblocks * 8 > tiles) nblocks = (tiles + 7) / 8;
  if (nblocks == 0) nblocks = 1;
  hipLaunchKernelGGL(k_node1, dim3(nblocks), dim3(512), kLdsNode1, s, m->S[0], m->degc, m->nu, n,
                     m->wb, m->params, m->Hb);
  hipLaunchKernelGGL(k_node_terms, dim3(nblocks), dim3(512), kLdsTerms, s, m->Hb, n, m->wb, m->PA,
                     m->PB);
  if (ev) (void)hipEventRecord(ev[2], s);
  hipLaunchKernelGGL((k_edge_fused<true>), dim3(blocks), dim3(256), 0, s, st->pos_s, st->vel_s,
                     st->rho_s, st->prs_s, st->nbr, st->nbr_count, n, d.p.h, d.p.fluid_density,
                     d.p.K, W(m, E2_W1), P(m, E2_B1), W(m, E2_W2), P(m, E2_B2), m->PA, m->PB,
                     m->S[1], m->degc, m->nu);
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL(k_node2, dim3(nblocks), dim3(512), kLdsNode2, s, m->S[1], m->Hb, m->degc, n,
                     m->wb, m->params, m->out, acc, gx, gy, gz);
  if (ev) (void)hipEventRecord(ev[4], s);
}

// Backward of a three-layer MLP given dy (f32, kout wide, in dY-compatible
// storage `dy` with stride ldy). Leaves d(input) in m->dY (f32, kin_used wide).
static void mlp3_back(clsph_gnn* m, int w1, const uint16_t* u, int kin, const float* pre1,
                      const uint16_t* act1, const float* pre2, const uint16_t* act2,
                      const uint16_t* dy_b, const float* dy, int ldy, int kout, int kin_used,
                      hipStream_t s) {
  int n = (int)m->n;
  // layer 3
  gemm(s, true, false, true, act2, kWidth, dy_b, kout, G(m, w1 + 4), kout, kWidth, kout, n, 1.f);
  hipLaunchKernelGGL((k_colsum<float>), dim3((n + 511) / 512), dim3(kout), 0, s, dy, ldy,
                     (size_t)n, kout, nullptr, G(m, w1 + 5));
  gemm(s, false, true, false, dy_b, kout, W(m, w1 + 4), kout, m->dY, kWidth, n, kWidth, kout, 1.f);
  hipLaunchKernelGGL(k_dsilu, grid_for((size_t)n * kWidth), dim3(256), 0, s, m->dY, kWidth, pre2,
                     m->dpb, (size_t)n, kWidth);
  // layer 2
  gemm(s, true, false, true, act1, kWidth, m->dpb, kWidth, G(m, w1 + 2), kWidth, kWidth, kWidth, n,
       1.f);
  hipLaunchKernelGGL((k_colsum<uint16_t>), dim3((n + 511) / 512), dim3(kWidth), 0, s, m->dpb,
                     kWidth, (size_t)n, kWidth, nullptr, G(m, w1 + 3));
  gemm(s, false, true, false, m->dpb, kWidth, W(m, w1 + 2), kWidth, m->dY, kWidth, n, kWidth,
       kWidth, 1.f);
  hipLaunchKernelGGL(k_dsilu, grid_for((size_t)n * kWidth), dim3(256), 0, s, m->dY, kWidth, pre1,
                     m->dpb, (size_t)n, kWidth);
  // layer 1
  gemm(s, true, false, true, u, kin, m->dpb, kWidth, G(m, w1), kWidth, kin, kWidth, n, 1.f);
  hipLaunchKernelGGL((k_colsum<uint16_t>), dim3((n + 511) / 512), dim3(kWidth), 0, s, m->dpb,
                     kWidth, (size_t)n, kWidth, nullptr, G(m, w1 + 1));
  gemm(s, false, true, false, m->dpb, kWidth, W(m, w1), kWidth, m->dY, kIn2, n, kin_used, kWidth,
       1.f);
}

// Backward of an edge round given dM (f32, stride ldm): gradients of W3/b3,
// W2/b2, W1/b1; leaves dz1 (bf16, e x 128) in m->dz1.
static void edge_back(clsph_gnn* m, int r, int w1, int w2, int w3, const float* dM, int ldm,
                      hipStream_t s) {
  int n = (int)m->n, e = (int)m->e;
  size_t ew = (size_t)e * kWidth;
  // bf16(dM) into dpb (n x 128)
  hipLaunchKernelGGL(k_cast_2d, grid_for((size_t)n * kWidth), dim3(256), 0, s, dM, ldm, m->dpb,
                     kWidth, (size_t)n, kWidth);
  gemm(s, true, false, true, m->S[r], kWidth, m->dpb, kWidth, G(m, w3), kWidth, kWidth, kWidth, n,
       1.f);
  hipLaunchKernelGGL((k_colsum<float>), dim3((n + 511) / 512), dim3(kWidth), 0, s, dM, ldm,
                     (size_t)n, kWidth, m->degc, G(m, w3 + 1));
  gemm(s, false, true, false, m->dpb, kWidth, W(m, w3), kWidth, m->dS, kWidth, n, kWidth, kWidth,
       1.f / kAggC);
  // edges
  hipLaunchKernelGGL(k_edge_dz, grid_for(ew), dim3(256), 0, s, m->dS, m->z2[r], m->dst, m->dz2,
                     (size_t)e);
  gemm(s, true, false, true, m->s1[r], kWidth, m->dz2, kWidth, G(m, w2), kWidth, kWidth, kWidth, e,
       1.f);
  hipLaunchKernelGGL((k_colsum<uint16_t>), dim3((e + 511) / 512), dim3(kWidth), 0, s, m->dz2,
                     kWidth, (size_t)e, kWidth, nullptr, G(m, w2 + 1));
  gemm(s, false, true, false, m->dz2, kWidth, W(m, w2), kWidth, m->ds1, kWidth, e, kWidth, kWidth,
       1.f);
  hipLaunchKernelGGL(k_dsilu, grid_for(ew), dim3(256), 0, s, m->ds1, kWidth, m->z1[r], m->dz1,
                     (size_t)e, kWidth);
  gemm(s, true, false, true, m->E1, kFeat, m->dz1, kWidth, G(m, w1), kWidth, kFeat, kWidth, e,
       1.f);
  hipLaunchKernelGGL((k_colsum<uint16_t>), dim3((e + 511) / 512), dim3(kWidth), 0, s, m->dz1,
                     kWidth, (size_t)e, kWidth, nullptr, G(m, w1 + 1));
}

void backward(clsph_gnn* m) {
  hipStream_t s = m->ctx->stream;
  int n = (int)m->n;
  size_t nw = (size_t)n * kWidth;
  fill(m->grad, kParamTotal, 0.f, s);
  fill(m->loss, 1, 0.f, s);
  hipLaunchKernelGGL(k_loss, grid_for(n), dim3(256), 0, s, m->out, m->target, m->n, m->dout,
                     m->loss);
  // dpb (n x 256 bf16): the first half is mlp3_back's scratch, the second
  // half holds the bf16 gradient fed into it.
  uint16_t* dy_b = m->dpb + nw;
  hipLaunchKernelGGL(k_cast, grid_for((size_t)n * kOut), dim3(256), 0, s, m->dout, dy_b,
                     (size_t)n * kOut);
  // node MLP 2: dY (stride 256) = [d H | d M2]
  mlp3_back(m, N2_W1, m->u2, kIn2, m->q1, m->w1, m->q2, m->w2, dy_b, m->dout, kOut, kOut, kIn2,
            s);
  hipLaunchKernelGGL(k_copy_2d, grid_for(nw), dim3(256), 0, s, m->dY, kIn2, m->dH, kWidth,
                     (size_t)n, kWidth);
  // edge round 2 and its node terms
  edge_back(m, 1, E2_W1, E2_W2, E2_W3, m->dY + kWidth, kIn2, s);
  hipLaunchKernelGGL(k_segsum_rows, dim3((n + 3) / 4), dim3(256), 0, s, m->dz1, m->off, m->n,
                     m->dPA);
  fill(m->dPB, nw, 0.f, s);
  hipLaunchKernelGGL(k_scatter_rows, grid_for((size_t)m->e * (kWidth / 2)), dim3(256), 0, s,
                     m->dz1, m->src, (size_t)m->e, m->dPB);
  hipLaunchKernelGGL(k_cast, grid_for(nw), dim3(256), 0, s, m->dPA, m->dPA_b, nw);
  hipLaunchKernelGGL(k_cast, grid_for(nw), dim3(256), 0, s, m->dPB, m->dPB_b, nw);
  gemm(s, true, false, true, m->Hb, kWidth, m->dPA_b, kWidth, G(m, E2_WA), kWidth, kWidth, kWidth,
       n, 1.f);
  gemm(s, true, false, true, m->Hb, kWidth, m->dPB_b, kWidth, G(m, E2_WB), kWidth, kWidth, kWidth,
       n, 1.f);
  gemm(s, false, true, true, m->dPA_b, kWidth, W(m, E2_WA), kWidth, m->dH, kWidth, n, kWidth,
       kWidth, 1.f);
  gemm(s, false, true, true, m->dPB_b, kWidth, W(m, E2_WB), kWidth, m->dH, kWidth, n, kWidth,
       kWidth, 1.f);
  // node MLP 1: dY[:, :128] = d M1
  hipLaunchKernelGGL(k_cast, grid_for(nw), dim3(256), 0, s, m->dH, dy_b, nw);
  mlp3_back(m, N1_W1, m->u1, kIn1, m->p1, m->y1, m->p2, m->y2, dy_b, m->dH, kWidth, kWidth,
            kWidth, s);
  // edge round 1
  edge_back(m, 0, E1_W1, E1_W2, E1_W3, m->dY, kIn2, s);
}

void adam(clsph_gnn* m, float lr, float beta1, float beta2, float eps) {
  m->adam_step += 1;
  float bc1 = 1.f - powf(beta1, (float)m->adam_step);
  float bc2 = 1.f - powf(beta2, (float)m->adam_step);
  hipLaunchKernelGGL(k_adam, grid_for(kParamTotal), dim3(256), 0, m->ctx->stream, m->params,
                     m->grad, m->adam_m, m->adam_v, (size_t)kParamTotal, lr, beta1, beta2, eps,
                     bc1, bc2, m->wb);
}

void cast_params(clsph_gnn* m) {
  hipLaunchKernelGGL(k_cast, grid_for(kParamTotal), dim3(256), 0, m->ctx->stream, m->params,
                     m->wb, (size_t)kParamTotal);
}

}  // namespace gnn
}  // namespace clsph
