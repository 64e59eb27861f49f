// Kernels of the learned force surrogate (gfx950): a bf16 matrix-core GEMM
// with float32 accumulation, the neighbour-graph and edge-input builders,
// activation / aggregation / loss kernels and Adam. The forward and
// backward passes are sequences of these launches on the context stream.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "nn.h"

namespace clsph {
namespace nn {

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

// ---------------------------------------------------------------------------
// GEMM: C (f32, M x N, ldc) = alpha * op(A) op(B) [+ rowscale[m] * bias[n]]
// (ADD: atomically added to C). op(A) is M x K: A row-major M x K (lda), or
// with TA stored K x M. op(B) is K x N: B row-major K x N (ldb), or with TB
// stored N x K. Tiles of 128 x 128, K steps of 32 through LDS, four waves of
// 2 x 2 32x32x16 bf16 matrix-core blocks. blockIdx.z splits K into chunks of
// k_chunk (a multiple of 32); with more than one chunk ADD must be set.
// The contiguous dimension of each operand must be a multiple of 8 and
// 16-byte aligned (checked by the host wrapper).
constexpr int kTile = 128;
constexpr int kStep = 32;
constexpr int kLds = kStep + 8;  // u16 per LDS row (padding against bank conflicts)

struct GemmArgs {
  const uint16_t* A;
  const uint16_t* B;
  float* C;
  int M, N, K, lda, ldb, ldc;
  float alpha;
  const float* bias;
  const float* rowscale;
  int k_chunk;
};

template <bool TA, bool TB, bool ADD>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kTile * kLds];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kTile * kLds];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * kTile, n0 = blockIdx.y * kTile;
  const int kb = blockIdx.z * g.k_chunk;
  const int ke = min(g.K, kb + g.k_chunk);
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  f32x16 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  const u32x4 zero = {0u, 0u, 0u, 0u};
  for (int k0 = kb; k0 < ke; k0 += kStep) {
    for (int it = 0; it < 2; ++it) {
      int idx = tid + 256 * it;
      if (!TA) {
        int m = idx >> 2, kq = (idx & 3) * 8;
        int gm = m0 + m, gk = k0 + kq;
        u32x4 v = zero;
        if (gm < g.M && gk < ke) v = *(const u32x4*)(g.A + (size_t)gm * g.lda + gk);
        *(u32x4*)(As + m * kLds + kq) = v;
      } else {
        int k = idx >> 4, mq = (idx & 15) * 8;
        int gm = m0 + mq, gk = k0 + k;
        u32x4 v = zero;
        if (gm < g.M && gk < ke) v = *(const u32x4*)(g.A + (size_t)gk * g.lda + gm);
        const uint16_t* e = (const uint16_t*)&v;
        for (int t = 0; t < 8; ++t) As[(mq + t) * kLds + k] = e[t];
      }
      if (!TB) {
        int k = idx >> 4, nq = (idx & 15) * 8;
        int gn = n0 + nq, gk = k0 + k;
        u32x4 v = zero;
        if (gn < g.N && gk < ke) v = *(const u32x4*)(g.B + (size_t)gk * g.ldb + gn);
        const uint16_t* e = (const uint16_t*)&v;
        for (int t = 0; t < 8; ++t) Bs[(nq + t) * kLds + k] = e[t];
      } else {
        int n = idx >> 2, kq = (idx & 3) * 8;
        int gn = n0 + n, gk = k0 + kq;
        u32x4 v = zero;
        if (gn < g.N && gk < ke) v = *(const u32x4*)(g.B + (size_t)gn * g.ldb + gk);
        *(u32x4*)(Bs + n * kLds + kq) = v;
      }
    }
    __syncthreads();
    for (int ks = 0; ks < kStep; ks += 16) {
      const int kk = ks + 8 * (lane >> 5);
      bf16x8 a[2], b[2];
      for (int i = 0; i < 2; ++i)
        a[i] = __builtin_bit_cast(bf16x8,
                                  *(const u32x4*)(As + (wm + 32 * i + (lane & 31)) * kLds + kk));
      for (int j = 0; j < 2; ++j)
        b[j] = __builtin_bit_cast(bf16x8,
                                  *(const u32x4*)(Bs + (wn + 32 * j + (lane & 31)) * kLds + kk));
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout of the 32x32 block: register r of lane l holds
  // row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 16; ++r) {
        int row = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        int col = n0 + wn + 32 * j + (lane & 31);
        if (row < g.M && col < g.N) {
          float v = g.alpha * acc[i][j][r];
          if (g.bias && blockIdx.z == 0) v += g.bias[col] * (g.rowscale ? g.rowscale[row] : 1.f);
          float* c = g.C + (size_t)row * g.ldc + col;
          if (ADD)
            atomicAdd(c, v);
          else
            *c = v;
        }
      }
}

// Host wrapper. `add` accumulates atomically; large-K products (K > 4096)
// are split over the K dimension, which requires `add`.
void gemm(hipStream_t s, bool ta, bool tb, bool add, const uint16_t* A, int lda,
          const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K, float alpha,
          const float* bias = nullptr, const float* rowscale = nullptr) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  GemmArgs g{A, B, C, M, N, K, lda, ldb, ldc, alpha, bias, rowscale, 0};
  int splits = 1;
  if (add && K > 4096) splits = (K + 4095) / 4096;
  if (splits > 1024) splits = 1024;
  int chunk = (K + splits - 1) / splits;
  chunk = (chunk + kStep - 1) / kStep * kStep;
  splits = (K + chunk - 1) / chunk;
  g.k_chunk = chunk;
  dim3 grid((M + kTile - 1) / kTile, (N + kTile - 1) / kTile, splits);
#define GEMM_CASE(a, b, c)                                                   \
  if (ta == a && tb == b && add == c) {                                      \
    hipLaunchKernelGGL((k_gemm<a, b, c>), grid, dim3(256), 0, s, g);         \
    return;                                                                  \
  }
  GEMM_CASE(false, false, false) GEMM_CASE(false, false, true)
  GEMM_CASE(false, true, false) GEMM_CASE(false, true, true)
  GEMM_CASE(true, false, false) GEMM_CASE(true, false, true)
  GEMM_CASE(true, true, false) GEMM_CASE(true, true, true)
#undef GEMM_CASE
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

// Constants of the edge inputs (one scene).
struct EdgeConsts {
  float h, h2, inv_h, rho0, inv_rho0, K, m, mu_over_m, c_spiky, c_visc, c_grad, c_lap;
  // the edge inputs' constant factors, multiplied out once (edge_features)
  float kp, kv, kg, kl, kr, kd, kx;
};

inline EdgeConsts edge_consts(const DevParams& d) {
  EdgeConsts k;
  k.h = d.p.h;
  k.h2 = d.p.h * d.p.h;
  k.inv_h = 1.f / d.p.h;
  k.rho0 = d.p.fluid_density;
  k.inv_rho0 = 1.f / d.p.fluid_density;
  k.K = d.p.K;
  k.m = d.p.particle_mass;
  k.mu_over_m = d.p.dynamic_viscosity / d.p.particle_mass;
  k.c_spiky = d.t.spiky;
  k.c_visc = d.t.viscosity;
  k.c_grad = d.t.poly_6_gradient;
  k.c_lap = d.t.poly_6_laplacian;
  const double inv_a = 1.0 / kAccScale, inv_v = 1.0 / kVelScale, h = d.p.h;
  const double m = d.p.particle_mass;
  k.kp = (float)(-m * d.t.spiky * inv_a);
  k.kv = (float)((double)d.p.dynamic_viscosity / m * d.t.viscosity * inv_a);
  k.kg = (float)((double)d.t.poly_6_gradient * h);
  k.kl = (float)((double)d.t.poly_6_laplacian * h * h);
  k.kr = (float)(-m * d.t.spiky * h * inv_v * inv_a);
  k.kd = (float)(m * d.t.poly_6_gradient * 1000.0 * h * inv_v / (double)d.p.fluid_density);
  k.kx = (float)(10.0 * inv_v / h);
  return k;
}

// Constants of the forces stage (forces.cl:45-112, sph.cl:53-58) for the
// fused inference, which sums the listed neighbours' pair terms itself.
struct ForceTail {
  float h2, spiky0, mu, sigma, thresh, gx, gy, gz;
};

inline ForceTail force_tail(const DevParams& d) {
  ForceTail t;
  t.h2 = d.p.h * d.p.h;
  const float h3 = t.h2 * d.p.h;
  t.spiky0 = -45.f / (3.14159274f * (h3 * h3));  // coincident particles (smoothing.cl:23-25)
  t.mu = d.p.dynamic_viscosity;
  t.sigma = d.p.surface_tension;
  t.thresh = d.p.surface_tension_threshold;
  t.gx = d.p.constant_acceleration.s[0];
  t.gy = d.p.constant_acceleration.s[1];
  t.gz = d.p.constant_acceleration.s[2];
  return t;
}

// Pair sums of the forces stage for particle i: pressure (without the
// -rho_i factor), viscosity (without mu), colour-field gradient and
// laplacian, as the forces kernel accumulates them (forces.cl:60-101).
struct PairSums {
  float px, py, pz, vx, vy, vz, nx, ny, nz, lap;
};

__device__ __attribute__((always_inline)) inline void pair_terms(PairSums& f, const EdgeConsts& k,
                                                                 const ForceTail& ft, float4 ai,
                                                                 float4 bi, float4 aj, float4 bj,
                                                                 float w) {
  float dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
  float d2 = dx * dx + dy * dy + dz * dz;
  float r = __builtin_amdgcn_sqrtf(d2);
  float t = k.h2 - d2;
  float hr = k.h - r;
  float ps = (aj.w + ai.w) * k.m;
  const bool coincident = r < 1e-7f;
  float g = coincident ? ps * ft.spiky0 : ps * k.c_spiky * hr * hr * __builtin_amdgcn_rcpf(r);
  g *= w;
  f.px += g * (coincident ? 1.f : dx);
  f.py += g * (coincident ? 1.f : dy);
  f.pz += g * (coincident ? 1.f : dz);
  float vw = bj.w * k.c_visc * hr * w;
  f.vx += (bj.x - bi.x) * vw;
  f.vy += (bj.y - bi.y) * vw;
  f.vz += (bj.z - bi.z) * vw;
  float gw = bj.w * k.c_grad * t * t * w;
  f.nx += gw * dx;
  f.ny += gw * dy;
  f.nz += gw * dz;
  f.lap += bj.w * k.c_lap * t * (3.f * k.h2 - 7.f * d2) * w;
}

// The 32 edge inputs of edge (i <- j) in float32 (27 used, 5 zero), from
// the density stage's per-particle force inputs a = (x, p / rho^2) and
// b = (v, m / rho) and the node inputs nu. The pair terms are the ones the
// forces stage sums (kernels/forces.cl:60-101): pressure and viscosity
// accelerations in units of A, the colour-field gradient term times h and
// its laplacian term times h^2. Then three rates along the pair's motion
// (x = x_i - x_j, v = v_i - v_j): the time derivative of the pressure pair
// acceleration at fixed pressures, -ps c_spiky [(h - r)^2 / r v -
// (x.v) (h - r)(h + r) / r^3 x], times h / (V A); the density-rate pair term
// m c_grad (h^2 - r^2)^2 (x.v) times 1000 h / (rho0 V); and x.v times
// 10 / (h V).
__device__ inline void edge_features(float4 ai, float4 bi, float4 aj, float4 bj, float2 ni,
                                     float2 nj, const EdgeConsts& k, float f[kFeat]) {
  static_assert(kVelScale == 1.f, "velocity inputs are taken unscaled");
  float dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
  float d2 = dx * dx + dy * dy + dz * dz;
  float r = __builtin_amdgcn_sqrtf(d2);  // the forces kernel's square root
  float q = r * k.inv_h;
  float w = 1.f - q * q;
  const float vx = bi.x - bj.x, vy = bi.y - bj.y, vz = bi.z - bj.z;
  f[0] = dx * k.inv_h, f[1] = dy * k.inv_h, f[2] = dz * k.inv_h;
  f[3] = vx, f[4] = vy, f[5] = vz;
  f[6] = q, f[7] = w * w * w;
  f[8] = ni.x, f[9] = ni.y, f[10] = nj.x, f[11] = nj.y;
  const float hr = k.h - r, t = k.h2 - d2, t2 = t * t;
  const float ws = ai.w + aj.w;  // p_i / rho_i^2 + p_j / rho_j^2
  const bool apart = r >= 1e-7f;
  const float ir = __builtin_amdgcn_rcpf(r);
  // pressure pair acceleration / A: -(ws m c_spiky (h - r)^2 / r) x / A
  const float g = apart ? ws * k.kp * hr * hr * ir : 0.f;
  f[12] = g * dx, f[13] = g * dy, f[14] = g * dz;
  // viscosity pair acceleration / A: (v_j - v_i) (m / rho_j) (mu / rho_i) c_visc (h - r) / A
  const float vw = bi.w * k.kv * bj.w * hr;
  f[15] = -vx * vw, f[16] = -vy * vw, f[17] = -vz * vw;
  // colour-field gradient term times h, laplacian term times h^2
  const float gw = bj.w * k.kg * t2;
  f[18] = gw * dx, f[19] = gw * dy, f[20] = gw * dz;
  f[21] = bj.w * k.kl * t * (3.f * k.h2 - 7.f * d2);
  const float xv = dx * vx + dy * vy + dz * vz;
  float ra = 0.f, rb = 0.f;
  if (apart) {
    ra = hr * hr * ir;
    rb = xv * hr * (k.h + r) * (ir * ir * ir);
  }
  const float cr = ws * k.kr;
  f[22] = cr * (ra * vx - rb * dx), f[23] = cr * (ra * vy - rb * dy);
  f[24] = cr * (ra * vz - rb * dz);
  f[25] = t2 * xv * k.kd;
  f[26] = xv * k.kx;
  for (int c = kFeatUsed; c < kFeat; ++c) f[c] = 0.f;
}

// nu_i = [rho_i / rho0 - 1, asinh(p_i / K) / 4] for every particle.
__global__ void k_node_inputs(const float* __restrict__ rho, const float* __restrict__ prs,
                              uint32_t n, float rho0, float K, float2* __restrict__ nu) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) nu[i] = node_in(rho[i], prs[i], rho0, K);
}

// Edge inputs e (32 bf16 per edge), after k_node_inputs.
__global__ __launch_bounds__(256) void k_edge_inputs(
    const float4* __restrict__ fab, const float2* __restrict__ nu,
    const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint32_t e,
    EdgeConsts kc, uint16_t* __restrict__ E1) {
  for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < e; k += gridDim.x * 256) {
    uint32_t i = dst[k], j = src[k];
    float f[kFeat];
    edge_features(fab[2u * i], fab[2u * i + 1u], fab[2u * j], fab[2u * j + 1u], nu[i], nu[j],
                  kc, f);
    u32x4* o = (u32x4*)(E1 + (size_t)k * kFeat);
    for (int q = 0; q < kFeat / 8; ++q) {
      uint32_t packed[4];
      for (int t = 0; t < 4; ++t)
        packed[t] = (uint32_t)to_bf16(f[8 * q + 2 * t]) |
                    ((uint32_t)to_bf16(f[8 * q + 2 * t + 1]) << 16);
      o[q] = u32x4{packed[0], packed[1], packed[2], packed[3]};
    }
  }
}

// ---------------------------------------------------------------------------
// Fused inference: one kernel walks the density stage's neighbour lists and
// evaluates the whole network.
//
// A wave owns tiles of 16 consecutive sorted particles; lane l serves
// particle c = l % 16 of the tile in lane group g = l / 16. The edge layer
// runs transposed on 16x16x32 bf16 matrix cores, z^T = W1^T e^T, with the
// tile's particles as the 16 columns: at step s, column c holds particle c's
// s-th neighbour edge, so each lane's accumulator sums its own particle's
// edges in list order and the neighbour sum needs no cross-lane reduction.
// A lane holds z for hidden units 16 ob + 4 g + r (r = 0..3) of its column.
//
// Edge inputs are computed once per edge. In each round of 4 steps, lane
// (g, c) computes the inputs of particle c's edge 4 k + g and stores its 32
// inputs as bf16 in the wave's LDS block; the B operand of step 4 k + s for
// lane (g, c) is inputs 8 g .. 8 g + 7 of edge 4 k + s of particle c. Per
// particle column the block holds 4 edges x 64 bytes plus 16 bytes of
// padding: 272 bytes = 68 dwords, so the 16 columns of one read start on 16
// distinct multiples of 4 banks.
//
// The node MLP follows in registers: the neighbour sums are the B operand of
// node layer 1 in the permuted K order (slot t of lane group g is feature
// 16 (t / 4) + 4 g + t % 4 of a 32-wide block), node inputs 32..63 come in
// the natural order (slot t of group g is feature 32 + 8 g + t), and each
// later layer takes its B operand from the previous layer's registers in the
// permuted order. Weights are staged once per work-group in LDS in the
// matching orders.
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTileP = 16;                 // particles per tile (MFMA columns)
constexpr int kSurWaves = 8;               // waves per work-group
constexpr int kColBytes = 4 * 64 + 16;     // LDS bytes per particle column
constexpr int kWaveLds = kTileP * kColBytes;

// Eight floats to bf16 (round to nearest even), as one vector conversion so
// that each pair becomes one two-operand v_cvt_pk_bf16_f32.
typedef float f32x8 __attribute__((ext_vector_type(8)));

__device__ __attribute__((always_inline)) inline u32x4 pack8(const float v[8]) {
  const f32x8 x = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
  return __builtin_bit_cast(u32x4, __builtin_convertvector(x, bf16x8));
}

// Stages W (row-major, rows = inputs, ld = outputs) for k-steps
// [ks0, ks0 + nks) of a layer with `ks_total` k-steps and `nob` 16-wide
// output blocks: dst[(ob * ks_total + ks) * 64 + lane] holds the lane's 8 K
// slots of W^T row 16 ob + lane % 16, natural or permuted K order; input
// rows at or beyond `kvalid` are zero. `krow0` is the W row of k-step ks0.
__device__ __attribute__((always_inline)) inline void stage_layer(
    u32x4* dst, int ks_total, int ks0, int nks, int nob, const uint16_t* __restrict__ Wm, int ld,
    int krow0, int kvalid, bool perm) {
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
__device__ __attribute__((always_inline)) inline void mfma_layer(const u32x4* __restrict__ w,
                                                                 const bf16x8* bop,
                                                                 const float* __restrict__ bias,
                                                                 f32x4* y) {
  const int lane = threadIdx.x & 63, g = lane >> 4;
#pragma unroll
  for (int ob = 0; ob < NOB; ++ob) {
    f32x4 a;
    for (int r = 0; r < 4; ++r) a[r] = bias[16 * ob + 4 * g + r];
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
      a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
          __builtin_bit_cast(bf16x8, w[(ob * NKS + ks) * 64 + lane]), bop[ks], a, 0, 0, 0);
    y[ob] = a;
  }
}

// Permuted-order B operands (NOB / 2 k-steps) from a layer's registers,
// through SiLU.
template <int NOB>
__device__ __attribute__((always_inline)) inline void silu_operands(const f32x4* y, bf16x8* bop) {
#pragma unroll
  for (int ks = 0; ks < NOB / 2; ++ks) {
    float v8[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v8[t] = silu(y[2 * ks + (t >> 2)][t & 3]);
    bop[ks] = __builtin_bit_cast(bf16x8, pack8(v8));
  }
}

// The neighbour of lane (g, c) in round k: edge 4 k + g of particle c (the
// particle itself when that edge is past its degree).
__device__ __attribute__((always_inline)) inline uint32_t edge_index(
    uint32_t k, uint32_t i, uint32_t n, uint32_t deg, const uint4* __restrict__ nbr) {
  const int g = (threadIdx.x & 63) >> 4;
  uint32_t j = i;
  if (4 * k + g < deg) {
    uint4 js = nbr[(size_t)k * n + i];
    j = g == 0 ? js.x : g == 1 ? js.y : g == 2 ? js.z : js.w;
  }
  return j;
}

struct EdgeFetch {
  float4 a, b;
  float2 nu;
  bool mine;
};

// The inputs of neighbour j, lane (g, c)'s edge of round k.
__device__ __attribute__((always_inline)) inline EdgeFetch fetch_edge(
    uint32_t k, uint32_t j, uint32_t deg, const float4* __restrict__ fab,
    const float2* __restrict__ nu) {
  const int g = (threadIdx.x & 63) >> 4;
  EdgeFetch e;
  e.mine = 4 * k + g < deg;
  e.a = fab[2u * j];
  e.b = fab[2u * j + 1u];
  e.nu = nu[j];
  return e;
}

// One round of 4 edge steps for the wave's tile (see above). MASKED rounds
// contain edges past some column's degree: those edges store zero inputs and
// their steps start from a zero accumulator, so they add relu(0) = 0.
template <bool MASKED, bool OWN>
__device__ __attribute__((always_inline)) inline void edge_round(
    uint32_t k, uint32_t deg, const float4& ai, const float4& bi, const float2& ni,
    const EdgeFetch& ej, const EdgeConsts& kc, const ForceTail& ft, uint8_t* wlds,
    const u32x4* __restrict__ w1s, const f32x4* zb, f32x4* s, PairSums& ps) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  if (OWN) pair_terms(ps, kc, ft, ai, bi, ej.a, ej.b, (MASKED && !ej.mine) ? 0.f : 1.f);
  float f[kFeat];
  edge_features(ai, bi, ej.a, ej.b, ni, ej.nu, kc, f);
  uint8_t* col = wlds + c * kColBytes;
  for (int q = 0; q < 4; ++q) {
    u32x4 v = pack8(f + 8 * q);
    if (MASKED && !ej.mine) v = u32x4{0u, 0u, 0u, 0u};
    *(u32x4*)(col + g * 64 + 16 * q) = v;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const bf16x8 eb = __builtin_bit_cast(bf16x8, *(const u32x4*)(col + st * 64 + g * 16));
    const bool live = !MASKED || 4 * k + st < deg;
#pragma unroll
    for (int ob = 0; ob < kEdgeW / 16; ++ob) {
      f32x4 c0 = zb[ob];
      if (MASKED && !live) c0 = f32x4{0.f, 0.f, 0.f, 0.f};
      f32x4 z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
          __builtin_bit_cast(bf16x8, w1s[ob * 64 + lane]), eb, c0, 0, 0, 0);
      // s + relu(z) as max(s + z, s): the same value and rounding, and the
      // operands of the max are canonical (one instruction, not two)
      for (int r = 0; r < 4; ++r) s[ob][r] = fmaxf(s[ob][r] + z[r], s[ob][r]);
    }
  }
  // the next round's stores come after this round's reads
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// OWN: the forces stage's acceleration a0 is summed here from the listed
// neighbours (particles whose list overflowed take acc[i], which the
// forces kernel wrote for them); otherwise a0 = acc[i].
template <bool OWN>
__global__ __launch_bounds__(kSurWaves * 64) void k_surrogate(
    const float4* __restrict__ fab, const float2* __restrict__ nu,
    const float* __restrict__ rho, const uint4* __restrict__ nbr,
    const uint32_t* __restrict__ count, uint32_t n, EdgeConsts kc, ForceTail ft,
    const uint16_t* __restrict__ wb, const float* __restrict__ params, float4* __restrict__ acc,
    float* __restrict__ out) {
  const float gx = ft.gx, gy = ft.gy, gz = ft.gz;
  __shared__ u32x4 w1s[2 * 64];          // edge layer: 2 output blocks, 1 k-step
  __shared__ u32x4 wn1[4 * 2 * 64];      // node layer 1: k-step 0 permuted, 1 natural
  __shared__ u32x4 wn2[4 * 2 * 64];
  __shared__ u32x4 wn3[2 * 2 * 64];
  __shared__ float bias[kEdgeW + 2 * kNodeW + kOut];
  __shared__ __attribute__((aligned(16))) uint8_t tiles_lds[kSurWaves * kWaveLds];
  stage_layer(w1s, 1, 0, 1, 2, wb + param_offset(E_W1), kEdgeW, 0, kFeatUsed, false);
  stage_layer(wn1, 2, 0, 1, 4, wb + param_offset(N_W1), kNodeW, 0, kEdgeW, true);
  stage_layer(wn1, 2, 1, 1, 4, wb + param_offset(N_W1), kNodeW, kEdgeW, kEdgeW + 6, false);
  stage_layer(wn2, 2, 0, 2, 4, wb + param_offset(N_W2), kNodeW, 0, kNodeW, true);
  stage_layer(wn3, 2, 0, 2, 2, wb + param_offset(N_W3), kOut, 0, kNodeW, true);
  for (int k = threadIdx.x; k < kEdgeW + 2 * kNodeW + kOut; k += blockDim.x) {
    float v;
    if (k < kEdgeW) v = params[param_offset(E_B1) + k];
    else if (k < kEdgeW + kNodeW) v = params[param_offset(N_B1) + k - kEdgeW];
    else if (k < kEdgeW + 2 * kNodeW) v = params[param_offset(N_B2) + k - kEdgeW - kNodeW];
    else v = params[param_offset(N_B3) + k - kEdgeW - 2 * kNodeW];
    bias[k] = v;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15, wave = threadIdx.x >> 6;
  uint8_t* wlds = tiles_lds + wave * kWaveLds;
  f32x4 zb[2];
  for (int ob = 0; ob < 2; ++ob)
    for (int r = 0; r < 4; ++r) zb[ob][r] = bias[16 * ob + 4 * g + r];
  const uint32_t waves = gridDim.x * kSurWaves;
  const uint32_t ntiles = (n + kTileP - 1) / kTileP;
  for (uint32_t tile = blockIdx.x * kSurWaves + wave; tile < ntiles; tile += waves) {
    const uint32_t ip = tile * kTileP + c;
    const bool valid = ip < n;
    const uint32_t i = valid ? ip : n - 1;
    const uint32_t deg = valid ? min(count[i], kMaxEdges) : 0u;
    uint32_t dmin = deg, dmax = deg;
    for (int o = 1; o < 16; o <<= 1) {
      dmin = min(dmin, (uint32_t)__shfl_xor((int)dmin, o, 64));
      dmax = max(dmax, (uint32_t)__shfl_xor((int)dmax, o, 64));
    }
    const uint32_t cnt = valid ? count[i] : 0u;
    const float4 ai = fab[2u * i], bi = fab[2u * i + 1u];
    const float2 ni = nu[i];
    f32x4 s[2];
    for (int ob = 0; ob < 2; ++ob) s[ob] = f32x4{0.f, 0.f, 0.f, 0.f};
    PairSums ps = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    // rounds inside every column's degree, then the masked tail; each
    // round's neighbour inputs are fetched one round ahead
    const uint32_t kfull = dmin >> 2, kend = (dmax + 3) >> 2;
    // neighbour indices are loaded two rounds ahead, so that a round's
    // gathers never wait for the index load that names them
    EdgeFetch cur;
    if (kend) cur = fetch_edge(0, edge_index(0, i, n, deg, nbr), deg, fab, nu);
    uint32_t jn = kend > 1 ? edge_index(1, i, n, deg, nbr) : i;
    for (uint32_t k = 0; k < kend; ++k) {
      EdgeFetch nxt = cur;
      if (k + 1 < kend) nxt = fetch_edge(k + 1, jn, deg, fab, nu);
      if (k + 2 < kend) jn = edge_index(k + 2, i, n, deg, nbr);
      if (k < kfull)
        edge_round<false, OWN>(k, deg, ai, bi, ni, cur, kc, ft, wlds, w1s, zb, s, ps);
      else
        edge_round<true, OWN>(k, deg, ai, bi, ni, cur, kc, ft, wlds, w1s, zb, s, ps);
      cur = nxt;
    }
    float4 a0;
    if (OWN) {
      // the four lane groups hold interleaved edges of the same particle
      auto group_sum = [](float& v) {
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
      };
      group_sum(ps.px), group_sum(ps.py), group_sum(ps.pz);
      group_sum(ps.vx), group_sum(ps.vy), group_sum(ps.vz);
      group_sum(ps.nx), group_sum(ps.ny), group_sum(ps.nz);
      group_sum(ps.lap);
      // the forces kernel's closing formula (forces.cl:95-111, sph.cl:53-58),
      // with the colour-field laplacian's self term
      const float rho_i = rho[i];
      const float lap = bi.w * kc.c_lap * ft.h2 * (3.f * ft.h2) + ps.lap;
      float fx = -rho_i * ps.px + ps.vx * ft.mu;
      float fy = -rho_i * ps.py + ps.vy * ft.mu;
      float fz = -rho_i * ps.pz + ps.vz * ft.mu;
      const float nl = sqrtf(ps.nx * ps.nx + ps.ny * ps.ny + ps.nz * ps.nz);
      if (nl > ft.thresh) {
        const float kk = -ft.sigma * lap / nl;
        fx += kk * ps.nx;
        fy += kk * ps.ny;
        fz += kk * ps.nz;
      }
      a0 = make_float4(fx / rho_i + gx, fy / rho_i + gy, fz / rho_i + gz, 0.f);
      if (cnt > kMaxEdges) a0 = acc[i];
    } else {
      a0 = acc[i];
    }

    // node layer 1: [S, nu, (a0 - g) / A, deg / C, 0...]
    bf16x8 bop[2];
    {
      float v8[8];
      for (int t = 0; t < 8; ++t) v8[t] = s[t >> 2][t & 3] * (1.f / kAggC);
      bop[0] = __builtin_bit_cast(bf16x8, pack8(v8));
      for (int t = 0; t < 8; ++t) v8[t] = 0.f;
      if (g == 0) {
        constexpr float inv_a = 1.f / kAccScale;
        v8[0] = ni.x, v8[1] = ni.y;
        v8[2] = (a0.x - gx) * inv_a, v8[3] = (a0.y - gy) * inv_a, v8[4] = (a0.z - gz) * inv_a;
        v8[5] = (float)deg * (1.f / kAggC);
      }
      bop[1] = __builtin_bit_cast(bf16x8, pack8(v8));
    }
    f32x4 y[4];
    mfma_layer<4, 2>(wn1, bop, bias + kEdgeW, y);
    silu_operands<4>(y, bop);
    mfma_layer<4, 2>(wn2, bop, bias + kEdgeW + kNodeW, y);
    silu_operands<4>(y, bop);
    f32x4 o[2];
    mfma_layer<2, 2>(wn3, bop, bias + kEdgeW + 2 * kNodeW, o);
    if (valid) {
      if (out) {
        float* orow = out + (size_t)i * kOut;
        for (int ob = 0; ob < 2; ++ob)
          *(float4*)(orow + 16 * ob + 4 * g) = make_float4(o[ob][0], o[ob][1], o[ob][2], o[ob][3]);
      }
      if (g == 0)
        acc[i] = make_float4(a0.x + kResScale * o[0][0], a0.y + kResScale * o[0][1],
                             a0.z + kResScale * o[0][2], a0.w);
    }
  }
}

// ---------------------------------------------------------------------------
// Training-path kernels (grid-stride loops unless noted).
__global__ void k_silu_bf16(const float* __restrict__ z, uint16_t* __restrict__ s, size_t count) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x)
    s[k] = to_bf16(silu(z[k]));
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

// S[i][c] = bf16(sum over i's edges of relu(z[e][c]) / C); thread per (i, c).
__global__ __launch_bounds__(256) void k_segsum_relu(const float* __restrict__ z,
                                                     const uint32_t* __restrict__ off, uint32_t n,
                                                     uint16_t* __restrict__ S) {
  const size_t count = (size_t)n * kEdgeW;
  for (size_t t = blockIdx.x * (size_t)256 + threadIdx.x; t < count;
       t += (size_t)gridDim.x * 256) {
    uint32_t i = (uint32_t)(t / kEdgeW);
    int c = (int)(t % kEdgeW);
    float a = 0.f;
    for (uint32_t k = off[i]; k < off[i + 1]; ++k) a += fmaxf(z[(size_t)k * kEdgeW + c], 0.f);
    S[t] = to_bf16(a * (1.f / kAggC));
  }
}

// u = bf16([S, nu, (a0 - g) / A, deg / C, 0...]) (64 wide).
__global__ void k_pack_u(const uint16_t* __restrict__ S, const float* __restrict__ nu,
                         const float4* __restrict__ acc0, const float* __restrict__ degc,
                         uint32_t n, float gx, float gy, float gz, uint16_t* __restrict__ u) {
  size_t count = (size_t)n * kIn;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t i = k / kIn;
    int c = (int)(k % kIn);
    if (c < kEdgeW) {
      u[k] = S[i * kEdgeW + c];
      continue;
    }
    float v = 0.f;
    const float4 a = acc0[i];
    switch (c - kEdgeW) {
      case 0: v = nu[2 * i]; break;
      case 1: v = nu[2 * i + 1]; break;
      case 2: v = (a.x - gx) * (1.f / kAccScale); break;
      case 3: v = (a.y - gy) * (1.f / kAccScale); break;
      case 4: v = (a.z - gz) * (1.f / kAccScale); break;
      case 5: v = degc[i]; break;
      default: break;
    }
    u[k] = to_bf16(v);
  }
}

// dz[e][c] = bf16(dY[dst[e]][c] / C) where z[e][c] > 0, else 0.
__global__ void k_edge_dz(const float* __restrict__ dY, int ldy, const float* __restrict__ z,
                          const uint32_t* __restrict__ dst, uint16_t* __restrict__ dz, size_t e) {
  size_t count = e * kEdgeW;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
       k += (size_t)gridDim.x * blockDim.x) {
    size_t ed = k / kEdgeW;
    int c = (int)(k % kEdgeW);
    float d = z[k] > 0.f ? dY[(size_t)dst[ed] * ldy + c] * (1.f / kAggC) : 0.f;
    dz[k] = to_bf16(d);
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
                         float* __restrict__ out) {
  int c = threadIdx.x;
  if (c >= cols) return;
  size_t r0 = (size_t)blockIdx.x * 512, r1 = r0 + 512 < rows ? r0 + 512 : rows;
  float a = 0.f;
  for (size_t r = r0; r < r1; ++r) a += load_val<T>(x + r * ldx + c);
  atomicAdd(out + c, a);
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

// loc[ids[k]] = k
__global__ void k_invert(const uint32_t* __restrict__ ids, uint32_t n, uint32_t* __restrict__ loc) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) {
    uint32_t id = ids[k];
    if (id < n) loc[id] = k;
  }
}

// target[i] = ((vh_T - vh_S) / dt - a0_i) / R for the learner's sorted i.
__global__ void k_target(const uint32_t* __restrict__ perm, const float4* __restrict__ vh_s,
                         const uint32_t* __restrict__ loc, const float4* __restrict__ vh_t,
                         const float4* __restrict__ acc0, uint32_t n, float inv_dt,
                         float* __restrict__ target) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t c = perm[i];
  float4 a = vh_s[c], b = vh_t[loc[c]], a0 = acc0[i];
  target[3 * i] = ((b.x - a.x) * inv_dt - a0.x) / kResScale;
  target[3 * i + 1] = ((b.y - a.y) * inv_dt - a0.y) / kResScale;
  target[3 * i + 2] = ((b.z - a.z) * inv_dt - a0.z) / kResScale;
}

// inv[perm[k]] = k
__global__ void k_invert_perm(const uint32_t* __restrict__ perm, uint32_t n,
                              uint32_t* __restrict__ inv) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) {
    uint32_t p = perm[k];
    if (p < n) inv[p] = k;
  }
}

// Density difference of particle (sorted position i of state a) against the
// same particle in state b: a's canonical index perm_a[i] carries identity
// ids_a[...]; b holds it at canonical loc_b[id], sorted inv_b[loc_b[id]].
// Grid-stride; each work-group writes its (max, sum) to part[2 blockIdx.x].
constexpr int kCmpBlocks = 1024;

__global__ __launch_bounds__(256) void k_compare(const uint32_t* __restrict__ perm_a,
                                                 const uint32_t* __restrict__ ids_a,
                                                 const float* __restrict__ rho_a,
                                                 const uint32_t* __restrict__ loc_b,
                                                 const uint32_t* __restrict__ inv_b,
                                                 const float* __restrict__ rho_b, uint32_t n,
                                                 float* __restrict__ part) {
  __shared__ float red[2][4];
  float mx = 0.f, sm = 0.f;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    uint32_t c = perm_a[i];
    uint32_t id = c < n ? ids_a[c] : 0u;
    uint32_t cb = id < n ? loc_b[id] : 0u;
    uint32_t ib = cb < n ? inv_b[cb] : 0u;
    float d = fabsf(rho_a[i] - rho_b[ib]);
    mx = fmaxf(mx, d);
    sm += d;
  }
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    sm += __shfl_xor(sm, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = mx;
    red[1][threadIdx.x >> 6] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = fmaxf(fmaxf(red[0][0], red[0][1]), fmaxf(red[0][2], red[0][3]));
    part[2 * blockIdx.x + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// res = (max, sum) over the `nparts` work-group partials; one work-group.
__global__ __launch_bounds__(256) void k_compare_final(const float* __restrict__ part,
                                                       uint32_t nparts, float* __restrict__ res) {
  __shared__ float red[2][4];
  float mx = 0.f, sm = 0.f;
  for (uint32_t k = threadIdx.x; k < nparts; k += 256) {
    mx = fmaxf(mx, part[2 * k]);
    sm += part[2 * k + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    sm += __shfl_xor(sm, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = mx;
    red[1][threadIdx.x >> 6] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    res[0] = fmaxf(fmaxf(red[0][0], red[0][1]), fmaxf(red[0][2], red[0][3]));
    res[1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// Adds the forces stage's acceleration of sorted particle j to its identity's
// running sum (each identity occurs once per substep).
__global__ void k_accum_acc(const float4* __restrict__ acc, const uint32_t* __restrict__ perm,
                            const uint32_t* __restrict__ ids, uint32_t n,
                            float* __restrict__ sum) {
  uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  uint32_t c = perm[j];
  uint32_t id = c < n ? ids[c] : n;
  if (id >= n) return;
  float4 a = acc[j];
  sum[3 * id] += a.x;
  sum[3 * id + 1] += a.y;
  sum[3 * id + 2] += a.z;
}

// target[i] = (sum[perm[i]] / nsub - a0_i) / R for the learner's sorted i.
__global__ void k_target_mean(const uint32_t* __restrict__ perm, const float* __restrict__ sum,
                              const float4* __restrict__ acc0, uint32_t n, float inv_nsub,
                              float* __restrict__ target) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t c = perm[i];
  if (c >= n) c = 0;
  float4 a0 = acc0[i];
  target[3 * i] = (sum[3 * c] * inv_nsub - a0.x) / kResScale;
  target[3 * i + 1] = (sum[3 * c + 1] * inv_nsub - a0.y) / kResScale;
  target[3 * i + 2] = (sum[3 * c + 2] * inv_nsub - a0.z) / kResScale;
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

void launch_graph(clsph_nn* m, clsph_state* st) {
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

void launch_edge_inputs(clsph_nn* m, clsph_state* st) {
  hipStream_t s = st->ctx->stream;
  uint32_t n = m->n, e = m->e;
  const DevParams& d = st->host_dps[0];
  hipLaunchKernelGGL(k_graph, grid_for(n), dim3(256), 0, s, st->nbr, st->nbr_count, m->off, n,
                     m->src, m->dst, m->degc);
  hipLaunchKernelGGL(k_node_inputs, grid_for(n), dim3(256), 0, s, st->rho_s, st->prs_s, n,
                     d.p.fluid_density, d.p.K, (float2*)m->nu);
  hipLaunchKernelGGL(k_edge_inputs, grid_for(e), dim3(256), 0, s, st->fab,
                     (const float2*)m->nu, m->src, m->dst, e, edge_consts(d), m->E1);
}

static inline const float* P(const clsph_nn* m, int id) { return m->params + param_offset(id); }
static inline const uint16_t* W(const clsph_nn* m, int id) { return m->wb + param_offset(id); }
static inline float* G(clsph_nn* m, int id) { return m->grad + param_offset(id); }

void forward(clsph_nn* m, const float4* acc0, float gx, float gy, float gz) {
  hipStream_t s = m->ctx->stream;
  int n = (int)m->n, e = (int)m->e;
  size_t nw = (size_t)n * kNodeW;
  gemm(s, false, false, false, m->E1, kFeat, W(m, E_W1), kEdgeW, m->z, kEdgeW, e, kEdgeW, kFeat,
       1.f, P(m, E_B1));
  hipLaunchKernelGGL(k_segsum_relu, grid_for((size_t)n * kEdgeW), dim3(256), 0, s, m->z, m->off,
                     m->n, m->S);
  hipLaunchKernelGGL(k_pack_u, grid_for((size_t)n * kIn), dim3(256), 0, s, m->S, m->nu, acc0,
                     m->degc, m->n, gx, gy, gz, m->u);
  gemm(s, false, false, false, m->u, kIn, W(m, N_W1), kNodeW, m->p1, kNodeW, n, kNodeW, kIn, 1.f,
       P(m, N_B1));
  hipLaunchKernelGGL(k_silu_bf16, grid_for(nw), dim3(256), 0, s, m->p1, m->y1, nw);
  gemm(s, false, false, false, m->y1, kNodeW, W(m, N_W2), kNodeW, m->p2, kNodeW, n, kNodeW,
       kNodeW, 1.f, P(m, N_B2));
  hipLaunchKernelGGL(k_silu_bf16, grid_for(nw), dim3(256), 0, s, m->p2, m->y2, nw);
  gemm(s, false, false, false, m->y2, kNodeW, W(m, N_W3), kOut, m->out, kOut, n, kOut, kNodeW,
       1.f, P(m, N_B3));
}

void forward_fused(clsph_nn* m, clsph_state* st, float4* acc, float* out, bool own_forces,
                   hipEvent_t* ev, bool nu_ready) {
  hipStream_t s = m->ctx->stream;
  uint32_t n = st->n;
  m->n = n;
  const DevParams& d = st->host_dps[0];
  if (ev) (void)hipEventRecord(ev[0], s);
  if (!nu_ready)
    hipLaunchKernelGGL(k_node_inputs, grid_for(n), dim3(256), 0, s, st->rho_s, st->prs_s, n,
                       d.p.fluid_density, d.p.K, (float2*)m->nu);
  // persistent grid: the work-groups that fit on the device at once (the
  // kernel's registers and LDS allow two per compute unit) stage the weights
  // once each
  static int per_cu = 0;
  if (per_cu <= 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_surrogate<true>,
                                                     kSurWaves * 64, 0) != hipSuccess ||
        per_cu <= 0)
      per_cu = 1;
  }
  uint32_t ntiles = (n + kTileP - 1) / kTileP;
  uint32_t blocks = (uint32_t)st->ctx->cus * (uint32_t)per_cu;
  if (blocks * kSurWaves > ntiles) blocks = (ntiles + kSurWaves - 1) / kSurWaves;
  if (blocks == 0) blocks = 1;
  if (own_forces)
    hipLaunchKernelGGL(k_surrogate<true>, dim3(blocks), dim3(kSurWaves * 64), 0, s, st->fab,
                       (const float2*)m->nu, st->rho_s, st->nbr, st->nbr_count, n,
                       edge_consts(d), force_tail(d), m->wb, m->params, acc, out);
  else
    hipLaunchKernelGGL(k_surrogate<false>, dim3(blocks), dim3(kSurWaves * 64), 0, s, st->fab,
                       (const float2*)m->nu, st->rho_s, st->nbr, st->nbr_count, n,
                       edge_consts(d), force_tail(d), m->wb, m->params, acc, out);
  if (ev) (void)hipEventRecord(ev[1], s);
}

// Backward of the node MLP and the edge layer.
void backward(clsph_nn* m) {
  hipStream_t s = m->ctx->stream;
  int n = (int)m->n, e = (int)m->e;
  size_t nw = (size_t)n * kNodeW;
  fill(m->grad, kParamTotal, 0.f, s);
  fill(m->loss, 1, 0.f, s);
  hipLaunchKernelGGL(k_loss, grid_for(n), dim3(256), 0, s, m->out, m->target, m->n, m->dout,
                     m->loss);
  hipLaunchKernelGGL(k_cast, grid_for((size_t)n * kOut), dim3(256), 0, s, m->dout, m->dout_b,
                     (size_t)n * kOut);
  const dim3 nchunks((n + 511) / 512);
  // layer 3
  gemm(s, true, false, true, m->y2, kNodeW, m->dout_b, kOut, G(m, N_W3), kOut, kNodeW, kOut, n,
       1.f);
  hipLaunchKernelGGL((k_colsum<float>), nchunks, dim3(kOut), 0, s, m->dout, kOut, (size_t)n, kOut,
                     G(m, N_B3));
  gemm(s, false, true, false, m->dout_b, kOut, W(m, N_W3), kOut, m->dY, kNodeW, n, kNodeW, kOut,
       1.f);
  hipLaunchKernelGGL(k_dsilu, grid_for(nw), dim3(256), 0, s, m->dY, kNodeW, m->p2, m->dpb,
                     (size_t)n, kNodeW);
  // layer 2
  gemm(s, true, false, true, m->y1, kNodeW, m->dpb, kNodeW, G(m, N_W2), kNodeW, kNodeW, kNodeW, n,
       1.f);
  hipLaunchKernelGGL((k_colsum<uint16_t>), nchunks, dim3(kNodeW), 0, s, m->dpb, kNodeW, (size_t)n,
                     kNodeW, G(m, N_B2));
  gemm(s, false, true, false, m->dpb, kNodeW, W(m, N_W2), kNodeW, m->dY, kNodeW, n, kNodeW,
       kNodeW, 1.f);
  hipLaunchKernelGGL(k_dsilu, grid_for(nw), dim3(256), 0, s, m->dY, kNodeW, m->p1, m->dpb,
                     (size_t)n, kNodeW);
  // layer 1, and d loss / d S (the first 32 inputs) into dY
  gemm(s, true, false, true, m->u, kIn, m->dpb, kNodeW, G(m, N_W1), kNodeW, kIn, kNodeW, n, 1.f);
  hipLaunchKernelGGL((k_colsum<uint16_t>), nchunks, dim3(kNodeW), 0, s, m->dpb, kNodeW, (size_t)n,
                     kNodeW, G(m, N_B1));
  gemm(s, false, true, false, m->dpb, kNodeW, W(m, N_W1), kNodeW, m->dY, kNodeW, n, kEdgeW,
       kNodeW, 1.f);
  // edge layer
  hipLaunchKernelGGL(k_edge_dz, grid_for((size_t)e * kEdgeW), dim3(256), 0, s, m->dY, kNodeW,
                     m->z, m->dst, m->dz, (size_t)e);
  gemm(s, true, false, true, m->E1, kFeat, m->dz, kEdgeW, G(m, E_W1), kEdgeW, kFeat, kEdgeW, e,
       1.f);
  hipLaunchKernelGGL((k_colsum<uint16_t>), dim3((e + 511) / 512), dim3(kEdgeW), 0, s, m->dz,
                     kEdgeW, (size_t)e, kEdgeW, G(m, E_B1));
}

void adam(clsph_nn* m, float lr, float beta1, float beta2, float eps) {
  m->adam_step += 1;
  float bc1 = 1.f - powf(beta1, (float)m->adam_step);
  float bc2 = 1.f - powf(beta2, (float)m->adam_step);
  hipLaunchKernelGGL(k_adam, grid_for(kParamTotal), dim3(256), 0, m->ctx->stream, m->params,
                     m->grad, m->adam_m, m->adam_v, (size_t)kParamTotal, lr, beta1, beta2, eps,
                     bc1, bc2, m->wb);
}

void cast_params(clsph_nn* m) {
  hipLaunchKernelGGL(k_cast, grid_for(kParamTotal), dim3(256), 0, m->ctx->stream, m->params,
                     m->wb, (size_t)kParamTotal);
}

void launch_target_mean(clsph_nn* m, clsph_state* learner, clsph_state* teacher, uint32_t nsub) {
  hipStream_t s = m->ctx->stream;
  uint32_t n = learner->n;
  fill(m->sum, 3 * (size_t)n, 0.f, s);
  for (uint32_t k = 0; k < nsub; ++k) {
    launch_grid_params(teacher);
    launch_hash(teacher);
    launch_sort(teacher);
    launch_density(teacher);
    launch_forces(teacher);
    hipLaunchKernelGGL(k_accum_acc, grid_for(n), dim3(256), 0, s, teacher->acc_s, teacher->val_b,
                       teacher->ids, n, m->sum);
    launch_advect(teacher);
    launch_track_ids(teacher);
  }
  hipLaunchKernelGGL(k_target_mean, grid_for(n), dim3(256), 0, s, learner->val_b, m->sum,
                     learner->acc_s, n, 1.f / (float)nsub, m->target);
}

void launch_target(clsph_nn* m, clsph_state* a, clsph_state* t, float dt) {
  hipStream_t s = m->ctx->stream;
  uint32_t n = a->n;
  hipLaunchKernelGGL(k_invert, grid_for(n), dim3(256), 0, s, t->ids, n, m->loc);
  hipLaunchKernelGGL(k_target, grid_for(n), dim3(256), 0, s, a->val_b, a->vh, m->loc, t->vh,
                     a->acc_s, n, 1.f / dt, m->target);
}

// Workspace of clsph_nn_gemm_noalloc: the device copies of its operands.
// Every GPU the library runs on holds one copy of this array until the
// process ends. Host code never passes the array's name to a kernel or a
// copy: the name is a host-side stand-in, and the device address comes from
// hipGetSymbolAddress below, per device.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

hipError_t scratch_address(int device, void** ptr) {
  static std::mutex lock;
  static std::vector<void*> addr;  // per device, looked up once
  std::lock_guard<std::mutex> g(lock);
  if (device < 0) return hipErrorInvalidDevice;
  if ((size_t)device >= addr.size()) addr.resize((size_t)device + 1, nullptr);
  if (!addr[device]) {
    // the caller has made `device` current
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

void gemm_public(hipStream_t s, bool ta, bool tb, bool add, const uint16_t* A, int lda,
                 const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K, float alpha,
                 const float* bias, const float* rowscale) {
  gemm(s, ta, tb, add, A, lda, B, ldb, C, ldc, M, N, K, alpha, bias, rowscale);
}

void launch_compare(clsph_nn* m, clsph_state* a, clsph_state* b, float res[2]) {
  hipStream_t s = m->ctx->stream;
  uint32_t n = a->n;
  hipLaunchKernelGGL(k_invert, grid_for(n), dim3(256), 0, s, b->ids, n, m->loc);
  hipLaunchKernelGGL(k_invert_perm, grid_for(n), dim3(256), 0, s, b->val_b, n, m->inv);
  uint32_t blocks = (n + 255) / 256;
  if (blocks > (uint32_t)kCmpBlocks) blocks = kCmpBlocks;
  if (blocks == 0) blocks = 1;
  // the partials go to m->sum: 2 ceil(n / 256) <= 2 n <= 3 n_cap floats
  hipLaunchKernelGGL(k_compare, dim3(blocks), dim3(256), 0, s, a->val_b, a->ids, a->rho_s, m->loc,
                     m->inv, b->rho_s, n, m->sum);
  hipLaunchKernelGGL(k_compare_final, dim3(1), dim3(256), 0, s, m->sum, blocks, m->cmp);
  (void)hipMemcpyAsync(res, m->cmp, 2 * sizeof(float), hipMemcpyDeviceToHost, s);
}

}  // namespace nn
}  // namespace clsph
