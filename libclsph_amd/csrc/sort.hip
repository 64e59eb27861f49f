// Stable LSD radix sort of (key, particle index) pairs.
//
// The reference sorts whole 80-byte particles with 4 passes of 8-bit digits
// over 128 work-items and a host scan (sort.cl:23-59,
// sph_simulation.cpp:107-156); its result is the stable order by key. Here
// the pairs are sorted in kSortPasses passes whose digit width is chosen on
// the device from the key width, so a captured step needs no host round
// trip. Each pass: per-tile digit counts laid out digit-major, an exclusive
// scan of each digit's counts over the tiles plus the digit totals (together
// the same prefix as the reference's host scan over that layout), and a
// scatter that ranks keys stably inside a tile with wave64 ballot matching.
#include "state.h"

namespace clsph {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 4;  // keys per thread
constexpr int kTile = kSortBlock * kSortItems;
constexpr int kWaves = kSortBlock / 64;

uint32_t sort_tiles_for(uint32_t n) { return (n + kTile - 1) / kTile; }

__device__ inline uint32_t digit_of(uint32_t key, uint32_t shift, uint32_t mask) {
  return (key >> shift) & mask;
}

// Digit width of this sort, clamped to max_db (the width the launch sized its
// shared memory for; a wider one only arises with the table-overflow error set).
__device__ inline void pass_digit(const BatchParams* __restrict__ bp, uint32_t max_db, int pass,
                                  uint32_t& nb, uint32_t& mask, uint32_t& shift) {
  uint32_t db = bp->digit_bits;
  db = db > max_db ? max_db : db;
  nb = 1u << db;
  mask = nb - 1;
  shift = db * (uint32_t)pass;
}

// Per-tile digit histogram, stored digit-major: hist[d * n_tiles + tile].
__global__ __launch_bounds__(kSortBlock) void k_sort_count(const uint32_t* __restrict__ keys,
                                                           uint32_t n,
                                                           const BatchParams* __restrict__ bp,
                                                           uint32_t max_db, int pass,
                                                           uint32_t* __restrict__ hist,
                                                           uint32_t n_tiles) {
  extern __shared__ uint32_t bins[];  // 1 << max_db entries
  uint32_t nb, mask, shift;
  pass_digit(bp, max_db, pass, nb, mask, shift);
  for (uint32_t d = threadIdx.x; d < nb; d += kSortBlock) bins[d] = 0;
  __syncthreads();
  uint32_t base = blockIdx.x * kTile;
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    uint32_t i = base + it * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&bins[digit_of(keys[i], shift, mask)], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < nb; d += kSortBlock)
    hist[d * n_tiles + blockIdx.x] = bins[d];
}

// Block-wide exclusive scan of one value per thread (blockDim.x = B).
template <int B>
__device__ inline uint32_t block_exclusive_scan(uint32_t v) {
  __shared__ uint32_t wsum[B / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t off = 0;
  for (int q = 0; q < w; ++q) off += wsum[q];
  __syncthreads();
  return off + x - v;
}

// One block per digit: exclusive scan of that digit's per-tile counts, in
// place, and the digit's total.
__global__ __launch_bounds__(kSortBlock) void k_sort_scan(uint32_t* __restrict__ hist,
                                                          uint32_t n_tiles,
                                                          const BatchParams* __restrict__ bp,
                                                          uint32_t max_db,
                                                          uint32_t* __restrict__ totals) {
  uint32_t d = blockIdx.x;
  uint32_t nb, mask, shift;
  pass_digit(bp, max_db, 0, nb, mask, shift);
  if (d >= nb) return;
  uint32_t* row = hist + (size_t)d * n_tiles;
  uint32_t per = (n_tiles + kSortBlock - 1) / kSortBlock;
  uint32_t s = threadIdx.x * per, e = min(s + per, n_tiles);
  uint32_t local = 0;
  for (uint32_t k = s; k < e; ++k) local += row[k];
  uint32_t run = block_exclusive_scan<kSortBlock>(local);
  for (uint32_t k = s; k < e; ++k) {
    uint32_t c = row[k];
    row[k] = run;
    run += c;
  }
  if (threadIdx.x == kSortBlock - 1) totals[d] = run;
}

// Stable scatter. Wave w of the tile owns keys [w*256, w*256+256) of the tile,
// visited in 4 rounds of 64 consecutive keys, so (wave, round, lane) order is
// input order. Within a round, lanes with equal digits are found by matching
// the digit bit by bit with ballots.
__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n,
    const BatchParams* __restrict__ bp, uint32_t max_db, int pass,
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ totals,
    uint32_t n_tiles) {
  // Shared memory: kWaves running counts per digit, then the digit starts.
  extern __shared__ uint32_t smem[];
  const uint32_t nbmax = 1u << max_db;
  uint32_t* wcnt = smem;                         // wcnt[q * nbmax + d]
  uint32_t* digit_start = smem + kWaves * nbmax;
  uint32_t nb, mask, shift;
  pass_digit(bp, max_db, pass, nb, mask, shift);
  uint32_t db = 31u - (uint32_t)__builtin_clz(nb);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t d = threadIdx.x; d < nb; d += kSortBlock)
    for (int q = 0; q < kWaves; ++q) wcnt[q * nbmax + d] = 0;
  // Exclusive scan of the digit totals: where each digit's run starts.
  {
    constexpr int per = kMaxBins / kSortBlock;
    uint32_t v[per];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
      uint32_t d = threadIdx.x * per + k;
      v[k] = d < nb ? totals[d] : 0u;
      local += v[k];
    }
    uint32_t run = block_exclusive_scan<kSortBlock>(local);
#pragma unroll
    for (int k = 0; k < per; ++k) {
      uint32_t d = threadIdx.x * per + k;
      if (d < nb) digit_start[d] = run;
      run += v[k];
    }
  }
  __syncthreads();

  uint32_t kk[kSortItems], vv[kSortItems], dd[kSortItems], rr[kSortItems];
  uint32_t base = blockIdx.x * kTile + w * (kTile / kWaves);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int t = 0; t < kSortItems; ++t) {
    uint32_t i = base + t * 64 + lane;
    bool valid = i < n;
    uint32_t key = valid ? keys_in[i] : 0u;
    uint32_t val = valid ? vals_in[i] : 0u;
    uint32_t dig = digit_of(key, shift, mask);
    uint64_t m = __builtin_amdgcn_ballot_w64(valid);
    for (uint32_t b = 0; b < db; ++b) {
      bool bit = (dig >> b) & 1u;
      uint64_t bal = __builtin_amdgcn_ballot_w64(bit);
      m &= bit ? bal : ~bal;
    }
    uint32_t peer = (uint32_t)__builtin_popcountll(m & lt);
    uint32_t cnt = (uint32_t)__builtin_popcountll(m);
    // The lowest lane of each digit group claims cnt slots of the wave's
    // running count and hands the old value to its peers.
    uint32_t leader = m ? (uint32_t)__builtin_ctzll(m) : 0u;
    uint32_t prev = 0;
    if (valid && peer == 0) prev = atomicAdd(&wcnt[w * nbmax + dig], cnt);
    prev = __shfl(prev, (int)leader, 64);
    kk[t] = key;
    vv[t] = val;
    dd[t] = valid ? dig : 0xFFFFFFFFu;
    rr[t] = prev + peer;
  }
  __syncthreads();
  // Exclusive prefix over waves, per digit.
  for (uint32_t d = threadIdx.x; d < nb; d += kSortBlock) {
    uint32_t run = 0;
    for (int q = 0; q < kWaves; ++q) {
      uint32_t c = wcnt[q * nbmax + d];
      wcnt[q * nbmax + d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kSortItems; ++t) {
    uint32_t dig = dd[t];
    if (dig == 0xFFFFFFFFu) continue;
    uint32_t p = digit_start[dig] + offsets[(size_t)dig * n_tiles + blockIdx.x] +
                 wcnt[w * nbmax + dig] + rr[t];
    keys_out[p] = kk[t];
    vals_out[p] = vv[t];
  }
}

void launch_sort(clsph_state* s) {
  if (s->n == 0) return;
  hipStream_t st = s->ctx->stream;
  uint32_t nt = sort_tiles_for(s->n);
  uint32_t* kin = s->key_a;
  uint32_t* vin = s->val_a;
  uint32_t* kout = s->key_b;
  uint32_t* vout = s->val_b;
  // Widest digit the state can need: keys stay below table_cap while the
  // cells fit the table.
  uint32_t kb = bitlen(s->table_cap);
  uint32_t max_db = (kb + kSortPasses - 1) / kSortPasses;
  max_db = max_db > (uint32_t)kMaxDigitBits ? (uint32_t)kMaxDigitBits : max_db;
  uint32_t nbmax = 1u << max_db;
  size_t count_lds = sizeof(uint32_t) * nbmax;
  size_t scatter_lds = sizeof(uint32_t) * nbmax * (kWaves + 1);
  for (int pass = 0; pass < kSortPasses; ++pass) {
    hipLaunchKernelGGL(k_sort_count, dim3(nt), dim3(kSortBlock), count_lds, st, kin, s->n,
                       s->bp, max_db, pass, s->hist, nt);
    hipLaunchKernelGGL(k_sort_scan, dim3(nbmax), dim3(kSortBlock), 0, st, s->hist, nt, s->bp,
                       max_db, s->digit_base);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nt), dim3(kSortBlock), scatter_lds, st, kin, vin,
                       kout, vout, s->n, s->bp, max_db, pass, s->hist, s->digit_base, nt);
    uint32_t* tk = kin;
    kin = kout;
    kout = tk;
    uint32_t* tv = vin;
    vin = vout;
    vout = tv;
  }
  static_assert(kSortPasses % 2 == 1, "sorted output must land in key_b/val_b");
}

}  // namespace clsph
