// Stable sort of the particles by cell id, then by octant within each cell;
// the half-cell table, the canonical slots and the gather of the state into
// sorted order.
//
// The reference sorts whole 80-byte particles with 4 LSD passes of 8-bit
// digits over 128 work-items and a host scan, then builds the cell table on
// the host (sort.cl:23-59, sph_simulation.cpp:107-171); its result is the
// order stable by cell id. Here:
//
//  1. Two (or three) LSD passes over the cell-id bits (key >> 3), each of
//     k_lsd_count (per-tile digit counts, one coalesced row per tile of 2048
//     keys), k_lsd_scan (each digit's counts scanned over the tiles) and
//     k_lsd_scatter (keys ranked in registers with wave64 ballot matching,
//     the tile reordered digit-major in LDS, each digit run written to
//     consecutive addresses). After the passes the particles are in the
//     reference's order, with their canonical indices. A state whose table
//     capacity needs three passes launches three, and its third pass does
//     nothing in a substep whose cell ids fit two (22 bits): the pass count
//     of the substep is set on the device with the digit width.
//  2. k_cells: each cell's particles are contiguous in that order; a
//     work-group takes the cells starting in its 256 positions and orders
//     each cell by octant (stably), which gives the sorted (half-cell)
//     order, the half-cell table table[f] = #{keys < f}, and for each sorted
//     position its position in the reference's order (the canonical slot,
//     where advection writes the particle back). It gathers the state into
//     sorted order on the way.
#include "state.h"

namespace clsph {

constexpr int kLsdBlock = 256;
constexpr int kLsdItems = 8;                         // keys per thread
constexpr int kLsdTile = kLsdBlock * kLsdItems;      // 2048 keys per tile
constexpr int kScanBlock = 1024;                     // 16 digits x 64 slices of the tiles

constexpr int kCellBlock = 256;

uint32_t sort_tiles_for(uint32_t n) { return (n + kLsdTile - 1) / kLsdTile; }

// Lanes of the wave whose `bits`-bit value equals this lane's (0 for lanes
// that are not `valid`). Every lane of the wave must call it.
__device__ inline uint64_t match_lanes(uint32_t v, uint32_t bits, bool valid) {
  uint64_t m = __builtin_amdgcn_ballot_w64(valid);
  for (uint32_t b = 0; b < bits; ++b) {
    const bool bit = (v >> b) & 1u;
    const uint64_t bal = __builtin_amdgcn_ballot_w64(bit);
    m &= bit ? bal : ~bal;
  }
  return valid ? m : 0ull;
}

__device__ inline uint64_t lanes_below(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Block-wide exclusive scan of one value per thread (blockDim.x = B).
template <int B>
__device__ inline uint32_t block_exclusive_scan(uint32_t v) {
  __shared__ uint32_t wsum[B / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t off = 0;
  for (int q = 0; q < w; ++q) off += wsum[q];
  __syncthreads();
  return off + x - v;
}

// Digit of pass `pass` of a half-cell key: bits [3 + db * pass, +db), the
// cell id's.
__device__ inline uint32_t pass_digit(uint32_t key, uint32_t db, uint32_t pass) {
  return (key >> (3u + db * pass)) & ((1u << db) - 1u);
}

// ---- 1. LSD passes ----------------------------------------------------------

// Digit counts of each tile: hist[tile * nb + d].
__global__ __launch_bounds__(kLsdBlock) void k_lsd_count(const uint32_t* __restrict__ keys,
                                                         uint32_t n,
                                                         const BatchParams* __restrict__ bp,
                                                         uint32_t pass,
                                                         uint32_t* __restrict__ hist,
                                                         uint32_t* __restrict__ runs_ctl) {
  __shared__ uint32_t bins[kMaxBins];
  // the first pass also empties the cell pass's list of long runs
  if (pass == 0 && blockIdx.x == 0 && threadIdx.x == 0) runs_ctl[0] = 0;
  if (pass >= lsd_passes_of(bp)) return;
  const uint32_t db = bp->digit_bits, nb = 1u << db;
  for (uint32_t d = threadIdx.x; d < nb; d += kLsdBlock) bins[d] = 0;
  const int lane = threadIdx.x & 63;
  const uint32_t base = blockIdx.x * kLsdTile;
  uint32_t kk[kLsdItems];
#pragma unroll
  for (int t = 0; t < kLsdItems; ++t) {
    const uint32_t i = base + t * kLsdBlock + threadIdx.x;
    kk[t] = i < n ? keys[i] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kLsdItems; ++t) {
    const uint32_t i = base + t * kLsdBlock + threadIdx.x;
    const bool valid = i < n;
    // one LDS add per run of equal digits among the wave's consecutive keys
    // (valid lanes are a prefix of the wave)
    const uint32_t d = pass_digit(kk[t], db, pass);
    const uint32_t prev = __shfl_up(d, 1, 64);
    const bool head = valid && (lane == 0 || prev != d);
    const uint64_t heads = __builtin_amdgcn_ballot_w64(head);
    const uint32_t nvalid = (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(valid));
    const uint64_t after = heads & ~((2ull << lane) - 1ull);
    const uint32_t next = after ? (uint32_t)__builtin_ctzll(after) : nvalid;
    if (head) atomicAdd(&bins[d], next - (uint32_t)lane);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < nb; d += kLsdBlock) hist[(size_t)blockIdx.x * nb + d] = bins[d];
}

// Each digit's counts scanned over the tiles, in place (exclusive), and the
// digit totals. A work-group takes 16 consecutive digits and cuts the tiles
// into 64 slices, one thread per (slice, digit): sums per slice, joined
// through LDS, then the running sums written back. REGS > 0: a slice has at
// most REGS tiles, whose counts are all loaded at once and kept in registers
// for the write-back (one round trip to memory each way); REGS = 0: any
// number of tiles, loaded twice.
template <uint32_t REGS>
__global__ __launch_bounds__(kScanBlock) void k_lsd_scan(uint32_t* __restrict__ hist,
                                                         uint32_t n_tiles,
                                                         const BatchParams* __restrict__ bp,
                                                         uint32_t pass,
                                                         uint32_t* __restrict__ totals) {
  constexpr uint32_t kDigits = 16, kSlices = kScanBlock / kDigits;
  __shared__ uint32_t part[kSlices][kDigits];
  const uint32_t nb = 1u << bp->digit_bits;
  if (blockIdx.x * kDigits >= nb || pass >= lsd_passes_of(bp)) return;
  const uint32_t dl = threadIdx.x % kDigits, g = threadIdx.x / kDigits;
  const uint32_t d = blockIdx.x * kDigits + dl;
  const bool on = d < nb;
  const uint32_t per = (n_tiles + kSlices - 1) / kSlices;
  const uint32_t t0 = min(g * per, n_tiles), t1 = min(t0 + per, n_tiles);
  uint32_t* col = hist + d;
  constexpr uint32_t R = REGS ? REGS : 1u;
  uint32_t c[R];
  uint32_t sum = 0;
  if (on) {
    if constexpr (REGS != 0) {
#pragma unroll
      for (uint32_t k = 0; k < R; ++k) {
        c[k] = t0 + k < t1 ? col[(size_t)(t0 + k) * nb] : 0u;
        sum += c[k];
      }
    } else {
#pragma unroll 8
      for (uint32_t t = t0; t < t1; ++t) sum += col[(size_t)t * nb];
    }
  }
  part[g][dl] = sum;
  __syncthreads();
  uint32_t run = 0, all = 0;
  for (uint32_t q = 0; q < kSlices; ++q) {
    const uint32_t v = part[q][dl];
    run += q < g ? v : 0u;
    all += v;
  }
  if (on) {
    if constexpr (REGS != 0) {
#pragma unroll
      for (uint32_t k = 0; k < R; ++k) {
        if (t0 + k < t1) col[(size_t)(t0 + k) * nb] = run;
        run += c[k];
      }
    } else {
#pragma unroll 8
      for (uint32_t t = t0; t < t1; ++t) {
        const size_t at = (size_t)t * nb;
        const uint32_t v = col[at];
        col[at] = run;
        run += v;
      }
    }
    if (g == 0) totals[d] = all;
  }
}

// Stable scatter of (key, value) by the digit of pass `pass`; vals_in ==
// nullptr stands for the identity (value = input index). A tile's 2048 keys
// are spread over 8 waves, so that each wave's ranking chain is short: wave
// w owns keys [w*256, w*256+256) of the tile, visited in 4 rounds of 64
// consecutive keys, so (wave, round, lane) order is input order.
//
// LDS, 45 KB (three work-groups, 24 waves per CU): 16-bit per-wave digit
// counts while the keys are ranked, the first half of the same bytes then
// holding the tile's keys and values in digit order; each digit's run start
// in that copy (16-bit), and its output start minus that run start.
constexpr int kScatBlock = 512;                         // scatter threads per tile
constexpr int kScatItems = kLsdTile / kScatBlock;       // keys per thread
constexpr int kScatWaves = kScatBlock / 64;
constexpr int kScatWaveKeys = kLsdTile / kScatWaves;    // consecutive keys per wave
__global__ __launch_bounds__(kScatBlock) void k_lsd_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t n,
    const BatchParams* __restrict__ bp, uint32_t pass, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ totals, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out) {
  static_assert(kScatWaves * kMaxBins * sizeof(uint16_t) >= 2 * kLsdTile * sizeof(uint32_t),
                "the tile copy fits in the bytes of the per-wave counts");
  constexpr int kStageQ = kScatWaves * kMaxBins * (int)sizeof(uint16_t) / 16;
  __shared__ uint4 stage[kStageQ];
  __shared__ uint16_t tile_start[kMaxBins];   // digit runs in the LDS copy of the tile
  __shared__ uint32_t dshift[kMaxBins];       // output start of a digit's run - tile_start
  uint16_t(*wcnt)[kMaxBins] = reinterpret_cast<uint16_t(*)[kMaxBins]>(stage);
  uint32_t* skey = reinterpret_cast<uint32_t*>(stage);
  uint32_t* sval = skey + kLsdTile;
  if (pass >= lsd_passes_of(bp)) return;
  const uint32_t db = bp->digit_bits, nb = 1u << db;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t base = blockIdx.x * kLsdTile + w * kScatWaveKeys;
  uint32_t kk[kScatItems], vv[kScatItems], rr[kScatItems];
#pragma unroll
  for (int t = 0; t < kScatItems; ++t) {
    const uint32_t i = base + t * 64 + lane;
    kk[t] = i < n ? keys_in[i] : 0u;
    vv[t] = i < n ? (vals_in ? vals_in[i] : i) : 0u;
  }
  for (uint32_t q = threadIdx.x; q < kStageQ; q += kScatBlock)
    stage[q] = make_uint4(0u, 0u, 0u, 0u);
  // where each digit's run of this tile starts in the output: the digit's
  // start (exclusive scan of the totals) plus the tiles before this one
  constexpr int per = kMaxBins / kScatBlock;
  {
    uint32_t v[per], o[per];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
      const uint32_t d = threadIdx.x * per + k;
      v[k] = d < nb ? totals[d] : 0u;
      o[k] = d < nb ? offsets[(size_t)blockIdx.x * nb + d] : 0u;
      local += v[k];
    }
    uint32_t run = block_exclusive_scan<kScatBlock>(local);
#pragma unroll
    for (int k = 0; k < per; ++k) {
      const uint32_t d = threadIdx.x * per + k;
      if (d < nb) dshift[d] = run + o[k];
      run += v[k];
    }
  }
  __syncthreads();
  const uint64_t lt = lanes_below(lane);
#pragma unroll
  for (int t = 0; t < kScatItems; ++t) {
    const uint32_t i = base + t * 64 + lane;
    const bool valid = i < n;
    const uint32_t d = pass_digit(kk[t], db, pass);
    const uint64_t m = match_lanes(d, db, valid);
    const uint32_t peer = (uint32_t)__builtin_popcountll(m & lt);
    // the lowest lane of each digit group takes its count from the wave's
    // running count (distinct groups touch distinct entries)
    uint32_t prev = 0;
    if (valid && peer == 0) {
      prev = wcnt[w][d];
      wcnt[w][d] = (uint16_t)(prev + (uint32_t)__builtin_popcountll(m));
    }
    prev = __shfl(prev, m ? (int)__builtin_ctzll(m) : 0, 64);
    rr[t] = prev + peer;
  }
  __syncthreads();
  // per digit: the waves' counts turned into their prefix, the tile's count,
  // and the digit's run start in the tile (at most 2048: 16 bits)
  {
    uint32_t c[per];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
      const uint32_t d = threadIdx.x * per + k;
      uint32_t run = 0;
      if (d < nb) {
#pragma unroll
        for (int q = 0; q < kScatWaves; ++q) {
          const uint32_t cq = wcnt[q][d];
          wcnt[q][d] = (uint16_t)run;
          run += cq;
        }
      }
      c[k] = run;
      local += run;
    }
    uint32_t run = block_exclusive_scan<kScatBlock>(local);
#pragma unroll
    for (int k = 0; k < per; ++k) {
      const uint32_t d = threadIdx.x * per + k;
      if (d < nb) {
        tile_start[d] = (uint16_t)run;
        dshift[d] -= run;
      }
      run += c[k];
    }
  }
  __syncthreads();
  // positions in the tile copy, then the copy (over the counts)
#pragma unroll
  for (int t = 0; t < kScatItems; ++t) {
    const uint32_t d = pass_digit(kk[t], db, pass);
    rr[t] += (uint32_t)tile_start[d] + (uint32_t)wcnt[w][d];
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kScatItems; ++t) {
    const uint32_t i = base + t * 64 + lane;
    if (i >= n) continue;
    skey[rr[t]] = kk[t];
    sval[rr[t]] = vv[t];
  }
  __syncthreads();
  const uint32_t t0 = blockIdx.x * kLsdTile;
  const uint32_t count = min((uint32_t)kLsdTile, n - t0);
  for (uint32_t j = threadIdx.x; j < count; j += kScatBlock) {
    const uint32_t key = skey[j];
    const uint32_t dst = dshift[pass_digit(key, db, pass)] + j;
    keys_out[dst] = key;
    vals_out[dst] = sval[j];
  }
}

// ---- 2. cells: octant order, table, slots, gather --------------------------

// Work-group g owns the cells whose first particle (in the reference's order
// sk / sv, stable by cell id) lies in [g * kCellRange, (g + 1) * kCellRange);
// [a, b) are those cells' particles, and also their sorted positions. The
// keys from just before the range to kCellAhead past it are staged in LDS
// (a cell running further is read from HBM). Two passes over [a, b), 256 at
// a time: the first records each cell's start and the octant counts before
// it; the second gives every particle its sorted position
//   cell start + particles of lower octants in the cell
//              + earlier particles of its octant in the cell,
// kept in LDS as the inverse map (sorted position - a -> reference
// position), so that keys, permutation, slots and the gathered state are
// written in sorted order, each output line by one work-group. (Owned spans
// longer than kCellSpan, cells of hundreds of particles straddling a range,
// write each particle where it goes instead.) The table entries of each
// owned cell and of the empty keys after it (up to the next cell, or through
// key_total) are written between the passes. With an error flag raised the
// table is left alone and slot[k] = k.
// 256 owned positions per work-group (rather than 512) keep more work-groups
// resident per CU to overlap the passes' load latencies: 5 us less per pass
// at 1M particles.
constexpr int kCellRange = 256;                     // positions whose cells a work-group owns
constexpr int kCellAhead = 256;                     // positions past them staged for the last cell
constexpr int kCellSpan = kCellRange + kCellAhead;  // [a, b) inverted in LDS up to this
constexpr int kLongRun = 4096;  // empty keys a work-group writes itself; longer runs are listed
constexpr int kFillBlocks = 512;

// Writes the particle at reference position k (key, canonical index src) to
// sorted position dst.
__device__ inline void put_sorted(uint32_t dst, uint32_t k, uint32_t key, uint32_t src,
                                  bool table_ok, uint32_t* __restrict__ keys_out,
                                  uint32_t* __restrict__ vals_out, uint32_t* __restrict__ slot,
                                  const float4* __restrict__ pos, const float4* __restrict__ vel,
                                  const float4* __restrict__ vh, float4* __restrict__ pos_s,
                                  float4* __restrict__ vel_s, float4* __restrict__ vh_s) {
  keys_out[dst] = key;
  vals_out[dst] = src;
  slot[dst] = table_ok ? k : dst;
  pos_s[dst] = pos[src];
  vel_s[dst] = vel[src];
  vh_s[dst] = vh[src];
}

// Pass 1, the table and pass 2 of k_cells over [a, b), with the keys read
// from the LDS copy (stk[p + 1 - r0]) or, when kStaged is false, from HBM.
template <bool kStaged>
__device__ __forceinline__ void cells_place(
    const uint32_t* stk, const uint32_t* __restrict__ sk, const uint32_t* __restrict__ sv,
    uint32_t n, uint32_t r0, uint32_t a, uint32_t b, uint32_t total, bool table_ok, uint32_t* inv,
    uint32_t* c_start, uint32_t* c_id, uint32_t (*c_oct)[8], uint32_t (*wtot)[9],
    uint32_t (*pre)[9], uint32_t* carry, uint4* __restrict__ runs, uint32_t* __restrict__ runs_ctl,
    uint32_t runs_cap,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t* __restrict__ slot,
    uint32_t* __restrict__ table, const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float4* __restrict__ vh, float4* __restrict__ pos_s, float4* __restrict__ vel_s,
    float4* __restrict__ vh_s) {
  constexpr int kMaxCells = kCellRange;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = lanes_below(lane);
  const uint64_t le = lt | (1ull << lane);
#define KEY_AT(p) (kStaged ? stk[(p) + 1u - r0] : sk[p])
#define STARTS(p) ((p) == 0 || (KEY_AT(p) >> 3) != (KEY_AT((p) - 1u) >> 3))
  // One chunk of 256 positions from k0: the key, whether the position starts
  // a cell, the wave's ballots of cell starts (sm) and of each octant (bq),
  // and in pre[w][0..7] / pre[w][8] the octant counts and cell starts in
  // [a, start of this wave's 64 positions) (carry[] holds those counts for
  // the chunk start and is advanced past the chunk).
#define CELLS_CHUNK(k0, end, key, valid, st, sm, bq)                                  \
  do {                                                                                \
    const uint32_t k_ = (k0) + threadIdx.x;                                          \
    valid = k_ < (end);                                                              \
    key = valid ? KEY_AT(k_) : 0u;                                              \
    st = valid && STARTS(k_);                                                   \
    const uint32_t o_ = key & 7u;                                                     \
    sm = __builtin_amdgcn_ballot_w64(st);                                             \
    _Pragma("unroll") for (uint32_t q = 0; q < 8; ++q) bq[q] =                        \
        __builtin_amdgcn_ballot_w64(valid && o_ == q);                                \
    if (lane == 0) {                                                                  \
      _Pragma("unroll") for (int q = 0; q < 8; ++q) wtot[w][q] =                      \
          (uint32_t)__builtin_popcountll(bq[q]);                                      \
      wtot[w][8] = (uint32_t)__builtin_popcountll(sm);                                \
    }                                                                                 \
    __syncthreads();                                                                  \
    if (threadIdx.x < 9) {                                                            \
      uint32_t run_ = carry[threadIdx.x];                                             \
      for (int v = 0; v < kCellBlock / 64; ++v) {                                     \
        pre[v][threadIdx.x] = run_;                                                   \
        run_ += wtot[v][threadIdx.x];                                                 \
      }                                                                               \
      carry[threadIdx.x] = run_;                                                      \
    }                                                                                 \
    __syncthreads();                                                                  \
  } while (0)

  // pass 1: cell starts and the octant counts before them
  if (threadIdx.x < 9) carry[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t k0 = a; k0 < b; k0 += kCellBlock) {
    uint32_t key;
    bool valid, st;
    uint64_t sm, bq[8];
    CELLS_CHUNK(k0, b, key, valid, st, sm, bq);
    if (st) {
      const uint32_t lc = pre[w][8] + (uint32_t)__builtin_popcountll(sm & lt);
      c_start[lc] = k0 + threadIdx.x;
      c_id[lc] = key >> 3;
#pragma unroll
      for (int q = 0; q < 8; ++q) c_oct[lc][q] = pre[w][q] + (uint32_t)__builtin_popcountll(bq[q] & lt);
    }
  }
  __syncthreads();
  const uint32_t ncells = carry[8];
  if (threadIdx.x == 0) {
    c_start[ncells] = b;
#pragma unroll
    for (int q = 0; q < 8; ++q) c_oct[ncells][q] = carry[q];
  }
  __syncthreads();

  // table over the work-group's keys: from its first cell (from 0 for the
  // first work-group) up to the next work-group's first cell (through
  // key_total for the last one), the owned cells' keys first, then the empty
  // keys after each cell. A key of an owned cell holds the start of its
  // octant run; an empty key the start of the next cell. Runs of more than
  // kLongRun empty keys (the sparse Morton range leaves a few of hundreds of
  // thousands) are listed for k_fill_runs, which spreads them over the chip.
  if (table_ok) {
    // the eight keys of every owned cell, one (cell, octant) per thread: the
    // cell's start plus its particles of lower octants
    for (uint32_t e = threadIdx.x; e < 8u * ncells; e += kCellBlock) {
      const uint32_t j = e >> 3, q = e & 7u;
      uint32_t r = c_start[j];
#pragma unroll
      for (uint32_t p = 0; p < 7u; ++p) r += p < q ? c_oct[j + 1][p] - c_oct[j][p] : 0u;
      table[(c_id[j] << 3) + q] = r;
    }
    // the empty keys before the first cell and after each cell
    const uint32_t kend = b < n ? (KEY_AT(b) & ~7u) : total + 1u;
    for (int j = -1; j < (int)ncells; ++j) {
      uint32_t f0, f1, v;
      if (j < 0) {
        if (a != 0) continue;  // the keys before the first cell: the first work-group's
        f0 = 0, f1 = c_id[0] << 3, v = 0;
      } else {
        f0 = (c_id[j] << 3) + 8u;
        f1 = j + 1 < (int)ncells ? (c_id[j + 1] << 3) : kend;
        v = c_start[j + 1];
      }
      if (f1 - f0 > (uint32_t)kLongRun && f1 > f0) {
        if (threadIdx.x == 0) {
          const uint32_t i = atomicAdd(&runs_ctl[0], 1u);
          if (i < runs_cap)
            runs[i] = make_uint4(f0, f1, v, 0u);
          else
            for (uint32_t f = f0; f < f1; ++f) table[f] = v;  // list full: written here
        }
      } else {
        for (uint32_t f = f0 + threadIdx.x; f < f1; f += kCellBlock) table[f] = v;
      }
    }
  }
  __syncthreads();

  // each cell's c_oct row becomes the offsets that turn the octant counts
  // before a particle into its sorted position
  uint32_t offs[kMaxCells / kCellBlock][8];
#pragma unroll
  for (int i = 0; i < kMaxCells / kCellBlock; ++i) {
    const uint32_t lc = threadIdx.x + i * kCellBlock;
    if (lc >= ncells) continue;
    uint32_t run = c_start[lc];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      offs[i][q] = run - c_oct[lc][q];
      run += c_oct[lc + 1][q] - c_oct[lc][q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kMaxCells / kCellBlock; ++i) {
    const uint32_t lc = threadIdx.x + i * kCellBlock;
    if (lc < ncells)
#pragma unroll
      for (int q = 0; q < 8; ++q) c_oct[lc][q] = offs[i][q];
  }
  if (threadIdx.x < 9) carry[threadIdx.x] = 0;
  __syncthreads();

  // pass 2: every particle's sorted position
  const bool through_lds = b - a <= (uint32_t)kCellSpan;
  for (uint32_t k0 = a; k0 < b; k0 += kCellBlock) {
    uint32_t key;
    bool valid, st;
    uint64_t sm, bq[8];
    CELLS_CHUNK(k0, b, key, valid, st, sm, bq);
    if (valid) {
      const uint32_t o = key & 7u;
      uint64_t mine = 0;
#pragma unroll
      for (uint32_t q = 0; q < 8; ++q) mine = q == o ? bq[q] : mine;
      const uint32_t lc = pre[w][8] + (uint32_t)__builtin_popcountll(sm & le) - 1u;
      const uint32_t dst = c_oct[lc][o] + pre[w][o] + (uint32_t)__builtin_popcountll(mine & lt);
      const uint32_t k = k0 + threadIdx.x;
      if (through_lds)
        inv[dst - a] = k;
      else
        put_sorted(dst, k, key, sv[k], table_ok, keys_out, vals_out, slot, pos, vel, vh, pos_s,
                   vel_s, vh_s);
    }
  }
  if (!through_lds) return;
  __syncthreads();
  // write in sorted order
#pragma unroll 4
  for (uint32_t i = a + threadIdx.x; i < b; i += kCellBlock) {
    const uint32_t k = inv[i - a];
    put_sorted(i, k, KEY_AT(k), sv[k], table_ok, keys_out, vals_out, slot, pos, vel, vh, pos_s,
               vel_s, vh_s);
  }
#undef CELLS_CHUNK
#undef STARTS
#undef KEY_AT
}

__global__ __launch_bounds__(kCellBlock) void k_cells(
    const uint32_t* __restrict__ sk_odd, const uint32_t* __restrict__ sv_odd,
    const uint32_t* __restrict__ sk_even, const uint32_t* __restrict__ sv_even, uint32_t n,
    const BatchParams* __restrict__ bp, uint4* __restrict__ runs, uint32_t* __restrict__ runs_ctl,
    uint32_t runs_cap, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t* __restrict__ slot, uint32_t* __restrict__ table,
    const float4* __restrict__ pos, const float4* __restrict__ vel, const float4* __restrict__ vh,
    float4* __restrict__ pos_s, float4* __restrict__ vel_s, float4* __restrict__ vh_s) {
  constexpr int kMaxCells = kCellRange;
  constexpr int kStage = kCellSpan + 1;          // keys at r0 - 1 .. r1 + kCellAhead
  __shared__ uint32_t stk[kStage];
  __shared__ uint32_t inv[kCellSpan];
  __shared__ uint32_t c_start[kMaxCells + 1];    // first position of each owned cell
  __shared__ uint32_t c_id[kMaxCells];
  __shared__ uint32_t c_oct[kMaxCells + 1][8];   // octant counts in [a, start of the cell),
                                                 // then sorted position offsets
  __shared__ uint32_t wtot[kCellBlock / 64][9];  // per wave of a chunk: octant counts, cell starts
  __shared__ uint32_t pre[kCellBlock / 64][9];   // the same before each wave, from a
  __shared__ uint32_t carry[9];                  // the same before the chunk, from a
  __shared__ uint32_t found;
  // the last LSD pass's output: (key_t, val_t) after an odd number of passes
  const bool odd = (lsd_passes_of(bp) & 1u) != 0u;
  const uint32_t* __restrict__ sk = odd ? sk_odd : sk_even;
  const uint32_t* __restrict__ sv = odd ? sv_odd : sv_even;
  const uint32_t total = bp->key_total;
  const bool table_ok = total <= bp->table_cap && !(bp->err & kErrFineKey);
  const uint32_t r0 = blockIdx.x * kCellRange;
  if (r0 >= n) return;
  const uint32_t r1 = min(n, r0 + kCellRange);
  const uint32_t hi = min(n, r1 + kCellAhead);  // staged: positions [r0 - 1, hi)
#pragma unroll
  for (int q = 0; q < (kStage + kCellBlock - 1) / kCellBlock; ++q) {
    const uint32_t j = q * kCellBlock + threadIdx.x;
    const uint32_t p = r0 + j - 1u;  // wraps for r0 = 0, j = 0: not read
    if (j < (uint32_t)kStage && (r0 + j) >= 1u && p < hi) stk[j] = sk[p];
  }
  if (threadIdx.x == 0) found = 0xFFFFFFFFu;
  __syncthreads();
  // staged keys and keys in HBM are read by separate macros, so that each
  // load keeps its address space
#define CELLS_STARTS_LDS(p) ((p) == 0 || (stk[(p) + 1u - r0] >> 3) != (stk[(p) - r0] >> 3))
#define CELLS_STARTS_HBM(p) ((p) == 0 || (sk[p] >> 3) != (sk[(p) - 1u] >> 3))

  // a: the first cell start in [r0, r1)
  {
    uint32_t first = 0xFFFFFFFFu;
    for (uint32_t k = r0 + threadIdx.x; k < r1; k += kCellBlock)
      if (first == 0xFFFFFFFFu && CELLS_STARTS_LDS(k)) first = k;
    if (first != 0xFFFFFFFFu) atomicMin(&found, first);
  }
  __syncthreads();
  const uint32_t a = found;
  if (a == 0xFFFFFFFFu) return;  // inside a cell that started earlier: nothing owned
  __syncthreads();
  if (threadIdx.x == 0) found = n;
  __syncthreads();
  // b: the first cell start at or after r1 (staged keys first, then HBM)
  {
    uint32_t first = 0xFFFFFFFFu;
    for (uint32_t k = r1 + threadIdx.x; k < hi; k += kCellBlock)
      if (first == 0xFFFFFFFFu && CELLS_STARTS_LDS(k)) first = k;
    if (first != 0xFFFFFFFFu) atomicMin(&found, first);
  }
  __syncthreads();
  for (uint32_t k0 = hi; k0 < n; k0 += kCellBlock) {
    if (found != n) break;  // read by every thread after the same barrier
    __syncthreads();
    const uint32_t k = k0 + threadIdx.x;
    if (k < n && CELLS_STARTS_HBM(k)) atomicMin(&found, k);
    __syncthreads();
  }
  const uint32_t b = found;

  // [a, b) through the staged keys when they hold every key it reads: the
  // keys at a - 1 .. b - 1, and the key at b unless b = n
  if (b < hi || (b == n && hi == n))
    cells_place<true>(stk, sk, sv, n, r0, a, b, total, table_ok, inv, c_start, c_id, c_oct, wtot,
                      pre, carry, runs, runs_ctl, runs_cap, keys_out, vals_out, slot, table, pos, vel, vh,
                      pos_s, vel_s, vh_s);
  else
    cells_place<false>(stk, sk, sv, n, r0, a, b, total, table_ok, inv, c_start, c_id, c_oct, wtot,
                       pre, carry, runs, runs_ctl, runs_cap, keys_out, vals_out, slot, table, pos, vel, vh,
                       pos_s, vel_s, vh_s);
#undef CELLS_STARTS_LDS
#undef CELLS_STARTS_HBM
}

// The long runs of empty keys k_cells listed (runs_ctl[0] of them, at most
// runs_cap): table[f] = value for f in [f0, f1), every work-group taking a
// share of every run.
__global__ __launch_bounds__(256) void k_fill_runs(const uint4* __restrict__ runs,
                                                   uint32_t* __restrict__ runs_ctl,
                                                   uint32_t runs_cap,
                                                   uint32_t* __restrict__ table) {
  __shared__ uint4 lr[256];
  const uint32_t count = min(runs_ctl[0], runs_cap);
  const uint32_t stride = gridDim.x * 256u;
  for (uint32_t i0 = 0; i0 < count; i0 += 256) {
    if (i0 + threadIdx.x < count) lr[threadIdx.x] = runs[i0 + threadIdx.x];
    __syncthreads();
    const uint32_t m = min(count - i0, 256u);
    for (uint32_t i = 0; i < m; ++i) {
      const uint4 r = lr[i];
      for (uint32_t f = r.x + blockIdx.x * 256u + threadIdx.x; f < r.y; f += stride) table[f] = r.z;
    }
    __syncthreads();
  }
}

// Empties the list of long runs (before a cell pass run on its own; the sort
// empties it in its first kernel).
__global__ void k_clear_runs(uint32_t* __restrict__ runs_ctl) { runs_ctl[0] = 0; }

void cell_pass(clsph_state* s) {
  hipLaunchKernelGGL(k_cells, dim3((s->n + kCellRange - 1) / kCellRange), dim3(kCellBlock), 0,
                     s->ctx->stream, s->key_t, s->val_t, s->key_u, s->val_u, s->n, s->bp, s->sort_runs, s->sort_runs_ctl,
                     s->sort_runs_cap, s->key_b, s->val_b, s->slot, s->table, s->pos, s->vel,
                     s->vh, s->pos_s, s->vel_s, s->vh_s);
  hipLaunchKernelGGL(k_fill_runs, dim3(kFillBlocks), dim3(256), 0, s->ctx->stream, s->sort_runs,
                     s->sort_runs_ctl, s->sort_runs_cap, s->table);
}

void launch_cell_table(clsph_state* s) {
  if (s->n == 0) return;
  hipLaunchKernelGGL(k_clear_runs, dim3(1), dim3(1), 0, s->ctx->stream, s->sort_runs_ctl);
  cell_pass(s);
}

uint32_t sort_runs_cap_for(uint32_t table_cap) { return table_cap / kLongRun + 64u; }

void launch_sort(clsph_state* s) {
  if (s->n == 0) return;
  hipStream_t st = s->ctx->stream;
  const uint32_t tiles = sort_tiles_for(s->n);
  // pass p reads the previous pass's output: key_a (identity values), then
  // (key_t, val_t) and (key_u, val_u) in turn
  const uint32_t* kin = s->key_a;
  const uint32_t* vin = nullptr;
  // tiles per slice of the scan (64 slices): held in registers up to 32
  const uint32_t per = (tiles + kScanBlock / 16 - 1) / (kScanBlock / 16);
  auto scan = [&](uint32_t p) {
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(kMaxBins / 16), dim3(kScanBlock), 0, st, s->sort_hist, tiles,
                         s->bp, p, s->sort_totals);
    };
    if (per <= 8) go(k_lsd_scan<8>);
    else if (per <= 16) go(k_lsd_scan<16>);
    else if (per <= 32) go(k_lsd_scan<32>);
    else go(k_lsd_scan<0>);
  };
  for (uint32_t p = 0; p < s->sort_passes; ++p) {
    uint32_t* kout = (p & 1) ? s->key_u : s->key_t;
    uint32_t* vout = (p & 1) ? s->val_u : s->val_t;
    hipLaunchKernelGGL(k_lsd_count, dim3(tiles), dim3(kLsdBlock), 0, st, kin, s->n, s->bp, p,
                       s->sort_hist, s->sort_runs_ctl);
    scan(p);
    hipLaunchKernelGGL(k_lsd_scatter, dim3(tiles), dim3(kScatBlock), 0, st, kin, vin, s->n, s->bp,
                       p, s->sort_hist, s->sort_totals, kout, vout);
    kin = kout;
    vin = vout;
  }
  cell_pass(s);
}

}  // namespace clsph
