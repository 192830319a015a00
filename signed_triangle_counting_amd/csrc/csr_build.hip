// On-device CSR build and degree orientation for gfx950.
//
// Restates, as a GPU pipeline, what the reference does serially on the host in
// build_forward_csr (Signed_TC/Multicore Implementations/graph_io.h:231-402):
//   canonicalise (min,max) + drop self-loops   graph_io.h:191-195
//   sort + dedupe                              graph_io.h:241-259
//   degree                                     graph_io.h:261-268
//   (deg, id) rank                             graph_io.h:270-317
//   orient low -> high rank, offsets, scatter  graph_io.h:319-394
//   per-list sort                              tc_merge_sign_enu.cpp:70-126
// with these MI355X-first choices:
//   * every sort is a project-written stable LSD radix sort (8-bit digits,
//     wave64 ballot match for in-tile ranks, tiles staged in LDS so each
//     digit's run is written to consecutive addresses), so duplicate lines
//     keep file order and the first line's sign wins;
//   * vertices are relabelled by rank, so an out-list holds only higher ranks
//     and the source's out-list has a compact span for the LDS bitmap;
//   * the sign travels as bit 0 of each 32-bit column word (rank << 1 | neg):
//     4 bytes per oriented edge instead of the reference's 5;
//   * one sort of (src_rank << (id bits + 1) | col) yields sorted out-lists
//     directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "internal.h"

namespace stc {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << __lane_id()) - 1ull;
}

// Exclusive scan of one 2048-element tile; block total to sums[blockIdx].
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(
    const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
    uint32_t *__restrict__ sums, uint64_t n) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  const int tid = threadIdx.x;
  uint32_t v[kScanItems];
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    uint64_t idx = base + (uint64_t)tid * kScanItems + i;
    v[i] = idx < n ? in[idx] : 0u;
    acc += v[i];
  }
  // inclusive wave scan of per-thread totals
  uint32_t x = acc;
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  const int w = tid >> 6;
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t wprefix = 0, total = 0;
#pragma unroll
  for (int j = 0; j < kScanThreads / 64; j++) {
    uint32_t s = wsum[j];
    if (j < w) wprefix += s;
    total += s;
  }
  uint32_t run = wprefix + x - acc;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    uint64_t idx = base + (uint64_t)tid * kScanItems + i;
    if (idx < n) out[idx] = run;
    run += v[i];
  }
  if (tid == 0 && sums) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(
    uint32_t *__restrict__ out, const uint32_t *__restrict__ add, uint64_t n) {
  const uint32_t a = add[blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  for (int i = threadIdx.x; i < kScanTile; i += kScanThreads) {
    uint64_t idx = base + i;
    if (idx < n) out[idx] += a;
  }
}

// ---- radix sort --------------------------------------------------------------
// Stable LSD radix sort of 64-bit keys with 8-bit digits, three steps per digit
// and no waiting between workgroups:
//   count    each 4096-key tile counts its digits (LDS) -> counts[tile][digit]
//   offsets  column sums over blocks of tiles, one scan over (digit, block),
//            then each block turns its tiles' counts into output offsets
//            (in place): offs[tile][digit] = keys of smaller digits anywhere
//            + keys of this digit in earlier tiles
//   scatter  each tile ranks its keys per digit (wave64 match; waves own
//            consecutive key ranges, so equal digits keep their order),
//            stages the tile in LDS in digit order and writes each digit's
//            run to consecutive output addresses.
// Per digit the keys are read twice and written once (24 B per key); the
// counts matrix adds 1 KB per tile.
constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / 64;
constexpr int kRsRounds = 16;
constexpr int kRsWaveKeys = 64 * kRsRounds;      // 1024 keys per wave
constexpr int kRsTile = kRsThreads * kRsRounds;  // 4096 keys per tile
constexpr int kRsRadix = 256;
constexpr uint32_t kRsColBlocks = 256;           // blocks of tiles in the offsets step
constexpr uint32_t kRsSmallTiles = 32;           // up to here one workgroup does the offsets

// Lanes of the wave holding the same 8-bit digit as this lane (built from
// eight ballots); empty for an invalid lane.
__device__ __forceinline__ uint64_t wave_peers8(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    peers &= bit ? bb : ~bb;
  }
  return valid ? peers : 0ull;
}

// Exclusive scan of one value per thread over a 256-thread workgroup; wsum
// holds kRsWaves words of LDS.  Every thread must call it.
__device__ __forceinline__ uint32_t block_scan_excl256(uint32_t x, uint32_t *wsum) {
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint32_t pre = 0;
#pragma unroll
  for (int v = 0; v < kRsWaves; v++) pre += v < w ? wsum[v] : 0u;
  return pre + incl - x;
}

// Digit counts of one tile held in registers (row r of the tile: keys
// tbase + r*256 + threadIdx.x): one LDS add per key, or one per row of 64
// keys when the row shares its digit (the high digits of small ids).
__device__ __forceinline__ void count_rows(const uint64_t (&k)[kRsRounds], uint64_t n,
                                           uint64_t tbase, int shift, uint32_t mask,
                                           uint32_t *h /* LDS, zeroed */) {
  const int tid = threadIdx.x, lane = __lane_id();
#pragma unroll
  for (int r = 0; r < kRsRounds; r++) {
    const uint64_t i = tbase + (uint64_t)r * kRsThreads + tid;
    const bool valid = i < n;
    const uint64_t act = __ballot(valid);
    if (!act) break;  // wave-uniform: later rows are past n too
    const uint32_t d = (uint32_t)(k[r] >> shift) & mask;
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);  // lane 0 is valid when any is
    if (__ballot(valid && d != d0) == 0) {
      if (lane == 0) atomicAdd(&h[d0], (uint32_t)__popcll(act));
    } else if (valid) {
      atomicAdd(&h[d], 1u);
    }
  }
}

__device__ __forceinline__ void tile_count(const uint64_t *__restrict__ keys, uint64_t n,
                                           uint64_t tbase, int shift, uint32_t mask,
                                           uint32_t *h /* LDS, zeroed */) {
  uint64_t k[kRsRounds];
#pragma unroll
  for (int r = 0; r < kRsRounds; r++) {
    const uint64_t i = tbase + (uint64_t)r * kRsThreads + threadIdx.x;
    k[r] = i < n ? keys[i] : 0ull;
  }
  count_rows(k, n, tbase, shift, mask, h);
}

__global__ __launch_bounds__(kRsThreads) void k_rs_count(const uint64_t *__restrict__ keys,
                                                         uint64_t n, int shift, uint32_t mask,
                                                         uint32_t *__restrict__ counts) {
  __shared__ uint32_t h[kRsRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  tile_count(keys, n, (uint64_t)blockIdx.x * kRsTile, shift, mask, h);
  __syncthreads();
  counts[(uint64_t)blockIdx.x * kRsRadix + threadIdx.x] = h[threadIdx.x];
}

// Loops over a column of the counts matrix (stride kRsRadix words) read
// kColBatch words at once, so the loads of a thread overlap.
constexpr int kColBatch = 8;

// Per-digit sum of the tiles [b*tpb, (b+1)*tpb), thread = digit.
__global__ __launch_bounds__(kRsRadix) void k_rs_colsum(const uint32_t *__restrict__ counts,
                                                        uint32_t ntiles, uint32_t tpb,
                                                        uint32_t *__restrict__ bsum) {
  const uint32_t t0 = blockIdx.x * tpb, t1 = min(ntiles, t0 + tpb);
  const uint32_t *col = counts + threadIdx.x;
  uint32_t s = 0, t = t0;
  for (; t + kColBatch <= t1; t += kColBatch) {
    uint32_t x[kColBatch];
#pragma unroll
    for (int u = 0; u < kColBatch; u++) x[u] = col[(uint64_t)(t + u) * kRsRadix];
#pragma unroll
    for (int u = 0; u < kColBatch; u++) s += x[u];
  }
  for (; t < t1; t++) s += col[(uint64_t)t * kRsRadix];
  bsum[(uint64_t)blockIdx.x * kRsRadix + threadIdx.x] = s;
}

// Running exclusive sums down rows [r0, r1) of one column, in place, from run.
__device__ __forceinline__ uint32_t column_excl(uint32_t *col, uint32_t r0, uint32_t r1,
                                                uint32_t run) {
  uint32_t r = r0;
  for (; r + kColBatch <= r1; r += kColBatch) {
    uint32_t x[kColBatch];
#pragma unroll
    for (int u = 0; u < kColBatch; u++) x[u] = col[(uint64_t)(r + u) * kRsRadix];
#pragma unroll
    for (int u = 0; u < kColBatch; u++) {
      col[(uint64_t)(r + u) * kRsRadix] = run;
      run += x[u];
    }
  }
  for (; r < r1; r++) {
    const uint32_t x = col[(uint64_t)r * kRsRadix];
    col[(uint64_t)r * kRsRadix] = run;
    run += x;
  }
  return run;
}

// One workgroup: rows of per-digit counts (block sums, or the tile counts
// themselves for a small sort) -> offsets, rows[b][d] = keys of digits < d +
// keys of digit d in rows < b.
__global__ __launch_bounds__(kRsRadix) void k_rs_colscan(uint32_t *__restrict__ rows,
                                                         uint32_t nrows) {
  __shared__ uint32_t wsum[kRsWaves];
  uint32_t *col = rows + threadIdx.x;
  uint32_t total = 0, b = 0;
  for (; b + kColBatch <= nrows; b += kColBatch) {
    uint32_t x[kColBatch];
#pragma unroll
    for (int u = 0; u < kColBatch; u++) x[u] = col[(uint64_t)(b + u) * kRsRadix];
#pragma unroll
    for (int u = 0; u < kColBatch; u++) total += x[u];
  }
  for (; b < nrows; b++) total += col[(uint64_t)b * kRsRadix];
  column_excl(col, 0, nrows, block_scan_excl256(total, wsum));
}

// Tile counts -> tile output offsets, in place.
__global__ __launch_bounds__(kRsRadix) void k_rs_coloff(uint32_t *__restrict__ counts,
                                                        uint32_t ntiles, uint32_t tpb,
                                                        const uint32_t *__restrict__ bsum) {
  const uint32_t t0 = blockIdx.x * tpb, t1 = min(ntiles, t0 + tpb);
  column_excl(counts + threadIdx.x, t0, t1, bsum[(uint64_t)blockIdx.x * kRsRadix + threadIdx.x]);
}

// Workgroup b -> tile: workgroups are dealt to the 8 XCDs round-robin, so XCD
// x = b % 8 gets a contiguous range of tiles.  Neighbouring tiles write
// neighbouring pieces of each digit's output run, which then meet in one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t ntiles) {
  const uint32_t x = b & 7u, q = ntiles >> 3, r = ntiles & 7u;
  return x * q + min(x, r) + (b >> 3);
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint64_t *__restrict__ in, uint64_t *__restrict__ out, uint32_t n, int shift,
    uint32_t mask, const uint32_t *__restrict__ offs) {
  __shared__ uint64_t skeys[kRsTile];
  __shared__ uint32_t wcnt[kRsWaves][kRsRadix];  // per wave: running count, then wave offset
  __shared__ uint32_t lbase[kRsRadix];           // digit run start in the tile
  __shared__ uint32_t gdelta[kRsRadix];          // output index - tile index, per digit
  __shared__ uint32_t wsum[kRsWaves];
  const int tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
  for (int d = lane; d < kRsRadix; d += 64) wcnt[w][d] = 0;
  const uint32_t goff = offs[(uint64_t)t * kRsRadix + tid];
  __syncthreads();
  const uint64_t tbase = (uint64_t)t * kRsTile;
  const uint64_t wbase = tbase + (uint64_t)w * kRsWaveKeys;
  // rank each key among the wave's earlier keys of its digit
  uint64_t key[kRsRounds];
  uint32_t pos[kRsRounds];
#pragma unroll
  for (int r = 0; r < kRsRounds; r++) {
    const uint64_t i = wbase + (uint64_t)r * 64 + lane;
    key[r] = i < n ? in[i] : 0ull;
  }
#pragma unroll
  for (int r = 0; r < kRsRounds; r++) {
    const bool valid = wbase + (uint64_t)r * 64 + lane < n;
    const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
    const uint64_t peers = wave_peers8(d, valid);
    const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
    uint32_t old = 0;
    if (valid && rank == 0) {
      old = wcnt[w][d];
      wcnt[w][d] = old + (uint32_t)__popcll(peers);
    }
    const int leader = valid ? (int)__builtin_ctzll(peers) : lane;
    pos[r] = (uint32_t)__shfl((int)old, leader, 64) + rank;
  }
  __syncthreads();
  // thread d: the waves' offsets within digit d and the tile's count of d
  const uint32_t d = tid;
  uint32_t c = 0;
#pragma unroll
  for (int v = 0; v < kRsWaves; v++) {
    const uint32_t x = wcnt[v][d];
    wcnt[v][d] = c;
    c += x;
  }
  const uint32_t lb = block_scan_excl256(c, wsum);
  lbase[d] = lb;
  gdelta[d] = goff - lb;
  __syncthreads();
  // stage the tile in digit order, then write each digit's run contiguously
#pragma unroll
  for (int r = 0; r < kRsRounds; r++) {
    if (wbase + (uint64_t)r * 64 + lane < n) {
      const uint32_t dk = (uint32_t)(key[r] >> shift) & mask;
      skeys[lbase[dk] + wcnt[w][dk] + pos[r]] = key[r];
    }
  }
  __syncthreads();
  const uint32_t nvalid = (uint32_t)min((uint64_t)kRsTile, (uint64_t)n - tbase);
  for (uint32_t j = tid; j < nvalid; j += kRsThreads) {
    const uint64_t k = skeys[j];
    const uint32_t o = gdelta[(uint32_t)(k >> shift) & mask] + j;
    if (o < n) out[o] = k;
  }
}

// ---- build kernels ----------------------------------------------------------
constexpr uint64_t kInvalidKey = ~0ull;

// Canonical keys (lo << (nb + 1)) | (hi << 1) | neg of the lines of sort
// tiles [t0, t0 + gridDim.x) (4096 lines each), self-loops and lines with an
// id out of range as kInvalidKey (the latter also set *bad), and the tiles'
// counts of the first sort digit (key bits 1..8).  For a batch (ebase
// non-null) line i of graph g (ebase[g] <= i < ebase[g+1]) has ids below
// vbase[g+1] - vbase[g], shifted by vbase[g] into the union's id space; a
// workgroup finds the graphs of its first and last line once and each line
// searches only between them.  Runs per staged upload chunk, overlapping the
// copies of the next one.
__global__ __launch_bounds__(kRsThreads) void k_canon_tiles(
    const uint32_t *__restrict__ u, const uint32_t *__restrict__ v,
    const int8_t *__restrict__ s, uint64_t m, uint32_t n, int nb, uint32_t t0,
    const unsigned long long *__restrict__ ebase, const uint32_t *__restrict__ vbase,
    uint32_t k, uint64_t *__restrict__ key, uint32_t *__restrict__ counts,
    uint32_t *__restrict__ bad) {
  __shared__ uint32_t h[kRsRadix];
  __shared__ uint32_t s_g[2];
  const int tid = threadIdx.x;
  const uint32_t t = t0 + blockIdx.x;
  const uint64_t l0 = (uint64_t)t * kRsTile;
  const uint64_t l1 = min(m, l0 + kRsTile);
  h[tid] = 0;
  if (ebase && tid < 2) {
    const uint64_t line = tid ? l1 - 1 : l0;
    uint32_t lo = 0, hi = k;  // the last g with ebase[g] <= line
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ebase[mid] <= line) lo = mid;
      else hi = mid;
    }
    s_g[tid] = lo;
  }
  __syncthreads();
  uint64_t kk[kRsRounds];
  uint32_t flag = 0;
#pragma unroll
  for (int r = 0; r < kRsRounds; r++) {
    const uint64_t i = l0 + (uint64_t)r * kRsThreads + tid;
    kk[r] = kInvalidKey;
    if (i < l1) {
      uint32_t a = u[i], b = v[i];
      bool ok;
      if (ebase) {
        uint32_t lo = s_g[0], hi = s_g[1] + 1;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (ebase[mid] <= i) lo = mid;
          else hi = mid;
        }
        const uint32_t base = vbase[lo], ng = vbase[lo + 1] - base;
        ok = a < ng && b < ng;
        a += base;
        b += base;
      } else {
        ok = a < n && b < n;
      }
      flag |= ok ? 0u : 1u;
      if (ok && a != b) {
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        kk[r] = ((uint64_t)lo << (nb + 1)) | ((uint64_t)hi << 1) | (s[i] < 0 ? 1ull : 0ull);
      }
      key[i] = kk[r];
    }
  }
  if (__ballot(flag != 0) && __lane_id() == 0) atomicOr(bad, 1u);
  count_rows(kk, m, l0, 1, (1u << min(8, nb)) - 1u, h);
  __syncthreads();
  counts[(uint64_t)t * kRsRadix + tid] = h[tid];
}

__global__ void k_dedupe_flag(const uint64_t *__restrict__ key, uint64_t m,
                              uint32_t *__restrict__ flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = key[i];
    bool keep = k != kInvalidKey && (i == 0 || (key[i - 1] >> 1) != (k >> 1));
    flag[i] = keep ? 1u : 0u;
  }
}

// Upper-endpoint degrees from keys sorted by their upper endpoint (the first
// half of the canonical sort): one atomic per run of equal ids in a wave.
// Every valid line is counted, duplicates too; k_compact_degree takes the
// duplicates back off.
__global__ void k_hi_degree(const uint64_t *__restrict__ key, uint64_t m, int nb, uint32_t n,
                            uint32_t *__restrict__ deg, uint32_t *__restrict__ err) {
  const uint64_t mask = (1ull << nb) - 1ull;
  const uint32_t lane = __lane_id();
  const uint64_t mround = (m + 63) & ~63ull;  // keep waves converged
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mround;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = i < m ? key[i] : kInvalidKey;
    bool valid = k != kInvalidKey;
    uint32_t hi = valid ? (uint32_t)((k >> 1) & mask) : 0xFFFFFFFFu;
    if (valid && hi >= n) {
      atomicOr(err, 2u);  // a sorted key that no input line made
      valid = false;
      hi = 0xFFFFFFFFu;
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)hi, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || prev != hi);
    const uint64_t vmask = __ballot(valid);
    if (valid && ((heads >> lane) & 1u)) {
      const uint64_t above = heads & ~((2ull << lane) - 1ull);
      const uint64_t run = (above ? (1ull << __builtin_ctzll(above)) - 1ull : ~0ull) &
                           ~((1ull << lane) - 1ull);
      atomicAdd(&deg[hi], (uint32_t)__popcll(vmask & run));
    }
  }
}

// Keys sorted by (lo, hi): compact the first line of each pair into ukey, add
// lower-endpoint degrees (lo is sorted: one atomic per run of equal lo in a
// wave) and take each duplicate line back off its upper endpoint's degree.
__global__ void k_compact_degree(const uint64_t *__restrict__ key,
                                 const uint32_t *__restrict__ flag,
                                 const uint32_t *__restrict__ pos, uint64_t m,
                                 int nb, uint32_t n, uint64_t *__restrict__ ukey,
                                 uint32_t *__restrict__ deg, uint32_t *__restrict__ err) {
  const uint64_t mask = (1ull << nb) - 1ull;
  const uint32_t lane = __lane_id();
  const uint64_t mround = (m + 63) & ~63ull;  // keep waves converged
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mround;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const bool in = i < m;
    bool keep = false;
    uint32_t lo = 0xFFFFFFFFu;
    if (in) {
      keep = flag[i] != 0;
      const uint64_t k = key[i];
      lo = (uint32_t)((k >> (nb + 1)) & mask);
      const uint32_t hi = (uint32_t)((k >> 1) & mask);
      if (k != kInvalidKey && (lo >= n || hi >= n)) {
        atomicOr(err, 2u);  // a sorted key that no input line made
        keep = false;
      } else if (keep) {
        ukey[pos[i]] = k;
      } else if (k != kInvalidKey) {
        atomicSub(&deg[hi], 1u);  // a repeated pair, counted once per line above
      }
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)lo, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || prev != lo);
    const uint64_t kept = __ballot(keep);
    if ((heads >> lane) & 1u) {
      const uint64_t above = heads & ~((2ull << lane) - 1ull);
      const uint64_t run = (above ? (1ull << __builtin_ctzll(above)) - 1ull : ~0ull) &
                           ~((1ull << lane) - 1ull);
      const uint32_t c = (uint32_t)__popcll(kept & run);
      if (c) atomicAdd(&deg[lo], c);
    }
  }
}

__global__ void k_graph_of(const uint32_t *__restrict__ vbase, uint32_t k,
                           uint32_t *__restrict__ graph_of) {
  uint32_t g = blockIdx.x;
  if (g >= k) return;
  for (uint32_t x = vbase[g] + threadIdx.x; x < vbase[g + 1]; x += blockDim.x)
    graph_of[x] = g;
}

__global__ void k_vertex_keys(const uint32_t *__restrict__ deg,
                              const uint32_t *__restrict__ graph_of,
                              uint32_t n, int bn, int bd,
                              uint64_t *__restrict__ vkey,
                              uint32_t *__restrict__ maxdeg) {
  uint32_t local = 0;
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n;
       x += gridDim.x * blockDim.x) {
    uint64_t g = graph_of ? graph_of[x] : 0u;
    uint32_t d = deg[x];
    if (d > local) local = d;
    vkey[x] = (g << (bd + bn)) | ((uint64_t)d << bn) | x;
  }
  if (maxdeg) atomicMax(maxdeg, local);
}

__global__ void k_degmax(const uint32_t *__restrict__ deg, uint32_t n,
                         uint32_t *__restrict__ maxdeg) {
  uint32_t local = 0;
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n;
       x += gridDim.x * blockDim.x)
    local = max(local, deg[x]);
  atomicMax(maxdeg, local);
}

__global__ void k_rank(const uint64_t *__restrict__ vkey, uint32_t n, int bn,
                       uint32_t *__restrict__ rank, uint32_t *__restrict__ perm,
                       uint32_t *__restrict__ err) {
  const uint64_t mask = (1ull << bn) - 1ull;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += gridDim.x * blockDim.x) {
    uint32_t x = (uint32_t)(vkey[r] & mask);
    if (x >= n) {
      atomicOr(err, 2u);
      continue;
    }
    rank[x] = r;
    perm[r] = x;
  }
}

__global__ void k_orient(const uint64_t *__restrict__ ukey, uint64_t m, int nb, uint32_t n,
                         const uint32_t *__restrict__ rank,
                         uint64_t *__restrict__ key2, uint32_t *__restrict__ err) {
  const uint64_t mask = (1ull << nb) - 1ull;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = ukey[i];
    const uint32_t a = (uint32_t)((k >> (nb + 1)) & mask), b = (uint32_t)((k >> 1) & mask);
    if (a >= n || b >= n) {
      atomicOr(err, 2u);
      key2[i] = 0;
      continue;
    }
    uint32_t ra = rank[a];
    uint32_t rb = rank[b];
    uint32_t s = ra < rb ? ra : rb, d = ra < rb ? rb : ra;
    key2[i] = ((uint64_t)s << (nb + 1)) | ((uint64_t)d << 1) | (k & 1ull);
  }
}

// Split the (src, dst)-sorted keys into col / src and count out-degrees: the
// sources are sorted, so each run of equal sources in a wave adds its length
// with one atomic.
__global__ void k_split_csr(const uint64_t *__restrict__ key2, uint64_t m, int nb, uint32_t n,
                            uint32_t *__restrict__ col,
                            uint32_t *__restrict__ src,
                            uint32_t *__restrict__ fdeg, uint32_t *__restrict__ err) {
  const uint32_t lane = __lane_id();
  const uint64_t mround = (m + 63) & ~63ull;  // keep waves converged
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mround;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const bool in = i < m;
    uint32_t a = 0xFFFFFFFFu;
    if (in) {
      const uint64_t k = key2[i];
      a = (uint32_t)(k >> (nb + 1));
      col[i] = (uint32_t)(k & ((2ull << nb) - 1ull));
      src[i] = a;
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)a, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || prev != a);
    const uint64_t valid = __ballot(in);
    if (in && ((heads >> lane) & 1u)) {
      const uint64_t above = heads & ~((2ull << lane) - 1ull);
      const uint64_t run = (above ? (1ull << __builtin_ctzll(above)) - 1ull : ~0ull) &
                           ~((1ull << lane) - 1ull);
      if (a < n) atomicAdd(&fdeg[a], (uint32_t)__popcll(valid & run));
      else atomicOr(err, 2u);
    }
  }
}

// Out-degree class of every rank: min(ceil_log2 d+(v), 15), 0xFF for an
// empty list (the binning pass's target side of an edge's bucket).
__global__ void k_dcls(const uint32_t *__restrict__ off, uint32_t n, uint8_t *__restrict__ dcls) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = off[v + 1] - off[v];
    const uint32_t c = d <= 1 ? 0u : 32u - (uint32_t)__clz(d - 1u);
    dcls[v] = d ? (uint8_t)min(c, 15u) : (uint8_t)0xFF;
  }
}

// Binning bucket of every oriented edge e = (u -> v), in edge order (the
// count's binning pass then reads two bytes per edge instead of gathering
// the target's class); the entries from m up to mp (m rounded up to 8) are
// kNoEdgeBucket, so the pass reads whole 16-byte groups.
__global__ void k_edge_buckets(const uint32_t *__restrict__ off, const uint32_t *__restrict__ col,
                               const uint32_t *__restrict__ src,
                               const uint8_t *__restrict__ dcls, uint64_t m, uint64_t mp,
                               uint16_t *__restrict__ ebk) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mp;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint16_t b = kNoEdgeBucket;
    if (e < m) {
      const uint32_t da = off[src[e] + 1] - (uint32_t)e - 1u;
      const uint32_t vc = da ? dcls[col[e] >> 1] : 0xFFu;
      if (vc != 0xFFu) {
        const uint32_t c = da <= 1 ? 0u : 32u - (uint32_t)__clz(da - 1u);
        b = (uint16_t)(min(c, (uint32_t)kBucketSide - 1u) * kBucketSide + vc);
      }
    }
    ebk[e] = b;
  }
}

// Keys of the transposed index: (target rank, edge id); sorted on the target
// bits (stably, so edge ids stay ascending within a target).
// Each edge's in-edge record word, (length of N+(u) after v) << 1 | sign,
// is written here in edge order (sequential reads), so the split after the
// sort gathers one word per in-edge instead of src, then off, then col.
__global__ void k_tkeys(const uint32_t *__restrict__ col, const uint32_t *__restrict__ src,
                        const uint32_t *__restrict__ off, uint64_t m, uint32_t n,
                        uint64_t *__restrict__ key, uint32_t *__restrict__ rec,
                        uint32_t *__restrict__ err) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t w = col[i], a = src[i];
    if (a >= n || (w >> 1) >= n) {
      atomicOr(err, 2u);
      key[i] = i;
      rec[i] = 0;
      continue;
    }
    key[i] = ((uint64_t)(w >> 1) << 32) | i;
    rec[i] = ((off[a + 1] - (uint32_t)i - 1u) << 1) | (w & 1u);
  }
}

// In-edge records of the sorted keys (edge e = (u -> v): the part of N+(u)
// after v and the edge's sign), in-degrees, and each
// target's suffix-class
// mask: bit c of smask[v] is set when some in-edge's part after v is
// non-empty and of class min(ceil_log2 len, 15) -- the source side of the
// binning bucket, so a counting pass can tell from the target alone whether
// any in-edge falls in a given bucket.  Targets are sorted, so a wave sees
// runs of equal targets: a run with a different target on both sides inside
// the wave holds all of that target's in-edges and is stored directly; a run
// touching lane 0 or lane 63 may continue in a neighbouring wave and is
// added (or or-ed) atomically into the zeroed arrays.
__global__ void k_tsplit(const uint64_t *__restrict__ key, uint64_t m, uint32_t n,
                         const uint32_t *__restrict__ rec, InEdge *__restrict__ tin,
                         uint32_t *__restrict__ tdeg, uint32_t *__restrict__ smask,
                         uint32_t *__restrict__ err) {
  const uint32_t lane = __lane_id();
  const uint64_t mround = (m + 63) & ~63ull;  // keep waves converged
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mround;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const bool in = i < m;
    uint32_t v = 0xFFFFFFFFu, cbit = 0;
    if (in) {
      const uint64_t k = key[i];
      v = (uint32_t)(k >> 32);
      uint32_t e = (uint32_t)k;
      if (v >= n || e >= m) {
        atomicOr(err, 2u);
        v = 0xFFFFFFFFu;
        e = 0;
      }
      const uint32_t r = rec[e];
      const uint32_t len = r >> 1;
      tin[i] = InEdge{e + 1u, r};
      const uint32_t c = len <= 1 ? 0u : 32u - (uint32_t)__clz(len - 1u);
      cbit = len ? 1u << min(c, 15u) : 0u;
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)v, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || prev != v);
    const uint64_t valid = __ballot(in);
    uint64_t has[16];
#pragma unroll
    for (int c = 0; c < 16; c++) has[c] = __ballot((cbit >> c) & 1u);
    if (in && ((heads >> lane) & 1u)) {
      const uint64_t above = heads & ~((2ull << lane) - 1ull);
      const uint64_t run = (above ? (1ull << __builtin_ctzll(above)) - 1ull : ~0ull) &
                           ~((1ull << lane) - 1ull);
      const uint32_t cnt = (uint32_t)__popcll(valid & run);
      uint32_t mask = 0;
#pragma unroll
      for (int c = 0; c < 16; c++) mask |= (has[c] & run) ? 1u << c : 0u;
      if (v >= n) {
        // (an out-of-range key, reported above)
      } else if (lane > 0 && !(run >> 63)) {
        tdeg[v] = cnt;
        smask[v] = mask;
      } else {
        atomicAdd(&tdeg[v], cnt);
        if (mask) atomicOr(&smask[v], mask);
      }
    }
  }
}

// Words of the merge model, sum over oriented edges (u -> v) of d+(u) + d+(v),
// and the maximum out-degree, per vertex instead of per edge: v is the source
// of d+(v) edges and the target of din(v) edges, each of which counts d+(v),
// so the sum is sum_v d+(v) * (d+(v) + din(v)).
// Merge-model words, largest out-degree, and the largest rank span of an
// out-list of 1,025..2,048 entries (class 11: whether such a target list
// fits the pull kernels' bitmap window depends on its span).
__global__ void k_balg(const uint32_t *__restrict__ off, const uint32_t *__restrict__ col,
                       const uint32_t *__restrict__ tin_off, uint32_t n,
                       unsigned long long *__restrict__ words, uint32_t *__restrict__ maxout,
                       uint32_t *__restrict__ maxspan) {
  unsigned long long acc = 0;
  uint32_t mo = 0, ms = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t dout = off[v + 1] - off[v], din = tin_off[v + 1] - tin_off[v];
    acc += (unsigned long long)dout * ((unsigned long long)dout + din);
    mo = max(mo, dout);
    if (dout > 1024u && dout <= 2048u)
      ms = max(ms, (col[off[v + 1] - 1u] >> 1) - (col[off[v]] >> 1) + 1u);
  }
  // wave reduce then one atomic per wave
  for (int d = 32; d > 0; d >>= 1) {
    acc += __shfl_down(acc, d, 64);
    mo = max(mo, (uint32_t)__shfl_down(mo, d, 64));
    ms = max(ms, (uint32_t)__shfl_down(ms, d, 64));
  }
  if (__lane_id() == 0) {
    atomicAdd(words, acc);
    atomicMax(maxout, mo);
    if (ms) atomicMax(maxspan, ms);
  }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

inline int grid_for(uint64_t n, int threads, int cap = 65536) {
  uint64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > (uint64_t)cap) g = cap;
  return (int)g;
}

inline int bits_for(uint64_t x) {  // bits needed to represent values < x
  int b = 0;
  while (b < 64 && (1ull << b) < x) b++;
  return b;
}

}  // namespace

#define STC_TRY(expr)                         \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) return e_;          \
  } while (0)

size_t scan_level_bytes(uint64_t n, int l) {
  uint64_t cur = n;
  for (int k = 0; n && k < 4; k++) {
    const uint64_t tiles = (cur + kScanTile - 1) / kScanTile;
    if (k == l) return 2 * tiles * sizeof(uint32_t) + 256;
    if (tiles <= 1) break;
    cur = tiles;
  }
  return 0;
}

hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint64_t n,
                              ScanScratch &scr, hipStream_t s) {
  return [&]() -> hipError_t {
    if (n == 0) return hipSuccess;
    // level sizes
    uint64_t sizes[5];
    int levels = 0;
    uint64_t cur = n;
    while (true) {
      uint64_t tiles = (cur + kScanTile - 1) / kScanTile;
      sizes[levels++] = tiles;
      if (tiles <= 1 || levels == 4) break;
      cur = tiles;
    }
    if (sizes[levels - 1] > 1) return hipErrorInvalidValue;  // > 2048^4
    // level buffers: sums and scanned sums
    for (int l = 0; l < levels; l++) STC_TRY(scr.level[l].ensure(scan_level_bytes(n, l)));
    // down: scan tiles of input, collect sums
    const uint32_t *lin = in;
    uint32_t *lout = out;
    uint64_t len = n;
    for (int l = 0; l < levels; l++) {
      uint32_t *sums = scr.level[l].as<uint32_t>();
      hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)sizes[l]), dim3(kScanThreads),
                         0, s, lin, lout, sums, len);
      STC_TRY(hipGetLastError());
      if (sizes[l] <= 1) break;
      lin = sums;
      lout = sums + sizes[l];
      len = sizes[l];
    }
    // up: add scanned sums back
    for (int l = levels - 2; l >= 0; l--) {
      uint32_t *target = (l == 0) ? out : scr.level[l - 1].as<uint32_t>() + sizes[l - 1];
      uint64_t tlen = (l == 0) ? n : sizes[l - 1];
      const uint32_t *scanned = scr.level[l].as<uint32_t>() + sizes[l];
      hipLaunchKernelGGL(k_scan_add, dim3((unsigned)sizes[l]), dim3(kScanThreads), 0,
                         s, target, scanned, tlen);
      STC_TRY(hipGetLastError());
    }
    return hipSuccess;
  }();
}

size_t radix_scratch_bytes(uint64_t n) {
  const uint64_t ntiles = (n + kRsTile - 1) / kRsTile;
  const uint64_t tpb = (ntiles + kRsColBlocks - 1) / kRsColBlocks;
  const uint64_t nblk = tpb ? (ntiles + tpb - 1) / tpb : 0;
  return (size_t)(ntiles + nblk) * kRsRadix * sizeof(uint32_t);
}

hipError_t radix_sort_u64(uint64_t **keys, uint64_t **tmp, uint64_t n, int lo_bit, int nbits,
                          DevBuf &scratch, hipStream_t s, bool first_counted) {
  if (n <= 1 || nbits <= 0) return hipSuccess;
  if (n > 0xFFFFFFFFull || lo_bit + nbits > 64) return hipErrorInvalidValue;
  const uint32_t ntiles = (uint32_t)((n + kRsTile - 1) / kRsTile);
  const uint32_t tpb = (ntiles + kRsColBlocks - 1) / kRsColBlocks;  // tiles per block
  const uint32_t nblk = (ntiles + tpb - 1) / tpb;
  STC_TRY(scratch.ensure(radix_scratch_bytes(n)));
  uint32_t *counts = scratch.as<uint32_t>();
  uint32_t *bsum = counts + (size_t)ntiles * kRsRadix;
  for (int lo = lo_bit; lo < lo_bit + nbits; lo += 8) {
    const uint32_t mask = (1u << std::min(8, lo_bit + nbits - lo)) - 1u;
    if (!(first_counted && lo == lo_bit))  // else the keys' producer counted this digit
      hipLaunchKernelGGL(k_rs_count, dim3(ntiles), dim3(kRsThreads), 0, s, *keys, n, lo, mask,
                         counts);
    if (ntiles <= kRsSmallTiles) {  // one workgroup turns the tile counts into offsets
      hipLaunchKernelGGL(k_rs_colscan, dim3(1), dim3(kRsRadix), 0, s, counts, ntiles);
    } else {
      hipLaunchKernelGGL(k_rs_colsum, dim3(nblk), dim3(kRsRadix), 0, s, counts, ntiles, tpb,
                         bsum);
      hipLaunchKernelGGL(k_rs_colscan, dim3(1), dim3(kRsRadix), 0, s, bsum, nblk);
      hipLaunchKernelGGL(k_rs_coloff, dim3(nblk), dim3(kRsRadix), 0, s, counts, ntiles, tpb,
                         bsum);
    }
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, s, *keys, *tmp,
                       (uint32_t)n, lo, mask, counts);
    STC_TRY(hipGetLastError());
    std::swap(*keys, *tmp);
  }
  return hipSuccess;
}

static int id_bits(uint32_t n) { return std::max(1, bits_for(n)); }  // ids < 2^nb

hipError_t build_begin(const BuildInput &in, BuildBuffers &bb, std::string &err, hipStream_t st) {
  const uint32_t n = in.n;
  const uint64_t m_in = in.m_in;
  if (id_bits(n) > 31) {
    err = "vertex ids must be < 2^31";
    return hipErrorInvalidValue;
  }
  STC_TRY(bb.misc.ensure(4096));
  // misc: [0] bad id, [1] maxdeg, [2] maxout, [3] largest class-11 span, [4..5] words,
  // [6] out-of-range sorted id
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, bb.misc.as<uint32_t>(), (uint64_t)16,
                     0u);
  // off/fdeg/etc. sized even for m == 0
  STC_TRY(bb.off.ensure(((uint64_t)n + 1) * 4));
  STC_TRY(bb.fdeg.ensure(((uint64_t)n + 1) * 4));
  STC_TRY(bb.deg.ensure(((uint64_t)n + 1) * 4));
  STC_TRY(bb.rank.ensure(((uint64_t)n + 1) * 4));
  STC_TRY(bb.perm.ensure(((uint64_t)n + 1) * 4));
  STC_TRY(bb.vkeyA.ensure(((uint64_t)n + 1) * 8));
  STC_TRY(bb.vkeyB.ensure(((uint64_t)n + 1) * 8));
  const uint64_t mm = m_in ? m_in : 1;
  STC_TRY(bb.keyA.ensure(mm * 8));
  STC_TRY(bb.keyB.ensure(mm * 8));
  STC_TRY(bb.flag.ensure(mm * 4));
  STC_TRY(bb.pos.ensure(mm * 4));
  // the canonical sort's tile counts: its first digit is counted by build_canon
  STC_TRY(bb.sort.ensure(radix_scratch_bytes(mm)));
  return hipSuccess;
}

hipError_t build_canon(const BuildInput &in, uint64_t l0, uint64_t l1, BuildBuffers &bb,
                       hipStream_t st) {
  if (l1 > in.m_in) l1 = in.m_in;
  if (l0 >= l1) return hipSuccess;
  if (l0 % kRsTile) return hipErrorInvalidValue;  // ranges start on a sort tile
  const uint32_t t0 = (uint32_t)(l0 / kRsTile);
  const uint32_t t1 = (uint32_t)((l1 + kRsTile - 1) / kRsTile);
  hipLaunchKernelGGL(k_canon_tiles, dim3(t1 - t0), dim3(kRsThreads), 0, st, in.u, in.v, in.s,
                     in.m_in, in.n, id_bits(in.n), t0,
                     (const unsigned long long *)(in.k ? in.ebase : nullptr), in.vbase, in.k,
                     bb.keyA.as<uint64_t>(), bb.sort.as<uint32_t>(), bb.misc.as<uint32_t>());
  return hipGetLastError();
}

hipError_t build_finish(const BuildInput &in, BuildBuffers &bb, BuildResult &res,
                        std::string &err, hipStream_t st) {
  const uint32_t n = in.n, k = in.k;
  const uint64_t m_in = in.m_in;
  const uint32_t *vbase_dev = in.vbase;
  res = BuildResult{};
  res.n = n;
  res.m_in = m_in;
  const int T = 256;
  const int nb = id_bits(n);
  uint32_t *misc = bb.misc.as<uint32_t>();
  uint32_t *id_err = misc + 6;  // set by the kernels that index by sorted ids
  uint64_t *kA = bb.keyA.as<uint64_t>(), *kB = bb.keyB.as<uint64_t>();
  // canonical sort in two halves: by hi (upper-endpoint degrees are then runs),
  // then stably by lo
  uint32_t *deg = bb.deg.as<uint32_t>();
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1, T)), dim3(T), 0, st, deg,
                     (uint64_t)n + 1, 0u);
  STC_TRY(radix_sort_u64(&kA, &kB, m_in, 1, nb, bb.sort, st, /*first_counted=*/true));
  hipLaunchKernelGGL(k_hi_degree, dim3(grid_for(m_in, T)), dim3(T), 0, st, kA, m_in, nb, n, deg,
                     id_err);
  STC_TRY(hipGetLastError());
  STC_TRY(radix_sort_u64(&kA, &kB, m_in, 1 + nb, nb, bb.sort, st, false));
  uint32_t *flag = bb.flag.as<uint32_t>(), *pos = bb.pos.as<uint32_t>();
  hipLaunchKernelGGL(k_dedupe_flag, dim3(grid_for(m_in, T)), dim3(T), 0, st, kA, m_in,
                     flag);
  STC_TRY(hipGetLastError());
  STC_TRY(scan_exclusive_u32(flag, pos, m_in, bb.scr, st));
  uint32_t hostv[2] = {0, 0};
  if (m_in) {
    STC_TRY(hipMemcpyAsync(&hostv[0], pos + (m_in - 1), 4, hipMemcpyDeviceToHost, st));
    STC_TRY(hipMemcpyAsync(&hostv[1], flag + (m_in - 1), 4, hipMemcpyDeviceToHost, st));
  }
  uint32_t flags0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  STC_TRY(hipMemcpyAsync(flags0, misc, sizeof flags0, hipMemcpyDeviceToHost, st));
  STC_TRY(hipStreamSynchronize(st));
  if (flags0[0]) {
    err = k ? "a batch graph has an endpoint >= its n" : "an edge endpoint is >= n";
    return hipErrorInvalidValue;
  }
  const uint64_t m = m_in ? (uint64_t)hostv[0] + hostv[1] : 0;
  res.m = m;
  hipLaunchKernelGGL(k_compact_degree, dim3(grid_for(m_in, T)), dim3(T), 0, st, kA,
                     flag, pos, m_in, nb, n, kB, deg, id_err);
  STC_TRY(hipGetLastError());
  std::swap(kA, kB);  // unique keys now in kA[0..m)

  // rank by (graph, degree, id)
  uint32_t *graph_of = nullptr;
  int bg = 0;
  if (vbase_dev && k >= 1) {
    STC_TRY(bb.graph_of.ensure(((uint64_t)n + 1) * 4));
    graph_of = bb.graph_of.as<uint32_t>();
    hipLaunchKernelGGL(k_graph_of, dim3(k), dim3(T), 0, st, vbase_dev, k, graph_of);
    STC_TRY(hipGetLastError());
    bg = bits_for(k);
  }
  hipLaunchKernelGGL(k_degmax, dim3(grid_for(n, T, 4096)), dim3(T), 0, st, deg, n,
                     misc + 1);
  uint32_t maxdeg = 0;
  STC_TRY(hipMemcpyAsync(&maxdeg, misc + 1, 4, hipMemcpyDeviceToHost, st));
  STC_TRY(hipStreamSynchronize(st));
  res.max_deg = maxdeg;
  const int bd = std::max(1, bits_for((uint64_t)maxdeg + 1));
  const int bn = nb;
  if (bg + bd + bn > 64) {
    err = "vertex key exceeds 64 bits";
    return hipErrorInvalidValue;
  }
  uint64_t *vA = bb.vkeyA.as<uint64_t>(), *vB = bb.vkeyB.as<uint64_t>();
  hipLaunchKernelGGL(k_vertex_keys, dim3(grid_for(n, T)), dim3(T), 0, st, deg, graph_of,
                     n, bn, bd, vA, (uint32_t *)nullptr);
  STC_TRY(hipGetLastError());
  STC_TRY(radix_sort_u64(&vA, &vB, n, 0, bg + bd + bn, bb.sort, st, false));
  uint32_t *rank = bb.rank.as<uint32_t>(), *perm = bb.perm.as<uint32_t>();
  hipLaunchKernelGGL(k_rank, dim3(grid_for(n, T)), dim3(T), 0, st, vA, n, bn, rank, perm,
                     id_err);
  STC_TRY(hipGetLastError());

  // orient and sort by (src, dst)
  uint32_t *fdeg = bb.fdeg.as<uint32_t>();
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1, T)), dim3(T), 0, st, fdeg,
                     (uint64_t)n + 1, 0u);
  hipLaunchKernelGGL(k_orient, dim3(grid_for(m, T)), dim3(T), 0, st, kA, m, nb, n, rank, kB,
                     id_err);
  STC_TRY(hipGetLastError());
  std::swap(kA, kB);
  STC_TRY(radix_sort_u64(&kA, &kB, m, 1, 2 * nb, bb.sort, st, false));
  STC_TRY(bb.col.ensure((m ? m : 1) * 4));
  STC_TRY(bb.src.ensure((m ? m : 1) * 4));
  hipLaunchKernelGGL(k_split_csr, dim3(grid_for(m, T)), dim3(T), 0, st, kA, m, nb, n,
                     bb.col.as<uint32_t>(), bb.src.as<uint32_t>(), fdeg, id_err);
  STC_TRY(hipGetLastError());
  STC_TRY(scan_exclusive_u32(fdeg, bb.off.as<uint32_t>(), (uint64_t)n + 1, bb.scr, st));
  // out-degree class per rank (binning reads one byte per edge's target)
  STC_TRY(bb.dcls.ensure((uint64_t)n + 1));
  hipLaunchKernelGGL(k_dcls, dim3(grid_for((uint64_t)n + 1, T)), dim3(T), 0, st,
                     bb.off.as<uint32_t>(), n, bb.dcls.as<uint8_t>());
  STC_TRY(hipGetLastError());
  const uint64_t mp = (m + 7) & ~7ull;
  STC_TRY(bb.ebk.ensure((mp ? mp : 8) * 2));
  hipLaunchKernelGGL(k_edge_buckets, dim3(grid_for(mp, T)), dim3(T), 0, st,
                     bb.off.as<uint32_t>(), bb.col.as<uint32_t>(), bb.src.as<uint32_t>(),
                     bb.dcls.as<uint8_t>(), m, mp, bb.ebk.as<uint16_t>());
  STC_TRY(hipGetLastError());
  // transposed index for the pull strategy: in-edge records grouped by target rank
  STC_TRY(bb.tin.ensure((m ? m : 1) * sizeof(InEdge)));
  STC_TRY(bb.tin_off.ensure(((uint64_t)n + 1) * 4));
  // the dedupe flags are dead by now: their buffer holds the in-edge record words
  uint32_t *rec = bb.flag.as<uint32_t>();
  hipLaunchKernelGGL(k_tkeys, dim3(grid_for(m, T)), dim3(T), 0, st, bb.col.as<uint32_t>(),
                     bb.src.as<uint32_t>(), bb.off.as<uint32_t>(), m, n, kA, rec, id_err);
  STC_TRY(hipGetLastError());
  STC_TRY(radix_sort_u64(&kA, &kB, m, 32, nb, bb.sort, st, false));
  STC_TRY(bb.smask.ensure(((uint64_t)n + 1) * 4));
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1, T)), dim3(T), 0, st, deg, (uint64_t)n + 1,
                     0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1, T)), dim3(T), 0, st,
                     bb.smask.as<uint32_t>(), (uint64_t)n + 1, 0u);
  hipLaunchKernelGGL(k_tsplit, dim3(grid_for(m, T)), dim3(T), 0, st, kA, m, n, rec,
                     bb.tin.as<InEdge>(), deg, bb.smask.as<uint32_t>(), id_err);
  STC_TRY(hipGetLastError());
  STC_TRY(scan_exclusive_u32(deg, bb.tin_off.as<uint32_t>(), (uint64_t)n + 1, bb.scr, st));
  // algorithmic bytes and max out-degree
  hipLaunchKernelGGL(k_balg, dim3(grid_for(n, T, 4096)), dim3(T), 0, st, bb.off.as<uint32_t>(),
                     bb.col.as<uint32_t>(), bb.tin_off.as<uint32_t>(), n,
                     (unsigned long long *)(misc + 4), misc + 2, misc + 3);
  STC_TRY(hipGetLastError());
  uint32_t fin[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  STC_TRY(hipMemcpyAsync(fin, misc, sizeof fin, hipMemcpyDeviceToHost, st));
  STC_TRY(hipStreamSynchronize(st));
  if (fin[6]) {  // a build kernel met (and skipped) a sorted key with an out-of-range id
    err = "csr build: a sorted key holds an out-of-range id";
    return hipErrorUnknown;
  }
  const uint32_t mo = fin[2];
  unsigned long long words = 0;
  std::memcpy(&words, fin + 4, 8);
  res.max_out = mo;
  res.max_span_c11 = fin[3];
  res.b_alg = 4.0 * (double)words + 4.0 * (double)m + 8.0 * ((double)n + 1.0);
  return hipSuccess;
}

}  // namespace stc
