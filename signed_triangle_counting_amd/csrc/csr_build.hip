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
//     wave64 ballot match for in-tile ranks), so duplicate lines keep file
//     order and the first line's sign wins;
//   * vertices are relabelled by rank, so an out-list holds only higher ranks
//     and the source's out-list has a compact span for the LDS bitmap;
//   * the sign travels as bit 0 of each 32-bit column word (rank << 1 | neg):
//     4 bytes per oriented edge instead of the reference's 5;
//   * one sort of (src_rank << 32 | col) yields sorted out-lists directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

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

// ---- radix sort -------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsRounds = 16;
constexpr int kRsTile = kRsThreads * kRsRounds;  // 4096 keys

// Rank of this lane among lanes of the wave holding the same 8-bit digit
// (wave64 match built from eight ballots), and the size of that group.
__device__ __forceinline__ void wave_match8(uint32_t d, bool valid,
                                            uint32_t &rank, uint32_t &cnt) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; b++) {
    bool bit = (d >> b) & 1u;
    uint64_t bb = __ballot(bit);
    peers &= bit ? bb : ~bb;
  }
  rank = (uint32_t)__popcll(peers & lanemask_lt());
  cnt = (uint32_t)__popcll(peers);
}

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(
    const uint64_t *__restrict__ keys, uint64_t n, int shift,
    uint32_t *__restrict__ hist, uint32_t ntiles) {
  __shared__ uint32_t h[256];
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  for (int r = 0; r < kRsRounds; r++) {
    uint64_t i = base + (uint64_t)r * kRsThreads + tid;
    bool valid = i < n;
    uint32_t d = valid ? (uint32_t)((keys[i] >> shift) & 255u) : 0u;
    uint32_t rank, cnt;
    wave_match8(d, valid, rank, cnt);
    if (valid && rank == 0) atomicAdd(&h[d], cnt);
  }
  __syncthreads();
  hist[(uint64_t)tid * ntiles + blockIdx.x] = h[tid];
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint64_t *__restrict__ in, uint64_t *__restrict__ out, uint64_t n,
    int shift, const uint32_t *__restrict__ offs, uint32_t ntiles) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[kRsThreads / 64][256];
  __shared__ uint32_t woff[kRsThreads / 64][256];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  run[tid] = offs[(uint64_t)tid * ntiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < kRsThreads / 64; w++) wcnt[w][tid] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  for (int r = 0; r < kRsRounds; r++) {
    uint64_t i = base + (uint64_t)r * kRsThreads + tid;
    bool valid = i < n;
    uint64_t key = valid ? in[i] : 0ull;
    uint32_t d = (uint32_t)((key >> shift) & 255u);
    uint32_t rank, cnt;
    wave_match8(d, valid, rank, cnt);
    if (valid && rank == 0) wcnt[wave][d] = cnt;
    __syncthreads();
    {
      uint32_t acc = run[tid];
#pragma unroll
      for (int w = 0; w < kRsThreads / 64; w++) {
        uint32_t c = wcnt[w][tid];
        wcnt[w][tid] = 0;
        woff[w][tid] = acc;
        acc += c;
      }
      run[tid] = acc;
    }
    __syncthreads();
    if (valid) out[woff[wave][d] + rank] = key;
  }
}

// ---- build kernels ----------------------------------------------------------
constexpr uint64_t kInvalidKey = ~0ull;

__global__ void k_canon(const uint32_t *__restrict__ u,
                        const uint32_t *__restrict__ v,
                        const int8_t *__restrict__ s, uint64_t m, uint32_t n,
                        int nb, uint64_t *__restrict__ key,
                        uint32_t *__restrict__ bad) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t a = u[i], b = v[i];
    uint64_t k;
    if (a >= n || b >= n) {
      atomicOr(bad, 1u);
      k = kInvalidKey;
    } else if (a == b) {
      k = kInvalidKey;
    } else {
      uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
      k = ((uint64_t)lo << (nb + 1)) | ((uint64_t)hi << 1) | (s[i] < 0 ? 1ull : 0ull);
    }
    key[i] = k;
  }
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

__global__ void k_compact_degree(const uint64_t *__restrict__ key,
                                 const uint32_t *__restrict__ flag,
                                 const uint32_t *__restrict__ pos, uint64_t m,
                                 int nb, uint64_t *__restrict__ ukey,
                                 uint32_t *__restrict__ deg) {
  const uint64_t mask = (1ull << nb) - 1ull;
  const uint32_t lane = __lane_id();
  const uint64_t mround = (m + 63) & ~63ull;  // keep waves converged
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mround;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const bool in = i < m;
    bool keep = false;
    uint32_t lo = 0xFFFFFFFFu, hi = 0xFFFFFFFFu;
    if (in) {
      keep = flag[i] != 0;
      const uint64_t k = key[i];
      lo = (uint32_t)((k >> (nb + 1)) & mask);
      if (keep) {
        ukey[pos[i]] = k;
        hi = (uint32_t)((k >> 1) & mask);
      }
    }
    // upper endpoints are scattered, and a hub is the upper endpoint of many
    // lanes of every wave: match equal values over the wave bit by bit (one
    // ballot per id bit), so each distinct value takes one atomic with its
    // lane count and a hub's counter sees one atomic per wave
    uint64_t peers = __ballot(keep);
    for (int b = 0; b < nb; b++) {
      const bool bit = (hi >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    if (keep && (peers & lanemask_lt()) == 0) atomicAdd(&deg[hi], (uint32_t)__popcll(peers));
    // lower endpoints are sorted: one atomic per run of equal lo in the wave
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
                       uint32_t *__restrict__ rank, uint32_t *__restrict__ perm) {
  const uint64_t mask = (1ull << bn) - 1ull;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += gridDim.x * blockDim.x) {
    uint32_t x = (uint32_t)(vkey[r] & mask);
    rank[x] = r;
    perm[r] = x;
  }
}

__global__ void k_orient(const uint64_t *__restrict__ ukey, uint64_t m, int nb,
                         const uint32_t *__restrict__ rank,
                         uint64_t *__restrict__ key2) {
  const uint64_t mask = (1ull << nb) - 1ull;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = ukey[i];
    uint32_t ra = rank[(k >> (nb + 1)) & mask];
    uint32_t rb = rank[(k >> 1) & mask];
    uint32_t s = ra < rb ? ra : rb, d = ra < rb ? rb : ra;
    key2[i] = ((uint64_t)s << 32) | ((uint64_t)d << 1) | (k & 1ull);
  }
}

// Split the (src, dst)-sorted keys into col / src and count out-degrees: the
// sources are sorted, so each run of equal sources in a wave adds its length
// with one atomic.
__global__ void k_split_csr(const uint64_t *__restrict__ key2, uint64_t m,
                            uint32_t *__restrict__ col,
                            uint32_t *__restrict__ src,
                            uint32_t *__restrict__ fdeg) {
  const uint32_t lane = __lane_id();
  const uint64_t mround = (m + 63) & ~63ull;  // keep waves converged
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mround;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const bool in = i < m;
    uint32_t a = 0xFFFFFFFFu;
    if (in) {
      const uint64_t k = key2[i];
      a = (uint32_t)(k >> 32);
      col[i] = (uint32_t)k;
      src[i] = a;
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)a, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || prev != a);
    const uint64_t valid = __ballot(in);
    if (in && ((heads >> lane) & 1u)) {
      const uint64_t above = heads & ~((2ull << lane) - 1ull);
      const uint64_t run = (above ? (1ull << __builtin_ctzll(above)) - 1ull : ~0ull) &
                           ~((1ull << lane) - 1ull);
      atomicAdd(&fdeg[a], (uint32_t)__popcll(valid & run));
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

// Keys of the transposed index: (target rank, edge id); sorted on the target
// bits (stably, so edge ids stay ascending within a target).
// Each edge's in-edge record word, (length of N+(u) after v) << 1 | sign,
// is written here in edge order (sequential reads), so the split after the
// sort gathers one word per in-edge instead of src, then off, then col.
__global__ void k_tkeys(const uint32_t *__restrict__ col, const uint32_t *__restrict__ src,
                        const uint32_t *__restrict__ off, uint64_t m,
                        uint64_t *__restrict__ key, uint32_t *__restrict__ rec) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t w = col[i];
    key[i] = ((uint64_t)(w >> 1) << 32) | i;
    rec[i] = ((off[src[i] + 1] - (uint32_t)i - 1u) << 1) | (w & 1u);
  }
}

// In-edge records of the sorted keys (edge e = (u -> v): the part of N+(u)
// after v and the edge's sign), in-degrees, and each target's suffix-class
// mask: bit c of smask[v] is set when some in-edge's part after v is
// non-empty and of class min(ceil_log2 len, 15) -- the source side of the
// binning bucket, so a counting pass can tell from the target alone whether
// any in-edge falls in a given bucket.  Targets are sorted, so a wave sees
// runs of equal targets: a run with a different target on both sides inside
// the wave holds all of that target's in-edges and is stored directly; a run
// touching lane 0 or lane 63 may continue in a neighbouring wave and is
// added (or or-ed) atomically into the zeroed arrays.
__global__ void k_tsplit(const uint64_t *__restrict__ key, uint64_t m,
                         const uint32_t *__restrict__ rec, InEdge *__restrict__ tin,
                         uint32_t *__restrict__ tdeg, uint32_t *__restrict__ smask) {
  const uint32_t lane = __lane_id();
  const uint64_t mround = (m + 63) & ~63ull;  // keep waves converged
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mround;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const bool in = i < m;
    uint32_t v = 0xFFFFFFFFu, cbit = 0;
    if (in) {
      const uint64_t k = key[i];
      v = (uint32_t)(k >> 32);
      const uint32_t e = (uint32_t)k;
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
      if (lane > 0 && !(run >> 63)) {
        tdeg[v] = cnt;
        smask[v] = mask;
      } else {
        atomicAdd(&tdeg[v], cnt);
        if (mask) atomicOr(&smask[v], mask);
      }
    }
  }
}

// Algorithmic bytes of the merge model and the maximum out-degree.
__global__ void k_balg(const uint32_t *__restrict__ off,
                       const uint32_t *__restrict__ col, const uint32_t *__restrict__ src,
                       uint64_t m, unsigned long long *__restrict__ words,
                       uint32_t *__restrict__ maxout) {
  unsigned long long acc = 0;
  uint32_t mo = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t a = src[i], b = col[i] >> 1;
    uint32_t da = off[a + 1] - off[a], db = off[b + 1] - off[b];
    acc += (unsigned long long)da + db;
    mo = max(mo, da);
  }
  // wave reduce then one atomic per wave
  for (int d = 32; d > 0; d >>= 1) {
    acc += __shfl_down(acc, d, 64);
    mo = max(mo, (uint32_t)__shfl_down(mo, d, 64));
  }
  if (__lane_id() == 0) {
    atomicAdd(words, acc);
    atomicMax(maxout, mo);
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

// Batched input: line i of the concatenated edge list belongs to graph
// g = upper_bound(ebase, i) - 1; its ids must be < vbase[g+1] - vbase[g] and
// are shifted by vbase[g] into the union's id space.
__global__ void k_shift_ids(uint32_t *__restrict__ u, uint32_t *__restrict__ v, uint64_t m,
                            const unsigned long long *__restrict__ ebase,
                            const uint32_t *__restrict__ vbase, uint32_t k,
                            uint32_t *__restrict__ bad) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = k;  // find the last g with ebase[g] <= i
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) >> 1;
      if (ebase[mid] <= i) lo = mid;
      else hi = mid;
    }
    const uint32_t n = vbase[lo + 1] - vbase[lo];
    const uint32_t a = u[i], b = v[i];
    if (a >= n || b >= n) atomicOr(bad, 1u);
    u[i] = a + vbase[lo];
    v[i] = b + vbase[lo];
  }
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

hipError_t radix_sort_u64(uint64_t **keys, uint64_t **tmp, uint64_t n,
                          const std::vector<int> &shifts, DevBuf &hist,
                          ScanScratch &scr, hipStream_t s) {
  if (n <= 1 || shifts.empty()) return hipSuccess;
  const uint32_t ntiles = (uint32_t)((n + kRsTile - 1) / kRsTile);
  const uint64_t hn = 256ull * ntiles;
  STC_TRY(hist.ensure(2 * hn * sizeof(uint32_t)));
  uint32_t *h = hist.as<uint32_t>();
  uint32_t *ho = h + hn;
  for (int sh : shifts) {
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kRsThreads), 0, s, *keys, n,
                       sh, h, ntiles);
    STC_TRY(hipGetLastError());
    STC_TRY(scan_exclusive_u32(h, ho, hn, scr, s));
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, s, *keys,
                       *tmp, n, sh, ho, ntiles);
    STC_TRY(hipGetLastError());
    std::swap(*keys, *tmp);
  }
  return hipSuccess;
}

static std::vector<int> shifts_for(int lo_bit, int nbits) {
  std::vector<int> v;
  for (int b = 0; b < nbits; b += 8) v.push_back(lo_bit + b);
  return v;
}

hipError_t build_csr(const uint32_t *u, const uint32_t *v, const int8_t *s,
                     uint64_t m_in, uint32_t n, const uint32_t *vbase_dev,
                     uint32_t k, BuildBuffers &bb, BuildResult &res,
                     std::string &err, hipStream_t st) {
  res = BuildResult{};
  res.n = n;
  res.m_in = m_in;
  const int T = 256;
  const int nb = std::max(1, bits_for(n));  // ids < 2^nb
  if (nb > 31) {
    err = "vertex ids must be < 2^31";
    return hipErrorInvalidValue;
  }
  STC_TRY(bb.misc.ensure(4096));
  uint32_t *misc = bb.misc.as<uint32_t>();  // [0] bad id, [1] maxdeg, [2] maxout, [4..5] words
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, misc, (uint64_t)16, 0u);

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

  uint64_t *kA = bb.keyA.as<uint64_t>(), *kB = bb.keyB.as<uint64_t>();
  hipLaunchKernelGGL(k_canon, dim3(grid_for(m_in, T)), dim3(T), 0, st, u, v, s, m_in,
                     n, nb, kA, misc + 0);
  STC_TRY(hipGetLastError());
  STC_TRY(radix_sort_u64(&kA, &kB, m_in, shifts_for(1, 2 * nb), bb.hist, bb.scr, st));
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
  uint32_t bad = 0;
  STC_TRY(hipMemcpyAsync(&bad, misc + 0, 4, hipMemcpyDeviceToHost, st));
  STC_TRY(hipStreamSynchronize(st));
  if (bad) {
    err = "an edge endpoint is >= n";
    return hipErrorInvalidValue;
  }
  const uint64_t m = m_in ? (uint64_t)hostv[0] + hostv[1] : 0;
  res.m = m;
  uint32_t *deg = bb.deg.as<uint32_t>();
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1, T)), dim3(T), 0, st, deg,
                     (uint64_t)n + 1, 0u);
  hipLaunchKernelGGL(k_compact_degree, dim3(grid_for(m_in, T)), dim3(T), 0, st, kA,
                     flag, pos, m_in, nb, kB, deg);
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
  STC_TRY(radix_sort_u64(&vA, &vB, n, shifts_for(0, bg + bd + bn), bb.hist, bb.scr, st));
  uint32_t *rank = bb.rank.as<uint32_t>(), *perm = bb.perm.as<uint32_t>();
  hipLaunchKernelGGL(k_rank, dim3(grid_for(n, T)), dim3(T), 0, st, vA, n, bn, rank, perm);
  STC_TRY(hipGetLastError());

  // orient and sort by (src, dst)
  uint32_t *fdeg = bb.fdeg.as<uint32_t>();
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1, T)), dim3(T), 0, st, fdeg,
                     (uint64_t)n + 1, 0u);
  hipLaunchKernelGGL(k_orient, dim3(grid_for(m, T)), dim3(T), 0, st, kA, m, nb, rank, kB);
  STC_TRY(hipGetLastError());
  std::swap(kA, kB);
  std::vector<int> sh = shifts_for(1, nb);
  for (int x : shifts_for(32, nb)) sh.push_back(x);
  STC_TRY(radix_sort_u64(&kA, &kB, m, sh, bb.hist, bb.scr, st));
  STC_TRY(bb.col.ensure((m ? m : 1) * 4));
  STC_TRY(bb.src.ensure((m ? m : 1) * 4));
  hipLaunchKernelGGL(k_split_csr, dim3(grid_for(m, T)), dim3(T), 0, st, kA, m,
                     bb.col.as<uint32_t>(), bb.src.as<uint32_t>(), fdeg);
  STC_TRY(hipGetLastError());
  STC_TRY(scan_exclusive_u32(fdeg, bb.off.as<uint32_t>(), (uint64_t)n + 1, bb.scr, st));
  // out-degree class per rank (binning reads one byte per edge's target)
  STC_TRY(bb.dcls.ensure((uint64_t)n + 1));
  hipLaunchKernelGGL(k_dcls, dim3(grid_for((uint64_t)n + 1, T)), dim3(T), 0, st,
                     bb.off.as<uint32_t>(), n, bb.dcls.as<uint8_t>());
  STC_TRY(hipGetLastError());
  // transposed index for the pull strategy: in-edge records grouped by target rank
  STC_TRY(bb.tin.ensure((m ? m : 1) * sizeof(InEdge)));
  STC_TRY(bb.tin_off.ensure(((uint64_t)n + 1) * 4));
  // the dedupe flags are dead by now: their buffer holds the in-edge record words
  uint32_t *rec = bb.flag.as<uint32_t>();
  hipLaunchKernelGGL(k_tkeys, dim3(grid_for(m, T)), dim3(T), 0, st, bb.col.as<uint32_t>(),
                     bb.src.as<uint32_t>(), bb.off.as<uint32_t>(), m, kA, rec);
  STC_TRY(hipGetLastError());
  STC_TRY(radix_sort_u64(&kA, &kB, m, shifts_for(32, nb), bb.hist, bb.scr, st));
  STC_TRY(bb.smask.ensure(((uint64_t)n + 1) * 4));
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1, T)), dim3(T), 0, st, deg, (uint64_t)n + 1,
                     0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1, T)), dim3(T), 0, st,
                     bb.smask.as<uint32_t>(), (uint64_t)n + 1, 0u);
  hipLaunchKernelGGL(k_tsplit, dim3(grid_for(m, T)), dim3(T), 0, st, kA, m, rec,
                     bb.tin.as<InEdge>(), deg, bb.smask.as<uint32_t>());
  STC_TRY(hipGetLastError());
  STC_TRY(scan_exclusive_u32(deg, bb.tin_off.as<uint32_t>(), (uint64_t)n + 1, bb.scr, st));
  // algorithmic bytes and max out-degree
  hipLaunchKernelGGL(k_balg, dim3(grid_for(m, T, 8192)), dim3(T), 0, st,
                     bb.off.as<uint32_t>(), bb.col.as<uint32_t>(), bb.src.as<uint32_t>(),
                     m, (unsigned long long *)(misc + 4), misc + 2);
  STC_TRY(hipGetLastError());
  uint32_t mo = 0;
  unsigned long long words = 0;
  STC_TRY(hipMemcpyAsync(&mo, misc + 2, 4, hipMemcpyDeviceToHost, st));
  STC_TRY(hipMemcpyAsync(&words, misc + 4, 8, hipMemcpyDeviceToHost, st));
  STC_TRY(hipStreamSynchronize(st));
  res.max_out = mo;
  res.b_alg = 4.0 * (double)words + 4.0 * (double)m + 8.0 * ((double)n + 1.0);
  return hipSuccess;
}

hipError_t shift_batch_ids(uint32_t *u, uint32_t *v, uint64_t m, const uint64_t *ebase_dev,
                           const uint32_t *vbase_dev, uint32_t k, uint32_t *bad_dev,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, bad_dev, (uint64_t)1, 0u);
  hipLaunchKernelGGL(k_shift_ids, dim3(grid_for(m, 256)), dim3(256), 0, st, u, v, m,
                     (const unsigned long long *)ebase_dev, vbase_dev, k, bad_dev);
  return hipGetLastError();
}

}  // namespace stc
